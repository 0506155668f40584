"""Rheologies (src/Physics.jl:5-47) as descriptors for the HIP library.

The reference evaluates `rheology(tauII, I)` inside update_V! (kernels_3d.jl:39).
Here a rheology is lowered to the fi_rheology record of the C ABI: an inner
law (linear constant / linear field / power law), an optional viscoplastic
cap and an optional relaxation toward the law.
"""
import numpy as np

from . import _abi as A


class IceRheology:
    pass


class LinearViscousRheology(IceRheology):
    """eta = eta[I]  (Physics.jl:7-11); eta is a number, an array or a callable f(x, y[, z])."""

    def __init__(self, eta):
        self.eta = eta


class PowerLawRheology(IceRheology):
    """eta = 0.5 / (A[I] * tauII^(n-1))  (Physics.jl:13-18)."""

    def __init__(self, A, n):
        self.A = A
        self.n = n


class ViscoplasticRegularisation(IceRheology):
    """eta = 1 / (1/viscosity(tauII, I) + 1/eta_reg)  (Physics.jl:20-26)."""

    def __init__(self, viscosity, eta_reg):
        if not np.isscalar(eta_reg) or eta_reg <= 0:
            raise ValueError("eta_reg must be a positive number")
        self.viscosity = viscosity
        self.eta_reg = float(eta_reg)
        self.inv_eta_reg = 1.0 / self.eta_reg


class Relaxation(IceRheology):
    """eta_new = field*(1-rate) + target*rate, optionally in log/log10 space (Physics.jl:28-47).

    The relaxed field is the model's lagged viscosity eta.
    """

    def __init__(self, field, target, rate, scale=None):
        if scale not in (None, "log", "log10", np.log, np.log10):
            raise ValueError("only `None`, `log` and `log10` are supported for the relaxation scale")
        self.field = field
        self.target = target
        self.rate = float(rate)
        self.scale = {np.log: "log", np.log10: "log10"}.get(scale, scale)


def _is_field_value(v):
    return callable(v) or isinstance(v, np.ndarray)


def lower(rheology):
    """Rheology -> (fi_rheology, field_value or None) for the C ABI."""
    r = A.fi_rheology()
    relax = None
    if isinstance(rheology, Relaxation):
        relax = rheology
        rheology = rheology.target
    reg = None
    if isinstance(rheology, ViscoplasticRegularisation):
        reg = rheology
        rheology = rheology.viscosity
    field = None
    if isinstance(rheology, LinearViscousRheology):
        if _is_field_value(rheology.eta):
            r.kind = A.FI_RHEO_LINEAR_FIELD
            field = rheology.eta
        else:
            r.kind = A.FI_RHEO_LINEAR_CONST
            r.eta = float(rheology.eta)
    elif isinstance(rheology, PowerLawRheology):
        r.kind = A.FI_RHEO_POWERLAW
        r.n = float(rheology.n)
        if _is_field_value(rheology.A):
            r.A_is_field = 1
            field = rheology.A
        else:
            r.A = float(rheology.A)
    else:
        raise TypeError("unsupported rheology %r" % (rheology,))
    if reg is not None:
        r.has_reg = 1
        r.eta_reg = reg.eta_reg
    if relax is not None:
        r.relax = {None: A.FI_RELAX_LIN, "log": A.FI_RELAX_LOG, "log10": A.FI_RELAX_LOG10}[relax.scale]
        r.rate = relax.rate
    return r, field

"""Shared set-up for the parity tests: a GPU model and an oracle model built
from the same descriptor and initial fields."""
import numpy as np


def rel_err(a, b):
    den = np.max(np.abs(b))
    num = np.max(np.abs(a - b))
    if den == 0.0:
        return float(num)
    return float(num / den)


def state_names(ndims):
    if ndims == 3:
        return ["P", "TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ", "VX", "VY", "VZ"]
    return ["P", "TXX", "TYY", "TXY", "VX", "VY"]


def make_pair(pkg, orc, kw, ib=False, eta=1.0, seed=None, amp=1e-2):
    """GPU model + oracle with identical inputs; optional seeded random state."""
    cls = pkg.IsothermalImmersedBoundaryFullStokesModel if ib else pkg.IsothermalFullStokesModel
    model = cls(**kw)
    ref = orc.OracleModel(model.desc)
    nd = kw["grid"].ndims
    for k in ("GX", "GY", "GZ")[:nd]:
        if model.desc.gravity_is_field:
            ref.set(k, model.field(k).parent())
    if model.desc.rheology.kind == 1 or (model.desc.rheology.kind == 2 and model.desc.rheology.A_is_field):
        ref.set("RHEO", model.field("RHEO").parent())
    if np.isscalar(eta):
        model.viscosity.eta.set(eta)
        model.viscosity.eta_next.set(eta)
        ref.set("ETA", eta)
        ref.set("ETA_NEXT", eta)
    else:
        model.viscosity.eta.set(eta)
        model.viscosity.eta_next.set(eta)
        ref.set("ETA", model.viscosity.eta.parent())
        ref.set("ETA_NEXT", model.viscosity.eta_next.parent())
    if seed is not None:
        rng = np.random.default_rng(seed)
        shape = kw["grid"].parent_shape()
        for name in state_names(nd):
            a = np.asfortranarray(amp * rng.standard_normal(shape))
            model.field(name).set(a)
            ref.set(name, a)
    return model, ref


def compare_state(model, ref, names, tol):
    worst = {}
    for name in names:
        b = ref.get(name)
        if not np.all(np.isfinite(b)):
            raise AssertionError("oracle state %s is not finite: the test case diverged" % name)
        worst[name] = rel_err(model.field(name).parent(), b)
    bad = {k: v for k, v in worst.items() if not v <= tol}
    return worst, bad

"""PT heat diffusion (src/Models/Diffusion/Heat) over libfasticehip.so.

HeatDiffusionModel(grid, boundary_conditions, physics, iter_params) follows
Heat.jl:16-40: a temperature T and its previous time level T_o at cell
centres, a heat flux q at the faces; boundary_conditions maps the sides
west / east / south / north / bottom / top to BoundaryCondition(Value, v)
(DirichletBC{HalfCell} on T) or BoundaryCondition(Flux, v)
(DirichletBC{FullCell} on q) as boundary_conditions.jl:1-69 does; a side
without an entry has none.  advance_iteration(model, t, dt) is one PT
iteration of update_q! / update_T! (kernels.jl:3-29, Heat.jl:42-58);
advance_timestep runs PT iterations to a residual tolerance and then
T_o <- T (Heat.jl exports it but does not define it).
"""
import ctypes as C
from types import SimpleNamespace

import numpy as np

from . import _abi as A
from .grids import UniformGrid


class Value:
    """Prescribed temperature on a side."""


class Flux:
    """Prescribed heat flux on a side."""


class BoundaryCondition:
    """BoundaryCondition{Kind}(condition) (boundary_conditions.jl:4-8)."""

    def __init__(self, kind, condition):
        if kind not in (Value, Flux):
            raise TypeError("kind must be Value or Flux")
        self.kind = kind
        self.condition = float(condition)


SIDES = (("west", "east"), ("south", "north"), ("bottom", "top"))


def _get(obj, *names):
    for n in names:
        if isinstance(obj, dict) and n in obj:
            return obj[n]
        if hasattr(obj, n):
            return getattr(obj, n)
    raise KeyError(names[0])


def build_heat_desc(grid, boundary_conditions, physics, iter_params):
    d = A.fi_heat_desc()
    nd = grid.ndims
    d.ndims = nd
    for q in range(3):
        d.n[q] = grid.dims[q] if q < nd else 0
        d.origin[q] = grid.origin[q] if q < nd else 0.0
        d.extent[q] = grid.extent_[q] if q < nd else 1.0
    for q in range(nd):
        for s in range(2):
            bc = (boundary_conditions or {}).get(SIDES[q][s])
            if bc is None:
                d.bc_kind[q][s] = A.FI_HEAT_BC_NONE
                continue
            d.bc_kind[q][s] = A.FI_HEAT_BC_VALUE if bc.kind is Value else A.FI_HEAT_BC_FLUX
            d.bc_value[q][s] = bc.condition
    d.lambda_rho_cp = float(_get(physics, "λ_ρCp", "lambda_rho_cp"))
    dtau = _get(iter_params, "Δτ", "dtau")
    d.dtau_q = float(_get(dtau, "q"))
    d.dtau_T = float(_get(dtau, "T"))
    return d


class HeatField:
    """One device field of a heat model (T, T_o, qx, qy or qz)."""

    def __init__(self, model, fid, bits):
        self.model = model
        self.fid = fid
        self.bits = bits

    def interior_shape(self):
        return tuple(n + ((self.bits >> d) & 1) for d, n in enumerate(self.model.grid.dims))

    def set(self, value):
        """set!(field, value): number, array (parent or interior shape) or callable f(x, y[, z])."""
        m = self.model
        if callable(value):
            value = np.asarray(value(*m.grid.parent_coords(self.bits)), dtype=np.float64)
        if np.isscalar(value):
            value = np.full(m.grid.parent_shape(), float(value))
        a = np.asfortranarray(np.asarray(value, dtype=np.float64))
        if a.shape == m.grid.parent_shape():
            layout = A.FI_LAYOUT_PARENT
        elif a.shape == self.interior_shape():
            layout = A.FI_LAYOUT_INTERIOR
        else:
            raise ValueError("array of shape %s fits neither the parent %s nor the interior %s"
                             % (a.shape, m.grid.parent_shape(), self.interior_shape()))
        A.check(A.lib().fi_heat_field_upload(m._h, self.fid, A.dptr(a), a.size, layout))
        return self

    def parent(self):
        m = self.model
        a = np.empty(m.grid.parent_shape(), dtype=np.float64, order="F")
        A.check(A.lib().fi_heat_field_download(m._h, self.fid, A.dptr(a), a.size, A.FI_LAYOUT_PARENT))
        return a

    def interior(self):
        m = self.model
        a = np.empty(self.interior_shape(), dtype=np.float64, order="F")
        A.check(A.lib().fi_heat_field_download(m._h, self.fid, A.dptr(a), a.size, A.FI_LAYOUT_INTERIOR))
        return a


class HeatDiffusionModel:
    """HeatDiffusionModel(; arch, grid, boundary_conditions, physics, iter_params, outer_width) (Heat.jl:29-40).

    physics: {"λ_ρCp": ...}; iter_params: {"Δτ": {"q": ..., "T": ...}}.
    outer_width has no equivalent (one device, nothing to overlap)."""

    def __init__(self, grid, boundary_conditions, physics, iter_params, arch=None, outer_width=None):
        if not isinstance(grid, UniformGrid):
            raise TypeError("grid must be a UniformGrid")
        self.arch = arch
        self.grid = grid
        self.boundary_conditions = boundary_conditions
        self.physics = physics
        self.iter_params = iter_params
        self.desc = build_heat_desc(grid, boundary_conditions, physics, iter_params)
        h = C.c_void_p()
        A.check(A.lib().fi_heat_create(C.byref(self.desc), C.byref(h)))
        self._h = h
        nd = grid.ndims
        self.temperature = HeatField(self, A.FI_HF_T, 0)
        self.temperature_old = HeatField(self, A.FI_HF_T_OLD, 0)
        q = [HeatField(self, A.FI_HF_QX + d, 1 << d) for d in range(nd)]
        self.heat_flux = SimpleNamespace(**dict(zip("xyz", q)))
        self.residual_norms_ = np.zeros(2)

    def apply_bc(self):
        """The Value conditions on T (after set!)."""
        A.check(A.lib().fi_heat_apply_bc(self._h))

    def advance(self, n_iter, dt):
        A.check(A.lib().fi_heat_advance_iterations(self._h, float(dt), int(n_iter)))

    def residual_norms(self, dt):
        out = np.zeros(2)
        A.check(A.lib().fi_heat_residual_norms(self._h, float(dt), A.dptr(out)))
        return out

    def step(self, dt, max_iter, ncheck, tol):
        it = C.c_int(0)
        out = np.zeros(2)
        A.check(A.lib().fi_heat_step(self._h, float(dt), int(max_iter), int(ncheck), float(tol), C.byref(it),
                                     A.dptr(out)))
        self.residual_norms_ = out
        return it.value, out

    def time_iterations(self, dt, n_iter):
        out = np.zeros(2)
        A.check(A.lib().fi_heat_time_iterations(self._h, float(dt), int(n_iter), A.dptr(out)))
        return out

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            A.lib().fi_heat_destroy(h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def advance_iteration(model, t, dt):
    """advance_iteration!(model, t, Δt) (Heat.jl:42-58); t is unused there too."""
    model.advance(1, dt)


def advance_timestep(model, t, dt, max_iter=None, ncheck=None, tol=1e-8):
    """PT iterations until both residual Inf-norms are below tol (or max_iter),
    then T_o <- T.  Returns (iterations, norms)."""
    n = max(model.grid.dims)
    return model.step(dt, max_iter if max_iter is not None else 1000 * n, ncheck or n, tol)

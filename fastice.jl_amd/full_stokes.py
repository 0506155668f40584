"""Isothermal full-Stokes models on the HIP library.

Mirrors FastIce.Models.FullStokes.Isothermal (src/Models/FullStokes/Isothermal/
Isothermal.jl:3-125) and FastIce.Models.ImmersedBoundaryFullStokes.Isothermal
(src/Models/ImmersedBoundaryFullStokes/Isothermal/Isothermal.jl:3-136):

    model = IsothermalFullStokesModel(grid=..., boundary_conditions=...,
                                      gravity=..., rheology=..., solver_params=...)
    advance_iteration(model, t, dt)       # advance_iteration!
    compute_residuals(model)              # compute_residuals! (+ Inf-norms)

Device fields are reached as in the reference: model.stress.P,
model.stress.tau.xx (alias model.stress.τ), model.velocity.V.x,
model.viscosity.eta / eta_next (aliases η / η_next), model.residual.r_P,
model.residual.r_V.x; each has set(value) and interior() / parent().
"""
import ctypes as C
from types import SimpleNamespace

import numpy as np

from . import _abi as A
from .grids import UniformGrid, bits_loc
from .physics import lower


class Traction:
    """Free surface / prescribed traction (boundary_conditions.jl:1)."""


class Velocity:
    """No-slip / prescribed velocity (boundary_conditions.jl:2)."""


class Slip:
    """Free slip (boundary_conditions.jl:3)."""


_KIND = {Traction: A.FI_BC_TRACTION, Velocity: A.FI_BC_VELOCITY, Slip: A.FI_BC_SLIP}


class BoundaryCondition:
    """BoundaryCondition{Kind}(components...)  (boundary_conditions.jl:8-13)."""

    def __init__(self, kind, *components):
        if kind not in _KIND:
            raise TypeError("kind must be Traction, Velocity or Slip")
        self.kind = kind
        if len(components) == 1 and isinstance(components[0], (tuple, list)):
            components = tuple(components[0])
        self.components = tuple(float(c) for c in components)

    def __repr__(self):
        return "BoundaryCondition{%s}%s" % (self.kind.__name__, self.components)


def TBC(*c):
    return BoundaryCondition(Traction, *c)


def VBC(*c):
    return BoundaryCondition(Velocity, *c)


def SBC(*c):
    return BoundaryCondition(Slip, *c)


def _get(d, *keys, default=None):
    for k in keys:
        if isinstance(d, dict) and k in d:
            return d[k]
        if hasattr(d, k):
            return getattr(d, k)
    return default


class DeviceField:
    """Handle of one device-resident field of a model."""

    def __init__(self, model, fid):
        self.model = model
        self.fid = fid
        self.bits = A.field_location(model.grid.ndims, fid)

    @property
    def location(self):
        return bits_loc(self.bits, self.model.grid.ndims)

    def interior_shape(self):
        g = self.model.grid
        return tuple(n + ((self.bits >> d) & 1) for d, n in enumerate(g.dims))

    def set(self, value):
        """set!(field, value): number, array (parent or interior shape) or callable f(x, y[, z])."""
        m = self.model
        if np.isscalar(value):
            A.check(A.lib().fi_field_fill(m._h, self.fid, float(value)))
            return self
        if callable(value):
            xs = m.grid.parent_coords(self.bits)
            value = np.asarray(value(*xs), dtype=np.float64)
            value = np.broadcast_to(value, m.grid.parent_shape())
        a = np.asfortranarray(np.asarray(value, dtype=np.float64))
        if a.shape == m.grid.parent_shape():
            layout = A.FI_LAYOUT_PARENT
        elif a.shape == self.interior_shape():
            layout = A.FI_LAYOUT_INTERIOR
        else:
            raise ValueError("array of shape %s fits neither the parent %s nor the interior %s"
                             % (a.shape, m.grid.parent_shape(), self.interior_shape()))
        A.check(A.lib().fi_field_upload(m._h, self.fid, A.dptr(a), a.size, layout))
        return self

    def parent(self):
        m = self.model
        a = np.empty(m.grid.parent_shape(), dtype=np.float64, order="F")
        A.check(A.lib().fi_field_download(m._h, self.fid, A.dptr(a), a.size, A.FI_LAYOUT_PARENT))
        return a

    def interior(self):
        m = self.model
        a = np.empty(self.interior_shape(), dtype=np.float64, order="F")
        A.check(A.lib().fi_field_download(m._h, self.fid, A.dptr(a), a.size, A.FI_LAYOUT_INTERIOR))
        return a

    def __repr__(self):
        return "DeviceField(%s, %s)" % (A.FIELD_NAMES[self.fid] if self.fid < len(A.FIELD_NAMES) else self.fid,
                                         self.location)


def _dtau(solver_params, ndims):
    dt = _get(solver_params, "Δτ", "dtau")
    if dt is None:
        raise ValueError("solver_params needs Δτ (dtau) = (Pr, τ=(xx, ...), V=(x, ...))")
    pr = float(_get(dt, "Pr", "P"))
    tau = _get(dt, "τ", "tau")
    v = _get(dt, "V")
    names = ("xx", "yy", "zz", "xy", "xz", "yz") if ndims == 3 else ("xx", "yy", "xy")
    taus = [float(_get(tau, k)) for k in names]
    vs = [float(_get(v, k)) for k in ("x", "y", "z")[:ndims]]
    return pr, taus, vs


def _bc_pairs(bcs, ndims):
    names = ("x", "y", "z")[:ndims]
    if isinstance(bcs, dict):
        pairs = [bcs[k] for k in names]
    else:
        pairs = [getattr(bcs, k) for k in names] if hasattr(bcs, "x") else list(bcs)
    if len(pairs) != ndims:
        raise ValueError("boundary_conditions needs one (low, high) pair per dimension")
    return pairs


def build_desc(grid, boundary_conditions, gravity, rheology, solver_params, immersed_boundary=False,
               precision=A.FI_PREC_FP64):
    """Lower the reference's model arguments to an fi_model_desc (+ field-valued parts)."""
    nd = grid.ndims
    d = A.fi_model_desc()
    d.ndims = nd
    for q in range(nd):
        d.n[q] = grid.dims[q]
        d.origin[q] = grid.origin[q]
        d.extent[q] = grid.extent_[q]
    for q, pair in enumerate(_bc_pairs(boundary_conditions, nd)):
        for s, bc in enumerate(pair):
            if not isinstance(bc, BoundaryCondition):
                raise TypeError("boundary condition must be a BoundaryCondition")
            if len(bc.components) != nd:
                raise ValueError("BoundaryCondition needs %d components" % nd)
            d.bc_kind[q][s] = _KIND[bc.kind]
            for c, v in enumerate(bc.components):
                d.bc_value[q][s][c] = v
    r, rfield = lower(rheology)
    d.rheology = r
    gfields = {}
    gnames = ("x", "y", "z")[:nd]
    for q, k in enumerate(gnames):
        gv = _get(gravity, k, default=0.0)
        if callable(gv) or isinstance(gv, np.ndarray):
            gfields[k] = gv
        else:
            d.gravity[q] = float(gv)
    d.gravity_is_field = 1 if gfields else 0
    pr, taus, vs = _dtau(solver_params, nd)
    d.dtau_P = pr
    for q, t in enumerate(taus):
        d.dtau_tau[q] = t
    for q, v in enumerate(vs):
        d.dtau_V[q] = v
    d.immersed_boundary = 1 if immersed_boundary else 0
    d.precision = precision
    return d, rfield, gfields


class _Vec(SimpleNamespace):
    def __iter__(self):
        return iter(v for k, v in self.__dict__.items())


class IsothermalFullStokesModel:
    """IsothermalFullStokesModel(; arch, grid, boundary_conditions, gravity, rheology,
    solver_params, outer_width)  (FullStokes Isothermal.jl:60-89).

    The model owns device fields P, tau, V, eta, eta_next, r_P, r_V (and the
    gravity / rheology fields when those are field valued).  `outer_width` is
    accepted for signature compatibility; a single device needs no
    inner/outer split.
    """

    _immersed = False

    def __init__(self, *, grid, boundary_conditions, gravity, rheology, solver_params, outer_width=None,
                 arch=None, level_sets=None):
        if not isinstance(grid, UniformGrid):
            raise TypeError("grid must be a UniformGrid")
        self.grid = grid
        self.arch = arch
        self.boundary_conditions = boundary_conditions
        self.gravity = gravity
        self.rheology = rheology
        self.solver_params = solver_params
        self.outer_width = outer_width
        self.desc, rfield, gfields = build_desc(grid, boundary_conditions, gravity, rheology, solver_params,
                                                immersed_boundary=self._immersed)
        lib = A.lib()
        h = C.c_void_p()
        A.check(lib.fi_model_create(C.byref(self.desc), C.byref(h)))
        self._h = h
        self._refine_mode = 1
        F = A.F
        fld = lambda name: DeviceField(self, F[name])
        nd = grid.ndims
        if nd == 3:
            tau = _Vec(xx=fld("TXX"), yy=fld("TYY"), zz=fld("TZZ"), xy=fld("TXY"), xz=fld("TXZ"), yz=fld("TYZ"))
            V = _Vec(x=fld("VX"), y=fld("VY"), z=fld("VZ"))
            rV = _Vec(x=fld("RVX"), y=fld("RVY"), z=fld("RVZ"))
        else:
            tau = _Vec(xx=fld("TXX"), yy=fld("TYY"), xy=fld("TXY"))
            V = _Vec(x=fld("VX"), y=fld("VY"))
            rV = _Vec(x=fld("RVX"), y=fld("RVY"))
        P = fld("P")
        self.stress = SimpleNamespace(P=P, Pr=P, tau=tau, τ=tau)
        self.velocity = SimpleNamespace(V=V)
        eta, eta_next = fld("ETA"), fld("ETA_NEXT")
        self.viscosity = SimpleNamespace(eta=eta, eta_next=eta_next, η=eta, η_next=eta_next)
        rP = fld("RP")
        self.residual = SimpleNamespace(r_P=rP, r_Pr=rP, r_V=rV, norms=None)
        if rfield is not None:
            DeviceField(self, F["RHEO"]).set(rfield)
        for q, k in enumerate(("x", "y", "z")[:nd]):
            if k in gfields:
                DeviceField(self, F["G" + k.upper()]).set(gfields[k])
        if self._immersed:
            self._init_ib(level_sets)

    def _init_ib(self, level_sets):
        pass

    # -- lifetime ---------------------------------------------------------
    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            A.lib().fi_model_destroy(h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- driver API -------------------------------------------------------
    def advance(self, n_iter=1):
        A.check(A.lib().fi_advance_iterations(self._h, int(n_iter)))

    def compute_residuals(self):
        out = np.zeros(4)
        A.check(A.lib().fi_compute_residuals(self._h, A.dptr(out)))
        self.residual.norms = out
        return out

    def solve(self, max_iter, ncheck, tol):
        """Loop advance_iteration! and check every ncheck iterations (tm_stokes_chmy.jl:96-108).

        Returns (iterations, Inf-norms of r_P, r_Vx, r_Vy[, r_Vz]).  Raises
        NonFiniteError when a residual is not finite.
        """
        out = np.zeros(4)
        it = C.c_int(0)
        A.check(A.lib().fi_solve(self._h, int(max_iter), int(ncheck), float(tol), C.byref(it), A.dptr(out)))
        self.residual.norms = out
        return it.value, out

    def refined_solve(self, tol, max_outer=None, max_inner=None, ncheck=None, inner_rtol=3e-2):
        """Solve with FP32 PT sweeps inside an FP64 defect-correction loop.

        Returns (outer steps, FP32 iterations, Inf-norms).  Same discrete
        solution as solve(); the outer residual is evaluated in FP64.
        Lagged PT (refine_mode 1, the default): ncheck = FP32 sweeps per
        outer step (default 32: longer chunks lag the viscosity enough to
        destabilise the Glen-law glaciers; the solver halves it, down to 8,
        when the residual stops halving); max_outer x ncheck caps the FP32
        sweeps in all (default 400 n), max_inner (default 100 n) the full-FP32
        nonlinear start.  Correction solves (refine_mode 0): ncheck = sweeps
        between inner residual checks (default n), max_outer 100, max_inner
        100 n, inner_rtol the inner solves' residual drop.
        """
        n = max(self.grid.dims)
        max_outer, max_inner, ncheck = refined_defaults(n, self._refine_mode, max_outer, max_inner, ncheck)
        counts = (C.c_int * 3)()
        out = np.zeros(4)
        A.check(A.lib().fi_refined_solve(self._h, int(max_outer), int(max_inner), int(ncheck), float(inner_rtol),
                                         float(tol), counts, A.dptr(out)))
        self.residual.norms = out
        self.refined_counts = dict(outer=counts[0], fp32_iterations=counts[1], nonlinear_phase=counts[2])
        return counts[0], counts[1], out

    def apply_bc(self, group):
        g = {"stress": A.FI_BCGROUP_STRESS, "velocity": A.FI_BCGROUP_VELOCITY, "viscosity": A.FI_BCGROUP_ETA,
             "eta": A.FI_BCGROUP_ETA, "residual": A.FI_BCGROUP_RESIDUAL}.get(group, group)
        A.check(A.lib().fi_apply_bc(self._h, int(g)))

    def compute_tau(self):
        A.check(A.lib().fi_compute_tau(self._h))

    def synchronize(self):
        A.check(A.lib().fi_synchronize(self._h))

    def field(self, name):
        return DeviceField(self, A.F[name])

    def set_option(self, key, value):
        """Kernel-path option (fused interior sweep on/off, z planes per workgroup)."""
        k = {"fused": A.FI_OPT_FUSED, "zchunk": A.FI_OPT_ZCHUNK, "refine_full_phase": A.FI_OPT_REFINE_FULL_PHASE,
             "refine_eta_rate_permil": A.FI_OPT_REFINE_ETA_RATE,
             "refine_phase1_digits": A.FI_OPT_REFINE_PHASE1_DIGITS, "fused_variant": A.FI_OPT_FUSED_VARIANT,
             "fused_variant_fp32": A.FI_OPT_FUSED_VARIANT_FP32, "tile_order": A.FI_OPT_TILE_ORDER,
             "refine_anderson": A.FI_OPT_REFINE_ANDERSON, "refine_mode": A.FI_OPT_REFINE_MODE,
             "fused_shell": A.FI_OPT_FUSED_SHELL, "refine_cols": A.FI_OPT_REFINE_COLS}.get(key, key)
        A.check(A.lib().fi_set_option(self._h, int(k), int(value)))
        if k == A.FI_OPT_REFINE_MODE:
            self._refine_mode = int(value)


def refined_defaults(n, mode, max_outer, max_inner, ncheck):
    """Default (max_outer, max_inner, ncheck) of a refined solve of a grid of
    n cells along its longest axis."""
    if mode == 1:
        # max_inner caps the full-FP32 nonlinear start
        ncheck = ncheck or 32
        return max_outer or -(-400 * n // ncheck), max_inner or 100 * n, ncheck
    return max_outer or 100, max_inner or 100 * n, ncheck or n


def advance_iteration(model, t=0.0, dt=1.0):
    """advance_iteration!(model, t, Δt) (Isothermal.jl:91-112); t and Δt are unused there too."""
    model.advance(1)


def compute_residuals(model):
    """compute_residuals!(model) (Isothermal.jl:114-125); returns the Inf-norms of r_P, r_V."""
    return model.compute_residuals()


def norm_inf(model):
    """The Inf-norms of the last compute_residuals (interior of r_P, r_Vx, r_Vy[, r_Vz])."""
    return model.residual.norms


class IsothermalImmersedBoundaryFullStokesModel(IsothermalFullStokesModel):
    """IsothermalImmersedBoundaryFullStokesModel(; ..., level_sets)
    (ImmersedBoundaryFullStokes Isothermal.jl:63-98).

    Adds level sets psi = (na, ns) at vertices and field masks
    omega = (na, ns), each with the eight staggered components
    ccc vvv vcc cvc ccv vvc vcv cvv (2D: cc vv vc cv).
    """

    _immersed = True

    def _init_ib(self, level_sets):
        nd = self.grid.ndims
        names = A.MASK3 if nd == 3 else A.MASK2

        def mask(base):
            ns = SimpleNamespace(**{n: DeviceField(self, base + q) for q, n in enumerate(names)})
            ns.phase = A.FI_PHASE_NA if base == A.FI_F_OM_NA0 else A.FI_PHASE_NS
            return ns

        self.field_masks = SimpleNamespace(na=mask(A.FI_F_OM_NA0), ns=mask(A.FI_F_OM_NS0))
        psi_na, psi_ns = DeviceField(self, A.F["PSI_NA"]), DeviceField(self, A.F["PSI_NS"])
        psi_na.phase, psi_ns.phase = A.FI_PHASE_NA, A.FI_PHASE_NS
        self.level_sets = SimpleNamespace(na=psi_na, ns=psi_ns)
        if level_sets is not None:
            for k in ("na", "ns"):
                v = _get(level_sets, k)
                if v is not None:
                    getattr(self.level_sets, k).set(v)

    def validity(self):
        """Validity bits of every location (fi_ib_validity): a uint8 array of
        the parent shape; bits 0-3 stress locations ccc vvc vcv cvv, bits 4-6
        velocity locations vcc cvc ccv (null_spaces.jl:1-43 as the masked
        kernels apply them)."""
        a = np.zeros(self.grid.parent_shape(), dtype=np.uint8, order="F")
        A.check(A.lib().fi_ib_validity(self._h, a.ctypes.data_as(C.POINTER(C.c_uint8)), a.size))
        return a

"""Problem set-ups of the reference drivers (the callers of the PT path).

pt_params:        PT parameters, stokes_manufactured_solution_free_slip_3D.jl:92-101
                  and test_vavilov.jl:145-156.
manufactured_3d:  confined free-slip flow, stokes_manufactured_solution_free_slip_3D.jl:25-48,70-89.
manufactured_2d:  stokes_manufactured_solution_free_slip_2D.jl:24-32,49-77.
manufactured_errors: relative L-inf/L1/L2 errors of |V| and tau_II against the
                  closed form, ..._3D.jl:140-167 / ..._2D.jl:116-139.
"""
import math

import numpy as np

from .full_stokes import SBC
from .grids import UniformGrid
from .physics import LinearViscousRheology

PI = math.pi


def pt_params(grid, r=0.7, re_mech=8 * PI, vdtau_div=1.1):
    """solver_params = (Δτ = (Pr, τ, V),) from the reference's PT parameter formulas."""
    nd = grid.ndims
    lt_re_m = min(grid.extent_) / re_mech
    vdtau = min(grid.spacing_) / math.sqrt(nd * vdtau_div)
    theta_dtau = lt_re_m * (r + 4 / 3) / vdtau
    dtau_r = 1.0 / (theta_dtau + 1.0)
    nudtau = vdtau * lt_re_m
    names = ("xx", "yy", "zz", "xy", "xz", "yz") if nd == 3 else ("xx", "yy", "xy")
    return {"Δτ": {"Pr": r / theta_dtau, "τ": {k: dtau_r for k in names},
                   "V": {k: nudtau for k in ("x", "y", "z")[:nd]}}}


# closed forms, 3D (..._3D.jl:27-48)
def _f(a, b):
    return np.cos(PI * a) * (b ** 2 - 1) ** 2 - np.cos(PI * b) * (a ** 2 - 1) ** 2


def _g(a, b):
    return np.sin(PI * b) * (a ** 2 - 1) * a - np.sin(PI * a) * (b ** 2 - 1) * b


def _p(a, b):
    return np.cos(PI * b) * (1 - 3 * a ** 2) * 2 - np.cos(PI * a) * (1 - 3 * b ** 2) * 2


EXACT3 = {
    "vx": lambda x, y, z: np.sin(PI * x) * _f(y, z),
    "vy": lambda x, y, z: np.sin(PI * y) * _f(z, x),
    "vz": lambda x, y, z: np.sin(PI * z) * _f(x, y),
    "txx": lambda x, y, z, eta=1.0: 2 * eta * PI * np.cos(PI * x) * _f(y, z),
    "tyy": lambda x, y, z, eta=1.0: 2 * eta * PI * np.cos(PI * y) * _f(z, x),
    "tzz": lambda x, y, z, eta=1.0: 2 * eta * PI * np.cos(PI * z) * _f(x, y),
    "txy": lambda x, y, z, eta=1.0: 4 * eta * np.cos(PI * z) * _g(x, y),
    "txz": lambda x, y, z, eta=1.0: 4 * eta * np.cos(PI * y) * _g(z, x),
    "tyz": lambda x, y, z, eta=1.0: 4 * eta * np.cos(PI * x) * _g(y, z),
    "gx": lambda x, y, z, eta=1.0: -2 * eta * np.sin(PI * x) * (_f(y, z) * PI ** 2 - _p(y, z)),
    "gy": lambda x, y, z, eta=1.0: -2 * eta * np.sin(PI * y) * (_f(z, x) * PI ** 2 - _p(z, x)),
    "gz": lambda x, y, z, eta=1.0: -2 * eta * np.sin(PI * z) * (_f(x, y) * PI ** 2 - _p(x, y)),
}

# closed forms, 2D (..._2D.jl:24-32)
EXACT2 = {
    "vx": lambda x, y: np.sin(PI * x) * np.cos(PI * y),
    "vy": lambda x, y: -np.cos(PI * x) * np.sin(PI * y),
    "txx": lambda x, y, eta=1.0: 2 * eta * PI * np.cos(PI * x) * np.cos(PI * y),
    "tyy": lambda x, y, eta=1.0: -2 * eta * PI * np.cos(PI * x) * np.cos(PI * y),
    "txy": lambda x, y, eta=1.0: np.zeros_like(x),
    "gx": lambda x, y, eta=1.0: -2 * eta * PI ** 2 * np.sin(PI * x) * np.cos(PI * y),
    "gy": lambda x, y, eta=1.0: 2 * eta * PI ** 2 * np.cos(PI * x) * np.sin(PI * y),
}


def manufactured_3d(n, origin=(-1.0, -1.0, -1.0), extent=(2.0, 2.0, 2.0)):
    """Model arguments and driver cadence of the 3D free-slip manufactured solution."""
    dims = (n, n, n) if np.isscalar(n) else tuple(n)
    grid = UniformGrid(origin=origin, extent=extent, dims=dims)
    fs = SBC(0.0, 0.0, 0.0)
    return dict(
        grid=grid,
        boundary_conditions=dict(x=(fs, fs), y=(fs, fs), z=(fs, fs)),
        gravity=dict(x=EXACT3["gx"], y=EXACT3["gy"], z=EXACT3["gz"]),
        rheology=LinearViscousRheology(1.0),
        solver_params=pt_params(grid, r=0.7, re_mech=8 * PI, vdtau_div=1.1),
    ), dict(niter=10 * max(dims), ncheck=max(dims))


def manufactured_2d(n, origin=(-1.0, -1.0), extent=(2.0, 2.0)):
    dims = (n, n) if np.isscalar(n) else tuple(n)
    grid = UniformGrid(origin=origin, extent=extent, dims=dims)
    fs = SBC(0.0, 0.0)
    return dict(
        grid=grid,
        boundary_conditions=dict(x=(fs, fs), y=(fs, fs)),
        gravity=dict(x=EXACT2["gx"], y=EXACT2["gy"]),
        rheology=LinearViscousRheology(1.0),
        solver_params=pt_params(grid, r=0.7, re_mech=8 * PI, vdtau_div=1.1),
    ), dict(niter=10 * max(dims), ncheck=max(dims))


def _avg(a, axis):
    s0 = [slice(None)] * a.ndim
    s1 = [slice(None)] * a.ndim
    s0[axis] = slice(0, -1)
    s1[axis] = slice(1, None)
    return 0.5 * (a[tuple(s0)] + a[tuple(s1)])


def _av4(a, ax0, ax1):
    s = lambda i0, i1: tuple(slice(i0, a.shape[d] - 1 + i0) if d == ax0 else
                             (slice(i1, a.shape[d] - 1 + i1) if d == ax1 else slice(None)) for d in range(a.ndim))
    return 0.25 * (a[s(0, 0)] + a[s(1, 0)] + a[s(1, 1)] + a[s(0, 1)])


def magnitudes(grid, V, tau):
    """|V| (face -> centre averages) and tau_II at cell centres (..._3D.jl:154-158)."""
    nd = grid.ndims
    if nd == 3:
        vm = np.sqrt(_avg(V["x"], 0) ** 2 + _avg(V["y"], 1) ** 2 + _avg(V["z"], 2) ** 2)
        tm = np.sqrt(0.5 * (tau["xx"] ** 2 + tau["yy"] ** 2 + tau["zz"] ** 2) + _av4(tau["xy"], 0, 1) ** 2 +
                     _av4(tau["xz"], 0, 2) ** 2 + _av4(tau["yz"], 1, 2) ** 2)
    else:
        vm = np.sqrt(_avg(V["x"], 0) ** 2 + _avg(V["y"], 1) ** 2)
        tm = np.sqrt(0.5 * (tau["xx"] ** 2 + tau["yy"] ** 2) + _av4(tau["xy"], 0, 1) ** 2)
    return vm, tm


def exact_fields(grid):
    """Closed-form V and tau sampled at the interior staggered locations."""
    nd = grid.ndims
    E = EXACT3 if nd == 3 else EXACT2
    locs3 = {"x": 0b001, "y": 0b010, "z": 0b100, "xx": 0, "yy": 0, "zz": 0, "xy": 0b011, "xz": 0b101, "yz": 0b110}
    locs2 = {"x": 0b01, "y": 0b10, "xx": 0, "yy": 0, "xy": 0b11}
    locs = locs3 if nd == 3 else locs2

    def axis(d, vertex):
        if vertex:
            return grid.origin[d] + (np.arange(1, grid.dims[d] + 2) - 1.0) * grid.spacing_[d]
        return grid.origin[d] + (np.arange(1, grid.dims[d] + 1) - 0.5) * grid.spacing_[d]

    def sample(fn, bits):
        axes = [axis(d, (bits >> d) & 1) for d in range(nd)]
        return fn(*np.meshgrid(*axes, indexing="ij"))

    V = {k: sample(E["v" + k], locs[k]) for k in ("x", "y", "z")[:nd]}
    tau = {k: sample(E["t" + k], locs[k]) for k in (("xx", "yy", "zz", "xy", "xz", "yz") if nd == 3 else ("xx", "yy", "xy"))}
    return V, tau


def manufactured_errors(grid, V, tau):
    """(Linf_v, Linf_tau, L1_v, L1_tau, L2_v, L2_tau), each relative to the closed form."""
    Ve, te = exact_fields(grid)
    vm, tm = magnitudes(grid, Ve, te)
    v, t = magnitudes(grid, V, tau)
    n = lambda a, p: np.linalg.norm(a.ravel(), p)
    return (n(vm - v, np.inf) / n(vm, np.inf), n(tm - t, np.inf) / n(tm, np.inf),
            n(vm - v, 1) / n(vm, 1), n(tm - t, 1) / n(tm, 1),
            n(vm - v, 2) / n(vm, 2), n(tm - t, 2) / n(tm, 2))


# PT parameters (r, re_mech, vdtau_div) of the glacier driver,
# scripts_future_API/test_vavilov.jl:145-156, which runs a linear viscosity
GLACIER_PT = (0.9, 5 * PI, 1.1)
# With Glen's law regularised at a strain rate of 1e-6 (a viscosity contrast
# of about 1e4) those parameters diverge; r = 0.5 and vdtau_div = 4 (a PT
# velocity scale vdtau 1.9 times smaller) converge.  Measured at
# 255x255x127, FP64: r 0.9, vdtau_div 1.1 and relaxation 0.01 or 0.05
# overflow within 4000 iterations; r 0.5, vdtau_div 2 grows 1e7-fold per
# 4000 iterations; r 0.9, re_mech 3 pi, vdtau_div 4 grows 10-fold; r 0.5,
# re_mech 5 pi, vdtau_div 4 with relaxation 0.01 falls about 5-fold per 4000
# iterations and reaches 1e-8 after 44-48 thousand.
GLACIER_PT_REG = (0.5, 5 * PI, 4.0)


def eta_reg_from_rate(eps0, A=1.0, n=3):
    """Viscosity of Glen's law eta = 1/2 A^(-1/n) edot^((1-n)/n) at the
    strain rate eps0: the cap that regularises it at that rate."""
    return 0.5 * A ** (-1.0 / n) * eps0 ** ((1.0 - n) / n)


def glacier_3d_spec(dims, seed=1, glen=True, eta_reg=30.0, relax_rate=0.05, lx=2.0, ly=2.0, H=1.0, eps0=None):
    """Model arguments, driver cadence and geometry of the seeded synthetic
    glacier (see glacier_3d), without allocating anything on the device.
    eps0 (with glen): regularise Glen's law at strain rate eps0 instead of
    eta_reg, with the PT parameters and log relaxation 0.01 that converge at
    that viscosity contrast (GLACIER_PT_REG)."""
    from .full_stokes import TBC
    from .geometries import GaussianGlacier
    from .physics import PowerLawRheology, Relaxation, ViscoplasticRegularisation

    dims = tuple(dims)
    gl = GaussianGlacier(seed=seed, lx=lx, ly=ly, H=H)
    grid = UniformGrid(origin=(-lx / 2, -ly / 2, 0.0), extent=(lx, ly, H), dims=dims)
    dgrid = UniformGrid(origin=(-lx / 2, -ly / 2), extent=(lx, ly), dims=dims[:2])
    tr = TBC(0.0, 0.0, 0.0)
    pt = GLACIER_PT
    if glen and eps0 is not None:
        eta_reg, relax_rate, pt = eta_reg_from_rate(eps0), 0.01, GLACIER_PT_REG
    if glen:
        rheology = Relaxation("eta", ViscoplasticRegularisation(PowerLawRheology(1.0, 3), eta_reg), relax_rate,
                              scale="log")
    else:
        rheology = LinearViscousRheology(1.0)
    kw = dict(grid=grid, boundary_conditions=dict(x=(tr, tr), y=(tr, tr), z=(tr, tr)),
              gravity=dict(x=0.0, y=0.0, z=1.0), rheology=rheology,
              solver_params=pt_params(grid, r=pt[0], re_mech=pt[1], vdtau_div=pt[2]))
    n = max(dims)
    return kw, dict(niter=50 * n, ncheck=5 * n, dem_grid=dgrid, glacier=gl)


def glacier_3d(dims, seed=1, glen=True, eta_reg=30.0, relax_rate=0.05, lx=2.0, ly=2.0, H=1.0, eps0=None):
    """Seeded synthetic glacier on [-lx/2, lx/2] x [-ly/2, ly/2] x [0, H].

    Set-up order of scripts_future_API/test_vavilov.jl:116-177: level sets of
    the ice surface (na) and of the bed (ns, inverted), volume fractions,
    Traction(0) on the box, gravity (0, 0, 1), PT parameters r = 0.9,
    re = 5 pi.  Rheology: Glen n = 3 (PowerLawRheology(1, 3)) capped by
    ViscoplasticRegularisation(eta_reg) and relaxed in log space, or linear
    eta = 1 when glen is False.  Returns (model, cadence).
    """
    from .full_stokes import IsothermalImmersedBoundaryFullStokesModel
    from .level_sets import compute_levelset_from_dem, invert_levelset, omega_from_psi

    kw, cad = glacier_3d_spec(dims, seed, glen, eta_reg, relax_rate, lx, ly, H, eps0)
    gl, dgrid, grid = cad["glacier"], cad["dem_grid"], kw["grid"]
    model = IsothermalImmersedBoundaryFullStokesModel(**kw)
    compute_levelset_from_dem(model.level_sets.na, gl.surface, dgrid, grid)
    compute_levelset_from_dem(model.level_sets.ns, gl.bed, dgrid, grid)
    invert_levelset(model.level_sets.ns)
    omega_from_psi(model.field_masks.ns, model.level_sets.ns)
    omega_from_psi(model.field_masks.na, model.level_sets.na)
    model.viscosity.eta.set(1.0)
    model.viscosity.eta_next.set(1.0)
    model.apply_bc("eta")
    return model, cad


def batch_glaciers(count=96, lo=48, hi=192, seed=2024):
    """The uneven glacier set of the batched workload: count instances whose
    cell counts are drawn uniformly in lo..hi per axis (seeded), cubic cells
    of side 2 / hi (so the domain's aspect ratio is the grid's), glacier
    seeds 100, 101, ...  Returns [(dims, dict(seed, lx, ly, H))]."""
    rng = np.random.default_rng(seed)
    h = 2.0 / hi
    out = []
    for i in range(count):
        d = tuple(int(x) for x in rng.integers(lo, hi + 1, 3))
        out.append((d, dict(seed=100 + i, lx=d[0] * h, ly=d[1] * h, H=d[2] * h)))
    return out


def glacier_2d_spec(dims, seed=1, glen=True, eta_reg=30.0, relax_rate=0.05, lx=2.0, H=1.0, y0=0.0):
    """Model arguments, cadence and the two polygons of a 2D glacier: the
    vertical section y = y0 of the seeded synthetic glacier on
    [-lx/2, lx/2] x [0, H].  na is the region below the ice surface, ns the
    region above the bed (both closed well outside the grid and its halo)."""
    from .full_stokes import TBC
    from .geometries import GaussianGlacier
    from .physics import PowerLawRheology, Relaxation, ViscoplasticRegularisation

    dims = tuple(dims)
    gl = GaussianGlacier(seed=seed, lx=lx, ly=lx, H=H)
    grid = UniformGrid(origin=(-lx / 2, 0.0), extent=(lx, H), dims=dims)
    tr = TBC(0.0, 0.0)
    if glen:
        rheology = Relaxation("eta", ViscoplasticRegularisation(PowerLawRheology(1.0, 3), eta_reg), relax_rate,
                              scale="log")
    else:
        rheology = LinearViscousRheology(1.0)
    kw = dict(grid=grid, boundary_conditions=dict(x=(tr, tr), y=(tr, tr)), gravity=dict(x=0.0, y=1.0),
              rheology=rheology, solver_params=pt_params(grid, r=0.9, re_mech=5 * PI, vdtau_div=1.1))
    pad = 3 * max(lx / dims[0], H / dims[1])
    xs = np.linspace(-lx / 2 - pad, lx / 2 + pad, 2 * dims[0] + 1)
    ys = np.full_like(xs, y0)
    surf = np.stack([xs, gl.surface(xs, ys)], axis=1)
    bed = np.stack([xs, gl.bed(xs, ys)], axis=1)
    na = np.concatenate([[[-lx / 2 - pad, -pad]], surf, [[lx / 2 + pad, -pad]]])
    ns = np.concatenate([[[-lx / 2 - pad, H + pad]], bed, [[lx / 2 + pad, H + pad]]])
    n = max(dims)
    return kw, dict(niter=50 * n, ncheck=5 * n, glacier=gl, poly_na=na, poly_ns=ns)


def glacier_2d(dims, seed=1, glen=True, **kw_spec):
    """2D immersed-boundary glacier section (glacier_2d_spec) with its level
    sets and volume fractions set up.  Returns (model, cadence)."""
    from .full_stokes import IsothermalImmersedBoundaryFullStokesModel
    from .level_sets import compute_levelset_from_polygon, omega_from_psi

    kw, cad = glacier_2d_spec(dims, seed, glen, **kw_spec)
    model = IsothermalImmersedBoundaryFullStokesModel(**kw)
    compute_levelset_from_polygon(model.level_sets.na, cad["poly_na"])
    compute_levelset_from_polygon(model.level_sets.ns, cad["poly_ns"])
    omega_from_psi(model.field_masks.ns, model.level_sets.ns)
    omega_from_psi(model.field_masks.na, model.level_sets.na)
    model.viscosity.eta.set(1.0)
    model.viscosity.eta_next.set(1.0)
    model.apply_bc("eta")
    return model, cad

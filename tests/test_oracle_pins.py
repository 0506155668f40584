"""Pins of the CPU oracle against what the reference's own tests and the
mathematics fix (no GPU).  Each test names what a plausible mistake in the
oracle (a dropped term, a wrong sign or index, a transposed operand) would
break.
"""
import math

import numpy as np
import pytest


# ---------------------------------------------------------------------------
# grid geometry: known answers of test/test_chmy_grids.jl:13-160
# ---------------------------------------------------------------------------
def test_grid_known_answers(pkg):
    C, V = pkg.Center(), pkg.Vertex()
    assert pkg.flip(C) is V and pkg.flip(V) is C
    nx, ny = 5, 20
    g = pkg.UniformGrid(origin=(-1, -2), extent=(2, 4), dims=(nx, ny))
    assert g.size(C) == (nx, ny) and g.size(V) == (nx + 1, ny + 1)
    assert g.size((C, V)) == (nx, ny + 1) and g.size((V, C)) == (nx + 1, ny)
    assert np.allclose(g.bounds(V, 0), (-1.0, 1.0)) and np.allclose(g.bounds(V, 1), (-2.0, 2.0))
    assert np.allclose(g.bounds(C, 0), (-0.8, 0.8)) and np.allclose(g.bounds(C, 1), (-1.9, 1.9))
    assert np.allclose(g.extent(V), (2.0, 4.0)) and np.allclose(g.extent(C), (1.6, 3.8))
    assert np.allclose(g.extent((C, V)), (1.6, 4.0)) and np.allclose(g.extent((V, C)), (2.0, 3.8))
    assert np.allclose(g.origin_of(C), (-0.8, -1.9)) and np.allclose(g.origin_of((V, C)), (-1.0, -1.9))
    assert np.allclose(g.spacing(), (0.4, 0.2)) and np.allclose(g.inv_spacing(), (2.5, 5.0))
    assert math.isclose(g.coord(V, 0, nx + 1), 1.0) and math.isclose(g.coord(V, 1, ny + 1), 2.0)
    assert math.isclose(g.coord(C, 0, nx), 0.8) and math.isclose(g.coord(C, 1, ny), 1.9)
    assert np.allclose((g.coord((V, C), 0, 1), g.coord((V, C), 1, ny)), (-1.0, 1.9))


def test_oracle_grid_spacing(pkg, orc):
    kw, _ = pkg.manufactured_2d((5, 20), origin=(-1, -2), extent=(2, 4))
    ref = orc.OracleModel(pkg.IsothermalFullStokesModel.__mro__[0] and _desc(pkg, kw))
    assert np.allclose(ref.m.dx[:2], (0.4, 0.2)) and np.allclose(ref.m.idx[:2], (2.5, 5.0))
    assert ref.m.sy == 5 + 3


def _desc(pkg, kw, ib=False):
    d, _, _ = pkg.build_desc(kw["grid"], kw["boundary_conditions"], kw["gravity"], kw["rheology"],
                             kw["solver_params"], immersed_boundary=ib)
    return d


def _oracle(pkg, orc, kw, ib=False):
    ref = orc.OracleModel(_desc(pkg, kw, ib))
    grid = kw["grid"]
    bits = {"GX": 0b001, "GY": 0b010, "GZ": 0b100}
    for q, k in enumerate(("x", "y", "z")[:grid.ndims]):
        gv = kw["gravity"][k]
        if callable(gv):
            ref.set("G" + k.upper(), gv, bits=bits["G" + k.upper()], grid=grid)
    ref.set("ETA", 1.0)
    ref.set("ETA_NEXT", 1.0)
    return ref


# ---------------------------------------------------------------------------
# staggered operators are exact on linear fields
# ---------------------------------------------------------------------------
def test_divergence_and_gradient_exact_on_linear_fields(pkg, orc):
    kw, _ = pkg.manufactured_3d((6, 5, 7), origin=(0.1, -0.2, 0.3), extent=(1.0, 0.7, 1.3))
    kw["gravity"] = dict(x=0.0, y=0.0, z=0.0)
    grid = kw["grid"]
    ref = _oracle(pkg, orc, kw)
    a = np.array([[0.3, -1.2, 0.7], [2.1, 0.4, -0.5], [-0.9, 1.5, 0.25]])
    for r, (name, bits) in enumerate((("VX", 0b001), ("VY", 0b010), ("VZ", 0b100))):
        ref.set(name, lambda x, y, z, r=r: a[r, 0] * x + a[r, 1] * y + a[r, 2] * z, bits=bits, grid=grid)
    p = (0.6, -0.8, 1.1)
    ref.set("P", lambda x, y, z: p[0] * x + p[1] * y + p[2] * z, bits=0, grid=grid)
    ref.compute_residuals()
    nx, ny, nz = grid.dims
    rP = ref.get("RP")[1:nx + 1, 1:ny + 1, 1:nz + 1]
    assert np.allclose(rP, a[0, 0] + a[1, 1] + a[2, 2], rtol=0, atol=1e-12)
    # r_V = -grad P with tau = 0, g = 0 (interior nodes away from residual BC faces)
    assert np.allclose(ref.get("RVX")[2:nx + 1, 1:ny + 1, 1:nz + 1], -p[0], atol=1e-12)
    assert np.allclose(ref.get("RVY")[1:nx + 1, 2:ny + 1, 1:nz + 1], -p[1], atol=1e-12)
    assert np.allclose(ref.get("RVZ")[1:nx + 1, 1:ny + 1, 2:nz + 1], -p[2], atol=1e-12)


def test_compute_tau_of_linear_velocity(pkg, orc):
    """tau = 2 eta (edot - div V/3 I) with edot_ij = (d_i V_j + d_j V_i)/2 for a linear V."""
    kw, _ = pkg.manufactured_3d((5, 6, 4))
    grid = kw["grid"]
    ref = _oracle(pkg, orc, kw)
    eta = 1.7
    ref.set("ETA", eta)
    G = np.array([[0.3, -1.2, 0.7], [2.1, 0.4, -0.5], [-0.9, 1.5, 0.25]])  # G[r, c] = dV_r/dx_c
    for r, (name, bits) in enumerate((("VX", 0b001), ("VY", 0b010), ("VZ", 0b100))):
        ref.set(name, lambda x, y, z, r=r: G[r, 0] * x + G[r, 1] * y + G[r, 2] * z, bits=bits, grid=grid)
    ref.compute_tau()
    div = np.trace(G)
    e = 0.5 * (G + G.T)
    sl = (slice(2, 5), slice(2, 5), slice(2, 4))
    for name, (r, c) in (("TXX", (0, 0)), ("TYY", (1, 1)), ("TZZ", (2, 2))):
        assert np.allclose(ref.get(name)[sl], 2 * eta * (e[r, c] - div / 3), atol=1e-12), name
    for name, (r, c) in (("TXY", (0, 1)), ("TXZ", (0, 2)), ("TYZ", (1, 2))):
        assert np.allclose(ref.get(name)[sl], 2 * eta * e[r, c], atol=1e-12), name


# ---------------------------------------------------------------------------
# discrete fixed point: an exact discrete solution is left unchanged
# ---------------------------------------------------------------------------
def _bump(t):
    """Smooth bump vanishing (with its derivatives) outside (0.3, 0.7)."""
    s = np.clip((t - 0.3) / 0.4, 0.0, 1.0)
    return (s * (1 - s)) ** 4 * 256.0


def test_fixed_point_invariance_3d(pkg, orc):
    n = 14
    kw, _ = pkg.manufactured_3d(n, origin=(0.0, 0.0, 0.0), extent=(1.0, 1.0, 1.0))
    kw["gravity"] = dict(x=0.0, y=0.0, z=0.0)
    grid = kw["grid"]
    ref = _oracle(pkg, orc, kw)
    # velocity = discrete curl of a vertex/edge potential: (A on edges) -> V on faces
    # use V = curl(0, 0, psi) with psi on (V,V,C) edges: Vx = d_y psi, Vy = -d_x psi
    X, Y, Z = grid.parent_coords(0b011)
    psi = _bump(X) * _bump(Y) * _bump(Z) * 0.01
    sy = psi.shape
    Vx = np.zeros(sy)
    Vy = np.zeros(sy)
    Vx[:, :-1, :] = (psi[:, 1:, :] - psi[:, :-1, :]) * (1 / grid.spacing_[1])
    Vy[:-1, :, :] = -(psi[1:, :, :] - psi[:-1, :, :]) * (1 / grid.spacing_[0])
    ref.set("VX", Vx)
    ref.set("VY", Vy)
    ref.set("P", lambda x, y, z: _bump(x) * _bump(y) * _bump(z) * 0.3, bits=0, grid=grid)
    ref.compute_tau()
    r0 = ref.compute_residuals()
    assert r0[0] < 1e-14                               # discrete div V = 0
    # gravity = -grad P + div tau of this state, so r_V vanishes
    kw2 = dict(kw)
    kw2["gravity"] = dict(x=lambda x, y, z: 0 * x, y=lambda x, y, z: 0 * x, z=lambda x, y, z: 0 * x)
    ref2 = _oracle(pkg, orc, kw2)
    for name in ("P", "TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ", "VX", "VY", "VZ"):
        ref2.set(name, ref.get(name))
    ref2.set("GX", ref.get("RVX"))
    ref2.set("GY", ref.get("RVY"))
    ref2.set("GZ", ref.get("RVZ"))
    assert max(ref2.compute_residuals()) < 1e-13
    groups = {"P": ("P",), "tau": ("TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ"), "V": ("VX", "VY", "VZ")}
    before = {k: ref2.get(k).copy() for g in groups.values() for k in g}
    scale = {g: max(np.max(np.abs(before[k])) for k in ks) for g, ks in groups.items()}
    assert min(scale.values()) > 0
    ref2.advance(3)
    for g, ks in groups.items():
        for k in ks:
            assert np.max(np.abs(ref2.get(k) - before[k])) / scale[g] < 1e-12, k


def test_linearity_in_forcing(pkg, orc):
    kw, _ = pkg.manufactured_3d(10)
    r1 = _oracle(pkg, orc, kw)
    g2 = {k: (lambda f: (lambda x, y, z: 2.0 * f(x, y, z)))(v) for k, v in kw["gravity"].items()}
    kw2 = dict(kw, gravity=g2)
    r2 = _oracle(pkg, orc, kw2)
    r1.advance(40)
    r2.advance(40)
    for k in ("P", "TXX", "TXY", "VX", "VZ"):
        a, b = r1.get(k), r2.get(k)
        assert np.max(np.abs(b - 2 * a)) <= 1e-12 * max(np.max(np.abs(b)), 1e-30), k


# ---------------------------------------------------------------------------
# boundary-condition invariants after iterations
# ---------------------------------------------------------------------------
def test_boundary_invariants(pkg, orc):
    grid = pkg.UniformGrid(origin=(0, 0, 0), extent=(1, 1, 1), dims=(8, 7, 6))
    tv = (0.2, -0.1, 0.3)
    bcs = dict(x=(pkg.SBC(0.0, 0.0, 0.0), pkg.SBC(0.0, 0.0, 0.0)),
               y=(pkg.VBC(0.1, 0.0, -0.2), pkg.VBC(0.0, 0.0, 0.0)),
               z=(pkg.TBC(*tv), pkg.TBC(0.0, 0.0, 0.0)))
    kw = dict(grid=grid, boundary_conditions=bcs, gravity=dict(x=0.0, y=0.0, z=1.0),
              rheology=pkg.LinearViscousRheology(1.0), solver_params=pkg.pt_params(grid))
    ref = _oracle(pkg, orc, kw)
    ref.advance(25)
    nx, ny, nz = grid.dims
    Vx, Vy, Vz = ref.get("VX"), ref.get("VY"), ref.get("VZ")
    # free slip on x: normal velocity and tangential stress vanish on the faces
    assert np.all(Vx[1, 1:ny + 1, 1:nz + 1] == 0) and np.all(Vx[nx + 1, 1:ny + 1, 1:nz + 1] == 0)
    # (edges shared with the z faces take the later z-face traction value: x, y, z order)
    assert np.all(ref.get("TXY")[1, 1:ny + 1, 1:nz + 1] == 0) and np.all(ref.get("TXZ")[nx + 1, 1:ny + 1, 2:nz + 1] == 0)
    assert np.all(ref.get("TXZ")[nx + 1, 1:ny + 1, 1] == tv[0])
    # no slip on y low with velocity (0.1, 0, -0.2): normal node and tangential face averages
    assert np.all(Vy[2:nx, 1, 2:nz] == 0.0)
    assert np.allclose(0.5 * (Vx[2:nx, 0, 2:nz] + Vx[2:nx, 1, 2:nz]), 0.1, atol=1e-14)
    assert np.allclose(0.5 * (Vz[2:nx, 0, 2:nz] + Vz[2:nx, 1, 2:nz]), -0.2, atol=1e-14)
    # traction on z low: pressure and tau_zz ghost averages equal the normal component,
    # tangential stresses on the face equal the tangential components
    P, tzz = ref.get("P"), ref.get("TZZ")
    assert np.allclose(0.5 * (P[2:nx, 2:ny, 0] + P[2:nx, 2:ny, 1]), tv[2], atol=1e-14)
    assert np.allclose(0.5 * (tzz[2:nx, 2:ny, 0] + tzz[2:nx, 2:ny, 1]), tv[2], atol=1e-14)
    assert np.all(ref.get("TXZ")[2:nx, 1:ny + 1, 1] == tv[0]) and np.all(ref.get("TYZ")[1:nx + 1, 2:ny, 1] == tv[1])
    # eta_next ghosts are Neumann copies
    e = ref.get("ETA")
    assert np.array_equal(e[0, 1:ny + 1, 1:nz + 1], e[1, 1:ny + 1, 1:nz + 1])


# ---------------------------------------------------------------------------
# manufactured solutions: O(h^2) convergence of |V| and tau_II
# ---------------------------------------------------------------------------
def _errors_2d(pkg, orc, n, box=((-1.0, -1.0), (2.0, 2.0)), tol=1e-10):
    kw, cad = pkg.manufactured_2d(n, origin=box[0], extent=box[1])
    ref = _oracle(pkg, orc, kw)
    it, norms = ref.solve(400 * n, cad["ncheck"], tol)
    assert max(norms) < tol
    g = kw["grid"]
    V = {"x": ref.get("VX")[1:n + 2, 1:n + 1], "y": ref.get("VY")[1:n + 1, 1:n + 2]}
    tau = {"xx": ref.get("TXX")[1:n + 1, 1:n + 1], "yy": ref.get("TYY")[1:n + 1, 1:n + 1],
           "xy": ref.get("TXY")[1:n + 2, 1:n + 2]}
    from fastice_jl_amd.problems import manufactured_errors
    return it, manufactured_errors(g, V, tau)


def test_manufactured_2d_second_order(pkg, orc):
    errs = [_errors_2d(pkg, orc, n)[1] for n in (16, 32, 64)]
    for q in range(6):
        r1, r2 = errs[0][q] / errs[1][q], errs[1][q] / errs[2][q]
        assert 3.3 < r1 < 4.7 and 3.5 < r2 < 4.5, (q, r1, r2)
    assert errs[2][0] < 1e-3


def test_manufactured_2d_unit_box_127(pkg, orc):
    """2D isoviscous free-slip manufactured solution, 127x127 cells on the unit box, FP64, to 1e-8."""
    it, e = _errors_2d(pkg, orc, 127, box=((0.0, 0.0), (1.0, 1.0)), tol=1e-8)
    assert it % 127 == 0 and it <= 40 * 127
    assert e[0] < 2e-4 and e[1] < 2e-4


def test_manufactured_3d_second_order(pkg, orc):
    errs = []
    for n in (8, 16):
        kw, cad = pkg.manufactured_3d(n)
        ref = _oracle(pkg, orc, kw)
        it, norms = ref.solve(300 * n, cad["ncheck"], 1e-10)
        assert max(norms) < 1e-10
        V = {"x": ref.get("VX")[1:n + 2, 1:n + 1, 1:n + 1], "y": ref.get("VY")[1:n + 1, 1:n + 2, 1:n + 1],
             "z": ref.get("VZ")[1:n + 1, 1:n + 1, 1:n + 2]}
        tau = {k: ref.get("T" + k.upper())[1:n + 1, 1:n + 1, 1:n + 1] for k in ("xx", "yy", "zz")}
        tau["xy"] = ref.get("TXY")[1:n + 2, 1:n + 2, 1:n + 1]
        tau["xz"] = ref.get("TXZ")[1:n + 2, 1:n + 1, 1:n + 2]
        tau["yz"] = ref.get("TYZ")[1:n + 1, 1:n + 2, 1:n + 2]
        from fastice_jl_amd.problems import manufactured_errors
        errs.append(manufactured_errors(kw["grid"], V, tau))
    for q in (2, 3, 4, 5):  # L1 and L2 norms
        r = errs[0][q] / errs[1][q]
        assert 3.0 < r < 5.0, (q, r)


# ---------------------------------------------------------------------------
# rheology closed forms (src/Physics.jl:11-47)
# ---------------------------------------------------------------------------
def _rheo_oracle(pkg, orc, rheology, eta_old=2.0):
    kw, _ = pkg.manufactured_3d(4)
    kw["rheology"] = rheology
    ref = _oracle(pkg, orc, kw)
    ref.set("ETA", eta_old)
    return ref


def test_rheology_closed_forms(pkg, orc):
    t = 0.37
    r = _rheo_oracle(pkg, orc, pkg.PowerLawRheology(0.8, 3))
    assert r.rheology_at(100, t) == 0.5 / (0.8 * (t * t))
    r = _rheo_oracle(pkg, orc, pkg.PowerLawRheology(0.8, 4.5))
    assert math.isclose(r.rheology_at(100, t), 0.5 / (0.8 * t ** 3.5), rel_tol=1e-15)
    r = _rheo_oracle(pkg, orc, pkg.ViscoplasticRegularisation(pkg.PowerLawRheology(1.0, 3), 10.0))
    assert r.rheology_at(100, 0.0) == 10.0                         # tau_II = 0: capped at eta_reg
    assert math.isclose(r.rheology_at(100, t), 1 / (1 / (0.5 / t ** 2) + 0.1), rel_tol=1e-15)
    r = _rheo_oracle(pkg, orc, pkg.ViscoplasticRegularisation(pkg.LinearViscousRheology(3.0), 1e300))
    assert math.isclose(r.rheology_at(100, t), 3.0, rel_tol=1e-15)  # eta_reg -> inf limit
    r = _rheo_oracle(pkg, orc, pkg.Relaxation("eta", pkg.LinearViscousRheology(5.0), 0.25), eta_old=2.0)
    assert r.rheology_at(100, t) == 2.0 * 0.75 + 5.0 * 0.25
    r = _rheo_oracle(pkg, orc, pkg.Relaxation("eta", pkg.LinearViscousRheology(8.0), 0.5, scale="log"), eta_old=2.0)
    assert math.isclose(r.rheology_at(100, t), 4.0, rel_tol=1e-14)   # geometric mean
    r = _rheo_oracle(pkg, orc, pkg.Relaxation("eta", pkg.LinearViscousRheology(1000.0), 1 / 3, scale="log10"),
                     eta_old=1.0)
    assert math.isclose(r.rheology_at(100, t), 10.0, rel_tol=1e-13)


# ---------------------------------------------------------------------------
# volume fractions (src/LevelSets/volume_fractions.jl)
# ---------------------------------------------------------------------------
def _box_corners(d):
    return np.array([[cx * d[0], cy * d[1], cz * d[2]] for cz in (0, 1) for cy in (0, 1) for cx in (0, 1)])


def test_volume_fraction_of_planes_is_exact(orc):
    rng = np.random.default_rng(1)
    d = np.array([0.3, 0.5, 0.2])
    P = _box_corners(d)
    for _ in range(200):
        z0 = rng.uniform(-0.05, 0.25)
        psi = P[:, 2] - z0                               # below the plane z = z0 is inside
        expect = min(max(z0 / d[2], 0.0), 1.0)
        assert abs(orc.psi2omega3(psi, d) - expect) < 1e-13
    for _ in range(200):
        nrm = rng.standard_normal(3)
        c = rng.uniform(-0.3, 0.3)
        psi = P @ nrm - c
        # inside fraction of a box cut by a plane: Monte-Carlo-free check via complement
        w = orc.psi2omega3(psi, d)
        wc = orc.psi2omega3(-psi, d)
        assert 0.0 <= w <= 1.0 and abs(w + wc - 1.0) < 1e-12
    assert orc.psi2omega3(-np.ones(8), d) == pytest.approx(1.0, abs=1e-15)
    assert orc.psi2omega3(np.ones(8), d) == 0.0


def test_volume_fraction_2d_planes_and_complement(orc):
    rng = np.random.default_rng(2)
    d = np.array([0.4, 0.25])
    P = np.array([[0, 0], [d[0], 0], [0, d[1]], [d[0], d[1]]])
    for _ in range(200):
        y0 = rng.uniform(-0.05, 0.3)
        assert abs(orc.psi2omega2(P[:, 1] - y0, d) - min(max(y0 / d[1], 0), 1)) < 1e-13
        nrm = rng.standard_normal(2)
        psi = P @ nrm - rng.uniform(-0.2, 0.2)
        assert abs(orc.psi2omega2(psi, d) + orc.psi2omega2(-psi, d) - 1.0) < 1e-12


def test_sphere_volume_converges_second_order(pkg, orc):
    errs = []
    for n in (12, 24):
        grid = pkg.UniformGrid(origin=(-1, -1, -1), extent=(2, 2, 2), dims=(n, n, n))
        kw, _ = pkg.manufactured_3d(n)
        ref = _oracle(pkg, orc, kw, ib=True)
        ref.set("PSI_NA", lambda x, y, z: np.sqrt(x * x + y * y + z * z) - 0.7, bits=0b111, grid=grid)
        ref.omega_from_psi(0)
        w = ref.get("OM_NA0")[1:n + 1, 1:n + 1, 1:n + 1]
        vol = w.sum() * (2.0 / n) ** 3
        errs.append(abs(vol - 4 / 3 * math.pi * 0.7 ** 3))
    assert errs[1] < errs[0] / 3.0


# ---------------------------------------------------------------------------
# signed distance to a DEM (src/LevelSets/signed_distances.jl)
# ---------------------------------------------------------------------------
def _ls_oracle(pkg, orc, dims=(10, 9, 8)):
    grid = pkg.UniformGrid(origin=(-1, -1, 0), extent=(2, 2, 1), dims=dims)
    dgrid = pkg.UniformGrid(origin=(-1, -1), extent=(2, 2), dims=dims[:2])
    tr = pkg.TBC(0.0, 0.0, 0.0)
    kw = dict(grid=grid, boundary_conditions=dict(x=(tr, tr), y=(tr, tr), z=(tr, tr)),
              gravity=dict(x=0.0, y=0.0, z=1.0), rheology=pkg.LinearViscousRheology(1.0),
              solver_params=pkg.pt_params(grid))
    return grid, dgrid, _oracle(pkg, orc, kw, ib=True)


def test_levelset_flat_and_inclined_dem(pkg, orc):
    from fastice_jl_amd.level_sets import dem_parent
    grid, dgrid, ref = _ls_oracle(pkg, orc)
    ref.levelset_from_dem(0, dem_parent(lambda x, y: 0.4 + 0 * x, dgrid), dgrid.dims, dgrid.origin, dgrid.extent_)
    X, Y, Z = grid.parent_coords(0b111)
    psi = ref.get("PSI_NA")
    cut = 4 * max(grid.spacing_)
    inner = (slice(1, 12), slice(1, 11), slice(1, 10))
    near = np.abs(Z - 0.4) < cut
    assert np.allclose(psi[inner][near[inner]], (Z - 0.4)[inner][near[inner]], atol=1e-14)
    # inclined plane z = 0.3 + 0.2 x: exact plane distance where the closest point is inside the DEM
    ref.levelset_from_dem(1, dem_parent(lambda x, y: 0.3 + 0.2 * x, dgrid), dgrid.dims, dgrid.origin, dgrid.extent_)
    psi = ref.get("PSI_NS")
    dist = (Z - 0.3 - 0.2 * X) / math.sqrt(1 + 0.04)
    ok = (np.abs(dist) < 0.5 * cut) & (np.abs(X) < 0.7) & (np.abs(Y) < 0.7)
    assert ok[inner].sum() > 20
    assert np.allclose(psi[inner][ok[inner]], dist[inner][ok[inner]], atol=1e-12)
    ref.invert_levelset(1)
    assert np.allclose(ref.get("PSI_NS")[inner][ok[inner]], -dist[inner][ok[inner]], atol=1e-12)


def test_levelset_rotated_flat_dem(pkg, orc):
    """init_levelset! takes the distance of R (x, y, z) (compute_levelsets.jl:1-8):
    a flat DEM z = 0.4 seen through a rotation by 0.3 about x gives
    psi = (R X)_z - 0.4 where the foot point lies inside the DEM."""
    from fastice_jl_amd.level_sets import dem_parent
    grid, dgrid, ref = _ls_oracle(pkg, orc)
    c, s = math.cos(0.3), math.sin(0.3)
    R = np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])
    ref.levelset_from_dem(0, dem_parent(lambda x, y: 0.4 + 0 * x, dgrid), dgrid.dims, dgrid.origin, dgrid.extent_,
                          R=R)
    X, Y, Z = grid.parent_coords(0b111)
    px, py, pz = X, c * Y - s * Z, s * Y + c * Z
    want = pz - 0.4
    cut = 4 * max(grid.spacing_)
    inner = (slice(1, 12), slice(1, 11), slice(1, 10))
    ok = (np.abs(want) < 0.5 * cut) & (np.abs(px) < 0.7) & (np.abs(py) < 0.7)
    assert ok[inner].sum() > 20
    psi = ref.get("PSI_NA")
    assert np.allclose(psi[inner][ok[inner]], want[inner][ok[inner]], atol=1e-13)
    # a transposed rotation (the plausible mistake) is far off
    assert np.max(np.abs(psi[inner][ok[inner]] - (c * Z - s * Y - 0.4)[inner][ok[inner]])) > 1e-2


def _closest_point_triangle(p, a, b, c):
    """Closest point on a triangle (region classification; C. Ericson, Real-Time Collision Detection, 5.1.5)."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    den = 1.0 / (va + vb + vc)
    return a + ab * (vb * den) + ac * (vc * den)


def test_levelset_matches_brute_force_distance(pkg, orc):
    from fastice_jl_amd.level_sets import dem_parent
    grid, dgrid, ref = _ls_oracle(pkg, orc, dims=(7, 6, 6))
    rng = np.random.default_rng(3)
    nx, ny = dgrid.dims
    dem = 0.5 + 0.08 * rng.standard_normal((nx + 1, ny + 1))
    parent = dem_parent(dem, dgrid)
    ref.levelset_from_dem(0, parent, dgrid.dims, dgrid.origin, dgrid.extent_)
    psi = ref.get("PSI_NA")
    # the reference's triangulated DEM: two triangles per vertex Iv in [1, nv] of the
    # parent (haloed) DEM, the last row/column reaching into the halo
    # (signed_distances.jl:58-66,84-90)
    xs = dgrid.origin[0] + (np.arange(nx + 3) - 1.0) * dgrid.spacing_[0]
    ys = dgrid.origin[1] + (np.arange(ny + 3) - 1.0) * dgrid.spacing_[1]
    tris = []
    for i in range(1, nx + 2):
        for j in range(1, ny + 2):
            v = lambda a, b: np.array([xs[a], ys[b], parent[a, b]])
            tris.append((v(i, j), v(i + 1, j), v(i, j + 1)))
            tris.append((v(i, j + 1), v(i + 1, j), v(i + 1, j + 1)))
    cut = 4 * max(grid.spacing_)
    checked = 0
    for k in range(1, grid.dims[2] + 2):
        for j in range(1, grid.dims[1] + 2):
            for i in range(1, grid.dims[0] + 2):
                p = np.array([grid.coord(pkg.Vertex(), 0, i), grid.coord(pkg.Vertex(), 1, j),
                              grid.coord(pkg.Vertex(), 2, k)])
                d = min(np.linalg.norm(p - _closest_point_triangle(p, *t)) for t in tris)
                if d < 0.5 * cut:
                    assert abs(abs(psi[i, j, k]) - d) < 1e-12, (i, j, k)
                    checked += 1
    assert checked > 30


def test_residual_norm_propagates_nan(pkg, orc):
    kw, _ = pkg.manufactured_3d(6)
    ref = _oracle(pkg, orc, kw)
    P = ref.get("P")
    P[3, 3, 3] = np.nan
    norms = ref.compute_residuals()
    assert np.isnan(norms[1])


def _seg_dist(p, a, b):
    """Point-segment distance from the endpoint distances and the
    perpendicular distance |e x w| / |e| when the foot lies on the segment."""
    e, w = b - a, p - a
    le = math.hypot(e[0], e[1])
    along = (e[0] * w[0] + e[1] * w[1]) / le
    if along <= 0:
        return math.hypot(w[0], w[1])
    if along >= le:
        return math.hypot(p[0] - b[0], p[1] - b[1])
    return abs(e[0] * w[1] - e[1] * w[0]) / le


def _winding(p, poly):
    """Total signed angle swept around p by the polygon's edges (2 pi inside, 0 outside)."""
    ang = 0.0
    for i in range(len(poly)):
        a, b = poly[i] - p, poly[(i + 1) % len(poly)] - p
        ang += math.atan2(a[0] * b[1] - a[1] * b[0], a[0] * b[0] + a[1] * b[1])
    return ang


def test_sd_poly_closed_forms_and_brute_force(orc):
    sq = np.array([[0.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.0, 1.0]])
    # inside: minus the distance to the nearest side; outside: to the side or the corner
    assert abs(orc.sd_poly((0.3, 0.4), sq) - (-0.3)) < 1e-15
    assert abs(orc.sd_poly((0.5, 0.9), sq) - (-0.1)) < 1e-15
    assert abs(orc.sd_poly((1.5, 0.5), sq) - 0.5) < 1e-15
    assert abs(orc.sd_poly((1.3, 1.4), sq) - 0.5) < 1e-15
    assert abs(orc.sd_poly((-0.6, -0.8), sq) - 1.0) < 1e-15
    # random star-shaped polygons: |psi| = brute-force segment distance, sign from the winding number
    rng = np.random.default_rng(21)
    for trial in range(6):
        n = int(rng.integers(3, 12))
        ang = np.sort(rng.uniform(0, 2 * math.pi, n))
        rad = rng.uniform(0.3, 1.0, n)
        poly = np.stack([0.1 + rad * np.cos(ang), -0.2 + rad * np.sin(ang)], axis=1)
        for p in rng.uniform(-1.3, 1.3, (40, 2)):
            d = min(_seg_dist(p, poly[i], poly[(i + 1) % n]) for i in range(n))
            inside = abs(_winding(p, poly)) > math.pi
            got = orc.sd_poly(p, poly)
            assert abs(abs(got) - d) < 1e-12, (trial, p)
            if d > 1e-9:
                assert (got < 0) == inside, (trial, p)


def test_levelset_from_polygon_half_plane(pkg, orc):
    """A polygon whose only edge near the grid is the line y = h: psi = y - h
    exactly, with Neumann halos."""
    grid = pkg.UniformGrid(origin=(-1, 0), extent=(2, 1), dims=(12, 9))
    tr = pkg.TBC(0.0, 0.0)
    kw = dict(grid=grid, boundary_conditions=dict(x=(tr, tr), y=(tr, tr)), gravity=dict(x=0.0, y=1.0),
              rheology=pkg.LinearViscousRheology(1.0), solver_params=pkg.pt_params(grid))
    ref = _oracle(pkg, orc, kw, ib=True)
    h = 0.37
    poly = np.array([[-50.0, -50.0], [50.0, -50.0], [50.0, h], [-50.0, h]])
    ref.levelset_from_polygon(0, poly)
    X, Y = grid.parent_coords(0b11)
    psi = ref.get("PSI_NA")
    inner = (slice(1, 14), slice(1, 11))
    assert np.allclose(psi[inner], (Y - h)[inner], atol=1e-14)
    # Neumann vertex ghosts mirror the node next to the boundary node
    assert np.array_equal(psi[0, 1:11], psi[2, 1:11]) and np.array_equal(psi[14, 1:11], psi[12, 1:11])
    assert np.array_equal(psi[1:14, 0], psi[1:14, 2]) and np.array_equal(psi[1:14, 11], psi[1:14, 9])


# --- PT heat diffusion (src/Models/Diffusion/Heat/kernels.jl:3-29) ---------

def _heat(orc, n, extent, kinds, vals, lam=1.0, dtq=1.0, dtT=None):
    nd = len(n)
    dx = [e / m for e, m in zip(extent, n)]
    if dtT is None:
        dtT = min(dx) ** 2 / (lam * 2.1 * nd)
    kinds = list(kinds) + [[0, 0]] * (3 - nd)
    vals = list(vals) + [[0.0, 0.0]] * (3 - nd)
    return orc.OracleHeat(nd, list(n) + [1] * (3 - nd), list(extent) + [1.0] * (3 - nd), kinds, vals, lam, dtq, dtT)


def test_heat_steady_linear_profile(orc):
    """Value conditions T = 1 (west) and 3 (east), no flux elsewhere, dt -> inf:
    the discrete steady state is T = 1 + 2 x / L at the cell centres."""
    n, ext = (16, 6, 5), (1.0, 0.5, 0.4)
    h = _heat(orc, n, ext, [[1, 1], [2, 2], [2, 2]], [[1.0, 3.0], [0.0, 0.0], [0.0, 0.0]])
    h.apply_bc()
    h.advance(1e30, 8000)
    x = (np.arange(n[0] + 3) - 0.5) * ext[0] / n[0]
    want = 1.0 + 2.0 * x
    T = h.get("T")
    assert np.max(np.abs(T[1:-2, 1:-2, 1:-2] - want[1:-2, None, None])) < 1e-10
    r = h.residual_norms(1e30)
    assert r[0] < 1e-9 and r[1] < 1e-9


def test_heat_time_term_decays_geometrically(orc):
    """No flux on any side: div q sums to zero over the cells, so each
    iteration scales sum(T - T_o) by exactly (1 - dtau_T / dt)."""
    n, ext = (9, 7, 6), (1.0, 0.8, 0.6)
    h = _heat(orc, n, ext, [[2, 2]] * 3, [[0.0, 0.0]] * 3)
    rng = np.random.default_rng(4)
    T = rng.standard_normal(h.shape)
    To = rng.standard_normal(h.shape)
    h.set("T", T)
    h.set("TO", To)
    inner = (slice(1, n[0] + 1), slice(1, n[1] + 1), slice(1, n[2] + 1))
    s0 = np.sum(T[inner] - To[inner])
    dt, it = 0.05, 25
    h.advance(dt, it)
    s = np.sum(h.get("T")[inner] - To[inner])
    want = s0 * (1.0 - h.h.dtT / dt) ** it
    assert abs(s - want) < 1e-11 * max(1.0, abs(s0))
    # a plausible mistake, the time term with the wrong sign, does not decay
    assert abs(s0 * (1.0 + h.h.dtT / dt) ** it - want) > 1e-3


def test_heat_backward_euler_cosine_mode(orc):
    """T_o = cos(pi x / L) with no flux on any side: one converged time step
    gives T = T_o / (1 + dt lam kappa^2), kappa^2 = (2/dx)^2 sin^2(pi dx / 2L),
    the discrete eigenvalue of the cell-centred Laplacian."""
    n, ext, lam, dt = (24, 4), (2.0, 0.5), 0.7, 0.01
    h = _heat(orc, n, ext, [[2, 2], [2, 2]], [[0.0, 0.0], [0.0, 0.0]], lam=lam)
    dx = ext[0] / n[0]
    x = (np.arange(n[0] + 3) - 0.5) * dx
    To = np.repeat(np.cos(math.pi * x / ext[0])[:, None], n[1] + 3, axis=1)
    h.set("TO", To)
    h.set("T", To)
    h.advance(dt, 6000)
    k2 = (2.0 / dx) ** 2 * math.sin(math.pi * dx / (2.0 * ext[0])) ** 2
    want = To / (1.0 + dt * lam * k2)
    T = h.get("T")
    assert np.max(np.abs(T[1:-2, 1:-2] - want[1:-2, 1:-2])) < 1e-11
    # the continuous eigenvalue (pi/L)^2 differs visibly at this resolution
    assert np.max(np.abs(To / (1.0 + dt * lam * (math.pi / ext[0]) ** 2) - want)[1:-2, 1:-2]) > 1e-7

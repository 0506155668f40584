"""GPU parity of the FP64 PT-Stokes path against the CPU oracle (through the C ABI).

Tolerance: relative 1e-12 per field (max |gpu - cpu| / max |cpu|) after a
fixed iteration count, equal iteration counts to a residual tolerance, as
the FP64 target of the path states.  With FP contraction off on both sides
the FP64 kernels are expected to agree bit for bit.
"""
import math

import numpy as np
import pytest

from helpers import compare_state, make_pair, rel_err, state_names

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(msg):
    print(msg, flush=True)


def test_manufactured_3d_fixed_iterations(pkg, orc):
    _say("case: 3D manufactured free slip 32^3, 2000 iterations")
    kw, cad = pkg.manufactured_3d(32)
    model, ref = make_pair(pkg, orc, kw)
    model.apply_bc("eta")
    ref.apply_bc("eta")
    model.advance(2000)
    ref.advance(2000)
    worst, bad = compare_state(model, ref, state_names(3) + ["ETA", "ETA_NEXT"], TOL)
    assert not bad, worst
    g = model.compute_residuals()
    c = ref.compute_residuals()
    for name in ("RP", "RVX", "RVY", "RVZ"):
        assert rel_err(model.field(name).parent(), ref.get(name)) <= TOL
    np.testing.assert_allclose(g, c, rtol=TOL, atol=0)
    model.close()


def test_manufactured_2d_converged_iteration_count(pkg, orc):
    _say("case: 2D manufactured free slip 48^2, solve to 1e-8")
    kw, cad = pkg.manufactured_2d(48)
    model, ref = make_pair(pkg, orc, kw)
    it_g, n_g = model.solve(200 * 48, cad["ncheck"], 1e-8)
    it_c, n_c = ref.solve(200 * 48, cad["ncheck"], 1e-8)
    assert it_g == it_c, (it_g, it_c)
    assert max(n_g) < 1e-8
    np.testing.assert_allclose(n_g, n_c, rtol=TOL, atol=0)
    worst, bad = compare_state(model, ref, state_names(2), TOL)
    assert not bad, worst
    model.close()


@pytest.mark.parametrize("dims,scale", [((20, 17, 12), None), ((9, 13, 6), None), ((20, 17, 12), "log"),
                                        ((20, 17, 12), "log10")])
def test_mixed_bcs_powerlaw_3d(pkg, orc, dims, scale):
    _say("case: 3D mixed BCs + regularised power law relaxed in scale %s, dims %s" % (scale, dims))
    grid = pkg.UniformGrid(origin=(-0.5, -0.3, 0.0), extent=(1.0, 0.8, 0.6), dims=dims)
    bcs = dict(x=(pkg.SBC(0.0, 0.1, -0.2), pkg.TBC(0.3, -0.1, 0.05)),
               y=(pkg.VBC(0.02, 0.0, -0.01), pkg.SBC(0.0, 0.2, 0.1)),
               z=(pkg.VBC(0.0, 0.0, 0.0), pkg.TBC(0.0, 0.0, 0.0)))
    rheo = pkg.Relaxation("eta", pkg.ViscoplasticRegularisation(pkg.PowerLawRheology(0.7, 3), 50.0), 0.3,
                          scale=scale)
    kw = dict(grid=grid, boundary_conditions=bcs, gravity=dict(x=0.1, y=-0.2, z=1.0), rheology=rheo,
              solver_params=pkg.pt_params(grid, r=0.9, re_mech=5 * math.pi))
    model, ref = make_pair(pkg, orc, kw, seed=7)
    for m in (model, ref):
        m.apply_bc("eta")
    model.advance(300)
    ref.advance(300)
    worst, bad = compare_state(model, ref, state_names(3) + ["ETA", "ETA_NEXT"], TOL)
    assert not bad, worst
    np.testing.assert_allclose(model.compute_residuals(), ref.compute_residuals(), rtol=TOL, atol=0)
    model.close()


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 1), (1, 6, 2), (3, 4, 5), (4, 4, 4), (7, 2, 3)])
def test_tiny_grids_3d(pkg, orc, dims):
    """Grids too small for the fused interior (fewer than 4 cells along some
    axis: the pointwise kernels sweep the whole box), and the smallest ones
    that have it: 40 iterations with every boundary-condition kind, bitwise
    to the oracle, ghosts and residuals included."""
    _say("case: tiny 3D grid %s, mixed BCs, relaxed power law" % (dims,))
    grid = pkg.UniformGrid(origin=(0.0, 0.0, 0.0), extent=(0.3, 0.5, 0.4), dims=dims)
    bcs = dict(x=(pkg.SBC(0.0, 0.1, -0.2), pkg.TBC(0.3, -0.1, 0.05)),
               y=(pkg.VBC(0.02, 0.0, -0.01), pkg.SBC(0.0, 0.2, 0.1)),
               z=(pkg.VBC(0.0, 0.0, 0.0), pkg.TBC(0.0, 0.0, 0.0)))
    rheo = pkg.Relaxation("eta", pkg.ViscoplasticRegularisation(pkg.PowerLawRheology(0.7, 3), 50.0), 0.3)
    kw = dict(grid=grid, boundary_conditions=bcs, gravity=dict(x=0.1, y=-0.2, z=1.0), rheology=rheo,
              solver_params=pkg.pt_params(grid, r=0.7, re_mech=2 * math.pi))
    model, ref = make_pair(pkg, orc, kw, seed=11)
    for m in (model, ref):
        m.apply_bc("eta")
    model.advance(40)
    ref.advance(40)
    worst, bad = compare_state(model, ref, state_names(3) + ["ETA", "ETA_NEXT"], TOL)
    assert not bad, worst
    np.testing.assert_allclose(model.compute_residuals(), ref.compute_residuals(), rtol=TOL, atol=0)
    model.close()


@pytest.mark.parametrize("dims", [(1, 1), (2, 5), (4, 1), (3, 3)])
def test_tiny_grids_2d(pkg, orc, dims):
    """2D sections of one to five cells per axis: 40 iterations bitwise to the
    oracle."""
    _say("case: tiny 2D grid %s, traction / slip / no-slip" % (dims,))
    grid = pkg.UniformGrid(origin=(0.0, 0.0), extent=(0.4, 0.3), dims=dims)
    bcs = dict(x=(pkg.SBC(0.0, 0.0), pkg.TBC(0.1, -0.05)), y=(pkg.VBC(0.0, 0.0), pkg.TBC(0.0, 0.0)))
    kw = dict(grid=grid, boundary_conditions=bcs, gravity=dict(x=0.2, y=1.0),
              rheology=pkg.LinearViscousRheology(1.0),
              solver_params=pkg.pt_params(grid, r=0.7, re_mech=2 * math.pi))
    model, ref = make_pair(pkg, orc, kw, seed=5)
    model.advance(40)
    ref.advance(40)
    worst, bad = compare_state(model, ref, state_names(2), TOL)
    assert not bad, worst
    np.testing.assert_allclose(model.compute_residuals(), ref.compute_residuals(), rtol=TOL, atol=0)
    model.close()


def test_mixed_bcs_linear_field_2d(pkg, orc):
    _say("case: 2D traction/no-slip with a viscosity field")
    grid = pkg.UniformGrid(origin=(0.0, 0.0), extent=(1.0, 0.5), dims=(40, 23))
    etaf = lambda x, y: 1.0 + 0.5 * np.sin(3 * x) * np.cos(5 * y)
    kw = dict(grid=grid, boundary_conditions=dict(x=(pkg.SBC(0, 0), pkg.SBC(0, 0)),
                                                   y=(pkg.VBC(0.0, 0.0), pkg.TBC(0.0, 0.0))),
              gravity=dict(x=0.0, y=1.0), rheology=pkg.LinearViscousRheology(etaf),
              solver_params=pkg.pt_params(grid, r=0.7, re_mech=2 * math.pi))
    model, ref = make_pair(pkg, orc, kw, eta=etaf, seed=3)
    for m in (model, ref):
        m.apply_bc("eta")
    model.advance(500)
    ref.advance(500)
    worst, bad = compare_state(model, ref, state_names(2) + ["ETA", "ETA_NEXT"], TOL)
    assert not bad, worst
    model.close()


@pytest.mark.parametrize("case", ["mixed3d", "noslip3d", "freeslip3d", "mixed2d"])
def test_bc_groups_bitwise(pkg, orc, case):
    """Every boundary-condition group, applied to a random state (ghosts
    included), equals the oracle's x-then-y-then-z passes bit for bit, edge
    and corner ghosts included."""
    _say("case: bc groups bitwise, %s" % case)
    if case.endswith("3d"):
        grid = pkg.UniformGrid(origin=(0.0, 0.0, 0.0), extent=(1.0, 0.8, 0.6), dims=(9, 7, 5))
        bcs = {"mixed3d": dict(x=(pkg.SBC(0.0, 0.1, -0.2), pkg.TBC(0.3, -0.1, 0.05)),
                               y=(pkg.VBC(0.02, 0.0, -0.01), pkg.SBC(0.0, 0.2, 0.1)),
                               z=(pkg.VBC(0.1, -0.3, 0.0), pkg.TBC(0.0, 0.4, 0.0))),
               "noslip3d": dict(x=(pkg.VBC(0.1, 0.2, 0.3), pkg.VBC(-0.1, 0.0, 0.5)),
                                y=(pkg.VBC(0.0, 0.7, 0.0), pkg.VBC(0.2, 0.0, -0.4)),
                                z=(pkg.VBC(0.3, 0.0, 0.0), pkg.VBC(0.0, -0.6, 0.1))),
               "freeslip3d": dict(x=(pkg.SBC(0.0, 0.0, 0.0), pkg.SBC(0.0, 0.0, 0.0)),
                                  y=(pkg.SBC(0.0, 0.0, 0.0), pkg.SBC(0.0, 0.0, 0.0)),
                                  z=(pkg.SBC(0.0, 0.0, 0.0), pkg.TBC(0.0, 0.0, 0.0)))}[case]
        nd = 3
    else:
        grid = pkg.UniformGrid(origin=(0.0, 0.0), extent=(1.0, 0.5), dims=(11, 6))
        bcs = dict(x=(pkg.VBC(0.3, -0.2), pkg.SBC(0.0, 0.1)), y=(pkg.TBC(0.2, 0.0), pkg.VBC(0.0, 0.4)))
        nd = 2
    kw = dict(grid=grid, boundary_conditions=bcs, gravity=dict(x=0.0, y=1.0), rheology=pkg.LinearViscousRheology(1.0),
              solver_params=pkg.pt_params(grid, r=0.7, re_mech=2 * math.pi))
    model, ref = make_pair(pkg, orc, kw, seed=11, amp=1.0)
    rng = np.random.default_rng(5)
    for name in ("ETA", "ETA_NEXT"):
        a = np.asfortranarray(1.0 + rng.random(grid.parent_shape()))
        model.field(name).set(a)
        ref.set(name, a)
    names = state_names(nd) + ["ETA", "ETA_NEXT"]
    for group in ("stress", "velocity", "eta"):
        model.apply_bc(group)
        ref.apply_bc(group)
        for name in names:
            a, b = model.field(name).parent(), ref.get(name)
            assert np.array_equal(a, b), (group, name, float(np.max(np.abs(a - b))))
    model.close()


def test_nonfinite_is_reported(pkg, orc):
    _say("case: NaN injected into one pressure cell")
    kw, cad = pkg.manufactured_3d(12)
    model, ref = make_pair(pkg, orc, kw)
    P = model.stress.P.parent()
    P[5, 6, 7] = np.nan
    model.stress.P.set(P)
    with pytest.raises(pkg.NonFiniteError):
        model.solve(100, 12, 1e-8)
    model.close()


def _glacier_ib_pair(pkg, orc, dims, seed=5):
    gl = pkg.GaussianGlacier(seed=seed)
    grid = pkg.UniformGrid(origin=(-1.0, -1.0, 0.0), extent=(2.0, 2.0, 1.0), dims=dims)
    dgrid = pkg.UniformGrid(origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=(dims[0], dims[1]))
    tr = pkg.TBC(0.0, 0.0, 0.0)
    kw = dict(grid=grid, boundary_conditions=dict(x=(tr, tr), y=(tr, tr), z=(tr, tr)),
              gravity=dict(x=0.0, y=0.0, z=1.0), rheology=pkg.LinearViscousRheology(1.0),
              solver_params=pkg.pt_params(grid, r=0.9, re_mech=5 * math.pi))
    model, ref = make_pair(pkg, orc, kw, ib=True)
    return gl, grid, dgrid, model, ref


def test_levelset_rotated_dem(pkg, orc):
    """compute_levelset_from_dem! with a rotation R (compute_levelsets.jl:1-8,
    the R of DataElevation): the GPU level sets of a glacier seen through a
    rotation about x and z match the oracle's."""
    _say("case: rotated level sets from DEM, 24x20x12 glacier")
    gl, grid, dgrid, model, ref = _glacier_ib_pair(pkg, orc, (24, 20, 12))
    from fastice_jl_amd.level_sets import dem_parent
    a, b = 0.2, 0.35
    Rx = np.array([[1.0, 0.0, 0.0], [0.0, math.cos(a), -math.sin(a)], [0.0, math.sin(a), math.cos(a)]])
    Rz = np.array([[math.cos(b), -math.sin(b), 0.0], [math.sin(b), math.cos(b), 0.0], [0.0, 0.0, 1.0]])
    R = Rz @ Rx
    for phase, fn, psi in ((0, gl.surface, model.level_sets.na), (1, gl.bed, model.level_sets.ns)):
        pkg.compute_levelset_from_dem(psi, fn, dgrid, grid, R=R)
        ref.levelset_from_dem(phase, dem_parent(fn, dgrid), dgrid.dims, dgrid.origin, dgrid.extent_, R=R)
    for name in ("PSI_NA", "PSI_NS"):
        assert rel_err(model.field(name).parent(), ref.get(name)) <= TOL, name
    model.close()


def test_levelset_and_volume_fractions_3d(pkg, orc):
    _say("case: level sets from DEM + volume fractions, 24x20x12 glacier")
    gl, grid, dgrid, model, ref = _glacier_ib_pair(pkg, orc, (24, 20, 12))
    from fastice_jl_amd.level_sets import dem_parent
    for phase, fn, psi, om in ((0, gl.surface, model.level_sets.na, model.field_masks.na),
                               (1, gl.bed, model.level_sets.ns, model.field_masks.ns)):
        pkg.compute_levelset_from_dem(psi, fn, dgrid, grid)
        ref.levelset_from_dem(phase, dem_parent(fn, dgrid), dgrid.dims, dgrid.origin, dgrid.extent_)
    pkg.invert_levelset(model.level_sets.ns)
    ref.invert_levelset(1)
    for name in ("PSI_NA", "PSI_NS"):
        assert rel_err(model.field(name).parent(), ref.get(name)) <= TOL, name
    pkg.omega_from_psi(model.field_masks.ns, model.level_sets.ns)
    pkg.omega_from_psi(model.field_masks.na, model.level_sets.na)
    ref.omega_from_psi(1)
    ref.omega_from_psi(0)
    for ph, base in (("NA", pkg._abi.FI_F_OM_NA0), ("NS", pkg._abi.FI_F_OM_NS0)):
        for q in range(8):
            g = pkg.full_stokes.DeviceField(model, base + q).parent()
            c = ref.get("OM_%s%d" % (ph, q))
            assert np.max(np.abs(g - c)) <= 1e-13, (ph, q)
    model.close()


@pytest.mark.parametrize("dims", [(24, 20, 12), (6, 5, 4), (3, 7, 3), (9, 2, 5)])
def test_immersed_boundary_iterations_3d(pkg, orc, dims):
    """400 immersed-boundary iterations of a glacier against the oracle, at a
    size with a fused interior and at tiny sizes swept pointwise."""
    _say("case: immersed-boundary PT iterations, %s glacier, 400 iterations" % (dims,))
    gl, grid, dgrid, model, ref = _glacier_ib_pair(pkg, orc, dims)
    from fastice_jl_amd.level_sets import dem_parent
    pkg.compute_levelset_from_dem(model.level_sets.na, gl.surface, dgrid, grid)
    pkg.compute_levelset_from_dem(model.level_sets.ns, gl.bed, dgrid, grid)
    pkg.invert_levelset(model.level_sets.ns)
    pkg.omega_from_psi(model.field_masks.ns, model.level_sets.ns)
    pkg.omega_from_psi(model.field_masks.na, model.level_sets.na)
    # hand the GPU masks to the oracle so the sweep is compared on identical inputs
    for ph, base in (("NA", pkg._abi.FI_F_OM_NA0), ("NS", pkg._abi.FI_F_OM_NS0)):
        for q in range(8):
            ref.set("OM_%s%d" % (ph, q), pkg.full_stokes.DeviceField(model, base + q).parent())
    for m in (model, ref):
        m.apply_bc("eta")
    model.advance(400)
    ref.advance(400)
    worst, bad = compare_state(model, ref, state_names(3) + ["ETA"], TOL)
    assert not bad, worst
    np.testing.assert_allclose(model.compute_residuals(), ref.compute_residuals(), rtol=TOL, atol=0)
    model.close()


# --- fused interior sweep (grids with at least one 62-wide x tile) ---------

@pytest.mark.parametrize("zchunk", [32, 5])
def test_fused_manufactured_3d(pkg, orc, zchunk):
    _say("case: fused sweep, 3D manufactured 70x36x24, zchunk %d, 300 iterations" % zchunk)
    kw, cad = pkg.manufactured_3d((70, 36, 24))
    model, ref = make_pair(pkg, orc, kw, seed=11)
    model.set_option("zchunk", zchunk)
    for m in (model, ref):
        m.apply_bc("eta")
    model.advance(300)
    ref.advance(300)
    worst, bad = compare_state(model, ref, state_names(3) + ["ETA", "ETA_NEXT"], TOL)
    assert not bad, worst
    np.testing.assert_allclose(model.compute_residuals(), ref.compute_residuals(), rtol=TOL, atol=0)
    model.close()


def test_fused_equals_pointwise_bitwise(pkg, orc):
    _say("case: fused vs pointwise sweep, mixed BCs + power law, 130x20x17")
    grid = pkg.UniformGrid(origin=(-0.5, -0.3, 0.0), extent=(1.3, 0.2, 0.17), dims=(130, 20, 17))
    bcs = dict(x=(pkg.SBC(0.0, 0.1, -0.2), pkg.TBC(0.3, -0.1, 0.05)),
               y=(pkg.VBC(0.02, 0.0, -0.01), pkg.SBC(0.0, 0.2, 0.1)),
               z=(pkg.VBC(0.0, 0.0, 0.0), pkg.TBC(0.0, 0.0, 0.0)))
    rheo = pkg.Relaxation("eta", pkg.ViscoplasticRegularisation(pkg.PowerLawRheology(0.7, 3), 50.0), 0.3)
    kw = dict(grid=grid, boundary_conditions=bcs, gravity=dict(x=0.1, y=-0.2, z=1.0), rheology=rheo,
              solver_params=pkg.pt_params(grid, r=0.9, re_mech=5 * math.pi))
    a, ref = make_pair(pkg, orc, kw, seed=4)
    b, _ = make_pair(pkg, orc, kw, seed=4)
    b.set_option("fused", 0)
    for m in (a, b, ref):
        m.apply_bc("eta")
    a.advance(200)
    b.advance(200)
    ref.advance(200)
    for name in state_names(3) + ["ETA"]:
        fa = a.field(name).parent()
        assert np.all(np.isfinite(fa)), name
        assert np.array_equal(fa, b.field(name).parent()), name
    worst, bad = compare_state(a, ref, state_names(3) + ["ETA"], TOL)
    assert not bad, worst
    a.close()
    b.close()


def test_fused_immersed_boundary(pkg, orc):
    _say("case: fused sweep, immersed boundary glacier 70x40x20, 300 iterations")
    gl, grid, dgrid, model, ref = _glacier_ib_pair(pkg, orc, (70, 40, 20))
    model.set_option("zchunk", 7)
    pkg.compute_levelset_from_dem(model.level_sets.na, gl.surface, dgrid, grid)
    pkg.compute_levelset_from_dem(model.level_sets.ns, gl.bed, dgrid, grid)
    pkg.invert_levelset(model.level_sets.ns)
    pkg.omega_from_psi(model.field_masks.ns, model.level_sets.ns)
    pkg.omega_from_psi(model.field_masks.na, model.level_sets.na)
    for ph, base in (("NA", pkg._abi.FI_F_OM_NA0), ("NS", pkg._abi.FI_F_OM_NS0)):
        for q in range(8):
            ref.set("OM_%s%d" % (ph, q), pkg.full_stokes.DeviceField(model, base + q).parent())
    for m in (model, ref):
        m.apply_bc("eta")
    model.advance(300)
    ref.advance(300)
    worst, bad = compare_state(model, ref, state_names(3) + ["ETA"], TOL)
    assert not bad, worst
    np.testing.assert_allclose(model.compute_residuals(), ref.compute_residuals(), rtol=TOL, atol=0)
    model.close()


@pytest.mark.parametrize("order", [0, 1, 2])
def test_every_fused_variant_equals_pointwise(pkg, order):
    """Every compiled fused-sweep variant (64-lane tiles at any alignment,
    66-lane line-aligned tiles, register limits) and both tile orders give the
    pointwise sweep's results bitwise, plain and immersed-boundary."""
    variants = pkg._abi.fused_variants(False)
    _say("case: %d FP64 fused variants, tile order %d, plain 97x23x19 and glacier 70x40x20" % (len(variants), order))
    grid = pkg.UniformGrid(origin=(-0.5, -0.3, 0.0), extent=(1.0, 0.3, 0.2), dims=(97, 23, 19))
    bcs = dict(x=(pkg.SBC(0.0, 0.1, -0.2), pkg.TBC(0.3, -0.1, 0.05)),
               y=(pkg.VBC(0.02, 0.0, -0.01), pkg.SBC(0.0, 0.2, 0.1)),
               z=(pkg.VBC(0.0, 0.0, 0.0), pkg.TBC(0.0, 0.0, 0.0)))
    rheo = pkg.Relaxation("eta", pkg.ViscoplasticRegularisation(pkg.PowerLawRheology(0.7, 3), 50.0), 0.3)
    kw = dict(grid=grid, boundary_conditions=bcs, gravity=dict(x=0.1, y=-0.2, z=1.0), rheology=rheo,
              solver_params=pkg.pt_params(grid, r=0.9, re_mech=5 * math.pi))
    rng = np.random.default_rng(9)
    init = {n: np.asfortranarray(1e-2 * rng.standard_normal(grid.parent_shape())) for n in state_names(3)}

    def plain():
        m = pkg.IsothermalFullStokesModel(**kw)
        m.viscosity.eta.set(1.0)
        m.viscosity.eta_next.set(1.0)
        for n, a in init.items():
            m.field(n).set(a)
        m.apply_bc("eta")
        return m

    def glacier():
        m, _ = pkg.glacier_3d((70, 40, 20), seed=6, glen=True)
        return m

    for make, n_it in ((plain, 60), (glacier, 60)):
        ref = make()
        ref.set_option("fused", 0)
        ref.advance(n_it)
        want = {n: ref.field(n).parent() for n in state_names(3) + ["ETA"]}
        ref.close()
        for v, cfg in enumerate(variants):
            m = make()
            m.set_option("fused_variant", v)
            m.set_option("tile_order", order)
            m.set_option("zchunk", 7)
            m.advance(n_it)
            for n, f in want.items():
                g = m.field(n).parent()
                assert np.array_equal(g, f), (make.__name__, cfg, n, float(np.nanmax(np.abs(g - f))))
            m.close()


def test_glacier_2d_levelsets_fractions_iterations(pkg, orc):
    _say("case: 2D glacier section 96x40: polygon level sets, volume fractions, 300 IB iterations")
    kw, cad = pkg.glacier_2d_spec((96, 40), seed=4, glen=True)
    model, ref = make_pair(pkg, orc, kw, ib=True)
    pkg.compute_levelset_from_polygon(model.level_sets.na, cad["poly_na"])
    pkg.compute_levelset_from_polygon(model.level_sets.ns, cad["poly_ns"])
    ref.levelset_from_polygon(0, cad["poly_na"])
    ref.levelset_from_polygon(1, cad["poly_ns"])
    for name in ("PSI_NA", "PSI_NS"):
        assert rel_err(model.field(name).parent(), ref.get(name)) <= TOL, name
    pkg.omega_from_psi(model.field_masks.ns, model.level_sets.ns)
    pkg.omega_from_psi(model.field_masks.na, model.level_sets.na)
    ref.omega_from_psi(1)
    ref.omega_from_psi(0)
    for ph, base in (("NA", pkg._abi.FI_F_OM_NA0), ("NS", pkg._abi.FI_F_OM_NS0)):
        for q in range(4):
            g = pkg.full_stokes.DeviceField(model, base + q).parent()
            assert rel_err(g, ref.get("OM_%s%d" % (ph, q))) <= TOL, (ph, q)
    om = ref.get("OM_NA0")
    assert 0.05 < om[1:97, 1:41].mean() < 0.95  # the section cuts the ice
    for m in (model, ref):
        m.apply_bc("eta")
    model.advance(300)
    ref.advance(300)
    worst, bad = compare_state(model, ref, state_names(2) + ["ETA"], TOL)
    assert not bad, worst
    np.testing.assert_allclose(model.compute_residuals(), ref.compute_residuals(), rtol=TOL, atol=0)
    model.close()


def test_shell_in_fused_launch_bitwise(pkg):
    """The boundary shell's stress update inside the fused launch gives the
    same fields, bit for bit, as its own launch (FP64 glacier iterations and
    FP32 correction sweeps)."""
    _say("case: shell stress in the fused launch vs separate, 70x40x20 glacier")
    out = []
    for fs in (0, 1):
        m, _ = pkg.glacier_3d((70, 40, 20), seed=6, glen=True)
        m.set_option("fused_shell", fs)
        m.advance(9)
        pkg._abi.check(pkg._abi.lib().fi_refined_advance_inner(m._h, 7))
        m.synchronize()
        out.append({n: m.field(n).parent() for n in state_names(3) + ["ETA"]})
        m.close()
    for n in out[0]:
        assert np.array_equal(out[0][n], out[1][n]), n


def test_ib_sweep_skips_empty_space_bitwise(pkg, orc):
    """The immersed-boundary sweep reads and writes nothing on the planes of a
    column that hold no valid location, reads no old stress at an invalid
    location, and skips velocity reads and stores of invalid locations only
    once the buffers are known to hold +0 there.  From a random state that is
    non-zero at invalid locations as well, the first iteration (which reads
    and writes everything), the second (which may skip reads) and the later
    ones (which skip both) equal the pointwise sweep bit for bit and the
    oracle; an upload of one field starts the sequence again; and a refined
    solve (FP32 sweeps after the full FP32 phase and after zero fills and
    stress lags) gives the same FP64 state with and without the fused sweep."""
    _say("case: empty-space skipping, 70x40x20 glacier, random state, fused vs pointwise vs oracle")
    import ctypes as C
    rng = np.random.default_rng(21)
    ms = []
    for fused in (1, 0):
        m, _ = pkg.glacier_3d((70, 40, 20), seed=6, glen=True)
        m.set_option("fused", fused)
        m.set_option("zchunk", 7)
        ms.append(m)
    a, b = ms
    ref = orc.OracleModel(a.desc)
    for ph, base in (("NA", pkg._abi.FI_F_OM_NA0), ("NS", pkg._abi.FI_F_OM_NS0)):
        for q in range(8):
            ref.set("OM_%s%d" % (ph, q), pkg.full_stokes.DeviceField(a, base + q).parent())
    shape = a.grid.parent_shape()
    for n in state_names(3) + ["ETA", "ETA_NEXT"]:
        x = np.asfortranarray((1.0 + 0.1 * rng.random(shape)) if n.startswith("ETA") else 1e-2 * rng.standard_normal(shape))
        for m in (a, b):
            m.field(n).set(x)
        ref.set(n, x)
    cells = (C.c_longlong * 2)()
    pkg._abi.check(pkg._abi.lib().fi_fused_active_cells(a._h, cells))
    _say("fused interior %d cells, %d on planes with a valid location" % (cells[0], cells[1]))
    assert 0 < cells[1] < cells[0]

    def same(tag):
        for name in state_names(3) + ["ETA"]:
            fa = a.field(name).parent()
            assert np.all(np.isfinite(fa)), (tag, name)
            assert np.array_equal(fa, b.field(name).parent()), (tag, name)
        worst, bad = compare_state(a, ref, state_names(3) + ["ETA"], TOL)
        assert not bad, (tag, worst)

    for step, n_it in enumerate((1, 1, 1, 4)):
        for m in (a, b, ref):
            m.advance(n_it)
        same("iteration block %d" % step)
    v = np.asfortranarray(1e-2 * rng.standard_normal(shape))
    for m in (a, b):
        m.field("VX").set(v)
    ref.set("VX", v)
    for step in range(3):
        for m in (a, b, ref):
            m.advance(1)
        same("after an upload, iteration %d" % step)
    oa = a.refined_solve(1e-12, max_outer=3, max_inner=140, ncheck=16)
    ob = b.refined_solve(1e-12, max_outer=3, max_inner=140, ncheck=16)
    assert oa[0] == ob[0] and oa[1] == ob[1]
    for name in state_names(3) + ["ETA"]:
        fa = a.field(name).parent()
        assert np.all(np.isfinite(fa)), name
        assert np.array_equal(fa, b.field(name).parent()), ("refined", name)
    a.close()
    b.close()


def test_ib_skipping_nonzero_traction_bitwise(pkg, orc):
    """Empty-space skipping with prescribed non-zero tractions and slip on
    the box faces, whose ghosts in empty space hold non-zero values: once a
    stress buffer's x-face ghosts are settled, the boundary conditions skip
    those whose point and source lie in empty space.  From a random state,
    fused and pointwise sweeps agree bit for bit, and with the oracle (ghosts
    included), over the iterations that skip nothing, skip reads, and skip
    both stores and settled ghosts."""
    _say("case: empty-space skipping, 40x32x16 glacier, non-zero traction and slip faces")
    from fastice_jl_amd.problems import glacier_3d_spec
    from fastice_jl_amd.level_sets import compute_levelset_from_dem, invert_levelset, omega_from_psi
    kw, cad = glacier_3d_spec((40, 32, 16), seed=6, glen=False)
    kw["boundary_conditions"] = dict(x=(pkg.TBC(0.1, -0.2, 0.3), pkg.TBC(-0.4, 0.05, 0.2)),
                                     y=(pkg.SBC(0.0, 0.3, -0.1), pkg.TBC(0.2, 0.1, -0.3)),
                                     z=(pkg.TBC(0.0, 0.25, 0.15), pkg.TBC(0.3, -0.2, 0.1)))
    gl, dgrid, grid = cad["glacier"], cad["dem_grid"], kw["grid"]
    ms = []
    for fused in (1, 0):
        m = pkg.IsothermalImmersedBoundaryFullStokesModel(**kw)
        compute_levelset_from_dem(m.level_sets.na, gl.surface, dgrid, grid)
        compute_levelset_from_dem(m.level_sets.ns, gl.bed, dgrid, grid)
        invert_levelset(m.level_sets.ns)
        omega_from_psi(m.field_masks.ns, m.level_sets.ns)
        omega_from_psi(m.field_masks.na, m.level_sets.na)
        m.set_option("fused", fused)
        ms.append(m)
    a, b = ms
    ref = orc.OracleModel(a.desc)
    for ph, base in (("NA", pkg._abi.FI_F_OM_NA0), ("NS", pkg._abi.FI_F_OM_NS0)):
        for q in range(8):
            ref.set("OM_%s%d" % (ph, q), pkg.full_stokes.DeviceField(a, base + q).parent())
    rng = np.random.default_rng(23)
    shape = grid.parent_shape()
    for n in state_names(3) + ["ETA", "ETA_NEXT"]:
        x = np.asfortranarray((1.0 + 0.1 * rng.random(shape)) if n.startswith("ETA") else 1e-2 * rng.standard_normal(shape))
        for m in (a, b):
            m.field(n).set(x)
        ref.set(n, x)
    for step, n_it in enumerate((1, 1, 1, 1, 4)):
        for m in (a, b, ref):
            m.advance(n_it)
        for name in state_names(3) + ["ETA"]:
            fa = a.field(name).parent()
            assert np.all(np.isfinite(fa)), (step, name)
            assert np.array_equal(fa, b.field(name).parent()), (step, name)
        worst, bad = compare_state(a, ref, state_names(3) + ["ETA"], TOL)
        assert not bad, (step, worst)
    a.close()
    b.close()


@pytest.mark.parametrize("law", ["powerlaw_A_field", "linear_eta_field"])
def test_ib_skipping_field_rheology_bitwise(pkg, orc, law):
    """Empty-space skipping with a rheology whose coefficient is a field: the
    viscosity of the unmarched planes (k_eta_rest) then evaluates the target
    per cell.  From a random state, fused and pointwise sweeps agree bit for
    bit, and with the oracle, over the iterations that skip nothing, skip
    reads, and skip both."""
    _say("case: empty-space skipping, 70x40x20 glacier, %s" % law)
    dims = (70, 40, 20)
    gl = pkg.GaussianGlacier(seed=6)
    grid = pkg.UniformGrid(origin=(-1.0, -1.0, 0.0), extent=(2.0, 2.0, 1.0), dims=dims)
    dgrid = pkg.UniformGrid(origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=dims[:2])
    tr = pkg.TBC(0.0, 0.0, 0.0)
    coef = lambda x, y, z: 1.0 + 0.5 * np.sin(3.0 * x) * np.cos(2.0 * y) + 0.2 * z
    if law == "powerlaw_A_field":
        rheo = pkg.Relaxation("eta", pkg.ViscoplasticRegularisation(pkg.PowerLawRheology(coef, 3), 30.0), 0.05,
                              scale="log")
    else:
        rheo = pkg.Relaxation("eta", pkg.LinearViscousRheology(coef), 0.2)
    kw = dict(grid=grid, boundary_conditions=dict(x=(tr, tr), y=(tr, tr), z=(tr, tr)),
              gravity=dict(x=0.0, y=0.0, z=1.0), rheology=rheo,
              solver_params=pkg.pt_params(grid, r=0.9, re_mech=5 * math.pi))
    ms = []
    for fused in (1, 0):
        m = pkg.IsothermalImmersedBoundaryFullStokesModel(**kw)
        m.set_option("fused", fused)
        m.set_option("zchunk", 7)
        pkg.compute_levelset_from_dem(m.level_sets.na, gl.surface, dgrid, grid)
        pkg.compute_levelset_from_dem(m.level_sets.ns, gl.bed, dgrid, grid)
        pkg.invert_levelset(m.level_sets.ns)
        pkg.omega_from_psi(m.field_masks.ns, m.level_sets.ns)
        pkg.omega_from_psi(m.field_masks.na, m.level_sets.na)
        ms.append(m)
    a, b = ms
    ref = orc.OracleModel(a.desc)
    ref.set("RHEO", a.field("RHEO").parent())
    for ph, base in (("NA", pkg._abi.FI_F_OM_NA0), ("NS", pkg._abi.FI_F_OM_NS0)):
        for q in range(8):
            ref.set("OM_%s%d" % (ph, q), pkg.full_stokes.DeviceField(a, base + q).parent())
    rng = np.random.default_rng(33)
    shape = grid.parent_shape()
    for n in state_names(3) + ["ETA", "ETA_NEXT"]:
        x = np.asfortranarray((1.0 + 0.1 * rng.random(shape)) if n.startswith("ETA") else 1e-2 * rng.standard_normal(shape))
        for m in (a, b):
            m.field(n).set(x)
        ref.set(n, x)
    for step, n_it in enumerate((1, 1, 1, 5)):
        for m in (a, b, ref):
            m.advance(n_it)
        for name in state_names(3) + ["ETA"]:
            fa = a.field(name).parent()
            assert np.all(np.isfinite(fa)), (step, name)
            assert np.array_equal(fa, b.field(name).parent()), (step, name)
        worst, bad = compare_state(a, ref, state_names(3) + ["ETA"], TOL)
        assert not bad, (step, worst)
    a.close()
    b.close()


def test_outer_width_changes_nothing(pkg):
    """outer_width (the reference's split of the domain into an interior
    launched first and boundary slabs overlapped with the halo exchange,
    Isothermal.jl:60-89) has nothing to overlap on one device: a model built
    with it iterates bit for bit as one built without."""
    _say("case: outer_width (16, 8, 4) vs none, 3D manufactured 40x36x24, 200 iterations")
    kw, _ = pkg.manufactured_3d((40, 36, 24))
    a = pkg.IsothermalFullStokesModel(**kw)
    b = pkg.IsothermalFullStokesModel(outer_width=(16, 8, 4), **kw)
    assert b.outer_width == (16, 8, 4)
    for m in (a, b):
        m.viscosity.eta.set(1.0)
        m.viscosity.eta_next.set(1.0)
        m.apply_bc("eta")
        m.advance(200)
    for name in state_names(3):
        assert np.array_equal(a.field(name).parent(), b.field(name).parent()), name
    assert np.array_equal(a.compute_residuals(), b.compute_residuals())
    a.close()
    b.close()

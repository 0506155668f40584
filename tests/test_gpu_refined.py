"""FP32 sweeps inside FP64 defect correction (through the C ABI).

The refined solve must reach the same FP64 residual tolerance as the plain
FP64 solve; the per-field difference of the two converged states is bounded
by what that tolerance allows (stated per test).
"""
import math

import numpy as np
import pytest

from helpers import state_names

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _field_errors(a, b, names):
    """Max difference per field, relative to the scale of its group (stress or
    velocity); pressure is compared modulo its mean (a free-slip or traction
    box leaves a constant pressure mode free)."""
    vals = {n: (a.field(n).interior(), b.field(n).interior()) for n in names}
    stress = [n for n in names if n == "P" or n.startswith("T")]
    vel = [n for n in names if n.startswith("V")]
    s_scale = max(np.max(np.abs(vals[n][1])) for n in stress) if stress else 1.0
    v_scale = max(np.max(np.abs(vals[n][1])) for n in vel) if vel else 1.0
    out = {}
    for n, (x, y) in vals.items():
        if n == "P":
            x, y = x - x.mean(), y - y.mean()
        out[n] = float(np.max(np.abs(x - y)) / (s_scale if n in stress else v_scale))
    return out


def test_refined_manufactured_3d_matches_fp64(pkg):
    _say("case: refined vs FP64, 3D manufactured 70x36x24 (fused interior), tol 1e-8")
    kw, cad = pkg.manufactured_3d((70, 36, 24))
    a = pkg.IsothermalFullStokesModel(**kw)
    b = pkg.IsothermalFullStokesModel(**kw)
    for m in (a, b):
        m.viscosity.eta.set(1.0)
        m.viscosity.eta_next.set(1.0)
    it, nb = b.solve(400 * 70, cad["ncheck"], 1e-8)
    assert max(nb) < 1e-8
    outer, inner, na = a.refined_solve(1e-8)
    assert max(na) < 1e-8, na
    errs = _field_errors(a, b, state_names(3))
    _say("outer %d inner %d fp64 iters %d errors %s" % (outer, inner, it, errs))
    # both states satisfy the discrete equations to 1e-8: measured differences ~1e-9
    assert max(errs.values()) < 1e-7, errs
    a.close()
    b.close()


def test_refined_manufactured_2d(pkg):
    _say("case: refined vs FP64, 2D manufactured 96^2, tol 1e-9")
    kw, cad = pkg.manufactured_2d(96)
    a = pkg.IsothermalFullStokesModel(**kw)
    b = pkg.IsothermalFullStokesModel(**kw)
    for m in (a, b):
        m.viscosity.eta.set(1.0)
        m.viscosity.eta_next.set(1.0)
    it, nb = b.solve(400 * 96, 96, 1e-9)
    outer, inner, na = a.refined_solve(1e-9)
    assert max(na) < 1e-9 and max(nb) < 1e-9
    errs = _field_errors(a, b, state_names(2))
    assert max(errs.values()) < 1e-7, errs


@pytest.mark.parametrize("dims", [(6, 5, 4), (12, 10, 6)])
def test_refined_tiny_glacier_matches_fp64(pkg, dims):
    """Refined solves of glaciers too small for the fused interior (every
    sweep pointwise, FP32 masks from the quantised fractions) reach the FP64
    solve's state."""
    _say("case: refined vs FP64, tiny immersed-boundary glacier %s, tol 1e-8" % (dims,))
    gl = pkg.GaussianGlacier(seed=5)
    grid = pkg.UniformGrid(origin=(-1.0, -1.0, 0.0), extent=(2.0, 2.0, 1.0), dims=dims)
    dgrid = pkg.UniformGrid(origin=(-1.0, -1.0), extent=(2.0, 2.0), dims=(dims[0], dims[1]))
    tr = pkg.TBC(0.0, 0.0, 0.0)
    kw = dict(grid=grid, boundary_conditions=dict(x=(tr, tr), y=(tr, tr), z=(tr, tr)),
              gravity=dict(x=0.0, y=0.0, z=1.0), rheology=pkg.LinearViscousRheology(1.0),
              solver_params=pkg.pt_params(grid, r=0.9, re_mech=5 * math.pi))
    models = []
    for _ in range(2):
        m = pkg.IsothermalImmersedBoundaryFullStokesModel(**kw)
        m.viscosity.eta.set(1.0)
        m.viscosity.eta_next.set(1.0)
        pkg.compute_levelset_from_dem(m.level_sets.na, gl.surface, dgrid, grid)
        pkg.compute_levelset_from_dem(m.level_sets.ns, gl.bed, dgrid, grid)
        pkg.invert_levelset(m.level_sets.ns)
        pkg.omega_from_psi(m.field_masks.ns, m.level_sets.ns)
        pkg.omega_from_psi(m.field_masks.na, m.level_sets.na)
        m.apply_bc("eta")
        models.append(m)
    a, b = models
    it, nb = b.solve(20000, 100, 1e-8)
    assert max(nb) < 1e-8, nb
    outer, inner, na = a.refined_solve(1e-8)
    _say("outer %d inner %d fp64 iters %d" % (outer, inner, it))
    assert max(na) < 1e-8, na
    errs = _field_errors(a, b, state_names(3))
    _say("errors %s" % errs)
    assert max(errs.values()) < 1e-6, errs
    a.close()
    b.close()


@pytest.mark.parametrize("mode", [1, 0])
@pytest.mark.parametrize("glen", [False, True])
def test_refined_glacier_matches_fp64(pkg, glen, mode):
    _say("case: refined (mode %d) vs FP64, immersed-boundary glacier 70x40x20, glen=%s" % (mode, glen))
    a, cad = pkg.glacier_3d((70, 40, 20), seed=2, glen=glen)
    b, _ = pkg.glacier_3d((70, 40, 20), seed=2, glen=glen)
    it, nb = b.solve(400 * 70, cad["ncheck"], 1e-8)
    assert max(nb) < 1e-8, nb
    a.set_option("refine_mode", mode)
    outer, inner, na = a.refined_solve(1e-8, max_outer=200 if mode == 0 else None)
    _say("outer %d inner %d fp64 iters %d" % (outer, inner, it))
    assert max(na) < 1e-8, na
    errs = _field_errors(a, b, ["P", "TXX", "TXY", "VX", "VY", "VZ"])
    _say("errors %s" % errs)
    # both states satisfy the discrete equations to residual 1e-8; measured
    # differences up to 1.2e-7 (pressure of the linear glacier, whose Stokes
    # operator on cut cells is the worst conditioned of these cases)
    assert max(errs.values()) < 1e-6, errs
    a.close()
    b.close()


def test_refined_fused_equals_pointwise(pkg):
    _say("case: FP32 correction sweep, fused vs pointwise, 3D manufactured 70x36x24")
    kw, cad = pkg.manufactured_3d((70, 36, 24))
    ms = []
    for fused in (1, 0):
        m = pkg.IsothermalFullStokesModel(**kw)
        m.viscosity.eta.set(1.0)
        m.viscosity.eta_next.set(1.0)
        m.set_option("fused", fused)
        out = m.refined_solve(1e-6, max_outer=3, max_inner=3 * 70, ncheck=35)
        ms.append((m, out))
    (a, oa), (b, ob) = ms
    assert oa[0] == ob[0] and oa[1] == ob[1]
    for name in state_names(3):
        fa = a.field(name).parent()
        assert np.all(np.isfinite(fa)), name
        assert np.array_equal(fa, b.field(name).parent()), name
    a.close()
    b.close()


def test_every_fp32_variant_equals_pointwise(pkg):
    """Every compiled FP32 fused-sweep variant gives the pointwise FP32
    correction sweep's results bitwise (glacier, immersed boundary)."""
    variants = pkg._abi.fused_variants(True)
    _say("case: %d FP32 fused variants, refined glacier 70x40x20, 2 outer steps" % len(variants))

    def run(option):
        m, _ = pkg.glacier_3d((70, 40, 20), seed=6, glen=True)
        option(m)
        out = m.refined_solve(1e-12, max_outer=2, max_inner=140, ncheck=35)
        fields = {n: m.field(n).parent() for n in state_names(3) + ["ETA"]}
        m.close()
        return out, fields

    (o_ref, want) = run(lambda m: m.set_option("fused", 0))
    for v, cfg in enumerate(variants):
        (o, got) = run(lambda m: m.set_option("fused_variant_fp32", v))
        assert o[0] == o_ref[0] and o[1] == o_ref[1], cfg
        for n in want:
            assert np.array_equal(got[n], want[n]), (cfg, n)


def test_refined_column_outer_steps_bitwise(pkg):
    """Lagged-PT outer steps restricted to the column ranges (the default on a
    3D immersed-boundary glacier with traction on every face) give the same
    bits as outer steps over every cell: state, viscosity and residuals."""
    _say("case: refined glacier 70x40x20, column-range outer steps vs whole grid, 6 outer steps")

    def run(cols):
        m, _ = pkg.glacier_3d((70, 40, 20), seed=3, glen=True)
        m.set_option("refine_cols", cols)
        out = m.refined_solve(1e-12, max_outer=6, max_inner=140, ncheck=32)
        fields = {n: m.field(n).parent() for n in state_names(3) + ["ETA", "RP", "RVX", "RVY", "RVZ"]}
        m.close()
        return out, fields

    (oa, fa), (ob, fb) = run(1), run(0)
    _say("outer %d inner %d" % (oa[0], oa[1]))
    assert oa[0] == ob[0] and oa[1] == ob[1]
    assert np.array_equal(np.asarray(oa[2]), np.asarray(ob[2]))
    for n in fa:
        assert np.all(np.isfinite(fa[n])), n
        assert np.array_equal(fa[n], fb[n]), n


def test_refined_glacier_configured_regularisation_255(pkg):
    """The precision workload as configured: the seeded glacier at
    255x255x127 with Glen's law regularised at strain rate 1e-6 (eta_reg =
    5000, a viscosity contrast of about 1e4; PT parameters r 0.5, re 5 pi,
    vdtau_div 4 and log relaxation 0.01, with which plain FP64 PT converges
    there).  FP64 and FP32-refined solves both reach FP64 residual 1e-8; the
    per-field difference of the converged states, relative to the FP64
    stress and velocity scales, is stated and bounded.  About a minute."""
    dims = (255, 255, 127)
    _say("case: refined vs FP64, glacier %s regularised at strain rate 1e-6, tol 1e-8" % (dims,))
    b, cad = pkg.glacier_3d(dims, seed=1, glen=True, eps0=1e-6)
    it, nb = b.solve(400 * max(dims), cad["ncheck"], 1e-8)
    _say("fp64: %d iterations, norms %s" % (it, nb))
    assert max(nb) < 1e-8, nb
    a, _ = pkg.glacier_3d(dims, seed=1, glen=True, eps0=1e-6)
    outer, inner, na = a.refined_solve(1e-8)
    _say("refined: %d outer steps, %d FP32 sweeps, norms %s" % (outer, inner, na))
    assert max(na) < 1e-8, na
    errs = _field_errors(a, b, state_names(3))
    _say("per-field difference (relative to the FP64 stress / velocity scale): %s" % errs)
    # measured at most 1.4e-9 (both states at residual 1e-8)
    assert max(errs.values()) < 1e-7, errs
    eta = a.field("ETA").interior()
    assert eta.max() <= 5000.0 * (1 + 1e-12) and eta.max() > 1000.0, (eta.min(), eta.max())
    a.close()
    b.close()

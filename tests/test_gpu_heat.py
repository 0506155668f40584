"""PT heat diffusion (fi_heat_*) against the oracle (or_heat_*), bitwise.

Random temperatures, previous temperatures and fluxes, boundary sides of
every kind (Value, Flux, none), 2D and 3D, grid sizes that do not fill the
64 x 4 thread blocks or the 32-plane z chunks of the 3D kernel: after a few
PT iterations the fused kernel's T and q equal the oracle's (update_q! +
flux conditions, update_T! + value conditions, in separate passes) bit for
bit, and so do the residual norms.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


CASES = [
    ((21, 13, 9), [[1, 2], [0, 1], [2, 0]]),
    ((70, 6, 5), [[2, 2], [2, 2], [2, 2]]),
    ((17, 9, 70), [[0, 2], [1, 2], [1, 2]]),  # z chunks of 32 planes: seams at 32 and 64
    ((33, 27), [[1, 1], [2, 0]]),
    ((1, 2, 3), [[1, 2], [2, 1], [1, 1]]),  # tiny grids: one or two cells along an axis
    ((2, 1), [[2, 1], [1, 2]]),
]


def _pair(pkg, orc, dims, kinds, seed):
    H = pkg.heat
    nd = len(dims)
    ext = (1.0, 0.8, 0.6)[:nd]
    grid = pkg.UniformGrid(origin=(0.0,) * nd, extent=ext, dims=dims)
    rng = np.random.default_rng(seed)
    vals = rng.uniform(-1.0, 1.0, (3, 2))
    bcs = {}
    for d in range(nd):
        for s in range(2):
            k = kinds[d][s]
            if k:
                bcs[H.SIDES[d][s]] = H.BoundaryCondition(H.Value if k == 1 else H.Flux, vals[d][s])
    dx = min(e / n for e, n in zip(ext, dims))
    params = dict(physics={"λ_ρCp": 0.9}, iter_params={"Δτ": {"q": 0.7, "T": dx * dx / (0.9 * 4.2 * nd)}})
    m = H.HeatDiffusionModel(grid, bcs, **params)
    d = m.desc
    ref = orc.OracleHeat(nd, list(dims) + [1] * (3 - nd), list(ext) + [1.0] * (3 - nd),
                         [[d.bc_kind[q][s] for s in range(2)] for q in range(3)],
                         [[d.bc_value[q][s] for s in range(2)] for q in range(3)], d.lambda_rho_cp, d.dtau_q, d.dtau_T)
    shape = grid.parent_shape()
    fields = {"T": m.temperature, "TO": m.temperature_old, "QX": m.heat_flux.x, "QY": m.heat_flux.y}
    if nd == 3:
        fields["QZ"] = m.heat_flux.z
    for name, f in fields.items():
        a = rng.standard_normal(shape)
        f.set(a)
        ref.set(name, a)
    m.apply_bc()
    ref.apply_bc()
    return m, ref, fields


@pytest.mark.parametrize("case", range(len(CASES)))
def test_heat_iterations_bitwise(pkg, orc, case):
    dims, kinds = CASES[case]
    _say("case: heat PT iterations, dims %s, boundary kinds %s" % (dims, kinds))
    m, ref, fields = _pair(pkg, orc, dims, kinds, seed=10 + case)
    dt = 0.3
    for n in (1, 6):
        m.advance(n, dt)
        ref.advance(dt, n)
        for name, f in fields.items():
            g, c = f.parent(), ref.get(name)
            assert np.array_equal(g, c), (name, float(np.max(np.abs(g - c))))
        assert np.array_equal(m.residual_norms(dt), ref.residual_norms(dt))
    m.close()


def test_heat_timestep_converges_like_oracle(pkg, orc):
    """advance_timestep: PT iterations to the tolerance, then T_o <- T; the
    state and the count match the oracle advanced by as many iterations."""
    _say("case: heat advance_timestep, 24x20x8, Value west/east, Flux elsewhere")
    m, ref, fields = _pair(pkg, orc, (24, 20, 8), [[1, 1], [2, 2], [2, 2]], seed=3)
    dt = 0.05
    it, norms = pkg.heat.advance_timestep(m, 0.0, dt, max_iter=200000, ncheck=50, tol=1e-9)
    assert max(norms) < 1e-9 and it % 50 == 0
    ref.advance(dt, it)
    assert np.array_equal(m.temperature.parent(), ref.get("T"))
    assert np.array_equal(m.temperature_old.parent(), ref.get("T"))
    assert np.allclose(ref.residual_norms(dt), norms, rtol=0, atol=0)
    m.close()

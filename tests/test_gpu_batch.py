"""Batched solves (fi_batch_*) against the same models solved alone.

Every instance of a batch must end bitwise equal to its solve alone
(fi_solve): same iteration count, same residual norms, same fields.  The
packed launches only widen the single-model launches to many instances and
the per-point arithmetic does not depend on the tile shape, so equality is
exact, whatever the tile options, graph replay length or the time at which
an instance joins the batch.
"""
import numpy as np
import pytest

from helpers import state_names

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _manufactured(pkg, dims):
    if len(dims) == 3:
        kw, cad = pkg.manufactured_3d(dims)
    else:
        kw, cad = pkg.manufactured_2d(dims)
    m = pkg.IsothermalFullStokesModel(**kw)
    m.viscosity.eta.set(1.0)
    m.viscosity.eta_next.set(1.0)
    return m, cad["ncheck"], 400 * max(dims)


def _glacier(pkg, dims, seed, glen):
    m, cad = pkg.glacier_3d(dims, seed=seed, glen=glen)
    return m, cad["ncheck"], 400 * max(dims)


# (kind, dims, seed): a mix of immersed-boundary Glen and linear glaciers,
# plain 3D and 2D manufactured problems; sizes give partial fused tiles, and
# two instances too small for a fused interior (swept pointwise)
CASES = [
    ("glacier_glen", (40, 34, 18), 3),
    ("manufactured", (22, 19, 25), None),
    ("manufactured", (48, 40), None),
    ("glacier_linear", (30, 26, 14), 4),
    ("glacier_glen", (26, 30, 20), 5),
    ("glacier_linear", (7, 6, 5), 6),
    ("manufactured", (3, 5, 2), None),
]


def _make(pkg, case):
    kind, dims, seed = case
    if kind == "manufactured":
        return _manufactured(pkg, dims)
    return _glacier(pkg, dims, seed, kind == "glacier_glen")


def _alone(pkg, cases, tol):
    out = []
    for case in cases:
        m, ncheck, max_iter = _make(pkg, case)
        it, norms = m.solve(max_iter, ncheck, tol)
        out.append((it, norms, {n: m.field(n).parent() for n in state_names(m.grid.ndims) + ["ETA"]}))
        m.close()
    return out


def _check_equal(pkg, batch, models, cases, ref):
    for m, case, (it, norms, fields) in zip(models, cases, ref):
        state, bit, bnorms = batch.status(m)
        assert state in ("converged", "max_iter"), (case, state)
        assert bit == it, (case, bit, it)
        assert np.array_equal(bnorms, norms), (case, bnorms, norms)
        for n, f in fields.items():
            g = m.field(n).parent()
            assert np.array_equal(g, f), (case, n, float(np.max(np.abs(g - f))))


@pytest.fixture(scope="module")
def alone(pkg):
    _say("case: batch reference, %d models solved alone to 1e-8" % len(CASES))
    return _alone(pkg, CASES, 1e-8)


@pytest.mark.parametrize("opts", [dict(), dict(autotune=False, variant=2, zchunk=5, graph_iters=0),
                                  dict(autotune=False, variant=1, zchunk=16, graph_iters=6),
                                  dict(autotune=False, variant=4, zchunk=32, graph_iters=8, order=2),
                                  dict(autotune=False, variant=5, zchunk=9, graph_iters=2, order=1)])
def test_batch_equals_alone(pkg, alone, opts):
    _say("case: batch of %d mixed models, options %s" % (len(CASES), opts))
    models = []
    with pkg.Batch(**opts) as batch:
        for case in CASES:
            m, ncheck, max_iter = _make(pkg, case)
            batch.add(m, max_iter, ncheck, 1e-8)
            models.append(m)
        batch.solve()
        _say("batch info %s" % batch.info())
        _check_equal(pkg, batch, models, CASES, alone)
    for m in models:
        m.close()


def test_batch_continuous_admission(pkg, alone):
    """Models join while others run (the serving pattern): each still ends
    bitwise equal to its solve alone."""
    _say("case: continuous admission, %d models joining one by one" % len(CASES))
    models = [None] * len(CASES)
    with pkg.Batch() as batch:
        nxt = 0
        steps = 0
        while True:
            if nxt < len(CASES) and steps % 2 == 0:
                m, ncheck, max_iter = _make(pkg, CASES[nxt])
                batch.add(m, max_iter, ncheck, 1e-8)
                models[nxt] = m
                nxt += 1
            batch.step()
            steps += 1
            if nxt == len(CASES) and batch.info()["running"] == 0:
                break
        _check_equal(pkg, batch, models, CASES, alone)
        # a finished model leaves the batch and can be solved further alone
        m0 = models[0]
        batch.remove(m0)
        m0.advance(3)
        m0.synchronize()
    for m in models:
        m.close()


def test_batch_max_iter_and_destroy(pkg):
    _say("case: batch iteration cap and model destruction while attached")
    m1, ncheck, _ = _manufactured(pkg, (20, 20, 20))
    m2, _, _ = _manufactured(pkg, (24, 20))
    with pkg.Batch() as batch:
        batch.add(m1, 7, 5, 1e-30)
        batch.add(m2, 0, 5, 1e-30)
        batch.solve()
        s1 = batch.status(m1)
        s2 = batch.status(m2)
        assert s1[0] == "max_iter" and s1[1] == 7
        assert s2[0] == "max_iter" and s2[1] == 0
        m2.close()  # detaches from the batch
        assert batch.info()["attached"] == 1
    m1.close()


REFINED_CASES = [
    ("glacier_glen", (40, 34, 18), 3),
    ("manufactured", (22, 19, 25), None),
    ("glacier_2d", (64, 32), 7),
    ("glacier_linear", (56, 40, 24), 2),
    ("glacier_linear", (7, 6, 5), 6),
]


def _make_refined(pkg, case):
    kind, dims, seed = case
    if kind == "glacier_2d":
        m, _ = pkg.glacier_2d(dims, seed=seed, glen=True)
        return m
    m, _, _ = _make(pkg, case)
    return m


@pytest.mark.parametrize("mode", [1, 0])
def test_batch_refined_equals_alone(pkg, mode):
    """FP32-refined solves packed in one batch end bitwise equal to the same
    refined solves alone: outer steps, FP32 sweeps, residual norms, fields.
    Lagged PT (mode 1) with a common chunk of 32 sweeps; correction solves
    (mode 0) with chunks of 60 sweeps between inner checks."""
    _say("case: refined batch (mode %d) of %d mixed models vs alone" % (mode, len(REFINED_CASES)))
    if mode == 1:
        args = dict(max_outer=400, ncheck=32)
    else:
        args = dict(max_outer=30, max_inner=3000, ncheck=60, inner_rtol=3e-2)

    def make(case):
        m = _make_refined(pkg, case)
        m.set_option("refine_mode", mode)
        return m

    alone = []
    for case in REFINED_CASES:
        m = make(case)
        outer, sweeps, norms = m.refined_solve(1e-8, **args)
        _say("alone %s: outer %d sweeps %d residual %.2e" % (case, outer, sweeps, max(norms)))
        alone.append((outer, sweeps, dict(m.refined_counts), norms.copy(),
                      {n: m.field(n).parent() for n in state_names(m.grid.ndims) + ["ETA"]}))
        m.close()
    models = []
    with pkg.Batch() as batch:
        for case in REFINED_CASES:
            m = make(case)
            batch.add_refined(m, 1e-8, **args)
            models.append(m)
        steps = 0
        while batch.info()["running"] > 0:
            batch.step()
            steps += 1
            if steps % 50 == 0:
                _say("batch step %d, running %d" % (steps, batch.info()["running"]))
            assert steps < 3000
        _say("batch done after %d steps" % steps)
        for m, case, (outer, sweeps, counts, norms, fields) in zip(models, REFINED_CASES, alone):
            state, it, bnorms = batch.status(m)
            bc = batch.refined_counts(m)
            _say("%s: %s outer %d sweeps %d (alone %d %d)" % (case, state, bc["outer"], bc["fp32_iterations"], outer, sweeps))
            assert state in ("converged", "max_iter"), (case, state)
            assert bc["outer"] == outer and bc["fp32_iterations"] == sweeps, (case, bc, outer, sweeps)
            assert bc["nonlinear_phase"] == counts["nonlinear_phase"], case
            assert np.array_equal(bnorms, norms), (case, bnorms, norms)
            for n, f in fields.items():
                assert np.array_equal(m.field(n).parent(), f), (case, n)
    for m in models:
        m.close()

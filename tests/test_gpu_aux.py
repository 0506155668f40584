"""Failure isolation and snapshot/restore on the GPU path.

  * A non-finite state in one batched model (the reference aborts its solve
    with error("simulation failed"), scripts_future_API/tm_stokes_chmy.jl:
    104-106) fails that model alone: it reports FI_BATCH_FAILED at its first
    residual check, and every other model of the batch, FP64 or FP32-refined,
    ends bitwise equal to its solve alone.
  * A solve interrupted, its state downloaded and uploaded into a fresh model
    (set!, interior/parent reads: tm_stokes_chmy.jl:86-91,110-113), then
    continued, ends bitwise equal to the uninterrupted solve.
"""
import numpy as np
import pytest

from helpers import state_names

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


CASES = [((40, 34, 18), 3, True), ((30, 26, 14), 4, False), ((26, 30, 20), 5, True)]


def _glacier(pkg, dims, seed, glen):
    m, cad = pkg.glacier_3d(dims, seed=seed, glen=glen)
    return m, cad["ncheck"], 400 * max(dims)


def _fields(m):
    return {n: m.field(n).parent() for n in state_names(m.grid.ndims) + ["ETA"]}


@pytest.mark.parametrize("refined", [False, True])
def test_nonfinite_model_fails_alone(pkg, refined):
    _say("case: NaN injected into one of %d batched glaciers, refined %s" % (len(CASES) + 1, refined))
    ref = []
    for dims, seed, glen in CASES:
        m, ncheck, max_iter = _glacier(pkg, dims, seed, glen)
        if refined:
            outer, sweeps, norms = m.refined_solve(1e-8)
            ref.append(((outer, sweeps), norms, _fields(m)))
        else:
            it, norms = m.solve(max_iter, ncheck, 1e-8)
            ref.append((it, norms, _fields(m)))
        m.close()
    models = []
    with pkg.Batch() as batch:
        bad, ncheck_bad, max_bad = _glacier(pkg, (28, 24, 14), 9, True)
        p = bad.field("P").parent()
        p[10, 10, 7] = np.nan
        bad.field("P").set(p)
        entries = [(bad, ncheck_bad, max_bad)]
        for dims, seed, glen in CASES:
            m, ncheck, max_iter = _glacier(pkg, dims, seed, glen)
            models.append(m)
            entries.append((m, ncheck, max_iter))
        for m, ncheck, max_iter in entries:
            if refined:
                batch.add_refined(m, 1e-8)
            else:
                batch.add(m, max_iter, ncheck, 1e-8)
        batch.solve()
        state, _, norms = batch.status(bad)
        assert state == "failed", state
        for m, (r_it, r_norms, r_fields), case in zip(models, ref, CASES):
            state, it, norms = batch.status(m)
            assert state == ("converged" if np.max(r_norms) < 1e-8 else "max_iter"), (case, state)
            if refined:
                c = batch.refined_counts(m)
                assert (c["outer"], c["fp32_iterations"]) == r_it, (case, c, r_it)
            else:
                assert it == r_it, (case, it, r_it)
            assert np.array_equal(norms, r_norms), case
            for n, f in r_fields.items():
                assert np.array_equal(m.field(n).parent(), f), (case, n)
    for m in [bad] + models:
        m.close()


@pytest.mark.parametrize("fused", [1, 0])
def test_snapshot_restore_bitwise(pkg, fused):
    _say("case: snapshot after 137 iterations, restore into a fresh model, 300 more, fused %d" % fused)
    dims, seed = (40, 34, 18), 3
    full, _ = pkg.glacier_3d(dims, seed=seed, glen=True)
    part, _ = pkg.glacier_3d(dims, seed=seed, glen=True)
    fresh, _ = pkg.glacier_3d(dims, seed=seed, glen=True)
    for m in (full, part, fresh):
        m.set_option("fused", fused)
    full.advance(437)
    part.advance(137)
    names = state_names(3) + ["ETA", "ETA_NEXT"]
    snap = {n: part.field(n).parent() for n in names}
    part.close()
    for n, a in snap.items():
        fresh.field(n).set(a)
    fresh.advance(300)
    for n in names:
        assert np.array_equal(fresh.field(n).parent(), full.field(n).parent()), n
    assert np.array_equal(fresh.compute_residuals(), full.compute_residuals())
    full.close()
    fresh.close()

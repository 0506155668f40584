"""The batched workload at its configured size: 96 seeded glaciers with
uneven grids, 48..192 cells per axis (aspect ratios 1 to 4), Glen n = 3,
FP32 sweeps inside FP64 defect correction, packed into one batch.  Every
instance must reach FP64 residual 1e-8 and end bitwise equal to the same
refined solve run alone (fi_refined_solve): same outer steps, same FP32
sweeps, same norms, same fields (compared through SHA-1 digests of every
field, to keep host memory small).  About two and a half minutes on one
MI355X (the batch, then the 96 solves one at a time)."""
import hashlib
import time

import numpy as np
import pytest

from helpers import state_names

pytestmark = pytest.mark.gpu

CHUNK = 32


def _digest(m):
    return {n: hashlib.sha1(m.field(n).parent().tobytes()).hexdigest() for n in state_names(3) + ["ETA"]}


def _args(d):
    return dict(max_outer=-(-400 * max(d) // CHUNK), ncheck=CHUNK)


def test_batch_of_96_glaciers_equals_alone(pkg):
    print("case: 96 glaciers 48-192 per axis, FP32-refined batch vs alone", flush=True)
    specs = pkg.batch_glaciers(96, 48, 192)
    t0 = time.perf_counter()
    models = []
    with pkg.Batch() as batch:
        for d, geo in specs:
            m, _ = pkg.glacier_3d(d, glen=True, **geo)
            models.append(m)
            batch.add_refined(m, 1e-8, **_args(d))
        batch.solve()
        t_batch = time.perf_counter() - t0
        got = []
        for m in models:
            state, it, norms = batch.status(m)
            c = batch.refined_counts(m)
            got.append((state, c["outer"], c["fp32_iterations"], norms, _digest(m)))
    for m in models:
        m.close()
    print("batch: %.1f s, %d converged" % (t_batch, sum(g[0] == "converged" for g in got)), flush=True)
    t0 = time.perf_counter()
    for q, ((d, geo), (state, outer, sweeps, norms, dig)) in enumerate(zip(specs, got)):
        assert state == "converged", (q, d, state)
        assert np.max(norms) < 1e-8, (q, d, norms)
        m, _ = pkg.glacier_3d(d, glen=True, **geo)
        o, s, n = m.refined_solve(1e-8, **_args(d))
        assert (o, s) == (outer, sweeps), (q, d, (o, s), (outer, sweeps))
        assert np.array_equal(n, norms), (q, d)
        assert _digest(m) == dig, (q, d)
        m.close()
        print("alone %d %s: %d outer steps, %d sweeps, equal; %.1f s so far" % (q, d, o, s, time.perf_counter() - t0),
              flush=True)
    print("alone: %.1f s" % (time.perf_counter() - t0), flush=True)

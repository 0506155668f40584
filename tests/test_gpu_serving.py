"""Serving loop: requests solved in the mix equal the same requests solved
alone (bitwise), and admission control refuses what cannot fit."""
import numpy as np
import pytest

from helpers import state_names

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _small_requests(pkg):
    # a deterministic mix of small 2D manufactured and 3D glacier requests
    return pkg.poisson_requests(8, 200.0, seed=11, range2d=(24, 72), range3d=(28, 40))


def test_serving_mix_equals_alone(pkg):
    reqs = _small_requests(pkg)
    _say("case: serving %d requests %s, continuous batching, deterministic admission" % (len(reqs), reqs))
    srv = pkg.Server(pkg, max_slots=5, keep_models=True)
    try:
        summary = srv.run(reqs, realtime=False)
        _say("summary %s" % summary)
        assert summary["requests"] == len(reqs)
        _say("states %s" % [(r["dims"], r["state"], r["iterations"]) for r in srv.results.values()])
        for rid, rec in srv.results.items():
            assert rec["state"] in ("converged", "max_iter"), rec
            req = reqs[rid]
            m = rec["model"]
            alone = pkg.serving.build_model(pkg, req)
            _, _, ncheck, max_iter = pkg.serving.request_spec(req)
            it, norms = alone.solve(max_iter, ncheck, srv.tol)
            assert it == rec["iterations"], (req, it, rec["iterations"])
            assert float(np.max(norms)) == rec["residual"]
            for n in state_names(req.ndims) + ["ETA"]:
                assert np.array_equal(m.field(n).parent(), alone.field(n).parent()), (req, n)
            alone.close()
    finally:
        srv.close()


def test_serving_realtime_and_admission(pkg):
    reqs = pkg.poisson_requests(6, 100.0, seed=5, range2d=(24, 48), range3d=(28, 36))
    _say("case: realtime serving of %d small requests, 2 slots" % len(reqs))
    srv = pkg.Server(pkg, max_slots=2)
    try:
        s = srv.run(reqs, realtime=True)
        _say("summary %s" % s)
        assert all(r["state"] in ("converged", "max_iter") for r in srv.results.values())
        assert s["latency_s"]["p50"] <= s["latency_s"]["p99"] <= s["makespan_s"]
        for r in srv.results.values():
            assert r["t_admit"] >= r["t_arrival"] - 1e-9 and r["t_done"] >= r["t_admit"]
    finally:
        srv.close()
    # a reserve larger than the device: nothing can be admitted
    free, total = pkg.mem_info()
    srv = pkg.Server(pkg, max_slots=2, hbm_reserve=2 * total)
    try:
        with pytest.raises(RuntimeError, match="does not fit"):
            srv.run(reqs[:1], realtime=False)
    finally:
        srv.close()


def test_serving_refined_mix_equals_alone(pkg):
    """Refined serving: every request converges to the tolerance, and each
    result equals the same refined solve alone (outer steps, sweeps, fields)."""
    reqs = pkg.poisson_requests(6, 200.0, seed=13, range2d=(24, 64), range3d=(28, 40))
    _say("case: refined serving of %d requests %s" % (len(reqs), reqs))
    srv = pkg.Server(pkg, max_slots=4, keep_models=True, precision="refined")
    try:
        summary = srv.run(reqs, realtime=False)
        _say("summary %s" % summary)
        assert summary["converged"] == len(reqs)
        for rid, rec in srv.results.items():
            req = reqs[rid]
            m = rec["model"]
            alone = pkg.serving.build_model(pkg, req)
            outer, sweeps, norms = alone.refined_solve(srv.tol)
            assert sweeps == rec["iterations"], (req, sweeps, rec["iterations"])
            assert float(np.max(norms)) == rec["residual"]
            for n in state_names(req.ndims) + ["ETA"]:
                assert np.array_equal(m.field(n).parent(), alone.field(n).parent()), (req, n)
            alone.close()
    finally:
        srv.close()


def test_serving_configured_sizes_mix_equals_alone(pkg):
    """Requests across the served size range (64 to 511 per 2D section, 64 to
    383 per 3D glacier: a 511x511x256 glacier does not reach 1e-8 and would
    run to its 100 n cap twice), solved in one mix to FP64 residual 1e-8:
    the large ones converge (a small glacier may stop at its 100 n cap, as it
    does alone), and every request ends bitwise equal to its solve alone."""
    R = pkg.serving.Request
    reqs = [R(0, 0.0, 2, (64, 32), 101), R(1, 0.0, 3, (383, 383, 192), 102), R(2, 0.01, 2, (511, 256), 103),
            R(3, 0.02, 3, (64, 64, 32), 104), R(4, 0.03, 3, (160, 160, 80), 105), R(5, 0.04, 2, (230, 115), 106)]
    _say("case: serving %d requests across the configured sizes %s" % (len(reqs), reqs))
    srv = pkg.Server(pkg, max_slots=6, keep_models=True)
    try:
        summary = srv.run(reqs, realtime=False)
        _say("summary %s" % summary)
        states = {rid: rec["state"] for rid, rec in srv.results.items()}
        assert all(v in ("converged", "max_iter") for v in states.values()), states
        assert states[1] == "converged" and states[2] == "converged", states
        for rid, rec in srv.results.items():
            req = reqs[rid]
            m = rec["model"]
            alone = pkg.serving.build_model(pkg, req)
            _, _, ncheck, max_iter = pkg.serving.request_spec(req)
            it, norms = alone.solve(max_iter, ncheck, srv.tol)
            _say("request %s: %d iterations in the mix, %d alone" % (req, rec["iterations"], it))
            assert it == rec["iterations"], (req, it, rec["iterations"])
            assert float(np.max(norms)) == rec["residual"]
            for n in state_names(req.ndims) + ["ETA"]:
                assert np.array_equal(m.field(n).parent(), alone.field(n).parent()), (req, n)
            alone.close()
    finally:
        srv.close()

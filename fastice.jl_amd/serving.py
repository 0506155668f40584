"""Serving loop: solve requests as they arrive, continuously batched.

No counterpart in the reference (its drivers solve one model per run,
scripts_future_API/tm_stokes_chmy.jl:96-108).  Requests arrive as a Poisson
process; each is a seeded synthetic glacier, 3D (test_vavilov.jl set-up) or
a 2D section of one (polygon level sets), solved to a residual tolerance:
in FP64 with the reference's glacier check cadence (ncheck = 5n) up to 100n
iterations, or
(precision="refined") by FP32 sweeps inside FP64 defect correction (lagged
PT, an outer step every 32 sweeps, at most 400n sweeps).  The server

  * admits a waiting request when a batch slot is free and the device has the
    model's bytes (fi_model_bytes) plus a reserve free (hipMemGetInfo),
  * builds its model (fields, level sets, volume fractions) and attaches it to
    the running Batch, which advances all running models together,
  * records its completion at the batch step in which it converges, then frees
    its model (its HBM is available to the next admission).

A model's result does not depend on what else is in the batch: every model
ends bitwise equal to its solve alone.
"""
import math
import time

import numpy as np

from .batch import Batch, model_bytes
from .problems import glacier_2d, glacier_2d_spec, glacier_3d, glacier_3d_spec


class Request:
    __slots__ = ("rid", "t_arrival", "ndims", "dims", "seed")

    def __init__(self, rid, t_arrival, ndims, dims, seed):
        self.rid, self.t_arrival, self.ndims, self.dims, self.seed = rid, t_arrival, ndims, tuple(dims), seed

    def __repr__(self):
        return "Request(%d, t=%.3f, %s)" % (self.rid, self.t_arrival, "x".join(map(str, self.dims)))


MAX_ITER_FACTOR = 100  # FP64 requests stop at 100 n iterations if not converged


def poisson_requests(n, rate, seed=0, p3d=0.5, range2d=(64, 511), range3d=(64, 511)):
    """n requests with exponential inter-arrival times (rate per second); 3D
    with probability p3d; sizes log-uniform in range2d (n x ceil(n/2)) or
    range3d (n x n x ceil(n/2))."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        three = rng.random() < p3d
        lo, hi = range3d if three else range2d
        s = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        dims = (s, s, (s + 1) // 2) if three else (s, (s + 1) // 2)
        out.append(Request(i, float(t[i]), 3 if three else 2, dims, int(rng.integers(1, 2 ** 31))))
    return out


def request_spec(req, iter_factor=MAX_ITER_FACTOR):
    """(model arguments, immersed boundary, ncheck, max_iter) of a request:
    the glacier driver's check cadence ncheck = 5 n (test_vavilov.jl:146),
    and an iteration cap of iter_factor n (the driver stops at 50 n,
    test_vavilov.jl:145; the glaciers here reach 1e-8 in about 35 n)."""
    spec = glacier_3d_spec if req.ndims == 3 else glacier_2d_spec
    kw, cad = spec(req.dims, seed=req.seed, glen=True)
    return kw, True, cad["ncheck"], iter_factor * max(req.dims)


def build_model(pkg, req):
    """The request's model with its initial state (on its own stream)."""
    build = glacier_3d if req.ndims == 3 else glacier_2d
    m, _ = build(req.dims, seed=req.seed, glen=True)
    return m


def percentiles(x, qs=(50, 95, 99)):
    x = np.asarray(x, dtype=float)
    if x.size == 0:
        return {("p%d" % q): None for q in qs}
    return {("p%d" % q): float(np.percentile(x, q)) for q in qs}


class Server:
    """Continuous-batching solver service over one device.

    max_slots: models attached to the batch at once; hbm_reserve: bytes kept
    free beyond the admitted models; tol: residual tolerance of every solve;
    keep_models: keep finished models open (tests read their fields);
    precision: "fp64" or "refined" (a refined model also holds its FP32
    workspace, counted here as half the FP64 model's bytes on top).
    """

    def __init__(self, pkg, max_slots=64, hbm_reserve=4 << 30, tol=1e-8, keep_models=False, batch_options=None,
                 precision="fp64"):
        if precision not in ("fp64", "refined"):
            raise ValueError("precision must be 'fp64' or 'refined'")
        self.precision = precision
        self.pkg = pkg
        self.max_slots = max_slots
        self.hbm_reserve = hbm_reserve
        self.tol = tol
        self.keep_models = keep_models
        self.batch = Batch(**(batch_options or {}))
        self.models = {}
        self.results = {}

    def _fits(self, req):
        kw, ib, _, _ = request_spec(req)
        need = model_bytes(immersed_boundary=ib, **kw)
        if self.precision == "refined":
            need += need // 2
        free, _ = self.pkg.mem_info()
        return free >= need + self.hbm_reserve

    def _admit(self, req, t0):
        ta = time.perf_counter() - t0
        _, _, ncheck, max_iter = request_spec(req)
        m = build_model(self.pkg, req)
        if self.precision == "refined":
            self.batch.add_refined(m, self.tol)
        else:
            self.batch.add(m, max_iter, ncheck, self.tol)
        self.models[req.rid] = (req, m)
        self.results[req.rid] = dict(rid=req.rid, ndims=req.ndims, dims=list(req.dims), t_arrival=req.t_arrival,
                                     t_admit=ta, setup_s=time.perf_counter() - t0 - ta)

    def _collect(self, t0):
        done = []
        for rid, (req, m) in self.models.items():
            state, iters, norms = self.batch.status(m)
            if state == "running":
                continue
            r = self.results[rid]
            r.update(t_done=time.perf_counter() - t0, state=state, iterations=iters,
                     residual=float(np.max(norms)))
            r["latency_s"] = r["t_done"] - r["t_arrival"]
            done.append(rid)
        for rid in done:
            req, m = self.models.pop(rid)
            self.batch.remove(m)
            if self.keep_models:
                self.results[rid]["model"] = m
            else:
                m.close()
        return len(done)

    def run(self, requests, realtime=True, log=None, offline=False):
        """Serve requests (sorted by arrival).  realtime: requests become
        eligible at their arrival time; otherwise one becomes eligible per
        batch step (a deterministic interleaving for tests); offline: all are
        eligible at once (the offline batched throughput of the same mix).
        Returns the summary; per-request records are in self.results."""
        reqs = sorted(requests, key=lambda r: r.t_arrival)
        if offline:
            reqs = [Request(r.rid, 0.0, r.ndims, r.dims, r.seed) for r in reqs]
        t0 = time.perf_counter()
        nxt = 0
        waiting = []
        steps = 0
        while True:
            now = time.perf_counter() - t0
            if realtime or offline:
                while nxt < len(reqs) and reqs[nxt].t_arrival <= now:
                    waiting.append(reqs[nxt])
                    nxt += 1
            elif nxt < len(reqs):
                r = reqs[nxt]
                r.t_arrival = now
                waiting.append(r)
                nxt += 1
            # admission: free slot and enough free HBM, first come first served
            while waiting and len(self.models) < self.max_slots and self._fits(waiting[0]):
                self._admit(waiting.pop(0), t0)
            if self.models:
                self.batch.step()
                steps += 1
                n = self._collect(t0)
                if log and n:
                    log("t=%.2f s done %d/%d running %d waiting %d" % (time.perf_counter() - t0, len(self.results) -
                        len(self.models), len(reqs), len(self.models), len(waiting)))
            elif waiting:
                raise RuntimeError("request %r does not fit in device memory" % (waiting[0],))
            elif nxt < len(reqs):
                time.sleep(max(0.0, reqs[nxt].t_arrival - (time.perf_counter() - t0)))
            else:
                break
        makespan = time.perf_counter() - t0
        recs = [r for r in self.results.values()]
        lat = [r["latency_s"] for r in recs]
        conv = sum(r["state"] == "converged" for r in recs)
        # solves/s counts the requests that reached the tolerance; those that
        # hit the iteration cap or failed are reported beside it
        return dict(requests=len(recs), makespan_s=makespan, solves_per_s=conv / makespan if makespan else 0.0,
                    requests_per_s=len(recs) / makespan if makespan else 0.0,
                    latency_s=percentiles(lat),
                    latency_s_converged=percentiles([r["latency_s"] for r in recs if r["state"] == "converged"]),
                    batch_steps=steps, converged=conv,
                    max_iter=sum(r["state"] == "max_iter" for r in recs),
                    failed=sum(r["state"] == "failed" for r in recs),
                    setup_s_mean=float(np.mean([r["setup_s"] for r in recs])) if recs else 0.0)

    def close(self):
        for _, m in self.models.values():
            m.close()
        self.models = {}
        if self.keep_models:
            for r in self.results.values():
                if "model" in r:
                    r["model"].close()
        self.batch.close()

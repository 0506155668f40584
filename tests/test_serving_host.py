"""Host-side logic of the batch / serving layer (no GPU needed): the request
generator, the latency percentiles and the device-byte estimate used by
admission control."""
import ctypes as C
import math

import numpy as np


def test_poisson_requests_deterministic_and_in_range(pkg):
    a = pkg.poisson_requests(400, 40.0, seed=7)
    b = pkg.poisson_requests(400, 40.0, seed=7)
    assert [(r.t_arrival, r.dims, r.seed) for r in a] == [(r.t_arrival, r.dims, r.seed) for r in b]
    t = np.array([r.t_arrival for r in a])
    assert np.all(np.diff(t) > 0)
    # mean inter-arrival 1/rate: 400 exponential samples, 5 standard errors
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 40.0) < 5 * (1 / 40.0) / math.sqrt(400)
    n3 = sum(r.ndims == 3 for r in a)
    assert abs(n3 / 400 - 0.5) < 5 * math.sqrt(0.25 / 400)
    for r in a:
        if r.ndims == 3:
            s = r.dims[0]
            assert 64 <= s <= 511 and r.dims == (s, s, (s + 1) // 2)
        else:
            s = r.dims[0]
            assert 64 <= s <= 511 and r.dims == (s, (s + 1) // 2)


def test_percentiles(pkg):
    x = np.arange(1, 101, dtype=float)
    p = pkg.percentiles(x)
    assert p == {"p50": 50.5, "p95": 95.05, "p99": 99.01}
    assert pkg.percentiles([]) == {"p50": None, "p95": None, "p99": None}


def _elems(dims):
    # padded layout: rows of n0+3 elements rounded up to 32, 31 leading
    # elements, 32 trailing
    n = list(dims)
    sy = (n[0] + 3 + 31) // 32 * 32
    rows = (n[1] + 3) * (n[2] + 3) if len(n) == 3 else n[1] + 3
    return sy, 31 + sy * rows + 32


def _pair(elems):
    # a double-buffered field: both buffers in one allocation, the second
    # starting a multiple of 64 elements after the first
    return 2 * ((elems + 63) // 64 * 64)


def _expected_bytes(dims, nfields_single, immersed):
    # the ten stress/velocity fields are double buffered
    n = list(dims)
    sy, elems = _elems(dims)
    b = nfields_single * elems * 8 + 10 * _pair(elems) * 8 + 32
    if immersed:
        b += elems
        if len(n) == 3:
            b += 2 * sy * (n[1] + 3) * 8  # per-column plane ranges (two int2 per column)
            # the z range each tile of an FP64 fused sweep marched (int2 per tile), and three scalars
            b += (((n[0] + 61) // 62 + 2) * ((n[1] + 1) // 2 + 1) * (n[2] + 1) + 3) * 8
    return b


def test_model_bytes_matches_layout(pkg):
    kw, _ = pkg.manufactured_3d((30, 20, 10))
    # P tau(6) V(3) double buffered; eta eta_next rP rV(3) gravity fields(3)
    assert pkg.model_bytes(**kw) == _expected_bytes((30, 20, 10), 2 + 4 + 3, False)
    kw, _ = pkg.manufactured_2d((50, 40))
    # P tau(3) V(2) double buffered; eta eta_next rP rV(2) gravity fields(2)
    _, elems = _elems((50, 40))
    assert pkg.model_bytes(**kw) == 6 * _pair(elems) * 8 + (2 + 3 + 2) * elems * 8 + 32
    kw, _ = pkg.glacier_3d_spec((40, 30, 20), seed=3)
    # + psi(2) + omega na/ns (8 each), constant gravity, no rheology field
    assert pkg.model_bytes(immersed_boundary=True, **kw) == _expected_bytes((40, 30, 20), 2 + 4 + 2 + 16, True)

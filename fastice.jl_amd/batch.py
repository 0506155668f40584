"""Batched solves over libfasticehip.so (fi_batch_* in include/fastice_hip.h).

No counterpart in the reference: FastIce.jl solves one model at a time in the
driver loop (scripts_future_API/tm_stokes_chmy.jl:96-108).  A Batch advances
every attached FP64 model with packed launches (one fused-sweep launch over
the tiles of all instances, one launch per pointwise shell kind, one per
boundary dimension) replayed as HIP graphs; each model keeps the semantics of
IsothermalFullStokesModel.solve exactly, so its result is bitwise the one it
gets alone.  A batch is either all FP64 (add) or all FP32-refined
(add_refined): refined models share the packed FP32 sweeps, and each runs
its own outer defect-correction steps between chunks, so it also ends
bitwise equal to IsothermalFullStokesModel.refined_solve alone.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from .full_stokes import build_desc, refined_defaults

STATES = {A.FI_BATCH_RUNNING: "running", A.FI_BATCH_CONVERGED: "converged", A.FI_BATCH_MAX_ITER: "max_iter",
          A.FI_BATCH_FAILED: "failed"}


def model_bytes(grid, boundary_conditions, gravity, rheology, solver_params, immersed_boundary=False):
    """Device bytes a model of this description allocates (admission control)."""
    desc, _, _ = build_desc(grid, boundary_conditions, gravity, rheology, solver_params,
                            immersed_boundary=immersed_boundary)
    return int(A.lib().fi_model_bytes(C.byref(desc)))


class Batch:
    """A set of models solved together.

    add(model, max_iter, ncheck, tol) attaches a model and starts its solve;
    step() advances the running models to the next residual check of the
    nearest one and returns how many finished; solve() steps until none runs;
    status(model) gives (state, iterations, Inf-norms).
    """

    def __init__(self, autotune=True, variant=None, zchunk=None, graph_iters=None, order=None):
        h = C.c_void_p()
        A.check(A.lib().fi_batch_create(C.byref(h)))
        self._h = h
        self._models = {}
        self.set_option(A.FI_BOPT_AUTOTUNE, 1 if autotune else 0)
        if variant is not None:
            self.set_option(A.FI_BOPT_VARIANT, variant)
        if zchunk is not None:
            self.set_option(A.FI_BOPT_ZCHUNK, zchunk)
        if graph_iters is not None:
            self.set_option(A.FI_BOPT_GRAPH_ITERS, graph_iters)
        if order is not None:
            self.set_option(A.FI_BOPT_TILE_ORDER, order)

    def set_option(self, key, value):
        A.check(A.lib().fi_batch_set_option(self._h, int(key), int(value)))

    def add(self, model, max_iter, ncheck, tol):
        A.check(A.lib().fi_batch_add(self._h, model._h, int(max_iter), int(ncheck), float(tol)))
        self._models[id(model)] = model

    def add_refined(self, model, tol, max_outer=None, max_inner=None, ncheck=None, inner_rtol=3e-2):
        """Attach a model and start its FP32-refined solve (the arguments and
        defaults of IsothermalFullStokesModel.refined_solve)."""
        n = max(model.grid.dims)
        max_outer, max_inner, ncheck = refined_defaults(n, model._refine_mode, max_outer, max_inner, ncheck)
        A.check(A.lib().fi_batch_add_refined(self._h, model._h, int(max_outer), int(max_inner), int(ncheck),
                                             float(inner_rtol), float(tol)))
        self._models[id(model)] = model

    def refined_counts(self, model):
        """{outer, fp32_iterations, nonlinear_phase} of a refined solve."""
        out = (C.c_int * 3)()
        A.check(A.lib().fi_batch_refined_counts(self._h, model._h, out))
        return dict(outer=out[0], fp32_iterations=out[1], nonlinear_phase=out[2])

    def remove(self, model):
        A.check(A.lib().fi_batch_remove(self._h, model._h))
        self._models.pop(id(model), None)

    def step(self):
        n = C.c_int(0)
        A.check(A.lib().fi_batch_step(self._h, C.byref(n)))
        return n.value

    def solve(self):
        A.check(A.lib().fi_batch_solve(self._h))

    def status(self, model):
        st, it = C.c_int(0), C.c_int(0)
        norms = np.zeros(4)
        A.check(A.lib().fi_batch_status(self._h, model._h, C.byref(st), C.byref(it), A.dptr(norms)))
        return STATES[st.value], it.value, norms

    def info(self):
        out = (C.c_int * 8)()
        A.check(A.lib().fi_batch_info(self._h, out))
        keys = ("attached", "running", "fused_items", "variant", "zchunk", "graph_iters", "autotune_runs",
                "rebuilds")
        d = dict(zip(keys, list(out)))
        v = d["variant"]  # (tile lanes in x, tile rows, min waves per SIMD)
        d["variant"] = (v // 16 // 32, v // 16 % 32, v % 16)
        d["zchunk"], d["tile_order"] = d["zchunk"] // 4, d["zchunk"] % 4
        return d

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            A.lib().fi_batch_destroy(h)
            self._h = C.c_void_p()
        self._models = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

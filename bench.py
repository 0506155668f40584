"""PT-Stokes benchmark: T_eff [GB/s] of the PT iteration on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu] [--no-solve] [--no-compare]

Workloads
  glacier_511x511x255 (default): 3D synthetic glacier, 511x511x255 cells,
      seeded 8-Gaussian-bump bed under a parabolic surface, level sets ->
      volume fractions (immersed boundary), Glen n = 3 regularised power law,
      gravity-driven.  A step is one cycle of the FP32-refined solver (32
      FP32 PT sweeps, then its FP64 outer step); the FP64 iteration of the
      same model is timed beside it, and a converged-solve comparison
      (FP32-refined vs plain FP64, both to FP64 residual 1e-8) runs on the
      same glacier at 255x255x127, regularised at strain rate 1e-6.
  stokes3d_fp64_511x511x255: isoviscous free-slip box with manufactured
      forcing, plain FP64 iteration.
  batch96_glacier: 96 seeded glaciers, each dimension uniform in 48..192
      (cubic cells, domain aspect 1 to 4), Glen n = 3, solved as one batch
      with packed launches, FP32 sweeps
      inside FP64 defect correction to FP64 residual 1e-8 (--batch-precision
      fp64: plain FP64 PT with HIP graphs); value = solves/s; the same models
      are then reset and a sample of them solved one at a time for the
      comparison.
  heat_diffusion: PT heat diffusion (Models/Diffusion/Heat) at 1023^3, FP64;
      value = T_eff by benchmark_diffusion_3D.jl's formula.
  h5_glacier_255x255x127: HDF5 + XDMF3 output (src/Writers.jl) of the fields
      the glacier driver saves, from device memory; value = GB/s of data.
  manufactured3d_255: 3D isoviscous free-slip manufactured solution, 255^3,
      FP64 (scripts_future_API/stokes_manufactured_solution_free_slip_3D.jl):
      2000 PT iterations timed, then a converged solve to residual 1e-8
      (ncheck 255) with its error norms against the closed form; value =
      T_eff of the iteration; the CPU oracle timed on a sample of iterations.
  refined_solve_255x255x127, fp64_solve_255x255x127: one of the two solves
      of the default line's converged-solve comparison, alone; value = seconds.
  converged_511x511x255: the 511 glacier solved by plain FP64 and by the
      FP32-refined solver until r_P and r_V outside its 16 margin faces
      (whose residual no iteration changes) are below 1e-8; value = the
      refined speedup, with the per-field difference.
  serving_poisson40: Poisson arrivals at 40/s, half 2D glacier sections
      (n x n/2, n in 64..511), half 3D glaciers (n x n x n/2, n in 64..511,
      --serving-3d), FP64 to 1e-8 or 100n iterations, continuously batched;
      value = solves/s of the requests that reached 1e-8, with p50/p95/p99
      latency and the counts that hit the cap or failed.

value = T_eff over the cells a step processes: 23 * sizeof(T) bytes per
processed cell and PT sweep (the reference's per-cell count,
scripts_future_API/tm_stokes_mpi_wip.jl:180) over the step's wall time,
inputs resident in HBM.  The glacier step is one cycle of the FP32-refined
solver (32 FP32 sweeps plus the FP64 outer step); processed cells are the
boundary shell plus the interior cells on planes of their column that hold a
valid location (the immersed-boundary sweep reads and writes nothing on the
other planes, whose state stays 0).  The reference's own formula, which
counts every cell of the box, is reported beside it as
teff_reference_formula_gbs.  roofline: the dominant kernel's algorithmic
bytes per launch (23 * sizeof(T) per processed interior cell; masks and the
shell points it also updates not credited) over its HIP-event duration on the
model's stream; traffic: its measured HBM bytes per launch from the rocprofv3
summary in profiles/.  cpu_baseline: the CPU oracle (OpenMP) on a bounded
number of FP64 iterations of the same grid, which it sweeps cell by cell.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

import __graft_entry__ as ge

HBM_PEAK_GBS = 8000.0
CYCLE = 32  # FP32 sweeps between FP64 outer steps of the refined solver (its default)
# Glen's law regularised at strain rate 1e-6 (eta_reg = 5000), with the PT
# parameters and log relaxation 0.01 that converge at that contrast
EPS0 = 1e-6
# The timed 511x511x255 model keeps eta_reg = 30 with the driver's PT
# parameters and log relaxation 0.01: with the strain-rate-1e-6
# regularisation its refined cycles diverged after about 50 outer steps
RELAX_511 = 0.01
# The converged comparison's glacier: on 511x511x255 neither plain FP64 PT nor
# the refined solver reaches residual 1e-8 (16 ice-margin velocity faces,
# valid by the reference's null-space rule but with an na fraction of 1.2e-6,
# keep a residual of 0.5 that no iteration changes)
SOLVE_DIMS = (255, 255, 127)
DEFAULT = "glacier_511x511x255"
WORKLOADS = {
    "glacier_511x511x255": (511, 511, 255),
    "stokes3d_fp64_511x511x255": (511, 511, 255),
    "batch96_glacier": None,
    "serving_poisson40": None,
    "h5_glacier_255x255x127": None,
    "heat_diffusion": None,
    "manufactured3d_255": None,
    "refined_solve_255x255x127": None,
    "converged_511x511x255": None,
    "fp64_solve_255x255x127": None,
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline")
    p.add_argument("--no-solve", action="store_true",
                   help="skip the plain FP64 converged solve (the refined solve runs: the timed cycles continue it)")
    p.add_argument("--no-compare", action="store_true",
                   help="skip the converged-solve comparison of the glacier line (profiling runs)")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--alone-sample", type=int, default=16,
                   help="batch96_glacier: models solved one at a time for the comparison")
    p.add_argument("--batch-precision", default="refined", choices=("refined", "fp64"),
                   help="batch96_glacier: FP32-refined solves to 1e-8 (default) or plain FP64")
    p.add_argument("--batch-graph-iters", type=int, default=None,
                   help="batch96_glacier: iterations per HIP graph replay (0: direct launches, for kernel-trace "
                        "profiling; default: the batch's own)")
    p.add_argument("--batch-chunk", type=int, default=32,
                   help="batch96_glacier refined: FP32 sweeps between FP64 outer steps, all models")
    p.add_argument("--heat-n", type=int, default=1023, help="heat_diffusion: cells per axis")
    p.add_argument("--serving-precision", default="fp64", choices=("fp64", "refined"),
                   help="serving_poisson40: plain FP64 with the driver's cadence (default) or FP32-refined solves")
    p.add_argument("--requests", type=int, default=200, help="serving_poisson40: number of requests")
    p.add_argument("--serving-3d", type=int, nargs=2, default=(64, 511), metavar=("LO", "HI"),
                   help="serving_poisson40: range of n of the 3D requests (n x n x ceil(n/2))")
    p.add_argument("--copy-calib", action="store_true",
                   help="run a 1 GiB device copy first (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return p.parse_args()


def profile_traffic(workload, role):
    """HBM bytes per launch of a kernel as measured by rocprofv3 (FETCH_SIZE x 2 +
    WRITE_SIZE, calibrated on a device copy) and recorded in profiles/*.md."""
    import glob
    here = os.path.dirname(os.path.abspath(__file__))
    for path in sorted(glob.glob(os.path.join(here, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 4 and parts[0] == "traffic_bytes_per_launch" and parts[1] == workload \
                        and parts[2] == role:
                    return int(parts[3]), os.path.relpath(path, here)
    return None, None


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def fused_cells(A, model):
    """(cells of the fused interior, those on planes of their column that
    hold a valid immersed-boundary location: the cells whose state the
    immersed-boundary sweep reads and writes; all of them without one)."""
    out = (C.c_longlong * 2)()
    A.check(A.lib().fi_fused_active_cells(model._h, out))
    return int(out[0]), int(out[1])


def event_times(A, fn, model, n):
    out = np.zeros(4)
    A.check(fn(model._h, n, A.dptr(out)))
    return out  # iteration, fused, shell, bc (ms)


def cpu_baseline(pkg, model, dims, seconds, ib, processed=None):
    """The CPU oracle (OpenMP, FP64) on a bounded number of PT iterations of
    the same model, in the headline's unit: 23 * 8 bytes per processed cell
    (processed: the cells the GPU step credits; the oracle sweeps every cell)
    and iteration; the reference formula over every cell beside it."""
    orc = ge.load_oracle()
    ref = orc.OracleModel(model.desc)
    names = ["P", "TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ", "VX", "VY", "VZ", "ETA", "ETA_NEXT"]
    if model.desc.gravity_is_field:
        names += ["GX", "GY", "GZ"]
    for k in names:
        ref.set(k, model.field(k).parent())
    if ib:
        A = pkg._abi
        for ph, base in (("NA", A.FI_F_OM_NA0), ("NS", A.FI_F_OM_NS0)):
            for q in range(8):
                ref.set("OM_%s%d" % (ph, q), pkg.full_stokes.DeviceField(model, base + q).parent())
    ref.advance(1)
    n, t0 = 0, time.perf_counter()
    while True:
        ref.advance(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 200:
            break
    cells = dims[0] * dims[1] * dims[2]
    credited = cells if processed is None else processed
    teff = 23 * 8 * credited * n / el / 1e9
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(teff, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "teff_reference_formula_gbs": round(23 * 8 * cells * n / el / 1e9, 3),
            "iterations_per_s": round(n / el, 3),
            "sample": "%d FP64 PT iterations of the same %dx%dx%d %s (CPU oracle, OpenMP, %.1f s); value credits "
                      "the %d cells the GPU step processes" % (n, dims[0], dims[1], dims[2],
                                                              "immersed-boundary glacier" if ib else "grid", el,
                                                              credited)}


def solve_comparison(pkg, dims, b=None, fp64=True, eps0=EPS0):
    """Converged glacier solves to FP64 residual 1e-8: plain FP64 (on model b,
    reset to the initial state, when given) vs FP32-refined (a new model,
    returned open with the result)."""
    a, cad = pkg.glacier_3d(dims, seed=1, glen=True, eps0=eps0)
    n = max(dims)
    out = {"grid": list(dims), "tol": 1e-8, "eps0": eps0}
    if fp64:
        own = b is None
        if own:
            b, _ = pkg.glacier_3d(dims, seed=1, glen=True, eps0=eps0)
        else:
            _reset(b)
        b.synchronize()
        t0 = time.perf_counter()
        it, nb = b.solve(400 * n, cad["ncheck"], 1e-8)
        t64 = time.perf_counter() - t0
        out["fp64"] = {"iterations": it, "seconds": round(t64, 3), "residual": float(max(nb)),
                       "cap": 400 * n, "ncheck": cad["ncheck"]}
        log("fp64 solve: %d iterations, %.2f s, residual %.3e" % (it, t64, max(nb)))
    a.synchronize()
    t0 = time.perf_counter()
    outer, inner, na = a.refined_solve(1e-8)
    t32 = time.perf_counter() - t0
    counts = dict(a.refined_counts)
    out["fp32_refined"] = {"outer_loop": "lagged PT: an FP64 outer step (residuals, viscosity, stress lag) "
                                         "every 32 FP32 sweeps, after a full-FP32 nonlinear start",
                           "outer_steps": outer, "fp32_iterations": inner,
                           "fp32_nonlinear_phase_iterations": counts["nonlinear_phase"],
                           "seconds": round(t32, 3), "residual": float(max(na))}
    log("refined solve: %d sweeps, %d outer steps, %.2f s, residual %.3e" % (inner, outer, t32, max(na)))
    if fp64:
        # per-field difference relative to the FP64 solution's stress / velocity scale
        names = ("P", "TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ", "VX", "VY", "VZ")
        vals = {k: (a.field(k).interior(), b.field(k).interior()) for k in names}
        s_scale = max(np.max(np.abs(vals[k][1])) for k in names[:7])
        v_scale = max(np.max(np.abs(vals[k][1])) for k in names[7:])
        errs = {k: float(np.max(np.abs(x - y)) / (s_scale if k[0] in "PT" else v_scale)) for k, (x, y) in vals.items()}
        out["speedup"] = round(t64 / t32, 3)
        out["rel_error_vs_fp64"] = {k: float("%.3e" % v) for k, v in errs.items()}
        if own:
            b.close()
    return out, a


def _line(metric, value, unit, ms_per_step, steps, warmup, dtype, config, higher=True, scaling="weak"):
    return {"metric": metric, "value": round(value, 4), "unit": unit, "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": higher,
            "scaling": scaling, "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


def _reset(model):
    """Back to the glacier's initial state: zero stress and velocity, eta = 1."""
    for name in ("P", "TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ", "VX", "VY", "VZ"):
        model.field(name).set(0.0)
    model.viscosity.eta.set(1.0)
    model.viscosity.eta_next.set(1.0)
    model.apply_bc("eta")


def bench_batch(pkg, args):
    """96 uneven glaciers solved as one batch vs one at a time.

    Cells are cubes of side 2/192 (a 192-cell axis spans the default 2.0),
    so the domain aspect ratio is the grid's (1 to 4).  refined (default):
    FP32 PT sweeps inside FP64 defect correction (lagged PT), every model to
    FP64 residual 1e-8 or 400n sweeps, all with one chunk length so their
    outer steps fall on the same batch steps; fp64: plain FP64 PT to 1e-8 or
    the driver's niter = 50n with ncheck = 5n."""
    refined = args.batch_precision == "refined"
    specs = pkg.batch_glaciers(96, 48, 192)
    dims = [d for d, _ in specs]
    cells = sum(d[0] * d[1] * d[2] for d in dims)
    log("setting up 96 glaciers, %.1f M cells in total (%s)" % (cells / 1e6, args.batch_precision))
    t0 = time.perf_counter()
    inst = []
    for d, geo in specs:
        m, cad = pkg.glacier_3d(d, glen=True, **geo)
        inst.append((m, cad["ncheck"], cad["niter"]))
    for m, _, _ in inst:
        m.synchronize()
    setup_s = time.perf_counter() - t0
    log("setup %.1f s" % setup_s)

    def refined_args(d):
        return dict(max_outer=-(-400 * max(d) // args.batch_chunk), ncheck=args.batch_chunk)

    batch = pkg.Batch(graph_iters=args.batch_graph_iters)
    for (m, nc, mi), d in zip(inst, dims):
        if refined:
            batch.add_refined(m, 1e-8, **refined_args(d))
        else:
            batch.add(m, mi, nc, 1e-8)
    t0 = time.perf_counter()
    last = t0
    while batch.info()["running"] > 0 or (not refined and batch.info()["rebuilds"] == 0):
        batch.step()
        if time.perf_counter() - last > 10.0:
            last = time.perf_counter()
            log("batch: %.1f s, %d running" % (last - t0, batch.info()["running"]))
    t_batch = time.perf_counter() - t0
    st = [batch.status(m) for m, _, _ in inst]
    rc = [batch.refined_counts(m) for m, _, _ in inst] if refined else None
    info = batch.info()
    batch.close()
    iters = [s[1] for s in st]
    log("batch: %.2f s, %.2f solves/s, info %s" % (t_batch, 96 / t_batch, info))
    # one at a time on a sample of the same models, reset to their initial state
    sample = list(range(0, 96, max(1, 96 // max(1, args.alone_sample))))[:args.alone_sample]
    for q in sample:
        _reset(inst[q][0])
    inst[0][0].synchronize()
    t0 = time.perf_counter()
    same = True
    for q in sample:
        m, nc, mi = inst[q]
        if refined:
            outer, it, _ = m.refined_solve(1e-8, **refined_args(dims[q]))
            same = same and it == iters[q] and outer == rc[q]["outer"]
        else:
            it, _ = m.solve(mi, nc, 1e-8)
            same = same and it == iters[q]
        log("alone: model %d %s, %d iterations, %.1f s so far" % (q, dims[q], it, time.perf_counter() - t0))
    t_alone = time.perf_counter() - t0
    alone_rate = len(sample) / t_alone
    sample_batch_iters = sum(iters[q] for q in sample)
    log("alone: %d models in %.2f s, %.2f solves/s" % (len(sample), t_alone, alone_rate))
    for m, _, _ in inst:
        m.close()
    states = {k: sum(s[0] == k for s in st) for k in ("converged", "max_iter", "failed")}
    if refined:
        how = ("FP32 PT sweeps inside FP64 defect correction (lagged PT, an FP64 outer step every %d sweeps) "
               "to FP64 residual 1e-8 or 400n sweeps" % args.batch_chunk)
    else:
        how = "FP64 PT to residual 1e-8 or the driver's niter = 50n, ncheck = 5n (test_vavilov.jl)"
    config = {"workload": "batch96_glacier", "precision": args.batch_precision,
              "instances": 96, "dims_range": [48, 192], "total_cells": cells,
              "description": "seeded synthetic glaciers (seeds 100..195), dims uniform in 48..192 per axis, cubic cells "
                             "of side 2/192 (domain n*2/192 per axis), Glen n=3 (A=1, eta_reg=30, log relaxation 0.05), "
                             + how,
              "batch_seconds": round(t_batch, 3), "states": states,
              "iterations": {"min": int(min(iters)), "max": int(max(iters)), "sum": int(sum(iters))},
              "one_at_a_time": {"models": len(sample), "seconds": round(t_alone, 3),
                                "solves_per_s": round(alone_rate, 4), "iterations": int(sample_batch_iters),
                                "same_iterations_as_batch": bool(same)},
              "speedup_vs_one_at_a_time": round((96 / t_batch) / alone_rate, 3),
              "batch_info": info, "setup_s": round(setup_s, 2)}
    if refined:
        config["outer_steps"] = {"min": min(r["outer"] for r in rc), "max": max(r["outer"] for r in rc)}
        config["max_residual"] = float("%.3e" % max(float(np.max(s[2])) for s in st))
        metric = "batched solves/s (96 glaciers 48-192 per axis, FP32-refined to FP64 residual 1e-8)"
        dtype = "f32+f64"
    else:
        metric = "batched solves/s (96 glaciers 48-192 per axis, FP64, tol 1e-8 / niter 50n)"
        dtype = "f64"
    line = _line(metric, 96 / t_batch, "solves/s", t_batch * 1e3, 1, 0, dtype, config)
    line["roofline"] = None
    line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def bench_h5(pkg, args):
    """Output of a solved glacier: write_h5 + write_xdmf of the fields the
    glacier driver saves (test_vavilov.jl:206, 237-247) at 255x255x127."""
    import shutil
    import tempfile
    dims = (255, 255, 127)
    m, cad = pkg.glacier_3d(dims, seed=1, glen=True)
    m.advance(200)
    m.synchronize()
    out = tempfile.mkdtemp(prefix="fastice_h5_")
    cells = dims[0] * dims[1] * dims[2]
    nbytes = 4 * 8 * cells
    reps = max(1, min(args.steps, 5))
    try:
        h5 = os.path.join(out, "results.h5")
        xdmf = os.path.join(out, "results.xdmf3")

        def once():
            fields = pkg.writers.glacier_output_fields(m)
            pkg.writers.write_h5(m.grid, h5, fields)
            pkg.writers.write_xdmf(m.grid, xdmf, fields, ("results.h5",))
            return fields

        once()
        t0 = time.perf_counter()
        for _ in range(reps):
            once()
        t = (time.perf_counter() - t0) / reps
        # the same arrays from host memory: the HDF5 write alone
        host = {k: (v.interior() if hasattr(v, "interior") else v) for k, v in pkg.writers.glacier_output_fields(m).items()}
        t0 = time.perf_counter()
        for _ in range(reps):
            pkg.writers.write_h5(m.grid, h5, host)
        th = (time.perf_counter() - t0) / reps
        ok = all(np.array_equal(pkg.writers.read_h5(h5, k), v) for k, v in host.items())
        size = os.path.getsize(h5)
    finally:
        shutil.rmtree(out, ignore_errors=True)
        m.close()
    log("h5: %.3f s per write (%.2f GB/s), host arrays %.3f s, file %d B, round trip %s" % (t, nbytes / t / 1e9, th,
                                                                                            size, ok))
    config = {"workload": "h5_glacier_255x255x127", "grid": list(dims),
              "fields": ["Pr", "Vm", "wt_ns_c", "wt_na_c"],
              "description": "device fields of a seeded glacier after 200 PT iterations -> interior download, "
                             "velocity magnitude, HDF5 datasets and XDMF3 (Writers.jl) on the box's local disk",
              "bytes_per_write": nbytes, "file_bytes": size, "round_trip_equal": bool(ok),
              "host_arrays_seconds": round(th, 4)}
    line = _line("HDF5+XDMF3 output GB/s (4 glacier fields, 255x255x127)", nbytes / t / 1e9, "GB/s", t * 1e3, reps, 1,
                 "f64", config)
    line["roofline"] = None
    line["cpu_baseline"] = {"value": round(nbytes / th / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
                            "sample": "the same four arrays written from host memory (HDF5 write only)"}
    print(json.dumps(line), flush=True)


def margin_residual(model, margin=None):
    """Inf-norms of a glacier's residuals (r_P, and r_V outside its margin
    faces) and the margin faces: velocity faces whose residual exceeds 1e-2
    (on the 511x511x255 glacier, 16 faces at the ice margin that the
    reference's null-space rule calls valid and non-empty with an na fraction
    of 1.2e-6; their residual is about -rho g and no PT iteration changes it).
    Pass the margin of a converged solve to use the same faces."""
    model.compute_residuals()
    r = {k: model.field(k).interior() for k in ("RP", "RVX", "RVY", "RVZ")}
    if margin is None:
        margin = {k: np.abs(r[k]) > 1e-2 for k in ("RVX", "RVY", "RVZ")}
    rv = max(float(np.max(np.abs(np.where(margin[k], 0.0, r[k])))) for k in ("RVX", "RVY", "RVZ"))
    return float(np.max(np.abs(r["RP"]))), rv, margin


# FP32 sweeps per outer step of the 511 comparison: the margin faces' velocity
# drifts, and with it the viscosity around them; a viscosity lagged over 32
# sweeps leaves r_V near them at about 1e-7, over 16 sweeps at about 3e-8
CHUNK_511 = 16


def bench_converged_511(pkg, args):
    """The precision workload at its configured size: the 511x511x255 glacier
    solved by plain FP64 PT and by the FP32-refined solver (16 sweeps per
    outer step), each checked every 1000 iterations / 1600 sweeps for r_P and
    r_V outside the margin faces.  The margin faces keep r_V = 0.5 and their
    velocity drifts, so the problem has no fixed point there: FP64 reaches
    1e-8 elsewhere, the refined solver's lagged viscosity holds it near 3e-8.
    value = the refined speedup to 1e-7 (a level both reach); the time and
    count to 1e-8 (FP64) and the refined solver's floor are in the config,
    with the per-field difference of the final states (velocity outside the
    margin faces) relative to the FP64 stress and velocity scales."""
    dims = (511, 511, 255)
    levels = (1e-7, 1e-8)
    b, _ = pkg.glacier_3d(dims, seed=1, glen=True, relax_rate=RELAX_511)
    t64, it64, margin, reach64 = 0.0, 0, None, {}
    while True:
        b.synchronize()
        t0 = time.perf_counter()
        it, _ = b.solve(1000, 1000, 1e-30)
        b.synchronize()
        t64 += time.perf_counter() - t0
        it64 += it
        if it64 < 8000:
            continue
        rp, rv, mg = margin_residual(b)
        r = max(rp, rv)
        for lv in levels:
            if r < lv and lv not in reach64:
                reach64[lv] = (it64, t64)
        log("fp64 %d iterations %.1f s: r_P %.3e, r_V outside %d margin faces %.3e"
            % (it64, t64, rp, sum(int(v.sum()) for v in mg.values()), rv))
        if r < levels[-1] or it64 >= 60000:
            margin = mg
            break
    a, _ = pkg.glacier_3d(dims, seed=1, glen=True, relax_rate=RELAX_511)
    t32, sweeps, outer, first, reach32, best, since = 0.0, 0, 0, True, {}, None, 0
    while True:
        a.synchronize()
        t0 = time.perf_counter()
        o, s, _ = a.refined_solve(1e-30, max_outer=1600 // CHUNK_511, ncheck=CHUNK_511)
        a.synchronize()
        t32 += time.perf_counter() - t0
        if first:
            a.set_option("refine_full_phase", 0)
            first = False
        sweeps += s
        outer += o
        rp32, rv32, _ = margin_residual(a, margin)
        r = max(rp32, rv32)
        for lv in levels:
            if r < lv and lv not in reach32:
                reach32[lv] = (sweeps, t32)
        log("refined %d sweeps %d outer steps %.1f s: r_P %.3e, r_V outside the margin faces %.3e"
            % (sweeps, outer, t32, rp32, rv32))
        # stop at the tolerance, or at the floor: no halving over 5 checks
        if best is None or r < 0.5 * best:
            best, since = r, 0
        else:
            since += 1
        if r < levels[-1] or since >= 5 or sweeps >= 60000:
            break
    names = ("P", "TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ", "VX", "VY", "VZ")
    vm = {"VX": margin["RVX"], "VY": margin["RVY"], "VZ": margin["RVZ"]}
    vals = {k: (a.field(k).interior(), b.field(k).interior()) for k in names}
    s_scale = max(float(np.max(np.abs(vals[k][1]))) for k in names[:7])
    v_scale = max(float(np.max(np.abs(np.where(vm[k], 0.0, vals[k][1])))) for k in names[7:])
    errs = {}
    for k in names:
        d = np.abs(vals[k][0] - vals[k][1])
        if k in vm:
            d = np.where(vm[k], 0.0, d)
        errs[k] = float("%.3e" % (float(np.max(d)) / (s_scale if k[0] in "PT" else v_scale)))
    a.close()
    b.close()
    to = lambda d: {("%g" % lv): {"count": d[lv][0], "seconds": round(d[lv][1], 3)} for lv in levels if lv in d}
    speedup = reach64[1e-7][1] / reach32[1e-7][1] if 1e-7 in reach64 and 1e-7 in reach32 else 0.0
    config = {"workload": "converged_511x511x255", "grid": list(dims),
              "criterion": "r_P and r_V outside the margin faces (which keep |r_V| = 0.5 and whose velocity drifts)",
              "margin_faces": {k: int(v.sum()) for k, v in margin.items()},
              "fp64": {"iterations": it64, "seconds": round(t64, 3), "r_P": rp, "r_V_outside_margin": rv,
                       "reached": to(reach64)},
              "fp32_refined": {"fp32_sweeps": sweeps, "outer_steps": outer, "sweeps_per_outer_step": CHUNK_511,
                               "seconds": round(t32, 3), "r_P": rp32, "r_V_outside_margin": rv32,
                               "reached": to(reach32)},
              "value_is": "FP64 time / refined time to r < 1e-7 outside the margin faces",
              "rel_error_vs_fp64": errs,
              "description": "synthetic glacier (seed 1), Glen n=3, eta_reg=30, log relaxation %g, PT r=0.9 re=5pi "
                             "vdtau_div=1.1, traction-free box; solver time only (residual checks not timed)"
                             % RELAX_511}
    line = _line("refined/FP64 converged-solve speedup at 511x511x255", speedup, "x", t32 * 1e3, 1, 0,
                 "f32 sweeps in f64 defect correction", config)
    line["roofline"] = None
    line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def bench_glacier_solve(pkg, args, precision):
    """One solve of the 255x255x127 glacier to FP64 residual 1e-8, alone:
    FP32-refined (lagged PT) or plain FP64, the two solves the default
    line's converged-solve comparison times one after the other (the
    workloads the solve profiles cover)."""
    dims = (255, 255, 127)
    model, cad = pkg.glacier_3d(dims, seed=1, glen=True, eps0=EPS0)
    model.synchronize()
    t0 = time.perf_counter()
    if precision == "refined":
        outer, sweeps, res = model.refined_solve(1e-8)
        counts = dict(model.refined_counts)
        info = {"outer_steps": outer, "fp32_sweeps": sweeps, "fp32_nonlinear_phase_sweeps": counts["nonlinear_phase"]}
        dtype = "f32 sweeps, f64 outer steps"
    else:
        it, res = model.solve(400 * max(dims), cad["ncheck"], 1e-8)
        info = {"fp64_iterations": it}
        dtype = "f64"
    model.synchronize()
    el = time.perf_counter() - t0
    model.close()
    log("%s solve: %s, %.3f s, residual %.3e" % (precision, info, el, max(res)))
    cfg = {"workload": "%s_solve_255x255x127" % precision, "grid": list(dims), "tol": 1e-8,
           "residual": float(max(res))}
    cfg.update(info)
    print(json.dumps(_line("converged_solve_time", el, "s", el * 1e3, 1, 0, dtype, cfg, higher=False)), flush=True)


def bench_manufactured(pkg, args):
    """The 3D manufactured free-slip solution at 255^3 in FP64: 2000 PT
    iterations (T_eff), then a solve to residual 1e-8 checked every 255
    iterations, with the relative error norms of |V| and tau_II against the
    closed form (..._3D.jl:140-167)."""
    A = pkg._abi
    n = 255
    kw, cad = pkg.manufactured_3d(n)
    model = pkg.IsothermalFullStokesModel(**kw)
    model.viscosity.eta.set(1.0)
    model.viscosity.eta_next.set(1.0)
    model.apply_bc("eta")
    model.synchronize()
    cells = n ** 3
    n_it = 2000
    log("manufactured 255^3: %d FP64 iterations" % n_it)
    t0 = time.perf_counter()
    model.advance(n_it)
    model.synchronize()
    ms = (time.perf_counter() - t0) / n_it * 1e3
    teff = 23 * 8 * cells / (ms * 1e-3) / 1e9
    ev = event_times(A, A.lib().fi_time_iterations, model, 20)
    icells, _ = fused_cells(A, model)
    shell_pts = (n + 1) ** 3 - icells if icells else 0
    kbytes = 25 * 8 * icells + 18 * 8 * shell_pts
    log("iteration %.3f ms, T_eff %.1f GB/s" % (ms, teff))
    # converged solve from the zero state
    for f in ("P", "TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ", "VX", "VY", "VZ"):
        model.field(f).set(0.0)
    model.viscosity.eta.set(1.0)
    model.viscosity.eta_next.set(1.0)
    model.apply_bc("eta")
    model.synchronize()
    t0 = time.perf_counter()
    it, norms = model.solve(200 * n, n, 1e-8)
    model.synchronize()
    solve_s = time.perf_counter() - t0
    V = {k: model.field("V" + k.upper()).interior() for k in ("x", "y", "z")}
    tau = {k: model.field("T" + k.upper()).interior() for k in ("xx", "yy", "zz", "xy", "xz", "yz")}
    err = pkg.problems.manufactured_errors(model.grid, V, tau)
    log("solve: %d iterations, %.2f s, residual %.2e, errors %s" % (it, solve_s, max(norms), err))
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        orc = ge.load_oracle()
        ref = orc.OracleModel(model.desc)
        for k in ("GX", "GY", "GZ"):
            ref.set(k, model.field(k).parent())
        ref.set("ETA", 1.0)
        ref.set("ETA_NEXT", 1.0)
        ref.advance(2)
        k_cpu, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < args.cpu_seconds:
            ref.advance(4)
            k_cpu += 4
        el = time.perf_counter() - t0
        cpu = {"value": round(23 * 8 * cells * k_cpu / el / 1e9, 3), "unit": "GB/s",
               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port",
               "sample": "%d FP64 PT iterations of the same 255^3 manufactured problem (CPU oracle, OpenMP, %.1f s)"
                         % (k_cpu, el)}
    config = {"workload": "manufactured3d_255", "grid": [n, n, n], "box": "[-1, 1]^3, free slip on every face",
              "iterations_timed": n_it, "iteration_ms": round(ms, 4),
              "converged": {"tol": 1e-8, "ncheck": n, "iterations": int(it), "seconds": round(solve_s, 3),
                            "residual": float(max(norms)),
                            "error_vs_closed_form": {"Linf_V": err[0], "Linf_tauII": err[1], "L1_V": err[2],
                                                     "L1_tauII": err[3], "L2_V": err[4], "L2_tauII": err[5]}}}
    line = _line("PT-Stokes T_eff GB/s at 255^3 (manufactured free slip, FP64)", teff, "GB/s", ms, n_it, 0, "f64",
                 config)
    achieved = kbytes / (ev[1] * 1e-3) / 1e9
    line["roofline"] = {"bound": "hbm", "kernel": "k_fused3<double>", "achieved": round(achieved, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": None, "kernel_ms": round(float(ev[1]), 4),
                        "algorithmic_bytes_per_launch": int(kbytes)}
    line["cpu_baseline"] = cpu
    model.close()
    print(json.dumps(line), flush=True)
    return 0


def bench_heat(pkg, args):
    """PT heat diffusion at 1023^3 in FP64 (the size of
    scripts_future_API/benchmark_diffusion_3D.jl:65): T_eff by that script's
    formula, 8 * cells * 8 B per iteration (:119-123), and the fused
    kernel's roofline with its 9 arrays per cell (T, T_o, q read; T, q
    written)."""
    H = pkg.heat
    n = args.heat_n
    dims = (n, n, n)
    L = 10.0
    grid = pkg.UniformGrid(origin=(-0.5 * L,) * 3, extent=(L,) * 3, dims=dims)
    dx = L / n
    lam = 1.0
    bcs = {side: H.BoundaryCondition(H.Flux, 0.0) for pair in H.SIDES for side in pair}
    m = H.HeatDiffusionModel(grid, bcs, physics={"λ_ρCp": lam},
                             iter_params={"Δτ": {"q": 1.0, "T": dx * dx / (lam * 6.3)}})
    t0 = time.perf_counter()
    xc = grid.coords(pkg.Center(), 0, 0, n + 2)
    g = np.exp(-xc * xc)
    T0 = np.multiply.outer(np.multiply.outer(g, g), g)  # exp(-x^2 - y^2 - z^2) on the parent box
    m.temperature.set(T0.T)  # symmetric: the transposed view is the Julia-order array, without a copy
    m.temperature_old.set(T0.T)
    del T0
    m.apply_bc()
    log("heat %s: set-up %.1f s" % (dims, time.perf_counter() - t0))
    dt = 0.1
    m.advance(args.warmup, dt)
    ms = m.time_iterations(dt, args.steps)
    cells = n ** 3
    teff = 8 * cells * 8 / (ms[0] * 1e-3) / 1e9
    alg = 9 * cells * 8
    achieved = alg / (ms[1] * 1e-3) / 1e9
    norms = m.residual_norms(dt)
    m.close()
    traffic, src = profile_traffic("heat_diffusion_%d" % n, "fused")
    config = {"workload": "heat_diffusion_%d" % n, "grid": list(dims), "dtype": "f64",
              "description": "PT heat diffusion (Models/Diffusion/Heat), Gaussian initial temperature, zero flux on "
                             "every side, one fused kernel per iteration + boundary kernels",
              "t_eff_formula": "8 * cells * sizeof(Float64) / t_iteration (benchmark_diffusion_3D.jl:119-123)",
              "iteration_ms": round(ms[0], 4), "kernel_ms": round(ms[1], 4),
              "residual_after": [float("%.3e" % x) for x in norms]}
    line = _line("PT heat diffusion T_eff GB/s at %d^3" % n, teff, "GB/s", ms[0], args.steps, args.warmup, "f64", config)
    line["roofline"] = {"bound": "hbm", "kernel": "k_heat_iter_z", "achieved": round(achieved, 2),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": traffic, "traffic_source": src, "algorithmic_bytes_per_launch": alg}
    if args.no_cpu:
        line["cpu_baseline"] = None
    else:
        line["cpu_baseline"] = heat_cpu_baseline(args.cpu_seconds)
    print(json.dumps(line), flush=True)


def heat_cpu_baseline(seconds):
    """The oracle's heat iteration on a 255^3 sample, T_eff by the same formula."""
    orc = ge.load_oracle()
    n = 255
    h = orc.OracleHeat(3, [n, n, n], [10.0] * 3, [[2, 2]] * 3, [[0.0, 0.0]] * 3, 1.0, 1.0, (10.0 / n) ** 2 / 6.3)
    x = (np.arange(n + 3) - 0.5) * (10.0 / n) - 5.0
    g = np.exp(-x * x)
    T0 = np.multiply.outer(np.multiply.outer(g, g), g)
    h.set("T", T0)
    h.set("TO", T0)
    h.advance(0.1, 1)
    it, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds and it < 200:
        h.advance(0.1, 1)
        it += 1
    el = time.perf_counter() - t0
    return {"value": round(8 * n ** 3 * 8 * it / el / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": "%d iterations of the CPU oracle's PT heat iteration at 255^3 (%.1f s)" % (it, el)}


def bench_serving(pkg, args):
    """Continuous-batching service under Poisson arrivals at 40/s."""
    reqs = pkg.poisson_requests(args.requests, 40.0, seed=7, range3d=tuple(args.serving_3d))
    log("serving %d requests (%d 3D), arrivals over %.1f s" % (len(reqs), sum(r.ndims == 3 for r in reqs),
                                                                reqs[-1].t_arrival))
    srv = pkg.Server(pkg, max_slots=96, precision=args.serving_precision)
    try:
        s = srv.run(reqs, realtime=True, log=log)
        recs = list(srv.results.values())
    finally:
        srv.close()
    # offline batched throughput of the same mix: every request present at t = 0
    log("offline batch of the same %d requests" % len(reqs))
    srv = pkg.Server(pkg, max_slots=96, precision=args.serving_precision)
    try:
        off = srv.run(reqs, offline=True, log=log)
    finally:
        srv.close()
    lat3 = [r["latency_s"] for r in recs if r["ndims"] == 3]
    lat2 = [r["latency_s"] for r in recs if r["ndims"] == 2]
    config = {"workload": "serving_poisson40", "requests": len(reqs), "arrival_rate_per_s": 40.0,
              "offered_window_s": round(reqs[-1].t_arrival, 3),
              "mix": "p=0.5 3D glacier (n log-uniform %d..%d, n x n x ceil(n/2)) / 2D glacier section "
                     "(n log-uniform 64..511, n x ceil(n/2)); Glen n=3, " % tuple(args.serving_3d)
                     + ("FP32 sweeps in FP64 defect correction (lagged PT, 32 sweeps per outer step) to residual "
                        "1e-8 or 400n sweeps" if args.serving_precision == "refined" else
                        "FP64 to residual 1e-8 or 100n iterations, ncheck 5n"),
              "precision": args.serving_precision,
              "latency_s": {k: (round(v, 4) if v is not None else None) for k, v in s["latency_s"].items()},
              "latency_s_3d": {k: round(v, 4) for k, v in pkg.percentiles(lat3).items()} if lat3 else None,
              "latency_s_2d": {k: round(v, 4) for k, v in pkg.percentiles(lat2).items()} if lat2 else None,
              "offline_batch": {"solves_per_s": round(off["solves_per_s"], 4),
                                "makespan_s": round(off["makespan_s"], 3)},
              "serving_vs_offline": round(s["solves_per_s"] / off["solves_per_s"], 3),
              "makespan_s": round(s["makespan_s"], 3), "converged": s["converged"],
              "max_iter": s["max_iter"], "failed": s["failed"],
              "requests_per_s": round(s["requests_per_s"], 4),
              "latency_s_converged": {k: (round(v, 4) if v is not None else None)
                                      for k, v in s["latency_s_converged"].items()},
              "value_counts": "requests that reached residual 1e-8",
              "batch_steps": s["batch_steps"], "setup_s_mean": round(s["setup_s_mean"], 4)}
    line = _line("serving solves/s (p99 latency in config), Poisson 40/s", s["solves_per_s"], "solves/s",
                 s["makespan_s"] * 1e3, 1, 0, "f32+f64" if args.serving_precision == "refined" else "f64", config)
    line["roofline"] = None
    line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: each solve is one device-resident domain and this build has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calib:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    pkg = ge.load_package()
    A = pkg._abi
    if args.workload == "batch96_glacier":
        return bench_batch(pkg, args)
    if args.workload == "serving_poisson40":
        return bench_serving(pkg, args)
    if args.workload == "h5_glacier_255x255x127":
        return bench_h5(pkg, args)
    if args.workload == "heat_diffusion":
        return bench_heat(pkg, args)
    if args.workload == "manufactured3d_255":
        return bench_manufactured(pkg, args)
    if args.workload == "converged_511x511x255":
        return bench_converged_511(pkg, args)
    if args.workload in ("refined_solve_255x255x127", "fp64_solve_255x255x127"):
        return bench_glacier_solve(pkg, args, args.workload.split("_")[0])
    dims = WORKLOADS[args.workload]
    cells = dims[0] * dims[1] * dims[2]
    glacier = args.workload.startswith("glacier")

    log("setting up %s" % args.workload)
    t0 = time.perf_counter()
    if glacier:
        model, cad = pkg.glacier_3d(dims, seed=1, glen=True, relax_rate=RELAX_511)
        model.advance(50)  # leave the zero initial state
    else:
        kw, _ = pkg.manufactured_3d(dims)
        model = pkg.IsothermalFullStokesModel(**kw)
        model.viscosity.eta.set(1.0)
        model.viscosity.eta_next.set(1.0)
        model.apply_bc("eta")
    model.synchronize()
    setup_s = time.perf_counter() - t0
    log("setup %.1f s" % setup_s)

    # FP64 iteration
    model.advance(args.warmup)
    model.synchronize()
    t0 = time.perf_counter()
    model.advance(args.steps)
    model.synchronize()
    ms64 = (time.perf_counter() - t0) / args.steps * 1e3
    teff64 = 23 * 8 * cells / (ms64 * 1e-3) / 1e9
    ev64 = event_times(A, A.lib().fi_time_iterations, model, max(10, min(50, args.steps // 4)))
    icells, acells = fused_cells(A, model)
    # stress points [1, n+1]^3 outside the fused interior: the fused launch
    # also updates their stress (the shell's stress update rides in it)
    shell_pts = (dims[0] + 1) * (dims[1] + 1) * (dims[2] + 1) - icells if icells else 0
    gfields = 3 if model.desc.gravity_is_field else 0
    log("fp64 %.3f ms/iter, T_eff %.1f GB/s" % (ms64, teff64))

    # cells a step processes: the fused interior's cells on planes of their
    # column that hold a valid location (the immersed-boundary sweep reads and
    # writes nothing else there), plus every cell of the boundary shell
    processed = acells + (cells - icells)
    solve = None
    if glacier:
        # The step is one cycle of the FP32-refined solver: 32 FP32 PT
        # sweeps, then the FP64 outer step (residuals, viscosity update,
        # stress lag, corrections added to the FP64 state).  The solve starts
        # as a solve does (full-FP32 nonlinear sweeps, then outer cycles);
        # the timed cycles continue it without a second full-FP32 start.
        # The reference scheme does not reach 1e-8 on this grid (see
        # SOLVE_DIMS): the converged comparison runs at 255x255x127.
        model32, _ = pkg.glacier_3d(dims, seed=1, glen=True, relax_rate=RELAX_511)
        model32.refined_solve(1e-30, max_outer=2, ncheck=CYCLE)
        model32.set_option("refine_full_phase", 0)
        model32.refined_solve(1e-30, max_outer=max(1, args.warmup), ncheck=CYCLE)
        model32.synchronize()
        t0 = time.perf_counter()
        outer, sweeps, _ = model32.refined_solve(1e-30, max_outer=args.steps, ncheck=CYCLE)
        model32.synchronize()
        t_total = time.perf_counter() - t0
        cycles = sweeps / CYCLE
        ms_step = t_total / cycles * 1e3
        ev32 = event_times(A, A.lib().fi_refined_time_inner, model32, max(10, min(50, args.steps)))
        model32.close()
        # converged solves to FP64 residual 1e-8, plain FP64 (unless --no-solve)
        # and FP32-refined, at the largest glacier size on which they converge
        if not args.no_compare:
            log("converged-solve comparison at %s" % (SOLVE_DIMS,))
            solve, a = solve_comparison(pkg, SOLVE_DIMS, fp64=not args.no_solve)
            a.close()
            log("solve: %s" % json.dumps(solve))
        sweep_ms = ev32[0]
        outer_ms = (t_total * 1e3 - sweeps * sweep_ms) / outer
        value = CYCLE * 23 * 4 * processed / (ms_step * 1e-3) / 1e9
        log("refined cycle %.3f ms (%d sweeps, %d outer steps in %.3f s), %.1f GB/s over processed cells"
            % (ms_step, sweeps, outer, t_total, value))
        dtype = "f32 sweeps in f64 defect correction"
        # dominant kernel, the fused FP32 correction sweep: 23 arrays of 4 B per
        # processed cell (the reference's per-cell count; masks, forcing and the
        # shell's stress points it also updates are not credited)
        kname, kms, kbytes = "k_fused3<float, IB>", ev32[1], 23 * 4 * acells
        role = "fp32_fused"
        events = {"fp32_iteration": ev32[0], "fp32_fused": ev32[1], "fp32_shell": ev32[2], "fp32_bc": ev32[3],
                  "fp64_iteration": ev64[0], "fp64_fused": ev64[1], "fp64_shell": ev64[2], "fp64_bc": ev64[3]}
        step_info = {"step": "one cycle of the FP32-refined solver: %d FP32 PT sweeps of the glacier, then the "
                             "FP64 outer step" % CYCLE,
                     "fp32_sweeps": int(sweeps), "outer_steps": int(outer), "timed_seconds": round(t_total, 4),
                     "fp32_sweep_ms": round(float(sweep_ms), 4), "fp64_outer_step_ms": round(float(outer_ms), 4),
                     "fp32_sweeps_per_s": round(sweeps / t_total, 2),
                     "teff_reference_formula_gbs": round(CYCLE * 23 * 4 * cells / (ms_step * 1e-3) / 1e9, 2)}
    else:
        value, ms_step, dtype = 23 * 8 * processed / (ms64 * 1e-3) / 1e9, ms64, "f64"
        # fused FP64 sweep: 23 arrays of 8 B per interior cell (the shell's
        # stress points it also updates are not credited)
        kname, kms, kbytes = "k_fused3<double>", ev64[1], 23 * 8 * icells
        role = "fp64_fused"
        events = {"fp64_iteration": ev64[0], "fp64_fused": ev64[1], "fp64_shell": ev64[2], "fp64_bc": ev64[3]}
        step_info = {"step": "one FP64 PT iteration of the whole grid"}
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    traffic, tsrc = profile_traffic(args.workload, role)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(pkg, model, dims, args.cpu_seconds, glacier, processed)

    config = {"workload": args.workload, "grid": list(dims),
              "value_is": "23 * sizeof(T) bytes (tm_stokes_mpi_wip.jl:180) per processed cell and PT sweep, over the "
                          "step's wall time",
              "teff_pct_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 2),
              "processed_cells": int(processed), "cells": int(cells),
              "fp64_iteration": {"ms": round(ms64, 4), "iterations_per_s": round(1e3 / ms64, 2),
                                 "teff_processed_gbs": round(23 * 8 * processed / (ms64 * 1e-3) / 1e9, 2),
                                 "teff_reference_formula_gbs": round(teff64, 2)},
              "setup_s": round(setup_s, 2)}
    config.update(step_info)
    if glacier:
        config["description"] = ("synthetic glacier, 8 Gaussian bumps (seed 1), parabolic surface, volume-fraction "
                                 "immersed boundary, Glen n=3 (A=1, eta_reg=30, log relaxation %g, PT r=0.9 re=5pi "
                                 "vdtau_div=1.1), traction-free box; the converged comparison: regularised at strain "
                                 "rate %g (eta_reg=%g), log relaxation 0.01, PT r=0.5 re=5pi vdtau_div=4"
                                 % (RELAX_511, EPS0, pkg.problems.eta_reg_from_rate(EPS0)))
        config["converged_solve"] = solve
        # the reference's T_eff (teff_reference_formula_gbs) counts every cell
        # of the box; the sweep reads and writes only the cells on planes of
        # their column that hold a valid location (the rest holds no valid
        # location and stays 0), and value counts only the processed cells
        config["empty_space"] = {"interior_cells": int(icells), "processed_interior_cells": int(acells),
                                 "processed_fraction": round(processed / cells, 4)}
    line = {
        "metric": "PT-Stokes T_eff GB/s (% HBM peak) at 511x511x255",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": config,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": tsrc,
                     "kernel_ms": round(float(kms), 4), "algorithmic_bytes_per_launch": int(kbytes),
                     "interior_cells": int(icells), "processed_interior_cells": int(acells),
                     "shell_stress_points": int(shell_pts),
                     "event_ms": {k: round(float(v), 4) for k, v in events.items()}},
        "cpu_baseline": cpu,
    }
    model.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

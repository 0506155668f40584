"""PT-Stokes benchmark: T_eff [GB/s] of the PT iteration on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu] [--no-solve]

Workloads
  glacier_511x511x255 (default): 3D synthetic glacier, 511x511x255 cells,
      seeded 8-Gaussian-bump bed under a parabolic surface, level sets ->
      volume fractions (immersed boundary), Glen n = 3 regularised power law,
      gravity-driven.  A step is one FP32 PT iteration of the defect-correction
      solver (the production sweep); the FP64 iteration of the same model is
      timed beside it, and a converged-solve comparison (FP32-refined vs plain
      FP64, both to FP64 residual 1e-8) runs on the same glacier at 127x127x63.
  stokes3d_fp64_511x511x255: isoviscous free-slip box with manufactured
      forcing, plain FP64 iteration.

value = T_eff = A_eff * K / t with A_eff = 23 * sizeof(T) * nx*ny*nz bytes per
iteration (scripts_future_API/tm_stokes_mpi_wip.jl:180), inputs resident in
HBM.  roofline: the dominant kernel's algorithmic bytes per launch (each
field it must read or write, once, over its cells; masks not credited) over
its HIP-event duration on the model's stream.  cpu_baseline: the CPU oracle
(OpenMP) on a bounded number of FP64 iterations of the same grid.
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
DEFAULT = "glacier_511x511x255"
WORKLOADS = {
    "glacier_511x511x255": (511, 511, 255),
    "stokes3d_fp64_511x511x255": (511, 511, 255),
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline")
    p.add_argument("--no-solve", action="store_true", help="skip the converged-solve comparison")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--solve-dims", default="255,255,127")
    p.add_argument("--copy-calib", action="store_true",
                   help="run a 1 GiB device copy first (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return p.parse_args()


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def fused_cells(A, model):
    lo, hi = (C.c_int * 3)(), (C.c_int * 3)()
    on = A.lib().fi_fused_region(model._h, lo, hi)
    return (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) if on else 0


def event_times(A, fn, model, n):
    out = np.zeros(4)
    A.check(fn(model._h, n, A.dptr(out)))
    return out  # iteration, fused, shell, bc (ms)


def cpu_baseline(pkg, model, dims, seconds, ib):
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
    teff = 23 * 8 * cells * n / el / 1e9
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(teff, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": "%d FP64 PT iterations of the same %dx%dx%d %s (CPU oracle, OpenMP, %.1f s)"
                      % (n, dims[0], dims[1], dims[2], "immersed-boundary glacier" if ib else "grid", el)}


def solve_comparison(pkg, dims):
    """Converged glacier solves to FP64 residual 1e-8: plain FP64 vs FP32-refined."""
    a, cad = pkg.glacier_3d(dims, seed=1, glen=True)
    b, _ = pkg.glacier_3d(dims, seed=1, glen=True)
    n = max(dims)
    b.synchronize()
    t0 = time.perf_counter()
    it, nb = b.solve(400 * n, cad["ncheck"], 1e-8)
    t64 = time.perf_counter() - t0
    a.synchronize()
    t0 = time.perf_counter()
    outer, inner, na = a.refined_solve(1e-8, max_outer=300)
    t32 = time.perf_counter() - t0
    # per-field difference relative to the FP64 solution's stress / velocity scale
    names = ("P", "TXX", "TYY", "TZZ", "TXY", "TXZ", "TYZ", "VX", "VY", "VZ")
    vals = {k: (a.field(k).interior(), b.field(k).interior()) for k in names}
    s_scale = max(np.max(np.abs(vals[k][1])) for k in names[:7])
    v_scale = max(np.max(np.abs(vals[k][1])) for k in names[7:])
    errs = {k: float(np.max(np.abs(x - y)) / (s_scale if k[0] in "PT" else v_scale)) for k, (x, y) in vals.items()}
    counts = dict(a.refined_counts)
    a.close()
    b.close()
    return {"grid": list(dims), "tol": 1e-8,
            "fp64": {"iterations": it, "seconds": round(t64, 3), "residual": float(max(nb))},
            "fp32_refined": {"outer_steps": outer, "fp32_iterations": inner,
                             "fp32_nonlinear_phase_iterations": counts["nonlinear_phase"],
                             "seconds": round(t32, 3),
                             "residual": float(max(na))},
            "speedup": round(t64 / t32, 3),
            "rel_error_vs_fp64": {k: float("%.3e" % v) for k, v in errs.items()}}


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
    dims = WORKLOADS[args.workload]
    cells = dims[0] * dims[1] * dims[2]
    glacier = args.workload.startswith("glacier")

    log("setting up %s" % args.workload)
    t0 = time.perf_counter()
    if glacier:
        model, cad = pkg.glacier_3d(dims, seed=1, glen=True)
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
    icells = fused_cells(A, model)
    gfields = 3 if model.desc.gravity_is_field else 0
    log("fp64 %.3f ms/iter, T_eff %.1f GB/s" % (ms64, teff64))

    if glacier:
        # FP32 correction iteration (inputs resident in HBM)
        A.check(A.lib().fi_refined_advance_inner(model._h, args.warmup))
        model.synchronize()
        t0 = time.perf_counter()
        A.check(A.lib().fi_refined_advance_inner(model._h, args.steps))
        model.synchronize()
        ms32 = (time.perf_counter() - t0) / args.steps * 1e3
        teff32 = 23 * 4 * cells / (ms32 * 1e-3) / 1e9
        ev32 = event_times(A, A.lib().fi_refined_time_inner, model, max(10, min(50, args.steps // 4)))
        log("fp32 %.3f ms/iter, T_eff %.1f GB/s" % (ms32, teff32))
        value, ms_step, dtype = teff32, ms32, "f32 sweeps in f64 defect correction"
        # fused FP32 correction sweep: R P tau(6) V(3) eta fP g(3), W P tau(6) V(3)
        kname, kms, kbytes = "k_fused3<float, IB>", ev32[1], 25 * 4 * icells
        events = {"fp32_iteration": ev32[0], "fp32_fused": ev32[1], "fp32_shell": ev32[2], "fp32_bc": ev32[3],
                  "fp64_iteration": ev64[0], "fp64_fused": ev64[1], "fp64_shell": ev64[2], "fp64_bc": ev64[3]}
    else:
        value, ms_step, dtype = teff64, ms64, "f64"
        # fused FP64 sweep: R P tau(6) V(3) eta [g(3)], W P tau(6) V(3) eta_next
        kname, kms, kbytes = "k_fused3<double>", ev64[1], (22 + gfields) * 8 * icells
        events = {"fp64_iteration": ev64[0], "fp64_fused": ev64[1], "fp64_shell": ev64[2], "fp64_bc": ev64[3]}
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0

    solve = None
    if glacier and not args.no_solve:
        sd = tuple(int(x) for x in args.solve_dims.split(","))
        log("converged-solve comparison at %s" % (sd,))
        solve = solve_comparison(pkg, sd)
        log("solve: %s" % json.dumps(solve))

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(pkg, model, dims, args.cpu_seconds, glacier)

    config = {"workload": args.workload, "grid": list(dims),
              "step": "one PT iteration of the whole grid" + (" (FP32 correction sweep)" if glacier else ""),
              "teff_pct_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 2),
              "iterations_per_s": round(1e3 / ms_step, 2),
              "fp64_iteration": {"ms": round(ms64, 4), "teff_gbs": round(teff64, 2),
                                 "teff_pct_hbm_peak": round(100.0 * teff64 / HBM_PEAK_GBS, 2)},
              "setup_s": round(setup_s, 2)}
    if glacier:
        config["description"] = ("synthetic glacier, 8 Gaussian bumps (seed 1), parabolic surface, volume-fraction "
                                 "immersed boundary, Glen n=3 (A=1, eta_reg=30, log relaxation 0.05), traction-free box")
        config["converged_solve"] = solve
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
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel_ms": round(float(kms), 4), "algorithmic_bytes_per_launch": int(kbytes),
                     "event_ms": {k: round(float(v), 4) for k, v in events.items()}},
        "cpu_baseline": cpu,
    }
    model.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

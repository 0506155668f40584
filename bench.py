"""PT-Stokes benchmark: T_eff [GB/s] of the PT iteration on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

A step is one PT iteration (advance_iteration!) over the whole grid, with
all fields resident in HBM.  value = T_eff = A_eff * K / t with
A_eff = 23 * sizeof(T) * nx*ny*nz bytes per iteration
(scripts_future_API/tm_stokes_mpi_wip.jl:180).  The roofline object reports
the dominant kernel's algorithmic bytes per launch over its HIP-event
duration on the model's stream; cpu_baseline times the CPU oracle on a
bounded number of iterations of the same grid on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

import __graft_entry__ as ge

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (dims, description)
    "stokes3d_fp64_511x511x255": ((511, 511, 255),
                                  "3D isoviscous PT-Stokes, FP64, free-slip box, manufactured-solution forcing"),
    "stokes3d_fp64_255": ((255, 255, 255), "3D manufactured free-slip solution, FP64 (255^3)"),
}
DEFAULT = "stokes3d_fp64_511x511x255"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--copy-calib", action="store_true",
                   help="run a 1 GiB device copy first (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return p.parse_args()


def setup_model(pkg, dims):
    kw, _ = pkg.manufactured_3d(dims, origin=(-1.0, -1.0, -1.0), extent=(2.0, 2.0, 2.0))
    model = pkg.IsothermalFullStokesModel(**kw)
    model.viscosity.eta.set(1.0)
    model.viscosity.eta_next.set(1.0)
    model.apply_bc("eta")
    return model, kw


def cpu_baseline(pkg, model, dims, seconds):
    orc = ge.load_oracle()
    ref = orc.OracleModel(model.desc)
    for k in ("GX", "GY", "GZ"):
        ref.set(k, model.field(k).parent())
    ref.set("ETA", 1.0)
    ref.set("ETA_NEXT", 1.0)
    ref.advance(1)  # warm-up (first touch)
    n = 0
    t0 = time.perf_counter()
    while True:
        ref.advance(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 200:
            break
    cells = dims[0] * dims[1] * dims[2]
    teff = 23 * 8 * cells * n / el / 1e9
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    del ref
    return {"value": round(teff, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": "%d PT iterations of the same %dx%dx%d FP64 grid (CPU oracle, OpenMP, %.1f s)"
                      % (n, dims[0], dims[1], dims[2], el)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each solve is one device-resident "
              "domain); run with --gpus 1", file=sys.stderr)
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
    dims, desc = WORKLOADS[args.workload]
    cells = dims[0] * dims[1] * dims[2]
    model, kw = setup_model(pkg, dims)

    model.advance(args.warmup)
    model.synchronize()
    t0 = time.perf_counter()
    model.advance(args.steps)
    model.synchronize()
    el = time.perf_counter() - t0
    ms_per_step = el / args.steps * 1e3
    a_eff = 23 * 8 * cells
    teff = a_eff / (el / args.steps) / 1e9

    # per-kernel HIP-event timing on the model's stream
    import ctypes as C
    A = pkg._abi
    out = np.zeros(4)
    nprof = max(10, min(50, args.steps // 4))
    A.check(A.lib().fi_time_iterations(model._h, nprof, A.dptr(out)))
    it_ms, fused_ms, shell_ms, bc_ms = out
    lo, hi = (C.c_int * 3)(), (C.c_int * 3)()
    fused_on = A.lib().fi_fused_region(model._h, lo, hi)
    nfields = 25 if model.desc.gravity_is_field else 22
    if fused_on:
        # fused sweep: R sigma(7) V(3) eta [g(3)], W sigma(7) V(3) eta_next, once per interior cell
        icells = (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1)
        kname, kms, kbytes = "k_fused3", fused_ms, nfields * 8 * icells
    else:
        kname, kms, kbytes = "pointwise sweeps", shell_ms, (nfields + 11) * 8 * cells
    achieved = kbytes / (kms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pkg, model, dims, args.cpu_seconds)

    line = {
        "metric": "PT-Stokes T_eff GB/s (% HBM peak) at 511x511x255",
        "value": round(teff, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": desc, "grid": list(dims),
                   "step": "one PT iteration (stress update + BCs + velocity/viscosity update + BCs)",
                   "teff_pct_hbm_peak": round(100.0 * teff / HBM_PEAK_GBS, 2),
                   "iterations_per_s": round(1e3 / ms_per_step, 2)},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel_ms": round(kms, 4), "algorithmic_bytes_per_launch": kbytes,
                     "event_ms": {"iteration": round(it_ms, 4), "fused": round(fused_ms, 4),
                                  "shell": round(shell_ms, 4), "bc": round(bc_ms, 4)}},
        "cpu_baseline": cpu,
    }
    model.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()

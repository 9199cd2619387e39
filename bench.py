#!/usr/bin/env python3
"""Throughput of the MPS time step on one MI355X.

Workload (default): the 3-D incompressible dam break, closed 1.61 x 0.15 x
0.6 m box of 12 triangles, fluid block 0.6 x 0.15 x 0.3 m at l0 = 0.0025 m
(240 x 60 x 120 = 1,728,000 particles), Jacobi-PCG pressure solve to a
relative residual of 1e-8, dt = 2e-4 s.  Synthetic input: lattice at cell
centres with seeded +-1% l0 jitter; the initial state and the first pressure
solve are built before timing.

A step is one full MPS time step (buckets, viscosity, prediction, collision,
wall distance, PND/NPCD, free surface, PPE assembly + solve, pressure
gradient, wall terms, correction).  `value` is particle-steps per second over
exactly --steps steps with all state resident in HBM.

The JSON line also carries the HBM roofline of the dominant kernel (the PPE
SpMV: algorithmic bytes per launch / mean launch time from HIP events on the
library's stream) and the CPU oracle timed on this host on one step of the
same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def cpu_baseline(w, steps: int, threads: int):
    from oracle.oracle import Oracle
    o = Oracle(w.params, w.types, w.pos, w.vel, mesh=w.mesh, threads=threads)
    o.call("init")
    t0 = time.perf_counter()
    o.steps(steps)
    dt = time.perf_counter() - t0
    o.close()
    return w.n * steps / dt, dt


def copy_calibration(nbytes: int, reps: int = 3):
    """Device-to-device copies of a known size (the runtime's copy kernel),
    so that HBM counters read for it can scale the counters of other kernels."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (src, dst):
        if hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    for _ in range(reps):
        if hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) != 0:   # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}))
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--l0", type=float, default=0.0025)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="relative tolerance of the inner FP32 solve of the mixed PPE")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (counter calibration)")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU path is not part of this build",
              file=sys.stderr)
        return 2

    from polymps_amd import configs
    from polymps_amd.lib import Simulation

    w = configs.dam3d_inc(l0=args.l0, steps=args.steps)
    w.params.precision_mode = 1 if args.precision == "mixed" else 0
    if args.inner_tol is not None:
        w.params.mixed_inner_tol = args.inner_tol
    sim = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh, device=0)
    t_init = time.perf_counter()
    sim.call("init")
    t_init = time.perf_counter() - t_init
    if args.warmup > 0:
        sim.steps(args.warmup)
    sim.set_scalar("profile", 1)
    sim.sync()
    t0 = time.perf_counter()
    sim.steps(args.steps)
    sim.sync()
    elapsed = time.perf_counter() - t0
    n = int(sim.scalar("num_particles"))
    value = n * args.steps / elapsed
    launches = sim.scalar("spmv_launches")
    spmv_ms = sim.scalar("spmv_ms_total") / max(launches, 1)
    spmv_bytes = sim.scalar("spmv_bytes_total") / max(launches, 1)
    achieved = spmv_bytes / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else 0.0
    result = {
        "metric": "particle-steps/s at 1.73M 3D INC (PPE Jacobi-PCG to 1e-8)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded jittered lattice, generated 12-triangle box)",
        "config": {"workload": f"dam3d_inc l0={args.l0:g} m, {n} particles, 12-triangle closed box",
                   "particles": n, "precision": args.precision, "solver_tol": w.params.solver_tol,
                   "dt": w.params.time_step},
        "ppe": {"last_solve_ms": sim.scalar("ppe_ms"), "last_iterations": sim.scalar("solver_iter"),
                "last_outer_iterations": sim.scalar("outer_iterations"),
                "true_residual": sim.scalar("true_residual"), "spmv_launches": launches},
        "roofline": {"kernel": "kCgSpmv" if args.precision == "fp64" else "kMxSpmv", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": spmv_bytes, "mean_launch_ms": spmv_ms},
        "init_s": t_init,
    }
    sim.close()
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        v, dt = cpu_baseline(w, args.cpu_steps, threads)
        result["cpu_baseline"] = {"value": v, "unit": "particle-steps/s", "cores": threads, "kind": "port",
                                  "sample": f"{args.cpu_steps} step(s) of the same {w.n}-particle workload after "
                                            f"its initial pressure solve, {dt:.1f} s"}
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())

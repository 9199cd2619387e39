#!/usr/bin/env python3
"""Throughput of the MPS time step on one MI355X.

Workload (default): the 3-D incompressible dam break, closed 1.61 x 0.15 x
0.6 m box of 12 triangles, fluid block 0.6 x 0.15 x 0.3 m at l0 = 0.0025 m
(240 x 60 x 120 = 1,728,000 particles), pressure solve to a relative
residual of 1e-8, dt = 2e-4 s.  Default precision "mixed": FP32 matrix,
SpMV and Jacobi preconditioner inside an FP64 iterative refinement on the
true residual (FP64 dot products and positions); "fp64" is the all-FP64
Jacobi-PCG.  Synthetic input: lattice at cell centres with seeded +-1% l0
jitter; the initial state and the first pressure solve are built before
timing.

A step is one full MPS time step (buckets, viscosity, prediction, collision,
wall distance, PND/NPCD, free surface, PPE assembly + solve, pressure
gradient, wall terms, correction).  `value` is particle-steps per second over
exactly --steps steps with all state resident in HBM; the default, 300 steps
(0.06 s of flow), is the length of the 1.73M mixed-precision run, so the
number covers the accelerating flow whose pressure solves need more
iterations as it develops.

The JSON line also carries the HBM roofline of the dominant kernel (the PPE
SpMV: algorithmic bytes per launch / mean launch time from HIP events on the
library's stream, around every eighth SpMV launch of the timed steps), the
PPE comparison (ms/solve to 1e-8 of plain FP64
Jacobi-PCG, the mixed solve and the learned+FP32 solve on the same in-step
systems, with the pressure difference to FP64) and the CPU oracle timed on
this host on one step of the same workload.
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
# HBM bytes / algorithmic bytes of the SpMV per launch on this workload, from
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs calibrated on a 1 GiB copy
# (FETCH_SIZE x 2, WRITE_SIZE x 1): profiles/spmv_traffic_dam3d_inc_1728k.md
SPMV_TRAFFIC_RATIO = {"fp64": 1.122, "mixed": 1.124}


def host_cpu():
    """CPU model, sockets, cores and SMT of this host (lscpu), and the
    physical cores among the CPUs this process may run on."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "logical_cpus"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
        topo = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
        mine = os.sched_getaffinity(0)
        cores = {(int(c), int(sk)) for line in topo.splitlines() if line and not line.startswith("#")
                 for cpu, c, sk in [line.split(",")[:3]] if int(cpu) in mine}
        info["affinity_cpus"] = len(mine)
        info["affinity_physical_cores"] = len(cores)
    except Exception as e:   # lscpu missing: report what Python knows
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
        info["lscpu_error"] = str(e)
    if "threads_per_core" in info:
        info["smt"] = "on" if info["threads_per_core"] != "1" else "off"
    return info


# oracle operators of a polygon-wall incompressible step before its PPE
ORACLE_BEFORE_PPE = ["update_buckets", "viscosity", "add_gravity", "prediction", "check_out_domain", "collision",
                     "closest_point", "update_near_mesh", "wall_npcd", "pnd_neigh_npcd", "update_bc"]


def cpu_baseline(w, steps: int, threads: int, ppe: bool):
    """The oracle on this host: `steps` full steps with `threads` OpenMP
    threads, then (ppe) the next step's pressure system solved the
    reference's way -- serial assembly, serial CG (Eigen's CG is serial,
    src/MpsPressure.cpp:424-445) to 1e-8 and to machine epsilon (Eigen's
    default tolerance) -- and by an OpenMP CSR Jacobi-PCG to 1e-8, each from
    the previous pressure."""
    from oracle.oracle import Oracle
    o = Oracle(w.params, w.types, w.pos, w.vel, mesh=w.mesh, threads=threads)
    o.call("init")
    t0 = time.perf_counter()
    o.steps(steps)
    dt = time.perf_counter() - t0
    out = {"value": w.n * steps / dt, "seconds": dt}
    if ppe:
        for op in ORACLE_BEFORE_PPE:
            o.call(op)
        x0 = o.get("ppe_x")
        res = {}
        o.set_scalar("threads", 1)
        t = time.perf_counter()
        o.call("assemble_ppe")
        res["serial_assembly_ms"] = (time.perf_counter() - t) * 1e3
        for name, tol, par, thr in (("serial_cg_1e-8", 1e-8, 0, 1), ("serial_cg_eps", 0.0, 0, 1),
                                    ("openmp_cg_1e-8", 1e-8, 1, threads)):
            o.set("ppe_x", x0)
            o.set_scalar("solver_tol", tol)
            o.set_scalar("parallel_cg", par)
            o.set_scalar("threads", thr)
            o.call("solve_ppe")
            res[name + "_ms"] = o.scalar("solve_iter_ms")
            res[name + "_iterations"] = o.scalar("solver_iter")
            res["matrix_setup_ms"] = max(res.get("matrix_setup_ms", 0.0), o.scalar("solve_setup_ms"))   # first solve
        res["openmp_threads"] = threads
        res["note"] = ("*_cg_* times are the CG iterations alone; matrix_setup_ms is the serial build of the CSR "
                       "iteration matrix (the transpose, row-sorted), which stands for the reference's "
                       "setFromTriplets and its by-value matrix copy; the reference's assembly also fills 238 zero "
                       "triplets per row, which serial_assembly_ms leaves out")
        out["ppe"] = res
    o.close()
    return out


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


# operators of the time step before and after its PPE (the order of mps_step
# for polygon walls, incompressible, no shifting)
BEFORE_PPE = ["bucket_update", "viscosity", "add_gravity", "prediction", "check_out_domain", "build_neighbor_list",
              "collision", "mesh_closest_point", "mesh_update_near", "wall_npcd", "pnd_neigh_npcd", "update_bc"]
AFTER_PPE = ["correction_matrix", "press_gradient", "wall_press_gradient", "wall_viscosity", "correction",
             "check_out_domain", "check_out_domain"]


def ppe_compare(sim, solves: int, gnn_weights):
    """PPE ms/solve to the tolerance on the same in-step systems: each step
    runs up to its PPE and the system is solved (1) from the previous
    pressure -- the reference's initial guess -- by plain FP64 Jacobi-PCG
    (the reference's method), the mixed-precision solve and (with weights)
    the learned warm start + learned flexible-CG preconditioner in the mixed
    solve, then (2) from the extrapolated initial guess of the time step
    (ppe_warm_start) by the same three; the step then finishes with the FP64
    solve from the extrapolated guess.  Ratios are against (1)'s FP64 time."""
    import numpy as np
    names = ("fp64", "mixed", "learned", "fp64_extrap", "mixed_extrap", "learned_extrap")
    rec = {k: [] for k in names}
    its = {k: [] for k in rec}
    err = {k: 0.0 for k in names if k != "fp64"}
    sim.set_scalar("profile", 0)
    if gnn_weights is not None:
        sim.gnn_load(gnn_weights)

    def run(name, x, learned, prec, press):
        sim.set("ppe_x", x)
        sim.set_scalar("learned_mode", learned)
        sim.set_scalar("precision_mode", prec)
        sim.call("ppe_assemble")
        sim.call("ppe_solve")
        rec[name].append(sim.scalar("ppe_ms"))
        its[name].append(sim.scalar("solver_iter"))
        press[name] = sim.get("press")

    for _ in range(solves):
        for op in BEFORE_PPE:
            sim.call(op)
        x0 = sim.get("ppe_x")
        press = {}
        modes = [("fp64", 0, 0), ("mixed", 0, 1)] + ([("learned", 2, 1)] if gnn_weights is not None else [])
        for name, learned, prec in modes:
            run(name, x0, learned, prec, press)
        # the time step's own initial guess from the history of solutions
        sim.set("ppe_x", x0)
        sim.call("ppe_warm_start")
        xg = sim.get("ppe_x")
        for name, learned, prec in modes:
            run(name + "_extrap", xg, learned, prec, press)
        scale = max(np.abs(press["fp64"]).max(), 1e-30)
        for name in err:
            if name in press:
                err[name] = max(err[name], float(np.abs(press[name] - press["fp64"]).max() / scale))
        # finish the step with the FP64 solve from the extrapolated guess
        sim.set("ppe_x", xg)
        sim.set_scalar("learned_mode", 0)
        sim.set_scalar("precision_mode", 0)
        sim.call("ppe_assemble")
        sim.call("ppe_solve")
        for op in AFTER_PPE:
            sim.call(op)
    out = {"solves": solves, "tolerance": "relative residual 1e-8"}
    for name in rec:
        if rec[name]:
            out[f"{name}_ms"] = float(np.mean(rec[name]))
            out[f"{name}_iterations"] = float(np.mean(its[name]))
    for name in err:
        if rec[name]:
            out[f"{name}_vs_fp64_time"] = out[f"{name}_ms"] / out["fp64_ms"]
            out[f"{name}_press_rel_err_vs_fp64"] = err[name]
    if rec["learned"]:
        out["learned_weights"] = str(gnn_weights.name)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the run length of the workload, 300 for dam3d, 200 for channel4m)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["dam3d", "channel4m"], default="dam3d",
                    help="dam3d: the 1.73M dam break (the metric's workload); channel4m: 4M particles over a "
                         "32k-triangle heightmap bed with obstacles")
    ap.add_argument("--l0", type=float, default=None, help="particle spacing (default 0.0025 dam3d, 0.005 channel4m)")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--ppe-solves", type=int, default=3,
                    help="in-step PPE systems solved by every method for the PPE comparison (0: skip)")
    ap.add_argument("--gnn", type=Path, default=ROOT / "polymps_amd" / "data" / "gnn_dam3d.bin",
                    help="learned-operator weights for the learned+FP32 PPE comparison")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--inner-tol", type=float, default=1e-7,
                    help="relative tolerance of the inner FP32 solve of the mixed PPE (at 1e-7 each inner "
                         "solve runs to half the FP64 target, which one pass reaches in most steps)")
    ap.add_argument("--warm-start", type=int, choices=[0, 1, 2, 3], default=2,
                    help="initial guess of each PPE solve: 0 previous solution, 1 linear, 2 quadratic extrapolation")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-ppe", action="store_true",
                    help="skip the CPU pressure-solve timings (serial to 1e-8 and to epsilon, OpenMP to 1e-8)")
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

    if args.steps is None:
        args.steps = 300 if args.workload == "dam3d" else 200

    if args.workload == "dam3d":
        args.l0 = args.l0 or 0.0025
        w = configs.dam3d_inc(l0=args.l0, steps=args.steps)
        wname = f"dam3d_inc l0={args.l0:g} m, {{n}} particles, 12-triangle closed box"
    else:
        args.l0 = args.l0 or 0.005
        w = configs.channel4m(l0=args.l0, steps=args.steps)
        wname = (f"channel4m l0={args.l0:g} m, {{n}} particles, {w.info['triangles']}-triangle heightmap bed, "
                 "channel and 3 obstacles")
    w.params.precision_mode = 1 if args.precision == "mixed" else 0
    w.params.ppe_warm_start = args.warm_start
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
        "metric": "particle-steps/s at 1.73M 3D INC; PPE ms to 1e-8, learned+FP32 vs plain FP64 CG",
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
        "data": ("synthetic (seeded jittered lattice, generated 12-triangle box)" if args.workload == "dam3d" else
                 "synthetic (seeded jittered lattice, generated fractal heightmap bed, channel and obstacles)"),
        "config": {"workload": wname.format(n=n),
                   "particles": n, "precision": args.precision, "solver_tol": w.params.solver_tol,
                   "ppe_warm_start": args.warm_start, "mixed_inner_tol": w.params.mixed_inner_tol,
                   "dt": w.params.time_step},
        "sizes": {"particles": n, "nnz": sim.scalar("nnz"), "nnz_escaped": sim.scalar("nnz_escaped"),
                  "window_escapes": sim.scalar("window_escapes"), "window_blocks": sim.scalar("ppe_window_blocks"),
                  "blocks": (n + 255) // 256,
                  "list_entries": sim.scalar("list_entries"), "list_entries_small": sim.scalar("list_entries_small"),
                  "buckets": sim.params.num_buckets_xyz},
        "ppe": {"last_solve_ms": sim.scalar("ppe_ms"), "last_iterations": sim.scalar("solver_iter"),
                "last_outer_iterations": sim.scalar("outer_iterations"),
                "true_residual": sim.scalar("true_residual"), "spmv_launches_timed": launches},
        "roofline": {"kernel": "kCgSpmv" if args.precision == "fp64" else "kMxSpmv", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (SPMV_TRAFFIC_RATIO[args.precision] * spmv_bytes
                                 if args.workload == "dam3d" and args.l0 == 0.0025 else None),
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE ratio measured at l0=0.0025",
                     "bytes_per_launch": spmv_bytes, "mean_launch_ms": spmv_ms,
                     "timing": f"HIP events around every eighth SpMV launch of the timed steps ({int(launches)} launches)"},
        "init_s": t_init,
    }
    if args.ppe_solves > 0:
        result["ppe_compare"] = ppe_compare(sim, args.ppe_solves, args.gnn if args.gnn.exists() else None)
    sim.close()
    if args.workload == "channel4m" and not args.no_cpu_baseline:
        # the oracle's brute-force wall distance (32k faces per particle) makes
        # even one CPU step of this workload take many minutes
        result["cpu_baseline"] = None
        args.no_cpu_baseline = True
    if not args.no_cpu_baseline:
        host = host_cpu()
        # OpenMP threads: the physical cores this process may use, capped by OMP_NUM_THREADS when set
        threads = int(host.get("affinity_physical_cores") or host["affinity_cpus"])
        if os.environ.get("OMP_NUM_THREADS"):
            threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
        cb = cpu_baseline(w, args.cpu_steps, threads, ppe=not args.no_cpu_ppe)
        result["cpu_baseline"] = {"value": cb["value"], "unit": "particle-steps/s", "cores": threads, "kind": "port",
                                  "sample": f"{args.cpu_steps} step(s) of the same {w.n}-particle workload after "
                                            f"its initial pressure solve, {cb['seconds']:.1f} s, {threads} OpenMP "
                                            f"threads", "host": host}
        if "ppe" in cb:
            result["cpu_baseline"]["ppe"] = cb["ppe"]
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Train and evaluate the learned PPE operator on the GPU.

Training data are generated on line by the plain solver: runs of the 3-D dam
break in the 1.61 x 0.15 x 0.6 m box with the fluid column's width drawn
from U[0.4, 0.8] m and its height from U[0.2, 0.4] m (rounded to the
particle spacing) and a seeded position jitter.  Every step of a run
assembles a PPE system and solves it with the Jacobi-PCG; the network then
takes one Adam step on that system (its loss is |A M v - v|^2 / |v|^2 for the
step's residual b - A x and a seeded random vector).  The forward and
backward passes are the library's own HIP kernels.

Evaluation runs held-out seeds and reports, per solve, the iterations and
time to the relative residual tolerance of (a) the mixed-precision
Jacobi-PCG, (b) the same with the learned warm start, (c) learned warm
start + learned preconditioner in flexible CG, and (d) FP64 Jacobi-PCG, all
from the same states.

    python -m polymps_amd.train_gnn --out polymps_amd/data/gnn_dam3d.bin
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

from polymps_amd import configs  # noqa: E402
from polymps_amd.lib import Simulation  # noqa: E402


def column_workload(seed: int, l0: float, steps: int):
    """Config-2 geometry with a random column (width U[0.4,0.8], height U[0.2,0.4])."""
    u = configs.splitmix64_uniform(10_000 + seed, 2)
    wdt = round((0.4 + 0.4 * u[0]) / l0) * l0
    hgt = round((0.2 + 0.2 * u[1]) / l0) * l0
    return configs.dam3d_inc(l0=l0, block=(wdt, 0.15, hgt), seed=seed, steps=steps)


def train(out: Path, runs: int, steps_per_run: int, l0: float, lr: float, seed0: int = 0, log=print):
    theta = None
    history = []
    for run in range(runs):
        w = column_workload(seed0 + run, l0, steps_per_run)
        w.params.precision_mode = 1
        g = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh, device=0)
        if theta is None:
            g.gnn_init(seed0)
        else:
            g.gnn_set_params(theta)
        g.call("init")
        t0 = time.perf_counter()
        losses = []
        for k in range(steps_per_run):
            g.call("step")
            losses.append(g.gnn_train_step(lr, 1_000_003 * (run + 1) + k))
        theta = g.gnn_params()
        history.append(float(np.mean(losses[-10:])))
        log(json.dumps({"run": run, "particles": w.n, "loss_first": losses[0], "loss_last10": history[-1],
                        "seconds": time.perf_counter() - t0}))
        g.close()
    g = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh, device=0)
    g.gnn_set_params(theta)
    out.parent.mkdir(parents=True, exist_ok=True)
    g.gnn_save(out)
    g.close()
    return history


def solve_stats(g: Simulation, x0, learned: int, precision: int):
    g.set("ppe_x", x0)
    g.set_scalar("learned_mode", learned)
    g.set_scalar("precision_mode", precision)
    g.call("ppe_assemble")
    g.call("ppe_solve")
    return dict(iterations=g.scalar("solver_iter"), ms=g.scalar("ppe_ms"), residual=g.scalar("true_residual"),
                fallback=g.scalar("fallback")), g.get("press")


# The time step of the polygon-wall incompressible case up to the PPE and
# after it (the order of mps_step).
_BEFORE_PPE = ["bucket_update", "viscosity", "add_gravity", "prediction", "check_out_domain", "build_neighbor_list",
               "collision", "mesh_closest_point", "mesh_update_near", "wall_npcd", "pnd_neigh_npcd", "update_bc"]
_AFTER_PPE = ["correction_matrix", "press_gradient", "wall_press_gradient", "wall_viscosity", "correction",
              "check_out_domain", "check_out_domain"]


def evaluate(weights: Path, seeds, l0: float, warm_steps: int, solves: int, log=print):
    """Per-solve comparison on held-out runs: each evaluated step is run up to
    its PPE, the same system is solved from the same initial guess (the
    previous pressure) by every method, and the step is finished with the
    plain solver's pressure.  Returns the list of records."""
    recs = []
    for sd in seeds:
        w = column_workload(sd, l0, warm_steps + solves)
        w.params.precision_mode = 1
        g = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh, device=0)
        g.gnn_load(weights)
        g.call("init")
        g.steps(warm_steps)
        for k in range(solves):
            for op in _BEFORE_PPE:
                g.call(op)
            x0 = g.get("ppe_x")
            base, p64 = solve_stats(g, x0, 0, 0)
            mixed, _ = solve_stats(g, x0, 0, 1)
            warm, _ = solve_stats(g, x0, 1, 1)
            full, pl = solve_stats(g, x0, 2, 1)
            scale = max(np.abs(p64).max(), 1e-30)
            rec = dict(seed=sd, solve=k, particles=w.n, fp64=base, mixed=mixed, learned_warm=warm,
                       learned_fcg=full, learned_press_err=float(np.abs(pl - p64).max() / scale))
            recs.append(rec)
            log(json.dumps(rec))
            g.set("ppe_x", x0)
            solve_stats(g, x0, 0, 0)
            for op in _AFTER_PPE:
                g.call(op)
        g.close()
    return recs


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", type=Path, default=ROOT / "polymps_amd" / "data" / "gnn_dam3d.bin")
    # 17 runs x 300 steps (the whole collapse and impact): 5,100 training systems
    ap.add_argument("--runs", type=int, default=17)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--l0", type=float, default=0.005)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-seeds", type=int, nargs="*", default=[901, 902])
    ap.add_argument("--eval-l0", type=float, default=0.005)
    ap.add_argument("--eval-only", action="store_true")
    args = ap.parse_args(argv)
    if not args.eval_only:
        train(args.out, args.runs, args.steps, args.l0, args.lr)
    recs = evaluate(args.out, args.eval_seeds, args.eval_l0, warm_steps=40, solves=3)
    summ = {k: float(np.mean([r[k]["ms"] for r in recs])) for k in ("fp64", "mixed", "learned_warm", "learned_fcg")}
    its = {k: float(np.mean([r[k]["iterations"] for r in recs])) for k in ("fp64", "mixed", "learned_warm", "learned_fcg")}
    print(json.dumps({"mean_ms": summ, "mean_iterations": its,
                      "max_learned_press_err": max(r["learned_press_err"] for r in recs)}))
    return 0


if __name__ == "__main__":
    sys.exit(main())

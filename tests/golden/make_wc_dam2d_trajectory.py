"""Write tests/golden/wc_dam2d_6332_trajectory.npz from the CPU oracle.

The 2-D weakly compressible dam break at its own size (50 x 100 fluid at
l0 = 0.005 m in a 1.0 x 0.6 m tank, 1 wall + 2 dummy layers: 6,332
particles; large radius 3.1 l0 and corrected gradient, configs.dam2d_wc_long)
over its whole 2,000 steps: the oracle's live count, positions
and velocities after 1,000 and 2,000 steps (regression data: the oracle is
pinned by tests/test_oracle_pinning.py and the other pin files).  The oracle
takes a few minutes for this run, so the device test compares against the
stored trajectory instead of running it.

    python tests/golden/make_wc_dam2d_trajectory.py [--threads T]
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from polymps_amd import configs  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

OUT = HERE / "wc_dam2d_6332_trajectory.npz"
CHECKPOINTS = (1000, 2000)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    args = ap.parse_args()
    w = configs.dam2d_wc_long()
    o = Oracle(w.params, w.types, w.pos, w.vel, mesh=w.mesh, threads=args.threads)
    o.call("init")
    out = {"n": np.int64(w.n)}
    done = 0
    for c in CHECKPOINTS:
        o.steps(c - done)
        done = c
        n = int(o.scalar("num_particles"))
        out[f"live_{c}"] = np.int64(n)
        out[f"pos_{c}"] = o.get("pos")
        out[f"vel_{c}"] = o.get("vel")
        print(f"step {c}: {n} live, vel max {np.abs(out[f'vel_{c}']).max():.4f}", flush=True)
    np.savez_compressed(OUT, **out)
    o.close()
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()

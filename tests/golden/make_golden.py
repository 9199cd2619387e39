"""Write the golden fixtures of this directory from the CPU oracle.

The reference ships no test vectors, so the fixtures are outputs of the
oracle (the C++ restatement of the reference's algorithm under oracle/) on
small seeded workloads, stored as .npz (no pickles).  The oracle itself is
pinned by closed forms, lattice sums and brute force in
tests/test_oracle_pinning.py; these fixtures freeze its outputs so that the
device path and later versions of the oracle are compared with fixed data.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from polymps_amd import configs  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

# (fixture name, workload generator, steps after the initial step)
CASES = [
    ("dam3d_inc_l0_0.025", lambda: configs.small_dam3d(l0=0.025), 3),
    ("dam2d_wc_l0_0.02", lambda: configs.small_dam2d(l0=0.02), 5),
]
FIELDS_INIT = ["pndi", "num_neigh", "bc", "press"]
FIELDS_STEP = ["pos", "vel", "press", "bc", "num_neigh", "pndi"]


def make(name, gen, steps):
    w = gen()
    o = Oracle(w.params, w.types, w.pos, w.vel, mesh=w.mesh, threads=1)
    o.call("init")
    out = {f"init_{f}": o.get(f) for f in FIELDS_INIT}
    o.steps(steps)
    out.update({f"step_{f}": o.get(f) for f in FIELDS_STEP})
    out["steps"] = np.array(steps)
    out["n"] = np.array(w.n)
    out["input_pos"] = w.pos
    out["input_types"] = w.types
    o.close()
    np.savez_compressed(HERE / f"{name}.npz", **out)
    print(name, w.n, "particles")


if __name__ == "__main__":
    for c in CASES:
        make(*c)

"""The device path against the committed golden fixtures (needs a GPU).

Integer fields and the PND are bit-exact; the PPE pressure differs by the
reduction order of the solver's dot products (1e-6 of the largest
pressure), and positions after the steps by what that difference moves
(1e-6 l0 for the incompressible case); the weakly-compressible case has no
solve and its positions agree to 1e-9 l0 (device pow in the equation of
state).
"""
from pathlib import Path

import numpy as np
import pytest

from polymps_amd import configs
from polymps_amd.lib import Simulation

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = {
    "dam3d_inc_l0_0.025": (lambda: configs.small_dam3d(l0=0.025), 1e-6),
    "dam2d_wc_l0_0.02": (lambda: configs.small_dam2d(l0=0.02), 1e-9),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_matches_golden(built, name):
    make, pos_tol = CASES[name]
    g = np.load(GOLDEN / f"{name}.npz")
    w = make()
    print(f"golden case {name}", flush=True)
    s = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh, device=0)
    s.call("init")
    for f in ("num_neigh", "bc", "pndi"):
        assert np.array_equal(s.get(f), g[f"init_{f}"]), f
    scale = max(np.abs(g["init_press"]).max(), 1e-30)
    assert np.max(np.abs(s.get("press") - g["init_press"])) <= 1e-6 * scale
    for _ in range(int(g["steps"])):
        s.call("step")
    l0 = w.params.part_dist
    assert np.max(np.abs(s.get("pos") - g["step_pos"])) <= pos_tol * l0
    assert np.array_equal(s.get("bc"), g["step_bc"])
    assert np.array_equal(s.get("num_neigh"), g["step_num_neigh"])
    scale = max(np.abs(g["step_press"]).max(), 1e-30)
    assert np.max(np.abs(s.get("press") - g["step_press"])) <= 1e-6 * scale

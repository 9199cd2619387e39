"""The committed golden fixtures (tests/golden/*.npz) against the oracle
and the workload generators, on the CPU."""
from pathlib import Path

import numpy as np
import pytest

from polymps_amd import configs

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = {
    "dam3d_inc_l0_0.025": lambda: configs.small_dam3d(l0=0.025),
    "dam2d_wc_l0_0.02": lambda: configs.small_dam2d(l0=0.02),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_generator_reproduces_fixture_inputs(name):
    g = np.load(GOLDEN / f"{name}.npz")
    w = CASES[name]()
    assert np.array_equal(g["input_pos"], w.pos)
    assert np.array_equal(g["input_types"], w.types)


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_fixture(built, name):
    from oracle.oracle import Oracle
    g = np.load(GOLDEN / f"{name}.npz")
    w = CASES[name]()
    o = Oracle(w.params, w.types, w.pos, w.vel, mesh=w.mesh, threads=1)
    o.call("init")
    for f in ("pndi", "num_neigh", "bc", "press"):
        assert np.array_equal(o.get(f), g[f"init_{f}"]), f
    o.steps(int(g["steps"]))
    for f in ("pos", "vel", "press", "bc", "num_neigh", "pndi"):
        assert np.array_equal(o.get(f), g[f"step_{f}"]), f
    o.close()


def test_oracle_reproduces_wc_trajectory(built):
    """The stored 2,000-step trajectory of the 6,332-particle WC dam break
    (the device test's reference): the oracle reproduces its first
    checkpoint bit for bit, with another thread count than the one that
    wrote it."""
    from oracle.oracle import Oracle
    g = np.load(GOLDEN / "wc_dam2d_6332_trajectory.npz")
    w = configs.dam2d_wc_long()
    assert w.n == int(g["n"])
    o = Oracle(w.params, w.types, w.pos, w.vel, mesh=w.mesh, threads=4)
    o.call("init")
    o.steps(1000)
    assert int(o.scalar("num_particles")) == int(g["live_1000"])
    assert np.array_equal(o.get("pos"), g["pos_1000"])
    assert np.array_equal(o.get("vel"), g["vel_1000"])
    o.close()

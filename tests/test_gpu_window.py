"""LDS windows of the CG solves' SpMVs (needs a gfx950 GPU).

With the windows on (scalar ppe_window 1, the default) every 256-row block
whose columns fit its window (the union of its rows' stencil-row slot
ranges) copies that part of the vector into LDS and gathers from it (the
mixed solve's FP32 SpMV, its FP64 true residual and the FP64 CG's SpMV); the
other blocks gather from global memory, as every block does with the windows
off.  Both sum the same products in the same order, so whole runs must agree
bit for bit, and at the benchmark's density almost every block must take
its window.
"""
import numpy as np
import pytest

from polymps_amd import configs
from polymps_amd.lib import MpsError, Simulation
from tests.parity_util import PERSISTENT

pytestmark = pytest.mark.gpu


def _pair(w):
    on = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh)
    off = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh)
    off.set_scalar("ppe_window", 0)
    assert on.scalar("ppe_window") == 1 and off.scalar("ppe_window") == 0
    return on, off


def _same(on, off, tag):
    for f in PERSISTENT:
        a, b = on.get(f), off.get(f)
        if not np.array_equal(a, b):
            d = np.nonzero(a != b)[0]
            raise AssertionError(f"{tag}: {f} differs at {len(d)} entries, first {d[:5]}")


@pytest.mark.parametrize("l0,steps,precision", [(0.0125, 40, 1), (0.005, 12, 1), (0.0125, 20, 0)])
def test_window_runs_are_bit_identical(built, l0, steps, precision):
    w = configs.dam3d_inc(l0=l0) if l0 == 0.005 else configs.small_dam3d(l0=l0)
    w.params.precision_mode = precision
    w.params.ppe_warm_start = 2
    print(f"window on/off: {w.n} particles, precision {precision}, {steps} steps", flush=True)
    on, off = _pair(w)
    on.call("init")
    off.call("init")
    nb = (w.n + 255) // 256
    for s in range(steps):
        on.call("step")
        off.call("step")
        if s % 4 == 3 or s == steps - 1:
            _same(on, off, f"step {s + 1}")
    used = on.scalar("ppe_window_blocks")
    print(f"  blocks with a window: {used:.0f} of {nb}", flush=True)
    assert off.scalar("ppe_window_blocks") == 0
    assert used >= 0.8 * nb
    assert on.scalar("solver_iter") == off.scalar("solver_iter")
    on.close()
    off.close()


def test_window_periodic_domain_uses_global_gathers(built):
    """Periodic domains keep every block on global gathers (the columns of
    rows at the seam are far apart in slot order)."""
    w = configs.periodic_channel3d_polygon(l0=0.02)
    w.params.precision_mode = 1
    on, off = _pair(w)
    on.call("init")
    off.call("init")
    for _ in range(5):
        on.call("step")
        off.call("step")
    assert on.scalar("ppe_window_blocks") == 0
    _same(on, off, "periodic")
    on.close()
    off.close()


def test_window_scalar_rejects_other_values(built):
    w = configs.small_dam3d(l0=0.025)
    g = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh)
    with pytest.raises(MpsError):
        g.set_scalar("ppe_window", 2)
    with pytest.raises(MpsError):
        g.set_scalar("ppe_window_blocks", 1)
    g.close()

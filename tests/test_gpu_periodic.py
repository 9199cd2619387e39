"""Periodic domains on the GPU against the oracle (needs a GPU).

The reference's periodic boundary (domain.boundary.type 1, one direction,
particle walls): ghost bucket layers that list the opposite interior layer
with a periodic-length shift, and the wrap of positions in
checkParticleOutDomain.  Per-operator lockstep as in test_gpu_parity.py,
before and after particles have crossed the seam: every integer and FP64
field bit-exact, the PPE solution within 1e-6 of the largest pressure.
"""
import numpy as np
import pytest

from polymps_amd import configs
from tests import test_gpu_parity as tp
from tests.parity_util import lockstep_step, new_pair, sync_state

pytestmark = pytest.mark.gpu

CASES = {
    "2d_wc": (lambda: configs.periodic_channel2d(u0=2.0), 0, 80),
    "2d_inc": (lambda: configs.periodic_channel2d(u0=2.0, mps_type=2), 0, 80),
    "3d_inc": (lambda: configs.periodic_channel3d(u0=2.5), 1, 40),
    "3d_inc_polygon": (lambda: configs.periodic_channel3d_polygon(u0=2.5), 1, 40),
}


@pytest.mark.parametrize("case", list(CASES))
def test_lockstep_periodic(built, case):
    mk, axis, advance = CASES[case]
    w = mk()
    print(f"periodic lockstep {w.name}", flush=True)
    g, o = new_pair(w)
    tp.lockstep_init(g, o)
    for s in range(2):
        lockstep_step(g, o, s)
    o.steps(advance)
    sync_state(g, o)
    n = int(o.scalar("num_particles"))
    x0 = w.pos.reshape(-1, 3)[:n, axis]
    crossed = np.abs(o.get("pos").reshape(-1, 3)[:n, axis] - x0) > 0.5 * o.params.periodic_length[axis]
    assert crossed.any()
    for s in range(3):
        lockstep_step(g, o, 10 + s)


def test_periodic_full_steps_track_oracle(built):
    """mps_step end to end on the 2-D weakly compressible channel: particles
    wrap on both paths and the trajectories agree."""
    w = configs.periodic_channel2d(u0=2.0)
    print(f"periodic full steps {w.name}", flush=True)
    g, o = new_pair(w)
    g.call("init")
    o.call("init")
    for _ in range(100):
        g.call("step")
        o.call("step")
    q = o.params
    pg, po = g.get("pos").reshape(-1, 3), o.get("pos").reshape(-1, 3)
    assert np.max(np.abs(pg - po)) < 1e-9
    assert pg[:, 0].min() >= q.phys_min[0] and pg[:, 0].max() <= q.phys_max[0]
    assert np.array_equal(g.get("bc"), o.get("bc"))

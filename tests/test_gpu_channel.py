"""The heightmap channel (a large polygon wall: fractal bed, channel sides,
box obstacles) on the device (needs a GPU).

The closest-point search over a mesh of more than 64 faces goes through the
triangle BVH; it must return exactly what the brute-force search over every
face returns (same distance, point, mirror point and face normal, with the
lowest face index among equally close faces).  The whole step then stays
bit-exact with the oracle (which searches every face).
"""
import numpy as np
import pytest

from polymps_amd import configs
from polymps_amd.lib import Simulation
from tests.parity_util import lockstep_step, new_pair, sync_state

pytestmark = pytest.mark.gpu


def test_bvh_matches_brute_force(built):
    w = configs.channel4m(l0=0.02)
    print(f"channel bvh {w.n} particles {w.info['triangles']} triangles", flush=True)
    g = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh, device=0)
    g.call("init")
    g.steps(3)
    fields = ["riw2", "mirror", "wall_pos", "polygon_normal"]
    out = {}
    for mode in (0, 1):
        g.set("riw2", np.full(w.n, 10e8 * w.params.part_dist))
        g.set_scalar("mesh_bvh", mode)
        g.call("mesh_closest_point")
        out[mode] = {f: g.get(f) for f in fields}
    for f in fields:
        assert np.array_equal(out[0][f], out[1][f]), f
    assert np.all(out[1]["riw2"] < 1.0)


def test_channel_lockstep_with_oracle(built):
    """A coarse channel (64 x 16 heightmap, 1,890 + 46 triangles) in lockstep
    with the oracle for two steps after a short developed run."""
    w = configs.channel4m(l0=0.05, nhx=64, nhy=16)
    print(f"channel lockstep {w.n} particles {w.info['triangles']} triangles", flush=True)
    w.params.precision_mode = 0
    g, o = new_pair(w)
    o.call("init")
    o.steps(10)
    sync_state(g, o)
    for s in range(2):
        lockstep_step(g, o, s)

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


def test_reference_terrain_closest_point(built):
    """The reference's own 46,215-triangle terrain mesh and parameter file
    (input/stl/BRUMADINHO_space10_model_lucas_top_0p50.stl,
    input/BRUMADINHO_space10_lo10p00.json, kept in tests/golden) with 4,000
    seeded query points over its bounding box: the BVH search, the
    brute-force search and the oracle's brute force agree bit for bit."""
    from pathlib import Path

    from oracle.oracle import Oracle
    from polymps_amd import lib as L
    data = Path(__file__).resolve().parent / "golden" / "reference_input"
    V, F, N = L.read_stl(data / "stl" / "BRUMADINHO_space10_model_lucas_top_0p50.stl")
    p, _, _ = L.params_from_json(data / "BRUMADINHO_space10_lo10p00.json")
    v = V.reshape(-1, 3)
    lo, hi = v.min(0), v.max(0)
    rng = np.random.default_rng(7)
    n = 4000
    pos = np.column_stack([rng.uniform(lo[0], hi[0], n), rng.uniform(lo[1], hi[1], n),
                           rng.uniform(lo[2], hi[2] + 50.0, n)]).reshape(-1)
    types = np.zeros(n, dtype=np.int32)
    print(f"terrain closest point: {len(F) // 3} triangles, {n} points", flush=True)
    g = Simulation(p, types, pos, np.zeros(3 * n), mesh=(V, F, N), device=0)
    o = Oracle(p, types, pos, np.zeros(3 * n), mesh=(V, F, N))
    far = np.full(n, 10e8 * p.part_dist)
    fields = ["riw2", "mirror", "wall_pos", "polygon_normal"]
    out = {}
    for mode in (0, 1):
        g.set("riw2", far)
        g.set_scalar("mesh_bvh", mode)
        g.call("mesh_closest_point")
        out[mode] = {f: g.get(f) for f in fields}
    o.set("riw2", far)
    o.call("closest_point")
    for f in fields:
        assert np.array_equal(out[0][f], out[1][f]), f
        assert np.array_equal(out[1][f], o.get(f)), f
    assert np.all(out[1]["riw2"] < 500.0 ** 2)
    g.close()
    o.close()

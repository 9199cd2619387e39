"""Three wall meshes: rigid tank, static slab (deformable slot) and a forced
paddle (needs a gfx950 GPU).

The closest-point search visits the meshes in type order
(src/main.cpp:296-300, src/PolygonMesh.cpp:615-650) and the forced mesh moves
before the corrector by updateForcedPolygonMesh (src/PolygonMesh.cpp:173-374).
Per-operator lockstep with the oracle, bit for bit (the PPE solution within
the solver tolerance), including the nearest mesh type and the moved nodes
and normals; the translation law also exercises the forced wall's velocity
in the no-slip and slip wall terms.
"""
import numpy as np
import pytest

from polymps_amd import configs
from tests.parity_util import check, lockstep_step, new_pair
from tests.parity_util import lockstep_init

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rotation,slip", [(1, 0), (0, 0), (0, 1)])
def test_forced_mesh_lockstep(built, rotation, slip):
    w = configs.forced_paddle3d(l0=0.025, rotation=rotation, velocity=-0.5)
    w.params.slip_condition = slip
    print(f"forced mesh lockstep rotation={rotation} slip={slip}", flush=True)
    g, o = new_pair(w)
    lockstep_init(g, o)
    check(g, o, ["near_mesh_type"], "init")
    kinds = set()
    for s in range(25):
        lockstep_step(g, o, s, time=s * w.params.time_step)
        kinds |= set(np.unique(g.get("near_mesh_type")).tolist())
    assert kinds == {0, 1, 2}
    V0 = configs.forced_paddle3d(l0=0.025).extra_meshes[2][0]
    assert np.abs(g.get_mesh(2)[0] - V0).max() > 1e-3      # the paddle has moved


@pytest.mark.parametrize("bvh", [0, 1])
def test_forced_mesh_full_steps(built, bvh):
    """mps_step (the library's own step order) against the oracle's step:
    identical mesh nodes every step, states within the solver tolerance."""
    w = configs.forced_paddle3d(l0=0.025)
    g, o = new_pair(w)
    g.set_scalar("mesh_bvh", bvh)
    g.call("init")
    o.call("init")
    for s in range(15):
        g.call("step")
        o.call("step")
        for a, b in zip(g.get_mesh(2), o.get_mesh(2)):
            assert np.array_equal(a, b), f"step {s}"
    assert np.array_equal(g.get("near_mesh_type"), o.get("near_mesh_type"))
    assert np.max(np.abs(g.get("pos") - o.get("pos"))) < 1e-6 * w.params.part_dist

"""Checkpoint and bit-exact restart (needs a gfx950 GPU).

A run of 2k steps and a run of k steps, a checkpoint, a new context that
reads it and k more steps must end in identical states, bit for bit: the
checkpoint carries every field a step reads from the step before (among them
the PPE solution that warm-starts the next solve, pressurePPE of
include/MpsParticle.h:205, and the extrapolation history), time, the velocity
maximum the wall force reads, the solve count and the moving meshes.
"""
import numpy as np
import pytest

from polymps_amd import configs
from polymps_amd.lib import MpsError, Simulation
from tests.parity_util import PERSISTENT, PERSISTENT_GRANULAR

pytestmark = pytest.mark.gpu


def _sim(w):
    return Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh, meshes=getattr(w, "extra_meshes", None))


def _leaving(w, k=6):
    """Add k particles outside the tank that leave the domain (compaction)."""
    extra = np.array([[-0.002 - 0.0007 * q, 0.05 + 0.01 * q, 0.1] for q in range(k)])
    w.pos = np.concatenate([w.pos.reshape(-1, 3), extra]).reshape(-1)
    w.vel = np.concatenate([w.vel.reshape(-1, 3), np.tile([-0.5, 0.0, 0.0], (k, 1))]).reshape(-1)
    w.types = np.concatenate([w.types, np.zeros(k, dtype=w.types.dtype) + w.types[0]])
    return w


def _case(name):
    if name == "dam3d_mixed_extrapolated":
        w = configs.small_dam3d(l0=0.0125, dt=5e-4)
        w.params.precision_mode = 1
        w.params.ppe_warm_start = 2
        return w, 15
    if name == "dam2d_wc":
        return configs.small_dam2d(l0=0.02), 30
    if name == "forced_paddle":
        return configs.forced_paddle3d(l0=0.025), 10
    if name == "granular2d":
        return configs.granular_dam2d(l0=0.02, mps_type=3), 15
    if name == "compaction":
        return _leaving(configs.small_dam3d(l0=0.0125)), 12
    raise KeyError(name)


@pytest.mark.parametrize("name", ["dam3d_mixed_extrapolated", "dam2d_wc", "forced_paddle", "granular2d",
                                  "compaction"])
def test_restart_is_bit_exact(built, tmp_path, name):
    w, k = _case(name)
    print(f"restart {name}: {k} + {k} steps", flush=True)
    fields = PERSISTENT + ["type"] + (PERSISTENT_GRANULAR if w.params.fluid_type == 1 else [])
    whole = _sim(w)
    whole.call("init")
    whole.steps(2 * k)
    first = _sim(w)
    first.call("init")
    first.steps(k)
    ck = tmp_path / "state.ckpt"
    first.checkpoint_write(ck)
    n_at_ck = first.scalar("num_particles")
    first.close()
    second = _sim(w)
    second.checkpoint_read(ck)
    assert second.scalar("num_particles") == n_at_ck
    assert second.scalar("time") == pytest.approx(k * w.params.time_step, rel=1e-12)
    second.steps(k)
    for f in fields:
        a, b = whole.get(f), second.get(f)
        if not np.array_equal(a, b):
            d = np.nonzero(a != b)[0]
            raise AssertionError(f"{name}: field {f} differs at {len(d)} entries, first {d[:5]}")
    assert whole.scalar("time") == second.scalar("time")
    assert whole.scalar("num_particles") == second.scalar("num_particles")
    if name == "forced_paddle":
        for a, b in zip(whole.get_mesh(2), second.get_mesh(2)):
            assert np.array_equal(a, b)
    if name == "compaction":
        assert whole.scalar("num_particles") < w.n
    whole.close()
    second.close()


def test_checkpoint_errors_leave_the_context_unchanged(built, tmp_path):
    w = configs.small_dam3d(l0=0.025)
    g = _sim(w)
    g.call("init")
    g.steps(3)
    ck = tmp_path / "a.ckpt"
    g.checkpoint_write(ck)
    pos = g.get("pos").copy()
    # not a checkpoint
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(b"not a checkpoint file at all" * 10)
    with pytest.raises(MpsError):
        g.checkpoint_read(bad)
    # truncated
    cut = tmp_path / "cut.ckpt"
    cut.write_bytes(ck.read_bytes()[:-100])
    with pytest.raises(MpsError):
        g.checkpoint_read(cut)
    with pytest.raises(MpsError):
        g.checkpoint_read(tmp_path / "missing.ckpt")
    assert np.array_equal(g.get("pos"), pos)
    # a context of another capacity
    w2 = configs.small_dam3d(l0=0.025)
    keep = np.ones(w2.n, dtype=bool)
    keep[-5:] = False
    small = Simulation(w2.params, w2.types[keep], w2.pos.reshape(-1, 3)[keep].reshape(-1),
                       w2.vel.reshape(-1, 3)[keep].reshape(-1), mesh=w2.mesh)
    with pytest.raises(MpsError):
        small.checkpoint_read(ck)
    # the good file still restores
    h = _sim(w)
    h.checkpoint_read(ck)
    assert np.array_equal(h.get("pos"), pos)
    for s in (g, small, h):
        s.close()

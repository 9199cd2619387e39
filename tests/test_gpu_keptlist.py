"""Neighbour list kept over the time step (needs a gfx950 GPU).

A step can build its list once, after the bucket sort, at the list radius
plus a skin, and use it for the viscosity at the old positions and for every
operator after the prediction.  Each consumer tests its own radius and the
list keeps slot order, so the kept list must give bit-identical states to the
list built after the prediction (list_skin 0), whether the step keeps it,
patches the particles whose stencil centre moved, or rebuilds it because a
particle moved farther than half the skin or the compaction moved slots.
"""
import numpy as np
import pytest

from polymps_amd import configs
from polymps_amd.lib import Simulation

pytestmark = pytest.mark.gpu

FIELDS = ["pos", "vel", "press", "pndi", "pnd_small", "bc", "num_neigh", "npcd", "ppe_x", "force_wall"]


def _state(g):
    return {f: g.get(f).copy() for f in FIELDS}


def _assert_same(a, b, where):
    for f in FIELDS:
        if not np.array_equal(a[f], b[f]):
            d = np.nonzero(a[f] != b[f])[0]
            raise AssertionError(f"{where}: field {f} differs at {len(d)} entries, first {d[:5]}")


def _pair(w, skin):
    plain = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh)
    kept = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh)
    plain.set_scalar("list_skin", 0)
    kept.set_scalar("list_skin", skin)
    # an automatic skin up to 0.25 l0 keeps the list in every step of this flow
    kept.set_scalar("list_skin_cap", 0.25)
    plain.call("init")
    kept.call("init")
    return plain, kept


@pytest.mark.parametrize("skin", [-1.0, 1e-4, 0.2])
def test_kept_list_matches_plain(built, skin):
    """Automatic skin (kept, a few particles patched per step), a skin so thin
    that most steps rebuild, and a wide fixed skin."""
    w = configs.small_dam3d(l0=0.0125, dt=5e-4)
    plain, kept = _pair(w, skin)
    print(f"kept list skin {skin}", flush=True)
    patched = rebuilt = 0
    for s in range(120):
        plain.call("step")
        kept.call("step")
        assert kept.scalar("list_skin") > 0
        patched += int(kept.scalar("list_patched"))
        rebuilt += int(kept.scalar("list_rebuilt"))
        if s % 20 == 19:
            _assert_same(_state(plain), _state(kept), f"step {s + 1}")
    print(f"patched {patched} rebuilt {rebuilt} steps of 120", flush=True)
    assert float(np.abs(plain.get("vel")).max()) > 0.1   # the flow has developed
    if skin == -1.0 or skin == 0.2:
        assert patched > 0
        assert rebuilt < 10
    else:
        assert rebuilt > 60
    plain.close()
    kept.close()


def test_kept_list_with_compaction(built):
    """Particles outside the tank moving out of the domain: the compaction
    moves slots and the kept list must be rebuilt in those steps."""
    w = configs.small_dam3d(l0=0.0125)
    extra = np.array([[-0.002 - 0.0007 * k, 0.05 + 0.01 * k, 0.1] for k in range(6)])
    pos = np.concatenate([w.pos.reshape(-1, 3), extra]).reshape(-1)
    vel = np.concatenate([w.vel.reshape(-1, 3), np.tile([-0.5, 0.0, 0.0], (6, 1))]).reshape(-1)
    types = np.concatenate([w.types, np.zeros(6, dtype=w.types.dtype) + w.types[0]])
    w.pos, w.vel, w.types = pos, vel, types
    plain, kept = _pair(w, 0.2)
    n0 = int(kept.scalar("num_particles"))
    compactions = 0
    for s in range(40):
        plain.call("step")
        kept.call("step")
        n = int(kept.scalar("num_particles"))
        assert n == int(plain.scalar("num_particles"))
        if n < n0:
            assert kept.scalar("list_rebuilt") == 1, f"step {s + 1}"
            compactions += 1
            n0 = n
        _assert_same(_state(plain), _state(kept), f"step {s + 1}")
    assert compactions >= 1
    plain.close()
    kept.close()


def test_overflowing_candidate_rows_match(built):
    """Stencil rows with more candidates than the bitmask window holds (one
    32-slot word here, so nearly every particle) are listed by a direct
    walk: the states stay bit-identical to the default windows."""
    w = configs.small_dam3d(l0=0.0125, dt=5e-4)
    wide = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh)
    narrow = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh)
    narrow.set_scalar("mask_words", 1)
    wide.call("init")
    narrow.call("init")
    for s in range(30):
        wide.call("step")
        narrow.call("step")
        if s % 10 == 9:
            _assert_same(_state(wide), _state(narrow), f"step {s + 1}")
    wide.close()
    narrow.close()


def test_skin_rule_settings_match_plain(built):
    """The automatic skin's cap and factor only move work between the kept
    list, its tightening after the prediction and the second build: a low cap
    (most steps build after the prediction) and the smallest factor (the
    refresh rebuilds whenever a particle outran half the skin) both give the
    plain list's states."""
    from polymps_amd.lib import MpsError
    w = configs.small_dam3d(l0=0.0125, dt=5e-4)
    plain = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh)
    plain.set_scalar("list_skin", 0)
    plain.call("init")
    sims = []
    for cap, factor in ((0.05, 3.0), (0.5, 2.0)):
        s = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh)
        s.set_scalar("list_skin_cap", cap)
        s.set_scalar("list_skin_factor", factor)
        assert s.scalar("list_skin_cap") == cap and s.scalar("list_skin_factor") == factor
        s.call("init")
        sims.append(s)
    for bad in (("list_skin_cap", -0.1), ("list_skin_cap", 1.5), ("list_skin_factor", 1.0)):
        with pytest.raises(MpsError):
            sims[0].set_scalar(*bad)
    kept = [0, 0]
    for step in range(80):
        plain.call("step")
        for k, s in enumerate(sims):
            s.call("step")
            kept[k] += s.scalar("list_skin") > 0
        if step % 20 == 19:
            ref = _state(plain)
            for s in sims:
                _assert_same(ref, _state(s), f"step {step + 1}")
    print(f"steps with a kept list: {kept}", flush=True)
    assert kept[1] > kept[0]
    for s in sims + [plain]:
        s.close()


@pytest.mark.parametrize("relax", [1.0, 0.5])
def test_start_list_matches_plain(built, relax):
    """A flow faster than the automatic skin's cap (cap 0 here): each step
    builds a list at its start positions for the operators before the
    prediction (the viscosity; with relax_press < 1 also the predicted
    pressure gradient, which reads the small-radius list) instead of
    walking the buckets, then the usual list after the prediction.  The
    states are bit-identical to the plain list's (bucket-walk viscosity)."""
    w = configs.small_dam3d(l0=0.0125, dt=5e-4)
    w.params.relax_press = relax
    plain = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh)
    fast = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh)
    plain.set_scalar("list_skin", 0)
    fast.set_scalar("list_skin_cap", 0.0)
    plain.call("init")
    fast.call("init")
    print(f"start list, relax_press {relax}", flush=True)
    for s in range(60):
        plain.call("step")
        fast.call("step")
        assert fast.scalar("list_start") == 1 and plain.scalar("list_start") == 0
        assert fast.scalar("list_skin") == 0
        if s % 20 == 19:
            _assert_same(_state(plain), _state(fast), f"step {s + 1}")
    assert float(np.abs(plain.get("vel")).max()) > 0.1
    plain.close()
    fast.close()

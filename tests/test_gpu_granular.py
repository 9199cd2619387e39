"""Granular (non-Newtonian) phase in lockstep with the oracle (needs a gfx950
GPU).

Cases after the reference's granular inputs: a 2-D dam break over a mobile
bed with particle walls, incompressible (input/2D_dam_fraccarollo_*_INC.json)
and weakly compressible (*_WC.json), and a 3-D granular bed in the polygon
box with no-slip and free-slip walls (input/damErosion3D_WC.json).  Every
operator of a step is compared with the oracle: the volume fraction, the
bed levels (through the normal stress), the strain rate, the smoothed
pressure, the inertial number, the grain volume fraction and the density
bit for bit; the viscosities MEU and MEU_Y, whose mu(I) law calls sin, cos,
tan, exp and pow, to 1e-12 relative (device library against glibc), after
which the oracle's values are handed to the device.  A free run of both
then stays within a stated drift bound.
"""
import numpy as np
import pytest

from polymps_amd import configs
from tests.parity_util import lockstep_init, lockstep_step, new_pair, sync_state

pytestmark = pytest.mark.gpu


def _case(mk):
    if mk == "2d_inc":
        return configs.granular_dam2d(l0=0.01)
    if mk == "2d_wc":
        return configs.granular_dam2d(l0=0.01, mps_type=1)
    return configs.granular_dam3d(l0=0.025, slip=0 if mk == "3d_noslip" else 1)


@pytest.mark.parametrize("mk", ["2d_inc", "2d_wc", "3d_noslip", "3d_slip"])
def test_granular_lockstep(built, mk):
    w = _case(mk)
    print(f"granular lockstep {mk}: {w.n} particles, {w.info.get('granular')} granular", flush=True)
    g, o = new_pair(w)
    lockstep_init(g, o)
    for s in range(3):
        lockstep_step(g, o, s)
    o.steps(40)
    sync_state(g, o)
    for s in range(2):
        lockstep_step(g, o, 50 + s)
    assert np.array_equal(g.get("ptype"), o.get("ptype"))
    assert float(np.abs(g.get("vel")).max()) > 0.0


@pytest.mark.parametrize("mk", ["2d_inc", "3d_noslip"])
def test_granular_free_run_drift(built, mk):
    """Both sides run 60 steps on their own: the viscosities differ in the
    last place (device libm against glibc), and the positions stay within
    1e-9 l0 of each other."""
    w = _case(mk)
    g, o = new_pair(w)
    g.call("init")
    o.call("init")
    g.steps(60)
    o.steps(60)
    l0 = w.params.part_dist
    d = np.abs(g.get("pos") - o.get("pos")).max()
    print(f"granular drift {mk}: {d / l0:.2e} l0", flush=True)
    assert d < 1e-9 * l0
    assert np.array_equal(g.get("bc"), o.get("bc"))
    assert abs(g.scalar("cfl_visc") - o.scalar("cfl_visc")) <= 1e-12 * max(o.scalar("cfl_visc"), 1e-30)

"""Extrapolated initial guess of the pressure solve (needs a gfx950 GPU).

With ppe_warm_start 1 or 2 each solve starts from the linear or quadratic
extrapolation of the last pressure solutions instead of the last one alone
(the reference's solveWithGuess, src/MpsPressure.cpp:433).  The solve still
runs to the same relative residual, so the trajectory stays within the
solver tolerance of the oracle's, while the solves take fewer iterations.
"""
import numpy as np
import pytest

from polymps_amd import configs
from polymps_amd.lib import Simulation
from tests.parity_util import new_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("precision", [0, 1])
def test_warm_start_tracks_oracle(built, order, precision):
    w = configs.small_dam3d(l0=0.025)
    w.params.precision_mode = precision
    w.params.ppe_warm_start = order
    g, o = new_pair(w)
    g.call("init")
    o.call("init")
    for _ in range(12):
        g.call("step")
        o.call("step")
        assert g.scalar("true_residual") < w.params.solver_tol
    assert np.max(np.abs(g.get("pos") - o.get("pos"))) < 1e-6 * w.params.part_dist
    assert np.array_equal(g.get("bc"), o.get("bc"))
    po = o.get("press")
    assert np.max(np.abs(g.get("press") - po)) < 1e-5 * np.abs(po).max()


def test_warm_start_cuts_iterations(built):
    """Same developed flow, 20 steps from each initial guess."""
    w = configs.small_dam3d(l0=0.0125)
    its = {}
    press = {}
    for order in (0, 1, 2):
        w.params.ppe_warm_start = order
        g = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh)
        g.call("init")
        g.steps(40)
        total = 0
        for _ in range(20):
            g.call("step")
            total += int(g.scalar("solver_iter"))
            assert g.scalar("true_residual") < w.params.solver_tol
        its[order] = total
        press[order] = g.get("press")
        g.close()
    print("PCG iterations over 20 steps by warm-start order:", its)
    assert its[1] < its[0] and its[2] < its[0]
    scale = np.abs(press[0]).max()
    for order in (1, 2):
        assert np.max(np.abs(press[order] - press[0])) < 1e-5 * scale

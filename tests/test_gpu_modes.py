"""Option matrix on the GPU against the oracle (needs a GPU): each case
changes one or two options of the small 3-D (polygon walls, incompressible)
or 2-D (particle walls, weakly compressible) dam break and runs the
per-operator lockstep from the initial state and again after the oracle has
advanced, bit-exact except the PPE solution (1e-6 of the largest pressure)."""
import pytest

from polymps_amd import configs
from tests.parity_util import lockstep_init, lockstep_step, new_pair, sync_state

pytestmark = pytest.mark.gpu

CASES = {
    # 3-D, polygon walls, PND source PPE
    "3d_gradient1": ("3d", {"gradient_type": 1}),
    "3d_gradient2": ("3d", {"gradient_type": 2}),
    "3d_gradient3": ("3d", {"gradient_type": 3}),
    "3d_weight1": ("3d", {"weight_type": 1}),
    "3d_weight2": ("3d", {"weight_type": 2}),
    "3d_weight3": ("3d", {"weight_type": 3}),
    "3d_weight4": ("3d", {"weight_type": 4}),
    "3d_dpc": ("3d", {"collision_type": 1}),
    "3d_free_surface_neigh": ("3d", {"free_surf_type": 0}),
    "3d_free_slip": ("3d", {"slip_condition": 1}),
    "3d_shifting1": ("3d", {"shifting_type": 1}),
    "3d_shifting2": ("3d", {"shifting_type": 2}),
    "3d_weakly_diffusive": ("3d", {"mps_type": 1, "pnd_type": 2}),
    "3d_weakly_diffusive_slip": ("3d", {"mps_type": 1, "pnd_type": 2, "slip_condition": 1}),
    "3d_weakly_mean_pnd": ("3d", {"mps_type": 1, "pnd_type": 1}),
    "3d_explicit": ("3d", {"mps_type": 0, "pnd_type": 1}),
    # 2-D, particle walls, weakly compressible
    "2d_explicit": ("2d", {"mps_type": 0}),
    "2d_mean_pnd": ("2d", {"pnd_type": 1}),
    "2d_shifting1": ("2d", {"shifting_type": 1}),
    "2d_dpc": ("2d", {"collision_type": 1}),
    "2d_gradient3_corrected": ("2d", {"gradient_type": 3, "gradient_correction": 1}),
    "2d_free_surface_npcd": ("2d", {"free_surf_type": 1}),
}


def make(case):
    base, over = CASES[case]
    w = configs.small_dam3d(l0=0.025) if base == "3d" else configs.small_dam2d(l0=0.02)
    for k, v in over.items():
        setattr(w.params, k, v)
    return w


@pytest.mark.parametrize("case", list(CASES))
def test_option_lockstep(built, case):
    w = make(case)
    print(f"option case {case}", flush=True)
    g, o = new_pair(w)
    lockstep_init(g, o)
    lockstep_step(g, o, 0)
    o.steps(30)
    sync_state(g, o)
    lockstep_step(g, o, 10)

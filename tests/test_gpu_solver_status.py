"""Pressure-solve outcome and call-order checks through the C ABI (needs a
gfx950 GPU).

The reference prints "Failed solving" when Eigen's CG ends above its
tolerance and carries on with the pressure it has (src/MpsPressure.cpp:
429-444).  Here the solve writes the pressure the same way and returns
MPS_E_SOLVER (Python: MpsSolverError); a step finishes first, and
mps_steps runs all its steps before it reports the failure.  A non-finite
residual is reported as well and ends a multi-step call at once.  The
failures are forced from the host (an iteration limit of 2, a NaN in the
initial guess), so every kernel still runs to its normal exit.
"""
import numpy as np
import pytest

from polymps_amd import configs
from polymps_amd.lib import MpsError, MpsSolverError, Simulation

pytestmark = pytest.mark.gpu

# operators of a polygon-wall incompressible step before its PPE
BEFORE_PPE = ["bucket_update", "viscosity", "add_gravity", "prediction", "check_out_domain", "build_neighbor_list",
              "collision", "mesh_closest_point", "mesh_update_near", "wall_npcd", "pnd_neigh_npcd", "update_bc"]


def mid_step(precision, max_iter=0, steps=5):
    w = configs.small_dam3d(l0=0.025)
    w.params.precision_mode = precision
    w.params.solver_max_iter = max_iter
    g = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh, device=0)
    for run in (lambda: g.call("init"), lambda: g.steps(steps)):
        try:
            run()
        except MpsSolverError:
            if not max_iter:   # only an iteration limit may stop a solve early here
                raise
    for op in BEFORE_PPE:
        g.call(op)
    return w, g


@pytest.mark.parametrize("precision", [0, 1])
def test_iteration_limit_reports_failure(built, precision):
    print(f"solver status: iteration limit, precision {precision}", flush=True)
    w, g = mid_step(precision)
    g.set_scalar("solver_failures", 0)
    # a converged solve first: status 0
    g.call("ppe_assemble")
    g.call("ppe_solve")
    assert g.scalar("solver_failed") == 0 and g.scalar("solver_failures") == 0
    assert g.scalar("solver_iter") > 2
    # an iteration limit of 2 (a parameter of the context, so a second one)
    _, h = mid_step(precision, max_iter=2)
    h.set_scalar("solver_failures", 0)
    h.call("ppe_assemble")
    with pytest.raises(MpsSolverError, match="Failed solving"):
        h.call("ppe_solve")
    assert h.scalar("solver_failed") == 1
    assert h.scalar("solver_failures") == 1
    assert h.scalar("solver_iter") <= 2 + (1 if precision else 0)
    assert h.scalar("true_residual") > w.params.solver_tol
    # the pressure is written anyway, as max(x, 0) (setZeroOnNegativePressure)
    x = h.get("ppe_x")
    assert np.array_equal(h.get("press"), np.where(x < 0, 0.0, x))
    assert np.all(np.isfinite(x))


def test_steps_finish_through_soft_failures(built):
    print("solver status: mps_steps with an iteration limit of 2", flush=True)
    w = configs.small_dam3d(l0=0.025)
    w.params.solver_max_iter = 2
    g = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh, device=0)
    with pytest.raises(MpsSolverError):
        g.call("init")
    it0 = g.scalar("iterations")
    with pytest.raises(MpsSolverError, match="step"):
        g.steps(3)
    # all three steps ran, each with a failed solve
    assert g.scalar("iterations") == it0 + 3
    assert g.scalar("solver_failures") == 4
    assert np.all(np.isfinite(g.get("pos")))


@pytest.mark.parametrize("precision", [0, 1])
def test_non_finite_guess_reports_failure(built, precision):
    print(f"solver status: NaN initial guess, precision {precision}", flush=True)
    w, g = mid_step(precision)
    x = g.get("ppe_x")
    fluid = np.flatnonzero(g.get("type") == 0)
    x[fluid[len(fluid) // 2]] = np.nan
    g.set("ppe_x", x)
    g.call("ppe_assemble")
    with pytest.raises(MpsSolverError, match="non-finite"):
        g.call("ppe_solve")
    assert g.scalar("solver_failed") == 2
    # a finite guess afterwards solves normally again
    x[np.isnan(x)] = 0.0
    g.set("ppe_x", x)
    g.call("ppe_assemble")
    g.call("ppe_solve")
    assert g.scalar("solver_failed") == 0
    assert g.scalar("true_residual") < w.params.solver_tol


def test_solve_needs_matching_assembly(built):
    print("solver status: call order", flush=True)
    w, g = mid_step(1)
    # a bucket update moves the slots: the last assembly no longer applies
    g.call("ppe_assemble")
    g.call("bucket_update")
    with pytest.raises(MpsError, match="needs a PPE assembly"):
        g.call("ppe_solve")
    # assembled for the mixed solve, then the precision changes
    g.call("ppe_assemble")
    g.set_scalar("precision_mode", 0)
    with pytest.raises(MpsError, match="assembled for mixed-precision CG"):
        g.call("ppe_solve")
    g.call("ppe_assemble")
    g.call("ppe_solve")
    assert g.scalar("true_residual") < w.params.solver_tol


@pytest.mark.parametrize("precision", [0, 1])
def test_first_residual_survives_other_operators(built, precision):
    """The assembly hands the solve its first residual b - A x0 and the
    block partials of |r|^2, |b|^2.  An operator with a reduction of its own
    (the corrector's velocity maximum) run between the assembly and the
    solve must not change the solve: the pressure is bit-identical to the
    solve that follows the assembly directly."""
    print(f"solver status: operator between assembly and solve, precision {precision}", flush=True)
    _, a = mid_step(precision)
    _, b = mid_step(precision)
    a.call("ppe_assemble")
    a.call("ppe_solve")
    b.call("ppe_assemble")
    b.call("correction")
    b.call("ppe_solve")
    assert a.scalar("solver_iter") == b.scalar("solver_iter")
    assert np.array_equal(a.get("press"), b.get("press"))

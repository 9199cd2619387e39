"""The learned PPE operator on the device (needs a gfx950 GPU).

Forward and backward passes are compared with the CPU reference
(oracle/gnn_oracle.py) on the graph, features and diagonal the device built,
with the reference rounding the same values to bfloat16 as the device.
Differences come from FP32 versus FP64 accumulation and the occasional
activation landing on the other side of a bfloat16 rounding boundary, so
the tolerances are a few bfloat16 ulps; a wrong lane layout or index gives
errors of order one.
"""
import numpy as np
import pytest

from oracle import gnn_oracle as G
from polymps_amd import configs
from polymps_amd.lib import Simulation

pytestmark = pytest.mark.gpu


def developed(l0=0.0125, steps=20, precision=1):
    w = configs.small_dam3d(l0=l0)
    w.params.precision_mode = precision
    g = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh, device=0)
    g.call("init")
    g.steps(steps)
    return w, g


def ref_graph(d):
    return G.Graph(d["indptr"], d["cols"], d["e"], d["w"])


def random_theta(seed):
    rng = np.random.default_rng(seed)
    th = G.init_params(seed)
    o = G.offsets()
    for k in ("v", "vd", "bd", "b0", "p") + tuple(f"c{l}" for l in range(3)) + tuple(f"d{l}" for l in range(3)):
        a, s = o[k]
        n = int(np.prod(s))
        th[a:a + n] = rng.normal(scale=0.2, size=n)
    return th


def test_initial_operator_is_jacobi(built):
    w, g = developed(steps=5)
    print("gnn jacobi start", flush=True)
    g.gnn_init(7)
    g.gnn_forward()
    d = g.gnn_graph()
    assert d["indptr"][-1] > 10 * len(d["diag"])    # about 32 edges per particle
    assert np.all(d["cs"] == 0.0)
    assert np.allclose(d["dsi"], 1.0 / d["diag"], rtol=1e-6, atol=0)


def test_forward_matches_reference(built):
    w, g = developed(steps=5)
    print("gnn forward", flush=True)
    th = random_theta(3)
    g.gnn_set_params(th)
    g.gnn_forward()
    d = g.gnn_graph()
    gr = ref_graph(d)
    cs, dsi, _ = G.forward(th, gr, d["feat"], d["diag"], rb=True)
    scale = np.abs(cs).max()
    assert scale > 0
    err = np.abs(d["cs"] - cs) / scale
    assert np.quantile(err, 0.999) < 2e-2 and err.mean() < 2e-3, (err.max(), err.mean())
    rel = np.abs(d["dsi"] - dsi) / np.abs(dsi)
    assert np.quantile(rel, 0.999) < 2e-2 and rel.mean() < 2e-3, (rel.max(), rel.mean())


def test_backward_matches_reference(built):
    w, g = developed(steps=5)
    print("gnn backward", flush=True)
    th = random_theta(4)
    g.gnn_set_params(th)
    g.gnn_forward()
    d = g.gnn_graph()
    gr = ref_graph(d)
    rng = np.random.default_rng(11)
    n = len(d["diag"])
    r = rng.normal(size=n)
    gz = rng.normal(size=n)
    grad = g.gnn_backward(r, gz)
    _, _, cache = G.forward(th, gr, d["feat"], d["diag"], rb=True)
    # the device holds r and gz in FP32
    ref = G.backward(th, gr, cache, r.astype(np.float32).astype(np.float64),
                     gz.astype(np.float32).astype(np.float64), rb=True)
    assert np.linalg.norm(grad - ref) < 3e-2 * np.linalg.norm(ref)
    for k, (a, s) in G.offsets().items():
        m = int(np.prod(s))
        gk, rk = grad[a:a + m], ref[a:a + m]
        nr = np.linalg.norm(rk)
        if nr > 1e-6 * np.linalg.norm(ref):
            assert np.linalg.norm(gk - rk) < 0.1 * nr, (k, np.linalg.norm(gk - rk) / nr)


def test_training_reduces_loss(built):
    w, g = developed(steps=5)
    print("gnn training", flush=True)
    g.gnn_init(1)
    losses = []
    for k in range(40):
        g.call("step")
        losses.append(g.gnn_train_step(2e-3, 1000 + k))
    first, last = np.mean(losses[:5]), np.mean(losses[-5:])
    assert np.all(np.isfinite(losses))
    assert last < 0.9 * first, (first, last)


@pytest.mark.parametrize("mode", [1, 2])
def test_learned_solve_reaches_tolerance(built, mode):
    w, g = developed(steps=10)
    print(f"gnn learned solve mode {mode}", flush=True)
    g.gnn_init(2)
    for k in range(20):
        g.call("step")
        g.gnn_train_step(2e-3, 2000 + k)
    x0 = g.get("ppe_x")
    g.set_scalar("learned_mode", mode)
    g.call("ppe_assemble")
    g.call("ppe_solve")
    pl = g.get("press")
    assert g.scalar("true_residual") < w.params.solver_tol
    g.set("ppe_x", x0)
    g.set_scalar("learned_mode", 0)
    g.set_scalar("precision_mode", 0)
    g.call("ppe_assemble")
    g.call("ppe_solve")
    pf = g.get("press")
    scale = np.abs(pf).max()
    assert np.max(np.abs(pl - pf)) <= 1e-6 * scale


def test_trained_network_gain(built):
    """The committed network (polymps_amd/data/gnn_dam3d.bin, trained by
    polymps_amd/train_gnn.py on 5,100 systems of columns of seeds 0-16)
    against the Jacobi operator it starts from, on a held-out column (seed
    950): its objective |A M v - v|^2 / |v|^2 is lower on the same states and
    vectors, the flexible CG it preconditions needs fewer iterations than
    the Jacobi PCG on the same system (to a 10x smaller residual) without
    falling back, and the learned steps reach the tolerance."""
    from pathlib import Path
    from polymps_amd.train_gnn import column_workload
    weights = Path(__file__).resolve().parent.parent / "polymps_amd" / "data" / "gnn_dam3d.bin"
    w = column_workload(950, 0.005, 40)
    w.params.precision_mode = 1
    g = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh, device=0)
    print(f"gnn gain, {w.n} particles", flush=True)
    g.call("init")
    g.steps(30)
    trained, jacobi = [], []
    for k in range(3):
        g.call("step")
        g.gnn_load(weights)
        trained.append(g.gnn_train_step(0.0, 77 + k))
        g.gnn_init(0)
        jacobi.append(g.gnn_train_step(0.0, 77 + k))
    print(f"objective trained {np.mean(trained):.4g} jacobi {np.mean(jacobi):.4g}", flush=True)
    assert np.mean(trained) < 0.5 * np.mean(jacobi)
    g.gnn_load(weights)
    x0 = g.get("ppe_x")
    g.set_scalar("learned_mode", 0)
    g.call("ppe_assemble")
    g.call("ppe_solve")
    jacobi_it = g.scalar("solver_iter")
    _learned_solve(g, x0)
    learned_it = g.scalar("solver_iter")
    print(f"iterations: Jacobi PCG {jacobi_it:.0f}, learned flexible CG {learned_it:.0f}", flush=True)
    assert g.scalar("fallback") == 0
    assert learned_it < jacobi_it
    g.set("ppe_x", x0)
    g.call("step")
    assert g.scalar("true_residual") < w.params.solver_tol


def _fp64_reference(g, x0):
    g.set("ppe_x", x0)
    g.set_scalar("learned_mode", 0)
    g.set_scalar("precision_mode", 0)
    g.call("ppe_assemble")
    g.call("ppe_solve")
    return g.get("press")


def _learned_solve(g, x0):
    g.set("ppe_x", x0)
    g.set_scalar("learned_mode", 2)
    g.set_scalar("precision_mode", 1)
    g.call("ppe_assemble")
    g.call("ppe_solve")
    return g.get("press")


def test_stall_hands_the_solve_to_jacobi(built):
    """A refinement pass of the learned flexible CG that leaves more than
    learned_stall_ratio of the true residual drops the learned
    preconditioner for Jacobi.  Forced from the host: a coarse inner
    tolerance (1e-3) from a zero initial guess makes two passes necessary
    and a ratio of 1e-12 makes the first one a stall; the solve still
    reaches the tolerance, reports the fallback and agrees with the FP64
    solve of the same system."""
    w = configs.small_dam3d(l0=0.0125)
    w.params.precision_mode = 1
    w.params.mixed_inner_tol = 1e-3
    g = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh, device=0)
    g.call("init")
    g.steps(10)
    print("gnn forced stall", flush=True)
    g.gnn_init(5)
    x0 = np.zeros_like(g.get("ppe_x"))
    g.set_scalar("learned_stall_ratio", 1e-12)
    assert g.scalar("learned_stall_ratio") == 1e-12
    pl = _learned_solve(g, x0)
    assert g.scalar("fallback") == 1
    assert g.scalar("outer_iterations") >= 2
    assert g.scalar("true_residual") < w.params.solver_tol
    pf = _fp64_reference(g, x0)
    assert np.max(np.abs(pl - pf)) <= 1e-6 * np.abs(pf).max()
    # the default ratio keeps the (Jacobi-initialised) learned operator
    g.set_scalar("learned_stall_ratio", 0.5)
    _learned_solve(g, x0)
    assert g.scalar("fallback") == 0
    assert g.scalar("true_residual") < w.params.solver_tol
    with pytest.raises(Exception):
        g.set_scalar("learned_stall_ratio", 1.0)
    g.close()


@pytest.mark.parametrize("scale", [1.0, 20.0])
def test_random_learned_operator_still_converges(built, scale):
    """Untrained random weights (output weights scaled up) make a poor, in
    general not positive, preconditioner: the breakdown test or the stall
    test hand the solve to Jacobi and it still reaches the tolerance with the
    FP64 solution's pressure."""
    w, g = developed(steps=10)
    th = random_theta(21)
    o = G.offsets()
    for k in ("v", "vd"):
        a, s = o[k]
        n = int(np.prod(s))
        th[a:a + n] *= scale
    g.gnn_set_params(th)
    x0 = g.get("ppe_x")
    pl = _learned_solve(g, x0)
    print(f"random operator x{scale}: fallback {g.scalar('fallback'):.0f}, "
          f"passes {g.scalar('outer_iterations'):.0f}", flush=True)
    assert g.scalar("true_residual") < w.params.solver_tol
    pf = _fp64_reference(g, x0)
    assert np.max(np.abs(pl - pf)) <= 1e-6 * np.abs(pf).max()
    g.close()

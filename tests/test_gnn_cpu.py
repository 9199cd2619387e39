"""The CPU reference of the learned PPE operator (oracle/gnn_oracle.py):
its hand-written backward pass against finite differences in FP64, and the
Jacobi start."""
import numpy as np

from oracle import gnn_oracle as G


def small_case(seed=0, n_side=4):
    rng = np.random.default_rng(seed)
    l0 = 1.0
    ix = np.stack(np.meshgrid(*[np.arange(n_side)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(float)
    pos = (ix + 0.5 + rng.uniform(-0.05, 0.05, ix.shape)) * l0
    g = G.graph_from_positions(pos, 2.1 * l0, l0)
    n = len(pos)
    f = rng.normal(size=(n, 4))
    diag = -rng.uniform(1.0, 2.0, n)
    return g, f, diag, rng


def test_initial_operator_is_jacobi():
    g, f, diag, rng = small_case()
    theta = G.init_params(3)
    cs, dsi, _ = G.forward(theta, g, f, diag)
    assert np.all(cs == 0.0)
    assert np.allclose(dsi, 1.0 / diag, rtol=0, atol=0)
    r = rng.normal(size=g.n)
    assert np.allclose(G.apply(g, cs, dsi, r), r / diag)


def test_backward_matches_finite_differences():
    g, f, diag, rng = small_case(1)
    theta = G.init_params(5)
    o = G.offsets()
    for k in ("v", "vd", "bd"):   # switch on the output heads
        a, s = o[k]
        theta[a:a + int(np.prod(s))] = rng.normal(scale=0.3, size=int(np.prod(s)))
    r = rng.normal(size=g.n)
    gz = rng.normal(size=g.n)

    def S(th):
        cs, dsi, _ = G.forward(th, g, f, diag)
        return float(gz @ G.apply(g, cs, dsi, r))

    cs, dsi, cache = G.forward(theta, g, f, diag)
    grad = G.backward(theta, g, cache, r, gz)
    idx = rng.choice(G.NPARAM, size=300, replace=False)
    idx = np.union1d(idx, [o[k][0] for k in o])   # at least one entry of every tensor
    eps = 1e-6
    fd = np.array([(S(theta + eps * np.eye(1, G.NPARAM, k)[0]) - S(theta - eps * np.eye(1, G.NPARAM, k)[0])) / (2 * eps)
                   for k in idx])
    scale = np.abs(grad).max()
    assert scale > 0
    err = np.abs(fd - grad[idx]) / scale
    # relu kinks can make an isolated difference quotient straddle a kink
    assert np.quantile(err, 0.99) < 1e-6, np.sort(err)[-5:]


def test_loss_gradient_chain():
    """dL/dz of L = |A z - r|^2 / |r|^2 against finite differences."""
    rng = np.random.default_rng(2)
    n = 30
    A = rng.normal(size=(n, n)) + n * np.eye(n)
    r = rng.normal(size=n)
    z = rng.normal(size=n)
    L, gz = G.loss_and_gz(A, z, r)
    eps = 1e-7
    for k in range(n):
        e = np.zeros(n); e[k] = eps
        fd = (G.loss_and_gz(A, z + e, r)[0] - G.loss_and_gz(A, z - e, r)[0]) / (2 * eps)
        assert abs(fd - gz[k]) < 1e-6 * max(1.0, abs(gz).max())

"""CPU reference of the learned PPE operator (test infrastructure).

The learned component has no counterpart in the reference code; this file is
the numpy statement of the network the device kernels implement
(polymps_amd/csrc/gnn.hip), used only by tests to check the device forward
and backward passes.  Its own backward pass is pinned by finite differences
in FP64 (tests/test_gnn_cpu.py).

Network (H = 32 hidden channels, 3 message-passing layers, Q = 16 edge-head
channels) on a graph whose edges (i <- j) are the neighbours within reS:

    h      = relu(W0 f_i + b0)                                    node encoder
    layer: G_j = B h_j
           q_ij = G_j + E e_ij + c
           m_i = sum_j om_ij relu(q_ij),   om_ij = w_ij / sum_k w_ik
           a_i = U [h_i; m_i] + d
           h_i = h_i + relu(a_i)
    head:  t_ij = relu(Pa h_i + Pb h_j + Pe e_ij + p)
           c_ij = v . t_ij
           s_i  = exp(clip(vd . h_i + bd, -3, 3))
    operator:  (M r)_i = s_i r_i / a_ii + sum_j c_ij r_j / sqrt(a_ii a_jj)

With v = 0, vd = 0, bd = 0 (the initial values) M is the Jacobi
preconditioner.  With bf16=True the values the device stores in bfloat16
(GEMM weights B, U, Pa|Pb and the activations h, G, m, Pa h|Pb h, and the
GEMM inputs of the backward data products) are rounded the same way.
"""
from __future__ import annotations

import numpy as np

H = 32
Q = 16
LAYERS = 3

# flat parameter layout (the device uses the same order)
SHAPES = [("W0", (H, 4)), ("b0", (H,))]
for _l in range(LAYERS):
    SHAPES += [(f"B{_l}", (H, H)), (f"E{_l}", (H, 4)), (f"c{_l}", (H,)), (f"U{_l}", (H, 2 * H)), (f"d{_l}", (H,))]
SHAPES += [("Pa", (Q, H)), ("Pb", (Q, H)), ("Pe", (Q, 4)), ("p", (Q,)), ("v", (Q,)), ("vd", (H,)), ("bd", (1,))]
NPARAM = int(sum(np.prod(s) for _, s in SHAPES))


def offsets():
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = (o, shp)
        o += n
    return out


def unpack(theta):
    return {k: np.asarray(theta[o:o + int(np.prod(s))], dtype=np.float64).reshape(s) for k, (o, s) in offsets().items()}


def pack(P):
    return np.concatenate([np.asarray(P[k], dtype=np.float64).ravel() for k, _ in SHAPES])


def init_params(seed: int = 0) -> np.ndarray:
    """Uniform +-sqrt(6/(fan_in+fan_out)) for matrices, zero biases, and zero
    output weights (v, vd, bd) so that the operator starts as Jacobi.  The
    device initialiser draws the same numbers from the same splitmix64
    stream."""
    theta = np.zeros(NPARAM)
    u = splitmix64_uniform(seed, NPARAM)
    for k, (o, s) in offsets().items():
        n = int(np.prod(s))
        if len(s) == 2 and k not in ("v", "vd"):
            lim = np.sqrt(6.0 / (s[0] + s[1]))
            theta[o:o + n] = (2.0 * u[o:o + n] - 1.0) * lim
    return theta


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def bf16(x):
    """Round to bfloat16 (nearest even) and back."""
    a = np.asarray(x, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def relu(x):
    return np.maximum(x, 0.0)


class Graph:
    """CSR graph: edges (i <- cols[k]) for k in indptr[i]:indptr[i+1]."""

    def __init__(self, indptr, cols, e, w):
        self.indptr = np.asarray(indptr, dtype=np.int64)
        self.cols = np.asarray(cols, dtype=np.int64)
        self.e = np.asarray(e, dtype=np.float64)      # [E, 4]
        self.w = np.asarray(w, dtype=np.float64)      # [E]
        self.n = len(self.indptr) - 1
        self.rows = np.repeat(np.arange(self.n), np.diff(self.indptr))
        ws = np.zeros(self.n)
        np.add.at(ws, self.rows, self.w)
        self.wsum = ws
        self.om = self.w / np.where(ws[self.rows] > 0, ws[self.rows], 1.0)


def forward(theta, g: Graph, f, diag, rb=False):
    """Returns (cs, dsi, cache): the operator is (M r)_i = dsi_i r_i + sum cs_e r_cols[e]."""
    P = unpack(theta)
    R = bf16 if rb else (lambda x: np.asarray(x, dtype=np.float64))
    cache = {"f": f}
    pre0 = f @ P["W0"].T + P["b0"]
    h = R(relu(pre0))
    cache["pre0"], cache["h0"] = pre0, h
    for l in range(LAYERS):
        G = R(h @ R(P[f"B{l}"]).T)
        q = G[g.cols] + g.e @ P[f"E{l}"].T + P[f"c{l}"]
        m = np.zeros((g.n, H))
        np.add.at(m, g.rows, g.om[:, None] * relu(q))
        m = R(m)
        x = np.concatenate([h, m], axis=1)
        a = x @ R(P[f"U{l}"]).T + P[f"d{l}"]
        hn = R(h + relu(a))
        cache[f"G{l}"], cache[f"q{l}"], cache[f"m{l}"], cache[f"a{l}"], cache[f"x{l}"] = G, q, m, a, x
        h = hn
        cache[f"h{l + 1}"] = h
    Pab = np.concatenate([P["Pa"], P["Pb"]], axis=0)
    AB = R(h @ R(Pab).T)
    PA, PB = AB[:, :Q], AB[:, Q:]
    tpre = PA[g.rows] + PB[g.cols] + g.e @ P["Pe"].T + P["p"]
    t = relu(tpre)
    c = t @ P["v"]
    da = 1.0 / np.sqrt(np.abs(diag))
    cs = c * da[g.rows] * da[g.cols]
    u = h @ P["vd"] + P["bd"][0]
    uc = np.clip(u, -3.0, 3.0)
    s = np.exp(uc)
    dsi = s / diag
    cache.update(dict(PA=PA, PB=PB, tpre=tpre, t=t, c=c, da=da, u=u, s=s, h=h, diag=diag))
    return cs, dsi, cache


def apply(g: Graph, cs, dsi, r):
    z = dsi * r
    np.add.at(z, g.rows, cs * r[g.cols])
    return z


def backward(theta, g: Graph, cache, r, gz, rb=False):
    """Gradient of sum_i gz_i (M r)_i with respect to theta."""
    P = unpack(theta)
    R = bf16 if rb else (lambda x: np.asarray(x, dtype=np.float64))
    D = {k: np.zeros(s) for k, s in SHAPES}
    da, diag, s, u = cache["da"], cache["diag"], cache["s"], cache["u"]
    h = cache["h"]
    # operator
    dcs = gz[g.rows] * r[g.cols]
    ddsi = gz * r
    dc = dcs * da[g.rows] * da[g.cols]
    ds = ddsi / diag
    du = ds * s * ((u > -3.0) & (u < 3.0))
    D["vd"] += du @ h
    D["bd"][0] += du.sum()
    dh = np.outer(du, P["vd"])
    # edge head
    t, tpre = cache["t"], cache["tpre"]
    D["v"] += dc @ t
    dt = np.outer(dc, P["v"]) * (tpre > 0)
    D["Pe"] += dt.T @ g.e
    D["p"] += dt.sum(axis=0)
    dPA = np.zeros((g.n, Q))
    dPB = np.zeros((g.n, Q))
    np.add.at(dPA, g.rows, dt)
    np.add.at(dPB, g.cols, dt)
    dAB = np.concatenate([dPA, dPB], axis=1)
    Pab = np.concatenate([P["Pa"], P["Pb"]], axis=0)
    dPab = dAB.T @ h
    D["Pa"] += dPab[:Q]
    D["Pb"] += dPab[Q:]
    dh = dh + R(dAB) @ R(Pab)
    for l in reversed(range(LAYERS)):
        a, x, q, hp = cache[f"a{l}"], cache[f"x{l}"], cache[f"q{l}"], cache[f"h{l}"]
        dA = dh * (a > 0)
        D[f"U{l}"] += dA.T @ x
        D[f"d{l}"] += dA.sum(axis=0)
        dx = R(dA) @ R(P[f"U{l}"])
        dhp = dh + dx[:, :H]
        dm = dx[:, H:]
        dq = g.om[:, None] * dm[g.rows] * (q > 0)
        D[f"E{l}"] += dq.T @ g.e
        D[f"c{l}"] += dq.sum(axis=0)
        dG = np.zeros((g.n, H))
        np.add.at(dG, g.cols, dq)
        D[f"B{l}"] += dG.T @ hp
        dhp = dhp + R(dG) @ R(P[f"B{l}"])
        dh = dhp
    dpre = dh * (cache["pre0"] > 0)
    D["W0"] += dpre.T @ cache["f"]
    D["b0"] += dpre.sum(axis=0)
    return pack(D)


def loss_and_gz(A, z, r):
    """L = |A z - r|^2 / |r|^2 and dL/dz (A a dense or scipy matrix)."""
    y = A @ z - r
    nr = float(r @ r)
    return float(y @ y) / nr, 2.0 * (A.T @ y) / nr


def graph_from_positions(pos, re, l0, weight=lambda r, re: re / r - 1.0):
    """Brute-force neighbour graph within re (for small tests): edges i <- j,
    e = ((x_j - x_i)/l0, w(r)/w(l0)), w = the MPS weight."""
    n = len(pos)
    d = pos[None, :, :] - pos[:, None, :]
    r2 = (d ** 2).sum(-1)
    indptr, cols, e, w = [0], [], [], []
    w0 = weight(l0, re)
    for i in range(n):
        js = np.flatnonzero((r2[i] < re * re) & (np.arange(n) != i))
        for j in js:
            rr = np.sqrt(r2[i, j])
            wij = weight(rr, re)
            cols.append(j)
            e.append([d[i, j, 0] / l0, d[i, j, 1] / l0, d[i, j, 2] / l0, wij / w0])
            w.append(wij)
        indptr.append(len(cols))
    return Graph(indptr, cols, np.array(e).reshape(-1, 4), np.array(w))

"""CPU FP64 restatement of the lambda-initial-guess message-passing network.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and the smoke check), never by the
product path.  The reference has no learned initial guess: its Krylov solvers
start from x0 = 0 (src/numeric/iterative_solvers.cpp:424-503, x.setZero()), so
this network has no reference counterpart and is pinned by mathematics instead:
finite differences of this FP64 restatement pin its analytic backward
(tests/test_oracle_pins.py), and the GPU kernels are compared with it.

Network (hidden H = 128, L = 3 layers, F = 16 input features, no biases):

    h0 = bf(X) bf(W_in)                                  X: (N, F)
    a_l = A h_{l-1},   A_ij = c_ij / deg_i               (signed edge coupling)
    u_l = bf(h_{l-1}) bf(Ws_l) + bf(a_l) bf(Wn_l)
    h_l = h_{l-1} + tanh(u_l)                            l = 1..L
    y   = h_L bf(w_out)

`bf` rounds to bfloat16 (round to nearest even from float32) where the GPU
feeds MFMA operands; `emulate_bf16=False` gives the exact FP64 network used for
finite-difference checks.  With no biases and an odd activation the network is
odd: y(-X) = -y(X), so the initial guess flips sign with the source.

Loss over a batch of B graphs (node u in sample b):
    L = (1/B) sum_u omega_b (y_u - tau_u)^2
"""
from __future__ import annotations

import numpy as np

H = 128
LAYERS = 3
F = 16


def param_shapes():
    shapes = [("W_in", (F, H))]
    for l in range(LAYERS):
        shapes.append((f"Ws{l}", (H, H)))
        shapes.append((f"Wn{l}", (H, H)))
    shapes.append(("w_out", (H,)))
    return shapes


def n_params():
    return sum(int(np.prod(s)) for _, s in param_shapes())


def unpack(theta):
    out, o = {}, 0
    for name, shp in param_shapes():
        k = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + k], dtype=np.float64).reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel()
                           for name, _ in param_shapes()])


def bf16(x):
    """Round to bfloat16 (nearest even) through float32; returns float64 values."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    nan = np.isnan(f)
    r = r.astype(np.uint32).view(np.float32)
    r = np.where(nan, f, r)
    return r.astype(np.float64)


class Graph:
    """Batched CSR graph: B copies of one n-node graph with coefficients c_ij."""

    def __init__(self, row_ptr, col, coef, batch=1):
        self.rp = np.asarray(row_ptr, dtype=np.int64)
        self.col = np.asarray(col, dtype=np.int64)
        self.coef = np.asarray(coef, dtype=np.float64)
        self.n = len(self.rp) - 1
        self.batch = batch
        deg = np.diff(self.rp).astype(np.float64)
        self.deg = np.maximum(deg, 1.0)
        rows = np.repeat(np.arange(self.n), np.diff(self.rp))
        self.rows = rows
        self.cA = self.coef / self.deg[rows]           # A_ij = c_ij / deg_i
        self.cT = self.coef / self.deg[self.col]       # (A^T)_ij = c_ji / deg_j

    def _apply(self, h, c):
        n = self.n
        out = np.zeros_like(h)
        for b in range(self.batch):
            hb = h[b * n:(b + 1) * n]
            ob = out[b * n:(b + 1) * n]
            np.add.at(ob, self.rows, c[:, None] * hb[self.col])
        return out

    def agg(self, h):
        return self._apply(h, self.cA)

    def agg_t(self, g):
        return self._apply(g, self.cT)


def forward(theta, X, graph, emulate_bf16=True, keep=False):
    p = unpack(theta)
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    X = np.asarray(X, dtype=np.float64)
    h = r(X) @ r(p["W_in"])
    hs, as_, ts = [h], [], []
    for l in range(LAYERS):
        a = graph.agg(h)
        u = r(h) @ r(p[f"Ws{l}"]) + r(a) @ r(p[f"Wn{l}"])
        t = np.tanh(u)
        h = h + t
        hs.append(h)
        as_.append(a)
        ts.append(t)
    y = h @ r(p["w_out"])
    if keep:
        return y, dict(h=hs, a=as_, t=ts)
    return y


def loss_and_grad(theta, X, graph, tau, omega, emulate_bf16=True):
    """Loss and its gradient w.r.t. theta (analytic backward).

    tau: (N,) targets; omega: (B,) per-sample weights."""
    p = unpack(theta)
    r = bf16 if emulate_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    y, c = forward(theta, X, graph, emulate_bf16, keep=True)
    B, n = graph.batch, graph.n
    w = np.repeat(np.asarray(omega, dtype=np.float64), n)
    diff = y - np.asarray(tau, dtype=np.float64)
    loss = float(np.sum(w * diff * diff) / B)
    gy = 2.0 * w * diff / B
    g = {}
    hL = c["h"][-1]
    g["w_out"] = hL.T @ gy
    gh = np.outer(gy, r(p["w_out"]))
    for l in reversed(range(LAYERS)):
        t, a, hp = c["t"][l], c["a"][l], c["h"][l]
        gu = gh * (1.0 - t * t)
        gub = r(gu)
        g[f"Ws{l}"] = r(hp).T @ gub
        g[f"Wn{l}"] = r(a).T @ gub
        gh = gh + gub @ r(p[f"Ws{l}"]).T + graph.agg_t(gub @ r(p[f"Wn{l}"]).T)
    g["W_in"] = r(X).T @ r(gh)
    return loss, pack(g)


def adam(theta, grad, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1); returns new (theta, m, v)."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v

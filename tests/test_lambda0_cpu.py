"""CPU pins of the learned-lambda0 pieces: the FP64 network oracle (finite
differences, oddness, Adam closed form, bf16 rounding), the interface graph the
library builds (host-only setup), and the seeded family sampler."""
import numpy as np
import pytest

from emagddm_amd import FETIDPSolver, workloads as W
from oracle import mpnn


def random_graph(n, rng, extra=3):
    """Symmetric random graph with coefficients in [-1, 1] (c_ij = c_ji)."""
    pairs = set()
    for i in range(n):
        for _ in range(extra):
            j = int(rng.integers(0, n))
            if j != i:
                pairs.add((min(i, j), max(i, j)))
    c = {p: float(rng.uniform(-1, 1)) for p in pairs}
    rows = [[] for _ in range(n)]
    for (i, j), v in c.items():
        rows[i].append((j, v))
        rows[j].append((i, v))
    rp, col, coef = [0], [], []
    for i in range(n):
        for j, v in sorted(rows[i]):
            col.append(j)
            coef.append(v)
        rp.append(len(col))
    return np.array(rp), np.array(col), np.array(coef)


def small_setup(seed=0, n=12, batch=2):
    rng = np.random.default_rng(seed)
    rp, col, coef = random_graph(n, rng)
    g = mpnn.Graph(rp, col, coef, batch)
    theta = rng.normal(size=mpnn.n_params()) * 0.08
    X = rng.normal(size=(batch * n, mpnn.F))
    tau = rng.normal(size=batch * n)
    omega = rng.uniform(0.5, 2.0, size=batch)
    return g, theta, X, tau, omega


def test_oracle_gradient_matches_finite_differences():
    g, theta, X, tau, omega = small_setup()
    loss, grad = mpnn.loss_and_grad(theta, X, g, tau, omega, emulate_bf16=False)
    rng = np.random.default_rng(1)
    # every parameter block, several random coordinates each
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in mpnn.param_shapes()])
    h = 1e-6
    for b in range(len(offs) - 1):
        for idx in rng.integers(offs[b], offs[b + 1], size=4):
            tp, tm = theta.copy(), theta.copy()
            tp[idx] += h
            tm[idx] -= h
            lp, _ = mpnn.loss_and_grad(tp, X, g, tau, omega, emulate_bf16=False)
            lm, _ = mpnn.loss_and_grad(tm, X, g, tau, omega, emulate_bf16=False)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grad[idx]) <= 1e-6 * max(1.0, abs(fd)), (b, idx, fd, grad[idx])
    # directional derivative along a random direction (all parameters at once)
    v = rng.normal(size=theta.size)
    lp, _ = mpnn.loss_and_grad(theta + h * v, X, g, tau, omega, emulate_bf16=False)
    lm, _ = mpnn.loss_and_grad(theta - h * v, X, g, tau, omega, emulate_bf16=False)
    assert abs((lp - lm) / (2 * h) - grad @ v) <= 1e-6 * max(1.0, abs(grad @ v))


def test_oracle_transpose_aggregation_is_adjoint():
    g, _, _, _, _ = small_setup(seed=3)
    rng = np.random.default_rng(4)
    a = rng.normal(size=(g.batch * g.n, 5))
    b = rng.normal(size=(g.batch * g.n, 5))
    assert abs(np.sum(g.agg(a) * b) - np.sum(a * g.agg_t(b))) <= 1e-12 * np.abs(a).sum()


def test_oracle_network_is_odd():
    g, theta, X, _, _ = small_setup(seed=5)
    for emulate in (False, True):
        y = mpnn.forward(theta, X, g, emulate)
        assert np.array_equal(mpnn.forward(theta, -X, g, emulate), -y)


def test_oracle_adam_first_step_closed_form():
    rng = np.random.default_rng(6)
    th = rng.normal(size=50)
    gr = rng.normal(size=50)
    new, m, v = mpnn.adam(th, gr, np.zeros(50), np.zeros(50), 1, 1e-3)
    # after one step m_hat = g and v_hat = g^2, so the update is lr g / (|g| + eps)
    assert np.allclose(new, th - 1e-3 * gr / (np.abs(gr) + 1e-8), rtol=0, atol=1e-15)


def test_oracle_bf16_rounding():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 3.0e38, 1e-40, 0.0])
    # ties go to the even significand; 8 significand bits (7 stored)
    expect = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5])
    got = mpnn.bf16(x)
    assert np.array_equal(got[:5], expect)
    assert got[7] == 0.0
    # every result is representable: the low 16 bits of its float32 image are zero
    f = got.astype(np.float32).view(np.uint32)
    assert np.all((f & 0xFFFF) == 0)
    # and no representable value lies strictly closer (brute force over neighbours)
    for v, r in zip(x.astype(np.float32), got.astype(np.float32)):
        u = int(r.view(np.uint32))
        for nb in (u - 0x10000, u + 0x10000):
            if not 0 <= nb < 2 ** 32:
                continue
            w = np.uint32(nb).view(np.float32)
            if np.isfinite(w) and (w >= 0) == (r >= 0):
                assert abs(float(w) - float(v)) >= abs(float(r) - float(v))


def _expected_dual_graph(s, prob):
    topo = s.topology()
    en = topo["edge_nodes"]
    tf = topo["tet_free"]
    # free dof -> global edge through the per-tet maps
    from oracle import topology as otopo
    edges, _ = otopo.edge_ids(prob.tet2node)
    free_edge = np.full(s.sizes["n_free"], -1)
    fl = tf.ravel()
    el = edges.ravel()
    free_edge[fl[fl >= 0]] = el[fl >= 0]
    lam_edges = free_edge[topo["lam_free"]]
    a, b = en[lam_edges, 0], en[lam_edges, 1]
    t = prob.xyz[b] - prob.xyz[a]
    t /= np.linalg.norm(t, axis=1)[:, None]
    inc = {}
    for l, (u, v) in enumerate(zip(a, b)):
        inc.setdefault(u, []).append(l)
        inc.setdefault(v, []).append(l)
    rp, col, coef = [0], [], []
    for l in range(len(a)):
        nb = sorted(set(inc[a[l]] + inc[b[l]]) - {l})
        for j in nb:
            c = float(t[l] @ t[j])
            if abs(c) < 1e-6:
                continue
            col.append(j)
            coef.append(c)
        rp.append(len(col))
    return np.array(rp), np.array(col), np.array(coef)


@pytest.mark.parametrize("n,parts,jitter", [(8, 2, None), (12, 3, (3, 5))])
def test_dual_graph_matches_host_construction(n, parts, jitter):
    p = W.make_problem(n, parts, jitter, seed=2)
    s = FETIDPSolver()
    s.setup_problem(p)
    rp, col, coef = s.dual_graph()
    erp, ecol, ecoef = _expected_dual_graph(s, p)
    assert np.array_equal(rp, erp) and np.array_equal(col, ecol)
    assert np.allclose(coef, ecoef, rtol=0, atol=1e-6)
    # symmetric pattern and coefficients
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    fwd = {(int(i), int(j)): c for i, j, c in zip(rows, col, coef)}
    assert all(abs(fwd[(j, i)] - c) == 0 for (i, j), c in fwd.items())


@pytest.mark.parametrize("name,seeds", [("cube16", [0, 1, 7]), ("cube32l", [0, 2])])
def test_family_parameters_match_make_workload(name, seeds):
    for seed in seeds:
        p = W.make_workload(name, seed)
        nu, kappa, loops = W.family_parameters(name, seed)
        assert np.array_equal(nu, p.nu) and np.array_equal(kappa, p.kappa)
        jq = W.loop_source(W.quadrature_points(p.xyz, p.tet2node), loops)
        assert np.array_equal(jq, p.jq)

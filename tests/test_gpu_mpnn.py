"""GPU parity of the learned-lambda0 network (bf16 MFMA, FP32 accumulate) with the
FP64 oracle, and of the solver path that starts PCPG from it (through the C ABI)."""
import ctypes as C

import numpy as np
import pytest

from emagddm_amd import FETIDPSolver, SolverStatus, _lib, workloads as W
from oracle import mpnn, solvers
from tests.test_lambda0_cpu import random_graph

pytestmark = pytest.mark.gpu

# bf16 operands (8-bit significand) with FP32 accumulation against FP64 with the
# same operand rounding: rounding-boundary flips between the two bound the gap.
FWD_TOL = 5e-3
GRAD_TOL = 2e-2


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def test_bf16_mfma_lane_layout(gpu_available):
    """16x16x32: lane l holds A[l%16][8(l/16)+j] and B[8(l/16)+j][l%16], j < 8;
    accumulator j of lane l is D[4(l/16)+j][l%16]."""
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 9, size=(16, 32)).astype(np.float32)
    B = rng.integers(-8, 9, size=(32, 16)).astype(np.float32)
    af = np.zeros(512, np.float32)
    bf = np.zeros(512, np.float32)
    for l in range(64):
        for j in range(8):
            af[8 * l + j] = A[l % 16, 8 * (l // 16) + j]
            bf[8 * l + j] = B[8 * (l // 16) + j, l % 16]
    d = np.zeros(256, np.float32)
    assert _lib.load().emag_mpnn_mfma_probe(_fp(af), _fp(bf), _fp(d)) == 0
    D = A.astype(np.float64) @ B.astype(np.float64)
    got = np.zeros((16, 16))
    for l in range(64):
        for j in range(4):
            got[4 * (l // 16) + j, l % 16] = d[4 * l + j]
    assert np.array_equal(got, D)


def _setup(seed, n, batch):
    import torch
    from emagddm_amd.lambda0 import LambdaMPNN
    rng = np.random.default_rng(seed)
    rp, col, coef = random_graph(n, rng, extra=4)
    coef = coef.astype(np.float32).astype(np.float64)
    theta = (rng.normal(size=mpnn.n_params()) * 0.06).astype(np.float32)
    X = rng.normal(size=(batch * n, mpnn.F)).astype(np.float32)
    tau = rng.normal(size=batch * n).astype(np.float32)
    om = rng.uniform(0.5, 2.0, size=batch).astype(np.float32)
    m = LambdaMPNN(params=theta)
    m.set_graph(rp, col, coef)
    dev = lambda a: torch.as_tensor(a, device="cuda")
    return m, mpnn.Graph(rp, col, coef, batch), theta.astype(np.float64), X, tau, om, dev


@pytest.mark.parametrize("n,batch", [(300, 3), (517, 2)])
def test_forward_matches_oracle(gpu_available, n, batch):
    print(f"[mpnn] forward n={n} batch={batch}", flush=True)
    m, g, theta, X, _, _, dev = _setup(1, n, batch)
    y = m.forward(dev(X)).cpu().numpy().astype(np.float64)
    yo = mpnn.forward(theta, X.astype(np.float64), g, emulate_bf16=True)
    assert np.linalg.norm(y - yo) <= FWD_TOL * np.linalg.norm(yo)
    # odd network: y(-X) = -y(X) (bf16 rounding and the sums are sign-symmetric)
    y2 = m.forward(dev(-X)).cpu().numpy().astype(np.float64)
    assert np.allclose(y2, -y, rtol=1e-6, atol=1e-7 * np.abs(y).max())


def test_gradient_matches_oracle_and_is_repeatable(gpu_available):
    print("[mpnn] gradient n=400 batch=3", flush=True)
    m, g, theta, X, tau, om, dev = _setup(2, 400, 3)
    loss, grad = m.gradient(dev(X), dev(tau), dev(om))
    lo, go = mpnn.loss_and_grad(theta, X.astype(np.float64), g, tau.astype(np.float64),
                                om.astype(np.float64), emulate_bf16=True)
    assert abs(loss - lo) <= FWD_TOL * abs(lo)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in mpnn.param_shapes()])
    for b, (name, _) in enumerate(mpnn.param_shapes()):
        gs, os_ = grad[offs[b]:offs[b + 1]], go[offs[b]:offs[b + 1]]
        assert np.linalg.norm(gs - os_) <= GRAD_TOL * np.linalg.norm(os_), name
    loss2, grad2 = m.gradient(dev(X), dev(tau), dev(om))
    assert loss2 == loss and np.array_equal(grad2, grad)


def test_train_step_is_adam_on_the_gradient(gpu_available):
    m, g, theta, X, tau, om, dev = _setup(3, 200, 2)
    _, grad = m.gradient(dev(X), dev(tau), dev(om))
    lr = 1e-3
    m.train_step(dev(X), dev(tau), dev(om), lr)
    got = m.get_params().astype(np.float64)
    exp, _, _ = mpnn.adam(theta, grad.astype(np.float64), np.zeros_like(theta),
                          np.zeros_like(theta), 1, lr)
    assert np.abs(got - exp).max() <= 1e-6 * lr + 1e-7 * np.abs(theta).max()
    # a few more steps on one batch reduce the loss
    l0 = m.gradient(dev(X), dev(tau), dev(om))[0]
    for _ in range(20):
        m.train_step(dev(X), dev(tau), dev(om), 3e-3)
    assert m.gradient(dev(X), dev(tau), dev(om))[0] < l0


def test_solver_features_and_learned_start(gpu_available):
    import torch
    from emagddm_amd.lambda0 import LambdaMPNN
    p = W.make_workload("cube16", 3)
    print(f"[mpnn] solver cube16 seed 3 tets={p.n_tets}", flush=True)
    s = FETIDPSolver()
    s.setup_problem(p)
    s.set_rhs_tet_J(p.jq)
    s.factor()
    X, sc = s.lambda_features()
    d = s.dual_rhs()
    z = s.apply_preconditioner(d)
    nl = d.size
    assert np.isclose(sc[0], np.sqrt(d @ d / nl), rtol=1e-12)
    assert np.isclose(sc[1], np.sqrt(z @ z / nl), rtol=1e-12)
    assert np.allclose(X[:, 0], d / sc[0], rtol=1e-6, atol=1e-7)
    assert np.allclose(X[:, 1], z / sc[1], rtol=1e-6, atol=1e-7)
    assert np.all(X[:, 14:] == 0)
    # tangent features: |t| components are a unit vector
    t = X[:, 6:9] / np.where(np.abs(X[:, 1:2]) > 1e-6, X[:, 1:2], np.inf)
    ok = np.abs(X[:, 1]) > 1e-3
    assert np.allclose(np.sum(t[ok] ** 2, axis=1), 1.0, atol=1e-4)
    # reference answer and PCPG from zero
    r0 = s.solve()
    assert r0.status == SolverStatus.SUCCESS and not r0.learned_guess
    rp, col, coef = s.dual_graph()
    rng = np.random.default_rng(7)
    theta = (rng.normal(size=mpnn.n_params()) * 0.05).astype(np.float32)
    m = LambdaMPNN(params=theta)
    m.set_graph(rp, col, coef)
    s.set_lambda0_model(m)
    r1 = s.solve()
    assert r1.status == SolverStatus.SUCCESS and r1.learned_guess
    assert r1.residual_norm <= 1e-8
    # the start is s_z * network(features), scaled to its F-norm optimum
    y = m.forward(torch.as_tensor(X, device="cuda")).cpu().numpy().astype(np.float64)
    lam0 = sc[1] * y
    q = s.apply_dual_operator(lam0)
    alpha = (lam0 @ d) / (lam0 @ q)
    assert np.isclose(r1.initial_dual_residual,
                      np.linalg.norm(d - alpha * q) / np.linalg.norm(d), rtol=1e-6)
    rg, cg, vg, fg = s.global_system()
    import scipy.sparse as sp
    A = sp.csr_matrix((vg, cg, rg), shape=(fg.size, fg.size))
    xd = solvers.direct(A, fg)
    for r in (r0, r1):
        assert np.linalg.norm(r.x - xd) <= 1e-7 * np.linalg.norm(xd)
    s.set_lambda0_model(None)
    r2 = s.solve()
    assert not r2.learned_guess and r2.iterations == r0.iterations


def test_new_source_reuses_factor(gpu_available):
    p = W.make_workload("cube16", 0)
    s = FETIDPSolver()
    s.setup_problem(p)
    s.set_rhs_tet_J(p.jq)
    s.factor()
    _, _, loops = W.family_parameters("cube16", 5)
    jq = W.loop_source(W.quadrature_points(p.xyz, p.tet2node), loops)
    s.set_rhs_tet_J(jq)  # no refactorization
    r = s.solve()
    t = FETIDPSolver()
    t.setup_problem(p)
    t.set_rhs_tet_J(jq)
    t.factor()
    rt = t.solve()
    assert r.status == SolverStatus.SUCCESS
    assert np.array_equal(r.x, rt.x)

"""Pins of the FP64 oracle against reference known answers, closed forms and
independent computations (CPU only)."""
import numpy as np
import pytest

from emagddm_amd import workloads as W
from oracle import assembly, fetidp, nedelec, solvers, topology


def test_reference_curl_known_answers():
    # test/test_shape_function.cpp:306-328: edge0 -> (0,0,-2), edge1 -> (-2,0,0), constant
    assert np.allclose(nedelec.curl_ref(0), [0, 0, -2], atol=1e-14)
    assert np.allclose(nedelec.curl_ref(1), [-2, 0, 0], atol=1e-14)
    # the other reference curls (src/numeric/nedelec_element.cpp:460-500)
    assert np.allclose(nedelec.curl_ref(2), [0, -2, 0])
    assert np.allclose(nedelec.curl_ref(3), [0, -2, 2])
    assert np.allclose(nedelec.curl_ref(4), [2, 0, -2])
    assert np.allclose(nedelec.curl_ref(5), [-2, 2, 0])


def test_reference_whitney_values():
    # W_0 = (l2, -l1, 0), W_3 = (l4 + l1, l1, l1)  (src/numeric/nedelec_element.cpp:421-446)
    xi = (0.1, 0.2, 0.3)
    assert np.allclose(nedelec.whitney_ref(0, xi), [0.2, -0.1, 0.0])
    assert np.allclose(nedelec.whitney_ref(3, xi), [0.4 + 0.1, 0.1, 0.1])
    assert np.allclose(nedelec.whitney_ref(5, xi), [0.3, 0.3, 0.4 + 0.3])


def test_whitney_circulation_second_to_first():
    # unit circulation along b -> a for edge (a, b) on the reference tet
    V = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0], [0, 0, 0]])
    s, w = np.polynomial.legendre.leggauss(4)
    s, w = 0.5 * (s + 1), 0.5 * w
    for e, (a, b) in enumerate(nedelec.REF_EDGES):
        for f, (c, d) in enumerate(nedelec.REF_EDGES):
            t = V[c] - V[d]                      # direction d -> c
            circ = sum(wi * nedelec.whitney_ref(e, V[d] + si * t) @ t for si, wi in zip(s, w))
            assert abs(circ - (1.0 if e == f else 0.0)) < 1e-13


def test_quadrature_rules():
    # the reference 1-point rule (test/test_em_element_integrator.cpp:229-239)
    P, w = nedelec.quad_points_bary("1pt")
    assert np.allclose(P, 0.25) and np.isclose(w.sum(), 1.0 / 6.0)
    # corrected 4-point rule: four points, total weight 1/6 (D4), exact for degree 2
    P, w = nedelec.quad_points_bary("4pt")
    assert P.shape == (4, 4) and np.isclose(w.sum(), 1.0 / 6.0)
    # int l_a l_b = (1 + delta_ab) / 120 on the reference tet
    for a in range(4):
        for b in range(4):
            assert np.isclose((w * P[:, a] * P[:, b]).sum(), (1 + (a == b)) / 120.0)


@pytest.fixture(scope="module")
def random_tets():
    rng = np.random.default_rng(7)
    X = rng.normal(size=(64, 4, 3))
    return X, rng.uniform(0.1, 3, 64), rng.uniform(0.01, 2, 64)


def test_element_closed_form_vs_quadrature(random_tets):
    X, nu, ka = random_tets
    K, M = nedelec.element_matrices(X, nu, ka)
    for rule in ("4pt", "duffy4"):
        Kq, Mq = nedelec.element_matrices_quadrature(X, nu, ka, rule)
        assert np.abs(K - Kq).max() <= 1e-12 * np.abs(K).max()
        assert np.abs(M - Mq).max() <= 1e-12 * np.abs(M).max()
    # D4: the reference's "4-point" rule keeps 3 points of weight 1/24 (src/numeric/
    # gauss_quadrature.cpp:200-219: (a,a,b), (a,b,a), then (b,a,a) replaced by (a,a,a)).
    # It misintegrates M_e, even rescaled to the right total weight 1/6, while K_e (constant
    # curls) only scales with the weight; the oracle's M_e is not that rule's.
    a, b = nedelec.QA, nedelec.QB
    xi3 = np.array([[a, a, b], [a, b, a], [a, a, a]])
    P3 = np.hstack([xi3, 1.0 - xi3.sum(1, keepdims=True)])
    K3, M3 = nedelec.element_matrices_quadrature(X, nu, ka, (P3, np.full(3, 1.0 / 24.0)))
    assert np.abs(K3 * (4.0 / 3.0) - K).max() <= 1e-12 * np.abs(K).max()
    rel = lambda A, B: np.abs(A - B).max() / np.abs(B).max()
    assert rel(M3, M) > 0.1
    assert rel(M3 * (4.0 / 3.0), M) > 1e-2


def test_element_invariants(random_tets):
    X, nu, ka = random_tets
    K, M = nedelec.element_matrices(X, nu, ka)
    G = nedelec.gradient_incidence()
    assert np.abs(K @ G).max() <= 1e-11 * np.abs(K).max()   # curl grad = 0
    assert np.allclose(K, np.transpose(K, (0, 2, 1)))
    assert np.allclose(M, np.transpose(M, (0, 2, 1)))
    for k in range(8):
        ev = np.linalg.eigvalsh(K[k])
        assert ev.min() > -1e-12 * ev.max()
        assert (ev > 1e-10 * ev.max()).sum() == 3              # rank 3
        assert np.linalg.eigvalsh(M[k]).min() > 0              # SPD


def test_element_rhs_brute_force(random_tets):
    X = random_tets[0]
    # linear J: the 4-point rule is exact, matches a 6^3 collapsed Gauss rule
    lin = lambda x: np.stack([x[:, 1] + 0.3, -2 * x[:, 0], 0.5 * x[:, 2] + x[:, 0]], 1)
    pts = W.quadrature_points(X.reshape(-1, 3), np.arange(X.shape[0] * 4).reshape(-1, 4))
    jq = lin(pts.reshape(-1, 3)).reshape(-1, 4, 3)
    F = nedelec.element_rhs(X, jq)
    Fd = nedelec.element_rhs_function(X, lin, "duffy6")
    assert np.abs(F - Fd).max() <= 1e-12 * np.abs(Fd).max()
    Fm = nedelec.element_rhs(X, -jq)
    assert np.allclose(Fm, -F)


def _node_divergence(p, jq):
    """(G^T f)_n = int grad(phi_n) . J over the mesh, for the interior nodes n."""
    asm = assembly.assemble(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.pec_pairs, jq)
    f, pairs, free = asm["f"], asm["pairs"], asm["free"]
    # grad(phi_n) = sum_e G[e, n] W_e with the global edge e running low -> high node id
    e = np.nonzero(free >= 0)[0]
    lo, hi = pairs[e].min(1), pairs[e].max(1)
    div = np.zeros(p.n_nodes)
    np.add.at(div, hi, f[free[e]])
    np.subtract.at(div, lo, f[free[e]])
    x = p.xyz
    interior = np.all((x > 1e-12) & (x < 1 - 1e-12), axis=1)
    return div[interior], np.linalg.norm(f)


def test_source_divergence_free():
    """G^T F -> 0 for a divergence-free J (div J = 0 and phi_n = 0 on the boundary):
    exactly for a linear J (the 4-point rule integrates grad(phi) . J exactly), and
    vanishing under refinement for the Gaussian current loops."""
    p = W.make_problem(6, 1, seed=3)
    lin = lambda x: np.stack([x[:, 1] - 0.5, 0.5 - x[:, 0] + 0.3 * x[:, 2], 0.2 * x[:, 0]], 1)
    pts = W.quadrature_points(p.xyz, p.tet2node)
    jq = lin(pts.reshape(-1, 3)).reshape(pts.shape)
    d, fn = _node_divergence(p, jq)
    assert np.abs(d).max() <= 1e-13 * fn
    # a J with divergence 1 is caught
    jq2 = jq + np.concatenate([pts[..., :1], 0 * pts[..., 1:]], -1)
    d2, fn2 = _node_divergence(p, jq2)
    assert np.abs(d2).max() > 1e-3 * fn2
    ratios = []
    for n in (8, 16):
        q = W.make_problem(n, 1, seed=5)
        d, fn = _node_divergence(q, q.jq)
        ratios.append(np.linalg.norm(d) / fn)
    assert ratios[1] < 0.5 * ratios[0] and ratios[1] < 0.2


def test_edge_ids_loop_equals_vectorised():
    p = W.make_problem(3, 1, seed=2)
    t2e, pairs = topology.edge_ids(p.tet2node)
    t2e2, pairs2 = topology.edge_ids_loop(p.tet2node)
    assert np.array_equal(t2e, t2e2) and np.array_equal(pairs, pairs2)
    # shuffled tet order changes ids but not the edge set
    perm = np.random.default_rng(0).permutation(p.n_tets)
    _, pairs3 = topology.edge_ids(p.tet2node[perm])
    assert {tuple(x) for x in pairs3} == {tuple(x) for x in pairs}


@pytest.mark.parametrize("n", [2, 3, 5, 16])
def test_kuhn_counts(n):
    xyz, tets = W.kuhn_cube(n)
    _, pairs = topology.edge_ids(tets)
    E, S = topology.kuhn_edge_counts(n, n, n)
    assert pairs.shape[0] == E
    assert W.pec_edges(n, tets).shape[0] == S


def test_free_numbering_semantics():
    # test/test_pyramid5_edge.cpp:291-360 semantics: a DIRICHLET edge maps to -1,
    # the remaining edges are numbered contiguously in prenumber order.
    tets = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], dtype=np.int32)
    t2e, pairs = topology.edge_ids(tets)
    free = topology.free_numbering(pairs, np.array([[1, 0]]))
    assert free[0] == -1                     # edge (0,1) is id 0
    assert list(free[1:]) == list(range(pairs.shape[0] - 1))
    free2 = topology.free_numbering(pairs, np.array([[3, 2]]))
    k = [tuple(x) for x in pairs].index((2, 3))
    assert free2[k] == -1 and (free2 >= 0).sum() == pairs.shape[0] - 1


def test_classification_box_partition():
    p = W.make_problem(8, 2, seed=0)
    l2g, pairs, free = topology.local_to_global(p.tet2node, p.pec_pairs)
    n = int(free.max()) + 1
    mult, dofs, subs = topology.classify(l2g, p.tet_subdomain, n)
    assert set(np.unique(mult)) <= {1, 2, 4}
    # interior coarse edges: 3 lines (x=y=4 etc.) through the cube minus boundary: 3*(P-1)^2*n
    assert (mult == 4).sum() == 3 * 1 * 8
    # dual count formula 3(P-1)(F2(n,n) - 4n) - 2*3(P-1)^2 n with P=2
    f2 = lambda a, b: a * (b + 1) + b * (a + 1) + a * b
    assert (mult == 2).sum() == 3 * 1 * (f2(8, 8) - 32) - 2 * 3 * 8


@pytest.fixture(scope="module")
def micro_system():
    p = W.make_problem(6, 2, seed=1)
    asm = assembly.assemble(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.pec_pairs, p.jq)
    return p, asm


def test_global_system_spd(micro_system):
    _, asm = micro_system
    A = asm["A"].toarray()
    assert np.allclose(A, A.T, atol=1e-14 * np.abs(A).max())
    assert np.linalg.eigvalsh(A).min() > 0


def test_global_system_matches_dense_assembly(micro_system):
    p, asm = micro_system
    # independent dense scatter with the reference per-element loop structure
    n = asm["A"].shape[0]
    A = np.zeros((n, n))
    K, M = asm["Ke"], asm["Me"]
    for t in range(p.n_tets):
        for i in range(6):
            for j in range(6):
                gi, gj = asm["l2g"][t, i], asm["l2g"][t, j]
                if gi >= 0 and gj >= 0:
                    A[gi, gj] += asm["signs"][t, i] * asm["signs"][t, j] * (K[t, i, j] + M[t, i, j])
    assert np.abs(A - asm["A"].toarray()).max() <= 1e-14 * np.abs(A).max()


def test_pcg_reference_semantics(micro_system):
    _, asm = micro_system
    x, st, it, res = solvers.pcg_jacobi(asm["A"], asm["f"], 1e-8, 1000)
    assert st == solvers.SUCCESS and res < 1e-8
    xd = solvers.direct(asm["A"], asm["f"])
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-6
    # b = 0 returns x = 0 immediately; max_iter=1 -> MAX_ITER_REACHED
    x0, st0, it0, _ = solvers.pcg_jacobi(asm["A"], 0 * asm["f"])
    assert st0 == solvers.SUCCESS and it0 == 0 and not x0.any()
    _, st1, it1, _ = solvers.pcg_jacobi(asm["A"], asm["f"], 1e-8, 1)
    assert st1 == solvers.MAX_ITER_REACHED and it1 == 1


def test_manufactured_convergence():
    # E = (sin pi y sin pi z, sin pi x sin pi z, sin pi x sin pi y) has zero tangential
    # trace on the cube; solve curl curl A + A = J with J = (2 pi^2 + 1) E and check
    # that the energy-norm distance to the edge interpolant of E falls with h.
    errs = []
    for n in (4, 8):
        xyz, tets = W.kuhn_cube(n)
        pec = W.pec_edges(n, tets)
        mat = np.zeros(tets.shape[0], dtype=np.int32)
        l2g, pairs, free = topology.local_to_global(tets, pec)

        def E(x):
            s = np.sin(np.pi * x)
            return np.stack([s[:, 1] * s[:, 2], s[:, 0] * s[:, 2], s[:, 0] * s[:, 1]], 1)

        def J(x):  # div E = 0, so curl curl E + E = (2 pi^2 + 1) E
            return (2 * np.pi ** 2 + 1) * E(x)

        jq = J(W.quadrature_points(xyz, tets).reshape(-1, 3)).reshape(-1, 4, 3)
        asm = assembly.assemble(xyz, tets, mat, [1.0], [1.0], pec, jq)
        u = solvers.direct(asm["A"], asm["f"])
        # edge interpolant of E (line integral along low->high), 4-point Gauss
        gs, gw = np.polynomial.legendre.leggauss(4)
        a, b = xyz[pairs[:, 0]], xyz[pairs[:, 1]]
        ie = sum(0.5 * w * np.einsum("ij,ij->i", E(a + 0.5 * (s + 1) * (b - a)), b - a)
                 for s, w in zip(gs, gw))
        ui = ie[free >= 0]
        d = u - ui
        errs.append(np.sqrt(d @ (asm["A"] @ d)) / np.sqrt(ui @ (asm["A"] @ ui)))
    assert errs[1] < 0.6 * errs[0]


def test_dense_fetidp_matches_direct():
    p = W.make_problem(6, 2, seed=4)
    F = fetidp.DenseFETIDP(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.tet_subdomain,
                           p.pec_pairs, p.jq)
    out = F.solve(tol=1e-12)
    xd = solvers.direct(F.A, F.f)
    assert out["status"] == solvers.SUCCESS
    assert np.linalg.norm(out["x"] - xd) <= 1e-9 * np.linalg.norm(xd)
    # F is symmetric positive definite on the dual space (brute force)
    Fd = F.dense_F()
    assert np.allclose(Fd, Fd.T, atol=1e-12 * np.abs(Fd).max())
    assert np.linalg.eigvalsh(0.5 * (Fd + Fd.T)).min() > 0
    # continuity of the exact solution: B u = 0 -> F lambda* = d holds for lambda from PCPG
    lam = out["lam"]
    assert np.linalg.norm(Fd @ lam - F.d) <= 1e-10 * np.linalg.norm(F.d)


def test_dense_fetidp_contrast():
    p = W.make_problem(8, 2, materials="contrast", seed=5)
    F = fetidp.DenseFETIDP(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.tet_subdomain,
                           p.pec_pairs, p.jq)
    out = F.solve(tol=1e-12)
    xd = solvers.direct(F.A, F.f)
    assert np.linalg.norm(out["x"] - xd) <= 1e-8 * np.linalg.norm(xd)


def test_deluxe_scaling_pins():
    """Deluxe preconditioner of the oracle: symmetric positive semidefinite; with
    diagonal Dirichlet blocks it is exactly stiffness scaling with rho = diag(P);
    on a coefficient-contrast problem it cuts PCPG from hundreds of iterations to
    the uniform-coefficient count and still reproduces the direct solution."""
    p = W.make_problem(8, 2, materials="contrast", seed=5)
    F = fetidp.DenseFETIDP(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.tet_subdomain,
                           p.pec_pairs, p.jq, scaling="deluxe")
    M = np.stack([F.apply_prec(e) for e in np.eye(F.n_lam)], axis=1)
    assert np.allclose(M, M.T, atol=1e-12 * np.abs(M).max())
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > -1e-10 * np.abs(M).max()
    # diagonal blocks: deluxe == stiffness scaling with rho = diag(P)
    for S in F.subs:
        S["P"] = np.diag(np.diag(S["P"]))
    F._deluxe_setup()
    for (s0, s1, lams, k0, k1, D0) in F.faces:
        r0 = np.diag(F.subs[s0]["P"])[k0]
        r1 = np.diag(F.subs[s1]["P"])[k1]
        F.subs[s0]["D"][k0] = r1 / (r0 + r1)
        F.subs[s1]["D"][k1] = r0 / (r0 + r1)
    rng = np.random.default_rng(0)
    r = rng.normal(size=F.n_lam)
    zd = F.apply_prec(r)
    F.scaling = "stiffness"
    zs = F.apply_prec(r)
    assert np.allclose(zd, zs, rtol=1e-12, atol=1e-14 * np.abs(zs).max())
    # contrast: deluxe converges like the uniform case
    q = W.make_problem(12, 3, None, "contrast", seed=7)
    out = {}
    for sc in ("stiffness", "deluxe"):
        G = fetidp.DenseFETIDP(q.xyz, q.tet2node, q.tet_material, q.nu, q.kappa, q.tet_subdomain,
                               q.pec_pairs, q.jq, scaling=sc)
        out[sc] = G.solve(tol=1e-10, max_iter=1000)
    assert out["stiffness"]["iterations"] > 200 and out["deluxe"]["iterations"] < 30
    xd = solvers.direct(G.A, G.f)
    assert np.linalg.norm(out["deluxe"]["x"] - xd) <= 1e-7 * np.linalg.norm(xd)


def test_dense_fetidp_irregular_partition():
    """Voronoi subdomains on a perturbed mesh: wire-basket primal edges still remove
    every local kernel (each local block is SPD after the primal constraint) and the
    FETI-DP answer is the direct one."""
    p = W.irregular_problem(6, 5, seed=2)
    assert p.n_sub >= 4
    F = fetidp.DenseFETIDP(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.tet_subdomain,
                           p.pec_pairs, p.jq)
    out = F.solve(tol=1e-12)
    xd = solvers.direct(F.A, F.f)
    assert out["status"] == solvers.SUCCESS
    assert np.linalg.norm(out["x"] - xd) <= 1e-9 * np.linalg.norm(xd)


def test_irregular_partition_needs_deluxe():
    """On Voronoi subdomains stiffness scaling stalls (hundreds of PCPG iterations) and
    deluxe scaling converges in a box-partition-like count; on an even box partition the
    two take the same count.  This is the measurement behind the automatic scaling rule."""
    p = W.irregular_problem(6, 5, seed=2)
    its = {}
    for sc in ("stiffness", "deluxe"):
        F = fetidp.DenseFETIDP(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa,
                               p.tet_subdomain, p.pec_pairs, p.jq, scaling=sc)
        its[sc] = F.solve(tol=1e-10, max_iter=400)["iterations"]
    assert its["deluxe"] <= 30 and its["stiffness"] >= 4 * its["deluxe"]
    q = W.make_problem(12, 3, seed=1)
    its = [fetidp.DenseFETIDP(q.xyz, q.tet2node, q.tet_material, q.nu, q.kappa,
                              q.tet_subdomain, q.pec_pairs, q.jq, scaling=sc)
           .solve(tol=1e-10)["iterations"] for sc in ("stiffness", "deluxe")]
    assert abs(its[0] - its[1]) <= 1

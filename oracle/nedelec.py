"""Lowest-order Nedelec (Whitney) tetrahedron -- FP64 restatement (oracle).

Conventions follow the reference:

* Reference tet: lambda_1..3 = xi, eta, zeta, lambda_4 = 1 - xi - eta - zeta;
  node 3 (0-based) at the origin (src/numeric/lagrange_element.cpp:658-683).
* Local edges: the NedelecElement TET4_EDGE list {0,1},{1,2},{2,0},{0,3},
  {1,3},{2,3} (src/numeric/nedelec_element.cpp:54-60).  For edge (a, b) the
  reference function is W = lambda_b grad lambda_a - lambda_a grad lambda_b
  (src/numeric/nedelec_element.cpp:410-458); its circulation is +1 from b to a.
  Reference curl = 2 grad lambda_b x grad lambda_a (:460-500).
* Geometry: J = [x0-x3, x1-x3, x2-x3] (columns dx/dxi_k), the base-class
  calcJacobian convention (include/numeric/shape_function_base.hpp:335-364)
  applied to the TET4 Lagrange gradients.
* K_e(i,j) = sum_q w_q |detJ| nu curl_i . curl_j and
  M_e(i,j) = sum_q w_q |detJ| kappa W_i . W_j
  (src/numeric/magnetic_vector_3d_integrator.cpp:139-323), with nu = 1/mu and
  kappa playing the role of sigma.

Corrections of the reference defects (each one would make the operator wrong):

* D1/D4: the tet rule is the correct 4-point rule (all four points, w=1/24),
  not the 3-point list of src/numeric/gauss_quadrature.cpp:200-219.
* D2: the Jacobian comes from the Lagrange TET4 gradients, not the zero
  NedelecElement::evalGradN (src/numeric/nedelec_element.cpp:127-133).
* D3: covariant Piola map W_phys = J^-T W_ref, without the extra 1/detJ of
  src/numeric/magnetic_vector_3d_integrator.cpp:429-430.
* D6: the orientation sign s_e = +1 if the global id of b is below that of a
  (global edges run low id -> high id), else -1.
* F_e(i) = sum_q w_q |detJ| W_i(x_q) . J(x_q) (the reference returns zero,
  :343-364), 4-point rule, point q at barycentric weight b on node q.
"""
from __future__ import annotations

import math

import numpy as np

REF_EDGES = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
GEOM_EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
GRAD_REF = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0], [-1.0, -1, -1]])
QA = (5.0 - math.sqrt(5.0)) / 20.0
QB = (5.0 + 3.0 * math.sqrt(5.0)) / 20.0


def barycentric_ref(xi):
    x, y, z = xi
    return np.array([x, y, z, 1.0 - x - y - z])


def whitney_ref(e: int, xi) -> np.ndarray:
    """Reference-element Whitney function of local edge e at xi (3-vector)."""
    a, b = REF_EDGES[e]
    lam = barycentric_ref(xi)
    return lam[b] * GRAD_REF[a] - lam[a] * GRAD_REF[b]


def curl_ref(e: int) -> np.ndarray:
    """Constant reference curl 2 grad(lambda_b) x grad(lambda_a)."""
    a, b = REF_EDGES[e]
    return 2.0 * np.cross(GRAD_REF[b], GRAD_REF[a])


def tet_geometry(X):
    """X (t,4,3) -> (detJ (t,), physical barycentric gradients (t,4,3))."""
    X = np.asarray(X, dtype=np.float64)
    J = np.stack([X[:, 0] - X[:, 3], X[:, 1] - X[:, 3], X[:, 2] - X[:, 3]], axis=2)
    det = np.linalg.det(J)
    Jinv = np.linalg.inv(J)                      # rows = grad lambda_0..2
    G = np.empty((X.shape[0], 4, 3))
    G[:, :3, :] = Jinv
    G[:, 3, :] = -Jinv.sum(axis=1)
    return det, G


def edge_signs(tet2node) -> np.ndarray:
    """(t,6) orientation signs of the reference edge functions vs low->high."""
    t = np.asarray(tet2node)
    s = np.empty((t.shape[0], 6))
    for e, (a, b) in enumerate(REF_EDGES):
        s[:, e] = np.where(t[:, b] < t[:, a], 1.0, -1.0)
    return s


def element_matrices(X, nu, kappa):
    """Closed-form K_e, M_e (t,6,6) for per-tet nu, kappa (arrays or scalars)."""
    det, G = tet_geometry(X)
    vol = np.abs(det) / 6.0
    nu = np.broadcast_to(np.asarray(nu, dtype=np.float64), vol.shape)
    kappa = np.broadcast_to(np.asarray(kappa, dtype=np.float64), vol.shape)
    curls = np.stack([2.0 * np.cross(G[:, b], G[:, a]) for a, b in REF_EDGES], axis=1)
    K = (nu * vol)[:, None, None] * np.einsum("tik,tjk->tij", curls, curls)
    g = np.einsum("tpk,tqk->tpq", G, G)          # grad lambda_p . grad lambda_q
    M = np.empty_like(K)
    for i, (a, b) in enumerate(REF_EDGES):
        for j, (c, d) in enumerate(REF_EDGES):
            M[:, i, j] = ((1 + (b == d)) * g[:, a, c] - (1 + (b == c)) * g[:, a, d]
                          - (1 + (a == d)) * g[:, b, c] + (1 + (a == c)) * g[:, b, d])
    M *= (kappa * vol / 20.0)[:, None, None]
    return K, M


def quad_points_bary(rule: str):
    """Barycentric points (q,4) and weights on the unit reference tet (sum 1/6)."""
    if rule == "1pt":
        return np.full((1, 4), 0.25), np.array([1.0 / 6.0])
    if rule == "4pt":
        P = np.full((4, 4), QA)
        np.fill_diagonal(P, QB)
        return P, np.full(4, 1.0 / 24.0)
    if rule.startswith("duffy"):
        m = int(rule[5:])
        x, w = np.polynomial.legendre.leggauss(m)
        x = 0.5 * (x + 1)
        w = 0.5 * w
        pts, wts = [], []
        for i in range(m):
            for j in range(m):
                for k in range(m):
                    u, v, s = x[i], x[j], x[k]
                    # collapsed map of the unit cube onto the tet
                    xi = u
                    eta = v * (1 - u)
                    zeta = s * (1 - u) * (1 - v)
                    pts.append([xi, eta, zeta, 1 - xi - eta - zeta])
                    wts.append(w[i] * w[j] * w[k] * (1 - u) ** 2 * (1 - v))
        return np.array(pts), np.array(wts)
    raise ValueError(rule)


def element_matrices_quadrature(X, nu, kappa, rule="4pt"):
    """K_e, M_e by quadrature with the reference-element map (D2/D3 fixed).  `rule` is a
    name of quad_points_bary or a (points (q,4) barycentric, weights (q,)) pair."""
    X = np.asarray(X, dtype=np.float64)
    det, _ = tet_geometry(X)
    J = np.stack([X[:, 0] - X[:, 3], X[:, 1] - X[:, 3], X[:, 2] - X[:, 3]], axis=2)
    JinvT = np.transpose(np.linalg.inv(J), (0, 2, 1))
    P, W = quad_points_bary(rule) if isinstance(rule, str) else rule
    t = X.shape[0]
    K = np.zeros((t, 6, 6))
    M = np.zeros((t, 6, 6))
    nu = np.broadcast_to(np.asarray(nu, dtype=np.float64), (t,))
    kappa = np.broadcast_to(np.asarray(kappa, dtype=np.float64), (t,))
    cref = np.stack([curl_ref(e) for e in range(6)])          # (6,3)
    cphys = np.einsum("tij,ej->tei", J, cref) / det[:, None, None]
    for q in range(P.shape[0]):
        xi = P[q, :3]
        wref = np.stack([whitney_ref(e, xi) for e in range(6)])   # (6,3)
        wphys = np.einsum("tij,ej->tei", JinvT, wref)
        f = W[q] * np.abs(det)
        K += (f * nu)[:, None, None] * np.einsum("tik,tjk->tij", cphys, cphys)
        M += (f * kappa)[:, None, None] * np.einsum("tik,tjk->tij", wphys, wphys)
    return K, M


def element_rhs(X, jq):
    """F_e (t,6) from J sampled at the 4-point rule nodes, jq (t,4,3)."""
    det, G = tet_geometry(X)
    vol = np.abs(det) / 6.0
    P, _ = quad_points_bary("4pt")
    F = np.zeros((X.shape[0], 6))
    for q in range(4):
        lam = P[q]
        for e, (a, b) in enumerate(REF_EDGES):
            w = lam[b] * G[:, a] - lam[a] * G[:, b]
            F[:, e] += (vol / 4.0) * np.einsum("tk,tk->t", w, jq[:, q])
    return F


def element_rhs_function(X, func, rule="duffy6"):
    """F_e with J given as a callable of physical points (brute-force check)."""
    X = np.asarray(X, dtype=np.float64)
    det, G = tet_geometry(X)
    P, W = quad_points_bary(rule)
    F = np.zeros((X.shape[0], 6))
    for q in range(P.shape[0]):
        lam = P[q]
        x = np.einsum("k,tkd->td", lam, X)
        jv = func(x)
        for e, (a, b) in enumerate(REF_EDGES):
            w = lam[b] * G[:, a] - lam[a] * G[:, b]
            F[:, e] += W[q] * np.abs(det) * np.einsum("tk,tk->t", w, jv)
    return F


def gradient_incidence() -> np.ndarray:
    """G_e (6,4): coefficients of grad(phi_node) in the reference edge basis.

    grad lambda_n = sum_e G[e, n] W_e, so K_e @ G = 0 (curl grad = 0).  For edge
    (a, b) with circulation +1 from b to a, the circulation of grad lambda_n
    is lambda_n(a) - lambda_n(b).
    """
    G = np.zeros((6, 4))
    for e, (a, b) in enumerate(REF_EDGES):
        G[e, a] += 1.0
        G[e, b] -= 1.0
    return G

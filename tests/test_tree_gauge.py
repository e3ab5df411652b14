"""Tree-cotree gauge for kappa = 0 (emag_tree_gauge; EMDOFManager::applyTreeGauge,
src/fe_em/em_dof_manager.cpp:448-495, TreeGauge include/fe_em/tree_gauge.hpp:131-293).

Pinned by what the mathematics fixes: with the PEC boundary as one contracted root the
tree has one edge per interior node and no cycle; the kappa = 0 curl-curl matrix has a
null space of exactly that dimension (the gradients of interior hat functions), and on
the cotree edges it is SPD; the gauged solution has the curl (face circulations) of the
kappa -> 0 limit of the ungauged problem, and gauging one edge too many breaks that."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from emagddm_amd import workloads as W
from emagddm_amd.assembly import tree_gauge_pairs
from oracle import assembly


def _interior(p):
    return np.all((p.xyz > 1e-12) & (p.xyz < 1 - 1e-12), axis=1)


def _face_circulations(p, asm, x):
    """Circulation of x around every tet face (node triples, gauge invariant)."""
    full = np.zeros(asm["pairs"].shape[0])
    e = np.nonzero(asm["free"] >= 0)[0]
    full[e] = x[asm["free"][e]]
    key = {tuple(pr): k for k, pr in enumerate(map(tuple, np.sort(asm["pairs"], axis=1)))}
    circ = []
    for tet in p.tet2node:
        for face in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
            a, b, c = tet[list(face)]
            s = 0.0
            for u, v in ((a, b), (b, c), (c, a)):  # oriented u -> v; dofs run low -> high
                s += full[key[(min(u, v), max(u, v))]] * (1.0 if u < v else -1.0)
            circ.append(s)
    return np.array(circ)


@pytest.mark.parametrize("n,parts", [(4, 2), (6, 3)])
def test_tree_structure(n, parts):
    p = W.make_problem(n, parts, seed=2)
    t = tree_gauge_pairs(p.n_nodes, p.tet2node, p.pec_pairs)
    assert len(t) == _interior(p).sum()
    # no cycle: union-find over the tree edges plus the contracted boundary
    parent = np.arange(p.n_nodes + 1)
    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i
    root = p.n_nodes
    for v in np.nonzero(~_interior(p))[0]:
        parent[find(v)] = find(root)
    pec = {tuple(sorted(e)) for e in p.pec_pairs.tolist()}
    for a, b in t.tolist():
        assert tuple(sorted((a, b))) not in pec
        ra, rb = find(a), find(b)
        assert ra != rb, "tree edge closes a cycle"
        parent[ra] = rb
    assert len({find(v) for v in range(p.n_nodes)}) == 1  # spans every node


def test_kappa0_nullspace_and_gauged_spd():
    p = W.make_problem(4, 2, seed=2)
    nu, zero = p.nu, np.zeros_like(p.kappa)
    A0 = assembly.assemble(p.xyz, p.tet2node, p.tet_material, nu, zero, p.pec_pairs, p.jq)["A"]
    ev = np.linalg.eigvalsh(A0.toarray())
    n_int = _interior(p).sum()
    assert (ev < 1e-10 * ev.max()).sum() == n_int  # gradients of interior hat functions
    t = tree_gauge_pairs(p.n_nodes, p.tet2node, p.pec_pairs)
    Ag = assembly.assemble(p.xyz, p.tet2node, p.tet_material, nu, zero,
                           np.vstack([p.pec_pairs, t]), p.jq)["A"]
    assert Ag.shape[0] == A0.shape[0] - n_int
    evg = np.linalg.eigvalsh(Ag.toarray())
    assert evg.min() > 1e-8 * evg.max()


def test_gauged_solution_reproduces_the_curl():
    """A compatible source f = K y (range of curl-curl, so G^T f = 0 exactly): the gauged
    kappa = 0 solution has the curl of y (face circulations, gauge invariant), whatever
    gradient y carries; gauging one free edge too many over-constrains the cotree space
    and the curl is lost.  (A loop source sampled at quadrature points is compatible only
    up to O(h^4), see tests/test_oracle_pins.py; its incompatible part has no kappa = 0
    solution.)"""
    p = W.make_problem(6, 2, seed=4)
    t = tree_gauge_pairs(p.n_nodes, p.tet2node, p.pec_pairs)
    zero = np.zeros_like(p.kappa)
    full = assembly.assemble(p.xyz, p.tet2node, p.tet_material, p.nu, zero, p.pec_pairs, p.jq)
    y = np.random.default_rng(0).standard_normal(full["A"].shape[0])
    f_full = full["A"] @ y
    g = assembly.assemble(p.xyz, p.tet2node, p.tet_material, p.nu, zero,
                          np.vstack([p.pec_pairs, t]), p.jq)
    # the gauged system's free dofs are a subset of the ungauged ones, in the same order
    keep = np.zeros(full["A"].shape[0], bool)
    e_g = np.nonzero(g["free"] >= 0)[0]
    keep[full["free"][e_g]] = True
    xg = spla.spsolve(g["A"].tocsc(), f_full[keep])
    cy = _face_circulations(p, full, y)
    cg = _face_circulations(p, g, xg)
    assert np.linalg.norm(cg - cy) <= 1e-9 * np.linalg.norm(cy)
    # K x_g reproduces f on every free edge, tree edges included
    xfull = np.zeros_like(y)
    xfull[keep] = xg
    assert np.linalg.norm(full["A"] @ xfull - f_full) <= 1e-9 * np.linalg.norm(f_full)
    # one free edge gauged too many: the curl of y is out of reach
    free_idx = np.nonzero(g["free"] >= 0)[0]
    extra = g["pairs"][free_idx[len(free_idx) // 2]][None, :]
    b = assembly.assemble(p.xyz, p.tet2node, p.tet_material, p.nu, zero,
                          np.vstack([p.pec_pairs, t, extra]), p.jq)
    keep_b = np.zeros(full["A"].shape[0], bool)
    e_b = np.nonzero(b["free"] >= 0)[0]
    keep_b[full["free"][e_b]] = True
    xb = spla.spsolve(b["A"].tocsc(), f_full[keep_b])
    cb = _face_circulations(p, b, xb)
    assert np.linalg.norm(cb - cy) > 1e-3 * np.linalg.norm(cy)

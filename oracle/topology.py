"""Global edge ids and free-dof numbering -- FP64/int restatement (oracle).

* ``edge_ids_loop``: the reference algorithm verbatim in behaviour: walk the
  tets in order and their local edges in ElementGeometry order
  (src/fe_em/element_geometry.cpp:119-126); key = sorted global node pair; a
  new key gets the next id (src/fe_em/global_edge_id_generator.cpp:30-110).
  Pure-Python dict; small meshes only.
* ``edge_ids``: the same result with numpy (unique + first appearance).
* ``free_numbering``: EMDOFManager::build semantics
  (src/fe_em/em_dof_manager.cpp:91-393): prenumber = edge id order, PEC edges
  (DIRICHLET markers with VECTOR_EDGE_ONLY node pairs) -> -1, the remaining
  edges numbered 0.. contiguously in prenumber order.
* ``classify``: FETI-DP edge classes from the subdomain multiplicity:
  1 -> interior, 2 -> dual (one Lagrange multiplier), >=3 -> primal.
"""
from __future__ import annotations

import numpy as np

from .nedelec import GEOM_EDGES


def edge_ids_loop(tet2node):
    """Reference-order edge ids with a dict; returns (tet2edge (t,6), pairs (E,2))."""
    emap = {}
    t2e = np.empty((len(tet2node), 6), dtype=np.int64)
    for t, nodes in enumerate(np.asarray(tet2node).tolist()):
        for le, (a, b) in enumerate(GEOM_EDGES):
            ga, gb = nodes[a], nodes[b]
            key = (min(ga, gb), max(ga, gb))
            if key not in emap:
                emap[key] = len(emap)
            t2e[t, le] = emap[key]
    pairs = np.array(sorted(emap, key=emap.get), dtype=np.int64).reshape(-1, 2)
    return t2e, pairs


def edge_ids(tet2node):
    """Vectorised first-appearance edge ids; same output as ``edge_ids_loop``."""
    t = np.asarray(tet2node, dtype=np.int64)
    a = np.stack([t[:, [p, q]] for p, q in GEOM_EDGES], axis=1)   # (t,6,2)
    lo = a.min(axis=2).ravel()
    hi = a.max(axis=2).ravel()
    big = int(t.max()) + 1
    key = lo * big + hi
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    t2e = rank[inv].reshape(-1, 6)
    u = uniq[order]
    pairs = np.stack([u // big, u % big], axis=1)
    return t2e, pairs


def free_numbering(pairs, pec_pairs):
    """Free index per global edge (-1 for PEC), in edge-id (prenumber) order."""
    pairs = np.asarray(pairs, dtype=np.int64)
    big = int(pairs.max()) + 1 if pairs.size else 1
    key = pairs[:, 0] * big + pairs[:, 1]
    pp = np.sort(np.asarray(pec_pairs, dtype=np.int64).reshape(-1, 2), axis=1)
    pkey = pp[:, 0] * big + pp[:, 1] if pp.size else np.zeros(0, np.int64)
    constrained = np.isin(key, pkey)
    free = np.full(pairs.shape[0], -1, dtype=np.int64)
    nf = int((~constrained).sum())
    free[~constrained] = np.arange(nf)
    return free


def local_to_global(tet2node, pec_pairs):
    """Per-tet free dof indices (t,6) in ElementGeometry order, -1 = PEC."""
    t2e, pairs = edge_ids(tet2node)
    free = free_numbering(pairs, pec_pairs)
    return free[t2e], pairs, free


def classify(l2g, tet_subdomain, n_free):
    """Multiplicity per free dof and the (dof, subdomain) incidence pairs."""
    l2g = np.asarray(l2g)
    sub = np.repeat(np.asarray(tet_subdomain, dtype=np.int64), 6)
    d = l2g.ravel()
    ok = d >= 0
    key = np.unique(d[ok] * (int(sub.max()) + 1) + sub[ok])
    dofs = key // (int(sub.max()) + 1)
    subs = key % (int(sub.max()) + 1)
    mult = np.bincount(dofs, minlength=n_free)
    return mult, dofs, subs


def kuhn_edge_counts(a, b, c):
    """Closed-form edge count of an a x b x c Kuhn box and its surface edges."""
    E = (a * (b + 1) * (c + 1) + b * (a + 1) * (c + 1) + c * (a + 1) * (b + 1)
         + (a + 1) * b * c + a * (b + 1) * c + a * b * (c + 1) + a * b * c)

    def f2(p, q):
        return p * (q + 1) + q * (p + 1) + p * q

    S = 2 * (f2(a, b) + f2(a, c) + f2(b, c)) - 4 * (a + b + c)
    return E, S

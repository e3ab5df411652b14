"""Global free-dof assembly -- FP64 restatement (oracle).

EMAssembly::assemble (src/numeric/em_assembly.cpp:144-411) scatters each
element matrix into COO skipping constrained (-1) dofs
(scatterElementMatrix, :654-718), builds a CSR that keeps duplicates
(include/numeric/csr_matrix_impl.hpp:80-128) and the solvers sum them
(src/numeric/sparse_converter.cpp:60).  Here the duplicates are summed by
scipy, and the two reference defects on this path are corrected: the real
Nedelec integrator replaces DummyEMIntegrator (D5, :39-125) and the
orientation sign s_i s_j multiplies every entry (D6).
System: A = K + M (curl-curl plus kappa mass), f = F.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import nedelec, topology


def element_data(xyz, tet2node, tet_material, nu, kappa, jq=None):
    X = np.asarray(xyz)[np.asarray(tet2node)]
    m = np.asarray(tet_material)
    K, M = nedelec.element_matrices(X, np.asarray(nu)[m], np.asarray(kappa)[m])
    F = nedelec.element_rhs(X, jq) if jq is not None else None
    return K, M, F


def assemble(xyz, tet2node, tet_material, nu, kappa, pec_pairs, jq=None):
    """Returns dict(A=csr (n_free^2), f, l2g, signs, Ke, Me, Fe, pairs, free)."""
    l2g, pairs, free = topology.local_to_global(tet2node, pec_pairs)
    K, M, F = element_data(xyz, tet2node, tet_material, nu, kappa, jq)
    s = nedelec.edge_signs(tet2node)
    n = int(free.max()) + 1
    Ae = (K + M) * s[:, :, None] * s[:, None, :]
    rows = np.repeat(l2g, 6, axis=1)
    cols = np.tile(l2g, (1, 6))
    ok = (rows >= 0) & (cols >= 0)
    A = sp.csr_matrix((Ae.reshape(-1, 36)[ok], (rows[ok], cols[ok])), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    f = None
    if F is not None:
        Fs = F * s
        okf = l2g >= 0
        f = np.bincount(l2g[okf], weights=Fs[okf], minlength=n)
    return dict(A=A, f=f, l2g=l2g, signs=s, Ke=K, Me=M, Fe=F, pairs=pairs, free=free)

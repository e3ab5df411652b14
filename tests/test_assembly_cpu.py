"""Symbolic part of the GPU global assembly (host only): the CSR pattern, the
Local2Global map and the orientation signs against the oracle
(EMAssembly::assemble with D5/D6 corrected, src/numeric/em_assembly.cpp:144-411;
EMDOFManager free numbering, src/fe_em/em_dof_manager.cpp:91-393)."""
import numpy as np
import pytest

from emagddm_amd import workloads as W
from emagddm_amd.assembly import GlobalAssembly
from oracle import assembly


@pytest.mark.parametrize("n,parts,jitter,perturb", [(4, 2, None, False), (8, 2, (3, 5), False),
                                                    (6, 3, None, True)])
def test_pattern_matches_oracle(n, parts, jitter, perturb):
    p = W.make_problem(n, parts, jitter=jitter, seed=7)
    t2n = p.tet2node
    if perturb:  # shuffled tets and odd node orders inside each tet
        rng = np.random.default_rng(1)
        t2n = t2n[rng.permutation(len(t2n))]
        t2n = t2n[:, [2, 0, 3, 1]]
    g = GlobalAssembly(p.n_nodes, t2n, p.pec_pairs)
    rp, col, tf, ts = g.pattern()
    asm = assembly.assemble(p.xyz, t2n, np.zeros(len(t2n), np.int32), p.nu[:1], p.kappa[:1],
                            p.pec_pairs)
    A = asm["A"]
    assert g.n_free == A.shape[0] and g.nnz == A.nnz
    assert np.array_equal(rp, A.indptr) and np.array_equal(col, A.indices)
    assert np.array_equal(tf, asm["l2g"]) and np.array_equal(ts, asm["signs"])


def test_create_rejects_bad_input():
    p = W.make_problem(4, 2, seed=1)
    t2n = p.tet2node.copy()
    t2n[0, 0] = p.n_nodes + 5
    with pytest.raises(RuntimeError, match="emag_assembly_create failed"):
        GlobalAssembly(p.n_nodes, t2n, p.pec_pairs)

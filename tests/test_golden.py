"""Oracle vs golden vectors written by the reference's own Eigen LLT and CGSolver
(tests/golden/make_golden.py, from oracle/_ref).  CPU only."""
import os

import numpy as np
import pytest

from emagddm_amd import workloads as W
from oracle import assembly, solvers

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = {
    "micro4": lambda: W.make_problem(4, 2, seed=3, name="micro"),
    "micro8c": lambda: W.make_problem(8, 2, materials="contrast", seed=5, name="micro_c"),
    "cube16": lambda: W.make_workload("cube16"),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_reference_solvers(name):
    g = np.load(os.path.join(HERE, f"{name}.npz"))
    p = CASES[name]()
    # the seeded generator still produces the same inputs
    assert np.isclose(p.jq.sum(), g["jq_sum"], rtol=1e-12, atol=1e-12)
    assert np.isclose(p.xyz.sum(), g["xyz_sum"], rtol=1e-14)
    asm = assembly.assemble(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.pec_pairs, p.jq)
    A, f = asm["A"], asm["f"]
    assert A.shape[0] == int(g["n_free"]) and A.nnz == int(g["nnz"])
    assert np.isclose(np.linalg.norm(f), g["f_norm"], rtol=1e-13)
    assert np.isclose(A.data.sum(), g["a_sum"], rtol=1e-12)
    x = solvers.direct(A, f)
    xr = g["x_llt"]
    assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)
    if int(g["cg_status"]) == 0:
        xc, st, it, res = solvers.pcg_jacobi(A, f, 1e-8, 5000)
        assert st == 0 and abs(it - int(g["cg_iterations"])) <= 3
        if "x_cg" in g:
            assert np.linalg.norm(xc - g["x_cg"]) <= 1e-6 * np.linalg.norm(g["x_cg"])

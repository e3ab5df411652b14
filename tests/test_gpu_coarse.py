"""Grouped coarse solve (S~ = L L^T, forward and backward sweeps over groups of tile
columns with explicit inverses of the groups' diagonal blocks) against the dense coarse
inverse and the oracle: the same dual operator F, the same solution."""
import numpy as np
import pytest

from emagddm_amd import FETIDPSolver, SolverStatus, workloads as W
from oracle import fetidp as ofetidp

pytestmark = pytest.mark.gpu


def _solver(p, **kw):
    s = FETIDPSolver(**kw).setup_problem(p)
    s.set_rhs_tet_J(p.jq)
    s.factor()
    return s


@pytest.mark.parametrize("group", [1, 2, 3, 8])
def test_grouped_coarse_equals_dense(gpu_available, group):
    # 16^3 hexes, 4^3 boxes: n_c = 432 primal edges, 7 coarse tile columns
    p = W.make_problem(16, 4, seed=3, name="c16x4")
    print(f"[coarse] group={group}", flush=True)
    sd = _solver(p, coarse_solve="dense")
    sg = _solver(p, coarse_solve="grouped", coarse_group=group)
    assert sd.sizes["n_primal"] == sg.sizes["n_primal"] == 432
    rng = np.random.default_rng(group)
    for _ in range(3):
        v = rng.standard_normal(sd.sizes["n_lambda"])
        qd, qg = sd.apply_dual_operator(v), sg.apply_dual_operator(v)
        assert np.linalg.norm(qg - qd) <= 1e-12 * np.linalg.norm(qd)
    rd, rg = sd.solve(), sg.solve()
    assert rd.status == rg.status == SolverStatus.SUCCESS
    assert rg.iterations == rd.iterations
    assert np.linalg.norm(rg.x - rd.x) <= 1e-10 * np.linalg.norm(rd.x)


def test_grouped_coarse_matches_oracle(gpu_available):
    p = W.make_problem(8, 2, seed=3, name="micro8")
    s = _solver(p, coarse_solve="grouped", coarse_group=1, scaling="stiffness")
    o = ofetidp.DenseFETIDP(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.tet_subdomain,
                            p.pec_pairs, p.jq, scaling="stiffness")
    rng = np.random.default_rng(0)
    v = rng.standard_normal(s.sizes["n_lambda"])
    q = s.apply_dual_operator(v)
    qo = o.apply_F(v)
    assert np.linalg.norm(q - qo) <= 1e-10 * np.linalg.norm(qo)


def test_grouped_coarse_config_switch_and_graphs(gpu_available):
    """Switching the coarse solve on a live handle rebuilds the plan; graph replay gives
    bitwise the eager iterates."""
    p = W.make_problem(16, 4, jitter=(3, 5), seed=5, name="c16j")
    s = FETIDPSolver(coarse_solve="dense", graphs=False).setup_problem(p)
    s.set_rhs_tet_J(p.jq)
    s.factor()
    xd = s.solve().x
    s.set_config_values(coarse_solve=2, coarse_group=2)
    s.factor()
    r_eager = s.solve()
    s.set_config_values(graphs=1)
    r_graph = s.solve()
    r_graph2 = s.solve()
    assert r_eager.status == SolverStatus.SUCCESS
    assert np.array_equal(r_eager.x, r_graph.x) and np.array_equal(r_graph.x, r_graph2.x)
    assert np.linalg.norm(r_eager.x - xd) <= 1e-10 * np.linalg.norm(xd)


@pytest.fixture(scope="module")
def c16x4_oracle():
    p = W.make_problem(16, 4, seed=3, name="c16x4")
    return p, ofetidp.DenseFETIDP(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa,
                                  p.tet_subdomain, p.pec_pairs, p.jq, scaling="stiffness")


@pytest.mark.parametrize("coarse", ["dense", "grouped"])
def test_coarse_solve_matches_oracle_several_tiles(gpu_available, c16x4_oracle, coarse):
    """n_c = 432 (7 coarse tile columns): F through the dense -S~^-1 tiles (diagonal tiles
    read through their lower triangle) or the grouped sweeps equals the oracle's F, whose
    coarse solve is a dense LAPACK solve with S~."""
    p, o = c16x4_oracle
    print(f"[coarse] oracle F, {coarse}", flush=True)
    s = _solver(p, coarse_solve=coarse, coarse_group=2, scaling="stiffness")
    assert s.sizes["n_primal"] == 432
    rng = np.random.default_rng(11)
    for _ in range(2):
        v = rng.standard_normal(s.sizes["n_lambda"])
        q, qo = s.apply_dual_operator(v), o.apply_F(v)
        assert np.linalg.norm(q - qo) <= 1e-10 * np.linalg.norm(qo)

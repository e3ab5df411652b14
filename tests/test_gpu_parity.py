"""GPU parity of the FETI-DP path against the FP64 oracle (through the C ABI)."""
import numpy as np
import pytest

from emagddm_amd import FETIDPSolver, SolverStatus, integrate_elements, workloads as W
from oracle import assembly, fetidp as ofetidp, nedelec, solvers

pytestmark = pytest.mark.gpu


def _problem(kind):
    if kind == "micro":
        return W.make_problem(4, 2, seed=3, name="micro")
    if kind == "micro_contrast":
        return W.make_problem(8, 2, materials="contrast", seed=5, name="micro_c")
    return W.make_workload(kind)


@pytest.mark.parametrize("kind", ["micro", "cube16"])
def test_element_kernel_matches_oracle(gpu_available, kind):
    p = _problem(kind)
    print(f"[element] {kind} tets={p.n_tets}", flush=True)
    Ke, Me, Fe = integrate_elements(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.jq)
    K, M = nedelec.element_matrices(p.xyz[p.tet2node], p.nu[p.tet_material], p.kappa[p.tet_material])
    F = nedelec.element_rhs(p.xyz[p.tet2node], p.jq)
    iu = np.triu_indices(6)
    Ko, Mo = K[:, iu[0], iu[1]], M[:, iu[0], iu[1]]
    Kg, Mg, Fg = Ke.cpu().numpy(), Me.cpu().numpy(), Fe.cpu().numpy()
    assert np.abs(Kg - Ko).max() <= 1e-13 * np.abs(Ko).max()
    assert np.abs(Mg - Mo).max() <= 1e-13 * np.abs(Mo).max()
    assert np.abs(Fg - F).max() <= 1e-13 * np.abs(F).max()


def test_global_system_matches_oracle(gpu_available):
    p = _problem("cube16")
    s = FETIDPSolver().setup_problem(p)
    s.set_rhs_tet_J(p.jq)
    s.factor()
    rp, col, val, f = s.global_system()
    asm = assembly.assemble(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.pec_pairs, p.jq)
    A = asm["A"]
    assert np.array_equal(rp, A.indptr) and np.array_equal(col, A.indices)
    assert np.abs(val - A.data).max() <= 1e-13 * np.abs(A.data).max()
    assert np.abs(f - asm["f"]).max() <= 1e-13 * np.abs(asm["f"]).max()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dual_operators_match_oracle(gpu_available, precision):
    p = _problem("micro")
    s = FETIDPSolver(precision=precision).setup_problem(p)
    s.set_rhs_tet_J(p.jq)
    s.factor()
    o = ofetidp.DenseFETIDP(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.tet_subdomain,
                            p.pec_pairs, p.jq)
    tol = 1e-11 if precision == "fp64" else 2e-5
    rng = np.random.default_rng(0)
    v = rng.normal(size=o.n_lam)
    Fv, Fo = s.apply_dual_operator(v), o.apply_F(v)
    assert np.linalg.norm(Fv - Fo) <= tol * np.linalg.norm(Fo)
    Mv, Mo = s.apply_preconditioner(v), o.apply_prec(v)
    assert np.linalg.norm(Mv - Mo) <= tol * np.linalg.norm(Mo)
    d = s.dual_rhs()
    assert np.linalg.norm(d - o.d) <= tol * np.linalg.norm(o.d)
    Sc = s.coarse_matrix()
    assert np.linalg.norm(Sc - o.Sc) <= tol * np.linalg.norm(o.Sc)


@pytest.mark.parametrize("kind,precision", [("micro", "fp64"), ("cube16", "fp64"),
                                            ("cube16", "fp32"), ("micro_contrast", "fp64")])
def test_solve_matches_direct(gpu_available, kind, precision):
    p = _problem(kind)
    print(f"[solve] {kind} {precision}", flush=True)
    s = FETIDPSolver(precision=precision).setup_problem(p)
    s.set_rhs_tet_J(p.jq)
    s.factor()
    res = s.solve()
    asm = assembly.assemble(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.pec_pairs, p.jq)
    xd = solvers.direct(asm["A"], asm["f"])
    err = np.linalg.norm(res.x - xd) / np.linalg.norm(xd)
    print(f"  status={res.status.name} its={res.iterations} ref={res.refinements} "
          f"res={res.residual_norm:.3e} err={err:.3e}", flush=True)
    assert res.status == SolverStatus.SUCCESS, res.error_msg
    assert res.residual_norm <= 1e-8
    assert solvers.rel_residual(asm["A"], res.x, asm["f"]) <= 1.01e-8
    assert err <= (1e-8 if precision == "fp64" else 1e-7)

"""Failure statuses of fetidp_factor / fetidp_solve on the GPU, each forced from the
host with bounded inputs so every kernel exits normally.

The contract is the reference CGSolver's (src/numeric/iterative_solvers.cpp:424-503):
b = 0 gives x = 0 after 0 iterations with residual 0 (:436-442); hitting the
iteration cap gives MAX_ITER_REACHED with iterations = cap (:487-495); a matrix the
Cholesky factorization cannot take is NUMERICAL_ERROR, as the LLT backend reports it
through UnifiedDirectSolver (include/numeric/eigen_solver_backends.h:159-170,
src/numeric/unified_direct_solver.cpp:122-136), and errors never raise across the ABI
(test/test_eigen_solver_backends_crtp.cpp:345-395 checks the same for the CPU backends).
"""
import numpy as np
import pytest

from emagddm_amd import FETIDPSolver, SolverStatus, workloads as W
from oracle import assembly, solvers

pytestmark = pytest.mark.gpu


def _small():
    return W.make_problem(8, 2, seed=1, name="status8")


def test_max_iterations_one(gpu_available):
    p = _small()
    print("[status] max_iterations=1", flush=True)
    s = FETIDPSolver(max_iterations=1).setup_problem(p)
    s.set_rhs_tet_J(p.jq)
    s.factor()
    r = s.solve()
    assert r.status == SolverStatus.MAX_ITER_REACHED, r
    assert r.iterations == 1
    assert r.residual_norm > 1e-8 and np.isfinite(r.residual_norm)
    assert r.error_msg == "" or isinstance(r.error_msg, str)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_zero_source(gpu_available, precision):
    p = _small()
    print(f"[status] b = 0 {precision}", flush=True)
    s = FETIDPSolver(precision=precision).setup_problem(p)
    s.set_rhs_free(np.zeros(s.sizes["n_free"]))
    s.factor()
    r = s.solve()
    assert r.status == SolverStatus.SUCCESS
    assert r.iterations == 0 and r.residual_norm == 0.0
    assert np.array_equal(r.x, np.zeros_like(r.x))
    # the same handle then solves a real source with the existing factor
    s.set_rhs_tet_J(p.jq)
    r2 = s.solve()
    assert r2.status == SolverStatus.SUCCESS and r2.residual_norm <= 1e-8


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_indefinite_material_is_numerical_error(gpu_available, precision):
    """kappa < 0 makes every K_rr indefinite (the gradient modes have energy kappa):
    the factorization reports NUMERICAL_ERROR, FP32 after its FP64 retry (fell_back)."""
    p = _small()
    print(f"[status] kappa < 0 {precision}", flush=True)
    s = FETIDPSolver(precision=precision).setup(p.xyz, p.tet2node, p.tet_material, p.nu,
                                                -0.5 * np.ones_like(p.kappa),
                                                p.tet_subdomain, p.pec_pairs)
    s.set_rhs_tet_J(p.jq)
    with pytest.raises(RuntimeError, match="NUMERICAL_ERROR.*not positive definite"):
        s.factor()
    r = s.solve()  # no factor: INVALID_INPUT, no exception
    assert r.status == SolverStatus.INVALID_INPUT and "factor" in r.error_msg
    assert r.fell_back_to_fp64 == (precision == "fp32")
    # valid materials on the same handle recover
    s.set_materials(p.nu, p.kappa)
    s.factor()
    r = s.solve()
    assert r.status == SolverStatus.SUCCESS and r.residual_norm <= 1e-8


def test_nonfinite_source_is_invalid_input(gpu_available):
    p = _small()
    print("[status] NaN source", flush=True)
    s = FETIDPSolver().setup_problem(p)
    f = np.zeros(s.sizes["n_free"])
    f[3] = np.nan
    s.set_rhs_free(f)
    s.factor()  # the matrix is fine; only the source carries the NaN
    r = s.solve()
    assert r.status == SolverStatus.INVALID_INPUT and "not finite" in r.error_msg


def test_fp32_gate_and_stall_fallback(gpu_available):
    """FP32_REFINE either converges or hands over to FP64 (fell_back) with the FP64
    answer: on a 24^3 cube (FP32 accuracy estimate 3.4e-3, under the gate), on an 8^3
    contrast problem (estimate 4.1e-3, under the gate) and on a 16^3 contrast problem
    whose estimate (6.6e-2) is far above it, which must factor in FP64 from the start."""
    for p in (W.make_problem(24, 2, seed=2, name="fp32_24"),
              W.make_problem(8, 2, materials="contrast", seed=5, name="fp32_c8"),
              W.make_problem(16, 2, materials="contrast", seed=5, name="fp32_c")):
        print(f"[status] fp32 {p.name}", flush=True)
        s = FETIDPSolver(precision="fp32").setup_problem(p)
        s.set_rhs_tet_J(p.jq)
        s.factor()
        r = s.solve()
        assert r.status == SolverStatus.SUCCESS and r.residual_norm <= 1e-8, r
        asm = assembly.assemble(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa,
                                p.pec_pairs, p.jq)
        xd = solvers.direct(asm["A"], asm["f"])
        err = np.linalg.norm(r.x - xd) / np.linalg.norm(xd)
        print(f"[status] {p.name}: its={r.iterations} fell_back={r.fell_back_to_fp64} "
              f"err={err:.2e}", flush=True)
        assert err <= 1e-7
        if p.name == "fp32_c":
            assert r.fell_back_to_fp64

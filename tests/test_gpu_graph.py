"""PCPG iterations replayed from a captured HIP graph give bitwise the same
iterates as the same kernels launched eagerly (every replay is checked)."""
import numpy as np
import pytest

from emagddm_amd import FETIDPSolver, SolverStatus, workloads as W

pytestmark = pytest.mark.gpu


def _solve(p, precision, graphs, repeats=1, tol=1e-8):
    s = FETIDPSolver(precision=precision, graphs=graphs, tolerance=tol)
    s.setup_problem(p)
    s.set_rhs_tet_J(p.jq)
    s.factor()
    out = []
    for _ in range(repeats):
        r = s.solve()
        out.append((r, s.history(), s.get_lambda()))
    return s, out


@pytest.mark.parametrize("kind,precision", [("cube16", "fp64"), ("contrast", "fp64"),
                                            ("cube16", "fp32")])
def test_graph_replays_match_eager(gpu_available, kind, precision):
    p = (W.make_workload("cube16") if kind == "cube16"
         else W.make_problem(8, 2, materials="contrast", seed=5, name="micro_c"))
    print(f"[graph] {kind} {precision}", flush=True)
    _, eager = _solve(p, precision, graphs=False)
    _, graph = _solve(p, precision, graphs=True, repeats=3)
    r0, h0, l0 = eager[0]
    assert r0.status == SolverStatus.SUCCESS
    for r, h, l in graph:  # the graph is captured once and replayed in every solve
        assert r.status == r0.status and r.iterations == r0.iterations
        assert np.array_equal(h, h0)
        assert np.array_equal(l, l0)
        assert np.array_equal(r.x, r0.x)


def test_graph_rebuilt_after_new_setup(gpu_available):
    a = W.make_problem(8, 2, seed=1, name="a")
    b = W.make_workload("cube16", 2)
    s = FETIDPSolver(graphs=True)
    for p in (a, b):
        s.setup_problem(p)
        s.set_rhs_tet_J(p.jq)
        s.factor()
        r = s.solve()
        _, ref = _solve(p, "fp64", graphs=False)
        assert r.status == SolverStatus.SUCCESS
        assert np.array_equal(r.x, ref[0][0].x)

"""Autotuner: every launch variant returns bitwise the same solution (checked inside
fetidp_autotune, which fails otherwise), the fastest is kept, and solves after tuning
equal an untuned solver's; for a grouped coarse solve the group widths are timed too
and the fastest that reaches the tolerance is kept."""
import numpy as np
import pytest

from emagddm_amd import FETIDPSolver, SolverStatus, workloads as W

pytestmark = pytest.mark.gpu


def test_autotune_keeps_results(gpu_available):
    p = W.make_workload("cube16", 1)
    print("[autotune] cube16", flush=True)
    s = FETIDPSolver()
    s.setup_problem(p)
    s.set_rhs_tet_J(p.jq)
    rows = s.autotune(repetitions=1)
    assert len(rows) == 12
    assert {(a, b, c) for a, b, c, _, _ in rows} == {(a, b, c) for a in (0, 1, 2) for b in (0, 1)
                                                      for c in (0, 4)}
    assert all(g == 0 for _, _, _, g, _ in rows)  # dense coarse inverse at this size
    best = min(rows, key=lambda r: r[4])
    cfg = s.get_config()
    assert (cfg.side_stream, cfg.graphs, cfg.tile_waves) == best[:3]
    r = s.solve()
    ref = FETIDPSolver()
    ref.setup_problem(p)
    ref.set_rhs_tet_J(p.jq)
    ref.factor()
    r0 = ref.solve()
    assert r.status == SolverStatus.SUCCESS and np.array_equal(r.x, r0.x)


def test_autotune_coarse_group_width(gpu_available):
    p = W.make_problem(24, 6, seed=3, name="c24x6")  # 1856 coarse rows: 29 tile columns
    print("[autotune] c24x6 grouped coarse", flush=True)
    s = FETIDPSolver(coarse_solve="grouped", coarse_group=4)
    s.setup_problem(p)
    s.set_rhs_tet_J(p.jq)
    rows = s.autotune(repetitions=1)
    assert len(rows) == 12 + 3 and all(r[3] == 4 for r in rows[:12])
    assert [r[3] for r in rows[12:]] == [8, 16, 24]
    launch = min(rows[:12], key=lambda r: r[4])
    assert all(r[:3] == launch[:3] for r in rows[12:])
    best = min(rows, key=lambda r: r[4])
    cfg = s.get_config()
    assert cfg.coarse_group == best[3]
    r = s.solve()
    ref = FETIDPSolver(coarse_solve="dense")
    ref.setup_problem(p)
    ref.set_rhs_tet_J(p.jq)
    ref.factor()
    r0 = ref.solve()
    assert r.status == SolverStatus.SUCCESS and r.residual_norm <= 1e-8
    # another summation order, the same converged solution (tolerance 1e-8 on the residual)
    assert np.linalg.norm(r.x - r0.x) <= 1e-6 * np.linalg.norm(r0.x)

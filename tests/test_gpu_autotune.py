"""Autotuner: every launch variant returns bitwise the same solution (checked inside
fetidp_autotune, which fails otherwise), the fastest is kept, and solves after tuning
equal an untuned solver's."""
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
    assert {(a, b, c) for a, b, c, _ in rows} == {(a, b, c) for a in (0, 1, 2) for b in (0, 1)
                                                   for c in (0, 4)}
    best = min(rows, key=lambda r: r[3])
    cfg = s.get_config()
    assert (cfg.side_stream, cfg.graphs, cfg.tile_waves) == best[:3]
    r = s.solve()
    ref = FETIDPSolver()
    ref.setup_problem(p)
    ref.set_rhs_tet_J(p.jq)
    ref.factor()
    r0 = ref.solve()
    assert r.status == SolverStatus.SUCCESS and np.array_equal(r.x, r0.x)

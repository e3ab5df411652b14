"""The host index arithmetic (topology.cpp, plan.cpp: edge numbering, classes, padded
layouts, envelopes, assembly plans, factorization schedules) and the problem-file
reader, built with AddressSanitizer + UndefinedBehaviorSanitizer (host code only, g++)
and run on box, jittered, Voronoi and single-subdomain partitions.  The sizes must match
the product library's setup of the same problem."""
import json
import os
import subprocess

import numpy as np
import pytest

from emagddm_amd import FETIDPSolver, workloads as W
from emagddm_amd.problem_io import write_problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "emagddm_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("san") / "plans_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           f"-I{ROOT}/include", f"-I{CSRC}",
           os.path.join(ROOT, "tests", "sanitize", "plans_check.cpp"),
           os.path.join(CSRC, "topology.cpp"), os.path.join(CSRC, "plan.cpp"),
           os.path.join(CSRC, "problem_io.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


CASES = {
    "box": lambda: W.make_problem(8, 2, seed=1),
    "jittered": lambda: W.make_problem(12, 3, jitter=(3, 5), seed=2),
    "contrast": lambda: W.make_problem(8, 2, materials="contrast", seed=5),
    "voronoi": lambda: W.irregular_problem(8, 5, seed=3),
    "one_subdomain": lambda: W.make_problem(6, 1, seed=4),
    "slabs": lambda: W.make_problem(8, 1, seed=6),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_plans_clean_under_sanitizers(checker, tmp_path, name):
    p = CASES[name]()
    if name == "slabs":  # four slabs along x: no primal dofs, empty coarse problem
        cent = p.xyz[p.tet2node].mean(axis=1)
        p.tet_subdomain = np.minimum((cent[:, 0] * 4).astype(np.int32), 3)
    path = str(tmp_path / f"{name}.emfd")
    write_problem(path, p)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([checker, path], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout, r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    got = json.loads(r.stdout.strip().splitlines()[-1])
    s = FETIDPSolver().setup_problem(p)
    z = s.sizes
    assert (got["n_edges"], got["n_free"], got["n_lambda"], got["n_primal"], got["n_sub"]) == \
        (z["n_edges"], z["n_free"], z["n_lambda"], z["n_primal"], z["n_sub"])
    assert got["pool_elems"] == z["pool_elems"]
    # the Schur updates write S_DD and W themselves: on box partitions every D-block tile
    # has interior coupling (no snapshot copies); on Voronoi subdomains (one without
    # interior edges) some D-block tiles are copied by the snapshot launches
    assert got["w_fused"] == 1
    if name == "voronoi":
        assert 0 < got["snap_tiles"] < got["d_tiles"]
    elif name in ("box", "jittered", "contrast"):
        assert got["snap_tiles"] == 0 and got["d_tiles"] > 0

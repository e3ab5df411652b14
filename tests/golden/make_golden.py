"""Writes golden vectors from the reference's own solvers (oracle/_ref).

Run in the build container only (it needs oracle/_ref/libref_solvers.so, built
from /root/reference by oracle/ref/Makefile).  The fixtures are data: the
seeded problem's input checksums and the reference solvers' outputs.

    python tests/golden/make_golden.py            # micro4, micro8c, cube16
    python tests/golden/make_golden.py cube48j    # full-size summary (minutes)
    python tests/golden/make_golden.py cg cube96j 20000   # reference CG iteration count
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from emagddm_amd import workloads as W  # noqa: E402
from oracle import assembly  # noqa: E402


def ref_lib():
    return C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_solvers.so"))


def run(lib, p, cg_max=5000):
    asm = assembly.assemble(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.pec_pairs, p.jq)
    A, f = asm["A"], asm["f"]
    n = A.shape[0]
    rp, col, val = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    x = np.zeros(n)
    fm, sm = C.c_double(), C.c_double()
    st = lib.ref_solve_llt(n, ip(rp), ip(col), dp(val), dp(f), dp(x), C.byref(fm), C.byref(sm))
    assert st == 0
    xc = np.zeros(n)
    it, res = C.c_int(), C.c_double()
    st2 = lib.ref_solve_cg(n, ip(rp), ip(col), dp(val), dp(f), C.c_double(1e-8), cg_max, dp(xc),
                           C.byref(it), C.byref(res), C.byref(sm))
    return dict(x_llt=x, x_cg=xc, cg_status=st2, cg_iterations=it.value, cg_residual=res.value,
                n_free=n, nnz=A.nnz, f_norm=np.linalg.norm(f), a_sum=A.data.sum(),
                jq_sum=p.jq.sum(), xyz_sum=p.xyz.sum())


N_SAMPLE, N_PROJ = 512, 8


def summary(x, seed=48):
    """A full-size solution as a small fixture: its 2-norm, its entries at N_SAMPLE
    seeded indices, and its inner products with N_PROJ seeded standard-normal vectors
    (regenerated by the test from the same seeds)."""
    n = x.size
    idx = np.sort(np.random.default_rng(seed).choice(n, N_SAMPLE, replace=False))
    proj = np.array([np.random.default_rng(seed + 1 + k).standard_normal(n) @ x
                     for k in range(N_PROJ)])
    return dict(x_norm=np.linalg.norm(x), sample_idx=idx, sample_x=x[idx], proj=proj,
                proj_seed=seed)


def main_large(name="cube48j"):
    """The reference Eigen LLT (through UnifiedDirectSolver) on the full 48^3 system,
    kept as a summary (the 753,552-entry solution itself is 6 MB)."""
    lib = ref_lib()
    p = W.make_workload(name)
    asm = assembly.assemble(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.pec_pairs, p.jq)
    A, f = asm["A"], asm["f"]
    n = A.shape[0]
    rp, col, val = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    x = np.zeros(n)
    fm, sm = C.c_double(), C.c_double()
    st = lib.ref_solve_llt(n, ip(rp), ip(col), dp(val), dp(f), dp(x), C.byref(fm), C.byref(sm))
    assert st == 0
    res = np.linalg.norm(f - A @ x) / np.linalg.norm(f)
    out = dict(n_free=n, nnz=A.nnz, f_norm=np.linalg.norm(f), a_sum=A.data.sum(),
               jq_sum=p.jq.sum(), xyz_sum=p.xyz.sum(), llt_residual=res,
               llt_factor_ms=fm.value, llt_solve_ms=sm.value, **summary(x))
    np.savez_compressed(os.path.join(HERE, f"{name}_llt.npz"), **out)
    print(name, n, res, fm.value, sm.value)


def main_cg(name, max_iterations):
    """The reference CGSolver (Jacobi, tol 1e-8, x0 = 0) on the full free-dof system of a
    workload, as <name>_cg.json: its status, iteration count and final relative residual
    (bench.py times a bounded prefix of these iterations on the GPU host)."""
    import json
    import time
    lib = ref_lib()
    p = W.make_workload(name)
    asm = assembly.assemble(p.xyz, p.tet2node, p.tet_material, p.nu, p.kappa, p.pec_pairs, p.jq)
    A, f = asm["A"], asm["f"]
    del asm
    n = A.shape[0]
    rp, col, val = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    del A
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    x = np.zeros(n)
    it, res, sm = C.c_int(), C.c_double(), C.c_double()
    t0 = time.perf_counter()
    st = lib.ref_solve_cg(n, ip(rp), ip(col), dp(val), dp(f), C.c_double(1e-8), int(max_iterations),
                          dp(x), C.byref(it), C.byref(res), C.byref(sm))
    out = {"workload": name, "solver": "reference CGSolver + Jacobi (oracle/_ref), tol 1e-8, x0 = 0",
           "status": int(st), "iterations": int(it.value), "residual": float(res.value),
           "solve_ms_container_1core": round(sm.value, 1),
           "wall_s": round(time.perf_counter() - t0, 1), "n_free": int(n),
           "note": ("full-run iteration count of the reference path" if st == 0 else
                    f"the reference path had not reached 1e-8 after {int(it.value)} iterations")}
    with open(os.path.join(HERE, f"{name}_cg.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(out)


def main():
    lib = ref_lib()
    cases = {
        "micro4": W.make_problem(4, 2, seed=3, name="micro"),
        "micro8c": W.make_problem(8, 2, materials="contrast", seed=5, name="micro_c"),
        "cube16": W.make_workload("cube16"),
    }
    for name, p in cases.items():
        out = run(lib, p)
        if out["n_free"] > 10000:
            del out["x_cg"]  # keep the large fixture small: the LLT solution pins it
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, out["n_free"], out["cg_iterations"], out["cg_residual"])


if __name__ == "__main__":
    if len(sys.argv) > 3 and sys.argv[1] == "cg":
        main_cg(sys.argv[2], int(sys.argv[3]))
    elif len(sys.argv) > 1:
        main_large(sys.argv[1])
    else:
        main()

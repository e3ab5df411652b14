"""Global solves -- FP64 restatement (oracle).

* ``pcg_jacobi``: CGSolver::solve_with_jacobi (src/numeric/iterative_solvers.cpp:
  424-503): x0 = 0, z = D^-1 r, stop when ||r||/||b|| < tol (default 1e-8,
  include/numeric/iterative_solvers.h:59-75), NUMERICAL_ERROR if |p.Ap| <
  1e-30, DIVERGED if after iteration 10 the residual grows 1000x in one step,
  MAX_ITER_REACHED otherwise; ||b|| < 1e-15 returns x = 0 at once.
* ``direct``: the sparse direct solve that UnifiedDirectSolver + Eigen LLT
  perform (src/numeric/unified_direct_solver.cpp:89-139), via SuperLU.
Status codes follow SolverStatus (include/numeric/em_linear_solver.h:26-32).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

SUCCESS, NUMERICAL_ERROR, DIVERGED, MAX_ITER_REACHED, INVALID_INPUT = range(5)


def pcg_jacobi(A, b, tol=1e-8, max_iter=1000):
    n = A.shape[0]
    x = np.zeros(n)
    bn = np.linalg.norm(b)
    if bn < 1e-15:
        return x, SUCCESS, 0, 0.0
    dinv = 1.0 / A.diagonal()
    r = b - A @ x
    z = dinv * r
    p = z.copy()
    rz = r @ z
    res = np.linalg.norm(r) / bn
    prev = res
    status = MAX_ITER_REACHED
    it = 0
    for it in range(1, max_iter + 1):
        Ap = A @ p
        pAp = p @ Ap
        if abs(pAp) < 1e-30:
            status = NUMERICAL_ERROR
            break
        alpha = rz / pAp
        x += alpha * p
        r -= alpha * Ap
        z = dinv * r
        rz_new = r @ z
        res = np.linalg.norm(r) / bn
        if res < tol:
            status = SUCCESS
            break
        if it > 10 and res > prev * 1000.0:
            status = DIVERGED
            break
        p = z + (rz_new / rz) * p
        rz = rz_new
        prev = res
    return x, status, min(it, max_iter), res


def direct(A, b):
    lu = spla.splu(A.tocsc())
    return lu.solve(np.asarray(b, dtype=np.float64))


def rel_residual(A, x, b):
    return float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))

// Thin C driver over the reference's own solvers, compiled from the reference
// sources where they lie (never copied): UnifiedDirectSolver + EigenLLTBackend
// (src/numeric/unified_direct_solver.cpp:51-139, include/numeric/
// eigen_solver_backends.h:338-395) and CGSolver with Jacobi
// (src/numeric/iterative_solvers.cpp:424-503).  Used only as the checker that
// pins the oracle's global solves, and as the CPU baseline.
#include <chrono>
#include <cstdint>
#include <memory>

#include "eigen_solver_backends.h"
#include "iterative_solvers.h"
#include "unified_direct_solver.h"

namespace {
template <typename M>
void fill(M& A, int n, const int32_t* rp, const int32_t* col, const double* val) {
  numeric::CooMatrix<double> coo(n, n);
  for (int i = 0; i < n; ++i)
    for (int32_t j = rp[i]; j < rp[i + 1]; ++j) coo.add_value(i, col[j], val[j]);
  A.build_from_coo(coo);
}
}  // namespace

extern "C" {

// Eigen LLT through UnifiedDirectSolver; returns SolverStatus, fills x and times.
int ref_solve_llt(int n, const int32_t* rp, const int32_t* col, const double* val,
                  const double* b, double* x, double* factor_ms, double* solve_ms) {
  numeric::CsrMatrix<double> A(n, n);
  fill(A, n, rp, col, val);
  numeric::UnifiedDirectSolver solver(std::make_unique<numeric::EigenLLTBackend>());
  auto t0 = std::chrono::steady_clock::now();
  solver.set_matrix(A);
  auto t1 = std::chrono::steady_clock::now();
  Eigen::VectorXd bb = Eigen::Map<const Eigen::VectorXd>(b, n);
  numeric::SolverResult r = solver.solve(bb);
  auto t2 = std::chrono::steady_clock::now();
  if (r.x.size() == n) Eigen::Map<Eigen::VectorXd>(x, n) = r.x;
  *factor_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  *solve_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
  return static_cast<int>(r.status);
}

// Jacobi-PCG through CGSolver; returns SolverStatus, fills x, iterations, residual.
int ref_solve_cg(int n, const int32_t* rp, const int32_t* col, const double* val,
                 const double* b, double tol, int max_iter, double* x, int* iterations,
                 double* residual, double* solve_ms) {
  numeric::CsrMatrix<double> A(n, n);
  fill(A, n, rp, col, val);
  numeric::CGConfig cfg;
  cfg.tolerance = tol;
  cfg.max_iterations = max_iter;
  cfg.preconditioner = numeric::CGConfig::PreconditionerType::JACOBI;
  numeric::CGSolver solver(cfg);
  solver.set_matrix(A);
  Eigen::VectorXd bb = Eigen::Map<const Eigen::VectorXd>(b, n);
  auto t0 = std::chrono::steady_clock::now();
  numeric::SolverResult r = solver.solve(bb);
  auto t1 = std::chrono::steady_clock::now();
  if (r.x.size() == n) Eigen::Map<Eigen::VectorXd>(x, n) = r.x;
  *iterations = r.iterations;
  *residual = r.residual_norm;
  *solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return static_cast<int>(r.status);
}
}

// Drives include/reference/numeric/fetidp_solver.h exactly as the reference tree would:
// the reference's own headers (EMMeshData, CsrMatrix, EMLinearSolverBase, read in place
// under /root/reference, never copied) around libemagddm_fetidp.so.  Test
// infrastructure, built into oracle/_ref/ by oracle/ref/Makefile.
//
//   binding_check setup  <problem.emfd>                      host only: prints n_free
//   binding_check solve  <problem.emfd> <system.bin> <x.bin> GPU: set_matrix + solves
// system.bin: int64 n, nnz; int32 row_ptr[n+1], col[nnz]; double val[nnz], f[n].
// Prints one JSON line with the statuses of every step.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "numeric/fetidp_solver.h"

namespace {

bool read_system(const char* path, numeric::CsrMatrix<double>& A, Eigen::VectorXd& f,
                 int64_t& n_out) {
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return false;
  int64_t hdr[2];
  bool ok = std::fread(hdr, sizeof(int64_t), 2, fp) == 2;
  const int64_t n = hdr[0], nnz = hdr[1];
  std::vector<int32_t> rp(static_cast<size_t>(n) + 1), col(static_cast<size_t>(nnz));
  std::vector<double> val(static_cast<size_t>(nnz));
  f.resize(n);
  ok = ok && std::fread(rp.data(), 4, rp.size(), fp) == rp.size();
  ok = ok && std::fread(col.data(), 4, col.size(), fp) == col.size();
  ok = ok && std::fread(val.data(), 8, val.size(), fp) == val.size();
  ok = ok && std::fread(f.data(), 8, size_t(n), fp) == size_t(n);
  std::fclose(fp);
  if (!ok) return false;
  numeric::CooMatrix<double> coo(static_cast<int>(n), static_cast<int>(n));
  for (int64_t i = 0; i < n; ++i)
    for (int32_t k = rp[size_t(i)]; k < rp[size_t(i) + 1]; ++k) coo.add_value(int(i), col[size_t(k)], val[size_t(k)]);
  A = numeric::CsrMatrix<double>(static_cast<int>(n), static_cast<int>(n));
  A.build_from_coo(coo);
  n_out = n;
  return true;
}

const char* name(numeric::SolverStatus s) {
  switch (s) {
    case numeric::SolverStatus::SUCCESS: return "SUCCESS";
    case numeric::SolverStatus::NUMERICAL_ERROR: return "NUMERICAL_ERROR";
    case numeric::SolverStatus::DIVERGED: return "DIVERGED";
    case numeric::SolverStatus::MAX_ITER_REACHED: return "MAX_ITER_REACHED";
    default: return "INVALID_INPUT";
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: binding_check setup|solve problem.emfd [system.bin x.bin]\n");
    return 2;
  }
  const std::string mode = argv[1];
  fe_em::FETIDPProblemReader reader;
  if (!reader.supports_format(argv[2])) {
    std::printf("{\"error\": \"unsupported file name\"}\n");
    return 1;
  }
  std::unique_ptr<fe_em::EMMeshData> mesh = reader.read(argv[2]);
  if (!mesh) {
    std::printf("{\"error\": \"read: %s\"}\n", reader.last_error().c_str());
    return 1;
  }
  // node ids of any base: shift every id by 1000 (the binding maps them back)
  for (auto& n : mesh->nodes) n.id += 1000;
  for (auto& e : mesh->elements)
    for (int& v : e.node_ids) v += 1000;
  for (auto& m : mesh->boundary_markers)
    for (auto& pr : std::get<std::vector<std::vector<int>>>(m.target_ids))
      for (int& v : pr) v += 1000;

  std::unique_ptr<numeric::EMLinearSolverBase> base;
  auto* solver = new numeric::FETIDPSolver();
  base.reset(solver);
  const numeric::SolverStatus st_setup = solver->setup(*mesh, reader.tet_subdomain(), reader.nu(),
                                                       reader.kappa());
  if (mode == "setup") {
    std::printf("{\"setup\": \"%s\", \"n_free\": %lld, \"name\": \"%s\", \"nodes\": %zu, "
                "\"elements\": %zu}\n",
                name(st_setup), (long long)solver->num_free_dofs(), base->get_solver_name().c_str(),
                mesh->nodes.size(), mesh->elements.size());
    return st_setup == numeric::SolverStatus::SUCCESS ? 0 : 1;
  }
  if (mode != "solve" || argc < 5) return 2;
  numeric::CsrMatrix<double> A;
  Eigen::VectorXd f;
  int64_t n = 0;
  if (!read_system(argv[3], A, f, n)) {
    std::printf("{\"error\": \"cannot read %s\"}\n", argv[3]);
    return 1;
  }
  // solve before set_matrix: INVALID_INPUT
  const numeric::SolverResult r0 = base->solve(f);
  base->set_matrix(A);
  const numeric::SolverResult r1 = base->solve(f);
  Eigen::VectorXd f2 = -2.5 * f;
  const numeric::SolverResult r2 = base->solve(f2);  // reuses the factor
  const double lin = (r2.x + 2.5 * r1.x).norm() / (2.5 * r1.x.norm());
  const Eigen::VectorXcd fc = f.cast<std::complex<double>>();
  const numeric::SolverResult rc = base->solve(fc);
  FILE* fx = std::fopen(argv[4], "wb");
  if (fx) {
    std::fwrite(r1.x.data(), 8, size_t(r1.x.size()), fx);
    std::fclose(fx);
  }
  const int factors = solver->factorizations();
  // a matrix of other values (one entry changed) is refused
  numeric::CsrMatrix<double> B = A;
  {
    numeric::CooMatrix<double> coo(static_cast<int>(n), static_cast<int>(n));
    const auto& rp = A.get_row_ptr();
    const auto& col = A.get_col_indices();
    const auto& val = A.get_values();
    for (int64_t i = 0; i < n; ++i)
      for (int k = rp[size_t(i)]; k < rp[size_t(i) + 1]; ++k)
        coo.add_value(int(i), col[size_t(k)], val[size_t(k)] * ((i == n / 2 && k == rp[size_t(i)]) ? 1.001 : 1.0));
    B = numeric::CsrMatrix<double>(static_cast<int>(n), static_cast<int>(n));
    B.build_from_coo(coo);
  }
  base->set_matrix(B);
  const numeric::SolverResult r3 = base->solve(f);
  numeric::CsrMatrix<double> C(static_cast<int>(n) - 1, static_cast<int>(n) - 1);  // wrong size
  base->set_matrix(C);
  const numeric::SolverResult r4 = base->solve(f);
  base->clear();
  const numeric::SolverResult r5 = base->solve(f);
  std::printf("{\"setup\": \"%s\", \"n_free\": %lld, \"before_set_matrix\": \"%s\", "
              "\"solve\": \"%s\", \"iterations\": %d, \"residual\": %.6e, \"solve2\": \"%s\", "
              "\"linearity\": %.3e, \"factorizations\": %d, \"complex\": \"%s\", "
              "\"other_values\": \"%s\", \"other_values_msg\": \"%s\", \"wrong_size\": \"%s\", "
              "\"after_clear\": \"%s\"}\n",
              name(st_setup), (long long)n, name(r0.status), name(r1.status), r1.iterations,
              r1.residual_norm, name(r2.status), lin, factors, name(rc.status), name(r3.status),
              r3.error_msg.c_str(), name(r4.status), name(r5.status));
  return 0;
}

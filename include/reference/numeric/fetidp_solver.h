// fetidp_solver.h -- reference-side C++ binding of the MI355X FETI-DP solver.
//
// A maintainer drops this header into the reference tree (next to
// include/numeric/em_linear_solver.h), puts this repository's include/ on the
// include path and links emagddm_amd/libemagddm_fetidp.so.  It provides
//   numeric::FETIDPSolver        an EMLinearSolverBase (include/numeric/
//                                em_linear_solver.h:82-235) over the C ABI of
//                                include/emagddm_fetidp.h, and
//   fe_em::FETIDPProblemReader   a fe_em::MeshReader (include/fe_em/mesh_reader.hpp:
//                                32-56, declared there without an implementation) for
//                                the problem files of include/emagddm_problem.h.
// Everything is inline; no exception leaves a member function that returns a status.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <variant>
#include <vector>

#include "em_linear_solver.h"
#include "em_mesh_data.hpp"
#include "mesh_reader.hpp"
#include "emagddm_fetidp.h"
#include "emagddm_problem.h"

namespace numeric {

// FETI-DP needs the unassembled subdomain matrices, so the mesh comes first (setup),
// then set_matrix() checks that the matrix is the assembled system of that mesh and
// factorizes once; every solve() after it reuses the factor (factor once, solve many,
// as the direct solvers behind this interface: include/numeric/em_linear_solver.h:96-108).
class FETIDPSolver : public EMLinearSolverBase {
 public:
  explicit FETIDPSolver(int32_t precision = FETIDP_FP64) {
    fetidp_config_default(&cfg_);
    cfg_.precision = precision;
    h_ = fetidp_create(&cfg_);
  }
  ~FETIDPSolver() override {
    if (h_) fetidp_destroy(h_);
  }
  FETIDPSolver(const FETIDPSolver&) = delete;
  FETIDPSolver& operator=(const FETIDPSolver&) = delete;

  // Symbolic setup from the reference mesh types: replaces GlobalEdgeIDGenerator::generate,
  // EMDOFManager::build and FETIDPContext::initialize.  Elements must be TET4 with
  // VECTOR_EDGE_ONLY dofs; element node_ids, Node::id and the PEC pairs of DIRICHLET
  // VECTOR_EDGE_ONLY markers use the same node ids, of any base (Node::id leaves the base
  // open, include/fe_em/em_mesh_data.hpp).  nu / kappa are indexed by material_id;
  // tet_subdomain[e] is the subdomain of element e.
  SolverStatus setup(const fe_em::EMMeshData& mesh, const std::vector<int>& tet_subdomain,
                     const std::vector<double>& nu, const std::vector<double>& kappa) {
    reset_state();
    if (!h_) return fail(SolverStatus::INVALID_INPUT, "fetidp_create failed");
    const size_t nn = mesh.nodes.size(), ne = mesh.elements.size();
    if (nn == 0 || ne == 0 || tet_subdomain.size() != ne || nu.empty() || nu.size() != kappa.size())
      return fail(SolverStatus::INVALID_INPUT, "setup: empty mesh or size mismatch");
    std::unordered_map<int, int32_t> index;
    index.reserve(nn);
    std::vector<double> xyz(3 * nn);
    for (size_t i = 0; i < nn; ++i) {
      const fe_em::Node& n = mesh.nodes[i];
      if (!index.emplace(n.id, int32_t(i)).second)
        return fail(SolverStatus::INVALID_INPUT, "setup: duplicate node id " + std::to_string(n.id));
      xyz[3 * i] = n.x;
      xyz[3 * i + 1] = n.y;
      xyz[3 * i + 2] = n.z;
    }
    auto node = [&](int id, int32_t& out) {
      auto it = index.find(id);
      if (it == index.end()) return false;
      out = it->second;
      return true;
    };
    std::vector<int32_t> t2n(4 * ne), mat(ne), sub(ne);
    for (size_t e = 0; e < ne; ++e) {
      const fe_em::Element& el = mesh.elements[e];
      if (el.elem_type != ElementType::TET4 || el.node_ids.size() != 4 ||
          el.dof_type != fe_em::DOFType::VECTOR_EDGE_ONLY)
        return fail(SolverStatus::INVALID_INPUT, "setup: element " + std::to_string(el.id) +
                                                     " is not a TET4 edge element");
      for (int k = 0; k < 4; ++k)
        if (!node(el.node_ids[size_t(k)], t2n[4 * e + size_t(k)]))
          return fail(SolverStatus::INVALID_INPUT, "setup: element " + std::to_string(el.id) +
                                                       " names an unknown node id");
      mat[e] = el.material_id;
      sub[e] = tet_subdomain[e];
    }
    std::vector<int32_t> pec;
    for (const fe_em::EMBoundaryMarker& m : mesh.boundary_markers) {
      if (m.bnd_type != fe_em::BndType::DIRICHLET || m.dof_type != fe_em::DOFType::VECTOR_EDGE_ONLY)
        continue;
      const auto* pairs = std::get_if<std::vector<std::vector<int>>>(&m.target_ids);
      if (!pairs) return fail(SolverStatus::INVALID_INPUT, "setup: edge marker without node pairs");
      for (const std::vector<int>& pr : *pairs) {
        int32_t a, b;
        if (pr.size() != 2 || !node(pr[0], a) || !node(pr[1], b))
          return fail(SolverStatus::INVALID_INPUT, "setup: bad PEC edge");
        pec.push_back(a);
        pec.push_back(b);
      }
    }
    const int32_t rc = fetidp_setup(h_, int64_t(nn), xyz.data(), int64_t(ne), t2n.data(), mat.data(),
                                    int32_t(nu.size()), nu.data(), kappa.data(), sub.data(),
                                    int64_t(pec.size() / 2), pec.data());
    if (rc != FETIDP_SUCCESS) return fail(SolverStatus(rc), fetidp_last_error(h_));
    int64_t sizes[10] = {0};
    fetidp_get_sizes(h_, sizes);
    n_free_ = sizes[1];
    setup_ = true;
    return SolverStatus::SUCCESS;
  }

  // Optional: the source J at the 4-point rule nodes of every tet (12 doubles per tet,
  // include/emagddm_fetidp.h fetidp_set_rhs_tet_J); solve(b) overrides it with b.
  SolverStatus set_source_J(const std::vector<double>& jq) {
    if (!setup_) return fail(SolverStatus::INVALID_INPUT, "set_source_J before setup");
    const int32_t rc = fetidp_set_rhs_tet_J(h_, jq.data());
    if (rc != FETIDP_SUCCESS) return fail(SolverStatus(rc), fetidp_last_error(h_));
    rhs_ = true;
    return SolverStatus::SUCCESS;
  }

  // A must be the assembled free-dof system of the mesh given to setup(), in EMDOFManager
  // order (duplicates allowed, as CsrMatrix keeps them: include/numeric/csr_matrix.hpp).
  // The factorization is built from the mesh; A is compared entry by entry with the
  // system the GPU assembles, and a matrix of another size, pattern or values is refused:
  // the next solve() returns INVALID_INPUT (as a failed set_matrix of the reference,
  // src/numeric/unified_direct_solver.cpp:97-102).
  void set_matrix(const CsrMatrix<double>& A) override {
    matrix_ok_ = false;
    if (!setup_) {
      error_ = "set_matrix before setup";
      return;
    }
    if (A.rows() != n_free_ || A.cols() != n_free_) {
      error_ = "set_matrix: size " + std::to_string(A.rows()) + "x" + std::to_string(A.cols()) +
               " is not the free-dof count " + std::to_string(n_free_) + " of the mesh";
      return;
    }
    if (!rhs_) {
      std::vector<double> zero(static_cast<size_t>(n_free_), 0.0);
      if (fetidp_set_rhs_free(h_, zero.data()) != FETIDP_SUCCESS) {
        error_ = fetidp_last_error(h_);
        return;
      }
      rhs_ = true;
    }
    if (fetidp_factor(h_) != FETIDP_SUCCESS) {
      error_ = fetidp_last_error(h_);
      factor_status_ = SolverStatus::NUMERICAL_ERROR;
      return;
    }
    ++n_factor_;
    const int64_t nnz = fetidp_get_global_nnz(h_);
    std::vector<int32_t> rp(static_cast<size_t>(n_free_) + 1), col(static_cast<size_t>(nnz));
    std::vector<double> val(static_cast<size_t>(nnz)), f(static_cast<size_t>(n_free_));
    fetidp_get_global_system(h_, rp.data(), col.data(), val.data(), f.data());
    double vmax = 0.0;
    for (double v : val) vmax = std::max(vmax, std::fabs(v));
    const std::vector<int>& arp = A.get_row_ptr();
    const std::vector<int>& acol = A.get_col_indices();
    const std::vector<double>& aval = A.get_values();
    std::vector<double> row(static_cast<size_t>(n_free_), 0.0);
    for (int64_t i = 0; i < n_free_; ++i) {
      for (int k = arp[size_t(i)]; k < arp[size_t(i) + 1]; ++k) row[size_t(acol[size_t(k)])] += aval[size_t(k)];
      bool ok = true;
      for (int32_t k = rp[size_t(i)]; k < rp[size_t(i) + 1]; ++k) {
        ok = ok && std::fabs(row[size_t(col[size_t(k)])] - val[size_t(k)]) <= 1e-12 * vmax;
        row[size_t(col[size_t(k)])] = 0.0;
      }
      for (int k = arp[size_t(i)]; k < arp[size_t(i) + 1]; ++k) {  // entries outside the pattern
        ok = ok && std::fabs(row[size_t(acol[size_t(k)])]) <= 1e-12 * vmax;
        row[size_t(acol[size_t(k)])] = 0.0;
      }
      if (!ok) {
        error_ = "set_matrix: row " + std::to_string(i) +
                 " differs from the curl-curl system of the mesh given to setup()";
        return;
      }
    }
    matrix_ok_ = true;
    error_.clear();
  }

  void set_matrix(const CsrMatrix<std::complex<double>>&) override {
    matrix_ok_ = false;
    error_ = "FETIDPSolver: complex (time-harmonic) systems are not supported: the solver "
             "takes real SPD curl-curl + kappa mass systems only";
  }

  SolverResult solve(const Eigen::VectorXd& b) override {
    SolverResult r;
    r.status = SolverStatus::INVALID_INPUT;
    if (!matrix_ok_) {
      r.error_msg = error_.empty() ? "solve before set_matrix" : error_;
      if (factor_status_ == SolverStatus::NUMERICAL_ERROR) r.status = factor_status_;
      return r;
    }
    if (b.size() != n_free_) {
      r.error_msg = "solve: right-hand side length does not match the matrix";
      return r;
    }
    int32_t rc = fetidp_set_rhs_free(h_, b.data());
    r.x = Eigen::VectorXd::Zero(b.size());
    fetidp_stats st{};
    if (rc == FETIDP_SUCCESS) rc = fetidp_solve(h_, nullptr, r.x.data(), &st);
    r.status = SolverStatus(rc);
    r.iterations = st.iterations;
    r.residual_norm = st.residual_norm;
    r.solve_time_ms = st.solve_ms;
    if (rc != FETIDP_SUCCESS) r.error_msg = fetidp_last_error(h_);
    last_ = st;
    return r;
  }

  SolverResult solve(const Eigen::VectorXcd&) override {
    SolverResult r;
    r.status = SolverStatus::INVALID_INPUT;
    r.error_msg = "FETIDPSolver: complex (time-harmonic) systems are not supported: the solver "
                  "takes real SPD curl-curl + kappa mass systems only";
    return r;
  }

  std::string get_solver_name() const override { return "FETIDP_MI355X"; }

  // Releases the device memory and the mesh; the configuration stays.
  void clear() override {
    if (h_) fetidp_destroy(h_);
    h_ = fetidp_create(&cfg_);
    reset_state();
  }

  int64_t num_free_dofs() const { return n_free_; }
  int factorizations() const { return n_factor_; }
  const fetidp_stats& last_stats() const { return last_; }
  fetidp_handle handle() const { return h_; }

 private:
  SolverStatus fail(SolverStatus s, const std::string& msg) {
    error_ = msg;
    return s;
  }
  void reset_state() {
    setup_ = rhs_ = matrix_ok_ = false;
    n_free_ = 0;
    n_factor_ = 0;
    factor_status_ = SolverStatus::SUCCESS;
    error_.clear();
  }

  fetidp_config cfg_{};
  fetidp_handle h_ = nullptr;
  bool setup_ = false, rhs_ = false, matrix_ok_ = false;
  int64_t n_free_ = 0;
  int n_factor_ = 0;
  SolverStatus factor_status_ = SolverStatus::SUCCESS;
  std::string error_;
  fetidp_stats last_{};
};

}  // namespace numeric

namespace fe_em {

// MeshReader for the problem files of include/emagddm_problem.h.  read() returns the
// mesh (Node::id = 0-based index, TET4 VECTOR_EDGE_ONLY elements with material ids, one
// DIRICHLET VECTOR_EDGE_ONLY marker holding the PEC edge node pairs); the partition, the
// material values and the source stay with the reader (tet_subdomain(), nu(), kappa(),
// jq()).  read() returns nullptr on a bad file; last_error() says why.
class FETIDPProblemReader : public MeshReader {
 public:
  FETIDPProblemReader() = default;

  std::unique_ptr<EMMeshData> read(const std::string& file_path) override {
    emag_problem p{};
    if (emag_problem_read(file_path.c_str(), &p) != 0) {
      error_ = emag_problem_last_error();
      return nullptr;
    }
    auto mesh = std::make_unique<EMMeshData>();
    mesh->nodes.resize(size_t(p.n_nodes));
    for (int64_t i = 0; i < p.n_nodes; ++i)
      mesh->nodes[size_t(i)] = Node{int(i), p.xyz[3 * i], p.xyz[3 * i + 1], p.xyz[3 * i + 2], 0};
    mesh->elements.resize(size_t(p.n_tets));
    for (int64_t t = 0; t < p.n_tets; ++t) {
      Element& e = mesh->elements[size_t(t)];
      e.id = int(t);
      e.node_ids.assign(p.tet2node + 4 * t, p.tet2node + 4 * t + 4);
      e.elem_type = ElemType::TET4;
      e.dof_type = DOFType::VECTOR_EDGE_ONLY;
      e.material_id = p.tet_material[t];
      e.region_id = p.tet_subdomain[t];
    }
    std::vector<std::vector<int>> pairs(static_cast<size_t>(p.n_pec));
    for (int64_t k = 0; k < p.n_pec; ++k)
      pairs[size_t(k)] = {p.pec_edge_nodes[2 * k], p.pec_edge_nodes[2 * k + 1]};
    EMBoundaryMarker pec;
    pec.id = 0;
    pec.bnd_type = BndType::DIRICHLET;
    pec.dof_type = DOFType::VECTOR_EDGE_ONLY;
    pec.target_ids = std::move(pairs);
    pec.value = 0.0;
    pec.name = "PEC";
    mesh->boundary_markers.push_back(std::move(pec));
    sub_.assign(p.tet_subdomain, p.tet_subdomain + p.n_tets);
    nu_.assign(p.nu, p.nu + p.n_mat);
    kappa_.assign(p.kappa, p.kappa + p.n_mat);
    if (p.jq) jq_.assign(p.jq, p.jq + 12 * p.n_tets);
    else jq_.clear();
    emag_problem_free(&p);
    error_.clear();
    return mesh;
  }

  bool supports_format(const std::string& file_path) const override {
    const std::string ext = ".emfd";
    return file_path.size() >= ext.size() &&
           file_path.compare(file_path.size() - ext.size(), ext.size(), ext) == 0;
  }

  const std::vector<int>& tet_subdomain() const { return sub_; }
  const std::vector<double>& nu() const { return nu_; }
  const std::vector<double>& kappa() const { return kappa_; }
  const std::vector<double>& jq() const { return jq_; }
  const std::string& last_error() const { return error_; }

 private:
  std::vector<int> sub_;
  std::vector<double> nu_, kappa_, jq_;
  std::string error_;
};

}  // namespace fe_em

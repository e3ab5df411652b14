// Host-only driver of the index arithmetic that builds every FETI-DP schedule
// (emagddm_amd/csrc/topology.cpp, plan.cpp) and of the problem-file reader
// (problem_io.cpp), compiled by tests/test_sanitize.py with AddressSanitizer and
// UndefinedBehaviorSanitizer.  Reads a problem file, builds the topology and every
// plan, prints their sizes as one JSON line.
#include <cstdio>

#include "emagddm_problem.h"
#include "plan.h"
#include "topology.h"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  emag_problem p{};
  if (emag_problem_read(argv[1], &p) != 0) {
    std::printf("{\"error\": \"%s\"}\n", emag_problem_last_error());
    return 1;
  }
  emag::MeshInput in{p.n_nodes, p.n_tets, p.tet2node, p.tet_subdomain, p.n_pec, p.pec_edge_nodes,
                     p.xyz};
  emag::Topology T;
  const std::string err = emag::build_topology(in, T);
  if (!err.empty()) {
    std::printf("{\"error\": \"%s\"}\n", err.c_str());
    emag_problem_free(&p);
    return 1;
  }
  emag::Plans P;
  emag::build_plans(T, P);
  std::printf("{\"n_edges\": %lld, \"n_free\": %lld, \"n_lambda\": %lld, \"n_primal\": %lld, "
              "\"n_sub\": %d, \"pool_elems\": %lld, \"csr_nnz\": %zu}\n",
              (long long)T.n_edges, (long long)T.n_free, (long long)T.n_lam, (long long)T.n_c,
              T.n_sub, (long long)P.pool_elems, P.csr.col.size());
  emag_problem_free(&p);
  return 0;
}

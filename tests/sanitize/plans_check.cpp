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
  // the grouped coarse plan first (groups of 2 and 3 tile columns), with its lists checked
  // against the bounds the sweep kernels assume, then the default plan
  long long cg_tiles = 0;
  for (int G : {2, 3}) {
    emag::build_plans(T, P, G);
    const int tn = int(P.coarse_ncp / emag::kTile);
    const int ng = int(P.cg_start.size()) - 1;
    bool ok = P.coarse_grouped && P.cg_start.front() == 0 && P.cg_start.back() == tn &&
              int(P.coarse_ftile.size()) == tn;
    for (int g = 0; g < ng && ok; ++g) {
      const int a = P.cg_start[g], b = P.cg_start[g + 1];
      ok = b - a >= 1 && b - a <= G;
      for (int64_t k = P.cg_fwd_off[g]; k < P.cg_fwd_off[g + 1] && ok; ++k) {
        const int I = P.cg_fwd_tiles[2 * k], J = P.cg_fwd_tiles[2 * k + 1];
        ok = I >= a && I < b && J >= P.coarse_ftile[I] && J < a && k >= P.cg_fwd_rowptr[I] &&
             k < P.cg_fwd_rowptr[I + 1];
      }
      for (int64_t k = P.cg_bwd_off[g]; k < P.cg_bwd_off[g + 1] && ok; ++k) {
        const int I = P.cg_bwd_tiles[2 * k], J = P.cg_bwd_tiles[2 * k + 1];
        ok = J >= a && J < b && I >= b && I < tn && P.coarse_ftile[I] <= J &&
             k >= P.cg_bwd_colptr[J] && k < P.cg_bwd_colptr[J + 1];
      }
      int64_t need = int64_t(b - a) * 64;
      ok = ok && P.cg_xoff[g] + need * need <= P.cg_x_elems;
    }
    for (int d = 0; d < G && ok; ++d)
      for (int64_t k = P.cg_inv_step[d]; k < P.cg_inv_step[d + 1] && ok; ++k) {
        const int i = P.cg_inv_items[2 * k], j = P.cg_inv_items[2 * k + 1];
        ok = i - j == d && P.cg_group_of[i] == P.cg_group_of[j];
      }
    if (!ok) {
      std::printf("{\"error\": \"grouped coarse plan out of bounds (G=%d)\"}\n", G);
      emag_problem_free(&p);
      return 1;
    }
    cg_tiles += (long long)(P.cg_fwd_tiles.size() + P.cg_bwd_tiles.size()) / 2;
  }
  emag::build_plans(T, P);
  // D-block tiles the Schur updates do not write into P themselves (copied by the
  // snapshot launches) and their total; whether the Schur updates write all of W
  long long d_tiles = 0;
  for (const emag::SubLayout& L : T.sub) d_tiles += (long long)(L.ndp / 64) * (L.ndp / 64 + 1) / 2;
  std::printf("{\"n_edges\": %lld, \"n_free\": %lld, \"n_lambda\": %lld, \"n_primal\": %lld, "
              "\"n_sub\": %d, \"pool_elems\": %lld, \"csr_nnz\": %zu, \"grouped_panel_tiles\": %lld, "
              "\"snap_tiles\": %zu, \"d_tiles\": %lld, \"w_fused\": %d}\n",
              (long long)T.n_edges, (long long)T.n_free, (long long)T.n_lam, (long long)T.n_c,
              T.n_sub, (long long)P.pool_elems, P.csr.col.size(), cg_tiles, P.snap_tiles.size(), d_tiles,
              int(P.w_fused));
  emag_problem_free(&p);
  return 0;
}

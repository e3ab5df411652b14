// FETI-DP solver orchestration behind the C ABI (include/emagddm_fetidp.h).
//
// Host C++ sequences the HIP kernels of kernels.hip; all numeric work runs on
// the GPU.  The reference has no FETI-DP implementation: the solver takes the
// place of FETIDPContext (src/fetidp/fetidp_context.cpp:19-35) and follows the
// EMLinearSolverBase contract (include/numeric/em_linear_solver.h:76-236):
// statuses instead of exceptions, factor once / solve many.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "emagddm_fetidp.h"
#include "emagddm_mpnn.h"
#include "kernels.h"
#include "mpnn.h"
#include "plan.h"
#include "topology.h"

using namespace emag;

namespace {

#define HIPCHECK(x)                                                                       \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      throw std::runtime_error(std::string(#x) + " failed: " + hipGetErrorString(e_));    \
  } while (0)

struct InputError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct NumericError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    n = count;
    if (count) HIPCHECK(hipMalloc(&p, count * sizeof(T)));
  }
  void upload(const std::vector<T>& v) {
    alloc(v.size());
    if (n) HIPCHECK(hipMemcpy(p, v.data(), n * sizeof(T), hipMemcpyHostToDevice));
  }
  void upload(const T* src, size_t count) {
    alloc(count);
    if (n) HIPCHECK(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
  }
  void zero(hipStream_t st) {
    if (n) HIPCHECK(hipMemsetAsync(p, 0, n * sizeof(T), st));
  }
};

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct fetidp_solver {
  fetidp_config cfg{};
  std::string err;
  Topology topo;
  Plans plans;
  bool setup_done = false, factored = false, rhs_from_J = false, rhs_set = false;
  hipStream_t st = nullptr;
  hipStream_t st2 = nullptr;  // side stream of the factorization (mode-3 launches)
  hipStream_t st3 = nullptr;  // operator extraction beside the coarse factorization
  hipEvent_t ev_main = nullptr, ev_side = nullptr, ev_extract = nullptr, ev_zero = nullptr,
             ev_early = nullptr, ev_face = nullptr;
  hipEvent_t ev_phi = nullptr, ev_coarse = nullptr;  // PCPG: coarse solve beside the Z product
  int32_t n_mat = 0;
  double symbolic_ms = 0.0, last_factor_ms = 0.0, upload_ms = 0.0;
  bool device_ready = false;
  std::vector<double> h_xyz_, h_nu_, h_kappa_;
  std::vector<int32_t> h_t2n_, h_mat_;

  DBuf<double> xyz, nu, kappa, jq;
  DBuf<int32_t> t2n, mat;
  DBuf<double> Ae, Fe;
  DBuf<int64_t> pool_target, pool_ptr;
  DBuf<int32_t> pool_src;
  DBuf<int64_t> rhs_target, rhs_ptr;
  DBuf<int32_t> rhs_src;
  DBuf<int32_t> csr_rp, csr_col, csr_src;
  DBuf<int64_t> csr_ptr;
  DBuf<double> csr_val;
  DBuf<int64_t> coarse_target, coarse_ptr, coarse_srcpos;
  DBuf<TileWork> local_items, local_items_ll, coarse_items;
  DBuf<int64_t> local_rowbase, coarse_rowbase, local_inv_base, coarse_inv_base;
  DBuf<int64_t> rowabs, coarse_rowabs;
  DBuf<int32_t> ftile, coarse_ftile;
  DBuf<DevSub> devsub;
  DBuf<SymTile> sym_tiles;
  DBuf<double> sympart;
  DBuf<SymTile> snap_tiles;  // S_DD snapshot tiles in snapshot-launch order
  // deluxe scaling (interface faces)
  bool deluxe = false;
  DBuf<FaceDev> face_dev;
  DBuf<int32_t> face_map_lam, face_map_self, d_self;
  DBuf<int64_t> face_slot0, face_slot1;
  DBuf<TileWork> face_items;
  DBuf<int64_t> face_rowbase, face_rowabs, face_inv_base;
  DBuf<double> fpool, finv, fops, fv1, fv2, xslot;
  // D_F of every face: FP32 (the default; M^-1 = sum R^T S R keeps its symmetric
  // positive semidefinite form whatever R's rounding) or FP64 (precond_precision 2)
  DBuf<float> fD32;
  DBuf<double> fD64;
  bool d64 = false;
  DBuf<int> ffail;
  DBuf<int32_t> vec_dof;
  DBuf<double> vec_w;
  DBuf<int64_t> local_vec_off;  // per subdomain: offset of its rows in vec / yvec
  // per pool tile: 1 if the factorization (FP64 schedule / FP32 schedule) ever writes it;
  // the others stay stale and read as zero in the extractions and the coarse gather
  DBuf<uint8_t> tw64, tw32, ts64, ts32;  // ts: written before the S_DD snapshot
  DBuf<int32_t> asm_tiles;  // the tiles that receive assembled entries (cleared per factor)
  const uint8_t* written() const { return fp32() ? tw32.p : tw64.p; }
  // vec / yvec hold the forward sweep of the current f_glob (done by the factorization
  // itself, or by a condensation of f_glob since): condensing f_glob again skips it
  bool fwd_valid = false;
  DBuf<int32_t> d_lam;
  DBuf<double> d_sgn, d_sgnD, d_half, rho;
  DBuf<int64_t> lam_pos, rho_pos;
  DBuf<int32_t> c_loc;
  DBuf<int64_t> c_ptr, c_ypos, c_vpos;
  DBuf<int8_t> free_kind;
  DBuf<int64_t> free_a;
  DBuf<double> pool64, inv64, W64;
  DBuf<float> pool32, inv32, W32;
  DBuf<float> Pf;   // Dirichlet preconditioner blocks in FP32 ...
  DBuf<double> Pd;  // ... or in FP64 when FP32 rounding could make them indefinite
  bool p64 = false;
  double h_min = 0.0;  // shortest mesh edge (conditioning estimate)
  // Largest overlap of neighbouring subdomains' bounding boxes (thinnest side of the
  // intersection box) in mean edge lengths: 0 for axis-aligned box partitions (about
  // 0.5 on a box partition of a mesh whose nodes moved by up to h/4), several for graph
  // or Voronoi partitions, whose staircase faces stiffness scaling handles badly.
  double box_overlap = 0.0;
  DBuf<int> fail;
  DBuf<double> cpool, cinv;
  DBuf<double> cpart;        // partial sums of the coarse S~^-1 products (tn * tn * 64)
  DBuf<int32_t> ctiles;      // (I, J <= I) tile pairs of the -S~^-1 block of the coarse pool
  int coarse_tn = 0;         // its tile order, coarse_ncp / 64
  // uc = S~^-1 x from the factored coarse pool, on stream cs (default: st)
  void coarse_apply(const double* x, double* y, hipStream_t cs = nullptr) {
    if (!cs) cs = st;
    if (!plans.coarse_grouped) {
      launch_coarse_symv(topo.n_c, coarse_tn, int64_t(ctiles.n / 2), ctiles.p, cpool.p, coarse_rowabs.p, x,
                         cpart.p, y, cs);
      return;
    }
    // S~^-1 x = L^-T L^-1 x: forward sweep over the groups, then backward; step A of a
    // group also writes the previous group's output blocks
    const std::vector<int32_t>& gs = plans.cg_start;
    const int ng = int(gs.size()) - 1;
    for (int g = 0; g < ng; ++g) {
      const int64_t o = plans.cg_fwd_off[g];
      const int pa = g > 0 ? gs[g - 1] : 0, pw = g > 0 ? gs[g] - gs[g - 1] : 0;
      launch_cg_step_a(false, pa, pw, cg_px.p, cg_yv.p, nullptr, 0, cg_fwd_tiles_d.p + 2 * o,
                       plans.cg_fwd_off[g + 1] - o, cpool.p, coarse_rowabs.p, cg_yv.p, cg_pf.p + o * kTile, cs);
      launch_cg_step_b(false, gs[g], gs[g + 1] - gs[g], cgX.p + plans.cg_xoff[g], x, topo.n_c,
                       cg_fwd_rowptr_d.p, cg_pf.p, cg_px.p, cs);
    }
    if (ng > 0)
      launch_cg_step_a(false, gs[ng - 1], gs[ng] - gs[ng - 1], cg_px.p, cg_yv.p, nullptr, 0, nullptr, 0,
                       cpool.p, coarse_rowabs.p, nullptr, nullptr, cs);
    for (int g = ng - 1; g >= 0; --g) {
      const int64_t o = plans.cg_bwd_off[g];
      const int pa = g + 1 < ng ? gs[g + 1] : 0, pw = g + 1 < ng ? gs[g + 2] - gs[g + 1] : 0;
      launch_cg_step_a(true, pa, pw, cg_px.p, cg_xv.p, y, topo.n_c, cg_bwd_tiles_d.p + 2 * o,
                       plans.cg_bwd_off[g + 1] - o, cpool.p, coarse_rowabs.p, cg_xv.p, cg_pb.p + o * kTile, cs);
      launch_cg_step_b(true, gs[g], gs[g + 1] - gs[g], cgX.p + plans.cg_xoff[g], cg_yv.p, plans.coarse_ncp,
                       cg_bwd_colptr_d.p, cg_pb.p, cg_px.p, cs);
    }
    if (ng > 0)
      launch_cg_step_a(true, gs[0], gs[1] - gs[0], cg_px.p, cg_xv.p, y, topo.n_c, nullptr, 0, cpool.p,
                       coarse_rowabs.p, nullptr, nullptr, cs);
  }
  DBuf<int> cfail;
  DBuf<double> f_glob, f_cur, vec, vout, U, ybuf, dbuf, fc, uc, gc, yvec, tbuf;
  DBuf<double> lam, r, z, p, q, dv, ubar, x, xtot, rglob, partial, pf, qf;
  std::vector<double> hist, ref_hist;
  double last_dual = 0.0, init_dual = 0.0;
  DBuf<double> dsc;                  // PCPG device scalars (kSc* slots)
  double* hsc = nullptr;             // pinned, mapped host copy written by the kernels
  double* hsc_dev = nullptr;
  hipGraphExec_t pcpg_exec = nullptr;
  const void* pcpg_key[3] = {nullptr, nullptr, nullptr};
  int pcpg_key_t = 0;

  // ---- optional event profiling (cfg.verbose & 2) ---------------------------
  // slot k accumulates milliseconds of one kernel class; see fetidp_get_profile.
  static constexpr int kProfSlots = 16;
  double prof_ms[kProfSlots] = {};
  double prof_cnt[kProfSlots] = {};
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_pending;
  std::vector<hipEvent_t> prof_pool;
  bool profiling() const { return (cfg.verbose & 2) != 0; }
  hipEvent_t prof_event() {
    if (!prof_pool.empty()) {
      hipEvent_t e = prof_pool.back();
      prof_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    return e;
  }
  template <typename Fn>
  void timed(int slot, Fn&& fn) {
    timed_on(slot, st, fn);
  }
  template <typename Fn>
  void timed_on(int slot, hipStream_t stream, Fn&& fn) {
    if (!profiling()) {
      fn();
      return;
    }
    hipEvent_t a = prof_event(), b = prof_event();
    HIPCHECK(hipEventRecord(a, stream));
    fn();
    HIPCHECK(hipEventRecord(b, stream));
    prof_pending.push_back({slot, {a, b}});
  }
  void prof_collect() {
    for (auto& pe : prof_pending) {
      HIPCHECK(hipEventSynchronize(pe.second.second));
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
      prof_ms[pe.first] += ms;
      prof_cnt[pe.first] += 1;
      prof_pool.push_back(pe.second.first);
      prof_pool.push_back(pe.second.second);
    }
    prof_pending.clear();
  }

  ~fetidp_solver() {
    for (auto& pe : prof_pending) {
      (void)hipEventDestroy(pe.second.first);
      (void)hipEventDestroy(pe.second.second);
    }
    for (hipEvent_t e : prof_pool) (void)hipEventDestroy(e);
    mp_graph_free(lam_graph);
    if (pcpg_exec) (void)hipGraphExecDestroy(pcpg_exec);
    for (hipEvent_t e : guess_ev)
      if (e) (void)hipEventDestroy(e);
    if (hsc) (void)hipHostFree(hsc);
    for (hipEvent_t e : {ev_main, ev_side, ev_extract, ev_zero, ev_early, ev_face, ev_phi, ev_coarse})
      if (e) (void)hipEventDestroy(e);
    if (st3) (void)hipStreamDestroy(st3);
    if (st2) (void)hipStreamDestroy(st2);
    if (st) (void)hipStreamDestroy(st);
  }

  // ---- learned initial multiplier (optional model, not owned) ---------------
  emag_mpnn* lam_model = nullptr;
  bool graph_built = false;
  std::vector<int32_t> g_rp, g_col;
  std::vector<float> g_coef, g_dir;  // coupling t_i . t_j; unit tangent per multiplier
  MpGraph lam_graph;
  DBuf<float> lam_X, lam_y, lam_dir;
  double last_guess_ms = 0.0;
  bool last_learned = false;

  // Interface graph: multipliers whose mesh edges share a node, coupled by the
  // cosine of their tangents (pairs with orthogonal tangents are left out).
  void build_dual_graph() {
    if (graph_built) return;
    const int64_t nl = topo.n_lam;
    std::vector<int32_t> edge_of_free(size_t(topo.n_free), -1);
    for (int64_t e = 0; e < topo.n_edges; ++e)
      if (topo.edge_free[e] >= 0) edge_of_free[topo.edge_free[e]] = int32_t(e);
    std::vector<int32_t> na(nl), nb(nl);
    g_dir.assign(size_t(3 * nl), 0.f);
    for (int64_t l = 0; l < nl; ++l) {
      const int32_t e = edge_of_free[topo.free_of_lam[l]];
      na[l] = topo.edge_nodes[2 * e];
      nb[l] = topo.edge_nodes[2 * e + 1];
      double t[3], n2 = 0.0;
      for (int c = 0; c < 3; ++c) {
        t[c] = h_xyz_[3 * size_t(nb[l]) + c] - h_xyz_[3 * size_t(na[l]) + c];
        n2 += t[c] * t[c];
      }
      const double inv = n2 > 0 ? 1.0 / std::sqrt(n2) : 0.0;
      for (int c = 0; c < 3; ++c) g_dir[3 * l + c] = float(t[c] * inv);
    }
    // node -> multipliers incidence (CSR)
    std::vector<int64_t> nptr(size_t(topo.n_nodes) + 1, 0);
    for (int64_t l = 0; l < nl; ++l) { ++nptr[na[l] + 1]; ++nptr[nb[l] + 1]; }
    for (int64_t v = 0; v < topo.n_nodes; ++v) nptr[v + 1] += nptr[v];
    std::vector<int32_t> nlam(static_cast<size_t>(nptr[topo.n_nodes]));
    std::vector<int64_t> fill(nptr.begin(), nptr.end() - 1);
    for (int64_t l = 0; l < nl; ++l) { nlam[fill[na[l]]++] = int32_t(l); nlam[fill[nb[l]]++] = int32_t(l); }
    g_rp.assign(size_t(nl) + 1, 0);
    g_col.clear();
    g_coef.clear();
    std::vector<int32_t> nbr;
    for (int64_t l = 0; l < nl; ++l) {
      nbr.clear();
      for (int32_t v : {na[l], nb[l]})
        for (int64_t k = nptr[v]; k < nptr[v + 1]; ++k)
          if (nlam[k] != l) nbr.push_back(nlam[k]);
      std::sort(nbr.begin(), nbr.end());
      nbr.erase(std::unique(nbr.begin(), nbr.end()), nbr.end());
      for (int32_t j : nbr) {
        const double c = double(g_dir[3 * l]) * g_dir[3 * j] + double(g_dir[3 * l + 1]) * g_dir[3 * j + 1] +
                         double(g_dir[3 * l + 2]) * g_dir[3 * j + 2];
        if (std::fabs(c) < 1e-6) continue;
        g_col.push_back(j);
        g_coef.push_back(float(c));
      }
      g_rp[l + 1] = int32_t(g_col.size());
    }
    graph_built = true;
  }

  void ensure_lam_device() {
    build_dual_graph();
    if (lam_graph.n != topo.n_lam) {
      mp_graph_upload(lam_graph, topo.n_lam, g_rp.data(), g_col.data(), g_coef.data());
      lam_dir.upload(g_dir);
      lam_X.alloc(size_t(topo.n_lam) * kMpF);
      lam_y.alloc(size_t(topo.n_lam));
    }
  }

  // network features from the condensed dual RHS dv, all on the stream:
  // z = M^-1 d, fsc = {1/rms(d), 1/rms(z), rms(z), rms(d)}, features into lam_X
  DBuf<double> fsc;
  hipEvent_t guess_ev[2] = {nullptr, nullptr};
  void lambda_features() {
    ensure_lam_device();
    const int64_t nl = topo.n_lam;
    fsc.alloc(4);
    apply_M(dv.p, z.p);
    const double* a[2] = {dv.p, z.p};
    launch_dots(nl, 2, a, a, partial.p, st);
    launch_feature_scales(partial.p, kDotBlocks, nl, fsc.p, st);
    launch_lambda_features(nl, dv.p, z.p, fsc.p, rho.p, lam_pos.p, lam_dir.p, lam_X.p, st);
  }

  // lam = s_z * network(features); timed with events, no host synchronization
  void learned_guess() {
    if (!guess_ev[0]) {
      HIPCHECK(hipEventCreate(&guess_ev[0]));
      HIPCHECK(hipEventCreate(&guess_ev[1]));
    }
    HIPCHECK(hipEventRecord(guess_ev[0], st));
    lambda_features();
    mpnn_forward(lam_model, lam_graph, 1, lam_X.p, lam_y.p, st);
    launch_lambda_from_y(topo.n_lam, lam_y.p, fsc.p, lam.p, st);
    HIPCHECK(hipEventRecord(guess_ev[1], st));
    last_learned = true;
  }

  // Coarse solve resolved from the configuration and n_c (0 = dense inverse, else the
  // grouped solve's tile columns per group).  Dense up to kDenseCoarseMax: there the
  // n_c^3 inverse costs little and one symmetric product per iteration is the fastest.
  static constexpr int64_t kDenseCoarseMax = 12288;
  int coarse_group_used = 0;
  int coarse_group_tiles() const {
    const bool grouped = topo.n_c > 0 && (cfg.coarse_solve == 2 ||
                                          (cfg.coarse_solve == 0 && topo.n_c > kDenseCoarseMax));
    return grouped ? (cfg.coarse_group > 0 ? cfg.coarse_group : 32) : 0;
  }
  // device data of the grouped coarse solve
  DBuf<int32_t> cg_start_d, cg_group_of_d, cg_fwd_tiles_d, cg_fwd_rowptr_d, cg_bwd_tiles_d,
      cg_bwd_colptr_d, cg_inv_items_d;
  DBuf<int64_t> cg_xoff_d;
  DBuf<double> cgX, cg_pf, cg_pb, cg_yv, cg_xv, cg_px;
  void upload_coarse_groups() {
    cg_start_d.upload(plans.cg_start);
    cg_group_of_d.upload(plans.cg_group_of);
    cg_fwd_tiles_d.upload(plans.cg_fwd_tiles);
    cg_fwd_rowptr_d.upload(plans.cg_fwd_rowptr);
    cg_bwd_tiles_d.upload(plans.cg_bwd_tiles);
    cg_bwd_colptr_d.upload(plans.cg_bwd_colptr);
    cg_inv_items_d.upload(plans.cg_inv_items);
    cg_xoff_d.upload(plans.cg_xoff);
    cgX.alloc(size_t(plans.cg_x_elems));
    cg_pf.alloc(size_t(std::max<int64_t>(plans.cg_fwd_off.back(), 1)) * kTile);
    cg_pb.alloc(size_t(std::max<int64_t>(plans.cg_bwd_off.back(), 1)) * kTile);
    cg_yv.alloc(size_t(plans.coarse_ncp));
    cg_xv.alloc(size_t(plans.coarse_ncp));
    cg_px.alloc(size_t(plans.coarse_group) * size_t(plans.coarse_group + 1) / 2 * kTile);
  }
  // Another group width for a grouped coarse solve already on the device: only the group
  // lists change (pool, schedule and factor do not depend on the width); the next factor
  // forms the group inverses.
  void set_coarse_group(int G) {
    if (!plans.coarse_grouped || G == plans.coarse_group) return;
    build_coarse_groups(plans, G);
    coarse_group_used = G;
    cfg.coarse_group = G;
    if (device_ready) upload_coarse_groups();
    factored = false;
    if (pcpg_exec) HIPCHECK(hipGraphExecDestroy(pcpg_exec));
    pcpg_exec = nullptr;
  }

  bool fell_back = false;  // FP32 requested but the factor fell back to FP64
  bool stalled = false;    // the last FP32 outer refinement stalled (solve hands over to FP64)
  std::string fallback_reason;
  bool fp32() const { return cfg.precision == FETIDP_FP32_REFINE && !fell_back; }

  // ---------------------------------------------------------------- setup
  void setup(int64_t n_nodes, const double* h_xyz, int64_t n_tets, const int32_t* h_t2n,
             const int32_t* h_mat, int32_t nmat, const double* h_nu, const double* h_kappa,
             const int32_t* h_sub, int64_t n_pec, const int32_t* h_pec) {
    if (n_nodes <= 0 || n_tets <= 0 || !h_xyz || !h_t2n || !h_mat || !h_nu || !h_kappa || !h_sub)
      throw InputError("setup: null or empty mesh input");
    if (n_pec < 0 || (n_pec > 0 && !h_pec)) throw InputError("setup: bad PEC input");
    if (nmat <= 0) throw InputError("setup: n_mat must be positive");
    for (int64_t t = 0; t < n_tets; ++t)
      if (h_mat[t] < 0 || h_mat[t] >= nmat) throw InputError("setup: material id out of range");
    // nu must be positive; kappa only finite: kappa < 0 (an indefinite system, e.g. a
    // -omega^2 eps term) is accepted here and reported as NUMERICAL_ERROR by the Cholesky
    // factorization, as the reference's LLT backend reports a non-SPD matrix
    // (include/numeric/eigen_solver_backends.h:159-170, src/numeric/unified_direct_solver.cpp:122-136)
    for (int32_t m = 0; m < nmat; ++m)
      if (!(h_nu[m] > 0.0) || !std::isfinite(h_nu[m]) || !std::isfinite(h_kappa[m]))
        throw InputError("setup: need finite nu > 0 and finite kappa");
    const double t0 = now_ms();
    setup_done = device_ready = factored = rhs_set = false;
    graph_built = false;
    mp_graph_free(lam_graph);
    if (pcpg_exec) HIPCHECK(hipGraphExecDestroy(pcpg_exec));  // sizes change with the plans
    pcpg_exec = nullptr;
    MeshInput in{n_nodes, n_tets, h_t2n, h_sub, n_pec, h_pec, h_xyz};
    std::string e = build_topology(in, topo);
    if (!e.empty()) throw InputError("setup: " + e);
    coarse_group_used = coarse_group_tiles();
    build_plans(topo, plans, coarse_group_used);
    n_mat = nmat;
    h_xyz_.assign(h_xyz, h_xyz + 3 * n_nodes);
    h_min = 1e300;
    for (int64_t e = 0; e < topo.n_edges; ++e) {
      const int64_t a = topo.edge_nodes[2 * e], b = topo.edge_nodes[2 * e + 1];
      double d2 = 0.0;
      for (int c = 0; c < 3; ++c) {
        const double d = h_xyz[3 * b + c] - h_xyz[3 * a + c];
        d2 += d * d;
      }
      h_min = std::min(h_min, std::sqrt(d2));
    }
    box_overlap = partition_box_overlap(h_xyz, h_t2n, h_sub);
    h_t2n_.assign(h_t2n, h_t2n + 4 * n_tets);
    h_mat_.assign(h_mat, h_mat + n_tets);
    h_nu_.assign(h_nu, h_nu + nmat);
    h_kappa_.assign(h_kappa, h_kappa + nmat);
    setup_done = true;
    symbolic_ms = now_ms() - t0;
  }

  // Device allocations and uploads, once per setup (first GPU-touching call).
  void ensure_device() {
    if (!setup_done) throw InputError("call fetidp_setup first");
    if (device_ready) return;
    const double t0 = now_ms();
    if (!st) HIPCHECK(hipStreamCreate(&st));
    if (!st2) {
      HIPCHECK(hipStreamCreate(&st2));
      HIPCHECK(hipStreamCreate(&st3));
      HIPCHECK(hipEventCreateWithFlags(&ev_main, hipEventDisableTiming));
      HIPCHECK(hipEventCreateWithFlags(&ev_side, hipEventDisableTiming));
      HIPCHECK(hipEventCreateWithFlags(&ev_extract, hipEventDisableTiming));
      HIPCHECK(hipEventCreateWithFlags(&ev_zero, hipEventDisableTiming));
      HIPCHECK(hipEventCreateWithFlags(&ev_early, hipEventDisableTiming));
      HIPCHECK(hipEventCreateWithFlags(&ev_face, hipEventDisableTiming));
      HIPCHECK(hipEventCreateWithFlags(&ev_phi, hipEventDisableTiming));
      HIPCHECK(hipEventCreateWithFlags(&ev_coarse, hipEventDisableTiming));
      HIPCHECK(hipEventRecord(ev_side, st2));  // a recorded, complete event to start from
    }
    xyz.upload(h_xyz_);
    t2n.upload(h_t2n_);
    mat.upload(h_mat_);
    nu.upload(h_nu_);
    kappa.upload(h_kappa_);
    const int64_t nt = topo.n_tets;
    Ae.alloc(size_t(21 * nt + 1));
    {
      const double one = 1.0;
      HIPCHECK(hipMemcpy(Ae.p + 21 * nt, &one, sizeof(double), hipMemcpyHostToDevice));
    }
    Fe.alloc(size_t(6 * nt));
    pool_target.upload(plans.pool.target);
    pool_ptr.upload(plans.pool.ptr);
    pool_src.upload(plans.pool.src);
    rhs_target.upload(plans.rhs_global.target);
    rhs_ptr.upload(plans.rhs_global.ptr);
    rhs_src.upload(plans.rhs_global.src);
    csr_rp.upload(plans.csr.row_ptr);
    csr_col.upload(plans.csr.col);
    csr_ptr.upload(plans.csr.ptr);
    csr_src.upload(plans.csr.src);
    csr_val.alloc(plans.csr.col.size());
    coarse_target.upload(plans.coarse.target);
    coarse_ptr.upload(plans.coarse.ptr);
    coarse_srcpos.upload(plans.coarse.srcpos);
    local_items.upload(plans.local_sched.items);
    local_items_ll.upload(plans.local_sched_ll.items);
    coarse_items.upload(plans.coarse_sched.items);
    {
      std::vector<int64_t> rb(plans.ftile_off.begin(), plans.ftile_off.end() - 1);
      local_rowbase.upload(rb);
      coarse_rowbase.upload(std::vector<int64_t>{0});
      rowabs.upload(plans.rowabs);
      coarse_rowabs.upload(plans.coarse_rowabs);
      local_inv_base.upload(plans.inv_off);
      tw64.upload(plans.tile_written);
      tw32.upload(plans.tile_written_ll);
      ts64.upload(plans.tile_written_snap);
      ts32.upload(plans.tile_written_snap_ll);
      asm_tiles.upload(plans.assembled_tiles);
      {
        std::vector<int64_t> vo(size_t(topo.n_sub));
        for (int64_t q = 0; q < topo.n_sub; ++q) vo[size_t(q)] = topo.sub[size_t(q)].vec_off;
        local_vec_off.upload(vo);
      }
      ftile.upload(plans.ftile);
      coarse_ftile.upload(plans.coarse_ftile);
      coarse_inv_base.upload(std::vector<int64_t>{0});
    }
    devsub.upload(plans.devsub);
    sym_tiles.upload(plans.sym_tiles);
    snap_tiles.upload(plans.snap_tiles);
    sympart.alloc(size_t(std::max<int64_t>(topo.sp_elems, 1)));
    {
      const FacePlan& F = plans.faces;
      face_dev.upload(F.dev);
      face_map_lam.upload(F.map_lam);
      face_map_self.upload(F.map_self);
      face_slot0.upload(F.slot0);
      face_slot1.upload(F.slot1);
      face_items.upload(F.sched.items);
      face_rowbase.upload(F.rowbase);
      face_rowabs.upload(F.rowabs);
      face_inv_base.upload(F.inv_off);
      d_self.upload(plans.d_self);
    }
    vec_dof.upload(plans.vec_dof);
    vec_w.upload(plans.vec_w);
    d_lam.upload(topo.d_lam);
    d_sgn.upload(topo.d_sgn);
    d_sgnD.alloc(topo.dvec_elems);
    d_half.alloc(topo.dvec_elems);
    rho.alloc(topo.dvec_elems);
    lam_pos.upload(topo.lam_pos);
    rho_pos.upload(plans.rho_pos);
    c_loc.upload(topo.c_loc);
    c_ptr.upload(topo.c_gather_ptr);
    c_ypos.upload(plans.c_ypos);
    c_vpos.upload(plans.c_vpos);
    free_kind.upload(topo.free_kind);
    free_a.upload(topo.free_src_a);
    fail.alloc(topo.n_sub);
    cfail.alloc(1);
    const int64_t nf = topo.n_free, nl = topo.n_lam, nc = std::max<int64_t>(topo.n_c, 1);
    f_glob.alloc(nf);
    f_cur.alloc(nf);
    vec.alloc(topo.vec_elems);
    vout.alloc(topo.vec_elems);
    yvec.alloc(topo.vec_elems);
    tbuf.alloc(topo.vec_elems);
    U.alloc(topo.vec_elems);
    ybuf.alloc(topo.y_elems);
    dbuf.alloc(topo.dvec_elems);
    fc.alloc(nc);
    uc.alloc(nc);
    gc.alloc(nc);
    for (DBuf<double>* b : {&lam, &r, &z, &p, &q, &dv, &ubar}) b->alloc(std::max<int64_t>(nl, 1));
    x.alloc(nf);
    xtot.alloc(nf);
    pf.alloc(nf);
    qf.alloc(nf);
    rglob.alloc(nf);
    partial.alloc(3 * kDotBlocks);
    dsc.alloc(kScCount);
    if (!hsc) {
      HIPCHECK(hipHostMalloc((void**)&hsc, sizeof(double) * kScCount, hipHostMallocMapped));
      HIPCHECK(hipHostGetDevicePointer((void**)&hsc_dev, hsc, 0));
    }
    // device and host sequence slots start equal (wait_scalars relies on it)
    dsc.zero(st);
    HIPCHECK(hipStreamSynchronize(st));
    std::memset(hsc, 0, sizeof(double) * kScCount);
    cpool.alloc(plans.coarse_pool_elems);
    cinv.alloc(plans.coarse_inv_elems);
    coarse_tn = int(plans.coarse_ncp / kTile);
    if (plans.coarse_grouped) {
      upload_coarse_groups();
      ctiles.release();
      cpart.release();
    } else {
      cgX.release();
      cg_pf.release();
      cg_pb.release();
    }
    if (!plans.coarse_grouped) {
      std::vector<int32_t> ct;
      ct.reserve(size_t(coarse_tn) * size_t(coarse_tn + 1));
      for (int I = 0; I < coarse_tn; ++I)
        for (int J = 0; J <= I; ++J) {
          ct.push_back(I);
          ct.push_back(J);
        }
      ctiles.upload(ct);
    }
    if (!plans.coarse_grouped)
      cpart.alloc(size_t(std::max(coarse_tn, 1)) * size_t(std::max(coarse_tn, 1)) * kTile);
    HIPCHECK(hipStreamSynchronize(st));
    device_ready = true;
    upload_ms = now_ms() - t0;
  }

  void set_materials(int32_t nmat, const double* h_nu, const double* h_kappa) {
    if (!setup_done) throw InputError("set_materials before setup");
    if (nmat != n_mat || !h_nu || !h_kappa) throw InputError("set_materials: n_mat mismatch");
    for (int32_t m = 0; m < nmat; ++m)
      if (!(h_nu[m] > 0.0) || !std::isfinite(h_nu[m]) || !std::isfinite(h_kappa[m]))
        throw InputError("set_materials: need finite nu > 0 and finite kappa");
    h_nu_.assign(h_nu, h_nu + nmat);
    h_kappa_.assign(h_kappa, h_kappa + nmat);
    if (device_ready) {
      nu.upload(h_nu_);
      kappa.upload(h_kappa_);
    }
    factored = false;
  }

  void set_rhs_J(const double* h_jq) {
    if (!setup_done) throw InputError("set_rhs before setup");
    if (!h_jq) throw InputError("set_rhs: null J");
    ensure_device();
    jq.upload(h_jq, size_t(12 * topo.n_tets));
    rhs_from_J = true;
    rhs_set = true;
    fwd_valid = false;
    if (factored) {
      // same matrix, new source: re-integrate (A_e is rewritten with identical
      // values) and re-assemble f, so no refactorization is needed
      launch_integrate(topo.n_tets, xyz.p, t2n.p, mat.p, nu.p, kappa.p, jq.p, 1, Ae.p, nullptr, Fe.p, st);
      launch_assemble_f64(plans.rhs_global.size(), rhs_target.p, rhs_ptr.p, rhs_src.p, Fe.p, f_glob.p, st);
      HIPCHECK(hipStreamSynchronize(st));
    }
  }

  void set_rhs_free(const double* h_f) {
    if (!setup_done) throw InputError("set_rhs before setup");
    if (!h_f) throw InputError("set_rhs: null f");
    ensure_device();
    HIPCHECK(hipMemcpy(f_glob.p, h_f, sizeof(double) * topo.n_free, hipMemcpyHostToDevice));
    rhs_from_J = false;
    rhs_set = true;
    fwd_valid = false;
  }

  // ---------------------------------------------------------------- factor
  // Launches run in order on st, except mode 3 (column item I = J + 1 plus the
  // factorization of the next diagonal tile), which runs on st2 beside the
  // mode-1 launch of the same column J:
  //   mode 3 of J waits for everything issued on st before it (column J-1,
  //   diagonal J); the mode-1 launch of J waits for mode 3 of J-1 only;
  //   every other launch, and the end of the schedule, wait for all of st2.
  // S_DD snapshot of the tiles snap_tiles[L.off, L.off + L.n) (a snapshot launch)
  template <typename T>
  void snapshot(const TileLaunch& L, T* pool, hipStream_t s) {
    if (L.n <= 0) return;
    const uint8_t* snap = sizeof(T) == 8 ? ts64.p : ts32.p;
    if (p64) launch_extract_P<T, double>(L.n, snap_tiles.p + L.off, devsub.p, pool, rowabs.p, ftile.p, snap, Pd.p, s);
    else launch_extract_P<T, float>(L.n, snap_tiles.p + L.off, devsub.p, pool, rowabs.p, ftile.p, snap, Pf.p, s);
  }

  // Dynamic LDS (unused) of the trailing early Schur launch, which runs beside the coarse
  // chain: two of its work-groups (33.3 KB + 24 KB each) per CU leave a chain work-group
  // (40 KB) room on every CU of the 160 KB, and fewer Schur waves on the chain's SIMDs
  static constexpr int kTailLdsPad = 24576;
  void* w_store = nullptr;  // W of the factorization in progress (written by the Schur updates)
  template <typename T>
  void run_schedule(const TileSchedule& S, const DBuf<TileWork>& items, const DBuf<int64_t>& rowbase,
                    const DBuf<int64_t>& rabs, const DBuf<int64_t>& inv_base, T* pool, T* inv,
                    int* failp, bool local, RhsArgs rhs = RhsArgs()) {
    const int slot0 = local ? 2 : 6;  // local: slots 2..4 (modes 0..2); coarse: 6..8
    bool after_side = false;      // the previous launch was a mode-3 launch on st2
    bool early = false;           // an early Schur launch (mode 4) is pending on st3
    bool tail = false;            // ... and only Schur launches (mode 0) follow it
    // the local Schur updates write the S_DD snapshot of their D-block tiles (flags 128)
    const SnapArgs snap = local ? SnapArgs{devsub.p, p64 ? (void*)Pd.p : (void*)Pf.p, p64 ? 0 : 1, w_store}
                                : SnapArgs();
    for (size_t li = 0; li < S.launches.size(); ++li) {
      const TileLaunch& L = S.launches[li];
      if (L.mode == 4) {
        tail = true;
        for (size_t k = li + 1; k < S.launches.size(); ++k) tail = tail && S.launches[k].mode == 0;
        // Schur updates of finished subdomains: beside the following column launches on
        // st3 (after everything queued on st so far); st meets them again before the
        // S_DD snapshot and at the end
        if (profiling()) {
          timed(slot0, [&] {
            launch_tile_mode<T>(0, items.p + L.off, L.n, rowbase.p, rabs.p, inv_base.p, pool, inv, failp, st,
                                cfg.tile_waves == 0 ? 0 : 4, rhs, 0, snap);
          });
          continue;
        }
        HIPCHECK(hipEventRecord(ev_main, st));
        HIPCHECK(hipStreamWaitEvent(st3, ev_main, 0));
        launch_tile_mode<T>(0, items.p + L.off, L.n, rowbase.p, rabs.p, inv_base.p, pool, inv, failp, st3,
                            cfg.tile_waves == 0 ? 0 : 4, rhs, tail ? kTailLdsPad : 0, snap);
        HIPCHECK(hipEventRecord(ev_early, st3));
        early = true;
        continue;
      }
      if (L.mode == -2) {  // S_DD snapshot of the early Schur's subdomains, beside the rest
        if (profiling()) {
          timed(5, [&] { snapshot<T>(L, pool, st); });
        } else {
          snapshot<T>(L, pool, st3);
          HIPCHECK(hipEventRecord(ev_early, st3));
        }
        continue;
      }
      if (L.mode < 0 && early) {
        HIPCHECK(hipStreamWaitEvent(st, ev_early, 0));
        early = false;
      }
      if (L.mode == 3 && cfg.side_stream == 2 && !profiling()) {
        // the factoring items (flags & 16) go at the front of the column-step launch
        // that follows them: dispatched first, no second stream
        int64_t n = L.n;
        if (li + 1 < S.launches.size() && S.launches[li + 1].mode == 1 &&
            S.launches[li + 1].off == L.off + L.n)
          n += S.launches[++li].n;
        timed(slot0 + 1, [&] {
          launch_tile_mode<T>(1, items.p + L.off, n, rowbase.p, rabs.p, inv_base.p, pool, inv, failp, st,
                              cfg.tile_waves == 0 ? 0 : 4, rhs);
        });
        continue;
      }
      if (L.mode == 3 && (profiling() || !cfg.side_stream)) {
        // profiled runs keep every launch on st, so event timings are per kernel
        timed(slot0 + 2, [&] {
          launch_tile_mode<T>(3, items.p + L.off, L.n, rowbase.p, rabs.p, inv_base.p, pool, inv, failp, st, 4, rhs);
        });
        continue;
      }
      if (L.mode == 3) {
        HIPCHECK(hipEventRecord(ev_main, st));
        HIPCHECK(hipStreamWaitEvent(st2, ev_main, 0));
        HIPCHECK(hipStreamWaitEvent(st, ev_side, 0));  // before ev_side moves on to this launch
        timed_on(slot0 + 2, st2, [&] {
          launch_tile_mode<T>(3, items.p + L.off, L.n, rowbase.p, rabs.p, inv_base.p, pool, inv, failp, st2, 4, rhs);
        });
        HIPCHECK(hipEventRecord(ev_side, st2));
        after_side = true;
        continue;
      }
      if (!after_side) HIPCHECK(hipStreamWaitEvent(st, ev_side, 0));
      after_side = false;
      if (L.mode < 0) {
        timed(5, [&] { snapshot<T>(L, pool, st); });
        continue;
      }
      timed(slot0 + L.mode, [&] {
        launch_tile_mode<T>(L.mode, items.p + L.off, L.n, rowbase.p, rabs.p, inv_base.p, pool, inv, failp, st,
                            cfg.tile_waves == 0 ? 0 : 4, rhs, 0, snap);
      });
    }
    HIPCHECK(hipStreamWaitEvent(st, ev_side, 0));
    // a trailing early launch (the D' rows of the dual Schur updates) is followed on st3
    // by the W extraction, which the end of the factorization waits for
    if (early && !tail) HIPCHECK(hipStreamWaitEvent(st, ev_early, 0));
  }

  // FP64 Dirichlet blocks when cond(S_DD) ~ max nu / (min kappa h^2) could exceed
  // what FP32 keeps positive definite (relative rounding 6e-8); cfg.precond_precision
  // 1 / 2 forces FP32 / FP64.
  // Deluxe scaling when the coefficient ratio kappa/nu varies by more than 1e4
  // between materials: diagonal stiffness weights then misjudge the gradient
  // (kappa-only) modes on the faces.  cfg.scaling 1 / 2 forces stiffness / deluxe.
  double partition_box_overlap(const double* h_xyz, const int32_t* h_t2n, const int32_t* h_sub) const {
    const int64_t ns = topo.n_sub;
    std::vector<double> lo(size_t(3 * ns), 1e300), hi(size_t(3 * ns), -1e300);
    double hsum = 0.0;
    for (int64_t t = 0; t < topo.n_tets; ++t) {
      const int64_t s = h_sub[t];
      for (int v = 0; v < 4; ++v)
        for (int c = 0; c < 3; ++c) {
          const double x = h_xyz[3 * int64_t(h_t2n[4 * t + v]) + c];
          lo[3 * s + c] = std::min(lo[3 * s + c], x);
          hi[3 * s + c] = std::max(hi[3 * s + c], x);
        }
    }
    for (int64_t e = 0; e < topo.n_edges; ++e) {
      const int64_t a = topo.edge_nodes[2 * e], b = topo.edge_nodes[2 * e + 1];
      double d2 = 0.0;
      for (int c = 0; c < 3; ++c) {
        const double d = h_xyz[3 * b + c] - h_xyz[3 * a + c];
        d2 += d * d;
      }
      hsum += std::sqrt(d2);
    }
    const double h_mean = hsum / double(std::max<int64_t>(topo.n_edges, 1));
    double worst = 0.0;
    for (const FaceDev& f : plans.faces.dev) {
      double thin = 1e300;
      for (int c = 0; c < 3; ++c)
        thin = std::min(thin, std::min(hi[3 * f.s0 + c], hi[3 * f.s1 + c]) -
                                  std::max(lo[3 * f.s0 + c], lo[3 * f.s1 + c]));
      worst = std::max(worst, thin);
    }
    return h_mean > 0.0 ? worst / h_mean : 0.0;
  }

  bool use_deluxe() const {
    if (cfg.scaling == 1) return false;
    if (cfg.scaling == 2) return true;
    if (box_overlap > 1.0) return true;  // not a box partition
    double lo = 1e300, hi = 0.0;
    for (int32_t m = 0; m < n_mat; ++m) {
      const double q = h_kappa_[m] / h_nu_[m];
      lo = std::min(lo, q);
      hi = std::max(hi, q);
    }
    if (!(lo > 0.0)) return hi > 0.0;
    return hi / lo > 1e4;
  }

  // Relative accuracy of an FP32 FETI-DP solve of this curl-curl system, estimated as
  // cond(K_rr) * 2^-24 with cond(K_rr) ~ nu_max / (kappa_min h_min^2): FP32 rounding of
  // the assembled entries, O(2^-24 nu / h^2), perturbs the discrete-gradient modes whose
  // energy is only kappa.  Measured outer contraction of the FP64 refinement per step
  // (uniform nu = 1, kappa = 1e-2): 0.19 at 16^3 (estimate 1.5e-3), 0.6 at 48^3 (1.4e-2).
  double fp32_estimate() const {
    double nu_max = 0.0, ka_min = 1e300;
    for (int32_t m = 0; m < n_mat; ++m) {
      nu_max = std::max(nu_max, h_nu_[m]);
      ka_min = std::min(ka_min, h_kappa_[m]);
    }
    if (!(ka_min > 0.0) || !(h_min > 0.0)) return 1e300;
    return nu_max / (ka_min * h_min * h_min) * 0x1p-24;
  }
  // Above this estimate FP32_REFINE factors in FP64 from the start (fell_back = 1).
  static constexpr double kFp32Limit = 5e-3;
  // FP32 outer refinement is abandoned for FP64 when two outer steps reduce the
  // residual by less than this factor (a contraction worse than 0.5 per step).
  static constexpr double kFp32Stall = 0.25;
  // FP64: a recovered residual this many times the dual residual is a floor (no tighter
  // dual tolerance); measured ratios: 474 on cube48j (not a floor), 5e3 on cube64c and
  // 2.6e4 on cube96j (floors, removed by one refinement step)
  static constexpr double kFloorRatio = 1e3;
  // Assumed amplification of a first refinement correction when the recovery takes the
  // deluxe average of the interface values; measured 1.4 (96^3 contrast cube) and 0.4
  // (64^3 contrast cube), against 470 with the plain mean of the two sides.  With the
  // correction's 4x margin a correction reaches the tolerance for any amplification up to
  // 12; a larger one costs a second correction, not accuracy.
  static constexpr double kCorrAmp = 3.0;

  bool precond_fp64(bool dl) const {
    if (dl) return true;  // the face blocks are formed and inverted in FP64
    if (cfg.precond_precision == 1) return false;
    if (cfg.precond_precision == 2) return true;
    double nu_max = 0.0, ka_min = 1e300;
    for (int32_t m = 0; m < n_mat; ++m) {
      nu_max = std::max(nu_max, h_nu_[m]);
      ka_min = std::min(ka_min, h_kappa_[m]);
    }
    if (!(ka_min > 0.0) || !(h_min > 0.0)) return true;
    return nu_max / (ka_min * h_min * h_min) > 1e6;
  }

  template <typename T>
  void factor_impl(DBuf<T>& pool, DBuf<T>& inv, DBuf<T>& W) {
    const int64_t nt = topo.n_tets;
    // subdomain pools
    pool.alloc(size_t(plans.pool_elems));
    inv.alloc(size_t(plans.inv_elems));
    W.alloc(size_t(topo.w_elems));
    if (p64) {
      Pf.release();
      Pd.alloc(size_t(topo.p_elems));
    } else {
      Pd.release();
      Pf.alloc(size_t(topo.p_elems));
    }
    // Only tiles with assembled entries are cleared: the schedule marks each item whose
    // tile (or diagonal tile) nothing has written yet, and the kernels read those as zero.
    // Outside profiled runs the clearing (st3) overlaps the element integration, the
    // global CSR values for the residual check follow on st3 beside the first column
    // launches, and the W extraction runs there beside the coarse factorization (which
    // leaves most of the GPU idle); st meets st3 again before the factor ends.
    const bool side = !profiling();
    hipStream_t sz = side ? st3 : st;
    if (side) {
      HIPCHECK(hipEventRecord(ev_main, st));  // st3 starts after what st has queued
      HIPCHECK(hipStreamWaitEvent(st3, ev_main, 0));
    }
    launch_zero_tiles<T>(int64_t(plans.assembled_tiles.size()), asm_tiles.p, pool.p, sz);
    if (side) HIPCHECK(hipEventRecord(ev_zero, st3));
    // element matrices A_e = K_e + M_e (and F_e when the RHS is given by J)
    timed(0, [&] {
      launch_integrate(nt, xyz.p, t2n.p, mat.p, nu.p, kappa.p, rhs_from_J ? jq.p : nullptr, 1, Ae.p,
                       nullptr, Fe.p, st);
    });
    // global CSR values (FP64) for the residual check
    if (!side)
      launch_assemble_f64(int64_t(plans.csr.col.size()), nullptr, csr_ptr.p, csr_src.p, Ae.p, csr_val.p, st);
    if (rhs_from_J)
      launch_assemble_f64(plans.rhs_global.size(), rhs_target.p, rhs_ptr.p, rhs_src.p, Fe.p, f_glob.p, st);
    if (side) HIPCHECK(hipStreamWaitEvent(st, ev_zero, 0));
    timed(1, [&] {
      if constexpr (sizeof(T) == 8)
        launch_assemble_f64(plans.pool.size(), pool_target.p, pool_ptr.p, pool_src.p, Ae.p, pool.p, st);
      else
        launch_assemble_f32(plans.pool.size(), pool_target.p, pool_ptr.p, pool_src.p, Ae.p, pool.p, st);
    });
    launch_scaling(topo.dvec_elems, topo.n_lam, rho_pos.p, pool.p, sizeof(T) == 4, rho.p, lam_pos.p,
                   d_sgn.p, d_sgnD.p, d_half.p, st);
    if (side) {
      // the global CSR values (needed first by the residual after the first PCPG solve) on
      // st3 beside the first column launches, which leave HBM bandwidth to spare
      HIPCHECK(hipEventRecord(ev_main, st));
      HIPCHECK(hipStreamWaitEvent(st3, ev_main, 0));
      launch_assemble_f64(int64_t(plans.csr.col.size()), nullptr, csr_ptr.p, csr_src.p, Ae.p, csr_val.p, st3);
    }
    fail.zero(st);
    // the forward sweep of f_glob rides along with the column steps (RhsArgs)
    launch_split_rhs(topo.vec_elems, vec_dof.p, vec_w.p, f_glob.p, vec.p, st);
    fwd_valid = false;
    const RhsArgs rhs{vec.p, yvec.p, local_vec_off.p};
    w_store = W.p;  // the D' Schur updates write W (SnapArgs)
    if constexpr (sizeof(T) == 8)
      run_schedule<T>(plans.local_sched, local_items, local_rowbase, rowabs, local_inv_base, pool.p, inv.p, fail.p, true, rhs);
    else
      run_schedule<T>(plans.local_sched_ll, local_items_ll, local_rowbase, rowabs, local_inv_base, pool.p, inv.p, fail.p, true, rhs);
    // W extraction (st3) and the coarse problem (st, st2) only read the local
    // factors, so they overlap; st waits for st3 before the end of the factor.
    // (the extraction only when the Schur updates do not cover all of W)
    if (profiling()) {
      if (!plans.w_fused)
        timed(5, [&] { launch_extract_W<T>(int64_t(plans.sym_tiles.size()), sym_tiles.p, topo.n_sub, plans.max_ncp, plans.max_ndp, devsub.p, pool.p, rowabs.p, ftile.p, (sizeof(T) == 8 ? tw64.p : tw32.p), W.p, st); });
    } else {
      HIPCHECK(hipEventRecord(ev_main, st));
      HIPCHECK(hipStreamWaitEvent(st3, ev_main, 0));
      if (!plans.w_fused)
        launch_extract_W<T>(int64_t(plans.sym_tiles.size()), sym_tiles.p, topo.n_sub, plans.max_ncp, plans.max_ndp, devsub.p, pool.p, rowabs.p, ftile.p, (sizeof(T) == 8 ? tw64.p : tw32.p), W.p, st3);
      HIPCHECK(hipEventRecord(ev_extract, st3));
      if (deluxe) {
        // the face factorizations only read the Dirichlet blocks: beside the coarse chain
        // (whose dependent single-matrix launches leave most of the GPU idle), on st2
        // when the coarse schedule leaves it free, else after the extraction on st3
        const hipStream_t fs = cfg.side_stream == 2 ? st2 : st3;
        if (fs == st2) HIPCHECK(hipStreamWaitEvent(st2, ev_main, 0));
        const hipStream_t main = st;
        st = fs;
        deluxe_setup();
        st = main;
        HIPCHECK(hipEventRecord(ev_face, fs));
      }
    }
    // coarse problem in FP64
    cpool.zero(st);
    launch_gather_pool<T>(plans.coarse.size(), coarse_target.p, coarse_ptr.p, coarse_srcpos.p,
                          pool.p, (sizeof(T) == 8 ? tw64.p : tw32.p), cpool.p, st);
    cfail.zero(st);
    run_schedule<double>(plans.coarse_sched, coarse_items, coarse_rowbase, coarse_rowabs, coarse_inv_base, cpool.p, cinv.p, cfail.p, false);
    if (plans.coarse_grouped)  // explicit inverses of the groups' diagonal blocks
      timed(8, [&] {
        launch_cg_invert(plans.coarse_group, plans.cg_inv_step.data(), cg_inv_items_d.p, cpool.p, coarse_rowabs.p, coarse_ftile.p, cinv.p, cg_start_d.p,
                         cg_group_of_d.p, cg_xoff_d.p, cgX.p, st);
      });
    if (deluxe && profiling()) deluxe_setup();
    if (!profiling()) {
      HIPCHECK(hipStreamWaitEvent(st, ev_extract, 0));
      if (deluxe) HIPCHECK(hipStreamWaitEvent(st, ev_face, 0));
    }
    HIPCHECK(hipStreamSynchronize(st));
    prof_collect();
    std::vector<int> hf(topo.n_sub);
    HIPCHECK(hipMemcpy(hf.data(), fail.p, sizeof(int) * topo.n_sub, hipMemcpyDeviceToHost));
    int hc = 0;
    HIPCHECK(hipMemcpy(&hc, cfail.p, sizeof(int), hipMemcpyDeviceToHost));
    for (int s = 0; s < topo.n_sub; ++s)
      if (hf[s]) throw NumericError("subdomain " + std::to_string(s) + ": K_rr not positive definite");
    if (hc) throw NumericError("coarse Schur complement not positive definite");
    fwd_valid = true;
    if (deluxe) {
      std::vector<int> ff(size_t(plans.faces.n));
      HIPCHECK(hipMemcpy(ff.data(), ffail.p, sizeof(int) * ff.size(), hipMemcpyDeviceToHost));
      for (size_t f = 0; f < ff.size(); ++f)
        if (ff[f]) throw NumericError("face " + std::to_string(f) + ": S_s + S_t not positive definite");
    }
  }

  double factor() {
    if (!setup_done) throw InputError("factor before setup");
    if (!rhs_set) throw InputError("factor needs a right-hand side (set_rhs_*) first");
    ensure_device();
    factored = false;
    const double t0 = now_ms();
    fell_back = false;
    deluxe = use_deluxe() && plans.faces.n > 0;
    p64 = precond_fp64(deluxe);
    d64 = cfg.precond_precision == 2;
    if (fp32() && fp32_estimate() > kFp32Limit) {
      // the FP32 solve could not reach the FP64 tolerance in a useful number of outer
      // steps: factor in FP64 from the start
      fell_back = true;
      char buf[160];
      std::snprintf(buf, sizeof buf, "FP32 accuracy estimate nu/(kappa h^2) 2^-24 = %.3g exceeds %.3g",
                    fp32_estimate(), kFp32Limit);
      fallback_reason = buf;
      if (cfg.verbose & 1) std::fprintf(stderr, "[fetidp] FP32 factor not used: %s\n", buf);
    }
    if (fp32()) {
      pool64.release(); inv64.release(); W64.release();
      try {
        factor_impl<float>(pool32, inv32, W32);
      } catch (const NumericError& e) {
        // FP32 factor lost positive definiteness: refactor in FP64 (solve follows)
        fell_back = true;
        fallback_reason = e.what();
        if (cfg.verbose & 1) std::fprintf(stderr, "[fetidp] FP32 factor fell back to FP64: %s\n", e.what());
        pool32.release(); inv32.release(); W32.release();
        factor_impl<double>(pool64, inv64, W64);
      }
    } else {
      pool32.release(); inv32.release(); W32.release();
      factor_impl<double>(pool64, inv64, W64);
    }
    factored = true;
    return now_ms() - t0;
  }

  // ---------------------------------------------------------------- operators
  double dots(int npairs, const double* const* a, const double* const* b, int64_t n, double* out) {
    launch_dots(n, npairs, a, b, partial.p, st);
    double h[3 * kDotBlocks];
    HIPCHECK(hipMemcpyAsync(h, partial.p, sizeof(double) * npairs * kDotBlocks, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    for (int k = 0; k < npairs; ++k) {
      double s = 0.0;
      for (int i = 0; i < kDotBlocks; ++i) s += h[k * kDotBlocks + i];
      out[k] = s;
    }
    return out[0];
  }

  template <typename T>
  void apply_F(const DBuf<T>& W, const double* pin, double* qout) {
    timed(9, [&] { apply_W<T>(W, pin); });
    timed(10, [&] { apply_F_rest<T>(W, qout); });
  }
  template <typename T>
  void apply_F_rest(const DBuf<T>& W, double* qout) {
    launch_coarse_gather(topo.n_c, c_ptr.p, c_ypos.p, ybuf.p, nullptr, gc.p, st);
    coarse_apply(gc.p, uc.p);
    launch_dual_correct<T>(topo.n_sub, plans.max_ndp, plans.max_ncp, devsub.p, W.p, sympart.p, c_loc.p, uc.p, vec.p, 1.0, 1.0, 0.0, dbuf.p, st);
    launch_lambda_scatter(topo.n_lam, lam_pos.p, d_sgn.p, dbuf.p, qout, st);
  }
  // Per face: [[A_F, I], [I, 0]] from the FP64 Dirichlet blocks, eliminate A_F with the
  // batched tile kernels, keep A_F^-1 and B_F as lower tiles and form D_F = A_F^-1 B_F.
  void deluxe_setup() {
    const FacePlan& F = plans.faces;
    fpool.alloc(size_t(F.pool_elems));
    finv.alloc(size_t(std::max<int64_t>(F.inv_elems, 1)));
    fops.alloc(size_t(F.op_elems));
    fv1.alloc(size_t(F.vec_elems));
    fv2.alloc(size_t(F.vec_elems));
    ffail.alloc(size_t(F.n));
    if (xslot.n != size_t(topo.dvec_elems)) {
      xslot.alloc(size_t(topo.dvec_elems));
      xslot.zero(st);  // padding slots stay zero; real slots are rewritten every application
    }
    launch_face_setup(F.n, F.max_T, face_dev.p, devsub.p, Pd.p, face_slot0.p, face_slot1.p, fpool.p, st);
    ffail.zero(st);
    run_schedule<double>(F.sched, face_items, face_rowbase, face_rowabs, face_inv_base, fpool.p, finv.p,
                         ffail.p, false);
    launch_face_extract(F.n, F.max_T, face_dev.p, devsub.p, Pd.p, face_slot1.p, fpool.p, fops.p, st);
    if (d64) {
      fD32.release();
      fD64.alloc(size_t(std::max<int64_t>(F.d_elems, 1)));
      launch_face_D<double>(F.n, F.max_T, face_dev.p, fops.p, fD64.p, st);
    } else {
      fD64.release();
      fD32.alloc(size_t(std::max<int64_t>(F.d_elems, 1)));
      launch_face_D<float>(F.n, F.max_T, face_dev.p, fops.p, fD32.p, st);
    }
  }

  // z = M^-1 r with deluxe weights: x_s0 = D_F r, x_s1 = D_F r - r (D_F = A_F^-1 B_F),
  // y = S_DD x per subdomain, z = D_F^T (y_s0 + y_s1) - y_s1 on every face.
  void apply_M_deluxe(const double* rin, double* zout) {
    const FacePlan& F = plans.faces;
    auto dmv = [&](bool trans, const int32_t* map, const double* src, double* out) {
      if (d64) launch_face_dmv<double>(trans, F.n, F.max_T, face_dev.p, fD64.p, map, src, out, st);
      else launch_face_dmv<float>(trans, F.n, F.max_T, face_dev.p, fD32.p, map, src, out, st);
    };
    dmv(false, face_map_lam.p, rin, fv2.p);                                         // D r
    launch_deluxe_split(F.vec_elems, face_map_lam.p, face_slot0.p, face_slot1.p, fv2.p, rin, xslot.p, st);
    launch_symv<double>(int64_t(plans.sym_tiles.size()), sym_tiles.p, topo.n_sub, plans.max_Td, devsub.p, Pd.p, 1,
                        d_self.p, nullptr, xslot.p, sympart.p, dbuf.p, st);       // y = S x
    launch_deluxe_sum(F.vec_elems, face_map_lam.p, face_slot0.p, face_slot1.p, dbuf.p, fv1.p, st);
    dmv(true, face_map_self.p, fv1.p, fv2.p);                                       // D^T u
    launch_deluxe_out(F.vec_elems, face_map_lam.p, face_slot1.p, fv2.p, dbuf.p, zout, st);
  }

  // W x with x = B^T p gathered per subdomain: Z x as per-tile partials in sympart (summed
  // by launch_dual_correct), Phi^T x into ybuf
  template <typename T>
  void apply_W(const DBuf<T>& W, const double* pin) {
    launch_apply_W<T>(int64_t(plans.sym_tiles.size()), sym_tiles.p, topo.n_sub, plans.max_ncp, plans.max_ndp,
                      devsub.p, W.p, d_lam.p, d_sgn.p, pin, sympart.p, ybuf.p, st);
  }
  // The same W x as two launches, Phi^T x first (ev_phi recorded after it), so that the
  // coarse solve it feeds can run on st2 beside the Z product (PCPG iterations)
  template <typename T>
  void apply_W_split(const DBuf<T>& W, const double* pin) {
    launch_gemv_phi<T>(topo.n_sub, plans.max_ncp, devsub.p, W.p, d_lam.p, d_sgn.p, pin, ybuf.p, st, plans.max_ndp);
    HIPCHECK(hipEventRecord(ev_phi, st));
    launch_symv<T>(int64_t(plans.sym_tiles.size()), sym_tiles.p, topo.n_sub, plans.max_Td, devsub.p, W.p, 0,
                   d_lam.p, d_sgn.p, pin, sympart.p, nullptr, st);
  }
  void apply_M(const double* rin, double* zout) {
    if (deluxe) {
      timed(11, [&] { apply_M_deluxe(rin, zout); });
      return;
    }
    timed(11, [&] {
      if (p64)
        launch_symv<double>(int64_t(plans.sym_tiles.size()), sym_tiles.p, topo.n_sub, plans.max_Td, devsub.p, Pd.p, 1,
                            d_lam.p, d_sgnD.p, rin, sympart.p, dbuf.p, st);
      else
        launch_symv<float>(int64_t(plans.sym_tiles.size()), sym_tiles.p, topo.n_sub, plans.max_Td, devsub.p, Pf.p, 1,
                           d_lam.p, d_sgnD.p, rin, sympart.p, dbuf.p, st);
    });
    launch_lambda_scatter(topo.n_lam, lam_pos.p, d_sgnD.p, dbuf.p, zout, st);
  }

  // condense the global free-dof RHS `fsrc` into the dual RHS dv (and fc, a_s);
  // is_f: fsrc holds f_glob, whose forward sweep may already be in vec / yvec
  template <typename T>
  // interface_only: fsrc vanishes on the interior dofs (to the caller's tolerance), which
  // are then taken as zero: the forward sweep skips every interior column
  void condense(const DBuf<T>& pool, const DBuf<T>& inv, const DBuf<T>& W, const double* fsrc,
                bool is_f, bool interface_only = false) {
    if (!(is_f && fwd_valid)) {
      launch_split_rhs(topo.vec_elems, vec_dof.p, vec_w.p, fsrc, vec.p, st);
      if (interface_only) {
        launch_zero_interior(topo.n_sub, plans.max_Ti, devsub.p, vec.p, st);
        launch_zero_interior(topo.n_sub, plans.max_Ti, devsub.p, yvec.p, st);
      }
      timed(12, [&] {
        launch_forward<T>(topo.n_sub, plans.max_T, plans.max_Tr, devsub.p, pool.p, rowabs.p, inv.p,
                          local_inv_base.p, ftile.p, vec.p, yvec.p, st, interface_only ? 1 : 0,
                          plans.max_dual, plans.max_after);
      });
      fwd_valid = is_f;
    }
    launch_coarse_gather(topo.n_c, c_ptr.p, c_vpos.p, vec.p, nullptr, fc.p, st);
    coarse_apply(fc.p, uc.p);
    // vec_D' holds -a_s: out = a_s - Phi u
    launch_dual_correct<T>(topo.n_sub, plans.max_ndp, plans.max_ncp, devsub.p, W.p, sympart.p, c_loc.p, uc.p, vec.p, 0.0, -1.0, -1.0, dbuf.p, st);
    launch_lambda_scatter(topo.n_lam, lam_pos.p, d_sgn.p, dbuf.p, dv.p, st);
  }

  // recover the global free-dof solution from lambda into xout
  template <typename T>
  void recover(const DBuf<T>& pool, const DBuf<T>& inv, const DBuf<T>& W, double* xout) {
    if (profiling()) {
      apply_W<T>(W, lam.p);
      launch_coarse_gather(topo.n_c, c_ptr.p, c_ypos.p, ybuf.p, fc.p, gc.p, st);
      coarse_apply(gc.p, uc.p);
    } else {  // the coarse solve beside the Z product, as in the PCPG iterations
      apply_W_split<T>(W, lam.p);
      HIPCHECK(hipStreamWaitEvent(st2, ev_phi, 0));
      launch_coarse_gather(topo.n_c, c_ptr.p, c_ypos.p, ybuf.p, fc.p, gc.p, st2);
      coarse_apply(gc.p, uc.p, st2);
      HIPCHECK(hipEventRecord(ev_coarse, st2));
      HIPCHECK(hipStreamWaitEvent(st, ev_coarse, 0));
    }
    launch_dual_correct<T>(topo.n_sub, plans.max_ndp, plans.max_ncp, devsub.p, W.p, sympart.p, c_loc.p, uc.p, vec.p, -1.0, -1.0, -1.0, dbuf.p, st);
    if (deluxe) {
      // deluxe average ubar = (S0 + S1)^-1 (S0 u0 + S1 u1) = u0 + D_F (u1 - u0) per face:
      // each side weighted by its own Dirichlet block, so a stiff side's value is not
      // spoiled by the soft side's (the interface residual then carries no contrast factor)
      const FacePlan& F = plans.faces;
      launch_deluxe_jump(F.vec_elems, face_map_lam.p, face_slot0.p, face_slot1.p, dbuf.p, fv1.p, st);
      if (d64) launch_face_dmv<double>(false, F.n, F.max_T, face_dev.p, fD64.p, face_map_self.p, fv1.p, fv2.p, st);
      else launch_face_dmv<float>(false, F.n, F.max_T, face_dev.p, fD32.p, face_map_self.p, fv1.p, fv2.p, st);
      launch_deluxe_avg(F.vec_elems, face_map_lam.p, face_slot0.p, dbuf.p, fv2.p, ubar.p, st);
    } else {
      launch_lambda_scatter(topo.n_lam, lam_pos.p, d_half.p, dbuf.p, ubar.p, st);
    }
    launch_fill_U(topo.n_sub, std::max(plans.max_ndp, plans.max_ncp), devsub.p, d_lam.p, ubar.p, c_loc.p, uc.p, U.p, st);
    timed(13, [&] {
      launch_interior<T>(topo.n_sub, plans.max_Ti, plans.bwd_width.data(), devsub.p, pool.p, rowabs.p, inv.p,
                         local_inv_base.p, ftile.p, yvec.p, U.p, tbuf.p, vout.p, st);
    });
    launch_gather_global(topo.n_free, free_kind.p, free_a.p, vout.p, ubar.p, uc.p, xout, st);
  }

  // ||rglob restricted to the interior dofs|| / ||f_glob|| (after global_residual)
  double interior_residual() {
    launch_kind_sq(topo.n_free, free_kind.p, 0, rglob.p, partial.p, st);
    double h[kDotBlocks];
    HIPCHECK(hipMemcpyAsync(h, partial.p, sizeof(double) * kDotBlocks, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    double ri = 0.0;
    for (int i = 0; i < kDotBlocks; ++i) ri += h[i];
    const double* af[1] = {f_glob.p};
    double o[1];
    const double fn = dots(1, af, af, topo.n_free, o);
    return fn > 0.0 ? std::sqrt(ri / fn) : 0.0;
  }
  // ||fref - K xin|| / ||fref||, residual vector left in rglob
  double global_residual(const double* xin, const double* fref) {
    launch_csr_residual(topo.n_free, csr_rp.p, csr_col.p, csr_val.p, xin, fref, rglob.p, st);
    const double* a[2] = {rglob.p, fref};
    double o[2];
    dots(2, a, a, topo.n_free, o);
    return o[1] > 0 ? std::sqrt(o[0] / o[1]) : 0.0;
  }

  // One PCPG iteration as a fixed kernel chain: the scalars alpha and beta stay
  // on the device, so the chain can be captured.  The lambda scatters carry the dot
  // products; alpha and beta are formed by the launches that use them.  Bitwise the
  // same arithmetic as apply_F / apply_M, k_dots, k_pcpg_alpha / beta and the updates.
  // head_queued: W p of this iteration is already on the stream (queued by the
  // previous iteration, see pcpg).
  template <typename T>
  void pcpg_iteration_kernels(const DBuf<T>& W, bool head_queued = false) {
    const int64_t nl = topo.n_lam;
    // outside event profiling the coarse solve (a chain of small launches) runs on st2
    // beside the Z product; profiled runs keep every launch on st
    const bool split = !profiling();
    if (!head_queued) {
      if (split) apply_W_split<T>(W, p.p);
      else timed(9, [&] { apply_W<T>(W, p.p); });
    }
    if (split) {
      HIPCHECK(hipStreamWaitEvent(st2, ev_phi, 0));
      launch_coarse_gather(topo.n_c, c_ptr.p, c_ypos.p, ybuf.p, nullptr, gc.p, st2);
      coarse_apply(gc.p, uc.p, st2);
      HIPCHECK(hipEventRecord(ev_coarse, st2));
      HIPCHECK(hipStreamWaitEvent(st, ev_coarse, 0));
    }
    timed(10, [&] {
      if (!split) {
        launch_coarse_gather(topo.n_c, c_ptr.p, c_ypos.p, ybuf.p, nullptr, gc.p, st);
        coarse_apply(gc.p, uc.p);
      }
      launch_dual_correct<T>(topo.n_sub, plans.max_ndp, plans.max_ncp, devsub.p, W.p, sympart.p, c_loc.p, uc.p, vec.p, 1.0, 1.0, 0.0,
                             dbuf.p, st);
      const double* a1[1] = {p.p};
      const double* b1[1] = {nullptr};
      launch_scatter_dot(nl, lam_pos.p, d_sgn.p, dbuf.p, q.p, 1, a1, b1, partial.p, st);
      launch_alpha_axpy2(nl, partial.p, dsc.p, hsc_dev, p.p, q.p, lam.p, r.p, st);
    });
    const double* a2[2] = {r.p, r.p};
    const double* b2[2] = {r.p, nullptr};
    if (deluxe) {
      apply_M(r.p, z.p);
      const double* b3[2] = {r.p, z.p};
      launch_dots(nl, 2, a2, b3, partial.p, st);
    } else {
      timed(11, [&] {
        if (p64)
          launch_symv<double>(int64_t(plans.sym_tiles.size()), sym_tiles.p, topo.n_sub, plans.max_Td, devsub.p,
                              Pd.p, 1, d_lam.p, d_sgnD.p, r.p, sympart.p, dbuf.p, st);
        else
          launch_symv<float>(int64_t(plans.sym_tiles.size()), sym_tiles.p, topo.n_sub, plans.max_Td, devsub.p,
                             Pf.p, 1, d_lam.p, d_sgnD.p, r.p, sympart.p, dbuf.p, st);
        launch_scatter_dot(nl, lam_pos.p, d_sgnD.p, dbuf.p, z.p, 2, a2, b2, partial.p, st);
      });
    }
    launch_beta_xpby(nl, partial.p, dsc.p, hsc_dev, z.p, p.p, st);
  }

  // Launch one iteration: replay of the captured graph, or the same kernels eagerly.
  template <typename T>
  void pcpg_iteration(const DBuf<T>& W, bool head_queued) {
    if (!cfg.graphs || profiling()) {
      pcpg_iteration_kernels<T>(W, head_queued);
      return;
    }
    if (head_queued) throw std::logic_error("graph iterations form their own W p");
    const void* key[3] = {W.p, p64 ? (const void*)Pd.p : (const void*)Pf.p, cpool.p};
    // the precision flags too: a released buffer's address can come back for the other one
    const int key_t = int(sizeof(T)) | (deluxe ? 256 : 0) | (p64 ? 512 : 0) | (deluxe && d64 ? 1024 : 0);
    if (!pcpg_exec || pcpg_key_t != key_t || std::memcmp(key, pcpg_key, sizeof(key))) {
      if (pcpg_exec) HIPCHECK(hipGraphExecDestroy(pcpg_exec));
      pcpg_exec = nullptr;
      hipGraph_t g = nullptr;
      HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      pcpg_iteration_kernels<T>(W);
      HIPCHECK(hipStreamEndCapture(st, &g));
      const hipError_t e = hipGraphInstantiate(&pcpg_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPCHECK(e);
      std::memcpy(pcpg_key, key, sizeof(key));
      pcpg_key_t = key_t;
    }
    HIPCHECK(hipGraphLaunch(pcpg_exec, st));
  }

  // PCPG on F lambda = d from the current lambda until ||r||/||d|| < tol.
  // `fresh` recomputes r, z, p from lambda; otherwise continues the recurrence
  // (r, z, p and the device scalar r.z persist between calls).
  template <typename T>
  // lam_zero: the start is lam = 0, so F lam = 0 and the first residual is d itself (the
  // F product of the start is skipped; bitwise the same r)
  int pcpg(const DBuf<T>& W, double tol, int max_it, bool fresh, int& its,
           double& dn, bool lam_zero = false) {
    const int64_t nl = topo.n_lam;
    volatile const double* hs = hsc;
    double o[3];
    if (fresh) {
      const double* a1[1] = {dv.p};
      dn = std::sqrt(dots(1, a1, a1, nl, o));
      if (dn < 1e-300) {
        last_dual = init_dual = 0.0;
        return FETIDP_SUCCESS;
      }
      if (!lam_zero) apply_F<T>(W, lam.p, q.p);
      if (last_learned && !lam_zero) {
        // scale the learned start by its F-norm optimum alpha = (lam.d)/(lam.F lam),
        // so ||lam0 - lam*||_F <= ||lam*||_F whatever the network predicted
        const double* a2[2] = {lam.p, lam.p};
        const double* b2[2] = {dv.p, q.p};
        dots(2, a2, b2, nl, o);
        const double alpha = o[1] > 0.0 ? o[0] / o[1] : 0.0;
        launch_scale2(nl, alpha, lam.p, q.p, st);
      }
      HIPCHECK(hipMemcpyAsync(r.p, dv.p, sizeof(double) * nl, hipMemcpyDeviceToDevice, st));
      if (!lam_zero) launch_axpy(nl, -1.0, q.p, r.p, st);
      apply_M(r.p, z.p);
      const double* a[2] = {r.p, r.p};
      const double* b[2] = {r.p, z.p};
      launch_dots(nl, 2, a, b, partial.p, st);
      launch_pcpg_beta(partial.p, dsc.p, hsc_dev, 1, st);
      HIPCHECK(hipMemcpyAsync(p.p, z.p, sizeof(double) * nl, hipMemcpyDeviceToDevice, st));
      HIPCHECK(hipStreamSynchronize(st));
      last_dual = init_dual = std::sqrt(hs[kScRR]) / dn;
      hist.push_back(last_dual);
      prev_ = last_dual;
    }
    if (last_dual < tol) return FETIDP_SUCCESS;
    // Eager launches queue the next iteration's W p (its first and longest launch) before
    // waiting for this iteration's scalars, so the GPU is not idle while the host polls
    // and launches.  Not within a factor 100 of the tolerance, where it would likely go
    // unused; when it does go unused it is only overwritten (recover() forms W lambda into
    // the same buffers before anything reads them).
    const bool queue_next = !cfg.graphs && !profiling();
    bool head = false;
    for (int k = 0; k < max_it; ++k) {
      const double seq = hs[kScSeq] + 1.0;
      pcpg_iteration<T>(W, head);
      head = queue_next && k + 1 < max_it && last_dual > 100.0 * tol;
      if (head) apply_W_split<T>(W, p.p);
      wait_scalars(seq);
      if (!(std::fabs(hs[kScPQ]) >= 1e-300)) return FETIDP_NUMERICAL_ERROR;
      ++its;
      last_dual = std::sqrt(hs[kScRR]) / dn;
      hist.push_back(last_dual);
      if (!std::isfinite(last_dual)) return FETIDP_NUMERICAL_ERROR;
      if (its > 10 && last_dual > prev_ * 1000.0) return FETIDP_DIVERGED;
      prev_ = last_dual;
      if (last_dual < tol) return FETIDP_SUCCESS;
    }
    return FETIDP_MAX_ITER_REACHED;
  }
  double prev_ = 0.0;
  double first_init = -1.0;  // initial dual residual of the first PCPG stage of a solve

  // Wait until the iteration's scalars are on the host: its last launch writes them to the
  // mapped host block and then bumps the sequence slot, so the host polls that slot (the
  // next iteration can be queued before the rest of the stream drains) and falls back on
  // the stream's completion.  The stream is queried at most once per 2 ms (none in the
  // first 2 ms): a query of a busy stream delayed the launch queued behind it by ~6 us.
  void wait_scalars(double seq) {
    volatile const double* hs = hsc;
    double t_next = -1.0;
    for (uint32_t spin = 1;; ++spin) {
      if (hs[kScSeq] == seq) break;
      if ((spin & 255) == 0) {
        const double t = now_ms();
        if (t_next < 0.0) t_next = t + 2.0;
        if (t < t_next) continue;
        t_next = t + 2.0;
        const hipError_t e = hipStreamQuery(st);
        if (e == hipSuccess) break;
        if (e != hipErrorNotReady) HIPCHECK(e);
      }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
  }

  template <typename T>
  int solve_impl(const DBuf<T>& pool, const DBuf<T>& inv, const DBuf<T>& W,
                 const double* h_lam0, fetidp_stats& S) {
    const int64_t nl = topo.n_lam, nf = topo.n_free;
    hist.clear();
    ref_hist.clear();
    int its = 0, status = FETIDP_SUCCESS;
    double dn = 0.0, res = 0.0;
    first_init = -1.0;
    stalled = false;
    {
      // ||b|| < 1e-15: x = 0, no iteration, residual 0 (src/numeric/iterative_solvers.cpp:436-442)
      double o[1];
      const double* af[1] = {f_glob.p};
      const double fn = std::sqrt(dots(1, af, af, nf, o));
      if (!std::isfinite(fn)) throw InputError("solve: right-hand side is not finite");
      if (fn < 1e-15) {
        xtot.zero(st);
        lam.zero(st);
        last_dual = init_dual = 0.0;
        S.iterations = 0;
        S.refinements = 0;
        S.residual_norm = 0.0;
        S.dual_residual = 0.0;
        S.initial_dual_residual = 0.0;
        return FETIDP_SUCCESS;
      }
    }
    if (!fp32()) {
      condense<T>(pool, inv, W, f_glob.p, true);
      if (h_lam0) HIPCHECK(hipMemcpyAsync(lam.p, h_lam0, sizeof(double) * nl, hipMemcpyHostToDevice, st));
      else if (lam_model) learned_guess();
      else lam.zero(st);
      const bool zero_start = !h_lam0 && !lam_model;
      double tol = cfg.dual_tolerance, prev = 1e300, amp = 1.0;
      bool fresh = true;
      for (int round = 0; round < 6; ++round) {
        status = pcpg<T>(W, tol, cfg.max_iterations - its, fresh, its, dn, fresh && zero_start);
        fresh = false;
        recover<T>(pool, inv, W, x.p);
        res = global_residual(x.p, f_glob.p);
        ref_hist.push_back(res);
        // global residual per unit of dual residual (an upper estimate of how much a dual
        // residual is amplified in the recovered solution)
        if (round == 0 && last_dual > 0.0) amp = std::max(1.0, res / last_dual);
        if (status != FETIDP_SUCCESS || res <= cfg.tolerance) break;
        if (res > 0.5 * prev) break;  // a tighter dual tolerance no longer helps
        // the recovered residual sits far above the dual one: a residual floor of the
        // ill-conditioned local problems, which only refinement removes
        if (res > kFloorRatio * last_dual) break;
        prev = res;
        tol *= 0.1;
      }
      HIPCHECK(hipMemcpyAsync(xtot.p, x.p, sizeof(double) * nf, hipMemcpyDeviceToDevice, st));
      first_init = init_dual;
      // FP64 iterative refinement with the same factor when the recovered solution has
      // a residual floor (ill-conditioned local problems): x += FETI-DP^-1 (f - K x);
      // rglob holds f - K xtot from the last residual evaluation
      int k = 0;
      // Amplification of a correction's dual residual into the recovered residual: a
      // first solve that ended on a floor measured the floor, not the amplification, so
      // the first correction assumes kCorrAmp (deluxe average in the recovery) or
      // kFloorRatio, and later ones what the previous correction measured
      double amp_ref = std::min(amp, deluxe ? kCorrAmp : kFloorRatio);
      while (status == FETIDP_SUCCESS && res > cfg.tolerance && k < cfg.max_refinements) {
        // the Dirichlet recovery re-solves every interior block, so the residual's interior
        // part is rounding (cube96j: 3e-11 of ||f||); below 1 % of the tolerance it is
        // dropped (it stays in the final residual, which is checked) and the correction's
        // condensation skips the interior columns
        condense<T>(pool, inv, W, rglob.p, false, interior_residual() <= 1e-2 * cfg.tolerance);
        lam.zero(st);
        // the correction needs only enough dual accuracy to bring ||f - K x|| under the
        // tolerance with a 4x margin
        const double rtol = std::min(1e-2, std::max(cfg.dual_tolerance, 0.25 * cfg.tolerance / (res * amp_ref)));
        status = pcpg<T>(W, rtol, cfg.max_iterations - its, true, its, dn, true);
        if (status != FETIDP_SUCCESS) break;
        recover<T>(pool, inv, W, x.p);
        launch_axpy(nf, 1.0, x.p, xtot.p, st);
        const double res_before = res;
        res = global_residual(xtot.p, f_glob.p);
        ref_hist.push_back(res);
        if (last_dual > 0.0) amp_ref = std::max(1.0, res / (res_before * last_dual));
        ++k;
      }
      S.refinements = k;
    } else {
      // FP64 flexible CG on the assembled system, preconditioned by the FP32
      // FETI-DP solve (inner PCPG to cfg.inner_tolerance): the outer recurrence,
      // inner products and residual are FP64, the subdomain work is FP32.
      xtot.zero(st);
      HIPCHECK(hipMemcpyAsync(f_cur.p, f_glob.p, sizeof(double) * nf, hipMemcpyDeviceToDevice, st));
      double o[3];
      const double* af[1] = {f_glob.p};
      const double fn2 = dots(1, af, af, nf, o);
      int k = 0;
      res = 1.0;
      double pq_prev = 0.0;
      for (k = 0; k < cfg.max_refinements; ++k) {
        condense<T>(pool, inv, W, f_cur.p, k == 0);
        if (k == 0 && h_lam0) HIPCHECK(hipMemcpyAsync(lam.p, h_lam0, sizeof(double) * nl, hipMemcpyHostToDevice, st));
        else if (k == 0 && lam_model) learned_guess();
        else lam.zero(st);
        status = pcpg<T>(W, cfg.inner_tolerance, cfg.max_iterations - its, true, its, dn,
                         k > 0 || (!h_lam0 && !lam_model));
        if (status == FETIDP_NUMERICAL_ERROR || status == FETIDP_DIVERGED) break;
        recover<T>(pool, inv, W, x.p);                      // z = M32^-1 r
        if (k == 0) {
          HIPCHECK(hipMemcpyAsync(pf.p, x.p, sizeof(double) * nf, hipMemcpyDeviceToDevice, st));
        } else {
          const double* a1[1] = {x.p};
          const double* b1[1] = {qf.p};
          const double beta = -dots(1, a1, b1, nf, o) / pq_prev;
          launch_xpby(nf, x.p, beta, pf.p, st);       // p = z + beta p
        }
        launch_csr_spmv(nf, csr_rp.p, csr_col.p, csr_val.p, pf.p, qf.p, st);  // q = K p
        const double* a2[2] = {pf.p, pf.p};
        const double* b2[2] = {f_cur.p, qf.p};
        dots(2, a2, b2, nf, o);
        pq_prev = o[1];
        if (!(pq_prev > 0.0)) { status = FETIDP_NUMERICAL_ERROR; break; }
        const double alpha = o[0] / pq_prev;
        launch_axpy2(nf, alpha, pf.p, qf.p, xtot.p, f_cur.p, st);  // x += a p, r -= a q
        const double* a3[1] = {f_cur.p};
        res = std::sqrt(dots(1, a3, a3, nf, o) / fn2);
        ref_hist.push_back(res);
        if (res <= 0.5 * cfg.tolerance) { status = FETIDP_SUCCESS; break; }
        if (its >= cfg.max_iterations) { status = FETIDP_MAX_ITER_REACHED; break; }
        const size_t nh = ref_hist.size();
        if (nh >= 3 && res > kFp32Stall * ref_hist[nh - 3]) {
          // the FP32 preconditioner gains less than 2x per outer step: hand over to FP64
          stalled = true;
          status = FETIDP_MAX_ITER_REACHED;
          break;
        }
      }
      if (stalled) {
        S.refinements = k + 1;
        S.iterations = its;
        return status;
      }
      // the recursive residual is confirmed with the true FP64 residual
      res = global_residual(xtot.p, f_glob.p);
      ref_hist.push_back(res);
      if (res > cfg.tolerance && status == FETIDP_SUCCESS) status = FETIDP_MAX_ITER_REACHED;
      S.refinements = k + 1;
    }
    if (status == FETIDP_SUCCESS && res > cfg.tolerance) status = FETIDP_MAX_ITER_REACHED;
    S.iterations = its;
    S.residual_norm = res;
    S.dual_residual = last_dual;
    S.initial_dual_residual = first_init >= 0.0 ? first_init : init_dual;
    return status;
  }

  int solve(const double* h_lam0, double* h_x, fetidp_stats& S) {
    if (!factored) throw InputError("solve before factor");
    const double t0 = now_ms();
    last_guess_ms = 0.0;
    last_learned = false;
    int status = fp32() ? solve_impl<float>(pool32, inv32, W32, h_lam0, S)
                        : solve_impl<double>(pool64, inv64, W64, h_lam0, S);
    if (stalled) {
      // FP32 refinement stalled: refactor in FP64 and solve again (fell_back = 1); the
      // FP32 iterations stay counted in the statistics
      const int its32 = S.iterations;
      char buf[160];
      std::snprintf(buf, sizeof buf, "FP32 refinement stalled at residual %.3g after %d outer steps",
                    ref_hist.empty() ? 0.0 : ref_hist.back(), S.refinements);
      fallback_reason = buf;
      if (cfg.verbose & 1) std::fprintf(stderr, "[fetidp] %s: refactoring in FP64\n", buf);
      fell_back = true;
      stalled = false;
      pool32.release(); inv32.release(); W32.release();
      const double tf = now_ms();
      factor_impl<double>(pool64, inv64, W64);
      last_factor_ms += now_ms() - tf;
      status = solve_impl<double>(pool64, inv64, W64, h_lam0, S);
      S.iterations += its32;
    }
    if (h_x) HIPCHECK(hipMemcpyAsync(h_x, xtot.p, sizeof(double) * topo.n_free, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (last_learned) {
      float ms = 0.f;
      HIPCHECK(hipEventElapsedTime(&ms, guess_ev[0], guess_ev[1]));
      last_guess_ms = ms;
    }
    prof_collect();
    S.solve_ms = now_ms() - t0;
    S.status = status;
    S.guess_ms = last_guess_ms;
    S.learned_guess = last_learned ? 1 : 0;
    S.precond_fp64 = p64 ? 1 : 0;
    S.deluxe = deluxe ? 1 : 0;
    return status;
  }

  // Autotuner: time factor + solve for every combination of the launch choices that
  // leave the arithmetic unchanged (where the diagonal factorizations run: same
  // stream, second stream, or at the front of the column-step launch; PCPG graph
  // replay; throughput tile kernels capped at 128 registers or allocated freely),
  // check that every combination returns bitwise the same solution, keep the fastest;
  // then, for a grouped coarse solve, time the group widths with that launch choice.
  // The width changes the order of the coarse sums and so the PCPG trajectory: a width
  // is kept only if its solve succeeds with the residual within the tolerance (its
  // solution is as valid as the first, not bitwise equal).  Per variant 5 values in out:
  // {side_stream, graphs, tile_waves, coarse group width (0 = dense inverse), ms}.
  int autotune(int reps, double* out, int cap) {
    if (!factored && !rhs_set) throw InputError("autotune needs a right-hand side (set_rhs_*) first");
    reps = std::max(reps, 1);
    const fetidp_config base = cfg;
    std::vector<double> ref, xh(size_t(topo.n_free));
    double best = 1e300;
    int best_v = -1, k = 0, nv = 0;
    auto pick = [&](int v) {
      cfg.side_stream = v % 3;
      cfg.graphs = (v / 3) & 1;
      cfg.tile_waves = (v / 6) & 1 ? 0 : 4;
    };
    fetidp_stats S{};
    auto time_it = [&]() {
      factor();
      solve(nullptr, nullptr, S);  // warm-up (graph capture, allocations)
      HIPCHECK(hipStreamSynchronize(st));
      const double t0 = now_ms();
      for (int r = 0; r < reps; ++r) {
        factor();
        solve(nullptr, nullptr, S);
      }
      const double ms = (now_ms() - t0) / reps;
      HIPCHECK(hipMemcpy(xh.data(), xtot.p, sizeof(double) * xh.size(), hipMemcpyDeviceToHost));
      return ms;
    };
    auto record = [&](double ms) {
      if (k + 5 <= cap && out) {
        out[k++] = cfg.side_stream;
        out[k++] = cfg.graphs;
        out[k++] = cfg.tile_waves;
        out[k++] = plans.coarse_grouped ? plans.coarse_group : 0;
        out[k++] = ms;
      }
      ++nv;
    };
    for (int v = 0; v < 12; ++v) {
      pick(v);
      const double ms = time_it();
      if (ref.empty()) ref = xh;
      else if (std::memcmp(ref.data(), xh.data(), sizeof(double) * xh.size()) != 0) {
        cfg = base;
        throw NumericError("autotune: launch variant changed the solution");
      }
      record(ms);
      if (ms < best) {
        best = ms;
        best_v = v;
      }
    }
    cfg = base;
    pick(best_v);
    if (plans.coarse_grouped) {
      const int tn = int(plans.coarse_ncp / kTile);
      const int g0 = plans.coarse_group;
      int best_g = g0;
      double best_ms = best;
      std::vector<int> widths{g0};
      for (int G : {8, 16, 24, 32, 48, 64})
        if (G != g0 && G <= std::max(tn, 1)) widths.push_back(G);
      for (size_t w = 1; w < widths.size(); ++w) {
        set_coarse_group(widths[w]);
        const double ms = time_it();
        record(ms);
        const bool valid = S.status == FETIDP_SUCCESS && S.residual_norm <= cfg.tolerance;
        if (valid && ms < best_ms) {
          best_ms = ms;
          best_g = widths[w];
        }
      }
      set_coarse_group(best_g);
    }
    factor();
    return nv;
  }

  template <typename F64, typename F32>
  void with_ops(F64 f64, F32 f32) {
    if (!factored) throw InputError("operators need fetidp_factor first");
    if (fp32()) f32(); else f64();
  }
};

namespace {
int32_t guard(fetidp_handle h, const std::function<int32_t()>& fn) {
  if (!h) return FETIDP_INVALID_INPUT;
  try {
    h->err.clear();
    return fn();
  } catch (const InputError& e) {
    h->err = e.what();
    return FETIDP_INVALID_INPUT;
  } catch (const NumericError& e) {
    h->err = e.what();
    return FETIDP_NUMERICAL_ERROR;
  } catch (const std::exception& e) {
    h->err = e.what();
    return FETIDP_NUMERICAL_ERROR;
  }
}
}  // namespace

template <typename T>
static int32_t tile_cholesky_test(int64_t m, int64_t pc, const double* A, double* out) {
  const int Ta = int(m / kTile), Tp = int(pc / kTile);
  const int64_t elems = int64_t(Ta) * (Ta + 1) / 2 * kTileElems;
  std::vector<T> pool(elems, T(0));
  for (int64_t i = 0; i < m; ++i)
    for (int64_t j = 0; j <= i; ++j) {
      pool[pool_pos(0, i, j)] = T(A[i * m + j]);
      if (i / kTile == j / kTile) pool[tile_offset(i / kTile, i / kTile) + (i % kTile) * kTile + j % kTile] = T(A[i * m + j]);
    }
  TileSchedule S = single_matrix_schedule(Tp, Ta, sizeof(T) == 8);
  DBuf<T> dpool, dinv;
  DBuf<TileWork> items;
  DBuf<int64_t> base, rabs;
  DBuf<int> fail;
  dpool.upload(pool);
  dinv.alloc(size_t(std::max(Tp, 1)) * kTileElems);
  items.upload(S.items);
  base.upload(std::vector<int64_t>{0});
  std::vector<int64_t> ra(Ta);
  for (int I = 0; I < Ta; ++I) ra[I] = tile_offset(I, 0);  // dense layout: no envelope
  rabs.upload(ra);
  fail.alloc(1);
  HIPCHECK(hipMemset(fail.p, 0, sizeof(int)));
  for (const TileLaunch& L : S.launches)
    launch_tile_mode<T>(L.mode, items.p + L.off, L.n, base.p, rabs.p, base.p, dpool.p, dinv.p, fail.p, nullptr);
  HIPCHECK(hipDeviceSynchronize());
  HIPCHECK(hipMemcpy(pool.data(), dpool.p, sizeof(T) * elems, hipMemcpyDeviceToHost));
  int hf = 0;
  HIPCHECK(hipMemcpy(&hf, fail.p, sizeof(int), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < m; ++i)
    for (int64_t j = 0; j < m; ++j) out[i * m + j] = j <= i ? double(pool[pool_pos(0, i, j)]) : 0.0;
  return hf ? FETIDP_NUMERICAL_ERROR : FETIDP_SUCCESS;
}


extern "C" {

void fetidp_config_default(fetidp_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->precision = FETIDP_FP64;
  cfg->max_iterations = 1000;
  cfg->tolerance = 1e-8;
  cfg->dual_tolerance = 1e-11;
  cfg->max_refinements = 12;
  cfg->verbose = 0;
  cfg->inner_tolerance = 1e-2;
  cfg->graphs = 1;
  cfg->side_stream = 2;
  cfg->tile_waves = 4;
}

// Copy of `in` with unset fields at their defaults; false for an unknown precision.
static bool normalized_config(const fetidp_config* in, fetidp_config& out) {
  if (in) out = *in;
  else fetidp_config_default(&out);
  if (out.precision != FETIDP_FP64 && out.precision != FETIDP_FP32_REFINE) return false;
  if (out.max_iterations <= 0) out.max_iterations = 1000;
  if (!(out.tolerance > 0)) out.tolerance = 1e-8;
  if (!(out.dual_tolerance > 0)) out.dual_tolerance = 1e-11;
  if (out.max_refinements <= 0) out.max_refinements = 12;
  if (!(out.inner_tolerance > 0)) out.inner_tolerance = 1e-2;
  if (out.tile_waves != 0) out.tile_waves = 4;
  if (out.side_stream < 0 || out.side_stream > 2) out.side_stream = 2;
  if (out.coarse_solve < 0 || out.coarse_solve > 2) return false;
  if (out.coarse_group < 0 || out.coarse_group > 64) return false;
  return true;
}

fetidp_handle fetidp_create(const fetidp_config* cfg) {
  try {
    fetidp_config c;
    if (!normalized_config(cfg, c)) return nullptr;
    auto* h = new fetidp_solver();
    h->cfg = c;
    return h;
  } catch (...) {
    return nullptr;
  }
}

int32_t fetidp_set_config(fetidp_handle h, const fetidp_config* cfg) {
  return guard(h, [&] {
    fetidp_config c;
    if (!cfg || !normalized_config(cfg, c)) throw InputError("set_config: bad configuration");
    // factors are per precision; the preconditioner blocks per scaling and precision
    if (c.precision != h->cfg.precision || c.scaling != h->cfg.scaling ||
        c.precond_precision != h->cfg.precond_precision)
      h->factored = false;
    h->cfg = c;
    if (h->setup_done && h->coarse_group_tiles() != h->coarse_group_used) {
      // another coarse solve: new coarse plan; the device data follow at the next factor
      h->coarse_group_used = h->coarse_group_tiles();
      build_coarse_plan(h->topo, h->plans, h->coarse_group_used);
      h->factored = false;
      h->device_ready = false;
      if (h->pcpg_exec) HIPCHECK(hipGraphExecDestroy(h->pcpg_exec));
      h->pcpg_exec = nullptr;
    }
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_get_config(fetidp_handle h, fetidp_config* cfg) {
  return guard(h, [&] {
    if (!cfg) throw InputError("get_config: null output");
    *cfg = h->cfg;
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_setup(fetidp_handle h, int64_t n_nodes, const double* xyz, int64_t n_tets,
                     const int32_t* tet2node, const int32_t* tet_material, int32_t n_mat,
                     const double* nu, const double* kappa, const int32_t* tet_subdomain,
                     int64_t n_pec, const int32_t* pec_edge_nodes) {
  return guard(h, [&] {
    h->setup(n_nodes, xyz, n_tets, tet2node, tet_material, n_mat, nu, kappa, tet_subdomain, n_pec,
             pec_edge_nodes);
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_set_materials(fetidp_handle h, int32_t n_mat, const double* nu, const double* kappa) {
  return guard(h, [&] {
    h->set_materials(n_mat, nu, kappa);
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_set_rhs_tet_J(fetidp_handle h, const double* jq) {
  return guard(h, [&] {
    h->set_rhs_J(jq);
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_set_rhs_free(fetidp_handle h, const double* f) {
  return guard(h, [&] {
    h->set_rhs_free(f);
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_factor(fetidp_handle h) {
  return guard(h, [&] {
    h->last_factor_ms = h->factor();
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_solve(fetidp_handle h, const double* lambda0, double* x_free, fetidp_stats* stats) {
  fetidp_stats S{};
  int32_t rc = guard(h, [&] { return int32_t(h->solve(lambda0, x_free, S)); });
  if (h) {
    S.status = rc;
    S.fell_back = h->fell_back ? 1 : 0;
    S.factor_ms = h->last_factor_ms;
    S.symbolic_ms = h->symbolic_ms;
  }
  if (stats) *stats = S;
  return rc;
}

int32_t fetidp_get_sizes(fetidp_handle h, int64_t* out) {
  return guard(h, [&] {
    if (!h->setup_done || !out) throw InputError("get_sizes before setup");
    const Topology& T = h->topo;
    out[0] = T.n_edges;
    out[1] = T.n_free;
    out[2] = T.n_lam;
    out[3] = T.n_c;
    out[4] = T.n_sub;
    out[5] = T.n_tets;
    out[6] = h->plans.pool_elems;
    out[7] = T.w_elems + T.p_elems;
    out[8] = h->plans.coarse_ncp;
    out[9] = h->plans.max_m;
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_get_choices(fetidp_handle h, int32_t* deluxe, int32_t* precond_fp64,
                           double* box_overlap) {
  return guard(h, [&] {
    if (!h->setup_done) throw InputError("get_choices before setup");
    const bool dl = h->use_deluxe() && h->plans.faces.n > 0;
    if (deluxe) *deluxe = dl ? 1 : 0;
    if (precond_fp64) *precond_fp64 = h->precond_fp64(dl) ? 1 : 0;
    if (box_overlap) *box_overlap = h->box_overlap;
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_get_topology(fetidp_handle h, int32_t* tet_free, int32_t* edge_nodes,
                            int32_t* lam_free, int32_t* primal_free) {
  return guard(h, [&] {
    if (!h->setup_done) throw InputError("get_topology before setup");
    const Topology& T = h->topo;
    if (tet_free) std::memcpy(tet_free, T.tet_free.data(), sizeof(int32_t) * T.tet_free.size());
    if (edge_nodes) std::memcpy(edge_nodes, T.edge_nodes.data(), sizeof(int32_t) * T.edge_nodes.size());
    if (lam_free) std::memcpy(lam_free, T.free_of_lam.data(), sizeof(int32_t) * T.free_of_lam.size());
    if (primal_free) std::memcpy(primal_free, T.free_of_c.data(), sizeof(int32_t) * T.free_of_c.size());
    return int32_t(FETIDP_SUCCESS);
  });
}

int64_t fetidp_get_global_nnz(fetidp_handle h) {
  if (!h || !h->setup_done) return -1;
  return int64_t(h->plans.csr.col.size());
}

int32_t fetidp_get_global_system(fetidp_handle h, int32_t* row_ptr, int32_t* col, double* val,
                                 double* f) {
  return guard(h, [&] {
    if (!h->factored) throw InputError("get_global_system needs fetidp_factor first");
    const CsrPlan& C = h->plans.csr;
    if (row_ptr) std::memcpy(row_ptr, C.row_ptr.data(), sizeof(int32_t) * C.row_ptr.size());
    if (col) std::memcpy(col, C.col.data(), sizeof(int32_t) * C.col.size());
    if (val) HIPCHECK(hipMemcpy(val, h->csr_val.p, sizeof(double) * C.col.size(), hipMemcpyDeviceToHost));
    if (f) HIPCHECK(hipMemcpy(f, h->f_glob.p, sizeof(double) * h->topo.n_free, hipMemcpyDeviceToHost));
    return int32_t(FETIDP_SUCCESS);
  });
}

static int32_t dual_io(fetidp_handle h, const double* in, double* out, int which) {
  return guard(h, [&] {
    if (!in || !out) throw InputError("null vector");
    const int64_t nl = h->topo.n_lam;
    h->with_ops(
        [&] {
          HIPCHECK(hipMemcpy(h->p.p, in, sizeof(double) * nl, hipMemcpyHostToDevice));
          if (which == 0) h->apply_F<double>(h->W64, h->p.p, h->q.p);
          else h->apply_M(h->p.p, h->q.p);
        },
        [&] {
          HIPCHECK(hipMemcpy(h->p.p, in, sizeof(double) * nl, hipMemcpyHostToDevice));
          if (which == 0) h->apply_F<float>(h->W32, h->p.p, h->q.p);
          else h->apply_M(h->p.p, h->q.p);
        });
    HIPCHECK(hipStreamSynchronize(h->st));
    HIPCHECK(hipMemcpy(out, h->q.p, sizeof(double) * nl, hipMemcpyDeviceToHost));
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_apply_dual_operator(fetidp_handle h, const double* p, double* q) { return dual_io(h, p, q, 0); }
int32_t fetidp_apply_preconditioner(fetidp_handle h, const double* r, double* z) { return dual_io(h, r, z, 1); }

int32_t fetidp_get_dual_rhs(fetidp_handle h, double* d) {
  return guard(h, [&] {
    if (!d) throw InputError("null vector");
    h->with_ops([&] { h->condense<double>(h->pool64, h->inv64, h->W64, h->f_glob.p, true); },
                [&] { h->condense<float>(h->pool32, h->inv32, h->W32, h->f_glob.p, true); });
    HIPCHECK(hipStreamSynchronize(h->st));
    HIPCHECK(hipMemcpy(d, h->dv.p, sizeof(double) * h->topo.n_lam, hipMemcpyDeviceToHost));
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_get_lambda(fetidp_handle h, double* lam) {
  return guard(h, [&] {
    if (!lam || !h->setup_done) throw InputError("get_lambda: bad call");
    HIPCHECK(hipMemcpy(lam, h->lam.p, sizeof(double) * h->topo.n_lam, hipMemcpyDeviceToHost));
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_get_history(fetidp_handle h, double* hist, int32_t cap) {
  if (!h) return -1;
  const int32_t n = int32_t(h->hist.size());
  for (int32_t i = 0; i < n && i < cap && hist; ++i) hist[i] = h->hist[i];
  return n;
}

int32_t fetidp_get_refinement_history(fetidp_handle h, double* hist, int32_t cap) {
  if (!h) return -1;
  const int32_t n = int32_t(h->ref_hist.size());
  for (int32_t i = 0; i < n && i < cap && hist; ++i) hist[i] = h->ref_hist[i];
  return n;
}

int32_t fetidp_get_coarse_matrix(fetidp_handle h, double* S) {
  return guard(h, [&] {
    if (!h->factored || !S) throw InputError("get_coarse_matrix needs fetidp_factor first");
    const int64_t nc = h->topo.n_c, ncp = h->plans.coarse_ncp;
    // the coarse pool keeps L after the factorization; rebuild S~ by re-gathering
    DBuf<double> tmp;
    tmp.alloc(h->plans.coarse_pool_elems);
    tmp.zero(h->st);
    if (h->fp32())
      launch_gather_pool<float>(h->plans.coarse.size(), h->coarse_target.p, h->coarse_ptr.p,
                                h->coarse_srcpos.p, h->pool32.p, h->written(), tmp.p, h->st);
    else
      launch_gather_pool<double>(h->plans.coarse.size(), h->coarse_target.p, h->coarse_ptr.p,
                                 h->coarse_srcpos.p, h->pool64.p, h->written(), tmp.p, h->st);
    HIPCHECK(hipStreamSynchronize(h->st));
    std::vector<double> pool(h->plans.coarse_pool_elems);
    HIPCHECK(hipMemcpy(pool.data(), tmp.p, sizeof(double) * pool.size(), hipMemcpyDeviceToHost));
    const std::vector<int64_t>& cra = h->plans.coarse_rowabs;
    const std::vector<int32_t>& cft = h->plans.coarse_ftile;
    for (int64_t a = 0; a < nc; ++a)
      for (int64_t b = 0; b < nc; ++b) {
        const int64_t hi = std::max(a, b), lo = std::min(a, b);
        S[a * nc + b] = lo / kTile < cft[hi / kTile] ? 0.0 : pool[env_pos(cra.data(), hi, lo)];
      }
    (void)ncp;
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_get_profile(fetidp_handle h, double* out, int32_t cap) {
  if (!h || !out) return -1;
  const int n = fetidp_solver::kProfSlots;
  int k = 0;
  for (int i = 0; i < n && k + 1 < cap; ++i) {
    out[k++] = h->prof_ms[i];
    out[k++] = h->prof_cnt[i];
  }
  const TileSchedule* sch[2] = {h->fp32() ? &h->plans.local_sched_ll : &h->plans.local_sched,
                                &h->plans.coarse_sched};
  // Real (unpadded) rows of tile row I: subdomain matrices [i | D | c | D'] pad each block
  // to 64; the coarse pool pads S~ (and, for the dense inverse, the identity block).
  const Topology& T = h->topo;
  auto real_rows = [&](int q, int mat, int I) -> double {
    int64_t r0 = int64_t(I) * kTile, real;
    if (q == 1) {
      if (r0 >= h->plans.coarse_ncp) r0 -= h->plans.coarse_ncp;
      real = T.n_c - r0;
    } else {
      const SubLayout& L = T.sub[size_t(mat)];
      if (r0 < L.o_d) real = L.ni - r0;
      else if (r0 < L.o_c) real = L.nd - (r0 - L.o_d);
      else if (r0 < L.o_a) real = L.nc - (r0 - L.o_c);
      else real = L.nd - (r0 - L.o_a);
    }
    return double(std::min<int64_t>(std::max<int64_t>(real, 0), kTile));
  };
  // executed: 2 * 64^3 per tile product as issued (the TRSM as a product with the stored
  // lower-triangular inverse, 160 of 256 MFMAs; the full diagonal-tile share); useful: the
  // same products on the real rows only, the TRSM as a triangular solve (r_I r_J^2) and
  // the lower half of the diagonal share (r_I^2 r_J)
  for (int useful = 0; useful < 2; ++useful)
    for (int q = 0; q < 2; ++q)
      for (int mode = 0; mode < 3; ++mode) {
        double fl = 0.0;
        for (const TileLaunch& L : sch[q]->launches)
          // mode 3 is timed with the diagonals, mode 4 (early Schur) with the Schur updates
          if (L.mode == mode || (mode == 2 && L.mode == 3) || (mode == 0 && L.mode == 4))
            for (int64_t i = 0; i < L.n; ++i) {
              const TileWork& w = sch[q]->items[L.off + i];
              const bool trsm = (L.mode == 1 || L.mode == 3) && !(w.flags & 2);  // not a lookahead
              if (!useful) {
                fl += 2.0 * kTile * kTile * kTile *
                      (double(w.k1 - w.k0) + (trsm ? 0.625 + (w.flags & 1) + ((w.flags & 64) ? 1.0 : 0.0) : 0.0));
                continue;
              }
              const double rI = real_rows(q, w.mat, w.I), rJ = real_rows(q, w.mat, w.J);
              double kr = 0.0;
              for (int K = w.k0; K < w.k1; ++K) kr += real_rows(q, w.mat, K);
              fl += 2.0 * rI * rJ * kr;
              if (trsm) fl += rI * rJ * rJ + ((w.flags & 1) ? rI * rI * rJ : 0.0);
              if (trsm && (w.flags & 64)) fl += rI * rI * real_rows(q, w.mat, w.I - 2);
            }
        if (k < cap) out[k++] = fl;
      }
  return k;
}

void fetidp_reset_profile(fetidp_handle h) {
  if (!h) return;
  for (int i = 0; i < fetidp_solver::kProfSlots; ++i) h->prof_ms[i] = h->prof_cnt[i] = 0.0;
}

int32_t fetidp_set_lambda0_model(fetidp_handle h, struct emag_mpnn* model) {
  return guard(h, [&] {
    h->lam_model = model;
    return int32_t(FETIDP_SUCCESS);
  });
}

int64_t fetidp_get_dual_graph(fetidp_handle h, int32_t* row_ptr, int32_t* col, float* coef) {
  if (!h) return -1;
  int64_t nnz = -1;
  const int32_t rc = guard(h, [&] {
    if (!h->setup_done) throw InputError("get_dual_graph before setup");
    h->build_dual_graph();
    nnz = int64_t(h->g_col.size());
    if (row_ptr) std::memcpy(row_ptr, h->g_rp.data(), sizeof(int32_t) * h->g_rp.size());
    if (col) std::memcpy(col, h->g_col.data(), sizeof(int32_t) * h->g_col.size());
    if (coef) std::memcpy(coef, h->g_coef.data(), sizeof(float) * h->g_coef.size());
    return int32_t(FETIDP_SUCCESS);
  });
  return rc == FETIDP_SUCCESS ? nnz : -1;
}

int32_t fetidp_get_lambda_features(fetidp_handle h, float* features, double* scales) {
  return guard(h, [&] {
    if (!features) throw InputError("get_lambda_features: null output");
    h->with_ops([&] { h->condense<double>(h->pool64, h->inv64, h->W64, h->f_glob.p, true); },
                [&] { h->condense<float>(h->pool32, h->inv32, h->W32, h->f_glob.p, true); });
    h->lambda_features();
    HIPCHECK(hipStreamSynchronize(h->st));
    HIPCHECK(hipMemcpy(features, h->lam_X.p, sizeof(float) * h->topo.n_lam * kMpF, hipMemcpyDeviceToHost));
    if (scales) {
      double f[4];
      HIPCHECK(hipMemcpy(f, h->fsc.p, sizeof(f), hipMemcpyDeviceToHost));
      scales[0] = f[3];
      scales[1] = f[2];
    }
    return int32_t(FETIDP_SUCCESS);
  });
}

int32_t fetidp_autotune(fetidp_handle h, int32_t repetitions, double* timings, int32_t cap) {
  int32_t n = 0;
  const int32_t rc = guard(h, [&] {
    n = h->autotune(repetitions, timings, cap);
    return int32_t(FETIDP_SUCCESS);
  });
  return rc == FETIDP_SUCCESS ? n : -rc;
}

const char* fetidp_last_error(fetidp_handle h) { return h ? h->err.c_str() : "null handle"; }

void fetidp_destroy(fetidp_handle h) { delete h; }

int32_t emag_mfma_layout_probe(int32_t precision, double* out) {
  if (!out) return FETIDP_INVALID_INPUT;
  try {
    DBuf<double> d;
    d.alloc(512);
    launch_mfma_probe(precision == FETIDP_FP32_REFINE, d.p);
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemcpy(out, d.p, sizeof(double) * 512, hipMemcpyDeviceToHost));
    return FETIDP_SUCCESS;
  } catch (...) {
    return FETIDP_NUMERICAL_ERROR;
  }
}

int32_t emag_tile_cholesky_test(int32_t precision, int64_t m, int64_t p_cols, const double* A,
                                double* out) {
  if (m <= 0 || m % kTile || p_cols < 0 || p_cols % kTile || p_cols > m || !A || !out)
    return FETIDP_INVALID_INPUT;
  try {
    return precision == FETIDP_FP32_REFINE ? tile_cholesky_test<float>(m, p_cols, A, out)
                                           : tile_cholesky_test<double>(m, p_cols, A, out);
  } catch (...) {
    return FETIDP_NUMERICAL_ERROR;
  }
}

int32_t emag_nedelec_integrate(int64_t n_tets, const double* d_xyz, const int32_t* d_tet2node,
                               const int32_t* d_tet_material, const double* d_nu,
                               const double* d_kappa, const double* d_jq, double* d_Ke,
                               double* d_Me, double* d_Fe, void* stream) {
  if (n_tets <= 0 || !d_xyz || !d_tet2node || !d_tet_material || !d_nu || !d_kappa || !d_Ke || !d_Me)
    return FETIDP_INVALID_INPUT;
  if (d_jq && !d_Fe) return FETIDP_INVALID_INPUT;
  launch_integrate(n_tets, d_xyz, d_tet2node, d_tet_material, d_nu, d_kappa, d_jq, 0, d_Ke, d_Me,
                   d_Fe, static_cast<hipStream_t>(stream));
  return hipGetLastError() == hipSuccess ? FETIDP_SUCCESS : FETIDP_NUMERICAL_ERROR;
}

}  // extern "C"

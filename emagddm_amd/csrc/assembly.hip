// GPU global Nedelec assembly and FP64 residual operator behind include/emagddm_assembly.h.
//
// EMAssembly::assemble (src/numeric/em_assembly.cpp:144-411) scatters element matrices
// into COO and builds a CSR on the host.  Here the host builds, once per mesh, the CSR
// pattern and a gather plan (for every CSR entry the signed list of element entries that
// sum into it, in a fixed order); the numeric part is two launches: k_integrate writes
// the element matrices (structure of arrays) and k_assemble gathers each CSR value in one
// thread with no atomics, so results are bitwise repeatable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "emagddm_assembly.h"
#include "emagddm_fetidp.h"
#include "kernels.h"
#include "plan.h"
#include "topology.h"

using namespace emag;

namespace {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
struct DevArray {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    n = count;
  }
  void upload(const std::vector<T>& v) {
    alloc(v.size());
    if (!v.empty()) check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevArray() { release(); }
};

// Workspace of emag_csr_residual_noalloc: the residual vector (when the caller passes
// no d_r) followed by the 2 x kDotBlocks partial sums, in a file-scope device array that
// every GPU the library runs on holds until the process ends (8 MiB + 4 KiB).  Host code
// never uses the name `scratch` as an address: each device's address comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr int64_t kNoAllocMaxRows = int64_t(1) << 20;
constexpr int64_t kScratchDoubles = kNoAllocMaxRows + 2 * kDotBlocks;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The current device's scratch address (set up at the first call on that device).
double* scratch_on_current_device() {
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw HipError("device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

struct emag_assembly {
  Topology topo;
  CsrPlan csr;
  NzPlan rhs;
  std::string err;
  DevArray<int32_t> t2n, csr_src, rhs_src;
  DevArray<int64_t> csr_ptr, rhs_target, rhs_ptr;
  DevArray<double> Ke, Me, Fe;  // element workspace: 21 + 21 + 6 doubles per tet
  std::vector<int32_t> h_t2n;
  bool on_device = false;

  // plans and workspace to the current device, at the first numeric call (creating the
  // handle touches no GPU)
  void ensure_device() {
    if (on_device) return;
    const int64_t nt = topo.n_tets;
    t2n.upload(h_t2n);
    csr_ptr.upload(csr.ptr);
    csr_src.upload(csr.src);
    rhs_target.upload(rhs.target);
    rhs_ptr.upload(rhs.ptr);
    rhs_src.upload(rhs.src);
    Ke.alloc(size_t(21 * nt));
    Me.alloc(size_t(21 * nt));
    Fe.alloc(size_t(6 * nt));
    on_device = true;
  }
};

extern "C" {

emag_assembly_handle emag_assembly_create(int64_t n_nodes, int64_t n_tets, const int32_t* tet2node,
                                          int64_t n_pec, const int32_t* pec_edge_nodes) {
  if (n_nodes <= 0 || n_tets <= 0 || !tet2node || n_pec < 0 || (n_pec > 0 && !pec_edge_nodes))
    return nullptr;
  auto* h = new (std::nothrow) emag_assembly;
  if (!h) return nullptr;
  try {
    // one subdomain holding every tet: the FETI-DP classes are not used here, only the
    // edge ids, orientation signs and free numbering the topology builds
    std::vector<int32_t> sub(size_t(n_tets), 0);
    MeshInput in{n_nodes, n_tets, tet2node, sub.data(), n_pec, pec_edge_nodes, nullptr};
    const std::string e = build_topology(in, h->topo);
    if (!e.empty()) throw std::invalid_argument(e);
    build_global_plan(h->topo, h->csr, h->rhs);
    h->h_t2n.assign(tet2node, tet2node + 4 * n_tets);
  } catch (const std::exception&) {
    delete h;
    return nullptr;
  }
  return h;
}

int32_t emag_assembly_sizes(emag_assembly_handle h, int64_t* n_free, int64_t* nnz) {
  if (!h) return FETIDP_INVALID_INPUT;
  if (n_free) *n_free = h->topo.n_free;
  if (nnz) *nnz = int64_t(h->csr.col.size());
  return FETIDP_SUCCESS;
}

int32_t emag_assembly_pattern(emag_assembly_handle h, int32_t* row_ptr, int32_t* col,
                              int32_t* tet_free, int8_t* tet_sign) {
  if (!h) return FETIDP_INVALID_INPUT;
  if (row_ptr) std::memcpy(row_ptr, h->csr.row_ptr.data(), sizeof(int32_t) * h->csr.row_ptr.size());
  if (col) std::memcpy(col, h->csr.col.data(), sizeof(int32_t) * h->csr.col.size());
  if (tet_free) std::memcpy(tet_free, h->topo.tet_free.data(), sizeof(int32_t) * h->topo.tet_free.size());
  if (tet_sign) std::memcpy(tet_sign, h->topo.tet_sign.data(), h->topo.tet_sign.size());
  return FETIDP_SUCCESS;
}

int32_t emag_assemble(emag_assembly_handle h, const double* d_xyz, const int32_t* d_tet_material,
                      const double* d_nu, const double* d_kappa, const double* d_jq, double* d_K,
                      double* d_M, double* d_A, double* d_f, void* stream) {
  if (!h) return FETIDP_INVALID_INPUT;
  h->err.clear();
  if (!d_xyz || !d_tet_material || !d_nu || !d_kappa || (d_f && !d_jq)) {
    h->err = "emag_assemble: null input (f needs jq)";
    return FETIDP_INVALID_INPUT;
  }
  try {
    h->ensure_device();
  } catch (const std::exception& ex) {
    h->err = ex.what();
    return FETIDP_NUMERICAL_ERROR;
  }
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t nt = h->topo.n_tets, nnz = int64_t(h->csr.col.size());
  const double* jq = d_f ? d_jq : nullptr;
  if (d_K || d_M) {  // K_e and M_e apart (mode 0)
    launch_integrate(nt, d_xyz, h->t2n.p, d_tet_material, d_nu, d_kappa, jq, 0, h->Ke.p, h->Me.p,
                     h->Fe.p, st);
    if (d_K) launch_assemble_f64(nnz, nullptr, h->csr_ptr.p, h->csr_src.p, h->Ke.p, d_K, st);
    if (d_M) launch_assemble_f64(nnz, nullptr, h->csr_ptr.p, h->csr_src.p, h->Me.p, d_M, st);
    jq = nullptr;  // F_e is done
  }
  if (d_A || jq) {  // K_e + M_e in one pass (mode 1), as the FETI-DP factorization forms it
    launch_integrate(nt, d_xyz, h->t2n.p, d_tet_material, d_nu, d_kappa, jq, 1, h->Ke.p, nullptr,
                     h->Fe.p, st);
    if (d_A) launch_assemble_f64(nnz, nullptr, h->csr_ptr.p, h->csr_src.p, h->Ke.p, d_A, st);
  }
  if (d_f)
    launch_assemble_f64(h->rhs.size(), h->rhs_target.p, h->rhs_ptr.p, h->rhs_src.p, h->Fe.p, d_f, st);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    h->err = std::string("emag_assemble: ") + hipGetErrorString(e);
    return FETIDP_NUMERICAL_ERROR;
  }
  return FETIDP_SUCCESS;
}

const char* emag_assembly_last_error(emag_assembly_handle h) { return h ? h->err.c_str() : "null handle"; }

void emag_assembly_destroy(emag_assembly_handle h) { delete h; }

int64_t emag_tree_gauge(int64_t n_nodes, int64_t n_tets, const int32_t* tet2node, int64_t n_pec,
                        const int32_t* pec_edge_nodes, int32_t* tree_edge_nodes) {
  if (n_nodes <= 0 || n_tets <= 0 || !tet2node || n_pec < 0 || (n_pec > 0 && !pec_edge_nodes)) return -1;
  try {
    std::vector<int32_t> sub(size_t(n_tets), 0);
    Topology T;
    MeshInput in{n_nodes, n_tets, tet2node, sub.data(), n_pec, pec_edge_nodes, nullptr};
    if (!build_topology(in, T).empty()) return -1;
    // node -> free edges incidence, edges in ascending global id
    std::vector<int64_t> ptr(size_t(n_nodes) + 1, 0);
    for (int64_t e = 0; e < T.n_edges; ++e)
      if (T.edge_free[e] >= 0) {
        ++ptr[size_t(T.edge_nodes[2 * e]) + 1];
        ++ptr[size_t(T.edge_nodes[2 * e + 1]) + 1];
      }
    for (int64_t v = 0; v < n_nodes; ++v) ptr[v + 1] += ptr[v];
    std::vector<int32_t> inc(static_cast<size_t>(ptr[n_nodes]));
    {
      std::vector<int64_t> pos(ptr.begin(), ptr.end() - 1);
      for (int64_t e = 0; e < T.n_edges; ++e)
        if (T.edge_free[e] >= 0) {
          inc[size_t(pos[T.edge_nodes[2 * e]]++)] = int32_t(e);
          inc[size_t(pos[T.edge_nodes[2 * e + 1]]++)] = int32_t(e);
        }
    }
    // breadth-first search from every PEC node at once: the PEC boundary (tangential A = 0)
    // acts as one contracted root, so each other node gets exactly one tree edge; further
    // components without PEC start from their lowest node
    std::vector<uint8_t> seen(size_t(n_nodes), 0);
    std::vector<int32_t> queue;
    queue.reserve(size_t(n_nodes));
    for (int64_t p = 0; p < 2 * n_pec; ++p)
      if (!seen[size_t(pec_edge_nodes[p])]) {
        seen[size_t(pec_edge_nodes[p])] = 1;
        queue.push_back(pec_edge_nodes[p]);
      }
    std::sort(queue.begin(), queue.end());
    int64_t n_tree = 0;
    size_t head = 0;
    for (int64_t root = 0;; ) {
      while (head < queue.size()) {
        const int32_t u = queue[head++];
        for (int64_t k = ptr[u]; k < ptr[u + 1]; ++k) {
          const int32_t e = inc[size_t(k)];
          const int32_t v = T.edge_nodes[2 * e] == u ? T.edge_nodes[2 * e + 1] : T.edge_nodes[2 * e];
          if (seen[size_t(v)]) continue;
          seen[size_t(v)] = 1;
          queue.push_back(v);
          if (tree_edge_nodes) {
            tree_edge_nodes[2 * n_tree] = T.edge_nodes[2 * e];
            tree_edge_nodes[2 * n_tree + 1] = T.edge_nodes[2 * e + 1];
          }
          ++n_tree;
        }
      }
      while (root < n_nodes && (seen[size_t(root)] || ptr[root] == ptr[root + 1])) ++root;
      if (root >= n_nodes) break;
      seen[size_t(root)] = 1;
      queue.push_back(int32_t(root));
    }
    return n_tree;
  } catch (const std::exception&) {
    return -1;
  }
}

int32_t emag_csr_residual(int64_t n, const int32_t* d_row_ptr, const int32_t* d_col,
                          const double* d_val, const double* d_x, const double* d_f, double* d_r,
                          double* rel_norm, void* stream) {
  if (n <= 0 || !d_row_ptr || !d_col || !d_val || !d_x || !d_f || !rel_norm) return FETIDP_INVALID_INPUT;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  try {
    DevArray<double> r, partial;
    double* rv = d_r;
    if (!rv) {
      r.alloc(size_t(n));
      rv = r.p;
    }
    partial.alloc(2 * kDotBlocks);
    launch_csr_residual(n, d_row_ptr, d_col, d_val, d_x, d_f, rv, st);
    const double* a[2] = {rv, d_f};
    launch_dots(n, 2, a, a, partial.p, st);
    double h[2 * kDotBlocks];
    check(hipMemcpyAsync(h, partial.p, sizeof h, hipMemcpyDeviceToHost, st), "residual copy");
    check(hipStreamSynchronize(st), "residual sync");
    double rr = 0.0, ff = 0.0;
    for (int b = 0; b < kDotBlocks; ++b) {
      rr += h[b];
      ff += h[kDotBlocks + b];
    }
    *rel_norm = ff > 0.0 ? std::sqrt(rr / ff) : 0.0;
  } catch (const std::exception&) {
    return FETIDP_NUMERICAL_ERROR;
  }
  return FETIDP_SUCCESS;
}

int32_t emag_csr_residual_noalloc(int64_t n, const int32_t* d_row_ptr, const int32_t* d_col,
                                  const double* d_val, const double* d_x, const double* d_f,
                                  double* d_r, double* rel_norm, void* stream) {
  if (n <= 0 || !d_row_ptr || !d_col || !d_val || !d_x || !d_f || !rel_norm) return FETIDP_INVALID_INPUT;
  if (!d_r && n > kNoAllocMaxRows) return FETIDP_INVALID_INPUT;  // residual would not fit
  const hipStream_t st = static_cast<hipStream_t>(stream);
  try {
    double* ws = scratch_on_current_device();
    double* partial = ws + kNoAllocMaxRows;
    double* rv = d_r ? d_r : ws;
    // the same launches and host sum as emag_csr_residual: bitwise the same result; both
    // kernels overwrite every scratch element they read, so nothing needs clearing
    launch_csr_residual(n, d_row_ptr, d_col, d_val, d_x, d_f, rv, st);
    const double* a[2] = {rv, d_f};
    launch_dots(n, 2, a, a, partial, st);
    double h[2 * kDotBlocks];
    check(hipMemcpyAsync(h, partial, sizeof h, hipMemcpyDeviceToHost, st), "residual copy");
    check(hipStreamSynchronize(st), "residual sync");
    double rr = 0.0, ff = 0.0;
    for (int b = 0; b < kDotBlocks; ++b) {
      rr += h[b];
      ff += h[kDotBlocks + b];
    }
    *rel_norm = ff > 0.0 ? std::sqrt(rr / ff) : 0.0;
  } catch (const std::exception&) {
    return FETIDP_NUMERICAL_ERROR;
  }
  return FETIDP_SUCCESS;
}

}  // extern "C"

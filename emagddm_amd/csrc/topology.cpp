// Host-side mesh topology (see topology.h).
#include "topology.h"

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>

namespace emag {

namespace {
// ElementGeometry TET4 edge list (src/fe_em/element_geometry.cpp:119-126);
// same node sets, same order, as NedelecElement TET4_EDGE
// (src/numeric/nedelec_element.cpp:54-60), whose edge 2 runs {2,0}.
const int kGeomEdge[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
const int kRefEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
}  // namespace

void radix_sort_pairs(std::vector<uint64_t>& key, std::vector<uint32_t>& val, uint64_t max_key) {
  const size_t n = key.size();
  if (n < 2) return;
  int bits = 0;
  while (bits < 64 && (max_key >> bits)) ++bits;
  std::vector<uint64_t> k2(n);
  std::vector<uint32_t> v2(n);
  std::vector<size_t> cnt(1 << 16);
  for (int shift = 0; shift < bits; shift += 16) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (size_t i = 0; i < n; ++i) cnt[(key[i] >> shift) & 0xFFFF]++;
    size_t s = 0;
    for (auto& c : cnt) { size_t t = c; c = s; s += t; }
    for (size_t i = 0; i < n; ++i) {
      size_t d = cnt[(key[i] >> shift) & 0xFFFF]++;
      k2[d] = key[i];
      v2[d] = val[i];
    }
    key.swap(k2);
    val.swap(v2);
  }
}

// Local elimination order inside each subdomain's class blocks [i | D | c].  The
// tile-envelope Cholesky costs about (rows) x (front width)^2 in the interior block and,
// for the boundary rows, their span back to the first interior neighbour.  So:
//  * interior dofs in layers across the longest side of the bounding box of their edge
//    midpoints (primary key: coordinate along that side, quantized to 1e-6 of the box
//    diagonal so the layers of a structured mesh stay exact; then the middle side, then
//    the shortest side, then the free id): the front is one cross-section, the smallest;
//  * dual dofs by the position of their first interior neighbour, ties by free id, so the
//    64 rows of a tile start their envelope at about one column.  Primal dofs keep the
//    global primal order, which the coarse assembly relies on.
// On the 48^3 / 6^3 jittered-box workload this removes about a third of the interior
// phase's tile products against free-id order.
static void order_subdomain_dofs(const double* xyz, Topology& T) {
  const int64_t ns = T.n_sub, nf = T.n_free;
  std::vector<int32_t> edge_of_free(size_t(nf), -1);
  for (int64_t e = 0; e < T.n_edges; ++e)
    if (T.edge_free[e] >= 0) edge_of_free[size_t(T.edge_free[e])] = int32_t(e);
  std::vector<int32_t> lpos(size_t(nf), -1);
  for (int64_t s = 0; s < ns; ++s) {
    const SubLayout& L = T.sub[s];
    int32_t* dofs = &T.sub_dof[T.sub_dof_ptr[s]];
    const int n = L.ni + L.nd + L.nc;
    // interior: layered order
    if (L.ni > 1) {
      std::vector<std::array<double, 3>> mid(size_t(L.ni));
      double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
      for (int k = 0; k < L.ni; ++k) {
        const int32_t e = edge_of_free[size_t(dofs[k])];
        const int64_t a = T.edge_nodes[2 * e], b = T.edge_nodes[2 * e + 1];
        for (int c = 0; c < 3; ++c) {
          mid[k][c] = 0.5 * (xyz[3 * a + c] + xyz[3 * b + c]);
          lo[c] = std::min(lo[c], mid[k][c]);
          hi[c] = std::max(hi[c], mid[k][c]);
        }
      }
      int ax[3] = {0, 1, 2};  // longest side first
      std::stable_sort(ax, ax + 3, [&](int p, int q) { return hi[p] - lo[p] > hi[q] - lo[q]; });
      double diag = 0.0;
      for (int c = 0; c < 3; ++c) diag += (hi[c] - lo[c]) * (hi[c] - lo[c]);
      const double q = 1e-6 * std::max(std::sqrt(diag), 1e-300);
      std::vector<int64_t> layer(size_t(L.ni));
      for (int k = 0; k < L.ni; ++k) layer[k] = int64_t(std::llround((mid[k][ax[0]] - lo[ax[0]]) / q));
      std::vector<int32_t> perm(size_t(L.ni));
      std::iota(perm.begin(), perm.end(), 0);
      std::sort(perm.begin(), perm.end(), [&](int32_t u, int32_t v) {
        if (layer[u] != layer[v]) return layer[u] < layer[v];
        if (mid[u][ax[1]] != mid[v][ax[1]]) return mid[u][ax[1]] < mid[v][ax[1]];
        if (mid[u][ax[2]] != mid[v][ax[2]]) return mid[u][ax[2]] < mid[v][ax[2]];
        return dofs[u] < dofs[v];
      });
      std::vector<int32_t> tmp(dofs, dofs + L.ni);
      for (int k = 0; k < L.ni; ++k) dofs[k] = tmp[perm[k]];
    }
    // dual rows: position of the first interior neighbour (INT32_MAX if none)
    for (int k = 0; k < n; ++k) lpos[size_t(dofs[k])] = k;
    std::vector<int32_t> fpos(size_t(n), INT32_MAX);
    for (int64_t j = T.sub_tet_ptr[s]; j < T.sub_tet_ptr[s + 1]; ++j) {
      const int64_t t = T.sub_tet[j];
      for (int a = 0; a < 6; ++a) {
        const int32_t da = T.tet_free[6 * t + a];
        if (da < 0) continue;
        const int32_t ka = lpos[size_t(da)];
        for (int b = 0; b < 6; ++b) {
          const int32_t db = T.tet_free[6 * t + b];
          if (db < 0) continue;
          const int32_t kb = lpos[size_t(db)];
          if (kb < L.ni) fpos[size_t(ka)] = std::min(fpos[size_t(ka)], kb);
        }
      }
    }
    auto order_block = [&](int k0, int k1) {
      std::vector<int32_t> perm(size_t(k1 - k0));
      std::iota(perm.begin(), perm.end(), k0);
      std::stable_sort(perm.begin(), perm.end(), [&](int32_t u, int32_t v) {
        if (fpos[size_t(u)] != fpos[size_t(v)]) return fpos[size_t(u)] < fpos[size_t(v)];
        return dofs[u] < dofs[v];
      });
      std::vector<int32_t> tmp(dofs + k0, dofs + k1);
      for (int k = k0; k < k1; ++k) dofs[k] = tmp[size_t(perm[size_t(k - k0)] - k0)];
    };
    order_block(L.ni, L.ni + L.nd);  // primal rows stay in global primal order (coarse assembly)
    for (int k = 0; k < n; ++k) lpos[size_t(dofs[k])] = -1;
  }
}

std::string build_topology(const MeshInput& in, Topology& T) {
  if (in.n_nodes <= 0 || in.n_tets <= 0) return "empty mesh";
  const int64_t nt = in.n_tets, nn = in.n_nodes;
  if (nt * 6 >= (int64_t(1) << 31)) return "too many tets";
  T = Topology();
  T.n_nodes = nn;
  T.n_tets = nt;
  int32_t max_sub = -1;
  for (int64_t t = 0; t < nt; ++t) {
    for (int a = 0; a < 4; ++a) {
      int32_t v = in.tet2node[4 * t + a];
      if (v < 0 || v >= nn) return "tet2node out of range at tet " + std::to_string(t);
    }
    int32_t s = in.tet_subdomain[t];
    if (s < 0) return "negative subdomain id at tet " + std::to_string(t);
    max_sub = std::max(max_sub, s);
  }
  T.n_sub = max_sub + 1;
  {
    std::vector<char> used(size_t(T.n_sub), 0);
    for (int64_t t = 0; t < nt; ++t) used[size_t(in.tet_subdomain[t])] = 1;
    for (int64_t s = 0; s < T.n_sub; ++s)
      if (!used[size_t(s)])
        return "subdomain ids must be 0..n_sub-1 with no empty subdomain (id " + std::to_string(s) +
               " has no tet)";
  }

  // ---- global edge ids in first-appearance order --------------------------
  const size_t N = size_t(nt) * 6;
  std::vector<uint64_t> key(N);
  std::vector<uint32_t> idx(N);
  for (int64_t t = 0; t < nt; ++t)
    for (int e = 0; e < 6; ++e) {
      uint64_t a = in.tet2node[4 * t + kGeomEdge[e][0]], b = in.tet2node[4 * t + kGeomEdge[e][1]];
      if (a == b) return "degenerate tet " + std::to_string(t);
      uint64_t lo = std::min(a, b), hi = std::max(a, b);
      key[6 * t + e] = lo * uint64_t(nn) + hi;
      idx[6 * t + e] = uint32_t(6 * t + e);
    }
  radix_sort_pairs(key, idx, uint64_t(nn) * uint64_t(nn));
  // group ids in key order; first appearance = idx of the group's first element (stable sort)
  std::vector<int32_t> group_of(N);
  std::vector<uint32_t> first;
  std::vector<uint64_t> gkey;
  for (size_t i = 0; i < N; ++i) {
    if (i == 0 || key[i] != key[i - 1]) {
      first.push_back(idx[i]);
      gkey.push_back(key[i]);
    }
    group_of[idx[i]] = int32_t(first.size() - 1);
  }
  const int64_t G = int64_t(first.size());
  std::vector<int32_t> id_of_group(G, -1);
  {
    std::vector<int32_t> group_at(N, -1);
    for (int64_t g = 0; g < G; ++g) group_at[first[g]] = int32_t(g);
    int32_t next = 0;
    for (size_t i = 0; i < N; ++i)
      if (group_at[i] >= 0) id_of_group[group_at[i]] = next++;
  }
  T.n_edges = G;
  T.tet_edge.resize(N);
  T.edge_nodes.resize(2 * G);
  for (size_t i = 0; i < N; ++i) T.tet_edge[i] = id_of_group[group_of[i]];
  for (int64_t g = 0; g < G; ++g) {
    int32_t id = id_of_group[g];
    T.edge_nodes[2 * id] = int32_t(gkey[g] / uint64_t(nn));
    T.edge_nodes[2 * id + 1] = int32_t(gkey[g] % uint64_t(nn));
  }
  // orientation sign of the reference function of local edge e = (a,b):
  // circulation +1 from b to a; global edges run low -> high node id.
  T.tet_sign.resize(N);
  for (int64_t t = 0; t < nt; ++t)
    for (int e = 0; e < 6; ++e) {
      int32_t ga = in.tet2node[4 * t + kRefEdge[e][0]], gb = in.tet2node[4 * t + kRefEdge[e][1]];
      T.tet_sign[6 * t + e] = gb < ga ? 1 : -1;
    }

  // ---- PEC elimination and free numbering -----------------------------------
  std::vector<uint64_t> pkey(in.n_pec);
  for (int64_t p = 0; p < in.n_pec; ++p) {
    int64_t a = in.pec_pairs[2 * p], b = in.pec_pairs[2 * p + 1];
    if (a < 0 || b < 0 || a >= nn || b >= nn) return "PEC pair out of range";
    pkey[p] = uint64_t(std::min(a, b)) * uint64_t(nn) + uint64_t(std::max(a, b));
  }
  std::sort(pkey.begin(), pkey.end());
  T.edge_free.assign(G, -1);
  {
    std::vector<uint8_t> constrained(G, 0);
    for (int64_t g = 0; g < G; ++g)
      if (std::binary_search(pkey.begin(), pkey.end(), gkey[g])) constrained[id_of_group[g]] = 1;
    int32_t nf = 0;
    for (int64_t e = 0; e < G; ++e)
      if (!constrained[e]) T.edge_free[e] = nf++;
    T.n_free = nf;
  }
  T.tet_free.resize(N);
  for (size_t i = 0; i < N; ++i) T.tet_free[i] = T.edge_free[T.tet_edge[i]];

  // ---- subdomain incidence, multiplicity, classes ----------------------------
  const int64_t nf = T.n_free, ns = T.n_sub;
  std::vector<uint64_t> dk;
  std::vector<uint32_t> dv;
  dk.reserve(N);
  for (int64_t t = 0; t < nt; ++t)
    for (int e = 0; e < 6; ++e) {
      int32_t d = T.tet_free[6 * t + e];
      if (d >= 0) dk.push_back(uint64_t(d) * uint64_t(ns) + uint64_t(in.tet_subdomain[t]));
    }
  dv.assign(dk.size(), 0);
  radix_sort_pairs(dk, dv, uint64_t(nf) * uint64_t(ns));
  dk.erase(std::unique(dk.begin(), dk.end()), dk.end());
  T.free_mult.assign(nf, 0);
  for (uint64_t k : dk) T.free_mult[k / ns]++;
  T.lam_of_free.assign(nf, -1);
  T.c_of_free.assign(nf, -1);
  for (int64_t d = 0; d < nf; ++d) {
    if (T.free_mult[d] == 0) return "free edge touched by no tet";
    if (T.free_mult[d] == 2) { T.lam_of_free[d] = int32_t(T.free_of_lam.size()); T.free_of_lam.push_back(int32_t(d)); }
    if (T.free_mult[d] >= 3) { T.c_of_free[d] = int32_t(T.free_of_c.size()); T.free_of_c.push_back(int32_t(d)); }
  }
  T.n_lam = int64_t(T.free_of_lam.size());
  T.n_c = int64_t(T.free_of_c.size());

  // per-subdomain dof lists: counting sort of (d, s) pairs by s keeps d ascending
  std::vector<int64_t> cnt(ns + 1, 0);
  for (uint64_t k : dk) cnt[k % ns + 1]++;
  for (int64_t s = 0; s < ns; ++s) cnt[s + 1] += cnt[s];
  std::vector<int32_t> sd(dk.size());
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (uint64_t k : dk) sd[pos[k % ns]++] = int32_t(k / ns);
  }
  T.sub.resize(ns);
  T.sub_dof_ptr.assign(ns + 1, 0);
  T.sub_dof.resize(sd.size());
  for (int64_t s = 0; s < ns; ++s) {
    SubLayout& L = T.sub[s];
    int64_t w = cnt[s];
    for (int cls = 0; cls < 3; ++cls)
      for (int64_t j = cnt[s]; j < cnt[s + 1]; ++j) {
        int32_t d = sd[j], m = T.free_mult[d];
        int c = m == 1 ? 0 : (m == 2 ? 1 : 2);
        if (c == cls) T.sub_dof[w++] = d;
      }
    for (int64_t j = cnt[s]; j < cnt[s + 1]; ++j) {
      int32_t m = T.free_mult[sd[j]];
      if (m == 1) L.ni++; else if (m == 2) L.nd++; else L.nc++;
    }
    T.sub_dof_ptr[s + 1] = cnt[s + 1];
  }
  // tets per subdomain
  T.sub_tet_ptr.assign(ns + 1, 0);
  for (int64_t t = 0; t < nt; ++t) T.sub_tet_ptr[in.tet_subdomain[t] + 1]++;
  for (int64_t s = 0; s < ns; ++s) T.sub_tet_ptr[s + 1] += T.sub_tet_ptr[s];
  T.sub_tet.resize(nt);
  {
    std::vector<int64_t> pos(T.sub_tet_ptr.begin(), T.sub_tet_ptr.end() - 1);
    for (int64_t t = 0; t < nt; ++t) T.sub_tet[pos[in.tet_subdomain[t]]++] = int32_t(t);
  }
  if (in.xyz) order_subdomain_dofs(in.xyz, T);

  // ---- padded layouts and offsets ------------------------------------------
  int64_t pool = 0, vec = 0, dvec = 0, wel = 0, pel = 0, yel = 0, cvec = 0, spel = 0;
  for (int64_t s = 0; s < ns; ++s) {
    SubLayout& L = T.sub[s];
    L.nip = int32_t(round_up(L.ni, kTile));
    L.ndp = int32_t(round_up(L.nd, kTile));
    L.ncp = int32_t(round_up(L.nc, kTile));
    L.o_d = L.nip;
    L.o_c = L.o_d + L.ndp;
    L.o_a = L.o_c + L.ncp;
    L.m = L.o_a + L.ndp;
    L.T = L.m / kTile;
    L.Ti = L.nip / kTile;
    L.Tr = L.o_c / kTile;
    L.Tc = L.o_a / kTile;
    L.pool_off = pool;
    pool += int64_t(L.T) * (L.T + 1) / 2 * kTile * kTile;
    L.vec_off = vec;
    vec += L.m;
    L.dvec_off = dvec;
    dvec += L.ndp;
    L.w_off = wel;
    wel += sym_tile_elems(L.ndp) + int64_t(L.ncp) * L.ndp;
    L.p_off = pel;
    pel += sym_tile_elems(L.ndp);
    L.sp_off = spel;
    spel += int64_t(L.ndp / kTile) * (L.ndp / kTile) * kTile;
    L.y_off = yel;
    yel += L.ndp + L.ncp;
    L.cvec_off = cvec;
    cvec += L.ncp;
  }
  T.pool_elems = pool;
  T.vec_elems = vec;
  T.dvec_elems = dvec;
  T.w_elems = wel;
  T.p_elems = pel;
  T.sp_elems = spel;
  T.y_elems = yel;
  T.cvec_elems = cvec;

  // ---- dual slots, primal slots ---------------------------------------------
  T.d_lam.assign(dvec, -1);
  T.d_sgn.assign(dvec, 0.0);
  T.lam_pos.assign(2 * T.n_lam, -1);
  T.c_loc.assign(cvec, -1);
  std::vector<int32_t> lam_seen(T.n_lam, 0);
  for (int64_t s = 0; s < ns; ++s) {
    const SubLayout& L = T.sub[s];
    const int32_t* dofs = &T.sub_dof[T.sub_dof_ptr[s]];
    for (int k = 0; k < L.nd; ++k) {
      int32_t lam = T.lam_of_free[dofs[L.ni + k]];
      int64_t slot = L.dvec_off + k;
      T.d_lam[slot] = lam;
      int side = lam_seen[lam]++;  // subdomains visited in ascending order
      if (side > 1) return "dual edge with more than two subdomains";
      T.d_sgn[slot] = side == 0 ? 1.0 : -1.0;
      T.lam_pos[2 * lam + side] = slot;
    }
    for (int k = 0; k < L.nc; ++k) T.c_loc[L.cvec_off + k] = T.c_of_free[dofs[L.ni + L.nd + k]];
  }
  for (int64_t l = 0; l < T.n_lam; ++l)
    if (lam_seen[l] != 2) return "dual edge without two sides";
  T.c_gather_ptr.assign(T.n_c + 1, 0);
  for (int64_t j = 0; j < cvec; ++j)
    if (T.c_loc[j] >= 0) T.c_gather_ptr[T.c_loc[j] + 1]++;
  for (int64_t c = 0; c < T.n_c; ++c) T.c_gather_ptr[c + 1] += T.c_gather_ptr[c];
  T.c_gather_slot.resize(T.c_gather_ptr[T.n_c]);
  {
    std::vector<int64_t> pos(T.c_gather_ptr.begin(), T.c_gather_ptr.end() - 1);
    for (int64_t j = 0; j < cvec; ++j)
      if (T.c_loc[j] >= 0) T.c_gather_slot[pos[T.c_loc[j]]++] = j;
  }

  // ---- global vector recovery map ------------------------------------------
  T.free_kind.assign(nf, -1);
  T.free_src_a.assign(nf, -1);
  T.free_src_b.assign(nf, -1);
  for (int64_t s = 0; s < ns; ++s) {
    const SubLayout& L = T.sub[s];
    const int32_t* dofs = &T.sub_dof[T.sub_dof_ptr[s]];
    for (int k = 0; k < L.ni; ++k) {
      T.free_kind[dofs[k]] = 0;
      T.free_src_a[dofs[k]] = L.vec_off + k;
    }
  }
  for (int64_t l = 0; l < T.n_lam; ++l) {
    int32_t d = T.free_of_lam[l];
    T.free_kind[d] = 1;
    T.free_src_a[d] = l;
  }
  for (int64_t c = 0; c < T.n_c; ++c) {
    int32_t d = T.free_of_c[c];
    T.free_kind[d] = 2;
    T.free_src_a[d] = c;
  }
  for (int64_t d = 0; d < nf; ++d)
    if (T.free_kind[d] < 0) return "unclassified free dof";
  return "";
}

}  // namespace emag

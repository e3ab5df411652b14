// Assembly plans and the batched factorization schedule (see plan.h).
#include "plan.h"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace emag {

int packed_index(int a, int b) {
  if (a > b) std::swap(a, b);
  int k = 0;
  for (int i = 0; i < a; ++i) k += 6 - i;
  return k + (b - a);
}

namespace {

// Group emissions (target, code) by target, keeping emission order inside a group.
void group_emissions(std::vector<uint64_t>& tgt, const std::vector<int32_t>& code, uint64_t max_t,
                     std::vector<int64_t>& target, std::vector<int64_t>& ptr,
                     std::vector<int32_t>& src) {
  std::vector<uint32_t> idx(tgt.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = uint32_t(i);
  radix_sort_pairs(tgt, idx, max_t);
  target.clear();
  ptr.clear();
  src.resize(tgt.size());
  for (size_t i = 0; i < tgt.size(); ++i) {
    if (i == 0 || tgt[i] != tgt[i - 1]) {
      target.push_back(int64_t(tgt[i]));
      ptr.push_back(int64_t(i));
    }
    src[i] = code[idx[i]];
  }
  ptr.push_back(int64_t(tgt.size()));
}

int32_t enc(int64_t elem_index, int sign) {
  return int32_t((elem_index << 1) | (sign < 0 ? 1 : 0));
}

void add_tiles(TileSchedule& S, int mode, const std::vector<TileWork>& list) {
  if (list.empty()) return;
  TileLaunch L{mode, int64_t(S.items.size()), int64_t(list.size())};
  S.items.insert(S.items.end(), list.begin(), list.end());
  S.launches.push_back(L);
  // flags & 16 marks a factoring item, so that it can also ride at the front of the
  // column-step launch that follows it
  if (mode == 3)
    for (size_t i = size_t(L.off); i < S.items.size(); ++i) S.items[i].flags |= 16;
  // modes 1 and 3 add the TRSM product (by the lower-triangular stored inverse: 160 of a
  // tile product's 256 MFMAs are issued) and the diagonal-tile share when flagged, with
  // flags & 64 also the diagonal's K = I - 2 term; lookahead items (flags & 2) are
  // update-only
  for (const TileWork& w : list) {
    const double extra = (mode == 1 || mode == 3) && !(w.flags & 2)
                             ? 0.625 + (w.flags & 1) + ((w.flags & 64) ? 1.0 : 0.0)
                             : 0.0;
    S.flops += 2.0 * kTile * kTile * kTile * (double(w.k1 - w.k0) + extra);
  }
}

// Longest K loops first (stable): a column launch ends with its last work-groups, so the
// long items go out early and the short ones fill the slots behind them.
void longest_first(std::vector<TileWork>& list) {
  std::stable_sort(list.begin(), list.end(), [](const TileWork& a, const TileWork& b) {
    return a.k1 - a.k0 > b.k1 - b.k0;
  });
}

// Work-groups of a launch go to the GPU's eight XCDs in turn (work-group b to XCD b % 8),
// and each XCD has its own L2.  Items that read the same panels (one matrix's tiles of
// one column J, or one matrix's Schur tiles, key = (mat, J) or mat) are therefore placed
// at positions of one residue mod 8, consecutively, so they run on one XCD at about the
// same time and the shared panel is fetched from HBM once per group instead of once per
// item.  Groups are dealt to the eight residues longest-total-work first, each to the
// residue with the least work so far; within a residue the groups keep longest-first
// order (by their longest item), and so do the items of a group.  Positions past the
// shortest residue's length hold the longer residues' leftovers in the same round-robin
// order.  Only the order changes: the arithmetic of every item is that of any other order.
void xcd_grouped(std::vector<TileWork>& list, bool by_matrix) {
  if (list.size() < 16) {
    longest_first(list);
    return;
  }
  longest_first(list);
  std::vector<int64_t> key(list.size());
  for (size_t i = 0; i < list.size(); ++i)
    key[i] = by_matrix ? int64_t(list[i].mat) : (int64_t(list[i].mat) << 20) | list[i].J;
  // groups in order of their longest item (the list is longest first)
  std::vector<int64_t> order(list.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = int64_t(i);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
  struct Group {
    std::vector<int32_t> idx;
    int first;      // position of its longest item in the longest-first list
    double work;
  };
  std::vector<Group> groups;
  for (size_t i = 0; i < order.size(); ++i) {
    const int64_t it = order[i];
    if (i == 0 || key[it] != key[order[i - 1]]) groups.push_back(Group{{}, int(it), 0.0});
    Group& g = groups.back();
    g.idx.push_back(int32_t(it));
    g.first = std::min(g.first, int(it));
    g.work += double(list[it].k1 - list[it].k0) + 2.0;
  }
  for (Group& g : groups) std::sort(g.idx.begin(), g.idx.end());  // longest first inside
  std::vector<int32_t> by_work(groups.size());
  for (size_t g = 0; g < groups.size(); ++g) by_work[g] = int32_t(g);
  std::stable_sort(by_work.begin(), by_work.end(),
                   [&](int32_t a, int32_t b) { return groups[a].work > groups[b].work; });
  constexpr int kXcd = 8;
  std::vector<double> load(kXcd, 0.0);
  std::vector<std::vector<int32_t>> res_groups(kXcd);
  for (int32_t g : by_work) {
    const int x = int(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += groups[g].work;
    res_groups[x].push_back(g);
  }
  std::vector<std::vector<int32_t>> queue(kXcd);
  for (int x = 0; x < kXcd; ++x) {
    std::sort(res_groups[x].begin(), res_groups[x].end(),
              [&](int32_t a, int32_t b) { return groups[a].first < groups[b].first; });
    for (int32_t g : res_groups[x])
      queue[x].insert(queue[x].end(), groups[g].idx.begin(), groups[g].idx.end());
  }
  std::vector<TileWork> out;
  out.reserve(list.size());
  std::vector<size_t> pos(kXcd, 0);
  while (out.size() < list.size())
    for (int x = 0; x < kXcd; ++x)
      if (pos[x] < queue[x].size()) out.push_back(list[size_t(queue[x][pos[x]++])]);
  list.swap(out);
}

}  // namespace

// Incremental diagonal (inc = true): the column-step workgroups (mode 1, flags & 1)
// also apply their share to the diagonal tile of their row, so diagonal items
// carry an empty K range and Schur lists skip the diagonal -- short critical path.
// Left-looking diagonal (inc = false): one rounding per diagonal entry and Schur
// step, used for FP32 factors where repeated rounding of the diagonal matters.
// With incremental diagonals, the item (j + 1, j) of column j gives C_{j+1,j+1}
// its last share, so it runs as a mode-3 item that also factors that tile for the
// next column (the diagonal launch of column j + 1 disappears).
// Per column: [mode 2][mode 3][mode 1].
// One matrix alone has few items per column, and the column launches are a chain
// whose links are the longest item of each.  So with incremental diagonals every
// tile's K range [k0, E) (E = J for a column-step tile, E = Tp for a Schur tile) is
// cut into pieces: its own item applies only the last p products (p = 1 for
// column-step tiles, 1 + (I + J) % kChunk for Schur tiles, which spreads their
// pieces over the launches), and update-only items (flags 2) of earlier column
// launches apply the rest in chunks of at most kChunk products, the chunk that ends
// at K = e riding in the launch of column e (columns < e are final there).  No item
// of a column launch loops over more than kChunk products.
TileSchedule single_matrix_schedule(int Tp, int Ta, bool inc, const int32_t* ft) {
  constexpr int kChunk = 8;
  TileSchedule S;
  auto f = [&](int I) { return ft ? ft[I] : 0; };
  auto k0_of = [&](int I, int J) { return std::max(f(I), f(J)); };
  // start of the part of tile (I, J)'s K range its own item applies
  auto tail = [&](int I, int J) {
    const int E = std::min(J, Tp), p = J < Tp ? 1 : 1 + (I + J) % kChunk;
    return std::max(k0_of(I, J), E - p);
  };
  // chunks of tile (I, J) end at tail, tail - kChunk, ...: the chunk ending at e
  // (e > k0) is [max(k0, e - kChunk), e) and rides in the launch of column e
  auto chunk_in = [&](int I, int J, int e, TileWork& w) {
    const int k0 = k0_of(I, J), t = tail(I, J);
    if (e <= k0 || e > t || (t - e) % kChunk) return false;
    w = TileWork{0, I, J, std::max(k0, e - kChunk), e, 2};
    return true;
  };
  for (int j = 0; j < Tp; ++j) {
    if (!(inc && j > 0 && f(j) <= j - 1))  // else factored by column j-1's mode-3 item
      add_tiles(S, 2, {TileWork{0, j, j, inc ? 0 : f(j), inc ? 0 : j, 0}});
    std::vector<TileWork> next, below;
    for (int I = j + 1; I < Ta; ++I)
      if (f(I) <= j) {
        TileWork w{0, I, j, inc ? tail(I, j) : k0_of(I, j), j, inc ? 1 : 0};
        (inc && I == j + 1 && I < Tp ? next : below).push_back(w);
      }
    if (inc)
      for (int J = j + 1; J < Ta; ++J)
        for (int I = J + 1; I < Ta; ++I) {
          TileWork w;
          if (f(I) < j && f(J) < j && chunk_in(I, J, j, w)) below.push_back(w);
        }
    add_tiles(S, 3, next);
    longest_first(below);
    add_tiles(S, 1, below);
  }
  std::vector<TileWork> schur;
  for (int J = Tp; J < Ta; ++J)
    for (int I = inc ? J + 1 : J; I < Ta; ++I) {
      const int k0 = inc ? tail(I, J) : k0_of(I, J);
      if (k0 < Tp) schur.push_back(TileWork{0, I, J, k0, Tp, 0});
    }
  longest_first(schur);
  add_tiles(S, 0, schur);
  return S;
}

TileSchedule batched_dense_schedule(const std::vector<int>& Tp) {
  TileSchedule S;
  int maxTp = 0;
  for (int t : Tp) maxTp = std::max(maxTp, t);
  for (int j = 0; j < maxTp; ++j) {
    std::vector<TileWork> diag, next, below;
    for (int m = 0; m < int(Tp.size()); ++m) {
      const int tp = Tp[m], ta = 2 * tp;
      if (j >= tp) continue;
      if (j == 0) diag.push_back(TileWork{m, 0, 0, 0, 0, 0});
      for (int I = j + 1; I < ta; ++I) {
        const TileWork w{m, I, j, 0, j, 1};
        (I == j + 1 && I < tp ? next : below).push_back(w);
      }
    }
    add_tiles(S, 2, diag);
    add_tiles(S, 3, next);
    longest_first(below);
    add_tiles(S, 1, below);
  }
  std::vector<TileWork> schur;
  for (int m = 0; m < int(Tp.size()); ++m) {
    const int tp = Tp[m], ta = 2 * tp;
    for (int J = tp; J < ta; ++J)
      for (int I = J + 1; I < ta; ++I) schur.push_back(TileWork{m, I, J, 0, tp, 0});
  }
  add_tiles(S, 0, schur);
  return S;
}

namespace {
// Faces of the partition and their dense factorization plan (see FacePlan).
void build_faces(const Topology& T, Plans& P) {
  FacePlan& F = P.faces;
  F = FacePlan();
  const int64_t nl = T.n_lam;
  std::vector<int32_t> slot_sub(T.dvec_elems, -1);
  for (int64_t s = 0; s < T.n_sub; ++s)
    for (int k = 0; k < T.sub[s].ndp; ++k) slot_sub[T.sub[s].dvec_off + k] = int32_t(s);
  // multipliers sorted by (s0, s1, lambda)
  std::vector<std::pair<std::pair<int32_t, int32_t>, int32_t>> key(nl);
  for (int64_t l = 0; l < nl; ++l)
    key[l] = {{slot_sub[T.lam_pos[2 * l]], slot_sub[T.lam_pos[2 * l + 1]]}, int32_t(l)};
  std::sort(key.begin(), key.end());
  std::vector<int> tp_list;
  int64_t vec = 0, op = 0, pool = 0, inv = 0, dd = 0;
  for (int64_t a = 0; a < nl;) {
    int64_t b = a;
    while (b < nl && key[b].first == key[a].first) ++b;
    FaceDev d{};
    d.s0 = key[a].first.first;
    d.s1 = key[a].first.second;
    d.nf = int32_t(b - a);
    const int nfp = int(round_up(d.nf, kTile));
    d.tf = nfp / kTile;
    d.vec_off = vec;
    d.pool_base = pool;
    d.ainv_off = op;
    d.b_off = op + sym_tile_elems(nfp);
    d.d_off = dd;
    for (int i = 0; i < nfp; ++i) {
      const bool real = i < d.nf;
      const int32_t l = real ? key[a + i].second : -1;
      F.map_lam.push_back(l);
      F.map_self.push_back(real ? int32_t(vec + i) : -1);
      F.slot0.push_back(real ? T.lam_pos[2 * l] : -1);
      F.slot1.push_back(real ? T.lam_pos[2 * l + 1] : -1);
    }
    // [[A, I], [I, 0]]: 2 tf tile rows, dense lower tiles, tile (I, J) at base + (I(I+1)/2 + J) 4096
    const int ta = 2 * d.tf;
    F.rowbase.push_back(int64_t(F.rowabs.size()));
    for (int I = 0; I < ta; ++I) F.rowabs.push_back(pool + tile_offset(I, 0));
    F.inv_off.push_back(inv);
    vec += nfp;
    op += 2 * sym_tile_elems(nfp);
    pool += int64_t(ta) * (ta + 1) / 2 * kTileElems;
    inv += int64_t(d.tf) * kTileElems;
    dd += int64_t(d.tf) * d.tf * kTileElems;
    F.max_T = std::max(F.max_T, d.tf);
    tp_list.push_back(d.tf);
    F.dev.push_back(d);
    a = b;
  }
  F.n = int64_t(F.dev.size());
  F.vec_elems = vec;
  F.op_elems = op;
  F.pool_elems = pool;
  F.inv_elems = inv;
  F.d_elems = dd;
  F.sched = batched_dense_schedule(tp_list);
}
}  // namespace

// Global free-dof CSR (A = K + M, duplicates summed in a fixed order) and RHS plans:
// every (tet, local entry) with both dofs free contributes s_a s_b elem to its entry.
void build_global_plan(const Topology& T, CsrPlan& csr, NzPlan& rhs) {
  const int64_t nt = T.n_tets, nf = T.n_free;
  {
    std::vector<uint64_t> tgt;
    std::vector<int32_t> code;
    tgt.reserve(size_t(nt) * 36);
    code.reserve(size_t(nt) * 36);
    for (int64_t t = 0; t < nt; ++t)
      for (int a = 0; a < 6; ++a) {
        const int32_t da = T.tet_free[6 * t + a];
        if (da < 0) continue;
        for (int b = 0; b < 6; ++b) {
          const int32_t db = T.tet_free[6 * t + b];
          if (db < 0) continue;
          tgt.push_back(uint64_t(da) * uint64_t(nf) + uint64_t(db));
          code.push_back(enc(int64_t(packed_index(a, b)) * nt + t, T.tet_sign[6 * t + a] * T.tet_sign[6 * t + b]));
        }
      }
    std::vector<int64_t> keys;
    group_emissions(tgt, code, uint64_t(nf) * uint64_t(nf), keys, csr.ptr, csr.src);
    csr.row_ptr.assign(nf + 1, 0);
    csr.col.resize(keys.size());
    for (size_t i = 0; i < keys.size(); ++i) {
      csr.row_ptr[keys[i] / nf + 1]++;
      csr.col[i] = int32_t(keys[i] % nf);
    }
    for (int64_t r = 0; r < nf; ++r) csr.row_ptr[r + 1] += csr.row_ptr[r];
    if (keys.size() >= (size_t(1) << 31)) throw std::runtime_error("global CSR too large for int32");

    tgt.clear();
    code.clear();
    for (int64_t t = 0; t < nt; ++t)
      for (int a = 0; a < 6; ++a) {
        const int32_t da = T.tet_free[6 * t + a];
        if (da < 0) continue;
        tgt.push_back(uint64_t(da));
        code.push_back(enc(int64_t(a) * nt + t, T.tet_sign[6 * t + a]));
      }
    group_emissions(tgt, code, uint64_t(nf), rhs.target, rhs.ptr, rhs.src);
  }
}

// Grouped coarse solve: S~ = L L^T (envelope); L^-1 v by a forward sweep over groups of
// tile columns g = [a, b): t_g = v_g - L_{g, <a} y (panel products), y_g = X_g t_g with
// X_g = (L_gg)^-1 explicit; L^-T y by the backward sweep with the panels below the groups
// and X_g^T.
void build_coarse_groups(Plans& P, int G) {
  const int tn = int(P.coarse_ncp / kTile);
  const int32_t* ft = P.coarse_ftile.data();
  P.coarse_grouped = true;
  P.coarse_group = G;
  P.cg_start.clear();
  for (int a = 0; a < tn; a += G) P.cg_start.push_back(a);
  P.cg_start.push_back(tn);
  const int ng = int(P.cg_start.size()) - 1;
  P.cg_xoff.assign(size_t(ng), 0);
  int64_t xo = 0;
  for (int g = 0; g < ng; ++g) {
    P.cg_xoff[g] = xo;
    const int64_t n = int64_t(P.cg_start[g + 1] - P.cg_start[g]) * kTile;
    xo += n * n;
  }
  P.cg_x_elems = std::max<int64_t>(xo, 1);
  P.cg_group_of.assign(size_t(tn), 0);
  for (int g = 0; g < ng; ++g)
    for (int i = P.cg_start[g]; i < P.cg_start[g + 1]; ++i) P.cg_group_of[i] = g;
  // forward panels, ordered by group, row I, column J: the slots of row I are contiguous
  P.cg_fwd_tiles.clear();
  P.cg_fwd_off.assign(size_t(ng) + 1, 0);
  P.cg_fwd_rowptr.assign(size_t(tn) + 1, 0);
  for (int g = 0; g < ng; ++g) {
    P.cg_fwd_off[g] = int64_t(P.cg_fwd_tiles.size() / 2);
    for (int I = P.cg_start[g]; I < P.cg_start[g + 1]; ++I) {
      P.cg_fwd_rowptr[I] = int32_t(P.cg_fwd_tiles.size() / 2);
      for (int J = ft[I]; J < P.cg_start[g]; ++J) {
        P.cg_fwd_tiles.push_back(I);
        P.cg_fwd_tiles.push_back(J);
      }
    }
  }
  P.cg_fwd_off[ng] = int64_t(P.cg_fwd_tiles.size() / 2);
  P.cg_fwd_rowptr[tn] = int32_t(P.cg_fwd_tiles.size() / 2);
  // backward panels, ordered by group, column J, row I: the slots of column J are contiguous
  P.cg_bwd_tiles.clear();
  P.cg_bwd_off.assign(size_t(ng) + 1, 0);
  P.cg_bwd_colptr.assign(size_t(tn) + 1, 0);
  for (int g = 0; g < ng; ++g) {
    P.cg_bwd_off[g] = int64_t(P.cg_bwd_tiles.size() / 2);
    for (int J = P.cg_start[g]; J < P.cg_start[g + 1]; ++J) {
      P.cg_bwd_colptr[J] = int32_t(P.cg_bwd_tiles.size() / 2);
      for (int I = P.cg_start[g + 1]; I < tn; ++I)
        if (ft[I] <= J) {
          P.cg_bwd_tiles.push_back(I);
          P.cg_bwd_tiles.push_back(J);
        }
    }
  }
  P.cg_bwd_off[ng] = int64_t(P.cg_bwd_tiles.size() / 2);
  P.cg_bwd_colptr[tn] = int32_t(P.cg_bwd_tiles.size() / 2);
  // items forming X_g by distance d = i - j from the diagonal (d = 0: X_ii = L_ii^-1)
  P.cg_inv_items.clear();
  P.cg_inv_step.assign(size_t(G) + 1, 0);
  for (int d = 0; d < G; ++d) {
    P.cg_inv_step[d] = int64_t(P.cg_inv_items.size() / 2);
    for (int g = 0; g < ng; ++g)
      for (int j = P.cg_start[g]; j + d < P.cg_start[g + 1]; ++j) {
        P.cg_inv_items.push_back(j + d);
        P.cg_inv_items.push_back(j);
      }
  }
  P.cg_inv_step[G] = int64_t(P.cg_inv_items.size() / 2);
}

// Coarse pool, coarse assembly plan and coarse schedule.  group_tiles = 0: the pool is
// [[S~, I], [I, 0]] and its partial factorization leaves -S~^-1 (dense) as lower tiles;
// group_tiles > 0: the pool is S~ alone, factored completely, and the solve sweeps over
// groups of group_tiles tile columns (build_coarse_groups).
void build_coarse_plan(const Topology& T, Plans& P, int group_tiles) {
  const int64_t ns = T.n_sub;
  P.coarse = PoolPlan();
  P.coarse_grouped = false;
  P.coarse_group = 0;
  {
    const int64_t nc = T.n_c;
    const int64_t ncp = round_up(std::max<int64_t>(nc, 1), kTile);
    P.coarse_ncp = ncp;
    const bool grouped = group_tiles > 0;
    const int64_t Tc = (grouped ? 1 : 2) * ncp / kTile;
    std::vector<int64_t> cfrow(size_t(Tc * kTile));
    for (int64_t r = 0; r < Tc * kTile; ++r) cfrow[r] = r;
    if (!grouped)
      for (int64_t k = 0; k < nc; ++k) cfrow[ncp + k] = k;
    for (int64_t s = 0; s < ns; ++s) {
      const SubLayout& L = T.sub[s];
      for (int ka = 0; ka < L.nc; ++ka) {
        const int64_t Ca = T.c_loc[L.cvec_off + ka];
        for (int kb = 0; kb <= ka; ++kb) {
          const int64_t Cb = T.c_loc[L.cvec_off + kb];
          if (Ca < Cb) throw std::runtime_error("primal order not monotone");
          cfrow[Ca] = std::min(cfrow[Ca], Cb);
        }
      }
    }
    P.coarse_ftile.assign(Tc, 0);
    P.coarse_rowabs.assign(Tc, 0);
    int64_t off = 0;
    for (int64_t I = 0; I < Tc; ++I) {
      int64_t f = I;
      for (int64_t r = I * kTile; r < (I + 1) * kTile; ++r) f = std::min(f, cfrow[r] / kTile);
      P.coarse_ftile[I] = int32_t(f);
      P.coarse_rowabs[I] = off - f * kTileElems;
      off += (I - f + 1) * kTileElems;
    }
    P.coarse_pool_elems = off;
    const int64_t* cra = P.coarse_rowabs.data();
    std::vector<uint64_t> tgt;
    std::vector<int64_t> srcpos;
    for (int64_t s = 0; s < ns; ++s) {
      const SubLayout& L = T.sub[s];
      const int64_t* ra = &P.rowabs[P.ftile_off[s]];
      for (int ka = 0; ka < L.nc; ++ka) {
        const int64_t Ca = T.c_loc[L.cvec_off + ka];
        for (int kb = 0; kb <= ka; ++kb) {
          const int64_t Cb = T.c_loc[L.cvec_off + kb];
          const int64_t sp = env_pos(ra, L.o_c + ka, L.o_c + kb);
          tgt.push_back(uint64_t(env_pos(cra, Ca, Cb)));
          srcpos.push_back(sp);
          if (Ca != Cb && Ca / kTile == Cb / kTile) {
            tgt.push_back(uint64_t(cra[Ca / kTile] + (Ca / kTile) * kTileElems + (Ca % kTile) * kTile + Cb % kTile));
            srcpos.push_back(sp);
          }
        }
      }
    }
    for (int64_t r = nc; r < ncp; ++r) {
      tgt.push_back(uint64_t(env_pos(cra, r, r)));
      srcpos.push_back(-1);
    }
    if (!grouped)
      for (int64_t k = 0; k < nc; ++k) {
        tgt.push_back(uint64_t(env_pos(cra, ncp + k, k)));
        srcpos.push_back(-1);
      }
    std::vector<uint32_t> idx(tgt.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = uint32_t(i);
    radix_sort_pairs(tgt, idx, uint64_t(P.coarse_pool_elems));
    for (size_t i = 0; i < tgt.size(); ++i) {
      if (i == 0 || tgt[i] != tgt[i - 1]) {
        P.coarse.target.push_back(int64_t(tgt[i]));
        P.coarse.ptr.push_back(int64_t(i));
      }
      P.coarse.srcpos.push_back(srcpos[idx[i]]);
    }
    P.coarse.ptr.push_back(int64_t(tgt.size()));
    // coarse schedule (single matrix, eliminate the first ncp/64 tile columns: the S~
    // block, which is all of a grouped pool)
    P.coarse_sched = single_matrix_schedule(int(ncp / kTile), int(Tc), true, P.coarse_ftile.data());
    if (grouped) build_coarse_groups(P, group_tiles);
  }
  P.coarse_inv_elems = P.coarse_ncp * kTile;  // (ncp / 64) tiles of 4096
}

void build_plans(const Topology& T, Plans& P, int coarse_group_tiles) {
  P = Plans();
  const int64_t nt = T.n_tets, ns = T.n_sub, nf = T.n_free;
  const int64_t one_code = enc(21 * nt, 1);  // trailing constant 1.0 of the element array

  // local index of every free dof inside its subdomain's padded layout
  std::vector<int32_t> loc(nf, -1);

  // ---- subdomain pools ------------------------------------------------------
  // Pass 1: envelope.  The first nonzero column of every local row (lower
  // triangle) bounds the Cholesky fill, so only tiles (I, J >= f_I) are stored:
  // tile row I is one contiguous run of (I - f_I + 1) tiles.
  P.ftile_off.resize(ns + 1, 0);
  for (int64_t s = 0; s < ns; ++s) P.ftile_off[s + 1] = P.ftile_off[s] + T.sub[s].T;
  P.ftile.assign(P.ftile_off[ns], 0);
  P.rowabs.assign(P.ftile_off[ns], 0);
  P.vec_dof.assign(T.vec_elems, -1);
  P.vec_w.assign(T.vec_elems, 0.0);
  P.rho_pos.assign(T.dvec_elems, -1);
  auto padded_of = [](const SubLayout& L, int k) {
    return k < L.ni ? k : (k < L.ni + L.nd ? L.o_d + (k - L.ni) : L.o_c + (k - L.ni - L.nd));
  };
  {
    int64_t off = 0;
    for (int64_t s = 0; s < ns; ++s) {
      const SubLayout& L = T.sub[s];
      const int32_t* dofs = &T.sub_dof[T.sub_dof_ptr[s]];
      const int n = L.ni + L.nd + L.nc;
      std::vector<int32_t> frow(L.m);
      for (int r = 0; r < L.m; ++r) frow[r] = r;
      for (int k = 0; k < n; ++k) loc[dofs[k]] = padded_of(L, k);
      for (int64_t j = T.sub_tet_ptr[s]; j < T.sub_tet_ptr[s + 1]; ++j) {
        const int64_t t = T.sub_tet[j];
        for (int a = 0; a < 6; ++a) {
          const int32_t da = T.tet_free[6 * t + a];
          if (da < 0) continue;
          for (int b = 0; b < 6; ++b) {
            const int32_t db = T.tet_free[6 * t + b];
            if (db < 0) continue;
            const int la = loc[da], lb = loc[db];
            if (la >= lb) frow[la] = std::min(frow[la], lb);
          }
        }
      }
      for (int k = 0; k < L.nd; ++k) frow[L.o_a + k] = std::min(frow[L.o_a + k], L.o_d + k);
      for (int k = 0; k < n; ++k) loc[dofs[k]] = -1;
      int32_t* ft = &P.ftile[P.ftile_off[s]];
      int64_t* ra = &P.rowabs[P.ftile_off[s]];
      P.pool_base.push_back(off);
      for (int I = 0; I < L.T; ++I) {
        int f = I;
        for (int r = I * kTile; r < (I + 1) * kTile; ++r) f = std::min(f, frow[r] / kTile);
        ft[I] = f;
        ra[I] = off - int64_t(f) * kTileElems;
        off += int64_t(I - f + 1) * kTileElems;
      }
    }
    P.pool_elems = off;
  }
  // Pass 2: contribution lists of every stored entry.
  {
    std::vector<uint64_t> tgt;
    std::vector<int32_t> code;
    tgt.reserve(size_t(nt) * 26);
    code.reserve(size_t(nt) * 26);
    for (int64_t s = 0; s < ns; ++s) {
      const SubLayout& L = T.sub[s];
      const int32_t* dofs = &T.sub_dof[T.sub_dof_ptr[s]];
      const int64_t* ra = &P.rowabs[P.ftile_off[s]];
      const int n = L.ni + L.nd + L.nc;
      for (int k = 0; k < n; ++k) {
        loc[dofs[k]] = padded_of(L, k);
        P.vec_dof[L.vec_off + padded_of(L, k)] = dofs[k];
        P.vec_w[L.vec_off + padded_of(L, k)] = 1.0 / T.free_mult[dofs[k]];
      }
      for (int k = 0; k < L.nd; ++k) P.rho_pos[L.dvec_off + k] = env_pos(ra, L.o_d + k, L.o_d + k);
      for (int64_t j = T.sub_tet_ptr[s]; j < T.sub_tet_ptr[s + 1]; ++j) {
        const int64_t t = T.sub_tet[j];
        for (int a = 0; a < 6; ++a) {
          const int32_t da = T.tet_free[6 * t + a];
          if (da < 0) continue;
          const int la = loc[da];
          for (int b = 0; b < 6; ++b) {
            const int32_t db = T.tet_free[6 * t + b];
            if (db < 0) continue;
            const int lb = loc[db];
            if (la < lb) continue;
            const int32_t c = enc(int64_t(packed_index(a, b)) * nt + t, T.tet_sign[6 * t + a] * T.tet_sign[6 * t + b]);
            tgt.push_back(uint64_t(env_pos(ra, la, lb)));
            code.push_back(c);
            if (la != lb && la / kTile == lb / kTile) {
              // upper half of a diagonal tile: element (lb, la) of tile (I, I)
              tgt.push_back(uint64_t(ra[la / kTile] + int64_t(la / kTile) * kTileElems + (la % kTile) * kTile + lb % kTile));
              code.push_back(c);
            }
          }
        }
      }
      auto diag_one = [&](int r) {
        tgt.push_back(uint64_t(env_pos(ra, r, r)));
        code.push_back(int32_t(one_code));
      };
      for (int r = L.ni; r < L.nip; ++r) diag_one(r);
      for (int r = L.o_d + L.nd; r < L.o_c; ++r) diag_one(r);
      for (int r = L.o_c + L.nc; r < L.o_a; ++r) diag_one(r);
      for (int k = 0; k < L.nd; ++k) {
        tgt.push_back(uint64_t(env_pos(ra, L.o_a + k, L.o_d + k)));
        code.push_back(int32_t(one_code));
      }
      for (int k = 0; k < n; ++k) loc[dofs[k]] = -1;
    }
    group_emissions(tgt, code, uint64_t(P.pool_elems), P.pool.target, P.pool.ptr, P.pool.src);
    P.tile_assembled.assign(size_t(P.pool_elems / kTileElems), 0);
    for (int64_t t : P.pool.target) P.tile_assembled[size_t(t / kTileElems)] = 1;
    for (size_t t = 0; t < P.tile_assembled.size(); ++t)
      if (P.tile_assembled[t]) P.assembled_tiles.push_back(int32_t(t));
  }

  // ---- global CSR and RHS ---------------------------------------------------
  build_global_plan(T, P.csr, P.rhs_global);

  // ---- coarse pool and its schedule ------------------------------------------
  build_coarse_plan(T, P, coarse_group_tiles);

  // ---- coarse gather positions ----------------------------------------------
  P.c_ypos.resize(T.c_gather_slot.size());
  P.c_vpos.resize(T.c_gather_slot.size());
  {
    std::vector<int32_t> sub_of_cslot(T.cvec_elems);
    for (int64_t s = 0; s < ns; ++s)
      for (int k = 0; k < T.sub[s].ncp; ++k) sub_of_cslot[T.sub[s].cvec_off + k] = int32_t(s);
    for (size_t j = 0; j < T.c_gather_slot.size(); ++j) {
      const int64_t slot = T.c_gather_slot[j];
      const SubLayout& L = T.sub[sub_of_cslot[slot]];
      const int64_t k = slot - L.cvec_off;
      P.c_ypos[j] = L.y_off + L.ndp + k;
      P.c_vpos[j] = L.vec_off + L.o_c + k;
    }
  }

  // ---- local factorization schedules (incremental diagonal / left-looking) ------
  // Only tiles inside the envelope (J >= f_I) are touched; the update of tile
  // (I, J) runs over K in [max(phase start, f_I, f_J), J).
  for (int variant = 0; variant < 2; ++variant) {
    const bool inc = variant == 0;
    TileSchedule& S = inc ? P.local_sched : P.local_sched_ll;
    int maxTi = 0, maxTd = 0;
    for (const SubLayout& L : T.sub) {
      maxTi = std::max(maxTi, L.Ti);
      maxTd = std::max(maxTd, L.Tr - L.Ti);
    }
    // phase 1: interior columns; rows of D' tiles are structurally zero here.
    // Lookahead (incremental variant): the mode-1 launch of column j also applies
    // K in [.., j) to the next column's mode-3 tile (j + 2, j + 1) (flags 2, update
    // only), so the side-stream item keeps one product and its diagonal factorization.
    // Diagonal tiles (FP64 variant): the same launch applies K in [f, j) to the diagonal
    // tile (j + 2, j + 2) (update only, left-looking over its row panel); the mode-3 item
    // of column j + 1 adds the K = j term (flags 64) and its own X X^T, so no other
    // column-step item touches a diagonal tile, and trailing diagonal tiles take their
    // sums in the Schur lists.
    // Schur updates of a subdomain need only its own interior columns; once the column
    // launches no longer fill the GPU (fewer items than kSideFill), the Schur updates of
    // the subdomains already done ride beside them (mode 4: a mode-0 launch that may run
    // on a second stream), the rest follow the last column.
    constexpr int kSideFill = 1024;
    constexpr bool kSplitSchur2 = true;
    // The Schur update of a D-block tile (flags 128) also writes its result, S_DD, into
    // the preconditioner store P; the snapshot launches copy only the D-block tiles that
    // no Schur item covers (no interior coupling: their assembled values, or zeros).
    auto schur1_of = [&](int64_t s, std::vector<TileWork>& out) {
      const SubLayout& L = T.sub[s];
      if (L.Ti == 0) return;
      const int32_t* ft = &P.ftile[P.ftile_off[s]];
      for (int J = L.Ti; J < L.Tc; ++J)
        for (int I = J; I < L.Tc; ++I) {
          const int k0 = std::max(ft[I], ft[J]);
          if (k0 < L.Ti) out.push_back(TileWork{int32_t(s), I, J, k0, L.Ti, I < L.Tr ? 128 : 0});
        }
    };
    std::vector<uint8_t> schur1_done(size_t(ns), 0);
    auto add_snap = [&](int64_t s) {
      const SubLayout& L = T.sub[s];
      const int32_t* ft = &P.ftile[P.ftile_off[s]];
      const int o = L.o_d / kTile;
      for (int I = 0; I < L.ndp / kTile; ++I)
        for (int J = 0; J <= I; ++J)
          if (L.Ti == 0 || std::max(ft[o + I], ft[o + J]) >= L.Ti)
            P.snap_tiles.push_back(SymTile{int32_t(s), I, J});
    };
    for (int j = 0; j < maxTi; ++j) {
      std::vector<TileWork> diag, next, below;
      for (int64_t s = 0; s < ns; ++s) {
        const SubLayout& L = T.sub[s];
        if (j >= L.Ti) continue;
        const int32_t* ft = &P.ftile[P.ftile_off[s]];
        if (!(inc && j > 0 && ft[j] <= j - 1))  // else factored by column j-1's mode-3 item
          diag.push_back(TileWork{int32_t(s), j, j, inc ? 0 : ft[j], inc ? 0 : j, 0});
        for (int I = j + 1; I < L.Tc; ++I)
          if (ft[I] <= j) {
            TileWork w{int32_t(s), I, j, std::max(ft[I], ft[j]), j, 0};
            const bool is_next = inc && I == j + 1 && I < L.Ti;
            if (is_next) {
              // diagonal share X X^T, and the K = j - 1 term of the diagonal when stored
              w.flags = 1 | (j >= 1 && ft[I] <= j - 1 ? 64 : 0);
              if (j >= 1 && w.k0 < j - 1) w.k0 = j - 1;  // rest applied by lookahead
            }
            (is_next ? next : below).push_back(w);
          }
        if (inc && j + 2 < L.Ti) {
          const int k0 = std::max(ft[j + 2], ft[j + 1]);
          if (k0 < j) below.push_back(TileWork{int32_t(s), j + 2, j + 1, k0, j, 2});
          if (ft[j + 2] < j) below.push_back(TileWork{int32_t(s), j + 2, j + 2, ft[j + 2], j, 2});
        }
      }
      if (inc && int64_t(diag.size() + next.size() + below.size()) < kSideFill &&
          std::find(schur1_done.begin(), schur1_done.end(), 1) == schur1_done.end()) {
        std::vector<TileWork> early;
        for (int64_t s = 0; s < ns; ++s)
          if (T.sub[s].Ti <= j) {
            schur1_of(s, early);
            schur1_done[size_t(s)] = 1;
          }
        xcd_grouped(early, true);
        add_tiles(S, 4, early);
        // their S_DD snapshot can follow at once (the later launches do not touch them)
        const int64_t off = int64_t(P.snap_tiles.size());
        for (int64_t s = 0; s < ns; ++s)
          if (schur1_done[size_t(s)]) add_snap(s);
        if (int64_t(P.snap_tiles.size()) > off)
          S.launches.push_back(TileLaunch{-2, off, int64_t(P.snap_tiles.size()) - off});
      }
      add_tiles(S, 2, diag);
      add_tiles(S, 3, next);
      xcd_grouped(below, false);
      add_tiles(S, 1, below);
    }
    std::vector<TileWork> schur1;
    for (int64_t s = 0; s < ns; ++s)
      if (!schur1_done[size_t(s)]) schur1_of(s, schur1);
    xcd_grouped(schur1, true);
    add_tiles(S, 0, schur1);
    // snapshot S_DD for the preconditioner (the rest of the subdomains; all of them in the
    // left-looking variant, over the same list)
    {
      if (inc) {
        const int64_t off = int64_t(P.snap_tiles.size());
        for (int64_t s = 0; s < ns; ++s)
          if (!schur1_done[size_t(s)]) add_snap(s);
        S.launches.push_back(TileLaunch{-1, off, int64_t(P.snap_tiles.size()) - off});
      } else {  // built after the incremental variant, whose lists cover every subdomain
        S.launches.push_back(TileLaunch{-1, 0, int64_t(P.snap_tiles.size())});
      }
    }
    // phase 2: dual columns (the Schur updates of finished subdomains again ride beside
    // the column launches once those no longer fill the GPU)
    auto schur2_of = [&](int64_t s, std::vector<TileWork>& out) {
      const SubLayout& L = T.sub[s];
      if (L.Tr == L.Ti) return;
      const int32_t* ft = &P.ftile[P.ftile_off[s]];
      for (int J = L.Tr; J < L.T; ++J)
        for (int I = J; I < L.T; ++I) {
          const int k0 = std::max(L.Ti, std::max(ft[I], ft[J]));
          // flags 32: a D' row's Schur update also writes its tile of W (Z or Phi^T)
          if (k0 < L.Tr) out.push_back(TileWork{int32_t(s), I, J, k0, L.Tr, I >= L.Tc ? 32 : 0});
        }
    };
    std::vector<uint8_t> schur2_done(size_t(ns), 0);
    bool early2 = false;
    for (int j = 0; j < maxTd; ++j) {
      std::vector<TileWork> diag, next, below;
      for (int64_t s = 0; s < ns; ++s) {
        const SubLayout& L = T.sub[s];
        const int col = L.Ti + j;
        if (col >= L.Tr) continue;
        const int32_t* ft = &P.ftile[P.ftile_off[s]];
        // the first dual diagonal keeps a launch of its own: the S_DD snapshot must
        // see it unfactored
        if (!(inc && col > L.Ti && ft[col] <= col - 1))
          diag.push_back(TileWork{int32_t(s), col, col, inc ? 0 : std::max(L.Ti, ft[col]), inc ? 0 : col, 0});
        for (int I = col + 1; I < L.T; ++I)
          if (ft[I] <= col) {
            TileWork w{int32_t(s), I, col, std::max(L.Ti, std::max(ft[I], ft[col])), col, 0};
            const bool is_next = inc && I == col + 1 && I < L.Tr;
            if (is_next) {
              w.flags = 1 | (col - 1 >= std::max(L.Ti, ft[I]) ? 64 : 0);
              if (col >= 1 && w.k0 < col - 1) w.k0 = col - 1;  // rest applied by lookahead
            }
            (is_next ? next : below).push_back(w);
          }
        if (inc && col + 2 < L.Tr) {
          const int k0 = std::max(L.Ti, std::max(ft[col + 2], ft[col + 1]));
          if (k0 < col) below.push_back(TileWork{int32_t(s), col + 2, col + 1, k0, col, 2});
          const int kd = std::max(L.Ti, ft[col + 2]);
          if (kd < col) below.push_back(TileWork{int32_t(s), col + 2, col + 2, kd, col, 2});
        }
      }
      if (inc && !kSplitSchur2 && !early2 && int64_t(diag.size() + next.size() + below.size()) < kSideFill) {
        early2 = true;
        std::vector<TileWork> early;
        for (int64_t s = 0; s < ns; ++s)
          if (T.sub[s].Tr - T.sub[s].Ti <= j) {
            schur2_of(s, early);
            schur2_done[size_t(s)] = 1;
          }
        xcd_grouped(early, true);
        add_tiles(S, 4, early);
      }
      add_tiles(S, 2, diag);
      add_tiles(S, 3, next);
      xcd_grouped(below, false);
      add_tiles(S, 1, below);
    }
    std::vector<TileWork> schur2;
    for (int64_t s = 0; s < ns; ++s)
      if (!schur2_done[size_t(s)]) schur2_of(s, schur2);
    if (kSplitSchur2) {
      // the coarse problem needs only the c rows: the D' rows (Z and Phi) go to a mode-4
      // launch that the coarse factorization does not wait for
      std::vector<TileWork> crow, drow;
      for (const TileWork& w : schur2) (w.I >= T.sub[w.mat].Tc ? drow : crow).push_back(w);
      xcd_grouped(drow, true);
      add_tiles(S, 4, drow);
      xcd_grouped(crow, true);
      add_tiles(S, 0, crow);
    } else {
      xcd_grouped(schur2, true);
      add_tiles(S, 0, schur2);
    }
  }

  // ---- W written by the Schur updates: complete when every Z tile and Phi^T block of
  // every subdomain has a flags-32 item in both schedules (else the extraction runs) -------
  P.w_fused = true;
  for (const TileSchedule* S : {&P.local_sched, &P.local_sched_ll}) {
    std::vector<int64_t> cnt(size_t(ns), 0);
    for (const TileWork& w : S->items)
      if (w.flags & 32) ++cnt[size_t(w.mat)];
    for (int64_t s = 0; s < ns; ++s) {
      const int64_t td = T.sub[s].ndp / kTile, tc = T.sub[s].ncp / kTile;
      if (cnt[size_t(s)] != td * (td + 1) / 2 + tc * td) P.w_fused = false;
    }
  }

  // ---- fresh tiles: mark items whose tile (flags 4) or diagonal tile (flags 8) nothing
  // has written yet (only assembled tiles are cleared before a factorization) ----------
  for (int variant = 0; variant < 2; ++variant) {
    TileSchedule& S = variant == 0 ? P.local_sched : P.local_sched_ll;
    std::vector<uint8_t> w = P.tile_assembled;
    std::vector<uint8_t>& snap = variant == 0 ? P.tile_written_snap : P.tile_written_snap_ll;
    for (const TileLaunch& L : S.launches) {
      if (L.mode < 0) {  // mode -2 tiles do not change before the mode -1 launch
        if (L.mode == -1) snap = w;
        continue;
      }
      for (int64_t i = L.off; i < L.off + L.n; ++i) {
        TileWork& it = S.items[size_t(i)];
        const int64_t* ra = &P.rowabs[P.ftile_off[it.mat]];
        const size_t c = size_t((ra[it.I] + int64_t(it.J) * kTileElems) / kTileElems);
        if (!w[c]) it.flags |= 4;
        const bool share = (L.mode == 1 && (it.flags & 1) && !(it.flags & 2)) || L.mode == 3;
        if (share) {
          const size_t d = size_t((ra[it.I] + int64_t(it.I) * kTileElems) / kTileElems);
          if (!w[d]) it.flags |= 8;
          w[d] = 1;
        }
      }
      for (int64_t i = L.off; i < L.off + L.n; ++i) {  // this launch's writes
        const TileWork& it = S.items[size_t(i)];
        const int64_t* ra = &P.rowabs[P.ftile_off[it.mat]];
        w[size_t((ra[it.I] + int64_t(it.J) * kTileElems) / kTileElems)] = 1;
      }
    }
    (variant == 0 ? P.tile_written : P.tile_written_ll) = w;
  }

  P.d_self.assign(T.dvec_elems, -1);
  for (int64_t k = 0; k < T.dvec_elems; ++k)
    if (T.d_lam[k] >= 0) P.d_self[k] = int32_t(k);
  build_faces(T, P);

  // ---- bounds validation of every index the kernels will use ------------------
  auto check = [](bool ok, const char* what) {
    if (!ok) throw std::runtime_error(std::string("plan out of range: ") + what);
  };
  for (int64_t t : P.pool.target) check(t >= 0 && t < P.pool_elems, "pool target");
  for (int64_t t : P.coarse.target) check(t >= 0 && t < P.coarse_pool_elems, "coarse target");
  for (int64_t sp : P.coarse.srcpos) check(sp < P.pool_elems, "coarse source");
  // every tile a work item touches (C, panels from k0, diagonal) lies in the envelope
  for (const TileSchedule* S : {&P.local_sched, &P.local_sched_ll})
    for (const TileWork& w : S->items) {
      check(w.mat >= 0 && w.mat < ns, "local tile matrix");
      const SubLayout& L = T.sub[w.mat];
      const int32_t* ft = &P.ftile[P.ftile_off[w.mat]];
      check(w.J <= w.I && w.I < L.T && w.k1 <= w.J && w.J >= ft[w.I], "local tile");
      check(w.k0 >= w.k1 || (w.k0 >= ft[w.I] && w.k0 >= ft[w.J]), "local tile panel");
    }
  // every stored tile (I > J) of an eliminated column J receives its column-step K range
  // ([max(f_I, f_J), J) for interior columns, [max(Ti, f_I, f_J), J) for dual columns)
  // exactly once, split between a lookahead item and its own item at most
  {
    const TileSchedule& S = P.local_sched;
    std::vector<int64_t> toff(size_t(ns) + 1, 0);
    for (int64_t s = 0; s < ns; ++s) toff[size_t(s) + 1] = toff[size_t(s)] + int64_t(T.sub[s].T) * T.sub[s].Tr;
    std::vector<int32_t> next_k(size_t(toff[size_t(ns)]), -1);
    for (const TileLaunch& L : S.launches) {
      if (L.mode != 1 && L.mode != 3) continue;
      for (int64_t i = L.off; i < L.off + L.n; ++i) {
        const TileWork& w = S.items[size_t(i)];
        if (w.I == w.J && (w.flags & 2)) continue;  // a diagonal lookahead (checked below)
        const SubLayout& Ls = T.sub[w.mat];
        check(w.J < Ls.Tr && w.I > w.J, "local column-step item");
        const int32_t* ft = &P.ftile[P.ftile_off[w.mat]];
        const int32_t start = std::max(w.J >= Ls.Ti ? Ls.Ti : 0, std::max(ft[w.I], ft[w.J]));
        int32_t& nk = next_k[size_t(toff[w.mat] + int64_t(w.I) * Ls.Tr + w.J)];
        if (nk < 0) nk = start;
        check(w.k0 == nk || w.k0 >= w.k1, "local K ranges contiguous");
        if (w.k1 > w.k0) nk = w.k1;
      }
    }
    for (int64_t s = 0; s < ns; ++s) {
      const SubLayout& Ls = T.sub[s];
      const int32_t* ft = &P.ftile[P.ftile_off[s]];
      for (int I = 0; I < Ls.T; ++I)
        for (int J = ft[I]; J < std::min(I, Ls.Tr); ++J) {
          if (J < Ls.Ti && I >= Ls.Tc) continue;  // D' rows are zero in interior columns
          const int32_t nk = next_k[size_t(toff[s] + int64_t(I) * Ls.Tr + J)];
          check(nk == J, "local K ranges complete (and every tile has its column step)");
        }
    }
  }
  // every diagonal tile receives its K range exactly once, in launch order: from the
  // start of its row's envelope (Schur items and diagonal lookaheads apply ranges, a
  // mode-3 item the K = I - 2 term when flagged and its own K = I - 1 share) to its own
  // column (eliminated rows) or to the end of the eliminated columns (trailing rows)
  {
    const TileSchedule& S = P.local_sched;
    std::vector<int64_t> roff(size_t(ns) + 1, 0);
    for (int64_t s = 0; s < ns; ++s) roff[size_t(s) + 1] = roff[size_t(s)] + T.sub[s].T;
    std::vector<int32_t> next_d(size_t(roff[size_t(ns)]), -1);
    auto start_of = [&](int64_t s, int I) {
      const SubLayout& Ls = T.sub[s];
      const int f = P.ftile[P.ftile_off[s] + I];
      return I >= Ls.Tc ? std::max(Ls.Ti, f) : f;  // D' rows are zero in interior columns
    };
    auto apply = [&](int64_t s, int I, int k0, int k1) {
      if (k1 <= k0) return;
      int32_t& nk = next_d[size_t(roff[s] + I)];
      if (nk < 0) nk = start_of(s, I);
      check(k0 == nk, "local diagonal K ranges contiguous");
      nk = k1;
    };
    for (const TileLaunch& L : S.launches) {
      if (L.mode < 0) continue;
      for (int64_t i = L.off; i < L.off + L.n; ++i) {
        const TileWork& w = S.items[size_t(i)];
        if (w.I == w.J) {
          apply(w.mat, w.I, w.k0, w.k1);
        } else if (L.mode == 3 || (w.flags & 16)) {
          if (w.flags & 64) apply(w.mat, w.I, w.I - 2, w.I - 1);
          apply(w.mat, w.I, w.J, w.J + 1);
        } else {
          check(!(w.flags & (1 | 64)), "diagonal share outside a mode-3 item");
        }
      }
    }
    for (int64_t s = 0; s < ns; ++s) {
      const SubLayout& Ls = T.sub[s];
      for (int I = 0; I < Ls.T; ++I) {
        const int end = std::min(I, Ls.Tr), st = start_of(s, I);
        if (st >= end) continue;
        const int32_t nk = next_d[size_t(roff[s] + I)];
        check(nk == end, "local diagonal K ranges complete");
      }
    }
  }
  // a mode-3 item completes and factors the diagonal tile of the next column
  for (const TileSchedule* S : {&P.local_sched, &P.local_sched_ll, &P.coarse_sched})
    for (const TileLaunch& L : S->launches)
      if (L.mode == 3)
        for (int64_t i = L.off; i < L.off + L.n; ++i)
          check(S->items[i].I == S->items[i].J + 1 && (S->items[i].flags & 1), "mode-3 item");
  {
    const int Ta = int(P.coarse_ftile.size());  // 2 ncp / 64 (dense inverse) or ncp / 64 (grouped)
    for (const TileWork& w : P.coarse_sched.items) {
      check(w.mat == 0 && w.J <= w.I && w.I < Ta && w.k1 <= w.J && w.J >= P.coarse_ftile[w.I],
            "coarse tile");
      check(w.k0 >= w.k1 || (w.k0 >= P.coarse_ftile[w.I] && w.k0 >= P.coarse_ftile[w.J]),
            "coarse tile panel");
    }
    // every stored tile (I, J) below the diagonal receives the K range
    // [max(f_I, f_J), min(J, Tp)) exactly once, in pieces listed in launch order
    // (update-only chunks, then its own item); no column-launch item other than a
    // first diagonal loops over more than 8 products
    const int Tp = int(P.coarse_ncp / kTile);
    std::vector<int32_t> next_k(size_t(Ta) * size_t(Ta), -1);
    for (const TileLaunch& L : P.coarse_sched.launches)
      for (int64_t i = L.off; i < L.off + L.n; ++i) {
        const TileWork& w = P.coarse_sched.items[size_t(i)];
        if (L.mode == 1 || L.mode == 3) check(w.k1 - w.k0 <= 8, "coarse chunk length");
        if (w.I == w.J) continue;
        int32_t& nk = next_k[size_t(w.I) * Ta + w.J];
        const int32_t start = std::max(P.coarse_ftile[w.I], P.coarse_ftile[w.J]);
        if (nk < 0) nk = start;
        check(w.k0 == nk || w.k0 >= w.k1, "coarse K ranges contiguous");
        if (w.k1 > w.k0) nk = w.k1;
      }
    for (int I = 0; I < Ta; ++I)
      for (int J = P.coarse_ftile[I]; J < I; ++J) {
        const int32_t nk = next_k[size_t(I) * Ta + J];
        const int32_t start = std::max(P.coarse_ftile[I], P.coarse_ftile[J]);
        if (J < Tp) check(nk == J, "coarse K ranges complete");
        else if (start < Tp) check(nk == Tp, "coarse Schur K ranges complete");
      }
  }
  for (int64_t y : P.c_ypos) check(y >= 0 && y < T.y_elems, "coarse y position");
  {
    const FacePlan& F = P.faces;
    for (const TileWork& w : F.sched.items) {
      check(w.mat >= 0 && w.mat < F.n, "face matrix");
      const int ta = 2 * F.dev[w.mat].tf;
      check(w.J <= w.I && w.I < ta && w.k1 <= w.J && w.k0 >= 0, "face tile");
    }
    int64_t dd = 0;
    for (const FaceDev& d : F.dev) {
      check(d.tf >= 1 && d.nf <= d.tf * kTile && d.d_off == dd, "face D_F offset");
      dd += int64_t(d.tf) * d.tf * kTileElems;
    }
    check(dd == F.d_elems, "face D_F size");
    int64_t nreal = 0;
    for (int32_t l : F.map_lam) nreal += l >= 0;
    check(nreal == T.n_lam, "faces cover every multiplier once");
  }
  for (int64_t v : P.c_vpos) check(v >= 0 && v < T.vec_elems, "coarse v position");

  // ---- inverse diagonal tiles: one per factored tile column ---------------------
  P.inv_off.resize(ns);
  {
    int64_t off = 0;
    for (int64_t s = 0; s < ns; ++s) {
      P.max_dual = std::max(P.max_dual, T.sub[s].Tr - T.sub[s].Ti);
      P.max_after = std::max(P.max_after, T.sub[s].T - T.sub[s].Ti);
      P.inv_off[s] = off;
      off += int64_t(T.sub[s].Tr) * kTileElems;
      P.max_T = std::max(P.max_T, T.sub[s].T);
      P.max_Tr = std::max(P.max_Tr, T.sub[s].Tr);
      P.max_Ti = std::max(P.max_Ti, T.sub[s].Ti);
    }
    P.inv_elems = std::max<int64_t>(off, kTileElems);
    // backward interior sweep: step J touches rows I in [f_J, J) of every subdomain with
    // J < Ti (and always I = J - 1); its launch needs the widest such range
    P.bwd_width.assign(size_t(std::max(P.max_Ti, 1)), 0);
    for (int64_t s = 0; s < ns; ++s) {
      const int32_t* ft = &P.ftile[P.ftile_off[s]];
      for (int J = 1; J < T.sub[s].Ti; ++J)
        P.bwd_width[size_t(J)] = std::max(P.bwd_width[size_t(J)], J - std::min(ft[J], J - 1));
    }
    P.coarse_inv_elems = P.coarse_ncp;  // (ncp / 64) tiles of 4096 = ncp * 64 elements
    P.coarse_inv_elems *= kTile;
  }

  // ---- device layout table ----------------------------------------------------
  P.devsub.resize(ns);
  for (int64_t s = 0; s < ns; ++s) {
    const SubLayout& L = T.sub[s];
    DevSub& D = P.devsub[s];
    D.nd = L.nd; D.nc = L.nc; D.ndp = L.ndp; D.ncp = L.ncp;
    D.o_d = L.o_d; D.o_c = L.o_c; D.o_a = L.o_a; D.m = L.m;
    D.T = L.T; D.Ti = L.Ti; D.Tr = L.Tr; D.Tc = L.Tc;
    D.pool_off = L.pool_off; D.vec_off = L.vec_off; D.dvec_off = L.dvec_off;
    D.w_off = L.w_off; D.p_off = L.p_off; D.y_off = L.y_off; D.cvec_off = L.cvec_off;
    D.ft_off = P.ftile_off[s];
    D.sp_off = L.sp_off;
    for (int I = 0; I < L.ndp / kTile; ++I)
      for (int J = 0; J <= I; ++J) P.sym_tiles.push_back(SymTile{int32_t(s), I, J});
    D.pool_off = P.pool_base[s];
    P.max_ndp = std::max(P.max_ndp, L.ndp);
    P.max_Td = std::max(P.max_Td, L.ndp / kTile);
    P.max_ncp = std::max(P.max_ncp, L.ncp);
    P.max_m = std::max(P.max_m, L.m);
    P.max_wrows = std::max(P.max_wrows, L.ndp + L.ncp);
  }
}

}  // namespace emag

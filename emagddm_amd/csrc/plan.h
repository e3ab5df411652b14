// Assembly plans and the batched factorization schedule (host side).
//
// Every GPU reduction in the solver is written as "each output sums its own
// contribution list in a fixed order", so results are bitwise repeatable.
// The plans below are those lists, built once per mesh/partition.
#pragma once
#include <cstdint>
#include <vector>

#include "topology.h"

namespace emag {

// out[target[i]] = sum_{c in [ptr[i], ptr[i+1])} +-elem[src[c] >> 1]
struct NzPlan {
  std::vector<int64_t> target, ptr;
  std::vector<int32_t> src;  // (element index << 1) | negative
  int64_t size() const { return int64_t(target.size()); }
};

// out[target[i]] = sum_c (srcpos[c] < 0 ? 1 : pool[srcpos[c]])
struct PoolPlan {
  std::vector<int64_t> target, ptr, srcpos;
  int64_t size() const { return int64_t(target.size()); }
};

struct CsrPlan {
  std::vector<int32_t> row_ptr, col;
  std::vector<int64_t> ptr;
  std::vector<int32_t> src;
};

// A run of the batched tile factorization: list of launches over `items`.
//   mode 0/1/2/3 as launch_tile_mode; mode 4: Schur updates (mode 0) of subdomains whose
//   columns are done, which may run beside the launches that follow them up to the next
//   snapshot or the end; mode -1: snapshot the Dirichlet Schur blocks (snap_tiles[off,
//   off + n)); mode -2: the same for the subdomains of the preceding mode-4 launch, which
//   may run beside it.
struct TileLaunch {
  int mode;
  int64_t off, n;
};
struct TileSchedule {
  std::vector<TileWork> items;
  std::vector<TileLaunch> launches;
  double flops = 0.0;  // algorithmic flops of the tile updates
};

// Interface faces for deluxe scaling: the multipliers shared by one subdomain pair
// (s0 < s1).  Per face F the solver forms A_F = S_s0,FF + S_s1,FF from the two
// Dirichlet Schur blocks, inverts it by a partial Cholesky of [[A_F, I], [I, 0]]
// (trailing block -A_F^-1), keeps B_F = S_s1,FF and forms D_F = A_F^-1 B_F explicitly;
// the scaled jump operator applies D_F and its transpose.
struct FacePlan {
  int64_t n = 0;
  std::vector<FaceDev> dev;
  std::vector<int32_t> map_lam;       // face-vector slot -> multiplier (-1 padding)
  std::vector<int32_t> map_self;      // face-vector slot -> itself (-1 padding)
  std::vector<int64_t> slot0, slot1;  // face-vector slot -> D slot in s0 / s1 (-1 padding)
  int64_t vec_elems = 0, op_elems = 0, pool_elems = 0, inv_elems = 0;
  int64_t d_elems = 0;                // explicit D_F of every face (tf^2 tiles each)
  int max_T = 0;                      // max padded order / 64
  std::vector<int64_t> rowabs, rowbase, inv_off;
  TileSchedule sched;                 // eliminates the A_F block of every face
};

struct Plans {
  NzPlan pool;        // element matrices -> subdomain factor pools (augmented, tiled)
  NzPlan rhs_global;  // element F -> global free-dof RHS
  CsrPlan csr;        // element matrices -> global free-dof CSR
  PoolPlan coarse;    // subdomain trailing blocks -> coarse augmented pool
  std::vector<int32_t> vec_dof;  // local-vector slot -> free dof (-1 none)
  std::vector<double> vec_w;     // 1 / multiplicity (RHS split weights)
  std::vector<int64_t> c_ypos;   // per coarse gather entry: position in W-output space
  std::vector<int64_t> c_vpos;   // per coarse gather entry: position in local-vector space
  std::vector<int64_t> rho_pos;  // per D slot: pool position of its diagonal (-1 padding)
  int64_t coarse_ncp = 0;        // padded coarse order
  int64_t coarse_pool_elems = 0;
  // Grouped coarse solve (coarse_grouped): the coarse pool holds L of S~ = L L^T and
  // S~^-1 v is a forward and a backward sweep over groups of coarse_group tile columns
  // (group g = [cg_start[g], cg_start[g+1])).  Per group, the panel tiles left of (fwd) or
  // below (bwd) its diagonal block, as (I, J) pairs whose products land in one partial
  // 64-vector slot each (slots of one row / column contiguous: cg_fwd_rowptr per tile row,
  // cg_bwd_colptr per tile column), and X_g = (L_gg)^-1, dense column-major of order
  // 64 * (group width) at cg_xoff[g].  cg_inv_items: the (i, j) tiles of every X_g by
  // distance i - j (step d at [cg_inv_step[d], cg_inv_step[d+1])).
  bool coarse_grouped = false;
  int coarse_group = 0;
  std::vector<int32_t> cg_start, cg_group_of;
  std::vector<int64_t> cg_xoff;
  int64_t cg_x_elems = 0;
  std::vector<int32_t> cg_fwd_tiles, cg_fwd_rowptr, cg_bwd_tiles, cg_bwd_colptr;
  std::vector<int64_t> cg_fwd_off, cg_bwd_off;
  std::vector<int32_t> cg_inv_items;
  std::vector<int64_t> cg_inv_step;
  TileSchedule local_sched, local_sched_ll, coarse_sched;
  std::vector<DevSub> devsub;
  int max_ndp = 0, max_ncp = 0, max_m = 0, max_wrows = 0;
  int max_T = 0, max_Tr = 0, max_Ti = 0;
  int max_dual = 0, max_after = 0;  // max over subdomains of Tr - Ti and of T - Ti
  bool w_fused = false;  // the Schur updates write all of W (no extraction launch)
  std::vector<int64_t> inv_off;  // per subdomain: offset of its inverse diagonal tiles
  std::vector<int32_t> bwd_width;  // per interior tile row J: max over subdomains of J - f_J
                                   // (rows a backward interior step touches)
  std::vector<int32_t> ftile;    // per subdomain and tile row: first nonzero tile column
  std::vector<int64_t> ftile_off;     // per subdomain: offset of its tile rows in ftile/rowabs
  std::vector<int64_t> rowabs;        // per (subdomain, tile row I): element offset of tile (I, 0)
                                      // in the packed pool; tile (I, J >= f_I) at rowabs + J*4096
  std::vector<int64_t> pool_base;     // per subdomain: first element of its packed tiles
  int64_t pool_elems = 0;
  // Pool tiles (index = element offset / 4096) that receive assembled entries: only they
  // are cleared before assembly; every other stored tile starts as "fresh" (its flag is 0)
  // and the tile kernels read it as zero until they first write it.
  std::vector<uint8_t> tile_assembled;
  std::vector<int32_t> assembled_tiles;
  // per pool tile: written by the end of the FP64 / FP32 schedule, and before its S_DD
  // snapshot; tiles never written keep stale data and are read as zero
  std::vector<uint8_t> tile_written, tile_written_ll, tile_written_snap, tile_written_snap_ll;
  std::vector<int32_t> coarse_ftile;
  std::vector<int64_t> coarse_rowabs;
  int64_t inv_elems = 0, coarse_inv_elems = 0;
  std::vector<SymTile> sym_tiles;  // every lower tile of every subdomain's Z / P
  // the same tiles for the S_DD snapshots: the subdomains whose Schur updates ran early
  // first (snapshot launch mode -2), then the rest (mode -1)
  std::vector<SymTile> snap_tiles;
  int max_Td = 0;                  // max ndp / 64
  std::vector<int32_t> d_self;     // D slot -> itself (-1 on padding)
  FacePlan faces;
};

// Packed index of (a, b), a <= b, in the 21-entry upper triangle of a 6x6.
int packed_index(int a, int b);

// coarse_group_tiles: 0 = dense coarse inverse, > 0 = grouped coarse solve with groups of
// that many tile columns (build_coarse_plan).
void build_plans(const Topology& T, Plans& P, int coarse_group_tiles = 0);
void build_coarse_plan(const Topology& T, Plans& P, int coarse_group_tiles);
// Group lists of a grouped coarse plan for another group width (the pool, its schedule and
// its factor do not depend on the width).
void build_coarse_groups(Plans& P, int group_tiles);
// Only the global free-dof CSR and RHS plans (also part of build_plans).
void build_global_plan(const Topology& T, CsrPlan& csr, NzPlan& rhs);

// Left-looking schedule of one matrix of Ta tile rows: eliminate the first Tp
// tile columns, then update the trailing (Ta-Tp)^2 block with the Schur term.
// ft (optional): first nonzero tile column of every tile row (envelope).
TileSchedule single_matrix_schedule(int Tp, int Ta, bool incremental_diagonal,
                                    const int32_t* ft = nullptr);

// Incremental-diagonal schedule of a batch of dense matrices: matrix m has 2 Tp[m]
// tile rows and its first Tp[m] tile columns are eliminated.
TileSchedule batched_dense_schedule(const std::vector<int>& Tp);

// Dense lower-triangular tile layout (used for matrices without an envelope).
inline int64_t tile_offset(int64_t I, int64_t J) { return (I * (I + 1) / 2 + J) * kTileElems; }
inline int64_t pool_pos(int64_t base, int64_t r, int64_t c) {  // r >= c
  return base + tile_offset(r / kTile, c / kTile) + (c % kTile) * kTile + (r % kTile);
}
// Envelope-packed layout: element (r, c), r >= c inside the envelope, of a
// matrix whose tile-row offsets are rowabs.
inline int64_t env_pos(const int64_t* rowabs, int64_t r, int64_t c) {
  return rowabs[r / kTile] + (c / kTile) * kTileElems + (c % kTile) * kTile + (r % kTile);
}

}  // namespace emag

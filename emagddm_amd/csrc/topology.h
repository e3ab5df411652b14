// Host-side mesh topology for the FETI-DP Nedelec solver.
//
// Builds, from the raw tet mesh, everything the GPU kernels index with:
// global edge ids (first-appearance order, as GlobalEdgeIDGenerator::generate,
// src/fe_em/global_edge_id_generator.cpp:30-110), orientation signs, free-dof
// numbering with PEC edges removed (EMDOFManager::build semantics,
// src/fe_em/em_dof_manager.cpp:91-393), the FETI-DP edge classes per
// subdomain (interior / dual / primal by multiplicity) and the padded,
// tile-aligned local layouts.  Integer work only; O(N) radix sorts.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace emag {

constexpr int kTile = 64;  // tile edge of the batched dense factorization
constexpr int kTileElems = kTile * kTile;

// One unit of batched tile work: tile (I, J) of matrix `mat`,
// C_IJ -= sum_{K in [k0,k1)} L_IK L_JK^T, followed by the launch's mode action.
struct TileWork {
  int32_t mat, I, J, k0, k1, flags;
};

// Sub-domain layout as the kernels see it (device copy of SubLayout fields).
struct DevSub {
  int32_t nd, nc, ndp, ncp, o_d, o_c, o_a, m, T, Ti, Tr, Tc;
  int64_t pool_off, vec_off, dvec_off, w_off, p_off, y_off, cvec_off;
  int64_t ft_off;  // offset of this subdomain's envelope (first nonzero tile per tile row)
  int64_t sp_off;  // offset of its (ndp/64)^2 partial 64-vectors of the symmetric products
};

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// One interface face (multipliers shared by subdomains s0 < s1) as the deluxe-scaling
// kernels see it.
struct FaceDev {
  int32_t s0, s1, nf, tf;       // subdomains, multipliers, padded order / 64
  int64_t pool_base, vec_off;   // factor pool (dense lower tiles), face-vector offset
  int64_t ainv_off, b_off;      // lower-tile stores of A_F^-1 and B_F
  int64_t d_off;                // D_F = A_F^-1 B_F, tf x tf full tiles (row-major tiles)
};

// One lower tile (I, J) of subdomain s's symmetric dual operators (Z and P).
struct SymTile {
  int32_t s, I, J;
};

// Elements of a symmetric n x n operator (n a multiple of 64) stored as its lower
// 64x64 tiles: tile (I, J <= I) at (I(I+1)/2 + J) * 4096, row-major inside the tile,
// diagonal tiles full.
inline int64_t sym_tile_elems(int64_t n) {
  const int64_t t = n / kTile;
  return t * (t + 1) / 2 * kTileElems;
}

// Local padded layout of one subdomain's augmented matrix:
//   [ i (interior) | D (dual) | c (primal) | D' (copy of D, identity-coupled) ]
// each block padded to a multiple of kTile.
struct SubLayout {
  int32_t ni = 0, nd = 0, nc = 0;      // real sizes
  int32_t nip = 0, ndp = 0, ncp = 0;   // padded sizes
  int32_t o_d = 0, o_c = 0, o_a = 0;   // block offsets (o_i = 0)
  int32_t m = 0;                       // padded order = nip + 2*ndp + ncp
  int32_t T = 0, Ti = 0, Tr = 0, Tc = 0;  // tiles: total, end of i, end of D, end of c
  int64_t pool_off = 0;                // element offset of tile (0,0) in the factor pool
  int64_t vec_off = 0;                 // offset of the length-m local vector
  int64_t dvec_off = 0;                // offset into the concatenated D space (ndp each)
  int64_t w_off = 0;                   // offset of W: Z (ndp x ndp, lower tiles) then Phi^T (ncp x ndp)
  int64_t p_off = 0;                   // offset of P = S_DD (ndp x ndp, lower tiles)
  int64_t sp_off = 0;                  // offset of the symmetric-product partials ((ndp/64)^2 x 64)
  int64_t y_off = 0;                   // offset into the W-output space ((ndp+ncp) each)
  int64_t cvec_off = 0;                // offset into the concatenated local primal space (ncp each)
};

struct Topology {
  int64_t n_nodes = 0, n_tets = 0, n_edges = 0, n_free = 0;
  int32_t n_sub = 0;
  int64_t n_lam = 0, n_c = 0;

  std::vector<int32_t> tet_edge;   // (t,6) global edge id, ElementGeometry order
  std::vector<int32_t> tet_free;   // (t,6) free dof or -1
  std::vector<int8_t> tet_sign;    // (t,6) +-1 of the reference Whitney function
  std::vector<int32_t> edge_nodes; // (E,2) low, high global node id
  std::vector<int32_t> edge_free;  // (E) free index or -1
  std::vector<int32_t> free_mult;  // (n_free) number of subdomains touching
  std::vector<int32_t> lam_of_free, c_of_free;  // -1 where not dual / primal
  std::vector<int32_t> free_of_lam, free_of_c;

  // per subdomain
  std::vector<SubLayout> sub;
  std::vector<int64_t> sub_dof_ptr;   // CSR: local dofs (free ids) in (i, D, c) order
  std::vector<int32_t> sub_dof;
  std::vector<int64_t> sub_tet_ptr;   // CSR: tets of each subdomain (ascending)
  std::vector<int32_t> sub_tet;
  int64_t pool_elems = 0, vec_elems = 0, dvec_elems = 0, w_elems = 0, p_elems = 0,
          y_elems = 0, cvec_elems = 0, sp_elems = 0;

  // dual maps
  std::vector<int32_t> d_lam;      // (dvec_elems) lambda index per D slot, -1 on padding
  std::vector<double> d_sgn;       // (dvec_elems) +1 lower subdomain, -1 higher, 0 padding
  std::vector<int64_t> lam_pos;    // (n_lam, 2) D-slot of the two sides
  std::vector<int32_t> c_loc;      // (cvec_elems) global primal index per local primal slot, -1 padding
  std::vector<int64_t> c_gather_ptr;  // (n_c+1) CSR over local primal slots per coarse dof
  std::vector<int64_t> c_gather_slot; // local primal slot ids (cvec space)

  // global vector recovery: for each free dof where its value comes from
  //   kind 0: interior  -> vec[a]      (a = index in the local-vector space)
  //   kind 1: dual      -> ubar[a]     (a = lambda index; mean of the two sides)
  //   kind 2: primal    -> uc[a]       (a = coarse index)
  std::vector<int8_t> free_kind;
  std::vector<int64_t> free_src_a, free_src_b;
};

struct MeshInput {
  int64_t n_nodes, n_tets;
  const int32_t* tet2node;
  const int32_t* tet_subdomain;
  int64_t n_pec;
  const int32_t* pec_pairs;
  const double* xyz = nullptr;  // node coordinates [3 n_nodes]: when given, every subdomain's
                                // interior dofs are ordered in layers across its longest
                                // extent (see order_subdomain_dofs)
};

// Returns empty string on success, else an error message.
std::string build_topology(const MeshInput& in, Topology& topo);

// Stable LSD radix sort of 64-bit keys carrying 32-bit payloads.
void radix_sort_pairs(std::vector<uint64_t>& key, std::vector<uint32_t>& val, uint64_t max_key);

}  // namespace emag

// Launch wrappers for the CDNA4 kernels of the FETI-DP Nedelec solver.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "topology.h"

namespace emag {


// Element integration: per tet the packed upper triangle (21) of
// K_e (+ M_e when mode == 1) and of M_e (mode 0), and F_e (6) if jq != null.
// Output arrays are structure-of-arrays: entry k of tet t at [k * nt + t].
void launch_integrate(int64_t nt, const double* xyz, const int32_t* t2n, const int32_t* mat,
                      const double* nu, const double* kappa, const double* jq, int mode,
                      double* outA, double* outM, double* outF, hipStream_t st);

// out[target[i]] = sum_{c in ptr[i]..ptr[i+1]} (+-) elem[src[c] >> 1]  (sign = src & 1)
void launch_assemble_f64(int64_t n, const int64_t* target, const int64_t* ptr, const int32_t* src,
                         const double* elem, double* out, hipStream_t st);
void launch_assemble_f32(int64_t n, const int64_t* target, const int64_t* ptr, const int32_t* src,
                         const double* elem, float* out, hipStream_t st);

// out[target[i]] = sum_{c} pool[srcpos[c]]  (coarse assembly from factor pools)
template <typename T>
void launch_gather_pool(int64_t n, const int64_t* target, const int64_t* ptr, const int64_t* srcpos,
                        const T* pool, const uint8_t* tflag, double* out, hipStream_t st);
// Clear the pool tiles listed by index (elements 4096 t .. 4096 t + 4095 of tile t).
template <typename T>
void launch_zero_tiles(int64_t n, const int32_t* list, T* pool, hipStream_t st);

// mode 0: Schur update list; 1: column step below the diagonal (update + trsm with
// the stored inverse diagonal tile + diagonal-tile share); 2: diagonal Cholesky + inverse.
// Inverse diagonal tile of column J of matrix m: inv + inv_base[m] + J * 4096.
// Matrix m's tile (I, J) lives at pool + rowabs[rowbase[m] + I] + J * 4096.
// Optional forward substitution carried by the factorization: with b != nullptr,
// every diagonal factorization (modes 2, 3) of column J of matrix m writes
// y_J = L_JJ^-1 b_J to y + off[m] + 64 J, and every column-step tile (modes 1, 3)
// applies b_I -= L_IJ y_J to b + off[m] + 64 I as soon as L_IJ is formed -- the
// sweep of launch_forward without reading the factor again.
struct RhsArgs {
  double* b = nullptr;
  double* y = nullptr;
  const int64_t* off = nullptr;
};
// Operator stores written by the Schur updates themselves.  A MODE-0 item with flags & 128
// (a tile of the D block, after its last interior product) also stores its result as
// tile (I - o_d/64, J - o_d/64) of the subdomain's lower-tile store P (row-major inside a
// tile, diagonal tiles from their lower half), FP32 when p32.  One with flags & 32 (a
// tile of a D' row, after its last dual product) stores the negated result into W: a D'D'
// tile as lower tile (I - o_a/64, J - o_a/64) of Z, a D'c tile as the 64x64 block
// (J - o_c/64, I - o_a/64) of the row-major Phi^T that follows Z.
struct SnapArgs {
  const DevSub* subs = nullptr;
  void* P = nullptr;
  int p32 = 0;
  void* W = nullptr;
};
template <typename T>
void launch_tile_mode(int mode, const TileWork* w, int64_t n, const int64_t* rowbase,
                      const int64_t* rowabs, const int64_t* inv_base, T* pool, T* inv, int* fail,
                      hipStream_t st, int waves = 4, RhsArgs rhs = RhsArgs(), int lds_pad = 0,
                      SnapArgs snap = SnapArgs());

// MFMA lane-layout probe (one wave, 512 doubles of output).
void launch_mfma_probe(int is_f32, double* out);

// P tiles (the D block of the pool) and W = [Z tiles, Phi^T rows] (the D' block and
// the (D', c) block, negated) from the partially factored pools; `tiles` lists every lower
// tile of every subdomain's D x D store.
template <typename T, typename TP>
void launch_extract_P(int64_t ntiles, const SymTile* tiles, const DevSub* subs, const T* pool,
                      const int64_t* rowabs, const int32_t* ftile, const uint8_t* tflag, TP* P,
                      hipStream_t st);
template <typename T>
void launch_extract_W(int64_t ntiles, const SymTile* tiles, int ns, int max_ncp, int max_ndp,
                      const DevSub* subs, const T* pool, const int64_t* rowabs, const int32_t* ftile,
                      const uint8_t* tflag, T* W, hipStream_t st);
// y = S~^-1 x from the factored coarse pool (its trailing block holds -S~^-1 as lower
// tiles): tiles lists the tn(tn+1)/2 tile pairs (I, J <= I) of that block, tn = ncp / 64;
// part holds tn * tn * 64 doubles.
void launch_coarse_symv(int64_t nc, int tn, int64_t ntiles, const int32_t* tiles, const double* cpool,
                        const int64_t* cra, const double* x, double* part, double* y, hipStream_t st);

// Grouped coarse solve (S~ = L L^T, L in the coarse pool; groups of tile columns
// [a, a + w) with X_g = (L_gg)^-1 dense column-major of order 64 w):
//  * launch_cg_invert forms every X_g after the factorization (one launch per distance
//    from the diagonal, max_group launches);
//  * S~^-1 v = L^-T L^-1 v by a forward sweep (groups in order: y_g = X_g (v_g -
//    L_{g, <a} y)) and a backward sweep (groups in reverse: x_g = X_g^T (y_g -
//    L_{>g, g}^T x)), two launches per group:
//    launch_cg_step_a: the output blocks of the previous group [pa, pa + pw) from xpart
//      into outv (and out2 for entries below nc; may be null), plus the panel products
//      of this group (npanel tiles, partial 64-vectors into part; vector blocks from vsrc,
//      or from xpart when they lie in the previous group);
//    launch_cg_step_b: the X_g (trans: X_g^T) products of the group [a, a + w), each on
//      its right-hand side block base - (panel slots [ptr[J], ptr[J+1]) of part) (base
//      entries past nb read as zero), partials into xpart (w(w+1)/2 64-vectors).
//    A step A with npanel = 0 writes the last group's output.
void launch_cg_invert(int max_group, const int64_t* step, const int32_t* items, const double* cpool,
                      const int64_t* cra, const int32_t* cft, const double* cinv,
                      const int32_t* cg_start, const int32_t* cg_group_of, const int64_t* xoff,
                      double* X, hipStream_t st);
void launch_cg_step_a(bool trans, int pa, int pw, const double* xpart, double* outv, double* out2,
                      int64_t nc, const int32_t* panel_tiles, int64_t npanel, const double* cpool,
                      const int64_t* cra, const double* vsrc, double* part, hipStream_t st);
void launch_cg_step_b(bool trans, int a, int w, const double* Xg, const double* base, int64_t nb,
                      const int32_t* ptr, const double* part, double* xpart, hipStream_t st);

// Forward solve, one launch per block column: yv_r = L_rr^-1 vec_r; afterwards
// vec_c = f_c - L_cr y and vec_D' = -L_D'r y (= -E K_rr^-1 f_r).
template <typename T>
void launch_forward(int ns, int max_T, int max_Tr, const DevSub* subs, const T* pool,
                    const int64_t* rowabs, const T* inv, const int64_t* inv_base,
                    const int32_t* ftile, double* vec, double* yv, hipStream_t st, int skip_int = 0,
                    int max_dual = 0, int max_after = 0);
// skip_int: the interior blocks of vec are zero and those of yv already hold zeros, so only
// the dual columns are swept: step J takes every subdomain's dual column Ti + J at once
// (max_dual = max (Tr - Ti) steps, max_after = max (T - Ti) rows below the first one).
// Zero the interior blocks (tile rows < Ti) of a local vector.
void launch_zero_interior(int ns, int max_Ti, const DevSub* subs, double* v, hipStream_t st);
// kDotBlocks partial sums of x_d^2 over the free dofs d of kind `want` (0 interior).
void launch_kind_sq(int64_t n, const int8_t* kind, int want, const double* x, double* partial,
                    hipStream_t st);
// Interior solve: xout_i = L_ii^-T (yv_i - L_{D,c; i}^T U); tbuf is scratch (vec space).
// width (host, max_Ti entries): the backward step of row J launches width[J] work-groups
// per subdomain (rows J - 1, J - 2, ...).
template <typename T>
void launch_interior(int ns, int max_Ti, const int32_t* width, const DevSub* subs, const T* pool,
                     const int64_t* rowabs, const T* inv, const int64_t* inv_base,
                     const int32_t* ftile, const double* yv, const double* U, double* tbuf,
                     double* xout, hipStream_t st);
// rho[slot] = pool diagonal at each D slot; stiffness scaling weights and 0.5 means.
void launch_scaling(int64_t ndv, int64_t nl, const int64_t* rho_pos, const void* pool, int is_f32,
                    double* rho, const int64_t* lam_pos, const double* sgn, double* sgnD,
                    double* half, hipStream_t st);

// y_s = M_s x_s with x_s[k] = scale(slot) * v[lam[slot]], rows_s = ndp (+ ncp if with_c)
// Phi^T rows of W times the gathered dual vector (out at y_off + ndp)
template <typename T>
void launch_gemv_phi(int ns, int max_ncp, const DevSub* subs, const T* W, const int32_t* d_lam,
                     const double* d_scale, const double* v, double* y, hipStream_t st, int max_ndp);
// symmetric operator (which 0: Z of W, out at y_off; which 1: P, out at dvec_off); the
// per-tile partials stay in part, and with out null their sum is left to the consumer
template <typename T>
void launch_symv(int64_t ntiles, const SymTile* tiles, int ns, int max_Td, const DevSub* subs,
                 const T* M, int which, const int32_t* d_lam, const double* d_scale,
                 const double* v, double* part, double* out, hipStream_t st);
// W x of every subdomain in one launch: Phi^T x into y (at y_off + ndp), Z x as the
// symmetric-product partials in part (summed by launch_dual_correct)
template <typename T>
void launch_apply_W(int64_t ntiles, const SymTile* tiles, int ns, int max_ncp, int max_ndp,
                    const DevSub* subs, const T* W, const int32_t* d_lam, const double* d_scale,
                    const double* v, double* part, double* y, hipStream_t st);
// out[slot] = alpha*(Z x)[k] + beta*(Phi u)[k] + gamma*vec[o_a + k], (Z x) summed from the
// symmetric-product partials zpart of launch_symv(which 0)
template <typename T>
void launch_dual_correct(int ns, int max_ndp, int max_ncp, const DevSub* subs, const T* W, const double* zpart,
                         const int32_t* c_loc, const double* uc, const double* vec, double alpha,
                         double beta, double gamma, double* out, hipStream_t st);
// out[l] = sum over the two sides of scale[slot] * w[slot]
void launch_lambda_scatter(int64_t nl, const int64_t* lam_pos, const double* scale, const double* w,
                           double* out, hipStream_t st);
void launch_coarse_gather(int64_t nc, const int64_t* ptr, const int64_t* pos, const double* src,
                          const double* add, double* out, hipStream_t st);
void launch_fill_U(int ns, int max_m, const DevSub* subs, const int32_t* d_lam, const double* ubar,
                   const int32_t* c_loc, const double* uc, double* U, hipStream_t st);
void launch_split_rhs(int64_t n, const int32_t* vdof, const double* vw, const double* f, double* vec,
                      hipStream_t st);
void launch_gather_global(int64_t nf, const int8_t* kind, const int64_t* a, const double* vec,
                          const double* ubar, const double* uc, double* x, hipStream_t st);
void launch_csr_residual(int64_t n, const int32_t* rp, const int32_t* col, const double* val,
                         const double* x, const double* f, double* r, hipStream_t st);
void launch_csr_spmv(int64_t n, const int32_t* rp, const int32_t* col, const double* val,
                     const double* x, double* y, hipStream_t st);
// partial[b] = sum over block b of a[i]*b[i] for up to 3 pairs; nblk partial sums each
constexpr int kDotBlocks = 256;
void launch_dots(int64_t n, int npairs, const double* const* a, const double* const* b,
                 double* partial, hipStream_t st);
void launch_axpy2(int64_t n, double alpha, const double* p, const double* q, double* lam, double* r,
                  hipStream_t st);
void launch_xpby(int64_t n, const double* z, double beta, double* p, hipStream_t st);
void launch_axpy(int64_t n, double alpha, const double* x, double* y, hipStream_t st);
void launch_scale2(int64_t n, double alpha, double* x, double* y, hipStream_t st);
// deluxe scaling (FP64 Dirichlet blocks P)
void launch_face_setup(int64_t nfaces, int max_tf, const FaceDev* faces, const DevSub* subs,
                       const double* P, const int64_t* slot0, const int64_t* slot1, double* fpool,
                       hipStream_t st);
void launch_face_extract(int64_t nfaces, int max_tf, const FaceDev* faces, const DevSub* subs,
                         const double* P, const int64_t* slot1, const double* fpool, double* ops,
                         hipStream_t st);
// D_F = A_F^-1 B_F of every face from the lower-tile stores (FP64 product, rounded to T).
template <typename T>
void launch_face_D(int64_t nfaces, int max_tf, const FaceDev* faces, const double* ops, T* D,
                   hipStream_t st);
// out[vec_off + k] = (D_F x)_k, or (D_F^T x)_k when trans, with x_k = src[map[vec_off + k]]
// (0 where map < 0), on every face.
template <typename T>
void launch_face_dmv(bool trans, int64_t nfaces, int max_tf, const FaceDev* faces, const T* D,
                     const int32_t* map, const double* src, double* out, hipStream_t st);
// Deluxe average of the two sides' values u (D-slot space) on every face:
// jmp = u_s1 - u_s0 (face-vector space, 0 on padding); ubar[l] = u_s0 + g with g = D_F jmp.
void launch_deluxe_jump(int64_t nv, const int32_t* map_lam, const int64_t* slot0, const int64_t* slot1,
                        const double* u, double* jmp, hipStream_t st);
void launch_deluxe_avg(int64_t nv, const int32_t* map_lam, const int64_t* slot0, const double* u,
                       const double* g, double* ubar, hipStream_t st);
void launch_deluxe_split(int64_t nv, const int32_t* map_lam, const int64_t* slot0, const int64_t* slot1,
                         const double* g, const double* r, double* xs, hipStream_t st);
void launch_deluxe_sum(int64_t nv, const int32_t* map_lam, const int64_t* slot0, const int64_t* slot1,
                       const double* y, double* u, hipStream_t st);
void launch_deluxe_out(int64_t nv, const int32_t* map_lam, const int64_t* slot1, const double* h,
                       const double* y, double* z, hipStream_t st);
// PCPG device scalars (slots of the scalar block)
enum { kScRZ = 0, kScPQ = 1, kScAlpha = 2, kScBeta = 3, kScRR = 4, kScRZold = 5, kScSeq = 6,
       kScCount = 8 };
// q = lambda scatter of w fused with the k_dots partials of npairs pairs (null b[k]: the
// pair's second vector is q)
void launch_scatter_dot(int64_t nl, const int64_t* lam_pos, const double* scale, const double* w,
                        double* q, int npairs, const double* const* a, const double* const* b,
                        double* partial, hipStream_t st);
// k_pcpg_alpha + k_axpy2_dev and k_pcpg_beta + k_xpby_dev as single launches; the
// latter bumps the sequence slot kScSeq after publishing the iteration's scalars to host
void launch_alpha_axpy2(int64_t n, const double* partial, double* sc, double* host, const double* p,
                        const double* q, double* lam, double* r, hipStream_t st);
void launch_beta_xpby(int64_t n, const double* partial, double* sc, double* host, const double* z,
                      double* p, hipStream_t st);
void launch_pcpg_alpha(const double* partial, double* sc, double* host, hipStream_t st);
void launch_pcpg_beta(const double* partial, double* sc, double* host, int init, hipStream_t st);
void launch_axpy2_dev(int64_t n, const double* sc, const double* p, const double* q, double* lam,
                      double* r, hipStream_t st);
void launch_xpby_dev(int64_t n, const double* z, const double* sc, double* p, hipStream_t st);

}  // namespace emag

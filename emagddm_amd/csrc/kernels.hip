// CDNA4 (gfx950) kernels of the FETI-DP Nedelec solver.
//
// Data layout in HBM
//  * element data: structure-of-arrays, entry k of tet t at [k * nt + t]
//    (coalesced per-thread-per-tet writes); one trailing 1.0 constant.
//  * factor pool: per matrix, the tiles of the lower triangle that lie inside the
//    envelope (J >= f_I), packed by tile row: tile (I, J) at rowabs[I] + J * 4096,
//    each 64x64 tile column-major; a tile row is one contiguous panel.
//  * iteration operators W_s = [Z; Phi^T] and P_s = S_DD: row-major dense.
//  * vectors: FP64 always.
//
// MFMA lane layout used (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32):
//   A operand  lane l -> A[l % 16][l / 16]           (16 x 4)
//   B operand  lane l -> B[l / 16][l % 16]           (4 x 16)
//   D          lane l, reg r -> D[Mfma<T>::row(l / 16, r)][l % 16]
//              (f64: row l/16 + 4r; f32: row 4(l/16) + r -- see Mfma<T>)
#include "kernels.h"

#include <hip/hip_runtime.h>

namespace emag {


// ---------------------------------------------------------------------------
// Element integration (closed form, FP64)
// ---------------------------------------------------------------------------
__constant__ int c_ref_edge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

__global__ __launch_bounds__(256) void k_integrate(int64_t nt, const double* __restrict__ xyz,
                                                   const int32_t* __restrict__ t2n,
                                                   const int32_t* __restrict__ mat,
                                                   const double* __restrict__ nu,
                                                   const double* __restrict__ kappa,
                                                   const double* __restrict__ jq, int mode,
                                                   double* __restrict__ outA,
                                                   double* __restrict__ outM,
                                                   double* __restrict__ outF) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int4 nd = reinterpret_cast<const int4*>(t2n)[t];
  const int ids[4] = {nd.x, nd.y, nd.z, nd.w};
  double X[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) X[a][d] = xyz[3 * (int64_t)ids[a] + d];
  // columns of J: c_k = x_k - x_3; rows of J^-1: (c1 x c2, c2 x c0, c0 x c1) / det
  double c[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) c[k][d] = X[k][d] - X[3][d];
  double G[4][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double* u = c[(k + 1) % 3];
    const double* v = c[(k + 2) % 3];
    G[k][0] = u[1] * v[2] - u[2] * v[1];
    G[k][1] = u[2] * v[0] - u[0] * v[2];
    G[k][2] = u[0] * v[1] - u[1] * v[0];
  }
  const double det = c[0][0] * G[0][0] + c[0][1] * G[0][1] + c[0][2] * G[0][2];
  const double idet = 1.0 / det;
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) G[k][d] *= idet;
#pragma unroll
  for (int d = 0; d < 3; ++d) G[3][d] = -(G[0][d] + G[1][d] + G[2][d]);
  const double vol = fabs(det) * (1.0 / 6.0);
  const int m = mat[t];
  const double ka = kappa[m] * vol * (1.0 / 20.0);
  const double kn = nu[m] * vol;
  double cu[6][3];
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const double* ga = G[c_ref_edge[e][0]];
    const double* gb = G[c_ref_edge[e][1]];
    cu[e][0] = 2.0 * (gb[1] * ga[2] - gb[2] * ga[1]);
    cu[e][1] = 2.0 * (gb[2] * ga[0] - gb[0] * ga[2]);
    cu[e][2] = 2.0 * (gb[0] * ga[1] - gb[1] * ga[0]);
  }
  double g[4][4];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) g[p][q] = G[p][0] * G[q][0] + G[p][1] * G[q][1] + G[p][2] * G[q][2];
  int k = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
#pragma unroll
    for (int j = i; j < 6; ++j, ++k) {
      const int a = c_ref_edge[i][0], b = c_ref_edge[i][1];
      const int cc = c_ref_edge[j][0], d = c_ref_edge[j][1];
      const double kij = kn * (cu[i][0] * cu[j][0] + cu[i][1] * cu[j][1] + cu[i][2] * cu[j][2]);
      const double mij = ka * ((1 + (b == d)) * g[a][cc] - (1 + (b == cc)) * g[a][d] -
                               (1 + (a == d)) * g[b][cc] + (1 + (a == cc)) * g[b][d]);
      if (mode == 1) {
        outA[k * nt + t] = kij + mij;
      } else {
        outA[k * nt + t] = kij;
        outM[k * nt + t] = mij;
      }
    }
  }
  if (jq != nullptr) {
    const double qa = 0.1381966011250105151795413165634361882280;  // (5 - sqrt 5) / 20
    const double qb = 0.5854101966249684544613760503096914353161;  // (5 + 3 sqrt 5) / 20
    const double* J = jq + 12 * t;
    double F[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double jx = J[3 * q], jy = J[3 * q + 1], jz = J[3 * q + 2];
      double gj[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) gj[p] = G[p][0] * jx + G[p][1] * jy + G[p][2] * jz;
#pragma unroll
      for (int e = 0; e < 6; ++e) {
        const int a = c_ref_edge[e][0], b = c_ref_edge[e][1];
        const double la = (a == q) ? qb : qa, lb = (b == q) ? qb : qa;
        F[e] += lb * gj[a] - la * gj[b];
      }
    }
#pragma unroll
    for (int e = 0; e < 6; ++e) outF[e * nt + t] = F[e] * (vol * 0.25);
  }
}

void launch_integrate(int64_t nt, const double* xyz, const int32_t* t2n, const int32_t* mat,
                      const double* nu, const double* kappa, const double* jq, int mode,
                      double* outA, double* outM, double* outF, hipStream_t st) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_integrate, dim3((nt + 255) / 256), dim3(256), 0, st, nt, xyz, t2n, mat, nu,
                     kappa, jq, mode, outA, outM, outF);
}

// ---------------------------------------------------------------------------
// Deterministic assembly: each output entry sums its contribution list in order
// ---------------------------------------------------------------------------
template <typename TO>
__global__ __launch_bounds__(256) void k_assemble(int64_t n, const int64_t* __restrict__ target,
                                                  const int64_t* __restrict__ ptr,
                                                  const int32_t* __restrict__ src,
                                                  const double* __restrict__ elem, TO* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int64_t c = ptr[i]; c < ptr[i + 1]; ++c) {
    const int32_t code = src[c];
    const double v = elem[(uint32_t)code >> 1];
    s += (code & 1) ? -v : v;
  }
  out[target ? target[i] : i] = (TO)s;
}

void launch_assemble_f64(int64_t n, const int64_t* target, const int64_t* ptr, const int32_t* src,
                         const double* elem, double* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_assemble<double>, dim3((n + 255) / 256), dim3(256), 0, st, n, target, ptr,
                     src, elem, out);
}
void launch_assemble_f32(int64_t n, const int64_t* target, const int64_t* ptr, const int32_t* src,
                         const double* elem, float* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_assemble<float>, dim3((n + 255) / 256), dim3(256), 0, st, n, target, ptr,
                     src, elem, out);
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather_pool(int64_t n, const int64_t* __restrict__ target,
                                                     const int64_t* __restrict__ ptr,
                                                     const int64_t* __restrict__ srcpos,
                                                     const T* __restrict__ pool,
                                                     const uint8_t* __restrict__ tflag,
                                                     double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int64_t c = ptr[i]; c < ptr[i + 1]; ++c) {
    const int64_t p = srcpos[c];
    s += p < 0 ? 1.0 : ((tflag && !tflag[p / kTileElems]) ? 0.0 : (double)pool[p]);
  }
  out[target[i]] = s;
}

template <typename T>
void launch_gather_pool(int64_t n, const int64_t* target, const int64_t* ptr, const int64_t* srcpos,
                        const T* pool, const uint8_t* tflag, double* out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather_pool<T>, dim3((n + 255) / 256), dim3(256), 0, st, n, target, ptr,
                     srcpos, pool, tflag, out);
}
template void launch_gather_pool<double>(int64_t, const int64_t*, const int64_t*, const int64_t*,
                                         const double*, const uint8_t*, double*, hipStream_t);
template void launch_gather_pool<float>(int64_t, const int64_t*, const int64_t*, const int64_t*,
                                        const float*, const uint8_t*, double*, hipStream_t);

// Clear the listed 64x64 tiles of a pool (tile t = elements 4096 t .. 4096 t + 4095).
// Each work-group clears kZeroGroup consecutive list entries with 16-byte stores (tile
// bases are 16-byte aligned: the pool comes from hipMalloc and tiles are 16 KB or
// 32 KB), so the launch is not bound by work-group dispatch.
constexpr int kZeroGroup = 4;
template <typename T>
__global__ __launch_bounds__(256) void k_zero_tiles(int64_t n, const int32_t* __restrict__ list,
                                                    T* __restrict__ pool) {
  constexpr int n4 = kTileElems * int(sizeof(T)) / 16;
  const int64_t t0 = int64_t(blockIdx.x) * kZeroGroup;
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int k = 0; k < kZeroGroup && t0 + k < n; ++k) {
    float4* t = reinterpret_cast<float4*>(pool + (int64_t)list[t0 + k] * kTileElems);
#pragma unroll
    for (int e = threadIdx.x; e < n4; e += 256) t[e] = z;
  }
}
template <typename T>
void launch_zero_tiles(int64_t n, const int32_t* list, T* pool, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_tiles<T>, dim3((n + kZeroGroup - 1) / kZeroGroup), dim3(256), 0, st, n, list,
                     pool);
}
template void launch_zero_tiles<double>(int64_t, const int32_t*, double*, hipStream_t);
template void launch_zero_tiles<float>(int64_t, const int32_t*, float*, hipStream_t);

// elements of a lower-tile symmetric store of order n (device copy of sym_tile_elems)
__device__ __forceinline__ int64_t sym_elems(int64_t n) {
  const int64_t t = n >> 6;
  return t * (t + 1) / 2 * kTileElems;
}

// ---------------------------------------------------------------------------
// Batched tile Cholesky on MFMA
// ---------------------------------------------------------------------------
template <typename T>
struct Mfma;
// row(lk, r): accumulator row of register r in lane group lk = lane / 16 (the
// column is lane % 16).  Measured on gfx950 by emag_mfma_layout_probe:
// f64 interleaves rows across lane groups, f32 keeps 4 consecutive rows per lane.
template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

constexpr int kLd = 65;  // padded LDS leading dimension for a 64x64 tile

// sum_{j<n} a[j*as] * b[j*bs] with four independent accumulators so that the LDS
// loads of consecutive terms overlap instead of serializing on load latency.
template <typename T>
__device__ __forceinline__ T dot_strided(const T* a, int as, const T* b, int bs, int n) {
  T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
  int j = 0;
#pragma unroll 2
  for (; j + 4 <= n; j += 4) {
    s0 += a[j * as] * b[j * bs];
    s1 += a[(j + 1) * as] * b[(j + 1) * bs];
    s2 += a[(j + 2) * as] * b[(j + 2) * bs];
    s3 += a[(j + 3) * as] * b[(j + 3) * bs];
  }
  for (; j < n; ++j) s0 += a[j * as] * b[j * bs];
  return (s0 + s1) + (s2 + s3);
}

// Lane-layout probe of the 16x16x4 MFMA builtins: with A[p][0] = p (A[p][k>0] = 0)
// and B[0][q] = 1 the product holds D[p][q] = p, with A[p][0] = 1 and B[0][q] = q
// it holds q; each lane writes its 4 accumulator values for both products.
// out: [2 products][64 lanes][4 regs], one wave, all indices in bounds.
template <typename T>
__global__ void k_mfma_probe(double* out) {
  using acc_t = typename Mfma<T>::acc_t;
  const int l = threadIdx.x;
  const int lr = l & 15, lk = l >> 4;
  acc_t z = acc_t{0, 0, 0, 0};
  acc_t d0 = Mfma<T>::op(lk == 0 ? T(lr) : T(0), lk == 0 ? T(1) : T(0), z);
  acc_t d1 = Mfma<T>::op(lk == 0 ? T(1) : T(0), lk == 0 ? T(lr) : T(0), z);
  for (int r = 0; r < 4; ++r) {
    out[(0 * 64 + l) * 4 + r] = (double)d0[r];
    out[(1 * 64 + l) * 4 + r] = (double)d1[r];
  }
}

void launch_mfma_probe(int is_f32, double* out) {
  if (is_f32)
    hipLaunchKernelGGL(k_mfma_probe<float>, dim3(1), dim3(64), 0, nullptr, out);
  else
    hipLaunchKernelGGL(k_mfma_probe<double>, dim3(1), dim3(64), 0, nullptr, out);
}

// acc (2x2 MFMA blocks per wave: rows p0 + 0..15 and p1 + 0..15, columns qb0 + 0..31)
// += A_p,k * B_k,q over one 64-deep tile (-= with NEG), A operand read from Ap
// (column-major 64x64, element (p,k) at k*lda + p), B operand from Bq (element (q,k) at
// k*ldb + q).
template <typename T, int LDA, int LDB, bool NEG = false>
__device__ __forceinline__ void tile_mma(typename Mfma<T>::acc_t (&acc)[2][2], const T* Ap,
                                         const T* Bq, int p0, int p1, int qb0, int lr, int lk) {
#pragma unroll 4
  for (int kk = 0; kk < 64; kk += 4) {
    const int kc = kk + lk;
    const T a0 = NEG ? -Ap[kc * LDA + p0 + lr] : Ap[kc * LDA + p0 + lr];
    const T a1 = NEG ? -Ap[kc * LDA + p1 + lr] : Ap[kc * LDA + p1 + lr];
    const T b0 = Bq[kc * LDB + qb0 + lr];
    const T b1 = Bq[kc * LDB + qb0 + 16 + lr];
    acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
    acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
    acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
    acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
  }
}

// The same product when the A tile is lower triangular (A(p, k) = 0 for k > p, as a
// stored inverse diagonal tile is, with exact zeros): the 16-row block b of a wave needs
// k < 16 (b + 1) only, so the steps past that are left out -- the products they would
// add are exact zeros.  The waves hold row blocks {0, 3} or {1, 2} (see tile_update), so
// each issues 40 of its 64 MFMAs; the same terms in the same order as tile_mma.
template <typename T, int LDA, int LDB, int B0, int B1>
__device__ __forceinline__ void tile_mma_lower_a_at(typename Mfma<T>::acc_t (&acc)[2][2],
                                                    const T* Ap, const T* Bq, int qb0, int lr,
                                                    int lk) {
#pragma unroll
  for (int kk = 0; kk < 16 * (B0 + 1); kk += 4) {
    const int kc = kk + lk;
    const T a0 = Ap[kc * LDA + 16 * B0 + lr];
    const T a1 = Ap[kc * LDA + 16 * B1 + lr];
    const T b0 = Bq[kc * LDB + qb0 + lr];
    const T b1 = Bq[kc * LDB + qb0 + 16 + lr];
    acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
    acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
    acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
    acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
  }
#pragma unroll
  for (int kk = 16 * (B0 + 1); kk < 16 * (B1 + 1); kk += 4) {
    const int kc = kk + lk;
    const T a1 = Ap[kc * LDA + 16 * B1 + lr];
    const T b0 = Bq[kc * LDB + qb0 + lr];
    const T b1 = Bq[kc * LDB + qb0 + 16 + lr];
    acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
    acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
  }
}
template <typename T, int LDA, int LDB>
__device__ __forceinline__ void tile_mma_lower_a(typename Mfma<T>::acc_t (&acc)[2][2],
                                                 const T* Ap, const T* Bq, int p0, int qb0,
                                                 int lr, int lk) {
  if (p0 == 0)  // uniform per wave: row blocks {0, 3}, else {1, 2}
    tile_mma_lower_a_at<T, LDA, LDB, 0, 3>(acc, Ap, Bq, qb0, lr, lk);
  else
    tile_mma_lower_a_at<T, LDA, LDB, 1, 2>(acc, Ap, Bq, qb0, lr, lk);
}

__device__ __forceinline__ double readlane(double v, int lane) {
  const unsigned long long u = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
  return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float readlane(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Blocked Cholesky of the 64x64 tile in s (LDS, column-major, ld kLd) by the whole
// work-group (call uniformly).  Panel b (columns 16b .. 16b+15): wave 0 factors it
// over all 64 rows (lane = row; each pivot and multiplier reaches every lane by
// readlane, and the panel's later columns take their update right after each pivot,
// so only the pivot chain is serial); the four waves then subtract the panel's
// rank-16 product from the trailing 16x16 blocks on MFMA.  On return s holds L in
// its lower triangle (zeros above the diagonal inside the diagonal blocks; the blocks
// above them are untouched) and rdg 1/diag.  Returns 1 (wave 0) when a pivot was not
// positive and finite; that column then uses pivot 1.
template <typename T>
__device__ __forceinline__ int tile_potrf(T* s, T* rdg, int tid) {
  using acc_t = typename Mfma<T>::acc_t;
  const int lane = tid & 63, wave = tid >> 6, lr = lane & 15, lk = lane >> 4;
  int bad = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int c0 = 16 * b;
    if (wave == 0) {
      T r[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) r[u] = s[(c0 + u) * kLd + lane];
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const T piv = readlane(r[k], c0 + k);
        const bool ok = piv > T(0) && piv < T(3.0e38);
        bad |= !ok;
        // rsqrt of the pivot or of a stand-in, then selects: no branch in the pivot chain
        const T rq = rsqrt(ok ? piv : T(1));
        const T rd = ok ? rq : T(1);
        const T d = ok ? piv * rq : T(1);
        r[k] = lane == c0 + k ? d : (lane > c0 + k ? r[k] * rd : T(0));
        rdg[c0 + k] = rd;  // the same value from every lane
#pragma unroll
        for (int j = k + 1; j < 16; ++j) r[j] -= r[k] * readlane(r[k], c0 + j);
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) s[(c0 + u) * kLd + lane] = r[u];
    }
    __syncthreads();
    if (b < 3) {
      // trailing blocks (I, J), b < J <= I <= 3, column by column; wave w takes
      // blocks w and w + 4.  D[p][q] = sum_k L(16J+p, k) L(16I+q, k) is the update
      // of C(16I+q, 16J+p).
      constexpr int nblk_tab[3] = {6, 3, 1};
      for (int m = wave; m < nblk_tab[b]; m += 4) {
        int J = b + 1, mm = m;
        while (mm >= 4 - J) {
          mm -= 4 - J;
          ++J;
        }
        const int I = J + mm;
        acc_t acc;
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] = s[(16 * J + Mfma<T>::row(lk, r)) * kLd + 16 * I + lr];
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
          const int kc = c0 + kk + lk;
          acc = Mfma<T>::op(-s[kc * kLd + 16 * J + lr], s[kc * kLd + 16 * I + lr], acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) s[(16 * J + Mfma<T>::row(lk, r)) * kLd + 16 * I + lr] = acc[r];
      }
      __syncthreads();
    }
  }
  return bad;
}

// In-place inverse of the lower-triangular factor in s (as left by tile_potrf; rdg
// holds 1/diag) by the whole work-group (call uniformly).  Diagonal blocks: wave w
// inverts block w by forward substitution (lane = column).  Block columns then go
// right to left, b = 2, 1, 0: with the trailing inverse Y already in place, block
// (I, b) of the inverse is -(sum_{K=b+1}^{I} Y_IK L_Kb) Y_bb, wave I-b-1 forming the
// sum on MFMA, staging it in tmp (3 x 16 x 17 elements of LDS) and multiplying by
// Y_bb; the blocks are written once every wave has read column b of L.  On return
// the lower triangle of s holds L^-1 (blocks above the diagonal blocks untouched).
template <typename T>
__device__ __forceinline__ void tile_trtri(T* s, const T* rdg, T* tmp, int tid) {
  using acc_t = typename Mfma<T>::acc_t;
  const int lane = tid & 63, wave = tid >> 6, lr = lane & 15, lk = lane >> 4;
  {
    const int c0 = 16 * wave;
    T y[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      T acc = T(0);
#pragma unroll
      for (int k = 0; k < i; ++k) acc += s[(c0 + k) * kLd + c0 + i] * y[k];
      const T rdi = rdg[c0 + i];
      y[i] = i < lr ? T(0) : (i == lr ? rdi : -acc * rdi);
    }
    wave_sync_lds();  // every lane has read block w of L
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if ((i >> 2) == lk) s[(c0 + lr) * kLd + c0 + i] = y[i];
  }
  __syncthreads();
#pragma unroll
  for (int b = 2; b >= 0; --b) {
    acc_t col = acc_t{0, 0, 0, 0};
    const int I = b + 1 + wave;
    if (I <= 3) {
      // W(i, c) = sum_K sum_k Y(16I+i, 16K+k) L(16K+k, 16b+c): p = i, q = c
      acc_t w = acc_t{0, 0, 0, 0};
      for (int K = b + 1; K <= I; ++K)
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
          const int kc = kk + lk;
          w = Mfma<T>::op(s[(16 * K + kc) * kLd + 16 * I + lr], s[(16 * b + lr) * kLd + 16 * K + kc], w);
        }
      T* tw = tmp + wave * (16 * 17);
#pragma unroll
      for (int r = 0; r < 4; ++r) tw[Mfma<T>::row(lk, r) * 17 + lr] = w[r];  // tw[i*17 + c]
      wave_sync_lds();
      // block (I, b) = -W Y_bb: A(p=i, k) = -W(i, k), B(q=c, k) = Y(16b+k, 16b+c)
#pragma unroll
      for (int kk = 0; kk < 16; kk += 4) {
        const int kc = kk + lk;
        col = Mfma<T>::op(-tw[lr * 17 + kc], s[(16 * b + lr) * kLd + 16 * b + kc], col);
      }
    }
    __syncthreads();  // column b of L has been read by every wave
    if (I <= 3) {
#pragma unroll
      for (int r = 0; r < 4; ++r) s[(16 * b + lr) * kLd + 16 * I + Mfma<T>::row(lk, r)] = col[r];
    }
    __syncthreads();
  }
}

// acc -= A_panel * B_panel^T over ktot (multiple of 16) columns.  A row panel of
// tiles (J, k0 .. k1) is contiguous: element (p, k) at k*64 + p, so the K loop is
// one panel product.  64x16 chunks of both panels are staged through LDS (two
// stages, next chunk prefetched into registers), so each element crosses L2 once
// per workgroup; the four waves then read their MFMA operands from LDS.
template <typename T, int DEPTH = 1>
__device__ __forceinline__ void panel_mma(typename Mfma<T>::acc_t (&acc)[2][2],
                                          const T* __restrict__ Ap, const T* __restrict__ Bq,
                                          int ktot, T* sA, T* sB, int tid, int p0, int p1,
                                          int qb0, int lr, int lk) {
  // DEPTH chunks are in flight in registers (2: one more chunk's loads overlap a chunk's
  // products; every mode uses 2, which fits the column step's 128-register budget since
  // only the factoring items hold a diagonal accumulator)
  constexpr int KC = 16, CH = 64 * KC;
  const int nst = ktot / KC;
  if (nst == 0) return;
  T ra0[4], rb0[4], ra1[4], rb1[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    ra0[u] = Ap[tid + 256 * u];
    rb0[u] = Bq[tid + 256 * u];
  }
  if (DEPTH == 2 && nst > 1) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      ra1[u] = Ap[CH + tid + 256 * u];
      rb1[u] = Bq[CH + tid + 256 * u];
    }
  }
  auto step = [&](int c, T (&ra)[4], T (&rb)[4]) {
    T* a = sA + (c & 1) * CH;
    T* bb = sB + (c & 1) * CH;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[tid + 256 * u] = -ra[u];
      bb[tid + 256 * u] = rb[u];
    }
    __syncthreads();
    if (c + DEPTH < nst) {
      const int64_t o = (int64_t)(c + DEPTH) * CH;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        ra[u] = Ap[o + tid + 256 * u];
        rb[u] = Bq[o + tid + 256 * u];
      }
    }
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      const int kc = kk + lk;
      const T a0 = a[kc * 64 + p0 + lr];
      const T a1 = a[kc * 64 + p1 + lr];
      const T b0 = bb[kc * 64 + qb0 + lr];
      const T b1 = bb[kc * 64 + qb0 + 16 + lr];
      acc[0][0] = Mfma<T>::op(a0, b0, acc[0][0]);
      acc[0][1] = Mfma<T>::op(a0, b1, acc[0][1]);
      acc[1][0] = Mfma<T>::op(a1, b0, acc[1][0]);
      acc[1][1] = Mfma<T>::op(a1, b1, acc[1][1]);
    }
  };
  if constexpr (DEPTH == 1) {
    for (int c = 0; c < nst; ++c) step(c, ra0, rb0);
  } else {
    int c = 0;
    for (; c + 1 < nst; c += 2) {
      step(c, ra0, rb0);
      step(c + 1, ra1, rb1);
    }
    if (c < nst) step(c, ra0, rb0);
  }
  __syncthreads();
}

// Cholesky of the 64x64 tile staged in s (LDS, ld kLd) and the inverse of the
// factor, in place, by the whole work-group (call uniformly); L (upper part zeroed)
// goes to C, L^-1 to Li, and a non-positive pivot is flagged in *fail.  On return s
// holds L^-1 with zeros above the diagonal.  rdg: 64 elements of LDS, tmp: 816.
template <typename T>
__device__ __forceinline__ void factor_diag_tile(T* s, T* rdg, T* tmp, T* __restrict__ C,
                                                 T* __restrict__ Li, int* fail, int tid) {
  const int bad = tile_potrf<T>(s, rdg, tid);
  if (bad && tid == 0) atomicOr(fail, 1);
  for (int e = tid; e < kTileElems; e += 256) {
    const int i = e & 63, j = e >> 6;
    C[e] = (i >= j) ? s[j * kLd + i] : T(0);
  }
  __syncthreads();  // L is stored before the inverse overwrites it
  tile_trtri<T>(s, rdg, tmp, tid);
  for (int e = tid; e < kTileElems; e += 256) {
    const int i = e & 63, j = e >> 6;
    if ((i >> 4) < (j >> 4)) s[j * kLd + i] = T(0);  // blocks above the diagonal blocks
  }
  __syncthreads();
  for (int e = tid; e < kTileElems; e += 256) Li[e] = s[(e >> 6) * kLd + (e & 63)];
}

// Row i of a 64x64 column-major block M (element (i, k) at k*ld + i) times v, in the
// one summation order every forward-substitution path uses: four 16-term fma chains
// (k = 16q .. 16q+15), then ((p0 + p1) + p2) + p3.  The factorization's own sweep and
// the separate launch_forward sweep therefore give bitwise the same vectors.
template <typename T>
__device__ __forceinline__ double part_dot16(const T* M, int ld, const double* v, int i, int q) {
  double acc = 0.0;
#pragma unroll
  for (int k = 16 * q; k < 16 * q + 16; ++k) acc = fma((double)M[k * ld + i], v[k], acc);
  return acc;
}

// The same product by a whole work-group: wave q forms the chain of part q for all 64
// rows, wave 0 adds the parts (result valid in wave 0, lane = row).  Call uniformly.
template <typename T>
__device__ __forceinline__ double block_row_dot64(const T* M, int ld, const double* v,
                                                  double (*rpart)[64], int wave, int lane) {
  rpart[wave][lane] = part_dot16(M, ld, v, lane, wave);
  __syncthreads();
  return ((rpart[0][lane] + rpart[1][lane]) + rpart[2][lane]) + rpart[3][lane];
}

// MODE 0: C_IJ -= sum_K L_IK L_JK^T                                (Schur updates)
// MODE 1: same, then X = C L_JJ^-T with the stored inverse diagonal tile; with
//         flags & 1 the diagonal tile of row I also takes its share C_II -= X X^T
//         (column step, I > J; every WG of a launch owns a distinct row I, apart
//         from the lookahead items, flags & 2, which only update the next column's
//         tile (J + 2, J + 1) or the diagonal tile (J + 2, J + 2) from K < J; in the
//         FP64 schedule only MODE 3 items take a share).  Items with flags & 16 do the
//         MODE 3 work; the
//         schedule puts them at the front of the launch so they are dispatched first,
//         and their waves run at raised priority (they are on the critical path).
// MODE 2: same (empty K range when the diagonal was updated incrementally), then
//         Cholesky of the diagonal tile and its inverse
// MODE 3: MODE 1 for the item I = J + 1, whose share completes C_II (with flags & 64
//         first the K = I - 2 term from the stored tile L_{I,I-2}), then the Cholesky
//         and inverse of C_II for the next column step
// The inverse diagonal tile of column J of matrix m lives at inv + inv_base[m] + J*4096.
template <typename T, int MODE>
__device__ __forceinline__ void tile_update(const TileWork* __restrict__ work,
                                            const int64_t* __restrict__ rowbase,
                                            const int64_t* __restrict__ rowabs,
                                            const int64_t* __restrict__ inv_base,
                                            T* __restrict__ pool, T* __restrict__ inv,
                                            int* __restrict__ fail, RhsArgs rhs, SnapArgs snap) {
  using acc_t = typename Mfma<T>::acc_t;
  const TileWork w = work[blockIdx.x];
  // uniform per work-group: this item completes and factors the next diagonal tile
  const bool fac = MODE == 3 || (MODE == 1 && (w.flags & 16));
  if (MODE == 1 && fac) __builtin_amdgcn_s_setprio(3);
  // envelope-packed storage: tile (I, J) at ra[I] + J * 4096
  const int64_t* ra = rowabs + rowbase[w.mat];
  T* C = pool + ra[w.I] + (int64_t)w.J * kTileElems;
  // flags & 4: nothing has written tile C yet (no assembled entries): it reads as zero
  const bool fresh = w.flags & 4;  // uniform per work-group
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  // D = C^T: rows p of D are columns j of C, columns q of D are rows i of C.
  // wave w holds D rows (C columns) in the 16-blocks {0, 3} (w < 2) or {1, 2} (w >= 2)
  // and the 32 D columns from qb0: the two row-block sets cost the same in the
  // triangular TRSM product (tile_mma_lower_a)
  const int qb0 = (wave & 1) * 32;
  const int px[2] = {(wave >> 1) ? 16 : 0, (wave >> 1) ? 32 : 48};
  // FP64: the accumulator starts from C (loaded before the panel loads are issued, so
  // its latency overlaps them) and takes -sum_K L_IK L_JK^T.  FP32 sums the products
  // from zero and adds C once at the end (one rounding against C; accumulating into C
  // costs the FP32 factors accuracy they do not have to spare).
  constexpr bool from_c = sizeof(T) == 8;
  acc_t acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        acc[x][y][r] = (!from_c || fresh) ? T(0) : C[(px[x] + Mfma<T>::row(lk, r)) * 64 + qb0 + 16 * y + lr];
  // C + acc: the updated tile (element (i, j) at j * 64 + i)
  auto updated = [&](int x, int y, int r, int j, int i) {
    return (from_c || fresh) ? acc[x][y][r] : C[j * 64 + i] + acc[x][y][r];
  };
  // LDS: the panel staging buffers are dead after panel_mma; the same space then
  // holds the staged tile (and, in place, its factor and inverse).  aux holds the
  // partial dot products of the right-hand side or the inverse's staging blocks,
  // never both at once.
  constexpr int kStage = 2 * 64 * 16, kPanelLds = 2 * kStage > 64 * kLd ? 2 * kStage : 64 * kLd;
  constexpr int kTmp = 3 * 16 * 17;
  // aux: partial dot products [4][64] + y_J [64], or the inverse's staging blocks
  constexpr int kAux = (kTmp * (int)sizeof(T) + 7) / 8 > 320 ? (kTmp * (int)sizeof(T) + 7) / 8 : 320;
  __shared__ T lds_panel[kPanelLds];
  __shared__ double aux[kAux];
  // right-hand side operands of a column-step item, fetched ahead: y_J into aux[256..320)
  // (read after later barriers), b_I into a register of wave 0
  double bpre = 0.0;
  if constexpr (MODE == 1 || MODE == 3) {
    if (rhs.b && !(MODE == 1 && (w.flags & 2))) {  // uniform per work-group
      const int64_t o = rhs.off[w.mat];
      if (tid < 64) aux[256 + tid] = rhs.y[o + (int64_t)w.J * 64 + tid];
      if (wave == 0) bpre = rhs.b[o + (int64_t)w.I * 64 + lane];
    }
  }
  panel_mma<T, 2>(acc, pool + ra[w.J] + (int64_t)w.k0 * kTileElems,
               pool + ra[w.I] + (int64_t)w.k0 * kTileElems, (w.k1 - w.k0) * 64, lds_panel,
               lds_panel + kStage, tid, px[0], px[1], qb0, lr, lk);

  // MODE 0, and lookahead items (flags & 2) of a column-step launch: update only
  if (MODE == 0 || (MODE == 1 && (w.flags & 2))) {  // uniform per workgroup
    // MODE 0 with flags & 128 / 32: also the S_DD snapshot / the W store (see SnapArgs)
    const bool to_p = MODE == 0 && (w.flags & 128) && snap.P;  // uniform per work-group
    const bool to_w = MODE == 0 && (w.flags & 32) && snap.W;   // uniform per work-group
    int64_t pbase = 0, wbase = 0;
    bool phi = false;
    int64_t ldw = 64;
    if (to_p) {
      const DevSub S = snap.subs[w.mat];
      const int o = S.o_d >> 6, tI = w.I - o, tJ = w.J - o;
      pbase = S.p_off + ((int64_t)tI * (tI + 1) / 2 + tJ) * kTileElems;
    }
    if (to_w) {
      const DevSub S = snap.subs[w.mat];
      const int oa = S.o_a >> 6, kI = w.I - oa;
      phi = w.J < oa;  // a D'c tile: block (J - o_c/64, I - o_a/64) of Phi^T
      if (phi) {
        const int cJ = w.J - (S.o_c >> 6);
        wbase = S.w_off + sym_elems(S.ndp) + (int64_t)(cJ * 64) * S.ndp + kI * 64;
        ldw = S.ndp;
      } else {
        const int kJ = w.J - oa;
        wbase = S.w_off + ((int64_t)kI * (kI + 1) / 2 + kJ) * kTileElems;
      }
    }
    auto put = [&](int64_t k, T v) {
      if (snap.p32) static_cast<float*>(snap.P)[k] = (float)v;
      else static_cast<double*>(snap.P)[k] = (double)v;
    };
    T* Wd = static_cast<T*>(snap.W);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int j = px[x] + Mfma<T>::row(lk, r), i = qb0 + 16 * y + lr;
          const T v = updated(x, y, r, j, i);
          C[j * 64 + i] = v;
          if (to_p && (w.I != w.J || i >= j)) {  // element (i, j); a diagonal tile mirrors
            put(pbase + i * 64 + j, v);           // its lower half
            if (w.I == w.J && i != j) put(pbase + j * 64 + i, v);
          }
          if (to_w) {
            if (phi) {
              Wd[wbase + (int64_t)j * ldw + i] = -v;  // Phi^T row j of the block, column i
            } else if (w.I != w.J || i >= j) {
              Wd[wbase + i * 64 + j] = -v;
              if (w.I == w.J && i != j) Wd[wbase + j * 64 + i] = -v;
            }
          }
        }
    return;
  }
  if constexpr (MODE != 0) {
    T* s = lds_panel;
    __shared__ T rdg[MODE >= 1 ? 64 : 1];
    T* tmp = reinterpret_cast<T*>(aux);
    double(*rpart)[64] = reinterpret_cast<double(*)[64]>(aux);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int j = px[x] + Mfma<T>::row(lk, r), i = qb0 + 16 * y + lr;
          s[j * kLd + i] = updated(x, y, r, j, i);
        }
    __syncthreads();
    if constexpr (MODE == 1 || MODE == 3) {
      // the diagonal share's accumulator starts from C_II, loaded ahead of the TRSM
      const bool share = fac || (w.flags & 1);  // uniform per work-group
      const bool dfresh = w.flags & 8;  // flags & 8: the diagonal tile is still unwritten
      T* Cd = pool + ra[w.I] + (int64_t)w.I * kTileElems;
      acc_t a3[2][2];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            a3[x][y][r] = (!from_c || !share || dfresh)
                              ? T(0)
                              : Cd[(px[x] + Mfma<T>::row(lk, r)) * 64 + qb0 + 16 * y + lr];
      auto shared_d = [&](int x, int y, int r, int j, int i) {
        return (from_c || dfresh) ? a3[x][y][r] : Cd[j * 64 + i] + a3[x][y][r];
      };
      if (fac && (w.flags & 64)) {
        // the diagonal's K = I - 2 term, C_II -= L_{I,I-2} L_{I,I-2}^T (tile final since
        // the previous column launch; the earlier terms came from a diagonal lookahead)
        const T* Tk = pool + ra[w.I] + (int64_t)(w.I - 2) * kTileElems;
        tile_mma<T, 64, 64, true>(a3, Tk, Tk, px[0], px[1], qb0, lr, lk);
      }
      // X[i][j] = sum_k C'[i][k] Linv[j][k]:  D2[p=j][q=i], A op = Linv (p,k), B op = C' (q,k)
      const T* Li = inv + inv_base[w.mat] + (int64_t)w.J * kTileElems;
      acc_t a2[2][2];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) a2[x][y] = acc_t{0, 0, 0, 0};
      tile_mma_lower_a<T, 64, kLd>(a2, Li, s, px[0], qb0, lr, lk);
      __syncthreads();  // every wave is done reading C' from s
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int j = px[x] + Mfma<T>::row(lk, r), i = qb0 + 16 * y + lr;
            C[j * 64 + i] = a2[x][y][r];
            s[j * kLd + i] = a2[x][y][r];
          }
      __shared__ double bl[64];
      if (rhs.b) {  // uniform: b_I -= L_IJ y_J (fixed summation order, wave 0 writes)
        __syncthreads();
        const double dot = block_row_dot64(s, kLd, aux + 256, rpart, wave, lane);
        if (wave == 0) {
          const double nb = bpre - dot;
          rhs.b[rhs.off[w.mat] + (int64_t)w.I * 64 + lane] = nb;
          if (fac) bl[lane] = nb;
        }
      }
      if (!fac && !(w.flags & 1)) return;  // uniform per workgroup
      __syncthreads();
      // C_II -= X X^T
      tile_mma<T, kLd, kLd, true>(a3, s, s, px[0], px[1], qb0, lr, lk);
      if (!fac) {
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int j = px[x] + Mfma<T>::row(lk, r), i = qb0 + 16 * y + lr;
              Cd[j * 64 + i] = shared_d(x, y, r, j, i);
            }
      } else {
        // C_II is complete: stage it in s once every wave has finished reading X
        __syncthreads();
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int j = px[x] + Mfma<T>::row(lk, r), i = qb0 + 16 * y + lr;
              s[j * kLd + i] = shared_d(x, y, r, j, i);
            }
        __syncthreads();
        factor_diag_tile<T>(s, rdg, tmp, Cd, inv + inv_base[w.mat] + (int64_t)w.I * kTileElems,
                            fail + w.mat, tid);
        if (rhs.b) {  // y_I = L_II^-1 b_I (b_I complete in bl; s holds L_II^-1)
          const double yv = block_row_dot64(s, kLd, bl, rpart, wave, lane);
          if (wave == 0) rhs.y[rhs.off[w.mat] + (int64_t)w.I * 64 + lane] = yv;
        }
      }
    } else {
      factor_diag_tile<T>(s, rdg, tmp, C, inv + inv_base[w.mat] + (int64_t)w.J * kTileElems,
                          fail + w.mat, tid);
      if (rhs.b) {  // y_J = L_JJ^-1 b_J (b_J complete in memory)
        const int64_t o = rhs.off[w.mat] + (int64_t)w.J * 64;
        const double yv = block_row_dot64(s, kLd, rhs.b + o, rpart, wave, lane);
        if (wave == 0) rhs.y[o + lane] = yv;
      }
    }
  }
}

// Modes 0 and 1 are the throughput kernels: capped at 128 registers per lane so four
// waves share each SIMD.  Modes 2 and 3 hold a 64x64 tile in one wave's registers.
template <typename T, int MODE>
__global__ __launch_bounds__(256, 4) void k_tile_update(const TileWork* __restrict__ work,
                                                        const int64_t* __restrict__ rowbase,
                                                        const int64_t* __restrict__ rowabs,
                                                        const int64_t* __restrict__ inv_base,
                                                        T* __restrict__ pool, T* __restrict__ inv,
                                                        int* __restrict__ fail, RhsArgs rhs, SnapArgs snap) {
  tile_update<T, MODE>(work, rowbase, rowabs, inv_base, pool, inv, fail, rhs, snap);
}
// Same code with the compiler's own register allocation (no occupancy floor): the
// autotuner's alternative for the throughput modes.
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_tile_update_free(const TileWork* __restrict__ work,
                                                           const int64_t* __restrict__ rowbase,
                                                           const int64_t* __restrict__ rowabs,
                                                           const int64_t* __restrict__ inv_base,
                                                           T* __restrict__ pool, T* __restrict__ inv,
                                                           int* __restrict__ fail, RhsArgs rhs, SnapArgs snap) {
  tile_update<T, MODE>(work, rowbase, rowabs, inv_base, pool, inv, fail, rhs, snap);
}
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_tile_factor(const TileWork* __restrict__ work,
                                                     const int64_t* __restrict__ rowbase,
                                                     const int64_t* __restrict__ rowabs,
                                                     const int64_t* __restrict__ inv_base,
                                                     T* __restrict__ pool, T* __restrict__ inv,
                                                     int* __restrict__ fail, RhsArgs rhs, SnapArgs snap) {
  tile_update<T, MODE>(work, rowbase, rowabs, inv_base, pool, inv, fail, rhs, snap);
}

template <typename T>
void launch_tile_mode(int mode, const TileWork* w, int64_t n, const int64_t* rowbase,
                      const int64_t* rowabs, const int64_t* inv_base, T* pool, T* inv, int* fail,
                      hipStream_t st, int waves, RhsArgs rhs, int lds_pad, SnapArgs snap) {
  if (n <= 0) return;
  if (mode == 0) rhs = RhsArgs();  // Schur updates do not touch the right-hand side
  else snap = SnapArgs();          // only Schur updates take the snapshot
  if (mode == 0 && waves == 0)
    hipLaunchKernelGGL((k_tile_update_free<T, 0>), dim3(n), dim3(256), lds_pad, st, w, rowbase, rowabs, inv_base, pool, inv, fail, rhs, snap);
  else if (mode == 1 && waves == 0)
    hipLaunchKernelGGL((k_tile_update_free<T, 1>), dim3(n), dim3(256), 0, st, w, rowbase, rowabs, inv_base, pool, inv, fail, rhs, snap);
  else if (mode == 0)
    hipLaunchKernelGGL((k_tile_update<T, 0>), dim3(n), dim3(256), lds_pad, st, w, rowbase, rowabs, inv_base, pool, inv, fail, rhs, snap);
  else if (mode == 1)
    hipLaunchKernelGGL((k_tile_update<T, 1>), dim3(n), dim3(256), 0, st, w, rowbase, rowabs, inv_base, pool, inv, fail, rhs, snap);
  else if (mode == 2)
    hipLaunchKernelGGL((k_tile_factor<T, 2>), dim3(n), dim3(256), 0, st, w, rowbase, rowabs, inv_base, pool, inv, fail, rhs, snap);
  else
    hipLaunchKernelGGL((k_tile_factor<T, 3>), dim3(n), dim3(256), 0, st, w, rowbase, rowabs, inv_base, pool, inv, fail, rhs, snap);
}
template void launch_tile_mode<double>(int, const TileWork*, int64_t, const int64_t*, const int64_t*,
                                       const int64_t*, double*, double*, int*, hipStream_t, int, RhsArgs, int, SnapArgs);
template void launch_tile_mode<float>(int, const TileWork*, int64_t, const int64_t*, const int64_t*,
                                      const int64_t*, float*, float*, int*, hipStream_t, int, RhsArgs, int, SnapArgs);

// ---------------------------------------------------------------------------
// Operator extraction after the partial factorization
// ---------------------------------------------------------------------------
// Symmetric read of element (a, b) from an envelope-packed matrix (ra: tile-row
// offsets, ft: first stored tile per row); entries outside the envelope are 0.
template <typename T>
__device__ __forceinline__ T pool_sym(const T* P, const int64_t* ra, const int32_t* ft, int a,
                                      int b) {
  const int hi = a > b ? a : b, lo = a > b ? b : a;
  if ((lo >> 6) < ft[hi >> 6]) return T(0);
  return P[ra[hi >> 6] + (int64_t)(lo >> 6) * kTileElems + (lo & 63) * 64 + (hi & 63)];
}

// (row, column) of element e of a lower-tile symmetric store (tile (I, J <= I) at
// (I(I+1)/2 + J) * 4096, row-major in the tile)
__device__ __forceinline__ void sym_tile_rc(int64_t e, int& row, int& col) {
  const int64_t t = e >> 12;
  int I = int((sqrt(8.0 * double(t) + 1.0) - 1.0) * 0.5);
  while (int64_t(I) * (I + 1) / 2 > t) --I;
  while (int64_t(I + 1) * (I + 2) / 2 <= t) ++I;
  const int J = int(t - int64_t(I) * (I + 1) / 2);
  const int w = int(e & 4095);
  row = I * 64 + (w >> 6);
  col = J * 64 + (w & 63);
}

// Lower 64x64 tiles of a symmetric block of the factor pool (the D block for P, the D'
// block for Z), copied tile by tile into a lower-tile store (row-major inside a tile):
// one work-group per stored tile, coalesced reads of the column-major pool tile, the
// transpose through LDS; diagonal tiles take their lower half; tiles outside the
// envelope are zero.  out tile (I, J) = sign * pool(o + 64 I .., o + 64 J ..).
template <typename T, typename TP>
__global__ __launch_bounds__(256) void k_extract_sym(const SymTile* __restrict__ tiles,
                                                     const DevSub* __restrict__ subs,
                                                     const T* __restrict__ pool,
                                                     const int64_t* __restrict__ rowabs,
                                                     const int32_t* __restrict__ ftile,
                                                     const uint8_t* __restrict__ tflag, int block,
                                                     double sign, TP* __restrict__ out) {
  const SymTile tl = tiles[blockIdx.x];
  const DevSub S = subs[tl.s];
  const int o = (block == 0 ? S.o_d : S.o_a) >> 6;
  const int I = o + tl.I, J = o + tl.J;
  TP* dst = out + (block == 0 ? S.p_off : S.w_off) + ((int64_t)tl.I * (tl.I + 1) / 2 + tl.J) * kTileElems;
  const int tid = threadIdx.x;
  const int64_t soff = rowabs[S.ft_off + I] + (int64_t)J * kTileElems;
  if (J < ftile[S.ft_off + I] || (tflag && !tflag[soff / kTileElems])) {  // outside the envelope
    for (int e = tid; e < kTileElems; e += 256) dst[e] = TP(0);         // or never written
    return;
  }
  __shared__ T t[64 * 65];
  const T* src = pool + soff;
  for (int e = tid; e < kTileElems; e += 256) t[(e >> 6) * 65 + (e & 63)] = src[e];  // t[c][r] = (r, c)
  __syncthreads();
  const bool diag = I == J;
  for (int e = tid; e < kTileElems; e += 256) {
    const int r = e >> 6, c = e & 63;
    const T v = (!diag || r >= c) ? t[c * 65 + r] : t[r * 65 + c];
    dst[e] = (TP)(sign * (double)v);
  }
}

// Phi^T = -A(c, D') as rows of n_cp x n_Dp (row-major) after Z in W: one work-group per
// 64 x 64 block; the pool tile (D' tile row, c tile column) is read coalesced and written
// in the same order (Phi^T row = pool column).
template <typename T>
__global__ __launch_bounds__(256) void k_extract_phi(const DevSub* __restrict__ subs,
                                                     const T* __restrict__ pool,
                                                     const int64_t* __restrict__ rowabs,
                                                     const int32_t* __restrict__ ftile,
                                                     const uint8_t* __restrict__ tflag,
                                                     T* __restrict__ W) {
  const DevSub S = subs[blockIdx.y];
  const int tc = S.ncp >> 6, td = S.ndp >> 6;
  const int b = blockIdx.x;
  if (b >= tc * td) return;
  const int cI = b / td, kJ = b % td;
  const int I = (S.o_a >> 6) + kJ, J = (S.o_c >> 6) + cI;  // pool tile (D' row, c column)
  T* dst = W + S.w_off + sym_elems(S.ndp) + (int64_t)(cI * 64) * S.ndp + kJ * 64;
  const int tid = threadIdx.x;
  const int64_t soff = rowabs[S.ft_off + I] + (int64_t)J * kTileElems;
  const bool in = J >= ftile[S.ft_off + I] && (!tflag || tflag[soff / kTileElems]);
  const T* src = pool + soff;
  for (int e = tid; e < kTileElems; e += 256) {
    const int c = e >> 6, r = e & 63;  // pool element (r, c) of the tile at c*64 + r
    dst[(int64_t)c * S.ndp + r] = in ? -src[e] : T(0);
  }
}

template <typename T, typename TP>
void launch_extract_P(int64_t ntiles, const SymTile* tiles, const DevSub* subs, const T* pool,
                      const int64_t* rowabs, const int32_t* ftile, const uint8_t* tflag, TP* P,
                      hipStream_t st) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL((k_extract_sym<T, TP>), dim3(ntiles), dim3(256), 0, st, tiles, subs, pool, rowabs,
                     ftile, tflag, 0, 1.0, P);
}
template void launch_extract_P<double, float>(int64_t, const SymTile*, const DevSub*, const double*,
                                              const int64_t*, const int32_t*, const uint8_t*, float*, hipStream_t);
template void launch_extract_P<float, float>(int64_t, const SymTile*, const DevSub*, const float*,
                                             const int64_t*, const int32_t*, const uint8_t*, float*, hipStream_t);
template void launch_extract_P<double, double>(int64_t, const SymTile*, const DevSub*, const double*,
                                               const int64_t*, const int32_t*, const uint8_t*, double*, hipStream_t);
template void launch_extract_P<float, double>(int64_t, const SymTile*, const DevSub*, const float*,
                                              const int64_t*, const int32_t*, const uint8_t*, double*, hipStream_t);

template <typename T>
void launch_extract_W(int64_t ntiles, const SymTile* tiles, int ns, int max_ncp, int max_ndp,
                      const DevSub* subs, const T* pool, const int64_t* rowabs, const int32_t* ftile,
                      const uint8_t* tflag, T* W, hipStream_t st) {
  if (ntiles > 0)
    hipLaunchKernelGGL((k_extract_sym<T, T>), dim3(ntiles), dim3(256), 0, st, tiles, subs, pool, rowabs,
                       ftile, tflag, 1, -1.0, W);
  const int nb = (max_ncp >> 6) * (max_ndp >> 6);
  if (nb > 0 && ns > 0)
    hipLaunchKernelGGL(k_extract_phi<T>, dim3(nb, ns), dim3(256), 0, st, subs, pool, rowabs, ftile, tflag, W);
}
template void launch_extract_W<double>(int64_t, const SymTile*, int, int, int, const DevSub*,
                                       const double*, const int64_t*, const int32_t*, const uint8_t*,
                                       double*, hipStream_t);
template void launch_extract_W<float>(int64_t, const SymTile*, int, int, int, const DevSub*,
                                      const float*, const int64_t*, const int32_t*, const uint8_t*, float*, hipStream_t);

// S~^-1 x for the coarse problem, straight from the factored coarse pool: the trailing
// block of [[S~, I], [I, 0]] holds -S~^-1 as lower 64x64 tiles (column-major inside a
// tile: element (r, c) of tile (I, J) at c * 64 + r; diagonal tiles full).  One work-group
// per lower tile (I, J) reads it once and writes two 64-vectors of partial sums into
// slot (I, J) (T_IJ x_J) and, for I > J, slot (J, I) (T_IJ^T x_I) of a (Tn x Tn) array;
// k_coarse_reduce sums each row of slots (in a fixed order) and negates.  Lane = tile row and
// wave w holds columns 16w .. 16w+15, so every load is one coalesced 512-byte column; the
// column sums use the same halving butterfly as k_symv_tile, the row sums a 4-way LDS
// reduction.  Diagonal tiles are read through their lower triangle only (the upper entry
// (r, c) as (c, r)), like every other reader of the pool, so the operator is symmetric by
// construction whatever the factorization left above the diagonal.  Entries of x past nc
// read as zero (padding rows of the block are zero).
__global__ __launch_bounds__(256) void k_coarse_symv(int64_t nc, int tn, const int32_t* __restrict__ tiles,
                                                     const double* __restrict__ cpool,
                                                     const int64_t* __restrict__ cra,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ part) {
  __shared__ double xI[64], xJ[64];
  __shared__ double colp[4][64];
  const int I = tiles[2 * blockIdx.x], J = tiles[2 * blockIdx.x + 1];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const double* A = cpool + cra[tn + I] + (int64_t)(tn + J) * kTileElems;
  double at[16];
  const bool diag = I == J;  // uniform per work-group
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int col = 16 * wv + q;
    at[q] = (diag && col > lane) ? A[lane * 64 + col] : A[col * 64 + lane];
  }
  if (tid < 128) {
    const int64_t idx = int64_t((tid >> 6) ? I : J) * 64 + lane;
    const double xv = idx < nc ? x[idx] : 0.0;
    if (tid >> 6) xI[lane] = xv;
    else xJ[lane] = xv;
  }
  __syncthreads();
  const double xr = xI[lane];
  double rv[16];
  double rs_part = 0.0;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    rv[q] = at[q] * xr;
    rs_part += at[q] * xJ[16 * wv + q];
  }
  colp[wv][lane] = rs_part;
  if (I != J) {  // uniform per work-group: column sums T_IJ^T x_I
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const bool hi = lane & 32;
      const double keep = hi ? rv[i + 8] : rv[i], send = hi ? rv[i] : rv[i + 8];
      rv[i] = keep + __shfl_xor(send, 32);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool hi = lane & 16;
      const double keep = hi ? rv[i + 4] : rv[i], send = hi ? rv[i] : rv[i + 4];
      rv[i] = keep + __shfl_xor(send, 16);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const bool hi = lane & 8;
      const double keep = hi ? rv[i + 2] : rv[i], send = hi ? rv[i] : rv[i + 2];
      rv[i] = keep + __shfl_xor(send, 8);
    }
    {
      const bool hi = lane & 4;
      const double keep = hi ? rv[1] : rv[0], send = hi ? rv[0] : rv[1];
      rv[0] = keep + __shfl_xor(send, 4);
    }
    double cs = rv[0];
    cs += __shfl_xor(cs, 2);
    cs += __shfl_xor(cs, 1);
    if ((lane & 3) == 0) {
      const int col = 16 * wv + ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                      ((lane >> 2) & 1);
      part[((int64_t)J * tn + I) * 64 + col] = cs;
    }
  }
  __syncthreads();
  if (tid < 64)
    part[((int64_t)I * tn + J) * 64 + tid] = (colp[0][tid] + colp[1][tid]) + (colp[2][tid] + colp[3][tid]);
}

// y[X*64 + e] = -sum_{Y < tn} part[(X, Y)][e] for X*64 + e < nc.  Wave w sums the slots
// Y = w (mod 4) with two accumulators (even and odd steps), so each lane keeps
// ~tn/4 independent loads in flight; the eight partials are then added in a fixed order.
__global__ __launch_bounds__(256) void k_coarse_reduce(int64_t nc, int tn, const double* __restrict__ part,
                                                       double* __restrict__ y) {
  __shared__ double red[4][64];
  const int X = blockIdx.x, e = threadIdx.x & 63, w = threadIdx.x >> 6;
  const double* p = part + (int64_t)X * tn * 64 + e;
  double s0 = 0.0, s1 = 0.0;
  int Y = w;
  for (; Y + 4 < tn; Y += 8) {
    s0 += p[(int64_t)Y * 64];
    s1 += p[(int64_t)(Y + 4) * 64];
  }
  if (Y < tn) s0 += p[(int64_t)Y * 64];
  red[w][e] = s0 + s1;
  __syncthreads();
  const int64_t i = int64_t(X) * 64 + e;
  if (w == 0 && i < nc) y[i] = -((red[0][e] + red[1][e]) + (red[2][e] + red[3][e]));
}

void launch_coarse_symv(int64_t nc, int tn, int64_t ntiles, const int32_t* tiles, const double* cpool,
                        const int64_t* cra, const double* x, double* part, double* y, hipStream_t st) {
  if (nc <= 0 || ntiles <= 0) return;
  hipLaunchKernelGGL(k_coarse_symv, dim3(ntiles), dim3(256), 0, st, nc, tn, tiles, cpool, cra, x, part);
  hipLaunchKernelGGL(k_coarse_reduce, dim3(tn), dim3(256), 0, st, nc, tn, part, y);
}

// ---------------------------------------------------------------------------
// Grouped coarse solve: S~ = L L^T with L in the envelope-packed coarse pool (tiles
// column-major), groups of tile columns g = [a, a + w), X_g = (L_gg)^-1 dense
// column-major of order n = 64 w (element (r, c) at c * n + r).
// ---------------------------------------------------------------------------

// X_g tile (i, j), i >= j: X_ii = L_ii^-1, X_ij = -L_ii^-1 sum_{k = max(j, f_i)}^{i-1} L_ik X_kj.
// One work-group per tile; a launch holds the tiles of one distance d = i - j, so every
// X_kj it reads (k - j < d) is complete.
__global__ __launch_bounds__(256) void k_cg_inv_step(const int32_t* __restrict__ items,
                                                     const double* __restrict__ cpool,
                                                     const int64_t* __restrict__ cra,
                                                     const int32_t* __restrict__ cft,
                                                     const double* __restrict__ cinv,
                                                     const int32_t* __restrict__ cg_start,
                                                     const int32_t* __restrict__ cg_group_of,
                                                     const int64_t* __restrict__ xoff,
                                                     double* __restrict__ X) {
  // FP64 MFMA: wave w holds rows 16 w .. 16 w + 15 of the 64x64 result (four 16x16
  // blocks); each product stages its left factor as sA[row][inner] and its right factor
  // as sB[inner][column] (leading dimension kLd)
  using acc_t = Mfma<double>::acc_t;
  __shared__ double sA[64 * kLd], sB[64 * kLd];
  const int i = items[2 * blockIdx.x], j = items[2 * blockIdx.x + 1];
  const int g = cg_group_of[i], a = cg_start[g];
  const int64_t n = int64_t(cg_start[g + 1] - a) * 64;
  double* Xg = X + xoff[g];
  const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6, lr = l & 15, lk = l >> 4;
  const double* Li = cinv + (int64_t)i * kTileElems;
  double* Xij = Xg + (int64_t)(j - a) * 64 * n + (int64_t)(i - a) * 64;
  if (i == j) {
    for (int e = tid; e < kTileElems; e += 256) Xij[(int64_t)(e >> 6) * n + (e & 63)] = Li[e];
    return;
  }
  auto mma = [&](acc_t (&acc)[4]) {
#pragma unroll 4
    for (int kk = 0; kk < 64; kk += 4) {
      const double av = sA[(16 * wv + lr) * kLd + kk + lk];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[cb] = Mfma<double>::op(av, sB[(kk + lk) * kLd + 16 * cb + lr], acc[cb]);
    }
  };
  acc_t m[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) m[cb] = acc_t{0.0, 0.0, 0.0, 0.0};
  // M = sum_k L_ik X_kj over the envelope of row i; the two tiles of step k + 1 are
  // loaded into registers while step k's products run
  const int k0 = max(j, cft[i]);
  double ra[16], rb[16];
  auto fetch = [&](int k) {
    const double* Lk = cpool + cra[i] + (int64_t)k * kTileElems;       // column-major tile
    const double* Xk = Xg + (int64_t)(j - a) * 64 * n + (int64_t)(k - a) * 64;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int e = tid + 256 * u;
      ra[u] = Lk[e];
      rb[u] = Xk[(int64_t)(e >> 6) * n + (e & 63)];
    }
  };
  if (k0 < i) fetch(k0);
  for (int k = k0; k < i; ++k) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int e = tid + 256 * u, hi = e >> 6, lo = e & 63;
      sA[lo * kLd + hi] = ra[u];  // L_ik(row lo, inner hi)
      sB[lo * kLd + hi] = rb[u];  // X_kj(inner lo, column hi)
    }
    __syncthreads();
    if (k + 1 < i) fetch(k + 1);
    mma(m);
    __syncthreads();
  }
  // X_ij = -L_ii^-1 M
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r) sB[(16 * wv + Mfma<double>::row(lk, r)) * kLd + 16 * cb + lr] = m[cb][r];
  for (int e = tid; e < kTileElems; e += 256) sA[(e & 63) * kLd + (e >> 6)] = Li[e];
  __syncthreads();
  acc_t o[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) o[cb] = acc_t{0.0, 0.0, 0.0, 0.0};
  mma(o);
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      Xij[(int64_t)(16 * cb + lr) * n + 16 * wv + Mfma<double>::row(lk, r)] = -o[cb][r];
}

void launch_cg_invert(int max_group, const int64_t* step, const int32_t* items, const double* cpool,
                      const int64_t* cra, const int32_t* cft, const double* cinv,
                      const int32_t* cg_start, const int32_t* cg_group_of, const int64_t* xoff,
                      double* X, hipStream_t st) {
  for (int d = 0; d < max_group; ++d) {
    const int64_t n = step[d + 1] - step[d];
    if (n > 0)
      hipLaunchKernelGGL(k_cg_inv_step, dim3(n), dim3(256), 0, st, items + 2 * step[d], cpool, cra, cft,
                         cinv, cg_start, cg_group_of, xoff, X);
  }
}

// Grouped sweeps, two launches per group (the fixed-order sums are recomputed where they
// are consumed instead of being launches of their own; every work-group that forms one of
// them does the same additions in the same order, so all copies agree bitwise):
//  * step A of group g: the output blocks of group g-1 (sum of its X-product slots,
//    written to outv / out2) and the panel products of group g (one work-group per tile
//    (I, J) of L, partial 64-vector slot = its index: forward L_IJ v_J, backward
//    L_IJ^T v_I; a block v of group g-1 is summed from the X-product slots in place);
//  * step B of group g: one work-group per lower tile (i, j) of X_g, which first forms its
//    right-hand side block t = base - (sum of that block's panel slots) and then writes
//    the partial X_ij t (forward) or X_ij^T t (backward) to slot i(i+1)/2 + j.

// The right-hand side block J of a group step: base_J - the sum of its panel slots
// [ptr[J], ptr[J+1]) (base entries past nb read as zero).  Four threads per entry each
// sum every fourth slot; the four sums are added in a fixed order.  256 threads; the
// result is in tout[0..63] for every thread after the call.
__device__ __forceinline__ void cg_rhs_block(int J, const double* __restrict__ base, int64_t nb,
                                             const int32_t* __restrict__ ptr,
                                             const double* __restrict__ part, double (*red)[64],
                                             double* tout) {
  const int e = threadIdx.x & 63, q = threadIdx.x >> 6;
  double s0 = 0.0, s1 = 0.0;
  int k = ptr[J] + q;
  const int k1 = ptr[J + 1];
  for (; k + 4 < k1; k += 8) {
    s0 += part[(int64_t)k * 64 + e];
    s1 += part[(int64_t)(k + 4) * 64 + e];
  }
  if (k < k1) s0 += part[(int64_t)k * 64 + e];
  red[q][e] = s0 + s1;
  __syncthreads();
  if (q == 0) {
    const int64_t gi = int64_t(J) * 64 + e;
    const double b = gi < nb ? base[gi] : 0.0;
    tout[e] = b - ((red[0][e] + red[1][e]) + (red[2][e] + red[3][e]));
  }
  __syncthreads();
}

// Output block b of a group of width w from its X-product slots: forward y_{a+b} = sum over
// j <= b of slot b(b+1)/2 + j, backward x_{a+b} = sum over i >= b of slot i(i+1)/2 + b.
// Same thread layout and order as cg_rhs_block; result in vout[0..63] after the call.
template <bool TRANS>
__device__ __forceinline__ void cg_out_block(int b, int w, const double* __restrict__ xpart,
                                             double (*red)[64], double* vout) {
  const int e = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int k0 = TRANS ? b : 0, k1 = TRANS ? w : b + 1;
  double s0 = 0.0;
  for (int k = k0 + q; k < k1; k += 4) {
    const int T = TRANS ? k * (k + 1) / 2 + b : b * (b + 1) / 2 + k;
    s0 += xpart[(int64_t)T * 64 + e];
  }
  red[q][e] = s0;
  __syncthreads();
  if (q == 0) vout[e] = (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
  __syncthreads();
}

// 64x64 tile (lane = row, wave w holds columns 16w .. 16w+15 in at[]) times a 64-vector:
// non-transposed: out[r] = sum_c A[r][c] xs[c] (column sums through LDS); transposed:
// out[c] = sum_r A[r][c] xr (xr = the lane's entry; halving butterfly over the lanes).
template <bool TRANS>
__device__ __forceinline__ void cg_tile_product(const double (&at)[16], const double* xs, double xr,
                                                double (*colp)[64], double* __restrict__ out) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if constexpr (!TRANS) {
    double rs = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) rs = fma(at[q], xs[16 * wv + q], rs);
    colp[wv][lane] = rs;
    __syncthreads();
    if (tid < 64) out[tid] = (colp[0][tid] + colp[1][tid]) + (colp[2][tid] + colp[3][tid]);
  } else {
    double rv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) rv[q] = at[q] * xr;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const bool hi = lane & 32;
      const double keep = hi ? rv[u + 8] : rv[u], send = hi ? rv[u] : rv[u + 8];
      rv[u] = keep + __shfl_xor(send, 32);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool hi = lane & 16;
      const double keep = hi ? rv[u + 4] : rv[u], send = hi ? rv[u] : rv[u + 4];
      rv[u] = keep + __shfl_xor(send, 16);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const bool hi = lane & 8;
      const double keep = hi ? rv[u + 2] : rv[u], send = hi ? rv[u] : rv[u + 2];
      rv[u] = keep + __shfl_xor(send, 8);
    }
    {
      const bool hi = lane & 4;
      const double keep = hi ? rv[1] : rv[0], send = hi ? rv[0] : rv[1];
      rv[0] = keep + __shfl_xor(send, 4);
    }
    double cs = rv[0];
    cs += __shfl_xor(cs, 2);
    cs += __shfl_xor(cs, 1);
    if ((lane & 3) == 0) {
      const int col = 16 * wv + ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                      ((lane >> 2) & 1);
      out[col] = cs;
    }
  }
}

// Step A.  Work-groups [0, pw): output block b of the previous group [pa, pa + pw) into
// outv (and out2 where the entry is below nc).  Work-groups [pw, pw + npanel): panel tile
// tiles[k] = (I, J); its vector block (J forward, I backward) comes from the previous
// group's X-product slots when it lies in that group, else from vsrc.
template <bool TRANS>
__global__ __launch_bounds__(256) void k_cg_step_a(int pa, int pw, const double* __restrict__ xpart,
                                                   double* __restrict__ outv, double* __restrict__ out2,
                                                   int64_t nc, const int32_t* __restrict__ tiles,
                                                   const double* __restrict__ cpool,
                                                   const int64_t* __restrict__ cra,
                                                   const double* __restrict__ vsrc,
                                                   double* __restrict__ part) {
  __shared__ double red[4][64];
  __shared__ double xs[64];
  __shared__ double colp[4][64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if ((int)blockIdx.x < pw) {
    const int b = blockIdx.x;
    cg_out_block<TRANS>(b, pw, xpart, red, xs);
    if (tid < 64) {
      const int64_t gi = int64_t(pa + b) * 64 + tid;
      outv[gi] = xs[tid];
      if (out2 && gi < nc) out2[gi] = xs[tid];
    }
    return;
  }
  const int k = blockIdx.x - pw;
  const int I = tiles[2 * k], J = tiles[2 * k + 1];
  const double* A = cpool + cra[I] + (int64_t)J * kTileElems;
  double at[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) at[q] = A[(16 * wv + q) * 64 + lane];
  const int vb = TRANS ? I : J;  // the vector block this tile multiplies
  if (vb >= pa && vb < pa + pw) {
    cg_out_block<TRANS>(vb - pa, pw, xpart, red, xs);
  } else {
    if (tid < 64) xs[tid] = vsrc[(int64_t)vb * 64 + tid];
    __syncthreads();
  }
  cg_tile_product<TRANS>(at, xs, xs[lane], colp, part + (int64_t)k * 64);
}

// Step B of the group [a, a + w), one work-group per lower tile (i, j) of X_g (row-major
// enumeration T = i(i+1)/2 + j): t = base - panel slots of block a + j (forward) or
// a + i (backward), then slot T = X_ij t (forward) or X_ij^T t (backward).
template <bool TRANS>
__global__ __launch_bounds__(256) void k_cg_step_b(int a, int64_t n, const double* __restrict__ Xg,
                                                   const double* __restrict__ base, int64_t nb,
                                                   const int32_t* __restrict__ ptr,
                                                   const double* __restrict__ part,
                                                   double* __restrict__ xpart) {
  __shared__ double red[4][64];
  __shared__ double xs[64];
  __shared__ double colp[4][64];
  const int T = blockIdx.x;
  int i = int((sqrt(8.0 * double(T) + 1.0) - 1.0) * 0.5);
  while ((i + 1) * (i + 2) / 2 <= T) ++i;
  while (i * (i + 1) / 2 > T) --i;
  const int j = T - i * (i + 1) / 2;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double at[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) at[q] = Xg[(int64_t)(j * 64 + 16 * wv + q) * n + i * 64 + lane];
  cg_rhs_block(a + (TRANS ? i : j), base, nb, ptr, part, red, xs);
  cg_tile_product<TRANS>(at, xs, xs[lane], colp, xpart + (int64_t)T * 64);
}

void launch_cg_step_a(bool trans, int pa, int pw, const double* xpart, double* outv, double* out2,
                      int64_t nc, const int32_t* panel_tiles, int64_t npanel, const double* cpool,
                      const int64_t* cra, const double* vsrc, double* part, hipStream_t st) {
  const int64_t nb = int64_t(pw) + npanel;
  if (nb <= 0) return;
  if (trans)
    hipLaunchKernelGGL(k_cg_step_a<true>, dim3(nb), dim3(256), 0, st, pa, pw, xpart, outv, out2, nc,
                       panel_tiles, cpool, cra, vsrc, part);
  else
    hipLaunchKernelGGL(k_cg_step_a<false>, dim3(nb), dim3(256), 0, st, pa, pw, xpart, outv, out2, nc,
                       panel_tiles, cpool, cra, vsrc, part);
}

void launch_cg_step_b(bool trans, int a, int w, const double* Xg, const double* base, int64_t nb,
                      const int32_t* ptr, const double* part, double* xpart, hipStream_t st) {
  if (w <= 0) return;
  const int64_t n = int64_t(64) * w;
  if (trans)
    hipLaunchKernelGGL(k_cg_step_b<true>, dim3(w * (w + 1) / 2), dim3(256), 0, st, a, n, Xg, base, nb, ptr,
                       part, xpart);
  else
    hipLaunchKernelGGL(k_cg_step_b<false>, dim3(w * (w + 1) / 2), dim3(256), 0, st, a, n, Xg, base, nb, ptr,
                       part, xpart);
}

// ---------------------------------------------------------------------------
// Triangular solves with the stored partial factor: one launch per block column,
// one workgroup per (row tile, subdomain); diagonal blocks use the stored inverse.
// ---------------------------------------------------------------------------
// y = Linv * b for a 64-vector b (Linv column-major); result in ys (LDS), all threads
template <typename T>
__device__ __forceinline__ void tile_matvec(const T* L, const double* b, double* ys,
                                            double (*part)[64], int tid) {
  const int i = tid & 63, q = tid >> 6;
  part[q][i] = part_dot16(L, 64, b, i, q);
  __syncthreads();
  if (tid < 64) ys[tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
  __syncthreads();
}

// Column sums of tile L (64x64 col-major) against u (64): thread t handles the rows
// q + 16m (q = t & 15, m = 0..3) of the columns c = (t >> 4) + 16p, p = 0..3, so every load
// instruction of a wave reads four whole 128-byte column segments; the 16 partial sums
// of a column are combined with shuffles.  out[p] = sum_r L(r, (t >> 4) + 16p) u(r), in
// all 16 lanes of the group.
// colsum4 in two parts: the thread's 16 tile values (v[4p + m] = L(q + 16m, cg + 16p)),
// which can be loaded ahead, and the sums over them.
template <typename T>
__device__ __forceinline__ void colsum4_load(const T* L, int tid, T (&v)[16]) {
  const int q = tid & 15, cg = tid >> 4;
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int m = 0; m < 4; ++m) v[4 * p + m] = L[(cg + 16 * p) * 64 + q + 16 * m];
}
template <typename T>
__device__ __forceinline__ void colsum4_sum(const T (&v)[16], const double* u, int tid, double (&out)[4]) {
  const int q = tid & 15;
  double uq[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) uq[m] = u[q + 16 * m];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    double sum = 0.0;
#pragma unroll
    for (int m = 0; m < 4; ++m) sum += (double)v[4 * p + m] * uq[m];
    sum += __shfl_xor(sum, 1);
    sum += __shfl_xor(sum, 2);
    sum += __shfl_xor(sum, 4);
    sum += __shfl_xor(sum, 8);
    out[p] = sum;
  }
}
template <typename T>
__device__ __forceinline__ void colsum4(const T* L, const double* u, int tid, double (&out)[4]) {
  T v[16];
  colsum4_load(L, tid, v);
  colsum4_sum(v, u, tid, out);
}

// Forward step: column J (J counted from the subdomain's first dual column Ti when
// skip_int) gets y_J = L_JJ^-1 b_J (written to yv by the row-J workgroup); rows I > J:
// b_I -= L_IJ y_J.  D' rows are structurally zero while J < Ti.  skip_int: the interior
// right-hand side is zero (and yv's interior blocks already are), so the interior columns
// change nothing and only the dual ones are swept, every subdomain's from the first step.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_step(const DevSub* __restrict__ subs,
                                                  const T* __restrict__ pool,
                                                  const int64_t* __restrict__ rowabs,
                                                  const T* __restrict__ inv,
                                                  const int64_t* __restrict__ inv_base,
                                                  const int32_t* __restrict__ ftile, int Jstep,
                                                  int skip_int, double* __restrict__ vec,
                                                  double* __restrict__ yv) {
  const DevSub S = subs[blockIdx.y];
  const int J = skip_int ? S.Ti + Jstep : Jstep;
  if (J >= S.Tr) return;
  const int I = J + blockIdx.x;
  if (I >= S.T || ftile[S.ft_off + I] > J) return;  // tile (I, J) outside the envelope
  __shared__ double part[4][64];
  __shared__ double ys[64];
  const int tid = threadIdx.x;
  double* v = vec + S.vec_off;
  tile_matvec(inv + inv_base[blockIdx.y] + (int64_t)J * kTileElems, v + J * 64, ys, part, tid);
  if (I == J) {
    if (tid < 64) yv[S.vec_off + J * 64 + tid] = ys[tid];
    return;
  }
  const T* L = pool + rowabs[S.ft_off + I] + (int64_t)J * kTileElems;
  const int i = tid & 63, q = tid >> 6;
  part[q][i] = part_dot16(L, 64, ys, i, q);
  __syncthreads();
  if (tid < 64) v[I * 64 + tid] -= part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
}

template <typename T>
void launch_forward(int ns, int max_T, int max_Tr, const DevSub* subs, const T* pool,
                    const int64_t* rowabs, const T* inv, const int64_t* inv_base,
                    const int32_t* ftile, double* vec, double* yv, hipStream_t st, int skip_int,
                    int max_dual, int max_after) {
  if (ns <= 0) return;
  const int steps = skip_int ? max_dual : max_Tr, rows = skip_int ? max_after : max_T;
  for (int J = 0; J < steps; ++J)
    hipLaunchKernelGGL(k_fwd_step<T>, dim3(rows - J, ns), dim3(256), 0, st, subs, pool, rowabs, inv,
                       inv_base, ftile, J, skip_int, vec, yv);
}
template void launch_forward<double>(int, int, int, const DevSub*, const double*, const int64_t*,
                                     const double*, const int64_t*, const int32_t*, double*,
                                     double*, hipStream_t, int, int, int);
template void launch_forward<float>(int, int, int, const DevSub*, const float*, const int64_t*,
                                    const float*, const int64_t*, const int32_t*, double*, double*,
                                    hipStream_t, int, int, int);

// Zero the interior blocks (tile rows < Ti) of a local vector, one work-group per
// (interior tile, subdomain).
__global__ __launch_bounds__(64) void k_zero_interior(const DevSub* __restrict__ subs,
                                                      double* __restrict__ v) {
  const DevSub S = subs[blockIdx.y];
  if ((int)blockIdx.x >= S.Ti) return;
  v[S.vec_off + (int64_t)blockIdx.x * 64 + threadIdx.x] = 0.0;
}
void launch_zero_interior(int ns, int max_Ti, const DevSub* subs, double* v, hipStream_t st) {
  if (ns <= 0 || max_Ti <= 0) return;
  hipLaunchKernelGGL(k_zero_interior, dim3(max_Ti, ns), dim3(64), 0, st, subs, v);
}

// Block partial sums of x_d^2 over the free dofs of one kind (0 interior, 1 dual, 2 primal),
// kDotBlocks partials (summed on the host in a fixed order).
__global__ __launch_bounds__(256) void k_kind_sq(int64_t n, const int8_t* __restrict__ kind, int want,
                                                 const double* __restrict__ x, double* __restrict__ partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    if (kind[i] == want) s += x[i] * x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
void launch_kind_sq(int64_t n, const int8_t* kind, int want, const double* x, double* partial,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_kind_sq, dim3(kDotBlocks), dim3(256), 0, st, n, kind, want, x, partial);
}

// Interface part of the interior solve, for every interior tile column K:
// t_K = y_K - sum_{I in [Ti, Tc)} L_IK^T U_I.  1024 threads per (K, subdomain): quarter q
// sums the rows I = Ti+q, Ti+q+4, ..., quarter 0 adds the partials in order.  The
// work-group of the last interior column also forms x_K = L_KK^-T t_K, which starts
// the backward steps.
template <typename T>
__global__ __launch_bounds__(1024) void k_int_prep(const DevSub* __restrict__ subs,
                                                   const T* __restrict__ pool,
                                                   const int64_t* __restrict__ rowabs,
                                                   const T* __restrict__ inv,
                                                   const int64_t* __restrict__ inv_base,
                                                   const int32_t* __restrict__ ftile,
                                                   const double* __restrict__ yv,
                                                   const double* __restrict__ U,
                                                   double* __restrict__ tbuf,
                                                   double* __restrict__ xout) {
  const DevSub S = subs[blockIdx.y];
  const int K = blockIdx.x;
  if (K >= S.Ti) return;  // uniform per work-group
  const int q = threadIdx.x >> 8, tid = threadIdx.x & 255;
  const double* u = U + S.vec_off;
  double sum[4] = {0.0, 0.0, 0.0, 0.0}, cs[4];
  for (int I = S.Ti + q; I < S.Tc; I += 4)
    if (ftile[S.ft_off + I] <= K) {
      colsum4(pool + rowabs[S.ft_off + I] + (int64_t)K * kTileElems, u + (int64_t)I * 64, tid, cs);
#pragma unroll
      for (int p = 0; p < 4; ++p) sum[p] += cs[p];
    }
  __shared__ double part[4][64];
  __shared__ double ts[64];
  if ((tid & 15) == 0)
#pragma unroll
    for (int p = 0; p < 4; ++p) part[q][(tid >> 4) + 16 * p] = sum[p];
  __syncthreads();
  if (q == 0 && tid < 64) {
    const int64_t o = S.vec_off + (int64_t)K * 64 + tid;
    const double t = yv[o] - (((part[0][tid] + part[1][tid]) + part[2][tid]) + part[3][tid]);
    tbuf[o] = t;
    ts[tid] = t;
  }
  if (K != S.Ti - 1) return;  // uniform per work-group
  __syncthreads();
  if (q == 0) {
    double x[4];
    colsum4(inv + inv_base[blockIdx.y] + (int64_t)K * kTileElems, ts, tid, x);
    if ((tid & 15) == 0)
#pragma unroll
      for (int p = 0; p < 4; ++p) xout[S.vec_off + (int64_t)K * 64 + (tid >> 4) + 16 * p] = x[p];
  }
}

// Backward step J >= 1 (x_J known): rows I < J of the envelope take t_I -= L_JI^T x_J
// (work-group x reads tile (J, I = J - 1 - x) of the contiguous tile row J); the
// work-group of row J - 1, whose t is then complete, also forms x_{J-1} =
// L_{J-1,J-1}^-T t_{J-1}.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_step(const DevSub* __restrict__ subs,
                                                  const T* __restrict__ pool,
                                                  const int64_t* __restrict__ rowabs,
                                                  const T* __restrict__ inv,
                                                  const int64_t* __restrict__ inv_base,
                                                  const int32_t* __restrict__ ftile, int J,
                                                  double* __restrict__ tbuf,
                                                  double* __restrict__ xout) {
  const DevSub S = subs[blockIdx.y];
  if (J >= S.Ti) return;
  const int I = J - 1 - int(blockIdx.x);
  if (I < 0) return;
  const bool in_env = I >= ftile[S.ft_off + J];  // tile (J, I) is stored
  if (I >= J || (!in_env && I != J - 1)) return;
  const int tid = threadIdx.x;
  const bool last = I == J - 1;  // uniform per work-group: also forms x_{J-1}
  // the inverse diagonal tile does not depend on this step: its loads go out first
  T vinv[16];
  if (last) colsum4_load(inv + inv_base[blockIdx.y] + (int64_t)I * kTileElems, tid, vinv);
  double* t = tbuf + S.vec_off;
  double cs[4] = {0.0, 0.0, 0.0, 0.0};
  if (in_env)
    colsum4(pool + rowabs[S.ft_off + J] + (int64_t)I * kTileElems, xout + S.vec_off + (int64_t)J * 64, tid, cs);
  __shared__ double ts[64];
  if ((tid & 15) == 0)
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int c = (tid >> 4) + 16 * p;
      const double nt = t[I * 64 + c] - cs[p];
      t[I * 64 + c] = nt;
      ts[c] = nt;
    }
  if (!last) return;
  __syncthreads();
  double x[4];
  colsum4_sum(vinv, ts, tid, x);
  if ((tid & 15) == 0)
#pragma unroll
    for (int p = 0; p < 4; ++p) xout[S.vec_off + (int64_t)I * 64 + (tid >> 4) + 16 * p] = x[p];
}

template <typename T>
void launch_interior(int ns, int max_Ti, const int32_t* width, const DevSub* subs, const T* pool,
                     const int64_t* rowabs, const T* inv, const int64_t* inv_base,
                     const int32_t* ftile, const double* yv, const double* U, double* tbuf,
                     double* xout, hipStream_t st) {
  if (ns <= 0 || max_Ti <= 0) return;
  hipLaunchKernelGGL(k_int_prep<T>, dim3(max_Ti, ns), dim3(1024), 0, st, subs, pool, rowabs, inv,
                     inv_base, ftile, yv, U, tbuf, xout);
  for (int J = max_Ti - 1; J >= 1; --J)
    hipLaunchKernelGGL(k_bwd_step<T>, dim3(std::max(1, std::min(J, width[J])), ns), dim3(256), 0, st,
                       subs, pool, rowabs, inv, inv_base, ftile, J, tbuf, xout);
}
template void launch_interior<double>(int, int, const int32_t*, const DevSub*, const double*,
                                      const int64_t*, const double*, const int64_t*, const int32_t*,
                                      const double*, const double*, double*, double*, hipStream_t);
template void launch_interior<float>(int, int, const int32_t*, const DevSub*, const float*,
                                     const int64_t*, const float*, const int64_t*, const int32_t*,
                                     const double*, const double*, double*, double*, hipStream_t);

__global__ __launch_bounds__(256) void k_rho(int64_t n, const int64_t* __restrict__ pos,
                                             const void* __restrict__ pool, int is_f32,
                                             double* __restrict__ rho) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t p = pos[i];
  double v = 0.0;
  if (p >= 0) v = is_f32 ? (double)((const float*)pool)[p] : ((const double*)pool)[p];
  rho[i] = v;
}

// Stiffness scaling: the D slot of subdomain s for a multiplier shared with t is
// weighted rho_t / (rho_s + rho_t); sgnD = sign * weight, half = 0.5 on real slots.
__global__ __launch_bounds__(256) void k_scaling(int64_t nl, const int64_t* __restrict__ lam_pos,
                                                 const double* __restrict__ rho,
                                                 const double* __restrict__ sgn,
                                                 double* __restrict__ sgnD, double* __restrict__ half) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nl) return;
  const int64_t p0 = lam_pos[2 * l], p1 = lam_pos[2 * l + 1];
  const double r0 = rho[p0], r1 = rho[p1];
  sgnD[p0] = sgn[p0] * (r1 / (r0 + r1));
  sgnD[p1] = sgn[p1] * (r0 / (r0 + r1));
  half[p0] = 0.5;
  half[p1] = 0.5;
}

void launch_scaling(int64_t ndv, int64_t nl, const int64_t* rho_pos, const void* pool, int is_f32,
                    double* rho, const int64_t* lam_pos, const double* sgn, double* sgnD,
                    double* half, hipStream_t st) {
  if (ndv > 0)
    hipLaunchKernelGGL(k_rho, dim3((ndv + 255) / 256), dim3(256), 0, st, ndv, rho_pos, pool, is_f32, rho);
  if (nl > 0)
    hipLaunchKernelGGL(k_scaling, dim3((nl + 255) / 256), dim3(256), 0, st, nl, lam_pos, rho, sgn,
                       sgnD, half);
}

// ---------------------------------------------------------------------------
// PCPG operators
// ---------------------------------------------------------------------------
constexpr int kGemvRows = 32;

// 16-byte row segment: 2 doubles or 4 floats per lane and load.
template <typename T> struct Seg;
template <> struct Seg<double> {
  static constexpr int V = 2;
  __device__ static double dot(const double* a, const double* x) {
    const double2 u = *reinterpret_cast<const double2*>(a);
    return u.x * x[0] + u.y * x[1];
  }
};
template <> struct Seg<float> {
  static constexpr int V = 4;
  __device__ static double dot(const float* a, const double* x) {
    const float4 u = *reinterpret_cast<const float4*>(a);
    return (double)u.x * x[0] + (double)u.y * x[1] + (double)u.z * x[2] + (double)u.w * x[3];
  }
};

// Phi^T rows of W (ncp x ndp, row-major after the Z tiles) times the gathered
// x = d_scale * v[d_lam] of each subdomain, out at y_off + ndp.  Rows are
// ndp-strided with ndp a multiple of 64 and row bases 256-byte aligned, so every
// lane reads whole 16-byte segments; columns in [nd, ndp) meet zeros in x.  Each
// wave walks two rows at once so four segment loads are in flight per lane.
template <typename T>
__device__ __forceinline__ void gemv_phi_block(int bx, int by, const DevSub* __restrict__ subs,
                                               const T* __restrict__ W,
                                               const int32_t* __restrict__ d_lam,
                                               const double* __restrict__ d_scale,
                                               const double* __restrict__ v, double* __restrict__ y,
                                               double* xs) {
  constexpr int V = Seg<T>::V;
  const DevSub S = subs[by];
  const int rows = S.ncp;
  const int r0 = bx * kGemvRows;
  if (r0 >= rows) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int k = tid; k < S.ndp; k += 256) {
    const int64_t slot = S.dvec_off + k;
    const int32_t l = d_lam[slot];
    xs[k] = l >= 0 ? d_scale[slot] * v[l] : 0.0;
  }
  __syncthreads();
  const T* base = W + S.w_off + sym_elems(S.ndp);
  const int kend = (S.nd + V - 1) / V * V;  // <= ndp
  double* out = y + S.y_off + S.ndp;
  for (int rr = 2 * wv; rr < kGemvRows; rr += 8) {
    const int row0 = r0 + rr;
    if (row0 >= rows) break;
    const bool two = row0 + 1 < rows;
    const T* a0 = base + (int64_t)row0 * S.ndp;
    const T* a1 = two ? a0 + S.ndp : a0;
    double s0 = 0.0, s1 = 0.0, t0 = 0.0, t1 = 0.0;
    int k = lane * V;
    for (; k + 64 * V < kend; k += 128 * V) {
      s0 += Seg<T>::dot(a0 + k, xs + k);
      s1 += Seg<T>::dot(a1 + k, xs + k);
      t0 += Seg<T>::dot(a0 + k + 64 * V, xs + k + 64 * V);
      t1 += Seg<T>::dot(a1 + k + 64 * V, xs + k + 64 * V);
    }
    if (k < kend) {
      s0 += Seg<T>::dot(a0 + k, xs + k);
      s1 += Seg<T>::dot(a1 + k, xs + k);
    }
    s0 += t0;
    s1 += t1;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      s0 += __shfl_xor(s0, o);
      s1 += __shfl_xor(s1, o);
    }
    if (lane == 0) {
      out[row0] = s0;
      if (two) out[row0 + 1] = s1;
    }
  }
}
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_phi(const DevSub* __restrict__ subs,
                                                  const T* __restrict__ W,
                                                  const int32_t* __restrict__ d_lam,
                                                  const double* __restrict__ d_scale,
                                                  const double* __restrict__ v, double* __restrict__ y) {
  extern __shared__ double xs[];
  gemv_phi_block<T>(blockIdx.x, blockIdx.y, subs, W, d_lam, d_scale, v, y, xs);
}

template <typename T>
void launch_gemv_phi(int ns, int max_ncp, const DevSub* subs, const T* W, const int32_t* d_lam,
                     const double* d_scale, const double* v, double* y, hipStream_t st, int max_ndp) {
  if (ns <= 0 || max_ncp <= 0) return;
  hipLaunchKernelGGL(k_gemv_phi<T>, dim3((max_ncp + kGemvRows - 1) / kGemvRows, ns), dim3(256),
                     max_ndp * sizeof(double), st, subs, W, d_lam, d_scale, v, y);
}
template void launch_gemv_phi<double>(int, int, const DevSub*, const double*, const int32_t*,
                                      const double*, const double*, double*, hipStream_t, int);
template void launch_gemv_phi<float>(int, int, const DevSub*, const float*, const int32_t*,
                                     const double*, const double*, double*, hipStream_t, int);

// Symmetric dual operator (Z of W, or P) times the gathered x, from its lower
// tiles: one work-group per tile (I, J).  It reads the tile once and writes two
// 64-vectors of partial sums, M_IJ x_J at slot (I, J) and (for I > J) M_IJ^T x_I
// at slot (J, I) of the subdomain's (ndp/64)^2 partial array; k_symv_reduce then
// sums slots (I, 0 .. Td-1) in order.  Wave w holds tile rows 16w .. 16w+15, lane
// = column, so each tile row is one coalesced load (512 B in FP64).  The 16 row sums
// of a wave are formed by a halving butterfly over the lanes (8 + 4 + 2 + 1 + 2
// shuffles, no LDS), the column sums by one 4-way LDS reduction across the waves.
template <typename T>
__device__ __forceinline__ void symv_tile_block(int b, const SymTile* __restrict__ tiles,
                                                const DevSub* __restrict__ subs,
                                                const T* __restrict__ M, int which,
                                                const int32_t* __restrict__ d_lam,
                                                const double* __restrict__ d_scale,
                                                const double* __restrict__ v,
                                                double* __restrict__ part) {
  __shared__ double xJ[64], xI[64];
  __shared__ double colp[4][64];
  const SymTile tl = tiles[b];
  const DevSub S = subs[tl.s];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // the tile's loads go out first; the indirect gather of x (d_lam, then v) overlaps them
  const T* A = M + (which == 0 ? S.w_off : S.p_off) + ((int64_t)tl.I * (tl.I + 1) / 2 + tl.J) * kTileElems;
  T at[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) at[q] = A[(16 * wv + q) * 64 + lane];
  if (tid < 128) {
    const int blk = (tid >> 6) ? tl.I : tl.J;
    const int64_t slot = S.dvec_off + blk * 64 + lane;
    const int32_t l = d_lam[slot];
    const double xv = l >= 0 ? (d_scale ? d_scale[slot] : 1.0) * v[l] : 0.0;
    if (tid >> 6) xI[lane] = xv;
    else xJ[lane] = xv;
  }
  __syncthreads();
  const double xc = xJ[lane];
  double rv[16];
  double cs = 0.0;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const double a = (double)at[q];
    rv[q] = a * xc;
    cs += a * xI[16 * wv + q];
  }
  // halving butterfly: after the steps over lane bits 5, 4, 3, 2 lane l holds the partial
  // of row 8 b5 + 4 b4 + 2 b3 + b2 (b = bits of l) over the lanes that differ in bits 0-1
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const bool hi = lane & 32;
    const double keep = hi ? rv[i + 8] : rv[i], send = hi ? rv[i] : rv[i + 8];
    rv[i] = keep + __shfl_xor(send, 32);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const bool hi = lane & 16;
    const double keep = hi ? rv[i + 4] : rv[i], send = hi ? rv[i] : rv[i + 4];
    rv[i] = keep + __shfl_xor(send, 16);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const bool hi = lane & 8;
    const double keep = hi ? rv[i + 2] : rv[i], send = hi ? rv[i] : rv[i + 2];
    rv[i] = keep + __shfl_xor(send, 8);
  }
  {
    const bool hi = lane & 4;
    const double keep = hi ? rv[1] : rv[0], send = hi ? rv[0] : rv[1];
    rv[0] = keep + __shfl_xor(send, 4);
  }
  double rs = rv[0];
  rs += __shfl_xor(rs, 2);
  rs += __shfl_xor(rs, 1);
  const int Td = S.ndp >> 6;
  if ((lane & 3) == 0) {
    const int row = 16 * wv + ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                    ((lane >> 2) & 1);
    part[S.sp_off + ((int64_t)tl.I * Td + tl.J) * 64 + row] = rs;
  }
  if (tl.I != tl.J) {  // uniform per work-group
    colp[wv][lane] = cs;
    __syncthreads();
    if (tid < 64)
      part[S.sp_off + ((int64_t)tl.J * Td + tl.I) * 64 + tid] =
          (colp[0][tid] + colp[1][tid]) + (colp[2][tid] + colp[3][tid]);
  }
}
template <typename T>
__global__ __launch_bounds__(256) void k_symv_tile(const SymTile* __restrict__ tiles,
                                                   const DevSub* __restrict__ subs,
                                                   const T* __restrict__ M, int which,
                                                   const int32_t* __restrict__ d_lam,
                                                   const double* __restrict__ d_scale,
                                                   const double* __restrict__ v,
                                                   double* __restrict__ part) {
  symv_tile_block<T>(blockIdx.x, tiles, subs, M, which, d_lam, d_scale, v, part);
}
// W x in one launch: the Phi^T row blocks of every subdomain first (nphi per subdomain,
// the longer work-groups), then the Z tiles; Z x is left as partials.
template <typename T>
__global__ __launch_bounds__(256) void k_symv_phi(int nphi, int ns, const SymTile* __restrict__ tiles,
                                                  const DevSub* __restrict__ subs,
                                                  const T* __restrict__ W,
                                                  const int32_t* __restrict__ d_lam,
                                                  const double* __restrict__ d_scale,
                                                  const double* __restrict__ v,
                                                  double* __restrict__ part,
                                                  double* __restrict__ y) {
  extern __shared__ double xs[];
  const int b = blockIdx.x, nb = nphi * ns;
  if (b < nb) gemv_phi_block<T>(b % nphi, b / nphi, subs, W, d_lam, d_scale, v, y, xs);
  else symv_tile_block<T>(b - nb, tiles, subs, W, 0, d_lam, d_scale, v, part);
}
template <typename T>
void launch_apply_W(int64_t ntiles, const SymTile* tiles, int ns, int max_ncp, int max_ndp,
                    const DevSub* subs, const T* W, const int32_t* d_lam, const double* d_scale,
                    const double* v, double* part, double* y, hipStream_t st) {
  if (ns <= 0) return;
  const int nphi = (max_ncp + kGemvRows - 1) / kGemvRows;
  const int64_t nb = int64_t(nphi) * ns + ntiles;
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_symv_phi<T>, dim3(nb), dim3(256), size_t(max_ndp) * sizeof(double), st, nphi,
                     ns, tiles, subs, W, d_lam, d_scale, v, part, y);
}
template void launch_apply_W<double>(int64_t, const SymTile*, int, int, int, const DevSub*,
                                     const double*, const int32_t*, const double*, const double*,
                                     double*, double*, hipStream_t);
template void launch_apply_W<float>(int64_t, const SymTile*, int, int, int, const DevSub*,
                                    const float*, const int32_t*, const double*, const double*,
                                    double*, double*, hipStream_t);

// out[I*64 + r] = sum_{J < Td} part[(I, J)][r], at y_off (which 0) or dvec_off (which 1)
__global__ __launch_bounds__(64) void k_symv_reduce(const DevSub* __restrict__ subs,
                                                    const double* __restrict__ part, int which,
                                                    double* __restrict__ out) {
  const DevSub S = subs[blockIdx.y];
  const int Td = S.ndp >> 6, I = blockIdx.x;
  if (I >= Td) return;
  const int r = threadIdx.x;
  const double* p = part + S.sp_off + (int64_t)I * Td * 64 + r;
  double sum = 0.0;
#pragma unroll 8
  for (int J = 0; J < Td; ++J) sum += p[(int64_t)J * 64];  // unrolled: loads ahead of the adds
  out[(which == 0 ? S.y_off : S.dvec_off) + I * 64 + r] = sum;
}

template <typename T>
void launch_symv(int64_t ntiles, const SymTile* tiles, int ns, int max_Td, const DevSub* subs,
                 const T* M, int which, const int32_t* d_lam, const double* d_scale,
                 const double* v, double* part, double* out, hipStream_t st) {
  if (ntiles <= 0 || ns <= 0) return;
  hipLaunchKernelGGL(k_symv_tile<T>, dim3(ntiles), dim3(256), 0, st, tiles, subs, M, which, d_lam,
                     d_scale, v, part);
  if (out)  // else the consumer sums the partials itself
    hipLaunchKernelGGL(k_symv_reduce, dim3(max_Td, ns), dim3(64), 0, st, subs, part, which, out);
}
template void launch_symv<double>(int64_t, const SymTile*, int, int, const DevSub*, const double*,
                                  int, const int32_t*, const double*, const double*, double*,
                                  double*, hipStream_t);
template void launch_symv<float>(int64_t, const SymTile*, int, int, const DevSub*, const float*, int,
                                 const int32_t*, const double*, const double*, double*, double*,
                                 hipStream_t);

// out_D = alpha y_D + beta Phi u_c(s) + gamma vec_D' per subdomain.  One work-group per
// 64 D slots: wave q sums the primal rows c = q, q+4, ... (Phi^T rows, coalesced over the
// slots), then the four partials are added in fixed order.
template <typename T>
__global__ __launch_bounds__(256) void k_dual_correct(const DevSub* __restrict__ subs,
                                                      const T* __restrict__ W,
                                                      const double* __restrict__ zpart,
                                                      const int32_t* __restrict__ c_loc,
                                                      const double* __restrict__ uc,
                                                      const double* __restrict__ vec, double alpha,
                                                      double beta, double gamma,
                                                      double* __restrict__ out) {
  const DevSub S = subs[blockIdx.y];
  const int k0 = blockIdx.x * 64;
  if (k0 >= S.ndp) return;
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int k = k0 + lane;
  __shared__ double part[4][64];
  extern __shared__ double ul[];  // the subdomain's primal values u_c(s), staged once
  for (int c = threadIdx.x; c < S.nc; c += 256) ul[c] = uc[c_loc[S.cvec_off + c]];
  // (Z x)_k from the symmetric-product partials, in k_symv_reduce's order (wave 0; its
  // loads overlap the staging)
  double zx = 0.0;
  if (q == 0 && alpha != 0.0) {
    const int Td = S.ndp >> 6;
    const double* zp = zpart + S.sp_off + (int64_t)blockIdx.x * Td * 64 + lane;
#pragma unroll 4
    for (int J = 0; J < Td; ++J) zx += zp[(int64_t)J * 64];
  }
  __syncthreads();
  const T* ph = W + S.w_off + sym_elems(S.ndp);  // Phi^T rows
  double s0 = 0.0, s1 = 0.0;
  int c = q;
#pragma unroll 4
  for (; c + 4 < S.nc; c += 8) {  // unrolled: eight row loads in flight per lane
    s0 = fma((double)ph[(int64_t)c * S.ndp + k], ul[c], s0);
    s1 = fma((double)ph[(int64_t)(c + 4) * S.ndp + k], ul[c + 4], s1);
  }
  if (c < S.nc) s0 = fma((double)ph[(int64_t)c * S.ndp + k], ul[c], s0);
  part[q][lane] = s0 + s1;
  __syncthreads();
  if (q == 0) {
    double r = 0.0;
    if (k < S.nd) {
      r = beta * ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]));
      if (alpha != 0.0) r += alpha * zx;
      if (gamma != 0.0) r += gamma * vec[S.vec_off + S.o_a + k];
    }
    out[S.dvec_off + k] = r;
  }
}

template <typename T>
void launch_dual_correct(int ns, int max_ndp, int max_ncp, const DevSub* subs, const T* W, const double* zpart,
                         const int32_t* c_loc, const double* uc, const double* vec, double alpha,
                         double beta, double gamma, double* out, hipStream_t st) {
  if (ns <= 0 || max_ndp <= 0) return;
  hipLaunchKernelGGL(k_dual_correct<T>, dim3((max_ndp + 63) / 64, ns), dim3(256),
                     size_t(std::max(max_ncp, 1)) * sizeof(double), st, subs, W, zpart, c_loc, uc, vec,
                     alpha, beta, gamma, out);
}
template void launch_dual_correct<double>(int, int, int, const DevSub*, const double*, const double*,
                                          const int32_t*, const double*, const double*, double,
                                          double, double, double*, hipStream_t);
template void launch_dual_correct<float>(int, int, int, const DevSub*, const float*, const double*,
                                         const int32_t*, const double*, const double*, double, double,
                                         double, double*, hipStream_t);

__global__ __launch_bounds__(256) void k_lambda_scatter(int64_t nl, const int64_t* __restrict__ pos,
                                                        const double* __restrict__ scale,
                                                        const double* __restrict__ w,
                                                        double* __restrict__ out) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nl) return;
  const int64_t p0 = pos[2 * l], p1 = pos[2 * l + 1];
  out[l] = scale[p0] * w[p0] + scale[p1] * w[p1];
}

void launch_lambda_scatter(int64_t nl, const int64_t* lam_pos, const double* scale, const double* w,
                           double* out, hipStream_t st) {
  if (nl <= 0) return;
  hipLaunchKernelGGL(k_lambda_scatter, dim3((nl + 255) / 256), dim3(256), 0, st, nl, lam_pos, scale,
                     w, out);
}

__global__ __launch_bounds__(256) void k_coarse_gather(int64_t nc, const int64_t* __restrict__ ptr,
                                                       const int64_t* __restrict__ pos,
                                                       const double* __restrict__ src,
                                                       const double* __restrict__ add,
                                                       double* __restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double s = add ? add[c] : 0.0;
  for (int64_t j = ptr[c]; j < ptr[c + 1]; ++j) s += src[pos[j]];
  out[c] = s;
}

void launch_coarse_gather(int64_t nc, const int64_t* ptr, const int64_t* pos, const double* src,
                          const double* add, double* out, hipStream_t st) {
  if (nc <= 0) return;
  hipLaunchKernelGGL(k_coarse_gather, dim3((nc + 255) / 256), dim3(256), 0, st, nc, ptr, pos, src,
                     add, out);
}

__global__ __launch_bounds__(256) void k_fill_U(const DevSub* __restrict__ subs,
                                                const int32_t* __restrict__ d_lam,
                                                const double* __restrict__ ubar,
                                                const int32_t* __restrict__ c_loc,
                                                const double* __restrict__ uc, double* __restrict__ U) {
  const DevSub S = subs[blockIdx.y];
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < S.ndp) {
    const int32_t l = d_lam[S.dvec_off + k];
    U[S.vec_off + S.o_d + k] = l >= 0 ? ubar[l] : 0.0;
  }
  if (k < S.ncp) {
    const int32_t c = c_loc[S.cvec_off + k];
    U[S.vec_off + S.o_c + k] = c >= 0 ? uc[c] : 0.0;
  }
}

void launch_fill_U(int ns, int max_m, const DevSub* subs, const int32_t* d_lam, const double* ubar,
                   const int32_t* c_loc, const double* uc, double* U, hipStream_t st) {
  if (ns <= 0 || max_m <= 0) return;
  hipLaunchKernelGGL(k_fill_U, dim3((max_m + 255) / 256, ns), dim3(256), 0, st, subs, d_lam, ubar,
                     c_loc, uc, U);
}

__global__ __launch_bounds__(256) void k_split_rhs(int64_t n, const int32_t* __restrict__ vdof,
                                                   const double* __restrict__ vw,
                                                   const double* __restrict__ f, double* __restrict__ vec) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t d = vdof[i];
  vec[i] = d >= 0 ? f[d] * vw[i] : 0.0;
}

void launch_split_rhs(int64_t n, const int32_t* vdof, const double* vw, const double* f, double* vec,
                      hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_split_rhs, dim3((n + 255) / 256), dim3(256), 0, st, n, vdof, vw, f, vec);
}

__global__ __launch_bounds__(256) void k_gather_global(int64_t nf, const int8_t* __restrict__ kind,
                                                       const int64_t* __restrict__ a,
                                                       const double* __restrict__ vec,
                                                       const double* __restrict__ ubar,
                                                       const double* __restrict__ uc,
                                                       double* __restrict__ x) {
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= nf) return;
  const int k = kind[d];
  const int64_t s = a[d];
  x[d] = k == 0 ? vec[s] : (k == 1 ? ubar[s] : uc[s]);
}

void launch_gather_global(int64_t nf, const int8_t* kind, const int64_t* a, const double* vec,
                          const double* ubar, const double* uc, double* x, hipStream_t st) {
  if (nf <= 0) return;
  hipLaunchKernelGGL(k_gather_global, dim3((nf + 255) / 256), dim3(256), 0, st, nf, kind, a, vec,
                     ubar, uc, x);
}

// CSR rows, eight lanes per row: lane g of a row's group takes the nonzeros rp[i] + g,
// + 8, ... (coalesced runs of the row), the eight partial sums are added by a fixed
// shuffle tree.  RESID: out = f - A x (f may be null), else out = A x.
template <bool RESID>
__global__ __launch_bounds__(256) void k_csr_rows(int64_t n, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ col,
                                                  const double* __restrict__ val,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ f,
                                                  double* __restrict__ out) {
  const int64_t i = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 3;
  const int g = threadIdx.x & 7;
  double s = 0.0;
  if (i < n)
    for (int32_t j = rp[i] + g; j < rp[i + 1]; j += 8) s += val[j] * x[col[j]];
  s += __shfl_xor(s, 4);
  s += __shfl_xor(s, 2);
  s += __shfl_xor(s, 1);
  if (i < n && g == 0) out[i] = RESID ? (f ? f[i] : 0.0) - s : s;
}

void launch_csr_residual(int64_t n, const int32_t* rp, const int32_t* col, const double* val,
                         const double* x, const double* f, double* r, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_csr_rows<true>, dim3((n * 8 + 255) / 256), dim3(256), 0, st, n, rp, col, val, x,
                     f, r);
}

void launch_csr_spmv(int64_t n, const int32_t* rp, const int32_t* col, const double* val,
                     const double* x, double* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_csr_rows<false>, dim3((n * 8 + 255) / 256), dim3(256), 0, st, n, rp, col, val,
                     x, nullptr, y);
}

struct DotArgs {
  const double* a[3];
  const double* b[3];
};

__global__ __launch_bounds__(256) void k_dots(int64_t n, int np, DotArgs args, double* __restrict__ partial) {
  __shared__ double red[3][256];
  double s[3] = {0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    for (int p = 0; p < np; ++p) s[p] += args.a[p][i] * args.b[p][i];
  for (int p = 0; p < np; ++p) red[p][threadIdx.x] = s[p];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int p = 0; p < np; ++p) red[p][threadIdx.x] += red[p][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int p = 0; p < np; ++p) partial[p * kDotBlocks + blockIdx.x] = red[p][0];
}

void launch_dots(int64_t n, int npairs, const double* const* a, const double* const* b,
                 double* partial, hipStream_t st) {
  DotArgs args{};
  for (int p = 0; p < npairs; ++p) {
    args.a[p] = a[p];
    args.b[p] = b[p];
  }
  hipLaunchKernelGGL(k_dots, dim3(kDotBlocks), dim3(256), 0, st, n, npairs, args, partial);
}

__global__ __launch_bounds__(256) void k_axpy2(int64_t n, double alpha, const double* __restrict__ p,
                                               const double* __restrict__ q, double* __restrict__ lam,
                                               double* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  lam[i] += alpha * p[i];
  r[i] -= alpha * q[i];
}

void launch_axpy2(int64_t n, double alpha, const double* p, const double* q, double* lam, double* r,
                  hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy2, dim3((n + 255) / 256), dim3(256), 0, st, n, alpha, p, q, lam, r);
}

__global__ __launch_bounds__(256) void k_xpby(int64_t n, const double* __restrict__ z, double beta,
                                              double* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[i] = z[i] + beta * p[i];
}

void launch_xpby(int64_t n, const double* z, double beta, double* p, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_xpby, dim3((n + 255) / 256), dim3(256), 0, st, n, z, beta, p);
}

__global__ __launch_bounds__(256) void k_axpy(int64_t n, double alpha, const double* __restrict__ x,
                                              double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] += alpha * x[i];
}

void launch_axpy(int64_t n, double alpha, const double* x, double* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy, dim3((n + 255) / 256), dim3(256), 0, st, n, alpha, x, y);
}

// ---------------------------------------------------------------------------
// PCPG scalars on the device (so one iteration is a fixed kernel chain that a
// HIP graph can hold): sc[kScRZ] = r.z of the current residual, then pq, alpha,
// beta, rr.  `host` (pinned, mapped; may be null) receives the same slots for
// the host's convergence checks.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, double* red) {
  red[threadIdx.x] = part[threadIdx.x];
  __syncthreads();
  for (int s = kDotBlocks / 2; s >= 1; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

__global__ __launch_bounds__(kDotBlocks) void k_pcpg_alpha(const double* __restrict__ partial,
                                                           double* __restrict__ sc, double* host) {
  __shared__ double red[kDotBlocks];
  const double pq = sum_partials(partial, red);
  if (threadIdx.x == 0) {
    const double alpha = fabs(pq) >= 1e-300 ? sc[kScRZ] / pq : 0.0;
    sc[kScPQ] = pq;
    sc[kScAlpha] = alpha;
    if (host) {
      host[kScPQ] = pq;
      host[kScAlpha] = alpha;
    }
  }
}

__global__ __launch_bounds__(kDotBlocks) void k_pcpg_beta(const double* __restrict__ partial,
                                                          double* __restrict__ sc, double* host,
                                                          int init) {
  __shared__ double red[kDotBlocks];
  const double rr = sum_partials(partial, red);
  __syncthreads();
  const double rz = sum_partials(partial + kDotBlocks, red);
  if (threadIdx.x == 0) {
    const double beta = init ? 0.0 : rz / sc[kScRZ];
    sc[kScBeta] = beta;
    sc[kScRZ] = rz;
    sc[kScRR] = rr;
    if (host) {
      host[kScBeta] = beta;
      host[kScRZ] = rz;
      host[kScRR] = rr;
    }
  }
}

__global__ __launch_bounds__(256) void k_axpy2_dev(int64_t n, const double* __restrict__ sc,
                                                   const double* __restrict__ p,
                                                   const double* __restrict__ q,
                                                   double* __restrict__ lam, double* __restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double alpha = sc[kScAlpha];
  lam[i] += alpha * p[i];
  r[i] -= alpha * q[i];
}

__global__ __launch_bounds__(256) void k_xpby_dev(int64_t n, const double* __restrict__ z,
                                                  const double* __restrict__ sc,
                                                  double* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  p[i] = z[i] + sc[kScBeta] * p[i];
}

// q[l] = scale[p0] w[p0] + scale[p1] w[p1] over the two D slots of multiplier l (as
// k_lambda_scatter), plus the block partial sums of a_k . y_k for np <= 2 pairs, y_k = b_k
// or, for a null b_k, q itself.  Blocks, threads and the reduction tree are those of
// k_dots, so the partials equal a separate k_dots' bitwise.
__global__ __launch_bounds__(256) void k_scatter_dot(int64_t nl, const int64_t* __restrict__ pos,
                                                     const double* __restrict__ scale,
                                                     const double* __restrict__ w,
                                                     double* __restrict__ q, int np, DotArgs args,
                                                     double* __restrict__ partial) {
  __shared__ double red[2][256];
  double s[2] = {0.0, 0.0};
  for (int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x; l < nl; l += (int64_t)gridDim.x * 256) {
    const int64_t p0 = pos[2 * l], p1 = pos[2 * l + 1];
    const double ql = scale[p0] * w[p0] + scale[p1] * w[p1];
    q[l] = ql;
    for (int p = 0; p < np; ++p) s[p] += args.a[p][l] * (args.b[p] ? args.b[p][l] : ql);
  }
  for (int p = 0; p < np; ++p) red[p][threadIdx.x] = s[p];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int p = 0; p < np; ++p) red[p][threadIdx.x] += red[p][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int p = 0; p < np; ++p) partial[p * kDotBlocks + blockIdx.x] = red[p][0];
}

void launch_scatter_dot(int64_t nl, const int64_t* lam_pos, const double* scale, const double* w,
                        double* q, int npairs, const double* const* a, const double* const* b,
                        double* partial, hipStream_t st) {
  DotArgs args{};
  for (int p = 0; p < npairs; ++p) {
    args.a[p] = a[p];
    args.b[p] = b[p];
  }
  hipLaunchKernelGGL(k_scatter_dot, dim3(kDotBlocks), dim3(256), 0, st, nl, lam_pos, scale, w, q,
                     npairs, args, partial);
}

// Every block forms pq and alpha = rz / pq from the k_dots partials (as k_pcpg_alpha) and
// applies lam += alpha p, r -= alpha q to its slice; block 0 records the scalars and keeps
// rz for the next beta (slot kScRZold, which no block of this launch reads).
__global__ __launch_bounds__(kDotBlocks) void k_alpha_axpy2(int64_t n, const double* __restrict__ partial,
                                                            double* __restrict__ sc, double* host,
                                                            const double* __restrict__ p,
                                                            const double* __restrict__ q,
                                                            double* __restrict__ lam,
                                                            double* __restrict__ r) {
  __shared__ double red[kDotBlocks];
  const double pq = sum_partials(partial, red);
  const double alpha = fabs(pq) >= 1e-300 ? sc[kScRZ] / pq : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[kScRZold] = sc[kScRZ];
    sc[kScPQ] = pq;
    sc[kScAlpha] = alpha;
    if (host) {
      host[kScPQ] = pq;
      host[kScAlpha] = alpha;
      __threadfence_system();
    }
  }
  const int64_t i = (int64_t)blockIdx.x * kDotBlocks + threadIdx.x;
  if (i >= n) return;
  lam[i] += alpha * p[i];
  r[i] -= alpha * q[i];
}

// Every block forms rr, rz and beta = rz / rz_old (as k_pcpg_beta) and applies
// p = z + beta p to its slice; block 0 records the scalars.
__global__ __launch_bounds__(kDotBlocks) void k_beta_xpby(int64_t n, const double* __restrict__ partial,
                                                          double* __restrict__ sc, double* host,
                                                          const double* __restrict__ z,
                                                          double* __restrict__ p) {
  __shared__ double red[kDotBlocks];
  const double rr = sum_partials(partial, red);
  __syncthreads();
  const double rz = sum_partials(partial + kDotBlocks, red);
  const double beta = rz / sc[kScRZold];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[kScBeta] = beta;
    sc[kScRZ] = rz;
    sc[kScRR] = rr;
    const double seq = sc[kScSeq] + 1.0;
    sc[kScSeq] = seq;
    if (host) {
      host[kScBeta] = beta;
      host[kScRZ] = rz;
      host[kScRR] = rr;
      __threadfence_system();  // the scalars reach the host before the sequence number
      *(volatile double*)(host + kScSeq) = seq;
    }
  }
  const int64_t i = (int64_t)blockIdx.x * kDotBlocks + threadIdx.x;
  if (i >= n) return;
  p[i] = z[i] + beta * p[i];
}

void launch_alpha_axpy2(int64_t n, const double* partial, double* sc, double* host, const double* p,
                        const double* q, double* lam, double* r, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_alpha_axpy2, dim3((n + kDotBlocks - 1) / kDotBlocks), dim3(kDotBlocks), 0, st,
                     n, partial, sc, host, p, q, lam, r);
}
void launch_beta_xpby(int64_t n, const double* partial, double* sc, double* host, const double* z,
                      double* p, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_beta_xpby, dim3((n + kDotBlocks - 1) / kDotBlocks), dim3(kDotBlocks), 0, st, n,
                     partial, sc, host, z, p);
}

void launch_pcpg_alpha(const double* partial, double* sc, double* host, hipStream_t st) {
  hipLaunchKernelGGL(k_pcpg_alpha, dim3(1), dim3(kDotBlocks), 0, st, partial, sc, host);
}
void launch_pcpg_beta(const double* partial, double* sc, double* host, int init, hipStream_t st) {
  hipLaunchKernelGGL(k_pcpg_beta, dim3(1), dim3(kDotBlocks), 0, st, partial, sc, host, init);
}
void launch_axpy2_dev(int64_t n, const double* sc, const double* p, const double* q, double* lam,
                      double* r, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy2_dev, dim3((n + 255) / 256), dim3(256), 0, st, n, sc, p, q, lam, r);
}
void launch_xpby_dev(int64_t n, const double* z, const double* sc, double* p, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_xpby_dev, dim3((n + 255) / 256), dim3(256), 0, st, n, z, sc, p);
}

// ---------------------------------------------------------------------------
// Deluxe scaling: per interface face F, A_F = S_s0,FF + S_s1,FF and B_F = S_s1,FF
// ---------------------------------------------------------------------------
// element (a, b) of subdomain s's FP64 Dirichlet block (lower-tile store), a, b
// absolute D slots of s
__device__ __forceinline__ double p_elem(const DevSub& S, const double* __restrict__ P, int64_t sa,
                                         int64_t sb) {
  int i = int(sa - S.dvec_off), j = int(sb - S.dvec_off);
  if (i < j) {
    const int t = i;
    i = j;
    j = t;
  }
  const int64_t I = i >> 6, J = j >> 6;
  return P[S.p_off + (I * (I + 1) / 2 + J) * kTileElems + (i & 63) * 64 + (j & 63)];
}

// [[A_F, I], [I, 0]] of every face as dense lower tiles, column-major inside a tile
// (the factor-pool convention); diagonal tiles are filled symmetric, padding rows of
// A_F get a unit diagonal and no coupling.
__global__ __launch_bounds__(256) void k_face_pool(const FaceDev* __restrict__ faces,
                                                   const DevSub* __restrict__ subs,
                                                   const double* __restrict__ P,
                                                   const int64_t* __restrict__ slot0,
                                                   const int64_t* __restrict__ slot1,
                                                   double* __restrict__ fpool) {
  const FaceDev F = faces[blockIdx.y];
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ta = 2 * F.tf;
  if (e >= ta * (ta + 1) / 2 * kTileElems) return;
  const int64_t t = e >> 12;
  int I = int((sqrt(8.0 * double(t) + 1.0) - 1.0) * 0.5);
  while (int64_t(I) * (I + 1) / 2 > t) --I;
  while (int64_t(I + 1) * (I + 2) / 2 <= t) ++I;
  const int J = int(t - int64_t(I) * (I + 1) / 2);
  const int w = int(e & 4095);
  const int r = I * 64 + (w & 63), c = J * 64 + (w >> 6);
  const int a = r > c ? r : c, b = r > c ? c : r;
  const int nfp = F.tf * 64;
  double v = 0.0;
  if (a < nfp) {
    if (a < F.nf && b < F.nf) {
      const int64_t ka = F.vec_off + a, kb = F.vec_off + b;
      v = p_elem(subs[F.s0], P, slot0[ka], slot0[kb]) + p_elem(subs[F.s1], P, slot1[ka], slot1[kb]);
    } else {
      v = a == b ? 1.0 : 0.0;
    }
  } else if (b < nfp) {
    v = (a - nfp == b && b < F.nf) ? 1.0 : 0.0;
  }
  fpool[F.pool_base + e] = v;
}

// A_F^-1 = -(trailing block of the factored pool) and B_F, as lower tiles row-major
// inside a tile (the symmetric-product layout)
__global__ __launch_bounds__(256) void k_face_extract(const FaceDev* __restrict__ faces,
                                                      const DevSub* __restrict__ subs,
                                                      const double* __restrict__ P,
                                                      const int64_t* __restrict__ slot1,
                                                      const double* __restrict__ fpool,
                                                      double* __restrict__ ops) {
  const FaceDev F = faces[blockIdx.y];
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nfp = F.tf * 64;
  if (e >= sym_elems(nfp)) return;
  int i, j;
  sym_tile_rc(e, i, j);
  const int a = (i > j ? i : j) + nfp, b = (i > j ? j : i) + nfp;
  ops[F.ainv_off + e] =
      -fpool[F.pool_base + ((int64_t)(a >> 6) * ((a >> 6) + 1) / 2 + (b >> 6)) * kTileElems +
             (int64_t)(b & 63) * 64 + (a & 63)];
  ops[F.b_off + e] = (i < F.nf && j < F.nf)
                         ? p_elem(subs[F.s1], P, slot1[F.vec_off + i], slot1[F.vec_off + j])
                         : 0.0;
}

void launch_face_setup(int64_t nfaces, int max_tf, const FaceDev* faces, const DevSub* subs,
                       const double* P, const int64_t* slot0, const int64_t* slot1, double* fpool,
                       hipStream_t st) {
  if (nfaces <= 0) return;
  const int64_t ta = 2 * int64_t(max_tf);
  const int64_t n = ta * (ta + 1) / 2 * kTileElems;
  hipLaunchKernelGGL(k_face_pool, dim3((n + 255) / 256, nfaces), dim3(256), 0, st, faces, subs, P,
                     slot0, slot1, fpool);
}

void launch_face_extract(int64_t nfaces, int max_tf, const FaceDev* faces, const DevSub* subs,
                         const double* P, const int64_t* slot1, const double* fpool, double* ops,
                         hipStream_t st) {
  if (nfaces <= 0) return;
  const int64_t n = sym_tile_elems(int64_t(max_tf) * 64);
  hipLaunchKernelGGL(k_face_extract, dim3((n + 255) / 256, nfaces), dim3(256), 0, st, faces, subs,
                     P, slot1, fpool, ops);
}

// D_F = A_F^-1 B_F of every face as full tiles (row-major inside a tile), one work-group
// per output tile (i, j) of a face: D_ij = sum_k (A^-1)_ik B_kj on FP64 MFMA (wave w
// holds output rows 16 w .. 16 w + 15, four 16x16 blocks), the two input tiles of each k
// staged in LDS; then rounded to T.
template <typename T>
__global__ __launch_bounds__(256) void k_face_D(int max_tf, const FaceDev* __restrict__ faces,
                                                const double* __restrict__ ops, T* __restrict__ D) {
  using acc_t = Mfma<double>::acc_t;
  const FaceDev F = faces[blockIdx.y];
  const int tf = F.tf;
  const int i = int(blockIdx.x) / max_tf, j = int(blockIdx.x) % max_tf;  // grid.x = max_tf^2
  if (i >= tf || j >= tf) return;
  __shared__ double sA[64 * kLd], sB[64 * kLd];
  const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6;
  const int lr = l & 15, lk = l >> 4;
  acc_t acc[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) acc[cb] = acc_t{0.0, 0.0, 0.0, 0.0};
  // the two tiles of step k + 1 are loaded into registers while step k's products run;
  // the operators are stored as lower tiles (row-major inside a tile), so a tile (r, c)
  // with r < c is staged as the transpose of the stored tile (c, r)
  double ra[16], rb[16];
  auto fetch = [&](int k) {
    const double* srcA = ops + F.ainv_off + (int64_t)(i >= k ? i * (i + 1) / 2 + k : k * (k + 1) / 2 + i) * kTileElems;
    const double* srcB = ops + F.b_off + (int64_t)(k >= j ? k * (k + 1) / 2 + j : j * (j + 1) / 2 + k) * kTileElems;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      ra[u] = srcA[tid + 256 * u];
      rb[u] = srcB[tid + 256 * u];
    }
  };
  fetch(0);
  for (int k = 0; k < tf; ++k) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int e = tid + 256 * u, r = e >> 6, c = e & 63;  // stored (row r, column c)
      if (i < k) sA[c * kLd + r] = ra[u];
      else sA[r * kLd + c] = ra[u];
      if (k < j) sB[c * kLd + r] = rb[u];
      else sB[r * kLd + c] = rb[u];
    }
    __syncthreads();
    if (k + 1 < tf) fetch(k + 1);
#pragma unroll 4
    for (int kk = 0; kk < 64; kk += 4) {
      const double a = sA[(16 * wv + lr) * kLd + kk + lk];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) acc[cb] = Mfma<double>::op(a, sB[(kk + lk) * kLd + 16 * cb + lr], acc[cb]);
    }
    __syncthreads();
  }
  T* out = D + F.d_off + ((int64_t)i * tf + j) * kTileElems;
#pragma unroll
  for (int cb = 0; cb < 4; ++cb)
#pragma unroll
    for (int r = 0; r < 4; ++r)
      out[(16 * wv + Mfma<double>::row(lk, r)) * 64 + 16 * cb + lr] = (T)acc[cb][r];
}

template <typename T>
void launch_face_D(int64_t nfaces, int max_tf, const FaceDev* faces, const double* ops, T* D,
                   hipStream_t st) {
  if (nfaces <= 0 || max_tf <= 0) return;
  hipLaunchKernelGGL(k_face_D<T>, dim3(max_tf * max_tf, nfaces), dim3(256), 0, st, max_tf, faces, ops, D);
}
template void launch_face_D<float>(int64_t, int, const FaceDev*, const double*, float*, hipStream_t);
template void launch_face_D<double>(int64_t, int, const FaceDev*, const double*, double*, hipStream_t);

// out = D_F x (TRANS: D_F^T x) on every face, one work-group per (output block, face):
// x_k = src[map[vec_off + k]] (0 where map < 0), out[vec_off + k].  Thread: column = lane,
// rows 16 w .. 16 w + 15 of each tile; FP64 sums over the tiles of the block row (column)
// in a fixed order, then over the lanes (D x: halving butterfly) or the waves (D^T x).
template <typename T, bool TRANS>
__global__ __launch_bounds__(256) void k_face_dmv(const FaceDev* __restrict__ faces,
                                                  const T* __restrict__ D,
                                                  const int32_t* __restrict__ map,
                                                  const double* __restrict__ src,
                                                  double* __restrict__ out) {
  __shared__ double xs[64];
  __shared__ double colp[4][64];
  const FaceDev F = faces[blockIdx.y];
  const int tf = F.tf, ob = blockIdx.x;
  if (ob >= tf) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const T* Db = D + F.d_off;
  const int32_t* mp = map + F.vec_off;
  if constexpr (!TRANS) {
    // row r = 16 w + q of block row ob: sum over column blocks jb and columns c = lane
    double rv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) rv[q] = 0.0;
    for (int jb = 0; jb < tf; ++jb) {
      const T* A = Db + ((int64_t)ob * tf + jb) * kTileElems;
      T at[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) at[q] = A[(16 * wv + q) * 64 + lane];
      const int32_t m = mp[jb * 64 + lane];
      const double xc = m >= 0 ? src[m] : 0.0;
#pragma unroll
      for (int q = 0; q < 16; ++q) rv[q] = fma((double)at[q], xc, rv[q]);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const bool hi = lane & 32;
      const double keep = hi ? rv[u + 8] : rv[u], send = hi ? rv[u] : rv[u + 8];
      rv[u] = keep + __shfl_xor(send, 32);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool hi = lane & 16;
      const double keep = hi ? rv[u + 4] : rv[u], send = hi ? rv[u] : rv[u + 4];
      rv[u] = keep + __shfl_xor(send, 16);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const bool hi = lane & 8;
      const double keep = hi ? rv[u + 2] : rv[u], send = hi ? rv[u] : rv[u + 2];
      rv[u] = keep + __shfl_xor(send, 8);
    }
    {
      const bool hi = lane & 4;
      const double keep = hi ? rv[1] : rv[0], send = hi ? rv[0] : rv[1];
      rv[0] = keep + __shfl_xor(send, 4);
    }
    double rs = rv[0];
    rs += __shfl_xor(rs, 2);
    rs += __shfl_xor(rs, 1);
    if ((lane & 3) == 0) {
      const int row = 16 * wv + ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 +
                      ((lane >> 2) & 1);
      out[F.vec_off + ob * 64 + row] = rs;
    }
  } else {
    // column c = lane of block column ob: sum over row blocks ib and rows 16 w + q
    double cs = 0.0;
    for (int ib = 0; ib < tf; ++ib) {
      const T* A = Db + ((int64_t)ib * tf + ob) * kTileElems;
      T at[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) at[q] = A[(16 * wv + q) * 64 + lane];
      if (tid < 64) {
        const int32_t m = mp[ib * 64 + tid];
        xs[tid] = m >= 0 ? src[m] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 16; ++q) cs = fma((double)at[q], xs[16 * wv + q], cs);
      __syncthreads();
    }
    colp[wv][lane] = cs;
    __syncthreads();
    if (tid < 64)
      out[F.vec_off + ob * 64 + tid] = (colp[0][tid] + colp[1][tid]) + (colp[2][tid] + colp[3][tid]);
  }
}

template <typename T>
void launch_face_dmv(bool trans, int64_t nfaces, int max_tf, const FaceDev* faces, const T* D,
                     const int32_t* map, const double* src, double* out, hipStream_t st) {
  if (nfaces <= 0 || max_tf <= 0) return;
  if (trans)
    hipLaunchKernelGGL((k_face_dmv<T, true>), dim3(max_tf, nfaces), dim3(256), 0, st, faces, D, map, src, out);
  else
    hipLaunchKernelGGL((k_face_dmv<T, false>), dim3(max_tf, nfaces), dim3(256), 0, st, faces, D, map, src, out);
}
template void launch_face_dmv<float>(bool, int64_t, int, const FaceDev*, const float*, const int32_t*,
                                     const double*, double*, hipStream_t);
template void launch_face_dmv<double>(bool, int64_t, int, const FaceDev*, const double*, const int32_t*,
                                      const double*, double*, hipStream_t);

// x (D slots) from g = D_F r on every face: side s0 takes g, side s1 takes g - r
__global__ __launch_bounds__(256) void k_deluxe_split(int64_t nv, const int32_t* __restrict__ map_lam,
                                                      const int64_t* __restrict__ slot0,
                                                      const int64_t* __restrict__ slot1,
                                                      const double* __restrict__ g,
                                                      const double* __restrict__ r,
                                                      double* __restrict__ xs) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= nv) return;
  const int32_t l = map_lam[k];
  if (l < 0) return;
  xs[slot0[k]] = g[k];
  xs[slot1[k]] = g[k] - r[l];
}

// u = y(s0 side) + y(s1 side) on every face (0 on padding)
__global__ __launch_bounds__(256) void k_deluxe_sum(int64_t nv, const int32_t* __restrict__ map_lam,
                                                    const int64_t* __restrict__ slot0,
                                                    const int64_t* __restrict__ slot1,
                                                    const double* __restrict__ y,
                                                    double* __restrict__ u) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= nv) return;
  u[k] = map_lam[k] >= 0 ? y[slot0[k]] + y[slot1[k]] : 0.0;
}

// z = h - y(s1 side), scattered to the multipliers
__global__ __launch_bounds__(256) void k_deluxe_out(int64_t nv, const int32_t* __restrict__ map_lam,
                                                    const int64_t* __restrict__ slot1,
                                                    const double* __restrict__ h,
                                                    const double* __restrict__ y,
                                                    double* __restrict__ z) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= nv) return;
  const int32_t l = map_lam[k];
  if (l >= 0) z[l] = h[k] - y[slot1[k]];
}

// Deluxe average of the two sides' interface values (recovery): jump j = u_s1 - u_s0 per
// face-vector slot (0 on padding), then ubar = u_s0 + D_F j, scattered to the multipliers.
__global__ __launch_bounds__(256) void k_deluxe_jump(int64_t nv, const int32_t* __restrict__ map_lam,
                                                     const int64_t* __restrict__ slot0,
                                                     const int64_t* __restrict__ slot1,
                                                     const double* __restrict__ u,
                                                     double* __restrict__ jmp) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= nv) return;
  jmp[k] = map_lam[k] >= 0 ? u[slot1[k]] - u[slot0[k]] : 0.0;
}
__global__ __launch_bounds__(256) void k_deluxe_avg(int64_t nv, const int32_t* __restrict__ map_lam,
                                                    const int64_t* __restrict__ slot0,
                                                    const double* __restrict__ u,
                                                    const double* __restrict__ g,
                                                    double* __restrict__ ubar) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= nv) return;
  const int32_t l = map_lam[k];
  if (l >= 0) ubar[l] = u[slot0[k]] + g[k];
}
void launch_deluxe_jump(int64_t nv, const int32_t* map_lam, const int64_t* slot0, const int64_t* slot1,
                        const double* u, double* jmp, hipStream_t st) {
  if (nv > 0)
    hipLaunchKernelGGL(k_deluxe_jump, dim3((nv + 255) / 256), dim3(256), 0, st, nv, map_lam, slot0, slot1,
                       u, jmp);
}
void launch_deluxe_avg(int64_t nv, const int32_t* map_lam, const int64_t* slot0, const double* u,
                       const double* g, double* ubar, hipStream_t st) {
  if (nv > 0)
    hipLaunchKernelGGL(k_deluxe_avg, dim3((nv + 255) / 256), dim3(256), 0, st, nv, map_lam, slot0, u, g,
                       ubar);
}

void launch_deluxe_split(int64_t nv, const int32_t* map_lam, const int64_t* slot0, const int64_t* slot1,
                         const double* g, const double* r, double* xs, hipStream_t st) {
  if (nv > 0)
    hipLaunchKernelGGL(k_deluxe_split, dim3((nv + 255) / 256), dim3(256), 0, st, nv, map_lam, slot0,
                       slot1, g, r, xs);
}
void launch_deluxe_sum(int64_t nv, const int32_t* map_lam, const int64_t* slot0, const int64_t* slot1,
                       const double* y, double* u, hipStream_t st) {
  if (nv > 0)
    hipLaunchKernelGGL(k_deluxe_sum, dim3((nv + 255) / 256), dim3(256), 0, st, nv, map_lam, slot0, slot1,
                       y, u);
}
void launch_deluxe_out(int64_t nv, const int32_t* map_lam, const int64_t* slot1, const double* h,
                       const double* y, double* z, hipStream_t st) {
  if (nv > 0)
    hipLaunchKernelGGL(k_deluxe_out, dim3((nv + 255) / 256), dim3(256), 0, st, nv, map_lam, slot1, h,
                       y, z);
}

// x *= alpha, y *= alpha
__global__ __launch_bounds__(256) void k_scale2(int64_t n, double alpha, double* __restrict__ x,
                                               double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] *= alpha;
  y[i] *= alpha;
}

void launch_scale2(int64_t n, double alpha, double* x, double* y, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale2, dim3((n + 255) / 256), dim3(256), 0, st, n, alpha, x, y);
}

}  // namespace emag

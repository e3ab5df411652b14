// Learned lambda0 network: gfx950 kernels (bf16 MFMA 16x16x32, FP32 accumulate)
// and the C ABI of include/emagddm_mpnn.h.
//
// Every reduction is a fixed-order sum (per-block partials, then a fixed-order
// pass), so forward, gradients and training are bitwise repeatable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "emagddm_mpnn.h"
#include "mpnn.h"

using namespace emag;

namespace {

#define MPCHECK(x)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      throw std::runtime_error(std::string(#x) + " failed: " + hipGetErrorString(e_));    \
  } while (0)

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16v8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// D[4(l/16)+j][l%16] += sum_k A[l%16][k] B[k][l%16] over k < 32 with lane l holding
// A[l%16][8(l/16)+j] and B[8(l/16)+j][l%16], j < 8 (checked by emag_mpnn_mfma_probe)
__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16v8, a),
                                                 __builtin_bit_cast(bf16v8, b), c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 pack8(float4 x, float4 y) {
  return bf16x8{(short)f2bf(x.x), (short)f2bf(x.y), (short)f2bf(x.z), (short)f2bf(x.w),
                (short)f2bf(y.x), (short)f2bf(y.y), (short)f2bf(y.z), (short)f2bf(y.w)};
}

// ---------------------------------------------------------------------------
// Row GEMM: out = [A0 | A1] (N x Kt, FP32, rounded to bf16 on load) times B
// (Kt x NC, given as Bt[n][k] in bf16).  One workgroup = 4 waves x 32 rows; K steps of
// 32 (k >= Kt read as zero; K0 and Kt are multiples of 8).  Bt is staged once per
// workgroup in LDS with the 16-byte k-chunks of row n XOR-swizzled by n, so the 16 lanes
// that read one k-chunk of 16 rows hit distinct banks.  LDS = NC * Kt * 2 bytes <= 64 KiB.
// ---------------------------------------------------------------------------
enum { kEpiStore = 0, kEpiTanhRes = 1, kEpiBwd = 2 };
constexpr int kMmWaveRows = 32, kMmRows = 128;

__device__ __forceinline__ int swz(int n, int q, int qmask) { return q ^ (n & qmask); }

template <int NC, int EPI>
__global__ __launch_bounds__(256) void k_mm(int64_t N, const float* __restrict__ A0, int K0,
                                            const float* __restrict__ A1, int K1,
                                            const unsigned short* __restrict__ Bt,
                                            float* __restrict__ out0, float* __restrict__ out1,
                                            const float* __restrict__ aux) {
  extern __shared__ unsigned short sB[];
  const int Kt = K0 + K1, nq = Kt / 8, qmask = (nq < 16 ? nq : 16) - 1;
  for (int e = threadIdx.x; e < NC * nq; e += 256) {
    const int n = e / nq, q = e % nq;
    const uint4 v = *reinterpret_cast<const uint4*>(Bt + (int64_t)n * Kt + 8 * q);
    *reinterpret_cast<uint4*>(sB + n * Kt + 8 * swz(n, q, qmask)) = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int64_t row0 = (int64_t)blockIdx.x * kMmRows + wv * kMmWaveRows;
  if (row0 >= N) return;
  constexpr int NT = NC / 16;
  f32x4 acc[2][NT];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int c = 0; c < NT; ++c) acc[h][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < Kt; k0 += 32) {
    const int k = k0 + 8 * lk;
    const bool kin = k < Kt;
    bf16x8 a[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t row = row0 + 16 * h + lr;
      float4 v0 = make_float4(0.f, 0.f, 0.f, 0.f), v1 = v0;
      if (row < N && kin) {
        const float* src = k < K0 ? A0 + row * K0 + k : A1 + row * K1 + (k - K0);
        v0 = *reinterpret_cast<const float4*>(src);
        v1 = *reinterpret_cast<const float4*>(src + 4);
      }
      a[h] = pack8(v0, v1);
    }
    const int q = k >> 3;
#pragma unroll
    for (int c = 0; c < NT; ++c) {
      const int n = 16 * c + lr;
      bf16x8 b = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
      if (kin) b = *reinterpret_cast<const bf16x8*>(sB + n * Kt + 8 * swz(n, q, qmask));
      acc[0][c] = mfma_bf16(a[0], b, acc[0][c]);
      acc[1][c] = mfma_bf16(a[1], b, acc[1][c]);
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int c = 0; c < NT; ++c)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t row = row0 + 16 * h + 4 * lk + j;
        if (row >= N) continue;
        const int n = 16 * c + lr;
        const float v = acc[h][c][j];
        if constexpr (EPI == kEpiStore) {
          out0[row * kMpH + n] = v;
        } else if constexpr (EPI == kEpiTanhRes) {
          const float t = tanhf(v);
          out1[row * kMpH + n] = t;
          out0[row * kMpH + n] = aux[row * kMpH + n] + t;
        } else {
          if (n < kMpH) out0[row * kMpH + n] = aux[row * kMpH + n] + v;
          else out1[row * kMpH + (n - kMpH)] = v;
        }
      }
}

template <int NC, int EPI>
void launch_mm(int64_t N, const float* A0, int K0, const float* A1, int K1, const unsigned short* Bt,
               float* out0, float* out1, const float* aux, hipStream_t st) {
  if (N <= 0) return;
  const size_t lds = size_t(NC) * (K0 + K1) * sizeof(unsigned short);
  hipLaunchKernelGGL((k_mm<NC, EPI>), dim3((N + kMmRows - 1) / kMmRows), dim3(256), lds, st, N, A0,
                     K0, A1, K1, Bt, out0, out1, aux);
}

// ---------------------------------------------------------------------------
// Aggregation over the batched graph: out[u] = base[u] + sum_k coef[k] in[b n + col[k]]
// (u = b n + i, k over row i in ascending order).  One wave per node, 2 floats per lane.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_agg(int64_t N, int64_t n, const int32_t* __restrict__ rp,
                                             const int32_t* __restrict__ col,
                                             const float* __restrict__ coef,
                                             const float* __restrict__ in,
                                             const float* __restrict__ base,
                                             float* __restrict__ out) {
  const int64_t u = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (u >= N) return;
  const int lane = threadIdx.x & 63;
  const int64_t b = u / n, i = u - b * n;
  float2 s = base ? reinterpret_cast<const float2*>(base + u * kMpH)[lane] : make_float2(0.f, 0.f);
  const int k1 = rp[i + 1];
  for (int k = rp[i]; k < k1; ++k) {
    const float c = coef[k];
    const float2 v = reinterpret_cast<const float2*>(in + (b * n + col[k]) * kMpH)[lane];
    s.x += c * v.x;
    s.y += c * v.y;
  }
  reinterpret_cast<float2*>(out + u * kMpH)[lane] = s;
}

void launch_agg(int64_t N, const MpGraph& g, const float* coef, const float* in, const float* base,
                float* out, hipStream_t st) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_agg, dim3((N + 3) / 4), dim3(256), 0, st, N, g.n, g.rp, g.col, coef, in,
                     base, out);
}

// y[u] = sum_n h[u][n] bf(w_out)[n]: one wave per node
__global__ __launch_bounds__(256) void k_head(int64_t N, const float* __restrict__ h,
                                              const unsigned short* __restrict__ wout,
                                              float* __restrict__ y) {
  const int64_t u = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (u >= N) return;
  const int lane = threadIdx.x & 63;
  const float2 v = reinterpret_cast<const float2*>(h + u * kMpH)[lane];
  float s = v.x * bf2f(wout[2 * lane]) + v.y * bf2f(wout[2 * lane + 1]);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) y[u] = s;
}

// gy[u] = 2 omega_b (y_u - tau_u) / B; lossp[block] = sum over the block of omega_b (y-tau)^2 / B
__global__ __launch_bounds__(256) void k_loss(int64_t N, int64_t n, double invB,
                                              const float* __restrict__ y,
                                              const float* __restrict__ tau,
                                              const float* __restrict__ omega,
                                              float* __restrict__ gy, double* __restrict__ lossp) {
  __shared__ double red[256];
  const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double l = 0.0;
  if (u < N) {
    const float w = omega[u / n];
    const float d = y[u] - tau[u];
    gy[u] = float(2.0 * invB) * w * d;
    l = double(w) * double(d) * double(d) * invB;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s >= 1; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) lossp[blockIdx.x] = red[0];
}

// gh[u][n] = gy[u] bf(w_out)[n]
__global__ __launch_bounds__(256) void k_head_bwd(int64_t N, const float* __restrict__ gy,
                                                  const unsigned short* __restrict__ wout,
                                                  float* __restrict__ gh) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= N * kMpH) return;
  gh[e] = gy[e / kMpH] * bf2f(wout[e % kMpH]);
}

// part[blk][n] = sum over the block's rows of gy[u] h[u][n] (FP32, fixed order)
__global__ __launch_bounds__(128) void k_colsum(int64_t N, int64_t rows_per_blk,
                                                const float* __restrict__ gy,
                                                const float* __restrict__ h,
                                                float* __restrict__ part) {
  const int n = threadIdx.x;
  const int64_t u0 = (int64_t)blockIdx.x * rows_per_blk;
  const int64_t u1 = u0 + rows_per_blk < N ? u0 + rows_per_blk : N;
  float s = 0.f;
  for (int64_t u = u0; u < u1; ++u) s += gy[u] * h[u * kMpH + n];
  part[(int64_t)blockIdx.x * kMpH + n] = s;
}

// gu = gh (1 - t^2)
__global__ __launch_bounds__(256) void k_gu(int64_t E, const float* __restrict__ gh,
                                            const float* __restrict__ t, float* __restrict__ gu) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= E) return;
  const float tv = t[e];
  gu[e] = gh[e] * (1.f - tv * tv);
}

// ---------------------------------------------------------------------------
// Weight gradient: part[blk][k][n] = sum_{i in block rows} bf(A[i][k]) bf(G[i][n]),
// A: N x KA, G: N x 128.  Chunks of 64 rows are staged transposed in LDS as
// bf16 ([k][i] and [n][i]) so both MFMA operands are 8-byte LDS reads; the
// (KA/16) x 8 output tiles are split over the 4 waves.
// ---------------------------------------------------------------------------
constexpr int kWgChunk = 64, kWgLd = kWgChunk + 8, kWgMaxBlocks = 256;

template <int KA>
__global__ __launch_bounds__(256) void k_wgrad(int64_t N, int64_t rows_per_blk,
                                               const float* __restrict__ A,
                                               const float* __restrict__ G,
                                               float* __restrict__ part) {
  __shared__ unsigned short sA[KA * kWgLd];
  __shared__ unsigned short sG[kMpH * kWgLd];
  constexpr int TW = (KA / 16) * 8 / 4;  // output tiles per wave
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  f32x4 acc[TW];
#pragma unroll
  for (int t = 0; t < TW; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int64_t i_begin = (int64_t)blockIdx.x * rows_per_blk;
  const int64_t i_end = i_begin + rows_per_blk < N ? i_begin + rows_per_blk : N;
  for (int64_t i0 = i_begin; i0 < i_end; i0 += kWgChunk) {
    __syncthreads();
    for (int e = tid; e < kWgChunk * KA; e += 256) {
      const int i = e / KA, k = e % KA;
      const int64_t row = i0 + i;
      sA[k * kWgLd + i] = f2bf(row < i_end ? A[row * KA + k] : 0.f);
    }
    for (int e = tid; e < kWgChunk * kMpH; e += 256) {
      const int i = e / kMpH, nn = e % kMpH;
      const int64_t row = i0 + i;
      sG[nn * kWgLd + i] = f2bf(row < i_end ? G[row * kMpH + nn] : 0.f);
    }
    __syncthreads();
#pragma unroll
    for (int ii = 0; ii < kWgChunk; ii += 32) {
#pragma unroll
      for (int t = 0; t < TW; ++t) {
        const int tile = wv * TW + t, kt = tile / 8, nt = tile % 8;
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(sA + (kt * 16 + lr) * kWgLd + ii + 8 * lk);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(sG + (nt * 16 + lr) * kWgLd + ii + 8 * lk);
        acc[t] = mfma_bf16(a, b, acc[t]);
      }
    }
  }
  float* out = part + (int64_t)blockIdx.x * KA * kMpH;
#pragma unroll
  for (int t = 0; t < TW; ++t) {
    const int tile = wv * TW + t, kt = tile / 8, nt = tile % 8;
#pragma unroll
    for (int j = 0; j < 4; ++j) out[(kt * 16 + 4 * lk + j) * kMpH + nt * 16 + lr] = acc[t][j];
  }
}

// out[e] = sum_{b < nb} part[b][e] (fixed order)
__global__ __launch_bounds__(256) void k_reduce_parts(int64_t nb, int64_t E,
                                                      const float* __restrict__ part,
                                                      float* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= E) return;
  float s = 0.f;
  for (int64_t b = 0; b < nb; ++b) s += part[b * E + e];
  out[e] = s;
}

// Adam (Kingma & Ba 2015, Algorithm 1) on the FP32 master weights
__global__ __launch_bounds__(256) void k_adam(int64_t P, float* __restrict__ w,
                                              const float* __restrict__ g, float* __restrict__ m,
                                              float* __restrict__ v, float lr, float b1, float b2,
                                              float eps, float bc1, float bc2) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= P) return;
  const float gv = g[e];
  const float mv = b1 * m[e] + (1.f - b1) * gv;
  const float vv = b2 * v[e] + (1.f - b2) * gv * gv;
  m[e] = mv;
  v[e] = vv;
  w[e] -= lr * (mv / bc1) / (sqrtf(vv / bc2) + eps);
}

// bf16 operand copies of the master weights:
//   inBt [128][16]: W_in^T;  fwd_l [128][256]: [Ws_l; Wn_l]^T;  bwd_l [256][128]: [Ws_l; Wn_l];
//   wout [128]
constexpr int64_t kBfIn = 0;
constexpr int64_t kBfFwd = kBfIn + int64_t(kMpH) * kMpF;
constexpr int64_t kBfLayer = 2 * int64_t(kMpH) * kMpH;
constexpr int64_t kBfBwd = kBfFwd + kMpL * kBfLayer;
constexpr int64_t kBfWout = kBfBwd + kMpL * kBfLayer;
constexpr int64_t kBfTotal = kBfWout + kMpH;

__global__ __launch_bounds__(256) void k_refresh_bf(const float* __restrict__ w,
                                                    unsigned short* __restrict__ bf) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= kBfTotal) return;
  float v;
  if (e < kBfFwd) {
    const int64_t n = e / kMpF, k = e % kMpF;
    v = w[kMpWin + k * kMpH + n];
  } else if (e < kBfBwd) {
    const int64_t r = e - kBfFwd, l = r / kBfLayer, q = r % kBfLayer;
    const int64_t n = q / (2 * kMpH), k = q % (2 * kMpH);
    const float* Wl = w + kMpLayer0 + l * kMpLayerStride;  // Ws then Wn, [H][H] each
    v = k < kMpH ? Wl[k * kMpH + n] : Wl[int64_t(kMpH) * kMpH + (k - kMpH) * kMpH + n];
  } else if (e < kBfWout) {
    const int64_t r = e - kBfBwd, l = r / kBfLayer, q = r % kBfLayer;
    v = w[kMpLayer0 + l * kMpLayerStride + q];  // row n of [Ws; Wn], column k
  } else {
    v = w[kMpWout + (e - kBfWout)];
  }
  bf[e] = f2bf(v);
}

// probe: lane l feeds a_frag/b_frag[8l..8l+7] and returns its accumulators
__global__ void k_mpnn_probe(const float* a, const float* b, float* d) {
  const int l = threadIdx.x;
  const bf16x8 av = pack8(make_float4(a[8 * l], a[8 * l + 1], a[8 * l + 2], a[8 * l + 3]),
                          make_float4(a[8 * l + 4], a[8 * l + 5], a[8 * l + 6], a[8 * l + 7]));
  const bf16x8 bv = pack8(make_float4(b[8 * l], b[8 * l + 1], b[8 * l + 2], b[8 * l + 3]),
                          make_float4(b[8 * l + 4], b[8 * l + 5], b[8 * l + 6], b[8 * l + 7]));
  const f32x4 r = mfma_bf16(av, bv, f32x4{0.f, 0.f, 0.f, 0.f});
  for (int j = 0; j < 4; ++j) d[4 * l + j] = r[j];
}

// Interface-edge features (nl x 16, FP32), all odd in (d, z):
//   dd = d / s_d, zz = z / s_z, w = rho_1 / (rho_0 + rho_1) - 1/2,
//   lr = log10(rho_0 / rho_1) / 3, |t| = |unit tangent| per axis
//   [dd, zz, dd w, zz w, dd lr, zz lr, zz|tx|, zz|ty|, zz|tz|, dd|tx|, dd|ty|, dd|tz|,
//    zz|lr|, dd|lr|, 0, 0]
// fsc = {1/s_d, 1/s_z, s_z, s_d} from the partial sums of d.d and z.z
// (nparts per pair; 0 in the inverses when a norm vanishes)
__global__ __launch_bounds__(256) void k_feature_scales(const double* __restrict__ partial, int nparts,
                                                        int64_t nl, double* __restrict__ fsc) {
  __shared__ double red[2][256];
  for (int q = 0; q < 2; ++q) {
    double v = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) v += partial[q * nparts + i];
    red[q][threadIdx.x] = v;
  }
  __syncthreads();
  for (int s = 128; s >= 1; s >>= 1) {
    if ((int)threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double sd = sqrt(red[0][0] / double(nl)), sz = sqrt(red[1][0] / double(nl));
    const bool ok = sd > 0.0 && sz > 0.0;
    fsc[0] = ok ? 1.0 / sd : 0.0;
    fsc[1] = ok ? 1.0 / sz : 0.0;
    fsc[2] = ok ? sz : 0.0;
    fsc[3] = sd;
  }
}

__global__ __launch_bounds__(256) void k_lambda_features(int64_t nl, const double* __restrict__ d,
                                                         const double* __restrict__ z,
                                                         const double* __restrict__ fsc,
                                                         const double* __restrict__ rho,
                                                         const int64_t* __restrict__ lam_pos,
                                                         const float* __restrict__ dir,
                                                         float* __restrict__ X) {
  const int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (l >= nl) return;
  const double dd = d[l] * fsc[0], zz = z[l] * fsc[1];
  const double r0 = rho[lam_pos[2 * l]], r1 = rho[lam_pos[2 * l + 1]];
  const double w = r1 / (r0 + r1) - 0.5;
  const double lr = log10(r0 / r1) / 3.0;
  const double tx = dir[3 * l], ty = dir[3 * l + 1], tz = dir[3 * l + 2];
  float* o = X + l * kMpF;
  o[0] = float(dd);
  o[1] = float(zz);
  o[2] = float(dd * w);
  o[3] = float(zz * w);
  o[4] = float(dd * lr);
  o[5] = float(zz * lr);
  o[6] = float(zz * fabs(tx));
  o[7] = float(zz * fabs(ty));
  o[8] = float(zz * fabs(tz));
  o[9] = float(dd * fabs(tx));
  o[10] = float(dd * fabs(ty));
  o[11] = float(dd * fabs(tz));
  o[12] = float(zz * fabs(lr));
  o[13] = float(dd * fabs(lr));
  o[14] = 0.f;
  o[15] = 0.f;
}

__global__ __launch_bounds__(256) void k_lambda_from_y(int64_t nl, const float* __restrict__ y,
                                                       const double* __restrict__ fsc,
                                                       double* __restrict__ lam) {
  const int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (l < nl) lam[l] = fsc[2] * double(y[l]);
}

template <typename T>
struct DevArr {
  T* p = nullptr;
  size_t n = 0;
  DevArr() = default;
  DevArr(const DevArr&) = delete;
  DevArr& operator=(const DevArr&) = delete;
  ~DevArr() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = count;
    if (count) MPCHECK(hipMalloc(&p, count * sizeof(T)));
  }
};

// splitmix64 -> uniform in [0, 1)
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

struct emag_mpnn {
  std::string err;
  DevArr<float> w, g, m, v;
  DevArr<unsigned short> bf;
  int64_t step = 0;
  MpGraph graph;
  // workspace for up to cap nodes
  int64_t cap = 0;
  DevArr<float> h, a, t, gh, o1, o2, gu, y, gy, part;
  DevArr<double> lossp;

  ~emag_mpnn() { mp_graph_free(graph); }

  void refresh(hipStream_t st) {
    hipLaunchKernelGGL(k_refresh_bf, dim3((kBfTotal + 255) / 256), dim3(256), 0, st, w.p, bf.p);
  }
  void reserve(int64_t N) {
    if (N <= cap) return;
    const size_t E = size_t(N) * kMpH;
    h.alloc(E * (kMpL + 1));
    a.alloc(E * kMpL);
    t.alloc(E * kMpL);
    gh.alloc(E);
    o1.alloc(E);
    o2.alloc(E);
    gu.alloc(E);
    y.alloc(size_t(N));
    gy.alloc(size_t(N));
    lossp.alloc(size_t((N + 255) / 256));
    part.alloc(size_t(kWgMaxBlocks) * kMpH * kMpH);  // every partial-sum layout fits
    cap = N;
  }
  const unsigned short* bf_in() const { return bf.p + kBfIn; }
  const unsigned short* bf_fwd(int l) const { return bf.p + kBfFwd + l * kBfLayer; }
  const unsigned short* bf_bwd(int l) const { return bf.p + kBfBwd + l * kBfLayer; }
  const unsigned short* bf_wout() const { return bf.p + kBfWout; }
  float* hl(int l) { return h.p + size_t(l) * size_t(cap) * kMpH; }
  float* al(int l) { return a.p + size_t(l) * size_t(cap) * kMpH; }
  float* tl(int l) { return t.p + size_t(l) * size_t(cap) * kMpH; }

  // forward keeping every layer's h, a, t for the backward pass
  void forward(const MpGraph& gr, int64_t N, const float* X, float* yout, hipStream_t st) {
    reserve(N);
    launch_mm<kMpH, kEpiStore>(N, X, kMpF, nullptr, 0, bf_in(), hl(0), nullptr, nullptr, st);
    for (int l = 0; l < kMpL; ++l) {
      launch_agg(N, gr, gr.cA, hl(l), nullptr, al(l), st);
      launch_mm<kMpH, kEpiTanhRes>(N, hl(l), kMpH, al(l), kMpH, bf_fwd(l), hl(l + 1), tl(l), hl(l), st);
    }
    hipLaunchKernelGGL(k_head, dim3((N + 3) / 4), dim3(256), 0, st, N, hl(kMpL), bf_wout(), yout);
  }

  void wgrad_into(int64_t N, int KA, const float* A, const float* G, float* dst, hipStream_t st) {
    const int64_t nb = std::min<int64_t>(kWgMaxBlocks, std::max<int64_t>(1, (N + kWgChunk - 1) / kWgChunk));
    const int64_t rpb = ((N + nb - 1) / nb + kWgChunk - 1) / kWgChunk * kWgChunk;
    const int64_t E = int64_t(KA) * kMpH;
    if (KA == kMpH)
      hipLaunchKernelGGL(k_wgrad<kMpH>, dim3(nb), dim3(256), 0, st, N, rpb, A, G, part.p);
    else
      hipLaunchKernelGGL(k_wgrad<kMpF>, dim3(nb), dim3(256), 0, st, N, rpb, A, G, part.p);
    hipLaunchKernelGGL(k_reduce_parts, dim3((E + 255) / 256), dim3(256), 0, st, nb, E, part.p, dst);
  }

  // loss + gradient into g (device); returns the loss
  double gradient(int64_t batch, const float* X, const float* tau, const float* omega, hipStream_t st) {
    if (graph.n <= 0) throw std::invalid_argument("emag_mpnn: set_graph first");
    const int64_t N = batch * graph.n;
    reserve(N);
    forward(graph, N, X, y.p, st);
    const int64_t nlb = (N + 255) / 256;
    hipLaunchKernelGGL(k_loss, dim3(nlb), dim3(256), 0, st, N, graph.n, 1.0 / double(batch), y.p, tau,
                       omega, gy.p, lossp.p);
    {  // w_out
      const int64_t nb = std::min<int64_t>(kWgMaxBlocks, N);
      const int64_t rpb = (N + nb - 1) / nb;
      hipLaunchKernelGGL(k_colsum, dim3(nb), dim3(128), 0, st, N, rpb, gy.p, hl(kMpL), part.p);
      hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(256), 0, st, nb, int64_t(kMpH), part.p,
                         g.p + kMpWout);
    }
    const int64_t E = N * kMpH;
    hipLaunchKernelGGL(k_head_bwd, dim3((E + 255) / 256), dim3(256), 0, st, N, gy.p, bf_wout(), gh.p);
    for (int l = kMpL - 1; l >= 0; --l) {
      hipLaunchKernelGGL(k_gu, dim3((E + 255) / 256), dim3(256), 0, st, E, gh.p, tl(l), gu.p);
      float* gl = g.p + kMpLayer0 + l * kMpLayerStride;
      wgrad_into(N, kMpH, hl(l), gu.p, gl, st);
      wgrad_into(N, kMpH, al(l), gu.p, gl + int64_t(kMpH) * kMpH, st);
      launch_mm<2 * kMpH, kEpiBwd>(N, gu.p, kMpH, nullptr, 0, bf_bwd(l), o1.p, o2.p, gh.p, st);
      launch_agg(N, graph, graph.cT, o2.p, o1.p, gh.p, st);
    }
    wgrad_into(N, kMpF, X, gh.p, g.p + kMpWin, st);
    std::vector<double> lp(static_cast<size_t>(nlb));
    MPCHECK(hipMemcpyAsync(lp.data(), lossp.p, sizeof(double) * nlb, hipMemcpyDeviceToHost, st));
    MPCHECK(hipStreamSynchronize(st));
    double loss = 0.0;
    for (double x : lp) loss += x;
    return loss;
  }
};

namespace emag {

void mp_graph_upload(MpGraph& g, int64_t n, const int32_t* rp, const int32_t* col, const float* coef) {
  if (n <= 0 || !rp || !col || !coef) throw std::invalid_argument("graph: empty or null input");
  const int64_t nnz = rp[n];
  if (rp[0] != 0 || nnz < 0) throw std::invalid_argument("graph: bad row_ptr");
  std::vector<float> cA(static_cast<size_t>(nnz)), cT(static_cast<size_t>(nnz));
  for (int64_t i = 0; i < n; ++i) {
    if (rp[i + 1] < rp[i]) throw std::invalid_argument("graph: row_ptr not monotone");
    const int64_t di = std::max<int64_t>(1, rp[i + 1] - rp[i]);
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
      const int64_t j = col[k];
      if (j < 0 || j >= n) throw std::invalid_argument("graph: column out of range");
      const int64_t dj = std::max<int64_t>(1, rp[j + 1] - rp[j]);
      cA[size_t(k)] = float(double(coef[k]) / double(di));
      cT[size_t(k)] = float(double(coef[k]) / double(dj));
    }
  }
  mp_graph_free(g);
  MPCHECK(hipMalloc(&g.rp, sizeof(int32_t) * (n + 1)));
  MPCHECK(hipMalloc(&g.col, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
  MPCHECK(hipMalloc(&g.cA, sizeof(float) * std::max<int64_t>(nnz, 1)));
  MPCHECK(hipMalloc(&g.cT, sizeof(float) * std::max<int64_t>(nnz, 1)));
  MPCHECK(hipMemcpy(g.rp, rp, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice));
  if (nnz) {
    MPCHECK(hipMemcpy(g.col, col, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
    MPCHECK(hipMemcpy(g.cA, cA.data(), sizeof(float) * nnz, hipMemcpyHostToDevice));
    MPCHECK(hipMemcpy(g.cT, cT.data(), sizeof(float) * nnz, hipMemcpyHostToDevice));
  }
  g.n = n;
  g.nnz = nnz;
}

void mp_graph_free(MpGraph& g) {
  for (void* p : {(void*)g.rp, (void*)g.col, (void*)g.cA, (void*)g.cT})
    if (p) (void)hipFree(p);
  g = MpGraph{};
}

void mpnn_forward(emag_mpnn* m, const MpGraph& g, int64_t batch, const float* X, float* y,
                  hipStream_t st) {
  m->forward(g, batch * g.n, X, y, st);
}

void launch_feature_scales(const double* partial, int nparts, int64_t nl, double* fsc, hipStream_t st) {
  hipLaunchKernelGGL(k_feature_scales, dim3(1), dim3(256), 0, st, partial, nparts, nl, fsc);
}

void launch_lambda_features(int64_t nl, const double* d, const double* z, const double* fsc,
                            const double* rho, const int64_t* lam_pos, const float* dir, float* X,
                            hipStream_t st) {
  if (nl <= 0) return;
  hipLaunchKernelGGL(k_lambda_features, dim3((nl + 255) / 256), dim3(256), 0, st, nl, d, z, fsc, rho,
                     lam_pos, dir, X);
}

void launch_lambda_from_y(int64_t nl, const float* y, const double* fsc, double* lam, hipStream_t st) {
  if (nl <= 0) return;
  hipLaunchKernelGGL(k_lambda_from_y, dim3((nl + 255) / 256), dim3(256), 0, st, nl, y, fsc, lam);
}

}  // namespace emag

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {
template <typename Fn>
int32_t mp_guard(emag_mpnn* m, Fn&& fn) {
  if (!m) return 4;
  try {
    m->err.clear();
    fn();
    return 0;
  } catch (const std::invalid_argument& e) {
    m->err = e.what();
    return 4;
  } catch (const std::exception& e) {
    m->err = e.what();
    return 1;
  }
}
}  // namespace

extern "C" {

emag_mpnn_handle emag_mpnn_create(uint64_t seed) {
  try {
    auto* m = new emag_mpnn();
    std::vector<float> w(size_t(kMpParams), 0.f);
    SplitMix rng{seed};
    auto fill = [&](int64_t off, int64_t fan_in, int64_t fan_out, int64_t count) {
      const double lim = std::sqrt(6.0 / double(fan_in + fan_out));
      for (int64_t i = 0; i < count; ++i) w[size_t(off + i)] = float((2.0 * rng.uniform() - 1.0) * lim);
    };
    fill(kMpWin, kMpF, kMpH, int64_t(kMpF) * kMpH);
    for (int l = 0; l < kMpL; ++l) {
      fill(kMpLayer0 + l * kMpLayerStride, 2 * kMpH, kMpH, 2 * int64_t(kMpH) * kMpH);
    }
    m->w.alloc(size_t(kMpParams));
    m->g.alloc(size_t(kMpParams));
    m->m.alloc(size_t(kMpParams));
    m->v.alloc(size_t(kMpParams));
    m->bf.alloc(size_t(kBfTotal));
    MPCHECK(hipMemcpy(m->w.p, w.data(), sizeof(float) * kMpParams, hipMemcpyHostToDevice));
    MPCHECK(hipMemset(m->m.p, 0, sizeof(float) * kMpParams));
    MPCHECK(hipMemset(m->v.p, 0, sizeof(float) * kMpParams));
    MPCHECK(hipMemset(m->g.p, 0, sizeof(float) * kMpParams));
    m->refresh(nullptr);
    MPCHECK(hipDeviceSynchronize());
    return m;
  } catch (...) {
    return nullptr;
  }
}

void emag_mpnn_destroy(emag_mpnn_handle m) { delete m; }

const char* emag_mpnn_last_error(emag_mpnn_handle m) { return m ? m->err.c_str() : "null handle"; }

int64_t emag_mpnn_num_params(void) { return kMpParams; }

int32_t emag_mpnn_get_params(emag_mpnn_handle m, float* out) {
  return mp_guard(m, [&] {
    if (!out) throw std::invalid_argument("get_params: null output");
    MPCHECK(hipMemcpy(out, m->w.p, sizeof(float) * kMpParams, hipMemcpyDeviceToHost));
  });
}

int32_t emag_mpnn_set_params(emag_mpnn_handle m, const float* in) {
  return mp_guard(m, [&] {
    if (!in) throw std::invalid_argument("set_params: null input");
    for (int64_t i = 0; i < kMpParams; ++i)
      if (!std::isfinite(in[i])) throw std::invalid_argument("set_params: non-finite weight");
    MPCHECK(hipMemcpy(m->w.p, in, sizeof(float) * kMpParams, hipMemcpyHostToDevice));
    MPCHECK(hipMemset(m->m.p, 0, sizeof(float) * kMpParams));
    MPCHECK(hipMemset(m->v.p, 0, sizeof(float) * kMpParams));
    m->step = 0;
    m->refresh(nullptr);
    MPCHECK(hipDeviceSynchronize());
  });
}

int32_t emag_mpnn_set_graph(emag_mpnn_handle m, int64_t n_nodes, const int32_t* row_ptr,
                            const int32_t* col, const float* coef) {
  return mp_guard(m, [&] { mp_graph_upload(m->graph, n_nodes, row_ptr, col, coef); });
}

int32_t emag_mpnn_forward(emag_mpnn_handle m, int64_t batch, const float* x_dev, float* y_dev,
                          void* stream) {
  return mp_guard(m, [&] {
    if (batch <= 0 || !x_dev || !y_dev) throw std::invalid_argument("forward: bad arguments");
    if (m->graph.n <= 0) throw std::invalid_argument("forward: set_graph first");
    hipStream_t st = static_cast<hipStream_t>(stream);
    m->forward(m->graph, batch * m->graph.n, x_dev, y_dev, st);
    MPCHECK(hipGetLastError());
    MPCHECK(hipStreamSynchronize(st));
  });
}

int32_t emag_mpnn_gradient(emag_mpnn_handle m, int64_t batch, const float* x_dev,
                           const float* target_dev, const float* omega_dev, float* grad_host,
                           double* loss, void* stream) {
  return mp_guard(m, [&] {
    if (batch <= 0 || !x_dev || !target_dev || !omega_dev || !grad_host)
      throw std::invalid_argument("gradient: bad arguments");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const double l = m->gradient(batch, x_dev, target_dev, omega_dev, st);
    MPCHECK(hipGetLastError());
    MPCHECK(hipMemcpyAsync(grad_host, m->g.p, sizeof(float) * kMpParams, hipMemcpyDeviceToHost, st));
    MPCHECK(hipStreamSynchronize(st));
    if (loss) *loss = l;
  });
}

int32_t emag_mpnn_train_step(emag_mpnn_handle m, int64_t batch, const float* x_dev,
                             const float* target_dev, const float* omega_dev, float lr,
                             double* loss, void* stream) {
  return mp_guard(m, [&] {
    if (batch <= 0 || !x_dev || !target_dev || !omega_dev || !(lr > 0.f))
      throw std::invalid_argument("train_step: bad arguments");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const double l = m->gradient(batch, x_dev, target_dev, omega_dev, st);
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = float(1.0 - std::pow(0.9, double(m->step)));
    const float bc2 = float(1.0 - std::pow(0.999, double(m->step)));
    hipLaunchKernelGGL(k_adam, dim3((kMpParams + 255) / 256), dim3(256), 0, st, kMpParams, m->w.p,
                       m->g.p, m->m.p, m->v.p, lr, b1, b2, 1e-8f, bc1, bc2);
    m->refresh(st);
    MPCHECK(hipGetLastError());
    MPCHECK(hipStreamSynchronize(st));
    if (loss) *loss = l;
  });
}

int32_t emag_mpnn_mfma_probe(const float* a_frag, const float* b_frag, float* d) {
  if (!a_frag || !b_frag || !d) return 4;
  float *da = nullptr, *db = nullptr, *dd = nullptr;
  int32_t rc = 0;
  try {
    MPCHECK(hipMalloc(&da, 512 * sizeof(float)));
    MPCHECK(hipMalloc(&db, 512 * sizeof(float)));
    MPCHECK(hipMalloc(&dd, 256 * sizeof(float)));
    MPCHECK(hipMemcpy(da, a_frag, 512 * sizeof(float), hipMemcpyHostToDevice));
    MPCHECK(hipMemcpy(db, b_frag, 512 * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mpnn_probe, dim3(1), dim3(64), 0, nullptr, da, db, dd);
    MPCHECK(hipDeviceSynchronize());
    MPCHECK(hipMemcpy(d, dd, 256 * sizeof(float), hipMemcpyDeviceToHost));
  } catch (...) {
    rc = 1;
  }
  for (float* p : {da, db, dd})
    if (p) (void)hipFree(p);
  return rc;
}

}  // extern "C"

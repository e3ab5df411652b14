// Message-passing network that predicts the FETI-DP initial multiplier lambda0
// from per-interface-edge features (device side).
//
// Fixed architecture: F = 16 input features, hidden H = 128, L = 3 residual
// tanh layers, no biases; MFMA operands in bfloat16, accumulation in FP32,
// FP32 master weights with Adam.  The reference has no learned initial guess
// (its Krylov solvers start from zero, src/numeric/iterative_solvers.cpp:424-503).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace emag {

constexpr int kMpF = 16, kMpH = 128, kMpL = 3;
// parameter offsets in the FP32 master vector
constexpr int64_t kMpWin = 0;                                   // W_in [F][H]
constexpr int64_t kMpLayer0 = int64_t(kMpF) * kMpH;             // Ws_l, Wn_l [H][H] each
constexpr int64_t kMpLayerStride = 2 * int64_t(kMpH) * kMpH;
constexpr int64_t kMpWout = kMpLayer0 + kMpL * kMpLayerStride;  // w_out [H]
constexpr int64_t kMpParams = kMpWout + kMpH;

// Device CSR graph of one n-node graph; a batch is B disjoint copies.
// cA[k] = c_ij / deg_i (forward aggregation), cT[k] = c_ij / deg_j (its transpose).
struct MpGraph {
  int64_t n = 0, nnz = 0;
  int32_t* rp = nullptr;
  int32_t* col = nullptr;
  float* cA = nullptr;
  float* cT = nullptr;
};
void mp_graph_upload(MpGraph& g, int64_t n, const int32_t* rp, const int32_t* col, const float* coef);
void mp_graph_free(MpGraph& g);

}  // namespace emag

struct emag_mpnn;

namespace emag {
// y[u] = network(X)[u] for u < batch * g.n, on stream st (model workspace grows as needed).
void mpnn_forward(emag_mpnn* m, const MpGraph& g, int64_t batch, const float* X, float* y,
                  hipStream_t st);
// Solver glue (all device-side, no host round trip): fsc = {1/s_d, 1/s_z, s_z, s_d}
// from the dot partials of d.d and z.z; node features of the interface
// multipliers (see kernel comment); lam = s_z * y.
void launch_feature_scales(const double* partial, int nparts, int64_t nl, double* fsc, hipStream_t st);
void launch_lambda_features(int64_t nl, const double* d, const double* z, const double* fsc,
                            const double* rho, const int64_t* lam_pos, const float* dir, float* X,
                            hipStream_t st);
void launch_lambda_from_y(int64_t nl, const float* y, const double* fsc, double* lam, hipStream_t st);
}  // namespace emag

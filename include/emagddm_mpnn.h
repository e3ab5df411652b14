/* Learned initial multiplier lambda0 for the FETI-DP solver: a message-passing
 * network on the interface graph (nodes = dual interface edges), trained and
 * evaluated on the GPU.
 *
 * There is no counterpart in the reference: its Krylov solvers start from
 * x0 = 0 (src/numeric/iterative_solvers.cpp:424-503, x.setZero()), and the
 * FETI-DP replacement keeps that default when no model is attached
 * (fetidp_solve with lambda0 == NULL).  A model only changes where PCPG starts;
 * the stopping rule and the FP64 residual check are unchanged.
 *
 * Architecture (fixed): F = 16 input features per node, hidden 128, 3 layers
 *     h0 = X W_in;  a = A h;  h <- h + tanh(h Ws_l + a Wn_l);  y = h w_out
 * with A_ij = c_ij / deg_i, no biases (the network is odd in X), MFMA operands
 * in bfloat16 (round to nearest even), FP32 accumulation, FP32 master weights.
 * Parameter vector order: W_in [16][128], then per layer Ws [128][128] and
 * Wn [128][128] (row = input index), then w_out [128]: 100,480 floats.
 *
 * Device pointers are HIP device addresses; `stream` is a hipStream_t (NULL =
 * default stream).  Every entry returns a status (0 = success, 4 = invalid
 * input, 1 = runtime failure); emag_mpnn_last_error holds the message. */
#ifndef EMAGDDM_MPNN_H
#define EMAGDDM_MPNN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct emag_mpnn* emag_mpnn_handle;

/* New model with seeded Glorot-uniform weights (w_out starts at zero so the
 * untrained network predicts lambda0 = 0). */
emag_mpnn_handle emag_mpnn_create(uint64_t seed);
void emag_mpnn_destroy(emag_mpnn_handle m);
const char* emag_mpnn_last_error(emag_mpnn_handle m);
int64_t emag_mpnn_num_params(void);

/* FP32 master weights (host arrays of emag_mpnn_num_params() floats); set also
 * refreshes the bfloat16 copies and resets the Adam state. */
int32_t emag_mpnn_get_params(emag_mpnn_handle m, float* out);
int32_t emag_mpnn_set_params(emag_mpnn_handle m, const float* in);

/* Graph of one sample (host CSR, symmetric pattern, c_ij = c_ji). */
int32_t emag_mpnn_set_graph(emag_mpnn_handle m, int64_t n_nodes, const int32_t* row_ptr,
                            const int32_t* col, const float* coef);

/* y = network(x) for `batch` stacked samples: x is (batch*n, 16), y is batch*n. */
int32_t emag_mpnn_forward(emag_mpnn_handle m, int64_t batch, const float* x_dev, float* y_dev,
                          void* stream);

/* Loss L = (1/batch) sum_u omega_b (y_u - target_u)^2 and its gradient
 * (host array of num_params floats), without updating the weights. */
int32_t emag_mpnn_gradient(emag_mpnn_handle m, int64_t batch, const float* x_dev,
                           const float* target_dev, const float* omega_dev, float* grad_host,
                           double* loss, void* stream);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the same loss. */
int32_t emag_mpnn_train_step(emag_mpnn_handle m, int64_t batch, const float* x_dev,
                             const float* target_dev, const float* omega_dev, float lr,
                             double* loss, void* stream);

/* Lane-layout probe of the 16x16x32 bfloat16 MFMA used by the network: lane l
 * passes a_frag[8l..8l+7] and b_frag[8l..8l+7] (values exact in bfloat16; host arrays
 * of 512 floats) and returns its four accumulators in d[4l..4l+3] (256 floats). */
int32_t emag_mpnn_mfma_probe(const float* a_frag, const float* b_frag, float* d);

#ifdef __cplusplus
}
#endif
#endif

/* C-ABI of the learned surrogate of the projected-kernel path: a small fully
 * connected network (SiLU hidden layers, linear output, FP32 parameters)
 * trained and evaluated on the GPU by hand-written HIP kernels
 * (sccm_amd/csrc/mlp_gpu.hip).
 *
 * No counterpart in the reference, which evaluates every mesh point exactly;
 * the network maps features of a mesh point (pair, Euler angles, gauge
 * angles) to the norm and energy kernels there, trained on samples written by
 * pnamp_mesh_samples (include/pnamp_gpu.h).  Status codes as pnamp_gpu.h:
 * 0 ok, < 0 misuse or device error. */
#ifndef PNAMP_MLP_H
#define PNAMP_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PNAMP_MLP_MAX_LAYERS = 8, PNAMP_MLP_MAX_WIDTH = 256 };

typedef struct pnamp_mlp pnamp_mlp;

/* nlayers weight layers, sizes[0] inputs .. sizes[nlayers] outputs (each <=
 * PNAMP_MLP_MAX_WIDTH).  Weights uniform in +-sqrt(6 / (fan_in + fan_out)) from
 * seed, biases zero.  Parameters are laid out W_0 [sizes[1]][sizes[0]]
 * row-major, b_0 [sizes[1]], W_1, b_1, ... */
int pnamp_mlp_create(int nlayers, const int32_t *sizes, uint32_t seed, int device, pnamp_mlp **out);
int64_t pnamp_mlp_n_params(pnamp_mlp *m);
int pnamp_mlp_get_params(pnamp_mlp *m, float *params);
int pnamp_mlp_set_params(pnamp_mlp *m, const float *params);

/* Training set: X [n][sizes[0]], Y [n][sizes[nlayers]] (row-major FP32),
 * per-output loss weights wout [sizes[nlayers]] (NULL: all 1).  Loss =
 * sum_o wout[o] (f_o(x) - y_o)^2 averaged over samples and outputs. */
int pnamp_mlp_set_data(pnamp_mlp *m, int n, const float *X, const float *Y, const float *wout);

/* Loss and its gradient (params layout) over the samples idx[0..n-1] of the
 * training set (idx NULL: the first n); grad may be NULL. */
int pnamp_mlp_loss_grad(pnamp_mlp *m, int n, const int32_t *idx, float *loss, float *grad);

/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) over shuffled minibatches of the
 * training set for `epochs` epochs, learning rate lr * decay^epoch; the mean
 * loss of each epoch is written to epoch_loss[epoch] when not NULL. */
int pnamp_mlp_train(pnamp_mlp *m, int epochs, int batch, float lr, float decay, uint32_t seed,
                    float *epoch_loss);

/* Predictions Y [n][sizes[nlayers]] for X [n][sizes[0]]; bf16 != 0 rounds
 * weights and activations to bfloat16 (FP32 accumulation). */
int pnamp_mlp_predict(pnamp_mlp *m, int n, const float *X, float *Y, int bf16);

/* Device time of the last pnamp_mlp_predict call (ms). */
float pnamp_mlp_last_ms(pnamp_mlp *m);
const char *pnamp_mlp_last_error(pnamp_mlp *m);
void pnamp_mlp_destroy(pnamp_mlp *m);

#ifdef __cplusplus
}
#endif

#endif

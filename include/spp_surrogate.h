/*
 * spp_surrogate.h -- C ABI of the parameter-sweep surrogate: a small MLP trained and evaluated on
 * the GPU with hand-written BF16 matrix-core (MFMA) kernels and FP32 accumulation.
 *
 * The reference has no learned component (its sweeps are plain per-point runs launched by shell
 * scripts, serial/mpi-main.cpp:157-160); every entry point here is new.  The network maps sweep
 * parameters (noise D, density, v0, log N) to steady-state statistics (mean polarisation, its
 * variance, the largest Lyapunov exponent); the sweep driver uses it only to decide where
 * simulation time goes -- every reported value comes from a plain simulation.
 *
 * Layout of the flat FP32 parameter vector (spp_mlp_get_params / set_params / gradients): for each
 * layer l = 0 .. n_hidden (n_hidden hidden layers, then the output layer) the weight matrix
 * W_l[k][j] (k over the layer's inputs, j over its outputs, row-major), then the bias b_l[j].
 * Hidden layers apply SiLU, x * sigmoid(x); the output layer is linear.
 *
 * Arithmetic: weights and activations are rounded to bfloat16 (round to nearest even) where they
 * enter a matrix product; products accumulate in FP32; biases, pre-activations, the loss, gradients
 * and the Adam moments are FP32.
 */
#ifndef SPP_SURROGATE_H
#define SPP_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spp_mlp spp_mlp;

typedef struct spp_mlp_config {
    int32_t n_in;       /* inputs (4: D, density, v0, log N)                                    */
    int32_t hidden;     /* width of every hidden layer (128)                                    */
    int32_t n_hidden;   /* hidden layers (3)                                                    */
    int32_t n_out;      /* outputs (3: mean polarisation, its variance, largest exponent)       */
    int32_t max_batch;  /* largest batch any call may use                                       */
    int32_t reserved;
    uint64_t seed;      /* weight initialisation (uniform +-sqrt(6 / (fan_in + fan_out)), zero bias) */
    float lr, beta1, beta2, eps;  /* Adam (1e-3, 0.9, 0.999, 1e-8)                               */
} spp_mlp_config;

void spp_mlp_config_default(spp_mlp_config* c);
int spp_mlp_create(const spp_mlp_config* c, spp_mlp** out);
void spp_mlp_destroy(spp_mlp* m);
const char* spp_mlp_last_error(const spp_mlp* m);
int64_t spp_mlp_param_count(const spp_mlp* m);
int spp_mlp_get_params(spp_mlp* m, float* params, int64_t n);
int spp_mlp_set_params(spp_mlp* m, const float* params, int64_t n);  /* also clears Adam state */

/* Inference: y[batch][n_out] from x[batch][n_in] (host arrays). */
int spp_mlp_forward(spp_mlp* m, const float* x, int32_t batch, float* y);

/* Loss L = sum_b sum_o w_o (y_bo - t_bo)^2 / (batch * n_out) and its gradient with respect to the
   flat parameter vector (grad may be null).  w: per-output weights, null = all ones.             */
int spp_mlp_loss_grad(spp_mlp* m, const float* x, const float* t, const float* w, int32_t batch,
                      float* loss, float* grad);

/* One Adam step on one batch (forward, backward and update on the device). */
int spp_mlp_train_step(spp_mlp* m, const float* x, const float* t, const float* w, int32_t batch,
                       float* loss);

/* `epochs` passes over n samples in minibatches of `batch` (the data stays on the device; sample
   order reshuffled every epoch from `shuffle_seed`).  loss: mean minibatch loss of the last epoch. */
int spp_mlp_train(spp_mlp* m, const float* x, const float* t, const float* w, int32_t n,
                  int32_t epochs, int32_t batch, uint64_t shuffle_seed, float* loss);

/* Device time of the last spp_mlp_train call (HIP events on the network's stream), milliseconds. */
int spp_mlp_last_train_ms(spp_mlp* m, double* ms);

#ifdef __cplusplus
}
#endif

#endif

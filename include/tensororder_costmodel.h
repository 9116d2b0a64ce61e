/*
 * tensororder_costmodel.h -- HIP kernels of the learned plan-cost model.
 *
 * A small FP32 multilayer perceptron, 16 -> 64 -> 64 -> 2 with ReLU, maps per-join
 * features (operand and output ranks, shared-bit count, log2 FLOPs, log2 bytes, kernel
 * variant, batch) to (log device time, log memory footprint).  It replaces the reference's
 * FLOP-count time model (performance_factor * FLOPs, planning.py:168-172; FLOPs from
 * contraction_tree.pyx:449-458) when choosing among candidate plans; the reference's
 * min-max-rank rule (planning.py:159-162) stays the fallback.  No counterpart exists in the
 * reference for the kernels themselves.
 *
 * Parameters are one float array of TO_MLP_NPARAMS entries:
 *   W1[64][16], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]   (row-major, out x in).
 * All entry points take device pointers, enqueue on `stream` and return TO_* codes.
 */
#ifndef TENSORORDER_COSTMODEL_H
#define TENSORORDER_COSTMODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TO_MLP_IN 16
#define TO_MLP_H1 64
#define TO_MLP_H2 64
#define TO_MLP_OUT 2
#define TO_MLP_NPARAMS (64 * 16 + 64 + 64 * 64 + 64 + 2 * 64 + 2)

/* h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), y = W3 h2 + b3 for n rows of x [n][16]. */
int to_mlp_forward(const float* params, const float* x, int64_t n, float* h1, float* h2, float* y,
                   void* stream);

/* Gradient of the mean squared error L = mean_{i,o} (y[i][o] - target[i][o])^2 with respect to
 * every parameter (grad, TO_MLP_NPARAMS floats) and L itself (loss, 1 float), given the
 * activations of to_mlp_forward.  dy [n][2], dh2 [n][64], dh1 [n][64] are scratch. */
int to_mlp_backward(const float* params, const float* x, const float* h1, const float* h2,
                    const float* y, const float* target, int64_t n, float* dy, float* dh2,
                    float* dh1, float* grad, float* loss, void* stream);

/* One Adam step (bias-corrected, step >= 1) on count parameters. */
int to_mlp_adam(float* params, const float* grad, float* m, float* v, int64_t count, float lr,
                float beta1, float beta2, float eps, int64_t step, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TENSORORDER_COSTMODEL_H */

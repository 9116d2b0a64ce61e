// Training and inference kernels of the plan-cost MLP (include/tensororder_costmodel.h).
//
// The network is tiny (5378 parameters) and the data are ~1e5 rows, so the kernels are
// written for clarity and determinism: one thread per activation for the forward pass and
// for activation gradients, one block per parameter for the weight gradients (a fixed-order
// reduction over the rows), elementwise Adam.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>

#include "tensororder_costmodel.h"

#define TO_OK 0
#define TO_ERR_INVALID 2
#define TO_ERR_HIP 3

namespace {

constexpr int kBlock = 256;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + TO_MLP_H1 * TO_MLP_IN;
constexpr int OFF_W2 = OFF_B1 + TO_MLP_H1;
constexpr int OFF_B2 = OFF_W2 + TO_MLP_H2 * TO_MLP_H1;
constexpr int OFF_W3 = OFF_B2 + TO_MLP_H2;
constexpr int OFF_B3 = OFF_W3 + TO_MLP_OUT * TO_MLP_H2;
static_assert(OFF_B3 + TO_MLP_OUT == TO_MLP_NPARAMS, "parameter layout");

// out[i][j] = act(b[j] + sum_k in[i][k] * W[j][k])
template <bool RELU>
__global__ __launch_bounds__(kBlock) void k_dense(const float* __restrict__ in, int64_t n, int din,
                                                  const float* __restrict__ W, const float* __restrict__ b,
                                                  int dout, float* __restrict__ out) {
    const int64_t total = n * dout;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        const int64_t i = t / dout;
        const int j = (int)(t - i * dout);
        float acc = b[j];
        for (int k = 0; k < din; ++k) acc += in[i * din + k] * W[j * din + k];
        out[t] = RELU ? fmaxf(acc, 0.0f) : acc;
    }
}

// dy[i][o] = 2 (y - target) / (n * OUT)
__global__ __launch_bounds__(kBlock) void k_out_grad(const float* __restrict__ y, const float* __restrict__ t,
                                                     int64_t n, float* __restrict__ dy) {
    const int64_t total = n * TO_MLP_OUT;
    const float scale = 2.0f / (float)total;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += (int64_t)gridDim.x * kBlock)
        dy[i] = scale * (y[i] - t[i]);
}

// g[i][j] = (sum_o gnext[i][o] * W[o][j]) * (h[i][j] > 0), W is [dnext][d]
__global__ __launch_bounds__(kBlock) void k_back_act(const float* __restrict__ gnext, int dnext,
                                                     const float* __restrict__ W, const float* __restrict__ h,
                                                     int64_t n, int d, float* __restrict__ g) {
    const int64_t total = n * d;
    for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total; t += (int64_t)gridDim.x * kBlock) {
        const int64_t i = t / d;
        const int j = (int)(t - i * d);
        float acc = 0.0f;
        for (int o = 0; o < dnext; ++o) acc += gnext[i * dnext + o] * W[o * d + j];
        g[t] = h[t] > 0.0f ? acc : 0.0f;
    }
}

__device__ __forceinline__ float block_sum_f(float v) {
    __shared__ float red[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = 0.0f;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBlock / 64; ++w) r += red[w];
    return r;
}

// Block p < dout*din: dW[j][k] = sum_i g[i][j] * in[i][k]; block dout*din + j: db[j] = sum_i g[i][j].
__global__ __launch_bounds__(kBlock) void k_weight_grad(const float* __restrict__ g, int dout,
                                                        const float* __restrict__ in, int din, int64_t n,
                                                        float* __restrict__ dW, float* __restrict__ db) {
    const int p = blockIdx.x;
    float acc = 0.0f;
    if (p < dout * din) {
        const int j = p / din, k = p - (p / din) * din;
        for (int64_t i = threadIdx.x; i < n; i += kBlock) acc += g[i * dout + j] * in[i * din + k];
        acc = block_sum_f(acc);
        if (threadIdx.x == 0) dW[p] = acc;
    } else {
        const int j = p - dout * din;
        for (int64_t i = threadIdx.x; i < n; i += kBlock) acc += g[i * dout + j];
        acc = block_sum_f(acc);
        if (threadIdx.x == 0) db[j] = acc;
    }
}

__global__ __launch_bounds__(kBlock) void k_loss(const float* __restrict__ y, const float* __restrict__ t,
                                                 int64_t n, float* __restrict__ loss) {
    const int64_t total = n * TO_MLP_OUT;
    float acc = 0.0f;
    for (int64_t i = threadIdx.x; i < total; i += kBlock) {
        const float e = y[i] - t[i];
        acc += e * e;
    }
    acc = block_sum_f(acc);
    if (threadIdx.x == 0) loss[0] = acc / (float)total;
}

__global__ __launch_bounds__(kBlock) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                                 float* __restrict__ m, float* __restrict__ v, int64_t count,
                                                 float lr, float b1, float b2, float eps, float c1, float c2) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += (int64_t)gridDim.x * kBlock) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

int grid_for(int64_t work) {
    const int64_t g = (work + kBlock - 1) / kBlock;
    return (int)(g < 1 ? 1 : (g > 65535 ? 65535 : g));
}

int status(const char*) {
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? TO_OK : TO_ERR_HIP;
}

}  // namespace

extern "C" {

int to_mlp_forward(const float* params, const float* x, int64_t n, float* h1, float* h2, float* y, void* stream) {
    if (!params || !x || !h1 || !h2 || !y || n <= 0) return TO_ERR_INVALID;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_dense<true>, dim3(grid_for(n * TO_MLP_H1)), dim3(kBlock), 0, st, x, n, TO_MLP_IN,
                       params + OFF_W1, params + OFF_B1, TO_MLP_H1, h1);
    hipLaunchKernelGGL(k_dense<true>, dim3(grid_for(n * TO_MLP_H2)), dim3(kBlock), 0, st, h1, n, TO_MLP_H1,
                       params + OFF_W2, params + OFF_B2, TO_MLP_H2, h2);
    hipLaunchKernelGGL(k_dense<false>, dim3(grid_for(n * TO_MLP_OUT)), dim3(kBlock), 0, st, h2, n, TO_MLP_H2,
                       params + OFF_W3, params + OFF_B3, TO_MLP_OUT, y);
    return status("to_mlp_forward");
}

int to_mlp_backward(const float* params, const float* x, const float* h1, const float* h2, const float* y,
                    const float* target, int64_t n, float* dy, float* dh2, float* dh1, float* grad, float* loss,
                    void* stream) {
    if (!params || !x || !h1 || !h2 || !y || !target || !dy || !dh2 || !dh1 || !grad || n <= 0)
        return TO_ERR_INVALID;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_out_grad, dim3(grid_for(n * TO_MLP_OUT)), dim3(kBlock), 0, st, y, target, n, dy);
    hipLaunchKernelGGL(k_back_act, dim3(grid_for(n * TO_MLP_H2)), dim3(kBlock), 0, st, dy, TO_MLP_OUT,
                       params + OFF_W3, h2, n, TO_MLP_H2, dh2);
    hipLaunchKernelGGL(k_back_act, dim3(grid_for(n * TO_MLP_H1)), dim3(kBlock), 0, st, dh2, TO_MLP_H2,
                       params + OFF_W2, h1, n, TO_MLP_H1, dh1);
    hipLaunchKernelGGL(k_weight_grad, dim3(TO_MLP_OUT * TO_MLP_H2 + TO_MLP_OUT), dim3(kBlock), 0, st, dy,
                       TO_MLP_OUT, h2, TO_MLP_H2, n, grad + OFF_W3, grad + OFF_B3);
    hipLaunchKernelGGL(k_weight_grad, dim3(TO_MLP_H2 * TO_MLP_H1 + TO_MLP_H2), dim3(kBlock), 0, st, dh2,
                       TO_MLP_H2, h1, TO_MLP_H1, n, grad + OFF_W2, grad + OFF_B2);
    hipLaunchKernelGGL(k_weight_grad, dim3(TO_MLP_H1 * TO_MLP_IN + TO_MLP_H1), dim3(kBlock), 0, st, dh1,
                       TO_MLP_H1, x, TO_MLP_IN, n, grad + OFF_W1, grad + OFF_B1);
    if (loss) hipLaunchKernelGGL(k_loss, dim3(1), dim3(kBlock), 0, st, y, target, n, loss);
    return status("to_mlp_backward");
}

int to_mlp_adam(float* params, const float* grad, float* m, float* v, int64_t count, float lr, float beta1,
                float beta2, float eps, int64_t step, void* stream) {
    if (!params || !grad || !m || !v || count <= 0 || step < 1) return TO_ERR_INVALID;
    const float c1 = 1.0f - powf(beta1, (float)step);
    const float c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(count)), dim3(kBlock), 0, reinterpret_cast<hipStream_t>(stream),
                       params, grad, m, v, count, lr, beta1, beta2, eps, c1, c2);
    return status("to_mlp_adam");
}

}  // extern "C"

// The MLP object shared by mlp.hip (network kernels) and surrogate.hip (collision surrogate).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "common.hpp"

namespace cfk {
// C[M x N] = op(A) op(B) (+ bias) with the GELU/bf16 epilogue of gemm_bf16_kernel (mlp.hip).
int gemm_bf16(bool ta, bool tb, bool gelu, const __bf16 *A, const __bf16 *B, float *C, __bf16 *Cb, const float *bias,
              int M, int N, int K, int64_t lda, int64_t ldb, int64_t ldc, hipStream_t st);
int fill_f32(float *x, float v, int64_t n, hipStream_t st);
int fill_f64(double *x, double v, int64_t n, hipStream_t st);
}  // namespace cfk

struct cfk_mlp_s {
    int n_in = 0, width = 0, depth = 0, n_out = 0;
    int64_t n_params = 0;
    std::vector<int64_t> w_off, b_off;  // per layer (depth + 1 layers) into the flat parameter arrays
    // flat parameter arrays, layer by layer: W_l [out][in] row-major, then b_l [out]
    cfk::DevBuf<float> w, m, v, g;     // FP32 master weights, Adam moments, gradients
    cfk::DevBuf<__bf16> wb;            // BF16 copy used by the products
    int64_t step = 0;
    // activations, sized for `cap` batch columns
    int64_t cap = 0;
    cfk::DevBuf<__bf16> X;             // [n_in][B]
    std::vector<std::unique_ptr<cfk::DevBuf<float>>> Z;     // pre-activations [width][B]
    std::vector<std::unique_ptr<cfk::DevBuf<__bf16>>> H;    // activations [width][B]
    cfk::DevBuf<float> Y;              // [n_out][B]
    cfk::DevBuf<__bf16> dY, dZ;
    cfk::DevBuf<float> dH;
    cfk::DevBuf<double> loss_acc;      // sum of squared errors of the last loss_grad call

    int create(int32_t n_in, int32_t width, int32_t depth, int32_t n_out, uint64_t seed);
    int ensure_batch(int64_t B);
    int input_f32(const float *Xf, int64_t B, hipStream_t st);
    int forward_bf16(int64_t B, bool train, hipStream_t st);
    int backward(const float *T, int64_t B, hipStream_t st);
    int adam(float lr, float b1, float b2, float eps, hipStream_t st);
    int upload_params(const float *host);
    // forward (train) + backward on the BF16 input in X; clears and accumulates the loss on device
    int loss_grad(int64_t B, const float *T, hipStream_t st);
    int read_loss(int64_t B, hipStream_t st, double *loss);  // mean squared error of the last loss_grad
};

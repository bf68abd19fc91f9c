// The MLP object shared by mlp.hip (network kernels) and surrogate.hip (collision surrogate).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "common.hpp"

namespace cfk {
// C[M][N] = A[M][K] B[N][K]^T (+ bias[n]) with both operands BF16 and K-contiguous ("NT"), FP32
// accumulation on MFMA.  lda, ldb multiples of 8, K a multiple of 32 with both operands zero beyond
// their true K.  Outputs (either may be null): C FP32 (ldc), Cb BF16 (ldcb) = bf16(gelu(v)) when
// gelu, else bf16(v).  Only m < M, n < N are written.
int gemm_nt(const __bf16 *A, const __bf16 *B, float *C, __bf16 *Cb, const float *bias, bool gelu, int M, int N, int K,
            int64_t lda, int64_t ldb, int64_t ldc, int64_t ldcb, hipStream_t st);
int fill_f32(float *x, float v, int64_t n, hipStream_t st);
int fill_f64(double *x, double v, int64_t n, hipStream_t st);
int fill_bf16(__bf16 *x, int64_t n, hipStream_t st);  // zeros
inline int pad32(int n) { return (n + 31) / 32 * 32; }
// out[c][r] = in[r][c] for r < rows, c < cols; rows <= r < rows_pad are written as zero
int transpose_f32(const float *in, int64_t ld_in, float *out, int64_t ld_out, int64_t rows, int64_t cols,
                  int64_t rows_pad, hipStream_t st);
}  // namespace cfk

// Layouts.  Parameters (the C-ABI view): one flat FP32 array, layer by layer, W_l [out][in]
// row-major then b_l [out].  Inside, activations are batch-major, [B][features] with the feature
// dimension padded to a multiple of 32 and the padding kept at zero, so every product is an NT
// product of K-contiguous operands; the BF16 weight copies are padded the same way, in both
// orientations (W_l [out][in] for the forward pass, W_l^T [in][out] for the backward pass).
struct cfk_mlp_s {
    int n_in = 0, width = 0, depth = 0, n_out = 0;
    int p_in = 0, p_w = 0, p_out = 0;   // padded feature dimensions
    int64_t n_params = 0;
    std::vector<int64_t> w_off, b_off;  // per layer (depth + 1 layers) into the flat parameter arrays
    cfk::DevBuf<float> w, m, v, g;      // FP32 master weights, Adam moments, gradients (flat layout)
    std::vector<std::unique_ptr<cfk::DevBuf<__bf16>>> Wb, Wt;  // padded BF16 copies per layer
    int64_t step = 0;
    // activations for up to `cap` rows (bp = cap rounded up to 32, the transposes' row length)
    int64_t cap = 0, bp = 0;
    cfk::DevBuf<__bf16> X;              // [cap][p_in]
    std::vector<std::unique_ptr<cfk::DevBuf<float>>> Z;     // pre-activations [cap][p_w]
    std::vector<std::unique_ptr<cfk::DevBuf<__bf16>>> H;    // activations [cap][p_w]
    cfk::DevBuf<float> Y;               // [cap][p_out]
    cfk::DevBuf<float> T;               // targets [cap][p_out]
    cfk::DevBuf<__bf16> dY;             // [cap][p_out]
    cfk::DevBuf<float> dH;              // [cap][p_w]
    cfk::DevBuf<__bf16> dZ;             // [cap][p_w]
    cfk::DevBuf<__bf16> dZt, Ht;        // transposes for the weight gradients [p_max][bp]
    cfk::DevBuf<float> colpart;         // partial column sums of the bias gradients
    cfk::DevBuf<double> loss_acc;       // sum of squared errors accumulated by backward()

    int create(int32_t n_in, int32_t width, int32_t depth, int32_t n_out, uint64_t seed);
    int ensure_batch(int64_t B);
    // X (bf16, [B][p_in]) and, for training, T ([B][p_out]) are filled by the caller
    int forward(int64_t B, bool train, hipStream_t st);
    int backward(int64_t B, hipStream_t st);  // adds the batch's squared error to loss_acc
    int adam(float lr, float b1, float b2, float eps, hipStream_t st);
    int upload_params(const float *host);
    int refresh_bf16(hipStream_t st);        // padded BF16 copies from the FP32 weights
    int read_loss(int64_t count, hipStream_t st, double *loss);  // loss_acc / count
};

// Multilayer perceptron trained and evaluated with this file's own HIP kernels (no BLAS), the
// network of the learned collision surrogate.
//
//   n_in -> width (GELU) -> ... -> width (GELU) -> n_out        (depth hidden layers)
//
// Products run on the CDNA4 matrix cores, v_mfma_f32_16x16x32_bf16: BF16 weights and activations,
// FP32 accumulation.  FP32 master weights, FP32 Adam moments.  Activations are feature-major,
// [features][batch] with the batch fastest, the layout of the batch-fastest distribution arrays,
// so no transposes are needed around the network.
//
// Training step (per layer l, H_0 = X):
//   Z_l = W_l H_{l-1} + b_l,  H_l = bf16(gelu(Z_l));  Y = W_L H_{L-1} + b_L
//   loss = mean((Y - T)^2),   dY = bf16(2 (Y - T) / numel)
//   gW_l = dZ_l H_{l-1}^T,    gb_l = sum_batch dZ_l,   dH_{l-1} = W_l^T dZ_l,
//   dZ_{l-1} = bf16(dH_{l-1} * gelu'(Z_{l-1}))
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <memory>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "mlp.hpp"

namespace cfk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GT = 256;  // threads per GEMM workgroup (4 waves)
constexpr int BM = 128, BN = 128, BK = 32;
constexpr int LDK = BK + 8;  // padded LDS row (elements)

__device__ inline float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad_f(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// C[M x N] = op(A)[M x K] op(B)[K x N] (+ bias[m]) with row-major storage:
//   TA = false: A is [M][K] (lda);  TA = true: A is stored [K][M] (lda) and A^T is used.
//   TB = false: B is [K][N] (ldb);  TB = true: B is stored [N][K] (ldb) and B^T is used.
// Epilogue: v = acc (+ bias[m] if bias); C (fp32, optional) = v; Cb (bf16, optional) = bf16(gelu(v))
// when GELU, else bf16(v).  A workgroup computes a 128 x 128 tile, each of its four waves a 64 x 64
// quarter as 4 x 4 blocks of 16 x 16 MFMA accumulators; K advances 32 per LDS stage.
//
// MFMA 16x16x32 operand layout: lane l supplies A[row l%16][8 consecutive k of group l/16] and
// B[the same 8 k][column l%16]; D[row 4*(l/16)+r][column l%16] comes back in element r.  (Which
// k values a lane's group covers does not matter for the product as long as A and B agree.)
template <bool TA, bool TB, bool GELU>
__global__ __launch_bounds__(GT) void gemm_bf16_kernel(const __bf16 *__restrict__ A, const __bf16 *__restrict__ B,
                                                       float *__restrict__ C, __bf16 *__restrict__ Cb,
                                                       const float *__restrict__ bias, int M, int N, int K, int64_t lda,
                                                       int64_t ldb, int64_t ldc) {
    __shared__ __align__(16) __bf16 As[BM][LDK];
    __shared__ __align__(16) __bf16 Bs[BN][LDK];
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int wm = (wave / 2) * 64, wn = (wave % 2) * 64;
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const __bf16 zero = (__bf16)0.0f;
    for (int k0 = 0; k0 < K; k0 += BK) {
        // stage op(A)[m0:+128, k0:+32] as As[m][k] and op(B)[k0:+32, n0:+128] as Bs[n][k]; consecutive
        // threads walk the contiguous dimension of the stored matrix
        for (int e = threadIdx.x; e < BM * BK; e += GT) {
            const int m = TA ? e % BM : e / BK;
            const int k = TA ? e / BM : e % BK;
            const int gm = m0 + m, gk = k0 + k;
            __bf16 v = zero;
            if (gm < M && gk < K) v = TA ? A[(int64_t)gk * lda + gm] : A[(int64_t)gm * lda + gk];
            As[m][k] = v;
        }
        for (int e = threadIdx.x; e < BN * BK; e += GT) {
            const int n = TB ? e / BK : e % BN;
            const int k = TB ? e % BK : e / BN;
            const int gn = n0 + n, gk = k0 + k;
            __bf16 v = zero;
            if (gn < N && gk < K) v = TB ? B[(int64_t)gn * ldb + gk] : B[(int64_t)gk * ldb + gn];
            Bs[n][k] = v;
        }
        __syncthreads();
        const int r = lane % 16, kq = (lane / 16) * 8;
        bf16x8 af[4], bfr[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) af[i] = *reinterpret_cast<const bf16x8 *>(&As[wm + 16 * i + r][kq]);
#pragma unroll
        for (int j = 0; j < 4; ++j) bfr[j] = *reinterpret_cast<const bf16x8 *>(&Bs[wn + 16 * j + r][kq]);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
            const int gm = m0 + wm + 16 * i + 4 * (lane / 16) + rr;
            if (gm >= M) continue;
            const float bm = bias ? bias[gm] : 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int gn = n0 + wn + 16 * j + lane % 16;
                if (gn < N) {
                    const float v = acc[i][j][rr] + bm;
                    if (C) C[(int64_t)gm * ldc + gn] = v;
                    if (Cb) Cb[(int64_t)gm * ldc + gn] = (__bf16)(GELU ? gelu_f(v) : v);
                }
            }
        }
}

// ---------------------------------------------------------------------------------------------
// Elementwise and reduction kernels of the training step
// ---------------------------------------------------------------------------------------------
__global__ void f32_to_bf16_kernel(const float *__restrict__ x, __bf16 *__restrict__ y, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (__bf16)x[i];
}

// dY = bf16(2 (Y - T) * scale); *loss += sum (Y - T)^2 (block partial sums, one FP64 atomic each)
__global__ void mse_grad_kernel(const float *__restrict__ Y, const float *__restrict__ T, __bf16 *__restrict__ dY,
                                double *__restrict__ loss, int64_t n, float scale) {
    __shared__ float red[256];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        const float d = Y[i] - T[i];
        l = d * d;
        dY[i] = (__bf16)(2.f * d * scale);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss) atomicAdd(loss, (double)red[0]);
}

// dZ = bf16(dH * gelu'(Z))
__global__ void gelu_back_kernel(const float *__restrict__ dH, const float *__restrict__ Z, __bf16 *__restrict__ dZ,
                                 int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dZ[i] = (__bf16)(dH[i] * gelu_grad_f(Z[i]));
}

// db[r] = sum_c dZ[r][c], one workgroup per row
__global__ void row_sum_kernel(const __bf16 *__restrict__ dZ, float *__restrict__ db, int64_t cols) {
    __shared__ float red[256];
    const int64_t r = blockIdx.x;
    float s = 0.f;
    for (int64_t c = threadIdx.x; c < cols; c += 256) s += (float)dZ[r * cols + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[r] = red[0];
}

// Adam on FP32 master weights (bias-corrected), refreshing the BF16 copy
__global__ void adam_kernel(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                            const float *__restrict__ g, __bf16 *__restrict__ wb, int64_t n, float lr, float b1, float b2,
                            float eps, float bc1, float bc2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    wb[i] = (__bf16)wi;
}

__global__ void fill_f32_kernel(float *__restrict__ x, float v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ void fill_f64_kernel(double *__restrict__ x, double v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

static inline unsigned nblk(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

int gemm_bf16(bool ta, bool tb, bool gelu, const __bf16 *A, const __bf16 *B, float *C, __bf16 *Cb, const float *bias,
              int M, int N, int K, int64_t lda, int64_t ldb, int64_t ldc, hipStream_t st) {
    if (M <= 0 || N <= 0) return CFK_OK;
    if ((M + BM - 1) / BM > 65535) return set_error(CFK_EINVAL, "GEMM too large (M=%d)", M);
    const dim3 grid((unsigned)((N + BN - 1) / BN), (unsigned)((M + BM - 1) / BM));
#define CFK_G(TA_, TB_, G_) \
    hipLaunchKernelGGL((gemm_bf16_kernel<TA_, TB_, G_>), grid, dim3(GT), 0, st, A, B, C, Cb, bias, M, N, K, lda, ldb, ldc)
    if (gelu) {
        if (ta || tb) return set_error(CFK_EUNSUP, "GELU epilogue with transposed operands");
        CFK_G(false, false, true);
    } else if (!ta && !tb) {
        CFK_G(false, false, false);
    } else if (ta && !tb) {
        CFK_G(true, false, false);
    } else if (!ta && tb) {
        CFK_G(false, true, false);
    } else {
        CFK_G(true, true, false);
    }
#undef CFK_G
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

int fill_f32(float *x, float v, int64_t n, hipStream_t st) {
    if (n <= 0) return CFK_OK;
    hipLaunchKernelGGL(fill_f32_kernel, dim3(nblk(n)), dim3(256), 0, st, x, v, n);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

int fill_f64(double *x, double v, int64_t n, hipStream_t st) {
    if (n <= 0) return CFK_OK;
    hipLaunchKernelGGL(fill_f64_kernel, dim3(nblk(n)), dim3(256), 0, st, x, v, n);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

}  // namespace cfk

// ---------------------------------------------------------------------------------------------
// The network object
// ---------------------------------------------------------------------------------------------
using namespace cfk;

namespace {
// splitmix64: seeded, platform-independent initialisation
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};
}  // namespace

int cfk_mlp_s::ensure_batch(int64_t B) {
    if (B <= cap) return CFK_OK;
    int rc;
    const size_t w = (size_t)width * B;
    if ((rc = X.alloc((size_t)n_in * B))) return rc;
    Z.clear();
    H.clear();
    for (int l = 0; l < depth; ++l) {
        Z.emplace_back(new DevBuf<float>());
        H.emplace_back(new DevBuf<__bf16>());
        if ((rc = Z.back()->alloc(w)) || (rc = H.back()->alloc(w))) return rc;
    }
    if ((rc = Y.alloc((size_t)n_out * B)) || (rc = dY.alloc((size_t)n_out * B)) || (rc = dH.alloc(w)) ||
        (rc = dZ.alloc(w)))
        return rc;
    cap = B;
    return CFK_OK;
}

// Forward pass on the BF16 input already in X (n_in x B).  train: keep Z for the backward pass.
int cfk_mlp_s::forward_bf16(int64_t B, bool train, hipStream_t st) {
    int rc;
    const __bf16 *in = X.p;
    int n_prev = n_in;
    for (int l = 0; l < depth; ++l) {
        if ((rc = gemm_bf16(false, false, true, wb.p + w_off[l], in, train ? Z[l]->p : nullptr, H[l]->p, w.p + b_off[l],
                            width, (int)B, n_prev, n_prev, B, B, st)))
            return rc;
        in = H[l]->p;
        n_prev = width;
    }
    return gemm_bf16(false, false, false, wb.p + w_off[depth], in, Y.p, nullptr, w.p + b_off[depth], n_out, (int)B,
                     n_prev, n_prev, B, B, st);
}

// Gradients of mean((Y - T)^2) for the batch whose forward pass (train = true) left Y and Z.
int cfk_mlp_s::backward(const float *T, int64_t B, hipStream_t st) {
    int rc;
    const int64_t n = (int64_t)n_out * B;
    hipLaunchKernelGGL(mse_grad_kernel, dim3(nblk(n)), dim3(256), 0, st, Y.p, T, dY.p, loss_acc.p, n, 1.0f / (float)n);
    CFK_HIP(hipGetLastError());
    const __bf16 *dcur = dY.p;  // gradient w.r.t. the pre-activation of layer l
    for (int l = depth; l >= 0; --l) {
        const int n_o = (l == depth) ? n_out : width;
        const int n_i = (l == 0) ? n_in : width;
        const __bf16 *hin = (l == 0) ? X.p : H[l - 1]->p;
        // gW_l = dZ_l H_{l-1}^T   (M = n_o, N = n_i, K = B)
        if ((rc = gemm_bf16(false, true, false, dcur, hin, g.p + w_off[l], nullptr, nullptr, n_o, n_i, (int)B, B, B, n_i,
                            st)))
            return rc;
        hipLaunchKernelGGL(row_sum_kernel, dim3((unsigned)n_o), dim3(256), 0, st, dcur, g.p + b_off[l], B);
        CFK_HIP(hipGetLastError());
        if (l == 0) break;
        // dH_{l-1} = W_l^T dZ_l   (M = n_i, N = B, K = n_o); W_l is stored [n_o][n_i]
        if ((rc = gemm_bf16(true, false, false, wb.p + w_off[l], dcur, dH.p, nullptr, nullptr, n_i, (int)B, n_o, n_i, B, B,
                            st)))
            return rc;
        const int64_t nh = (int64_t)width * B;
        hipLaunchKernelGGL(gelu_back_kernel, dim3(nblk(nh)), dim3(256), 0, st, dH.p, Z[l - 1]->p, dZ.p, nh);
        CFK_HIP(hipGetLastError());
        dcur = dZ.p;
    }
    return CFK_OK;
}

int cfk_mlp_s::adam(float lr, float b1, float b2, float eps, hipStream_t st) {
    ++step;
    const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(nblk(n_params)), dim3(256), 0, st, w.p, m.p, v.p, g.p, wb.p, n_params, lr, b1, b2,
                       eps, bc1, bc2);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

int cfk_mlp_s::input_f32(const float *Xf, int64_t B, hipStream_t st) {
    const int64_t n = (int64_t)n_in * B;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(nblk(n)), dim3(256), 0, st, Xf, X.p, n);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

int cfk_mlp_s::upload_params(const float *host) {
    CFK_HIP(hipMemcpy(w.p, host, (size_t)n_params * sizeof(float), hipMemcpyHostToDevice));
    int rc;
    if ((rc = fill_f32(m.p, 0.0f, n_params, nullptr)) || (rc = fill_f32(v.p, 0.0f, n_params, nullptr))) return rc;
    step = 0;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(nblk(n_params)), dim3(256), 0, nullptr, w.p, wb.p, n_params);
    CFK_HIP(hipGetLastError());
    CFK_HIP(hipDeviceSynchronize());
    return CFK_OK;
}

int cfk_mlp_s::loss_grad(int64_t B, const float *T, hipStream_t st) {
    int rc;
    if ((rc = fill_f64(loss_acc.p, 0.0, 1, st))) return rc;
    if ((rc = forward_bf16(B, true, st))) return rc;
    return backward(T, B, st);
}

int cfk_mlp_s::read_loss(int64_t B, hipStream_t st, double *loss) {
    double h = 0.0;
    CFK_HIP(hipMemcpyAsync(&h, loss_acc.p, sizeof(double), hipMemcpyDeviceToHost, st));
    CFK_HIP(hipStreamSynchronize(st));
    *loss = h / (double)((int64_t)n_out * B);
    return CFK_OK;
}

int cfk_mlp_s::create(int32_t n_in_, int32_t width_, int32_t depth_, int32_t n_out_, uint64_t seed) {
    n_in = n_in_;
    width = width_;
    depth = depth_;
    n_out = n_out_;
    int64_t off = 0;
    w_off.clear();
    b_off.clear();
    for (int l = 0; l <= depth; ++l) {
        const int64_t fi = (l == 0) ? n_in : width, fo = (l == depth) ? n_out : width;
        w_off.push_back(off);
        off += fi * fo;
        b_off.push_back(off);
        off += fo;
    }
    n_params = off;
    int rc;
    if ((rc = w.alloc(off)) || (rc = m.alloc(off)) || (rc = v.alloc(off)) || (rc = g.alloc(off)) || (rc = wb.alloc(off)) ||
        (rc = loss_acc.alloc(1)))
        return rc;
    // He-uniform hidden layers (GELU), Glorot-uniform output layer, zero biases
    std::vector<float> host((size_t)off, 0.0f);
    SplitMix rng{seed};
    for (int l = 0; l <= depth; ++l) {
        const int64_t fi = (l == 0) ? n_in : width, fo = (l == depth) ? n_out : width;
        const double a = (l == depth) ? std::sqrt(6.0 / (double)(fi + fo)) : std::sqrt(6.0 / (double)fi);
        for (int64_t i = 0; i < fi * fo; ++i) host[w_off[l] + i] = (float)((2.0 * rng.uniform() - 1.0) * a);
    }
    return upload_params(host.data());
}

extern "C" {

int cfk_mlp_create(int32_t n_in, int32_t width, int32_t depth, int32_t n_out, uint64_t seed, cfk_mlp *out) {
    if (!out) return set_error(CFK_EINVAL, "null output handle");
    *out = nullptr;
    if (n_in < 1 || width < 1 || n_out < 1 || depth < 1 || depth > 16 || n_in > 65535 || width > 65535 || n_out > 65535)
        return set_error(CFK_EINVAL, "mlp sizes out of range (n_in=%d width=%d depth=%d n_out=%d)", n_in, width, depth,
                         n_out);
    int rc;
    if ((rc = check_device())) return rc;
    std::unique_ptr<cfk_mlp_s> p(new cfk_mlp_s());
    if ((rc = p->create(n_in, width, depth, n_out, seed))) return rc;
    *out = p.release();
    return CFK_OK;
}

int cfk_mlp_destroy(cfk_mlp mlp) {
    delete mlp;
    return CFK_OK;
}

int cfk_mlp_num_params(cfk_mlp mlp, int64_t *n) {
    if (!mlp || !n) return set_error(CFK_EINVAL, "null argument");
    *n = mlp->n_params;
    return CFK_OK;
}

int cfk_mlp_get_params(cfk_mlp mlp, float *host) {
    if (!mlp || !host) return set_error(CFK_EINVAL, "null argument");
    CFK_HIP(hipDeviceSynchronize());
    CFK_HIP(hipMemcpy(host, mlp->w.p, (size_t)mlp->n_params * sizeof(float), hipMemcpyDeviceToHost));
    return CFK_OK;
}

int cfk_mlp_set_params(cfk_mlp mlp, const float *host) {
    if (!mlp || !host) return set_error(CFK_EINVAL, "null argument");
    CFK_HIP(hipDeviceSynchronize());
    return mlp->upload_params(host);
}

int cfk_mlp_get_grads(cfk_mlp mlp, float *host) {
    if (!mlp || !host) return set_error(CFK_EINVAL, "null argument");
    CFK_HIP(hipDeviceSynchronize());
    CFK_HIP(hipMemcpy(host, mlp->g.p, (size_t)mlp->n_params * sizeof(float), hipMemcpyDeviceToHost));
    return CFK_OK;
}

int cfk_mlp_forward(cfk_mlp mlp, float *Y_dev, const float *X_dev, int64_t B, void *stream) {
    if (!mlp || !Y_dev || !X_dev) return set_error(CFK_EINVAL, "null argument");
    if (B < 1 || B > 0x7fffffff) return set_error(CFK_EINVAL, "batch %lld out of range", (long long)B);
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = mlp->ensure_batch(B)) || (rc = mlp->input_f32(X_dev, B, st)) || (rc = mlp->forward_bf16(B, false, st)))
        return rc;
    CFK_HIP(hipMemcpyAsync(Y_dev, mlp->Y.p, (size_t)mlp->n_out * B * sizeof(float), hipMemcpyDeviceToDevice, st));
    return CFK_OK;
}

int cfk_mlp_loss_grad(cfk_mlp mlp, const float *X_dev, const float *T_dev, int64_t B, double *loss, void *stream) {
    if (!mlp || !X_dev || !T_dev) return set_error(CFK_EINVAL, "null argument");
    if (B < 1 || B > 0x7fffffff) return set_error(CFK_EINVAL, "batch %lld out of range", (long long)B);
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = mlp->ensure_batch(B)) || (rc = mlp->input_f32(X_dev, B, st)) || (rc = mlp->loss_grad(B, T_dev, st)))
        return rc;
    return loss ? mlp->read_loss(B, st, loss) : CFK_OK;
}

int cfk_mlp_adam_step(cfk_mlp mlp, float lr, float beta1, float beta2, float eps, void *stream) {
    if (!mlp) return set_error(CFK_EINVAL, "null argument");
    if (!(lr >= 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
        return set_error(CFK_EINVAL, "bad Adam hyper-parameters");
    return mlp->adam(lr, beta1, beta2, eps, (hipStream_t)stream);
}

}  // extern "C"

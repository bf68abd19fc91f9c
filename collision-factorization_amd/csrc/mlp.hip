// Multilayer perceptron trained and evaluated with this file's own HIP kernels (no BLAS), the
// network of the learned collision surrogate.
//
//   n_in -> width (GELU) -> ... -> width (GELU) -> n_out        (depth hidden layers)
//
// Every product runs on the CDNA4 matrix cores, v_mfma_f32_16x16x32_bf16: BF16 weights and
// activations, FP32 accumulation.  FP32 master weights and Adam moments.  Activations are
// batch-major, [B][features] (feature dimension padded to 32 with zeros), so each product is an NT
// product whose operands are both contiguous along the reduction: tiles move global -> registers ->
// LDS as 16-byte vectors, double-buffered.
//
// Training step (per layer l, H_0 = X, rows = batch):
//   Z_l = H_{l-1} W_l^T + b_l,  H_l = bf16(gelu(Z_l));  Y = H_{L-1} W_L^T + b_L
//   loss = mean((Y - T)^2),     dY = bf16(2 (Y - T) / numel)
//   gW_l = dZ_l^T H_{l-1}  (as the NT product of the transposes),  gb_l = column sums of dZ_l,
//   dH_{l-1} = dZ_l W_l,  dZ_{l-1} = bf16(dH_{l-1} * gelu'(Z_{l-1}))
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <vector>

#include "common.hpp"
#include "mlp.hpp"

namespace cfk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GT = 256;  // threads per GEMM workgroup (4 waves, 2 x 2 wave tiles)
constexpr int BK = 32;   // reduction depth per LDS stage (one MFMA k-step)
constexpr int LDK = BK + 8;  // padded LDS row (elements): 80 bytes keeps 16-byte rows spread over banks

// erf by Abramowitz & Stegun 7.1.26 (|error| <= 1.5e-7, far below the BF16 rounding of the activations
// that follows): erf(x) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-x^2), t = 1 / (1 + p x), x >= 0.
// A dozen instructions instead of the library erff's range reduction: the epilogue of the network's
// products evaluates it for every activation.
__device__ inline float erf_as(float x) {
    const float ax = fabsf(x);
    const float t = __frcp_rn(fmaf(0.3275911f, ax, 1.0f));
    float p = fmaf(t, 1.061405429f, -1.453152027f);
    p = fmaf(t, p, 1.421413741f);
    p = fmaf(t, p, -0.284496736f);
    p = fmaf(t, p, 0.254829592f);
    const float y = 1.0f - t * p * __expf(-ax * ax);
    return copysignf(y, x);
}
__device__ inline float gelu_f(float z) { return 0.5f * z * (1.0f + erf_as(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad_f(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// C[M][N] = A[M][K] B[N][K]^T (+ bias[n]).  A BM x BN workgroup tile, four waves of (BM/2) x (BN/2),
// each a grid of 16 x 16 MFMA accumulators.  MFMA 16x16x32 operands: lane l supplies A[row l%16][8
// consecutive k of group l/16] and B[column l%16][the same 8 k]; D[row 4*(l/16)+r][column l%16]
// comes back in element r.
template <int BM, int BN, bool GELU>
__global__ __launch_bounds__(GT) void gemm_nt_kernel(const __bf16 *__restrict__ A, const __bf16 *__restrict__ B,
                                                     float *__restrict__ C, __bf16 *__restrict__ Cb,
                                                     const float *__restrict__ bias, int M, int N, int K, int64_t lda,
                                                     int64_t ldb, int64_t ldc, int64_t ldcb) {
    constexpr int FM = BM / 32, FN = BN / 32;          // MFMA blocks per wave and dimension
    constexpr int AV = BM * BK / 8 / GT, BV = BN * BK / 8 / GT;  // 16-byte vectors per thread and stage
    static_assert(AV >= 1 && BV >= 1, "tile too small for the loader");
    __shared__ __align__(16) __bf16 As[2][BM][LDK];
    __shared__ __align__(16) __bf16 Bs[2][BN][LDK];
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave / 2) * (BM / 2), wn = (wave % 2) * (BN / 2);
    uint4 ra[AV], rb[BV];
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < AV; ++i) {
            const int v = tid + i * GT, row = v >> 2, kv = (v & 3) * 8;
            const int gm = m0 + row;
            ra[i] = gm < M ? *reinterpret_cast<const uint4 *>(A + (int64_t)gm * lda + k0 + kv) : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < BV; ++i) {
            const int v = tid + i * GT, row = v >> 2, kv = (v & 3) * 8;
            const int gn = n0 + row;
            rb[i] = gn < N ? *reinterpret_cast<const uint4 *>(B + (int64_t)gn * ldb + k0 + kv) : make_uint4(0, 0, 0, 0);
        }
    };
    auto swrite = [&](int buf) {
#pragma unroll
        for (int i = 0; i < AV; ++i) {
            const int v = tid + i * GT;
            *reinterpret_cast<uint4 *>(&As[buf][v >> 2][(v & 3) * 8]) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < BV; ++i) {
            const int v = tid + i * GT;
            *reinterpret_cast<uint4 *>(&Bs[buf][v >> 2][(v & 3) * 8]) = rb[i];
        }
    };
    f32x4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nk = K / BK;
    gload(0);
    swrite(0);
    __syncthreads();
    const int r = lane % 16, kq = (lane / 16) * 8;
    for (int s = 0; s < nk; ++s) {
        const int buf = s & 1;
        if (s + 1 < nk) gload((s + 1) * BK);
        bf16x8 af[FM], bfr[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = *reinterpret_cast<const bf16x8 *>(&As[buf][wm + 16 * i + r][kq]);
#pragma unroll
        for (int j = 0; j < FN; ++j) bfr[j] = *reinterpret_cast<const bf16x8 *>(&Bs[buf][wn + 16 * j + r][kq]);
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        // the other buffer was last read in iteration s-1, before the barrier that ended it
        if (s + 1 < nk) swrite(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
        const int gn = n0 + wn + 16 * j + lane % 16;
        if (gn >= N) continue;
        const float bn = bias ? bias[gn] : 0.0f;
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int gm = m0 + wm + 16 * i + 4 * (lane / 16) + rr;
                if (gm < M) {
                    const float v = acc[i][j][rr] + bn;
                    if (C) C[(int64_t)gm * ldc + gn] = v;
                    if (Cb) Cb[(int64_t)gm * ldcb + gn] = (__bf16)(GELU ? gelu_f(v) : v);
                }
            }
    }
}

// Large products (the surrogate's inference over thousands of cells): C[M][N] = A[M][K] B[N][K]^T (+ bias)
// with a 128 x 128 workgroup tile, four waves of 64 x 64, and a ring of PS stages of BK = 32 in LDS that
// the loads fill directly (global_load_lds, 16 bytes per lane, no registers and no LDS stores on the way),
// PS - 1 stages ahead of the one being multiplied, so the loads' latency hides behind PS - 1 stages of
// MFMAs.  A stage holds 128 rows x 64 bytes of A and of B; a wave's load instruction fills 1 KB of it,
// lane L landing at 16-byte slot L of its kilobyte.  The slot of (row r, 8-element chunk c) is
// 4 r + (c ^ ((r >> 2) & 3)): the 16 rows one MFMA operand read touches then start in 16 different bank
// groups (no LDS bank conflicts) although rows are not padded.
constexpr int PT = 128;  // workgroup tile (both dimensions)
constexpr int PS = 4;    // LDS stages in the ring
constexpr int PSTAGE = PT * BK;  // bf16 elements of one operand per stage (8 KB)

__device__ __forceinline__ int pipe_slot(int r, int c) { return 4 * r + (c ^ ((r >> 2) & 3)); }

// The ring's stages are separate LDS arrays and the main loop is unrolled over them, so every load and
// every operand read names its stage at compile time.
__shared__ __align__(1024) __bf16 pipe_a0[PSTAGE], pipe_a1[PSTAGE], pipe_a2[PSTAGE], pipe_a3[PSTAGE];
__shared__ __align__(1024) __bf16 pipe_b0[PSTAGE], pipe_b1[PSTAGE], pipe_b2[PSTAGE], pipe_b3[PSTAGE];

template <int SLOT>
__device__ __forceinline__ __bf16 *pipe_a() {
    return SLOT == 0 ? pipe_a0 : SLOT == 1 ? pipe_a1 : SLOT == 2 ? pipe_a2 : pipe_a3;
}
template <int SLOT>
__device__ __forceinline__ __bf16 *pipe_b() {
    return SLOT == 0 ? pipe_b0 : SLOT == 1 ? pipe_b1 : SLOT == 2 ? pipe_b2 : pipe_b3;
}

template <bool GELU>
__global__ __launch_bounds__(GT) void gemm_nt_pipe_kernel(const __bf16 *__restrict__ A, const __bf16 *__restrict__ B,
                                                          float *__restrict__ C, __bf16 *__restrict__ Cb,
                                                          const float *__restrict__ bias, int M, int N, int K,
                                                          int64_t lda, int64_t ldb, int64_t ldc, int64_t ldcb) {
    static_assert(PS == 4, "the unrolled loop below is written for four stages");
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    // Tile of this workgroup.  Workgroups go to the 8 XCDs round-robin in launch order; when the row tiles
    // divide among them, XCD x takes the row tiles t = x (mod 8), each with all column tiles in a row, so
    // an XCD's L2 holds its A rows and all of B instead of every XCD streaming all of A.
    const int NT = (N + PT - 1) / PT, MT = (M + PT - 1) / PT;
    int mt, nt;
    if (MT % 8 == 0) {
        const int idx = blockIdx.x / 8;
        nt = idx % NT;
        mt = (idx / NT) * 8 + blockIdx.x % 8;
    } else {
        nt = blockIdx.x % NT;
        mt = blockIdx.x / NT;
    }
    const int m0 = mt * PT, n0 = nt * PT;
    const int wm = (wave / 2) * 64, wn = (wave % 2) * 64;
    // Loader: this wave fills kilobytes 2*wave and 2*wave+1 of each operand's stage (rows 32*wave .. +32);
    // lane L of kilobyte j holds slot 64 j + L = (row, chunk) with row = 16 j + L / 4 and chunk from the swizzle.
    const __bf16 *ga[2], *gb[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int kb = 2 * wave + j, r = 16 * kb + lane / 4, c = (lane % 4) ^ ((r >> 2) & 3);
        const int gm = min(m0 + r, M - 1), gn = min(n0 + r, N - 1);  // rows past the end load a valid row
        ga[j] = A + (int64_t)gm * lda + 8 * c;                       // (their products are never stored)
        gb[j] = B + (int64_t)gn * ldb + 8 * c;
    }
    // One 16-byte-per-lane load straight into LDS: M0 holds the wave's LDS destination, lane L lands at
    // M0 + 16 L.  Written as assembly (M0 bound as an input operand) because the compiler, unable to tell
    // the ring's stages apart, would otherwise wait for every load in flight before each operand read; the
    // waits are explicit below instead.  The instruction has no destination register: its address
    // registers are read at issue.
    auto dma = [&](const __bf16 *src, __bf16 *dst) {
        const uint32_t lds = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) void *)dst;
        asm volatile("global_load_lds_dwordx4 %0, off" ::"v"(src), "{m0}"(__builtin_amdgcn_readfirstlane(lds)) : "memory");
    };
    auto issue = [&](__bf16 *as, __bf16 *bs, int k0) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            dma(ga[j] + k0, as + (2 * wave + j) * 512);
            dma(gb[j] + k0, bs + (2 * wave + j) * 512);
        }
    };
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // operand reads: lane l takes row (16-row block base + l % 16), chunk l / 16; the swizzle term depends
    // on l % 16 only, because block bases are multiples of 16
    const int r16 = lane % 16, cq = lane / 16;
    const int rd = 4 * r16 + (cq ^ ((r16 >> 2) & 3));  // slot within a 16-row block
    // Operand fragments of two stages in registers: the reads of stage s + 1 are issued before the MFMAs
    // of stage s, which do not wait for them, so the LDS latency hides behind the matrix work.
    bf16x8 fa[2][4], fb[2][4];
    auto frags = [&](const __bf16 *as, const __bf16 *bs, bf16x8 (&a)[4], bf16x8 (&b)[4]) {
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const bf16x8 *>(as + ((wm + 16 * i) * 4 + rd) * 8);
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const bf16x8 *>(bs + ((wn + 16 * j) * 4 + rd) * 8);
    };
    auto mult = [&](const bf16x8 (&a)[4], const bf16x8 (&b)[4]) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    };
    // Stage t has landed once at most the loads of the younger stages in flight (at most `ymax` of them)
    // are outstanding (4 loads per stage and thread; loads retire in order); the barrier then makes every
    // wave's part visible.
    const int nk = K / BK;
    auto wait_stage = [&](int t, int ymax) {
        const int younger = min(nk - 1 - t, ymax);
        if (younger >= 2)
            __builtin_amdgcn_s_waitcnt(0x0F70 | 8);
        else if (younger == 1)
            __builtin_amdgcn_s_waitcnt(0x0F70 | 4);
        else
            __builtin_amdgcn_s_waitcnt(0x0F70);
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");  // no operand read moves above the wait and the barrier
    };
    if (nk <= 0) return;
    issue(pipe_a<0>(), pipe_b<0>(), 0);
    if (1 < nk) issue(pipe_a<1>(), pipe_b<1>(), BK);
    if (2 < nk) issue(pipe_a<2>(), pipe_b<2>(), 2 * BK);
    wait_stage(0, 2);
    frags(pipe_a<0>(), pipe_b<0>(), fa[0], fb[0]);
    // Step s: stage s + 1 is waited for, stage s + 3 is requested into the slot of stage s - 1 (read in step
    // s - 2, before the barrier of step s - 1), the fragments of s + 1 are read, and stage s is multiplied.
    // The last step waits for nothing in flight and reads a stale slot into fragments it never uses: no
    // branch around the waits and reads, so the compiler can track which reads each MFMA depends on.
#define CFK_PIPE_STEP(U, NEXT, REFILL)                                              \
    {                                                                               \
        const int s = s0 + U;                                                       \
        if (s < nk) {                                                               \
            wait_stage(s + 1, 1);                                                   \
            if (s + 3 < nk) issue(pipe_a<REFILL>(), pipe_b<REFILL>(), (s + 3) * BK); \
            frags(pipe_a<NEXT>(), pipe_b<NEXT>(), fa[(U + 1) % 2], fb[(U + 1) % 2]); \
            mult(fa[U % 2], fb[U % 2]);                                             \
        }                                                                           \
    }
    for (int s0 = 0; s0 < nk; s0 += 4) {
        CFK_PIPE_STEP(0, 1, 3)
        CFK_PIPE_STEP(1, 2, 0)
        CFK_PIPE_STEP(2, 3, 1)
        CFK_PIPE_STEP(3, 0, 2)
    }
#undef CFK_PIPE_STEP
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int gn = n0 + wn + 16 * j + lane % 16;
        if (gn >= N) continue;
        const float bn = bias ? bias[gn] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int gm = m0 + wm + 16 * i + 4 * (lane / 16) + rr;
                if (gm < M) {
                    const float v = acc[i][j][rr] + bn;
                    if (C) C[(int64_t)gm * ldc + gn] = v;
                    if (Cb) Cb[(int64_t)gm * ldcb + gn] = (__bf16)(GELU ? gelu_f(v) : v);
                }
            }
    }
}

// ---------------------------------------------------------------------------------------------
// Elementwise, transpose and reduction kernels
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ void fill_kernel(T *__restrict__ x, T v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// out[c][r] = in[r][c] for r < rows, c < cols; rows <= r < rows_pad are written as zero.
// 32 x 32 tiles through LDS (coalesced on both sides).
template <typename TI, typename TO>
__global__ void transpose_kernel(const TI *__restrict__ in, int64_t ld_in, TO *__restrict__ out, int64_t ld_out,
                                 int64_t rows, int64_t cols, int64_t rows_pad) {
    __shared__ float t[32][33];
    const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 256 threads: 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int64_t r = r0 + k, c = c0 + tx;
        t[k][tx] = (r < rows && c < cols) ? (float)in[r * ld_in + c] : 0.0f;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int64_t c = c0 + k, r = r0 + tx;
        if (c < cols && r < rows_pad) out[c * ld_out + r] = (TO)t[tx][k];
    }
}

// dY = bf16(2 (Y - T) * scale); *loss += sum (Y - T)^2 (grid-stride, one FP64 atomic per workgroup)
__global__ void mse_grad_kernel(const float *__restrict__ Y, const float *__restrict__ T, __bf16 *__restrict__ dY,
                                double *__restrict__ loss, int64_t n, float scale) {
    __shared__ float red[256];
    float l = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float d = Y[i] - T[i];
        l += d * d;
        dY[i] = (__bf16)(2.f * d * scale);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, (double)red[0]);
}

// dZ = bf16(dH * gelu'(Z))
__global__ void gelu_back_kernel(const float *__restrict__ dH, const float *__restrict__ Z, __bf16 *__restrict__ dZ,
                                 int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dZ[i] = (__bf16)(dH[i] * gelu_grad_f(Z[i]));
}

// Column sums gb[c] = sum_r dZ[r][c] (rows x ld layout) in two deterministic passes: workgroup
// (column block of 64, row chunk) sums its chunk with 4 threads per column into part[chunk][c],
// then a thread per column adds the chunks in order.
constexpr int CS_CHUNKS = 32;
__global__ void col_sum_part_kernel(const __bf16 *__restrict__ dZ, int64_t ld, float *__restrict__ part, int64_t rows,
                                    int cols) {
    __shared__ float red[4][64];
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    const int c = blockIdx.x * 64 + tx;
    const int64_t per = (rows + CS_CHUNKS - 1) / CS_CHUNKS;
    const int64_t r0 = (int64_t)blockIdx.y * per, r1 = std::min<int64_t>(rows, r0 + per);
    float s = 0.f;
    if (c < cols)
        for (int64_t r = r0 + ty; r < r1; r += 4) s += (float)dZ[r * ld + c];
    red[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && c < cols) part[(int64_t)blockIdx.y * cols + c] = red[0][tx] + red[1][tx] + red[2][tx] + red[3][tx];
}

__global__ void col_sum_final_kernel(const float *__restrict__ part, float *__restrict__ gb, int cols) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int k = 0; k < CS_CHUNKS; ++k) s += part[(int64_t)k * cols + c];
    gb[c] = s;
}

// Adam on FP32 master weights (bias-corrected)
__global__ void adam_kernel(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                            const float *__restrict__ g, int64_t n, float lr, float b1, float b2, float eps, float bc1,
                            float bc2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// padded BF16 copies of one layer: Wb[r][c] = Wt[c][r] = bf16(w[r*in + c]) for r < out, c < in
__global__ void refresh_kernel(const float *__restrict__ w, __bf16 *__restrict__ Wb, __bf16 *__restrict__ Wt, int out,
                               int in, int p_out, int p_in) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)out * in) return;
    const int r = (int)(i / in), c = (int)(i % in);
    const __bf16 b = (__bf16)w[i];
    Wb[(int64_t)r * p_in + c] = b;
    Wt[(int64_t)c * p_out + r] = b;
}

static inline unsigned nblk(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

int gemm_nt(const __bf16 *A, const __bf16 *B, float *C, __bf16 *Cb, const float *bias, bool gelu, int M, int N, int K,
            int64_t lda, int64_t ldb, int64_t ldc, int64_t ldcb, hipStream_t st) {
    if (M <= 0 || N <= 0) return CFK_OK;
    if (K % BK || lda % 8 || ldb % 8) return set_error(CFK_EINVAL, "gemm_nt: K=%d lda=%lld ldb=%lld not aligned", K,
                                                       (long long)lda, (long long)ldb);
    // the pipelined 128 x 128 kernel when its tiles fill the GPU twice over, 64 x 64 tiles otherwise
    const int64_t big = (int64_t)((M + PT - 1) / PT) * ((N + PT - 1) / PT);
    if (big >= 512) {
        const dim3 grid((unsigned)big);
        if (gelu)
            hipLaunchKernelGGL((gemm_nt_pipe_kernel<true>), grid, dim3(GT), 0, st, A, B, C, Cb, bias, M, N, K, lda, ldb,
                               ldc, ldcb);
        else
            hipLaunchKernelGGL((gemm_nt_pipe_kernel<false>), grid, dim3(GT), 0, st, A, B, C, Cb, bias, M, N, K, lda, ldb,
                               ldc, ldcb);
        CFK_HIP(hipGetLastError());
        return CFK_OK;
    }
    const int TM = 64, TN = 64;
    if ((M + TM - 1) / TM > 65535) return set_error(CFK_EINVAL, "gemm_nt: M=%d too large", M);
    const dim3 grid((unsigned)((N + TN - 1) / TN), (unsigned)((M + TM - 1) / TM));
#define CFK_G(BM_, BN_, G_) \
    hipLaunchKernelGGL((gemm_nt_kernel<BM_, BN_, G_>), grid, dim3(GT), 0, st, A, B, C, Cb, bias, M, N, K, lda, ldb, ldc, ldcb)
    if (gelu) CFK_G(64, 64, true); else CFK_G(64, 64, false);
#undef CFK_G
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

template <typename TI, typename TO>
int transpose(const TI *in, int64_t ld_in, TO *out, int64_t ld_out, int64_t rows, int64_t cols, int64_t rows_pad,
              hipStream_t st) {
    if (cols <= 0 || rows_pad <= 0) return CFK_OK;
    const dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((std::max(rows, rows_pad) + 31) / 32));
    hipLaunchKernelGGL((transpose_kernel<TI, TO>), grid, dim3(256), 0, st, in, ld_in, out, ld_out, rows, cols, rows_pad);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

int transpose_f32(const float *in, int64_t ld_in, float *out, int64_t ld_out, int64_t rows, int64_t cols,
                  int64_t rows_pad, hipStream_t st) {
    return transpose(in, ld_in, out, ld_out, rows, cols, rows_pad, st);
}

int fill_f32(float *x, float v, int64_t n, hipStream_t st) {
    if (n <= 0) return CFK_OK;
    hipLaunchKernelGGL(fill_kernel<float>, dim3(nblk(n)), dim3(256), 0, st, x, v, n);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

int fill_f64(double *x, double v, int64_t n, hipStream_t st) {
    if (n <= 0) return CFK_OK;
    hipLaunchKernelGGL(fill_kernel<double>, dim3(nblk(n)), dim3(256), 0, st, x, v, n);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

int fill_bf16(__bf16 *x, int64_t n, hipStream_t st) {
    if (n <= 0) return CFK_OK;
    hipLaunchKernelGGL(fill_kernel<__bf16>, dim3(nblk(n)), dim3(256), 0, st, x, (__bf16)0.0f, n);
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

template <typename T>
int zalloc(DevBuf<T> &b, size_t n) {  // allocation with zeroed contents (pads must read as zero)
    int rc = b.alloc(n);
    if (rc) return rc;
    if constexpr (std::is_same<T, float>::value) return fill_f32(b.p, 0.0f, (int64_t)n, nullptr);
    else if constexpr (std::is_same<T, double>::value) return fill_f64(b.p, 0.0, (int64_t)n, nullptr);
    else return fill_bf16(b.p, (int64_t)n, nullptr);
}
}  // namespace

int cfk_mlp_s::ensure_batch(int64_t B) {
    if (B <= cap) return CFK_OK;
    int rc;
    const int64_t nb = std::max<int64_t>(B, 32);
    const int64_t nbp = (nb + 31) / 32 * 32;
    const int pmax = std::max(p_in, std::max(p_w, p_out));
    CFK_HIP(hipDeviceSynchronize());  // buffers may still be in use by queued work
    if ((rc = zalloc(X, (size_t)nb * p_in))) return rc;
    Z.clear();
    H.clear();
    for (int l = 0; l < depth; ++l) {
        Z.emplace_back(new DevBuf<float>());
        H.emplace_back(new DevBuf<__bf16>());
        if ((rc = zalloc(*Z.back(), (size_t)nb * p_w)) || (rc = zalloc(*H.back(), (size_t)nb * p_w))) return rc;
    }
    if ((rc = zalloc(Y, (size_t)nb * p_out)) || (rc = zalloc(T, (size_t)nb * p_out)) ||
        (rc = zalloc(dY, (size_t)nb * p_out)) || (rc = zalloc(dH, (size_t)nb * p_w)) || (rc = zalloc(dZ, (size_t)nb * p_w)) ||
        (rc = zalloc(dZt, (size_t)pmax * nbp)) || (rc = zalloc(Ht, (size_t)pmax * nbp)) ||
        (rc = zalloc(colpart, (size_t)CS_CHUNKS * pmax)))
        return rc;
    CFK_HIP(hipDeviceSynchronize());
    cap = nb;
    bp = nbp;
    return CFK_OK;
}

int cfk_mlp_s::forward(int64_t B, bool train, hipStream_t st) {
    int rc;
    const __bf16 *in = X.p;
    int kin = p_in;
    for (int l = 0; l < depth; ++l) {
        if ((rc = gemm_nt(in, Wb[l]->p, train ? Z[l]->p : nullptr, H[l]->p, w.p + b_off[l], true, (int)B, width, kin, kin,
                          kin, p_w, p_w, st)))
            return rc;
        in = H[l]->p;
        kin = p_w;
    }
    return gemm_nt(in, Wb[depth]->p, Y.p, nullptr, w.p + b_off[depth], false, (int)B, n_out, kin, kin, kin, p_out, 0, st);
}

int cfk_mlp_s::backward(int64_t B, hipStream_t st) {
    int rc;
    const int64_t n = B * p_out;
    hipLaunchKernelGGL(mse_grad_kernel, dim3(std::min<unsigned>(nblk(n), 1024u)), dim3(256), 0, st, Y.p, T.p, dY.p,
                       loss_acc.p, n, 1.0f / (float)(B * (int64_t)n_out));
    CFK_HIP(hipGetLastError());
    const int64_t kb = (B + 31) / 32 * 32;  // reduction length of the weight gradients (zero-padded batch)
    const __bf16 *dcur = dY.p;
    int p_o = p_out, n_o = n_out;
    for (int l = depth; l >= 0; --l) {
        const int n_i = (l == 0) ? n_in : width, p_i = (l == 0) ? p_in : p_w;
        const __bf16 *hin = (l == 0) ? X.p : H[l - 1]->p;
        // gW_l[o][i] = sum_b dZ[b][o] H[b][i]: NT product of the transposes, written in the flat layout
        if ((rc = transpose(dcur, p_o, dZt.p, bp, B, n_o, kb, st)) || (rc = transpose(hin, p_i, Ht.p, bp, B, n_i, kb, st)))
            return rc;
        if ((rc = gemm_nt(dZt.p, Ht.p, g.p + w_off[l], nullptr, nullptr, false, n_o, n_i, (int)kb, bp, bp, n_i, 0, st)))
            return rc;
        hipLaunchKernelGGL(col_sum_part_kernel, dim3((unsigned)((n_o + 63) / 64), CS_CHUNKS), dim3(256), 0, st, dcur,
                           (int64_t)p_o, colpart.p, B, n_o);
        hipLaunchKernelGGL(col_sum_final_kernel, dim3(nblk(n_o)), dim3(256), 0, st, colpart.p, g.p + b_off[l], n_o);
        CFK_HIP(hipGetLastError());
        if (l == 0) break;
        // dH[b][i] = sum_o dZ[b][o] W[o][i] = NT product with W^T [i][o]
        if ((rc = gemm_nt(dcur, Wt[l]->p, dH.p, nullptr, nullptr, false, (int)B, n_i, p_o, p_o, p_o, p_w, 0, st)))
            return rc;
        const int64_t nh = B * p_w;
        hipLaunchKernelGGL(gelu_back_kernel, dim3(nblk(nh)), dim3(256), 0, st, dH.p, Z[l - 1]->p, dZ.p, nh);
        CFK_HIP(hipGetLastError());
        dcur = dZ.p;
        p_o = p_w;
        n_o = width;
    }
    return CFK_OK;
}

int cfk_mlp_s::refresh_bf16(hipStream_t st) {
    for (int l = 0; l <= depth; ++l) {
        const int fi = (l == 0) ? n_in : width, fo = (l == depth) ? n_out : width;
        const int pi = (l == 0) ? p_in : p_w, po = (l == depth) ? p_out : p_w;
        hipLaunchKernelGGL(refresh_kernel, dim3(nblk((int64_t)fi * fo)), dim3(256), 0, st, w.p + w_off[l], Wb[l]->p,
                           Wt[l]->p, fo, fi, po, pi);
        CFK_HIP(hipGetLastError());
    }
    return CFK_OK;
}

int cfk_mlp_s::adam(float lr, float b1, float b2, float eps, hipStream_t st) {
    ++step;
    const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(nblk(n_params)), dim3(256), 0, st, w.p, m.p, v.p, g.p, n_params, lr, b1, b2, eps,
                       bc1, bc2);
    CFK_HIP(hipGetLastError());
    return refresh_bf16(st);
}

int cfk_mlp_s::upload_params(const float *host) {
    CFK_HIP(hipMemcpy(w.p, host, (size_t)n_params * sizeof(float), hipMemcpyHostToDevice));
    int rc;
    if ((rc = fill_f32(m.p, 0.0f, n_params, nullptr)) || (rc = fill_f32(v.p, 0.0f, n_params, nullptr)) ||
        (rc = refresh_bf16(nullptr)))
        return rc;
    step = 0;
    CFK_HIP(hipDeviceSynchronize());
    return CFK_OK;
}

int cfk_mlp_s::read_loss(int64_t count, hipStream_t st, double *loss) {
    double h = 0.0;
    CFK_HIP(hipMemcpyAsync(&h, loss_acc.p, sizeof(double), hipMemcpyDeviceToHost, st));
    CFK_HIP(hipStreamSynchronize(st));
    *loss = h / (double)count;
    return CFK_OK;
}

int cfk_mlp_s::create(int32_t n_in_, int32_t width_, int32_t depth_, int32_t n_out_, uint64_t seed) {
    n_in = n_in_;
    width = width_;
    depth = depth_;
    n_out = n_out_;
    p_in = pad32(n_in);
    p_w = pad32(width);
    p_out = pad32(n_out);
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
    if ((rc = w.alloc(off)) || (rc = m.alloc(off)) || (rc = v.alloc(off)) || (rc = g.alloc(off)) ||
        (rc = loss_acc.alloc(1)))
        return rc;
    Wb.clear();
    Wt.clear();
    for (int l = 0; l <= depth; ++l) {
        const int pi = (l == 0) ? p_in : p_w, po = (l == depth) ? p_out : p_w;
        Wb.emplace_back(new DevBuf<__bf16>());
        Wt.emplace_back(new DevBuf<__bf16>());
        if ((rc = zalloc(*Wb.back(), (size_t)po * pi)) || (rc = zalloc(*Wt.back(), (size_t)pi * po))) return rc;
    }
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
    if ((rc = mlp->ensure_batch(B)) || (rc = transpose(X_dev, B, mlp->X.p, (int64_t)mlp->p_in, (int64_t)mlp->n_in, B, (int64_t)mlp->n_in, st)))
        return rc;
    if ((rc = mlp->forward(B, false, st))) return rc;
    return transpose(mlp->Y.p, (int64_t)mlp->p_out, Y_dev, B, B, (int64_t)mlp->n_out, B, st);
}

int cfk_mlp_loss_grad(cfk_mlp mlp, const float *X_dev, const float *T_dev, int64_t B, double *loss, void *stream) {
    if (!mlp || !X_dev || !T_dev) return set_error(CFK_EINVAL, "null argument");
    if (B < 1 || B > 0x7fffffff) return set_error(CFK_EINVAL, "batch %lld out of range", (long long)B);
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = mlp->ensure_batch(B))) return rc;
    if ((rc = transpose(X_dev, B, mlp->X.p, (int64_t)mlp->p_in, (int64_t)mlp->n_in, B, (int64_t)mlp->n_in, st)) ||
        (rc = transpose(T_dev, B, mlp->T.p, (int64_t)mlp->p_out, (int64_t)mlp->n_out, B, (int64_t)mlp->n_out, st)) ||
        (rc = fill_f64(mlp->loss_acc.p, 0.0, 1, st)) || (rc = mlp->forward(B, true, st)) || (rc = mlp->backward(B, st)))
        return rc;
    return loss ? mlp->read_loss(B * (int64_t)mlp->n_out, st, loss) : CFK_OK;
}

int cfk_mlp_adam_step(cfk_mlp mlp, float lr, float beta1, float beta2, float eps, void *stream) {
    if (!mlp) return set_error(CFK_EINVAL, "null argument");
    if (!(lr >= 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
        return set_error(CFK_EINVAL, "bad Adam hyper-parameters");
    return mlp->adam(lr, beta1, beta2, eps, (hipStream_t)stream);
}

}  // extern "C"

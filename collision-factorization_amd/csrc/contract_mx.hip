// Mixed-precision angular-first contraction with the radial stage on the matrix cores, operands
// moved between lanes in registers.
//
// Work decomposition as af_item (contract.hip): one wave = one output row q1 x 64 distributions of
// the batch, every lane walking the same symmetrized, grouped routing-table entries; stage 1 builds
// lane b's Phi_b[k2][k3] = sum_z g_z f_b[q2_z,k2] f_b[q3_z,k3] in FP32 on the VALU.
//
// Stage 2, Q[k1][b] += sum_p R_t[k1][p] Phi_b[p] with p = (k2,k3), is a product over p:
//   D[k1][b] += A[k1][p] B[p][b],  A = R_t (k1 padded to 16 rows), B = Phi (one column per distribution),
// run as v_mfma_f32_16x16x4_f32 steps of 4 values of p, one MFMA per 16-distribution column block T.
// The operand layout wants lane (g, n) = (lane / 16, lane % 16) to hold B[4s+g][16T+n], i.e. a value of
// ANOTHER lane (16T+n)'s Phi.  For each step the 4 Phi registers x_0..x_3 of p = 4s..4s+3 are transposed
// across the four 16-lane rows of the wave -- register index <-> row index -- by two
// v_permlane32_swap and two v_permlane16_swap (gfx950), after which register y_T is exactly the B
// operand of block T.  The A operand of step s (R_t[k1 = n][p = 4s+g] in lane (g, n)) is shared by the
// 4 blocks and comes pre-arranged from memory, 4 steps per 16-byte load.
//
// The 4 FP32 accumulators (one 16x16 block each: lane (g,n) holds D[4g+r][16T+n], r = 0..3) stay in the
// matrix unit's registers for the whole row: every slice of row q1 adds into the same outputs.  Slices
// whose partner block is l = 0 (flag SLICE_PRECISE: large, cancelling Maxwellian-times-perturbation
// terms) run their radial stage in FP64 on the VALU instead (FP64 R through the scalar path, one lane =
// one distribution, sums kept in LDS), as in the VALU kernel.  At the end of the row the precise part is moved into the
// matrix layout by the same register transpose, the two parts are added in FP64 and the conservation
// projection zeroes the five invariants.  On (k2,k3)-symmetric slices
// (l2 == l3) Phi is folded onto k2 <= k3 first and only those p are stepped (14 instead of 25 steps at
// K = 10).
//
// On gfx950 FP32 MFMAs of one wave and FP32 VALU work of another wave on the same SIMD do not run
// concurrently (measured with register-only loops: both kinds together take the sum of their times), so
// the gain over the VALU kernel is the cheaper radial stage (VALU FMAs that read R from SGPRs issue one
// wave at a time), not overlap; at K = 10 the two kernels are equally fast,
// from K = 12 this one is faster (the 16-row tiles are fuller and the VALU kernel needs half chunks).
#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "af_device.hpp"
#include "common.hpp"

namespace cfk {

typedef float mx_f32x4 __attribute__((ext_vector_type(4)));

// Step shapes of one k3 chunk of width KC: P = K*KC values of p (NSF steps), and on symmetric
// blocks (full chunk only) the K(K+1)/2 values k2 <= k3 (NSS steps).  A-operand rows are stored in
// groups of 4 steps (one 16-byte load per lane).
template <int K, int KC>
struct MxShape {
    static constexpr int P = K * KC;
    static constexpr int NSF = (P + 3) / 4;
    static constexpr int PS = K * (K + 1) / 2;
    static constexpr int NSS = (PS + 3) / 4;
    static constexpr int NG = (NSF + 3) / 4;  // 4-step groups per chunk
    static constexpr int NCH = (K + KC - 1) / KC;
};

// LDS region of one wave, in doubles: the A operands of the current chunk [NG][64] float4 (fetched
// straight into LDS at the chunk start) while the matrix-core slices run, then the FP64 sums [K][64] of
// the trailing precise slices.  With a row split the second wave's region, free by then, carries its
// [16][64] partial sums to the first.
template <int K, int KC, bool SPLIT>
constexpr int mx_region_doubles() {
    constexpr int a = MxShape<K, KC>::NG * WAVE * 2, q = K * WAVE;
    constexpr int r = a > q ? a : q;
    return SPLIT && r < 16 * WAVE ? 16 * WAVE : r;
}

// p -> (k2, j) of a full (non-folded) chunk: j = p / K (k3 = c0 + j), k2 = p % K.
// p -> (k2, k3) of the folded triangle: k3-major, k2 = 0..k3.
template <int K>
struct TriIndex {
    int k2[K * (K + 1) / 2], k3[K * (K + 1) / 2];
    constexpr TriIndex() : k2(), k3() {
        int p = 0;
        for (int c = 0; c < K; ++c)
            for (int r = 0; r <= c; ++r) {
                k2[p] = r;
                k3[p] = c;
                ++p;
            }
    }
};

// 4x4 transpose of (register index) x (16-lane row index): afterwards x[T] in row g holds what
// x[g] held in row T.  v_permlane32_swap exchanges rows {2,3} of its first operand with rows {0,1}
// of its second; v_permlane16_swap exchanges rows {1,3} of the first with rows {0,2} of the second.
__device__ __forceinline__ void transpose4(float &x0, float &x1, float &x2, float &x3) {
#if defined(__HIP_DEVICE_COMPILE__)
    auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(x0), __float_as_uint(x2), false, false);
    auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(x1), __float_as_uint(x3), false, false);
    auto c = __builtin_amdgcn_permlane16_swap(a[0], b[0], false, false);
    auto d = __builtin_amdgcn_permlane16_swap(a[1], b[1], false, false);
    x0 = __uint_as_float(c[0]);
    x1 = __uint_as_float(c[1]);
    x2 = __uint_as_float(d[0]);
    x3 = __uint_as_float(d[1]);
#endif
}

// The same transpose of four FP64 registers (each as two 32-bit halves).
__device__ __forceinline__ void transpose4_f64(double (&x)[4]) {
    uint32_t lo[4], hi[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint64_t u = __double_as_longlong(x[i]);
        lo[i] = (uint32_t)u;
        hi[i] = (uint32_t)(u >> 32);
    }
    float l0 = __uint_as_float(lo[0]), l1 = __uint_as_float(lo[1]), l2 = __uint_as_float(lo[2]), l3 = __uint_as_float(lo[3]);
    float h0 = __uint_as_float(hi[0]), h1 = __uint_as_float(hi[1]), h2 = __uint_as_float(hi[2]), h3 = __uint_as_float(hi[3]);
    transpose4(l0, l1, l2, l3);
    transpose4(h0, h1, h2, h3);
    const float lv[4] = {l0, l1, l2, l3}, hv[4] = {h0, h1, h2, h3};
#pragma unroll
    for (int i = 0; i < 4; ++i)
        x[i] = __longlong_as_double((long long)(((uint64_t)__float_as_uint(hv[i]) << 32) | __float_as_uint(lv[i])));
}

template <int K, int KC, bool SPLIT>
__device__ __forceinline__ void af_mx_item(const float *__restrict__ fT_in, double *__restrict__ QT_in,
                                           const Slice *__restrict__ slices, const int32_t *__restrict__ row_ptr,
                                           const int2 *__restrict__ zq, const float *__restrict__ zg,
                                           const double *__restrict__ R, const float *__restrict__ Rmx, int64_t B,
                                           int N_Q, int accumulate, int row, int64_t tile,
                                           const int32_t *__restrict__ row_mid, int half, double *lds) {
    using S = MxShape<K, KC>;
    constexpr int K2 = K * K, K3 = K2 * K;
    constexpr bool FULL = KC == K;  // one k3 chunk
    // f as packed FP32 rows: chunk ch of row (q, b) at float ((q NCH + ch) B + b) KCP
    constexpr int KCP = kcp_of(KC), NCH = S::NCH;
    const int lane = threadIdx.x % WAVE;
    const gptr<double> QT = (gptr<double>)QT_in;
    const cptr<Slice> slices_c = (cptr<Slice>)slices;
    const cptr<int32_t> row_ptr_c = (cptr<int32_t>)row_ptr;
    const cptr<int2> zq_c = (cptr<int2>)zq;
    const cptr<float> zg_c = (cptr<float>)zg;
    const cptr<double> R_c = (cptr<double>)R;
    const int64_t b = tile * WAVE + lane;
    const bool act = b < B;
    const uint32_t lo = act ? (uint32_t)b * (uint32_t)(KCP * 4) : 0u;
    const frsrc_t frs = make_frsrc(fT_in, (uint32_t)((int64_t)N_Q * B * NCH * KCP * 4));
    // the KC coefficients of chunk ch of row q (the f3 of an entry; zero past K)
    auto load_chunk = [&](int q, int ch, float (&dst)[KC]) {
        const int soff = (q * NCH + ch) * (int)B * (KCP * 4);
#pragma unroll
        for (int c = 0; c < KCP / 4; ++c) {
            const float4 v = fload4(frs, lo + 16u * c, soff);
            const float ve[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (4 * c + e < KC) dst[4 * c + e] = ve[e];
        }
    };
    // all K coefficients of row q (the f2 of a q2 group)
    auto load_row = [&](int q, float (&dst)[K]) {
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const int soff = (q * NCH + ch) * (int)B * (KCP * 4);
#pragma unroll
            for (int c = 0; c < KCP / 4; ++c) {
                const float4 v = fload4(frs, lo + 16u * c, soff);
                const float ve[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int j = 4 * c + e, k = ch * KC + j;
                    if (j < KC && k < K) dst[k] = ve[e];
                }
            }
        }
    };
    int s0 = row_ptr_c[row], s1 = row_ptr_c[row + 1];
    if (SPLIT) {
        const int mid = ((cptr<int32_t>)row_mid)[row];
        if (half == 0)
            s1 = mid;
        else
            s0 = mid;
    }
    mx_f32x4 acc[4];
#pragma unroll
    for (int T = 0; T < 4; ++T) acc[T] = mx_f32x4{0.f, 0.f, 0.f, 0.f};
    constexpr int RD = mx_region_doubles<K, KC, SPLIT>();
    mx_f32x4 *const abuf = reinterpret_cast<mx_f32x4 *>(lds + __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE) * RD);  // [NG][64]
    constexpr TriIndex<K> tri{};
    // ---- stage 1 of chunk c0 of slice sl: Phi[k2][j] (k3 = c0 + j) over the slice's q2 groups, FP32.
    // A group accumulates h = sum g f3 and is closed (phi += f2 (x) h) by the next group's first entry or
    // the end of the slice.  With a full chunk the first group is peeled off and its close writes phi by
    // a multiply (0 + x*y rounds as x*y); with k3 chunks phi is zero-filled (the peel's second copy of
    // the entry loop costs registers there).
    auto build_phi = [&](const Slice &sl, int c0, float (&phi)[K][KC]) {
        const int ch = c0 / KC;
        float f2[K], h[KC];
        int zmain = sl.z0;
#pragma unroll
        for (int j = 0; j < KC; ++j) h[j] = 0.f;
        if constexpr (FULL) {
            const int zg2 = sl.z0 + (sl.q1 >> SLICE_G1_SHIFT);
            load_row(~zq_c[sl.z0].x, f2);
            for (int zb = sl.z0; zb < zg2; zb += ZB) {
                int2 qv[ZB];
                float gv[ZB];
#pragma unroll
                for (int e = 0; e < ZB; ++e) {
                    qv[e] = zq_c[zb + e];
                    gv[e] = zg_c[zb + e];
                }
#pragma unroll
                for (int e = 0; e < ZB; ++e) {
                    if (zb + e < zg2) {
                        float f3[KC];
                        load_chunk(qv[e].y, ch, f3);
                        const float gg = in_vgpr(gv[e]);
#pragma unroll
                        for (int j = 0; j < KC; ++j) h[j] = fmaf(gg, f3[j], h[j]);
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < KC; ++j) {
#pragma unroll
                for (int k2 = 0; k2 < K; ++k2) phi[k2][j] = f2[k2] * h[j];
                h[j] = 0.f;
            }
            zmain = zg2;
        } else {
#pragma unroll
            for (int k2 = 0; k2 < K; ++k2)
#pragma unroll
                for (int j = 0; j < KC; ++j) phi[k2][j] = 0.f;
        }
        bool have = false;
        for (int zb = zmain; zb < sl.z1; zb += ZB) {
            int2 qv[ZB];
            float gv[ZB];
#pragma unroll
            for (int e = 0; e < ZB; ++e) {
                qv[e] = zq_c[zb + e];
                gv[e] = zg_c[zb + e];
            }
#pragma unroll
            for (int e = 0; e < ZB; ++e) {
                if (zb + e < sl.z1) {
                    const int2 qq = qv[e];
                    // The entry's f3 chunk and, for a new group, its f2 row are requested before the
                    // close of the previous group, whose products then cover their latency (the new
                    // row goes to f2n while the close still reads f2).  At K = 12 the second row
                    // costs the fourth wave per SIMD (121 -> 133 VGPRs) and measured slower: there the
                    // row is loaded after the close.
                    constexpr bool F2N = K != 12;
                    const float gg = in_vgpr(gv[e]);
                    float f3[KC];
                    load_chunk(qq.y, ch, f3);
                    if (qq.x < 0) {
                        float f2n[K];
                        if constexpr (F2N) load_row(~qq.x, f2n);
                        if (have) {
#pragma unroll
                            for (int j = 0; j < KC; ++j) {
#pragma unroll
                                for (int k2 = 0; k2 < K; ++k2) phi[k2][j] = fmaf(f2[k2], h[j], phi[k2][j]);
                                h[j] = 0.f;
                            }
                        }
                        have = true;
#pragma unroll
                        for (int j = 0; j < KC; ++j) h[j] = fmaf(gg, f3[j], h[j]);
                        if constexpr (F2N) {
#pragma unroll
                            for (int k2 = 0; k2 < K; ++k2) f2[k2] = f2n[k2];
                        } else {
                            load_row(~qq.x, f2);
                        }
                    } else {
#pragma unroll
                        for (int j = 0; j < KC; ++j) h[j] = fmaf(gg, f3[j], h[j]);
                    }
                }
            }
        }
        if (have) {
#pragma unroll
            for (int j = 0; j < KC; ++j)
#pragma unroll
                for (int k2 = 0; k2 < K; ++k2) phi[k2][j] = fmaf(f2[k2], h[j], phi[k2][j]);
        }
    };
    // The host orders each row's precise slices (partner block l = 0, FP64 radial stage) after its other
    // slices: this wave's range [s0, s1) ends with its precise slices [sp, s1), which run after the
    // matrix-core loop with their FP64 sums in registers only then.
    int sp = s1;
    while (sp > s0 && (slices_c[sp - 1].q1 & SLICE_PRECISE) != 0) --sp;
    for (int s = s0; s < sp; ++s) {
        const Slice sl = slices_c[s];
        const bool symm = FULL && (sl.q1 & SLICE_SYMMETRIC) != 0;
#pragma unroll 1
        for (int c0 = 0; c0 < K; c0 += KC) {
            // The chunk's A operands go straight from memory into this wave's LDS area (one 16-byte
            // element per lane and 4-step group, landing at lane order) while stage 1 runs; no
            // registers are held for them until stage 2 reads them back.
            const int64_t gbase = ((int64_t)sl.t * S::NCH + c0 / KC) * S::NG;  // first 4-step group of the chunk
            {
                const int ng = symm ? (S::NSS + 3) / 4 : S::NG;
#pragma unroll
                for (int g4 = 0; g4 < S::NG; ++g4)
                    if (g4 < ng)
                        __builtin_amdgcn_global_load_lds(
                            (const __attribute__((address_space(1))) void *)(Rmx + ((gbase + g4) * WAVE + lane) * 4),
                            (__attribute__((address_space(3))) void *)(abuf + g4 * WAVE), 16, 0, 0);
            }
            float phi[K][KC];
            build_phi(sl, c0, phi);
            // ---- stage 2 on the matrix cores.  The LDS fills of the A operands are vector-memory
            // operations: wait for them (the f loads of stage 1 have all been consumed by now).
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
            if (symm) {
                if constexpr (FULL) {
#pragma unroll
                    for (int k3 = 1; k3 < K; ++k3)
#pragma unroll
                        for (int k2 = 0; k2 < k3; ++k2) phi[k2][k3] += phi[k3][k2];
#pragma unroll
                    for (int g4 = 0; g4 < (S::NSS + 3) / 4; ++g4) {
                        const mx_f32x4 a4 = abuf[g4 * WAVE + lane];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int st = 4 * g4 + u;
                            if (st < S::NSS) {
                                float x[4];
#pragma unroll
                                for (int r = 0; r < 4; ++r) {
                                    const int p = 4 * st + r;
                                    x[r] = p < S::PS ? phi[tri.k2[p]][tri.k3[p]] : 0.f;
                                }
                                transpose4(x[0], x[1], x[2], x[3]);
#pragma unroll
                                for (int T = 0; T < 4; ++T)
                                    acc[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[u], x[T], acc[T], 0, 0, 0);
                            }
                        }
                    }
                }
            } else {
#pragma unroll
                for (int g4 = 0; g4 < S::NG; ++g4) {
                    const mx_f32x4 a4 = abuf[g4 * WAVE + lane];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int st = 4 * g4 + u;
                        if (st < S::NSF) {
                            float x[4];
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int p = 4 * st + r;
                                x[r] = (p < S::P && c0 + p / K < K) ? phi[p % K][p / K] : 0.f;
                            }
                            transpose4(x[0], x[1], x[2], x[3]);
#pragma unroll
                            for (int T = 0; T < 4; ++T)
                                acc[T] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[u], x[T], acc[T], 0, 0, 0);
                        }
                    }
                }
            }
        }
    }
    // Their FP64 sums (lane = distribution) stay in this wave's LDS region, free of A operands by now,
    // across the stage-1 passes; each lane reads and writes only its own column.
    double *const qpl = lds + __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE) * RD;  // [K][64]
    if (sp < s1) {
        __builtin_amdgcn_s_waitcnt(0x0F70);  // the last chunk's LDS fills have landed (vmcnt(0))
#pragma unroll
        for (int k = 0; k < K; ++k) qpl[k * WAVE + lane] = 0.0;
    }
    for (int s = sp; s < s1; ++s) {
        const Slice sl = slices_c[s];
        const cptr<double> Rt = R_c + (int64_t)sl.t * K3;
#pragma unroll 1
        for (int c0 = 0; c0 < K; c0 += KC) {
            float phi[K][KC];
            build_phi(sl, c0, phi);
            // ---- stage 2 in FP64 on the VALU: one k3-plane of R_t per k3, scalar batches
            constexpr int RB = (K2 + 1) / 2 < 25 ? (K2 + 1) / 2 : 25;
            double qp[K];
#pragma unroll
            for (int k = 0; k < K; ++k) qp[k] = qpl[k * WAVE + lane];
#pragma unroll
            for (int j = 0; j < KC; ++j) {
                if (c0 + j < K) {
                    const cptr<double> plane = Rt + (c0 + j) * K2;
#pragma unroll
                    for (int e0 = 0; e0 < K2; e0 += RB) {
                        double rr[RB];
#pragma unroll
                        for (int e = 0; e < RB; ++e)
                            if (e0 + e < K2) rr[e] = plane[e0 + e];
#pragma unroll
                        for (int e = 0; e < RB; ++e)
                            if (e0 + e < K2) {
                                const int k2 = (e0 + e) / K, k1 = (e0 + e) % K;
                                qp[k1] = fma(rr[e], (double)phi[k2][j], qp[k1]);
                            }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < K; ++k) qpl[k * WAVE + lane] = qp[k];
        }
    }
    // ---- combine.  The precise part qp (lane = distribution 16T+n, register k1) moves into the matrix
    // layout (lane (g,n), block T, register r, k1 = 4g + r) by the same 4x4 transpose, once per r:
    // registers qp[4G + r], G = 0..3, become pv[T][r] = qp_{16T+n}[4g + r].
    double qp[K];
#pragma unroll
    for (int k = 0; k < K; ++k) qp[k] = sp < s1 ? qpl[k * WAVE + lane] : 0.0;
    double pv[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        double x[4];
#pragma unroll
        for (int G = 0; G < 4; ++G) x[G] = (4 * G + r < K) ? qp[4 * G + r] : 0.0;
        transpose4_f64(x);
#pragma unroll
        for (int T = 0; T < 4; ++T) pv[T][r] = x[T] + (double)acc[T][r];
    }
    if constexpr (SPLIT) {
        // the second half hands its sums to the first through LDS ([16][64] doubles per tile); the
        // first adds them after its own (a fixed order: deterministic)
        double *const part = lds + (__builtin_amdgcn_readfirstlane(threadIdx.x / WAVE) | 1) * RD;  // the second wave's region
        if (half == 1) {
#pragma unroll
            for (int T = 0; T < 4; ++T)
#pragma unroll
                for (int r = 0; r < 4; ++r) part[(4 * T + r) * WAVE + lane] = pv[T][r];
        }
        __syncthreads();
        if (half == 1) return;
#pragma unroll
        for (int T = 0; T < 4; ++T)
#pragma unroll
            for (int r = 0; r < 4; ++r) pv[T][r] += part[(4 * T + r) * WAVE + lane];
    }
    const int g = lane / 16, n = lane % 16;
#pragma unroll
    for (int T = 0; T < 4; ++T) {
        const int64_t bb = tile * WAVE + 16 * T + n;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k1 = 4 * g + r;
            if (k1 < K && bb < B) {
                double v = pv[T][r];
                if ((row < 4 && k1 == 0) || (row == 0 && k1 == 1)) v = 0.0;  // conservation projection
                const gptr<double> o = QT + ((int64_t)k1 * N_Q + row) * B + bb;
                *o = accumulate ? *o + v : v;
            }
        }
    }
}


// blockDim.x / 64 waves per workgroup.  Without a row split each wave is one batch tile of the same
// output row; with it, consecutive wave pairs are one tile, each wave half of the row's slices.
template <int K, int KC, bool SPLIT>
__global__ __launch_bounds__(4 * WAVE) __attribute__((amdgpu_waves_per_eu(3))) void af_mx_kernel(const float *__restrict__ fT, double *__restrict__ QT,
                                                       const Slice *__restrict__ slices,
                                                       const int32_t *__restrict__ row_ptr,
                                                       const int32_t *__restrict__ row_order,
                                                       const int2 *__restrict__ zq, const float *__restrict__ zg,
                                                       const double *__restrict__ R, const float *__restrict__ Rmx,
                                                       int64_t B, int N_Q, int accumulate,
                                                       const int32_t *__restrict__ row_mid, int nx, int gx) {
    extern __shared__ double mx_lds[];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    // Workgroup index -> (column workgroup xw, output row): the nx column workgroups run in groups of gx,
    // each group's rows heaviest first, so the waves in flight read the f rows of one group only (they
    // stay in L2).
    const int64_t per = (int64_t)gx * N_Q, g = (int64_t)blockIdx.x / per, rem = (int64_t)blockIdx.x - g * per;
    const int64_t x0 = g * gx;
    const int gw = (int)(nx - x0 < gx ? nx - x0 : gx);
    const int rix = (int)(rem / gw);
    const int64_t xw = x0 + (rem - (int64_t)rix * gw);
    int64_t tile;
    int half = -1;
    if (SPLIT) {
        // every wave reaches the barrier of the combine (a tile past the batch end runs with all
        // lanes inactive)
        tile = xw * (blockDim.x / (2 * WAVE)) + (w >> 1);
        half = w & 1;
    } else {
        tile = xw * (blockDim.x / WAVE) + w;
        if (tile * WAVE >= B) return;  // wave-uniform; no barrier without a row split
    }
    af_mx_item<K, KC, SPLIT>(fT, QT, slices, row_ptr, zq, zg, R, Rmx, B, N_Q, accumulate, row_order[rix], tile,
                             row_mid, half, mx_lds);
}

// k3 chunk of the matrix-core kernel: full width up to K = 10 (Phi in 100 registers), half above.
constexpr int mx_kc(int K) { return K <= 10 ? K : (K + 1) / 2; }

// Batched uneven sizes: one work item (group, output row, 64-distribution tile) per workgroup of one
// wave, or of two waves taking the halves of the row's slices (SPLIT), in the order the host sorted
// the items (heaviest first; the dispatcher hands out workgroups in index order).
template <int K, bool SPLIT>
__global__ __launch_bounds__(2 * WAVE) __attribute__((amdgpu_waves_per_eu(3))) void multi_mx_kernel(const MultiGroupDev *__restrict__ groups,
                                                           const int4 *__restrict__ items, int i0, int i1) {
    extern __shared__ double mx_lds[];
    const int it = i0 + (int)blockIdx.x;
    if (it >= i1) return;  // uniform over the workgroup
    const int4 w = items[it];
    const MultiGroupDev &G = groups[w.x];
    const int half = SPLIT ? __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE) : -1;
    af_mx_item<K, mx_kc(K), SPLIT>(G.fT, G.Q, G.slices, G.row_ptr, G.zq, G.zg, G.R, G.Rmx, G.B, G.N_Q, G.accumulate, w.y,
                                   (int64_t)w.z, G.row_mid, half, mx_lds);
}

bool mx_supports_k(int K) { return K == 6 || K == 8 || K == 10 || K == 12 || K == 14; }

int mx_chunk(int K) { return mx_kc(K); }

// Host: the A operands of every symmetrized block t and k3 chunk c, in 4-step groups:
//   Rmx[(((t*NCH + c)*NG + g4)*64 + lane)*4 + u] = R'_t[k1 = lane % 16][p] for step s = 4 g4 + u,
//   p = 4 s + lane / 16,
// with p enumerating (k2, k3 = c*KC + p / K) on plain blocks and the triangle k2 <= k3 (k3-major) on
// (k2,k3)-symmetric ones (full chunk only); zero for k1 >= K and past the last p.
int build_mx_tables(cfk_plan_s *p, std::vector<float> &out, int64_t &n_groups) {
    const int K = p->K;
    if (!mx_supports_k(K)) return set_error(CFK_EUNSUP, "matrix-core kernel: K=%d", K);
    const int KC = mx_kc(K), P = K * KC, NSF = (P + 3) / 4, NG = (NSF + 3) / 4, NCH = (K + KC - 1) / KC;
    const int64_t K3 = (int64_t)K * K * K;
    const int64_t nblk = (int64_t)p->sym_R.n / K3;
    std::vector<double> R((size_t)p->sym_R.n);
    CFK_HIP(hipMemcpy(R.data(), p->sym_R.p, R.size() * sizeof(double), hipMemcpyDeviceToHost));
    // which blocks are (k2,k3)-symmetric: from the device slice flags' host mirror is not kept, so
    // test the merged block itself (exact symmetry holds for l2 == l3 by construction)
    std::vector<char> symm((size_t)nblk, 0);
    {
        std::vector<Slice> sl((size_t)p->sym_slices.n);
        CFK_HIP(hipMemcpy(sl.data(), p->sym_slices.p, sl.size() * sizeof(Slice), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < p->sym_S; ++i)
            if (sl[(size_t)i].q1 & SLICE_SYMMETRIC) symm[(size_t)sl[(size_t)i].t] = 1;
    }
    n_groups = nblk * NCH * NG;
    out.assign((size_t)(n_groups * 64 * 4), 0.0f);
    for (int64_t t = 0; t < nblk; ++t) {
        const double *Rt = R.data() + (size_t)(t * K3);
        for (int c = 0; c < NCH; ++c)
            for (int s = 0; s < 4 * NG; ++s)
                for (int l = 0; l < 64; ++l) {
                    const int k1 = l % 16, pp = 4 * s + l / 16;
                    if (k1 >= K) continue;
                    int k2, k3;
                    if (symm[(size_t)t] && KC == K) {
                        if (pp >= K * (K + 1) / 2) continue;
                        int rem = pp, cc = 0;
                        while (rem > cc) {
                            rem -= cc + 1;
                            ++cc;
                        }
                        k2 = rem;
                        k3 = cc;
                    } else {
                        if (pp >= P) continue;
                        k2 = pp % K;
                        k3 = c * KC + pp / K;
                        if (k3 >= K) continue;
                    }
                    const size_t g4 = (size_t)(((t * NCH + c) * NG) + s / 4);
                    out[(g4 * 64 + l) * 4 + s % 4] = (float)Rt[k1 + k2 * K + (int64_t)k3 * K * K];
                }
    }
    return CFK_OK;
}

namespace {
template <int K, bool SPLIT>
void launch_mx_k(const cfk_plan_s *p, const float *Rmx, int wpb, const float *fT, double *QT, int64_t B, int acc,
                 hipStream_t st) {
    constexpr int KC = mx_kc(K);
    const int64_t tiles = (B + WAVE - 1) / WAVE;
    int tpw, waves;  // batch tiles and waves per workgroup
    if (SPLIT) {
        tpw = wpb >= 2 ? 2 : 1;
        waves = 2 * tpw;
    } else {
        tpw = wpb == 2 || wpb == 4 ? wpb : 1;
        waves = tpw;
    }
    // column groups of 1024 distributions (whole workgroups), as the VALU kernels of the grouped tables
    const int nx = (int)((tiles + tpw - 1) / tpw), gx = std::max(1, 1024 / (WAVE * tpw));
    const dim3 grid((unsigned)((int64_t)nx * p->N_Q));
    constexpr int rd = mx_region_doubles<K, KC, SPLIT>();
    const size_t lds_bytes = sizeof(double) * (size_t)waves * rd;
    hipLaunchKernelGGL((af_mx_kernel<K, KC, SPLIT>), grid, dim3(WAVE * waves), lds_bytes, st, fT, QT,
                       p->sym_slices.p, p->sym_row_ptr.p, p->sym_row_order.p, p->sym_zq.p, p->sym_zg32.p, p->sym_R.p,
                       Rmx, B, p->N_Q, acc, SPLIT ? p->sym_row_mid.p : nullptr, nx, gx);
}
}  // namespace

namespace {
template <int K, bool SPLIT>
void launch_multi_mx_k(const MultiGroupDev *groups, const int4 *items, int i0, int i1, hipStream_t st) {
    constexpr int waves = SPLIT ? 2 : 1;
    constexpr int rd = mx_region_doubles<K, mx_kc(K), SPLIT>();
    const size_t lds_bytes = sizeof(double) * waves * (size_t)rd;
    hipLaunchKernelGGL((multi_mx_kernel<K, SPLIT>), dim3((unsigned)(i1 - i0)), dim3(WAVE * waves), lds_bytes, st, groups,
                       items, i0, i1);
}
}  // namespace

// Items [i0, i1) of one K-class of the batched path on the matrix-core kernel.
int launch_multi_mx(const MultiGroupDev *groups, const int4 *items, int K, bool split, int i0, int i1,
                    hipStream_t st) {
    if (i1 <= i0) return CFK_OK;
#define CFK_MMX(KK)                                                     \
    case KK:                                                            \
        if (split)                                                      \
            launch_multi_mx_k<KK, true>(groups, items, i0, i1, st);     \
        else                                                            \
            launch_multi_mx_k<KK, false>(groups, items, i0, i1, st);    \
        break;
    switch (K) {
        CFK_MMX(6)
        CFK_MMX(8)
        CFK_MMX(10)
        CFK_MMX(12)
        CFK_MMX(14)
        default:
            return set_error(CFK_EUNSUP, "matrix-core multi kernel: K=%d", K);
    }
#undef CFK_MMX
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

int launch_af_mx(const cfk_plan_s *p, const float *Rmx, const Tile &tl, const float *fT, double *QT, int64_t B,
                 int accumulate, hipStream_t st) {
    if (B <= 0) return CFK_OK;
#define CFK_MX(KK)                                                                  \
    case KK:                                                                        \
        if (tl.split)                                                               \
            launch_mx_k<KK, true>(p, Rmx, tl.wpb, fT, QT, B, accumulate, st);       \
        else                                                                        \
            launch_mx_k<KK, false>(p, Rmx, tl.wpb, fT, QT, B, accumulate, st);      \
        break;
    switch (p->K) {
        CFK_MX(6)
        CFK_MX(8)
        CFK_MX(10)
        CFK_MX(12)
        CFK_MX(14)
        default:
            return set_error(CFK_EUNSUP, "matrix-core kernel: K=%d", p->K);
    }
#undef CFK_MX
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

}  // namespace cfk

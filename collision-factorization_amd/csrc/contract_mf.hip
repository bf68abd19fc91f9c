// Mixed-precision angular-first contraction with the radial stage on the matrix cores.
//
// Same work decomposition and stage 1 as af_item in contract.hip (one wave = 64 distributions x one
// output row; Phi = sum g f2 (x) f3 built per lane in FP32 on the VALU over grouped, symmetrized
// tables).  Stage 2, Q[k1] += sum_{k2,k3} R_t[k1,k2,k3] Phi[k2,k3], runs for each k3 chunk as a
// product on v_mfma_f32_16x16x4_f32 (FP32 operands, FP32 accumulation):
//   Out[b][k1] = sum_p Phi[b][p] R[k1][p],   p = (k2, k3 in the chunk),
// with M = 16 distributions per block (4 blocks), N = k1 (K padded to 16), and 4 values of p per
// instruction.  The MFMA needs Phi of 16 distributions spread over lane groups by p, while stage 1
// leaves each lane with all of its own distribution's Phi, so each lane stages its Phi row in LDS
// (in the operand order) and reads back its operands; R comes pre-arranged in the operand order, one
// contiguous 256-byte row per instruction.  The per-slice FP32 sums are added into FP64 row
// accumulators, as on the VALU path; the precise slices (partner block l = 0) run in FP64 on the VALU
// as there.  The matrix unit executes stage 2 while the vector unit runs other waves' stage 1.
#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "af_device.hpp"
#include "common.hpp"

namespace cfk {

typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int K, int KC>
struct MfShape {
    static constexpr int P = K * KC;            // (k2, k3) pairs of one chunk
    static constexpr int TQ = (P + 3) / 4;      // MFMA steps per chunk (4 values of p each)
    static constexpr int SEG = (TQ + 3) / 4 * 4;  // floats per lane group in a staged row (16-byte multiple)
    static constexpr int PS = 4 * SEG + 4;      // staged row stride in floats
    static constexpr int NCH = (K + KC - 1) / KC;
    static constexpr size_t LDS = sizeof(float) * 64 * PS;
    static_assert(LDS >= sizeof(double) * 64 * K, "staging rows double as the output combine buffer");
};

template <int K, int KC>
__global__ __launch_bounds__(WAVE) void af_mf_kernel(const float *__restrict__ fT_in, double *__restrict__ QT_in,
                                                     const Slice *__restrict__ slices,
                                                     const int32_t *__restrict__ row_ptr,
                                                     const int32_t *__restrict__ row_order, const int2 *__restrict__ zq,
                                                     const float *__restrict__ zg, const double *__restrict__ R,
                                                     const float *__restrict__ Rmf, int64_t B, int N_Q, int accumulate) {
    using S = MfShape<K, KC>;
    constexpr int K2 = K * K, K3 = K2 * K;
    extern __shared__ float lds[];
    const int row = row_order[blockIdx.y];
    const int64_t tile = blockIdx.x;
    const int lane = threadIdx.x % WAVE;
    const int g = lane / 16, m = lane % 16;  // MFMA lane group and row/column within a 16-block
    const gptr<double> QT = (gptr<double>)QT_in;
    const cptr<Slice> slices_c = (cptr<Slice>)slices;
    const cptr<int32_t> row_ptr_c = (cptr<int32_t>)row_ptr;
    const cptr<int2> zq_c = (cptr<int2>)zq;
    const cptr<float> zg_c = (cptr<float>)zg;
    const cptr<double> R_c = (cptr<double>)R;
    const gptr<const float> Rmf_g = (gptr<const float>)Rmf;
    const int64_t b = tile * WAVE + lane;
    const bool act = b < B;
    const uint32_t lo = act ? (uint32_t)b * 4u : 0u;
    const int64_t kstride = (int64_t)N_Q * B;
    const frsrc_t frs = make_frsrc(fT_in, (uint32_t)(K * kstride * 4));
    const int s0 = row_ptr_c[row], s1 = row_ptr_c[row + 1];
    double qv[K];        // precise slices, lane = distribution
    double qm[4][4];     // MFMA slices, lane (g, m): distribution 16 blk + 4 g + r, k1 = m
#pragma unroll
    for (int k = 0; k < K; ++k) qv[k] = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) qm[i][r] = 0.0;
    float *myrow = lds + lane * S::PS;
    for (int s = s0; s < s1; ++s) {
        const Slice sl = slices_c[s];
        const cptr<double> Rt = R_c + (int64_t)sl.t * K3;
        const bool precise = (sl.q1 & SLICE_PRECISE) != 0;
        f32x4 acc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int c = 0; c < S::NCH; ++c) {
            const int c0 = c * KC;
            // ---- stage 1: Phi over the grouped entries (as af_item, GROUPED, one distribution per lane)
            float phi[K][KC];
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int j = 0; j < KC; ++j) phi[k][j] = 0.f;
            {
                float f2[K], h[KC];
#pragma unroll
                for (int j = 0; j < KC; ++j) h[j] = 0.f;
                bool have = false;
                for (int zb = sl.z0; zb < sl.z1; zb += ZB) {
                    int2 qv8[ZB];
                    float gv[ZB];
#pragma unroll
                    for (int e = 0; e < ZB; ++e) {
                        qv8[e] = zq_c[zb + e];
                        gv[e] = zg_c[zb + e];
                    }
#pragma unroll
                    for (int e = 0; e < ZB; ++e) {
                        if (zb + e < sl.z1) {
                            const int2 qq = qv8[e];
                            const float gg = gv[e];
                            if (qq.x < 0) {
                                if (have) {
#pragma unroll
                                    for (int j = 0; j < KC; ++j) {
#pragma unroll
                                        for (int k2 = 0; k2 < K; ++k2) phi[k2][j] = fmaf(f2[k2], h[j], phi[k2][j]);
                                        h[j] = 0.f;
                                    }
                                }
                                have = true;
                                const int64_t u2 = (int64_t)(~qq.x) * B;
#pragma unroll
                                for (int k2 = 0; k2 < K; ++k2) f2[k2] = fload<float>(frs, lo, u2 + k2 * kstride);
                            }
                            const int64_t u3 = (int64_t)qq.y * B;
#pragma unroll
                            for (int j = 0; j < KC; ++j)
                                if (c0 + j < K) h[j] = fmaf(gg, fload<float>(frs, lo, u3 + (c0 + j) * kstride), h[j]);
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < KC; ++j)
#pragma unroll
                    for (int k2 = 0; k2 < K; ++k2) phi[k2][j] = fmaf(f2[k2], h[j], phi[k2][j]);
            }
            if (precise) {
                // ---- stage 2 in FP64 on the VALU (R planes through the scalar path)
                constexpr int RB = (K2 + 1) / 2 < 25 ? (K2 + 1) / 2 : 25;
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
                                    qv[k1] = fma(rr[e], (double)phi[k2][j], qv[k1]);
                                }
                        }
                    }
                }
            } else {
                // ---- stage 2 on the matrix cores: stage Phi row p = k2*KC + j at (p%4)*SEG + p/4
#pragma unroll
                for (int k2 = 0; k2 < K; ++k2)
#pragma unroll
                    for (int j = 0; j < KC; ++j) {
                        const int p = k2 * KC + j;
                        myrow[(p % 4) * S::SEG + p / 4] = (c0 + j < K) ? phi[k2][j] : 0.f;
                    }
#pragma unroll
                for (int p = S::P; p < 4 * S::TQ; ++p) myrow[(p % 4) * S::SEG + p / 4] = 0.f;
                __syncthreads();
                // B operand of step t: lane (g, m) holds R[k1 = m][p = 4t + g] (zero padded)
                const gptr<const float> rb = Rmf_g + (((int64_t)sl.t * S::NCH + c) * S::TQ) * WAVE + lane;
                float bt[S::TQ];
#pragma unroll
                for (int t = 0; t < S::TQ; ++t) bt[t] = rb[t * WAVE];
#pragma unroll
                for (int blk = 0; blk < 4; ++blk) {
                    const float *src = lds + (16 * blk + m) * S::PS + g * S::SEG;
                    float at[S::SEG];
#pragma unroll
                    for (int t = 0; t < S::SEG; t += 4) {
                        const float4 v = *reinterpret_cast<const float4 *>(src + t);
                        at[t] = v.x;
                        at[t + 1] = v.y;
                        at[t + 2] = v.z;
                        at[t + 3] = v.w;
                    }
#pragma unroll
                    for (int t = 0; t < S::TQ; ++t)
                        acc[blk] = __builtin_amdgcn_mfma_f32_16x16x4f32(at[t], bt[t], acc[blk], 0, 0, 0);
                }
                __syncthreads();  // the next chunk overwrites the staged rows
            }
        }
        if (!precise) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) qm[i][r] += (double)acc[i][r];
        }
    }
    // combine the two accumulator layouts through LDS, then write lane = distribution
    double *qc = reinterpret_cast<double *>(lds);  // [64][K]
#pragma unroll
    for (int k1 = 0; k1 < K; ++k1) qc[lane * K + k1] = qv[k1];
    __syncthreads();
    if (m < K) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) qc[(16 * i + 4 * g + r) * K + m] += qm[i][r];
    }
    __syncthreads();
    if (!act) return;
#pragma unroll
    for (int k1 = 0; k1 < K; ++k1) {
        double v = qc[lane * K + k1];
        if ((row < 4 && k1 == 0) || (row == 0 && k1 == 1)) v = 0.0;  // conservation projection
        const gptr<double> o = lane_ptr(QT, ((int64_t)k1 * N_Q + row) * B, (uint32_t)b * 8u);
        *o = accumulate ? *o + v : v;
    }
}

bool mf_supports_k(int K) { return K == 6 || K == 8 || K == 10 || K == 12 || K == 14; }

constexpr int mf_kc(int K) { return (K + 1) / 2; }

// Host: R in the B-operand order of af_mf_kernel, per symmetrized block t and chunk c:
// Rmf[((t*NCH + c)*TQ + tq)*64 + lane] = R_t[k1 = lane%16][k2][k3 = c*KC + j] with
// p = 4*tq + lane/16 = k2*KC + j, zero outside the block.
int build_mf_tables(cfk_plan_s *p) {
    if (p->sym_Rmf.p) return CFK_OK;
    const int K = p->K;
    if (!mf_supports_k(K)) return set_error(CFK_EUNSUP, "matrix-core variant: K=%d", K);
    const int KC = mf_kc(K), P = K * KC, TQ = (P + 3) / 4, NCH = (K + KC - 1) / KC;
    const int64_t K3 = (int64_t)K * K * K;
    const int64_t nblk = (int64_t)p->sym_R.n / K3;
    std::vector<double> R((size_t)p->sym_R.n);
    CFK_HIP(hipMemcpy(R.data(), p->sym_R.p, R.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<float> out((size_t)(nblk * NCH * TQ * 64), 0.0f);
    for (int64_t t = 0; t < nblk; ++t)
        for (int c = 0; c < NCH; ++c)
            for (int tq = 0; tq < TQ; ++tq)
                for (int l = 0; l < 64; ++l) {
                    const int k1 = l % 16, pp = 4 * tq + l / 16;
                    if (k1 >= K || pp >= P) continue;
                    const int k2 = pp / KC, k3 = c * KC + pp % KC;
                    if (k3 >= K) continue;
                    out[(size_t)(((t * NCH + c) * TQ + tq) * 64 + l)] = (float)R[(size_t)(t * K3 + k1 + k2 * K + (int64_t)k3 * K * K)];
                }
    return p->sym_Rmf.upload(out.data(), out.size());
}

int launch_af_mf(cfk_plan_s *p, const float *fT, double *QT, int64_t B, int accumulate, hipStream_t st) {
    int rc;
    if ((rc = build_mf_tables(p))) return rc;
    if (B <= 0) return CFK_OK;
    const dim3 grid((unsigned)((B + WAVE - 1) / WAVE), (unsigned)p->N_Q);
#define CFK_MF(KK)                                                                                              \
    case KK: {                                                                                                  \
        constexpr int KCC = mf_kc(KK);                                                                          \
        constexpr size_t lds = MfShape<KK, KCC>::LDS;                                                           \
        hipLaunchKernelGGL((af_mf_kernel<KK, KCC>), grid, dim3(WAVE), lds, st, fT, QT,                          \
                           p->sym_slices.p, p->sym_row_ptr.p, p->sym_row_order.p, p->sym_zq.p, p->sym_zg32.p,   \
                           p->sym_R.p, p->sym_Rmf.p, B, p->N_Q, accumulate);                                    \
        break;                                                                                                  \
    }
    switch (p->K) {
        CFK_MF(6)
        CFK_MF(8)
        CFK_MF(10)
        CFK_MF(12)
        CFK_MF(14)
        default:
            return set_error(CFK_EUNSUP, "matrix-core variant: K=%d", p->K);
    }
#undef CFK_MF
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

}  // namespace cfk

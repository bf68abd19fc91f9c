// FP64 contraction kernels for gfx950 (wave64).
//
// Work decomposition shared by the sparse variants: one wavefront owns 64 distributions of the
// batch (one per lane) and one output row q1.  Every lane walks the SAME routing-table entries
// (the (t,q1) slices of that row and their Gaunt nonzeros), so the table entries, Gaunt values and
// the radial block R_t are wave-uniform and are fetched with scalar loads into SGPRs; each lane
// only loads its own distribution's coefficients, which sit batch-fastest so a
// wave's load of one (q,k) coefficient is one contiguous 512-byte row ([K][N_Q][B] layout).  Each row's output
// Q[q1,:,b] is owned by exactly one lane, so no atomics are needed and results are deterministic.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "common.hpp"
#include "af_device.hpp"

namespace cfk {

int table_padding() { return ZB; }

#ifndef CFK_K_CASES
#define CFK_K_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)
#endif

// Device pointers of one set of routing tables.
template <typename T>
struct SparseTables {
    const Slice *slices;
    const int32_t *row_ptr;
    const int32_t *row_order;
    const int2 *zq;
    const T *zg;
    const double *R;
    const float *R32;  // FP32 copy of R (mixed path); may be null for double tables
    const int32_t *row_mid = nullptr;  // split point of each row's slices (row-split launches only)
};

// ---------------------------------------------------------------------------------------------
// Angular-first (angular_first_collision_kernel_mex.cpp:24-75).  Per slice (t,q1):
//   Phi[k2,k3] = sum_z g_z f[q2_z,k2] f[q3_z,k3]            (type T: double or float, per lane)
//   Q[q1,k1]  += sum_{k2,k3} R_t[k1,k2,k3] Phi[k2,k3]         (always FP64)
// The R.Phi stage streams one contiguous row R_t[0:K,k2,k3] (wave-uniform, scalar loads) per
// (k2,k3) and FMAs it into the K double accumulators of the output row, so only K values of R are
// live at a time.  k3 is processed in chunks of KC columns so Phi fits in registers.
//
// GROUPED tables (symmetrized path) store the entries of a slice sorted by q2; an entry whose q2
// is stored bit-inverted (negative) starts a new q2 group.  Per group the kernel accumulates
//   h[k3] = sum_{entries of the group} g f[q3,k3]   and then   Phi += f[q2,:] (x) h,
// one outer product per distinct (t,q1,q2) instead of one per nonzero.
// PROJ forces the five collision invariants of the contribution to exactly zero.
//
// Mixed precision (T = float): Phi is built in FP32; the R.Phi stage runs in FP32 with an FP32
// per-slice sum (R32, one FP32 row per (k2,k3)) except on slices flagged SLICE_PRECISE (the blocks
// whose partner is the l=0 block, which carry the large, cancelling Maxwellian-times-perturbation
// terms), where it runs in FP64 as for T = double.  On slices flagged SLICE_SYMMETRIC (l2 == l3, where
// the merged block is symmetric in (k2,k3)) the FP32 stage folds Phi onto k2 <= k3 and reads that half
// of R_t only (K(K+1)/2 instead of K^2 rows).
// ---------------------------------------------------------------------------------------------
// One work item of the angular-first contraction: output row `row` for batch tile `tile` (64*NB
// distributions), lane = threadIdx.x % 64 of a one-wave workgroup.

// FP32 stage R_t : Phi of one k3 chunk: a[i][k1] = sum_{k3 in chunk} sum_{k2} R_t[k1,k2,k3] phi[i][k2][k3-c0],
// one k3-plane of R_t (K*K contiguous FP32 values, rows k2, k1 fastest) per k3, fetched through the
// scalar path in batches of RB values that are all in flight together (scalar loads return out of
// order, so each batch costs one wait), each value feeding NB FMAs.  With SYM only the rows k2 <= k3 of
// each plane are read (the caller has folded Phi onto that triangle).  a is written by its first product.
template <int K, int KC, int NB, bool SYM, typename PhiT>
__device__ __forceinline__ void r_phi_f32(cptr<float> Rt32, const PhiT (&phi)[NB][K][KC], float (&a)[NB][K], int c0) {
    constexpr int K2 = K * K;
    constexpr int RB = (K * K + 1) / 2 < 50 ? (K * K + 1) / 2 : 50;  // floats per batch
#pragma unroll
    for (int j = 0; j < KC; ++j) {
        if (c0 + j < K) {
            const cptr<float> plane = Rt32 + (c0 + j) * K2;
            const int n = SYM ? (j + 1) * K : K2;  // compile-time after unrolling
#pragma unroll
            for (int e0 = 0; e0 < K2; e0 += RB) {
                if (e0 < n) {
                    float rr[RB];
#pragma unroll
                    for (int e = 0; e < RB; ++e)
                        if (e0 + e < n) rr[e] = plane[e0 + e];
#pragma unroll
                    for (int e = 0; e < RB; ++e) {
                        if (e0 + e < n) {
                            const int k2 = (e0 + e) / K, k1 = (e0 + e) % K;
#pragma unroll
                            for (int i = 0; i < NB; ++i) {
                                if (j == 0 && k2 == 0)
                                    a[i][k1] = rr[e] * (float)phi[i][k2][j];
                                else
                                    a[i][k1] = fmaf(rr[e], (float)phi[i][k2][j], a[i][k1]);
                            }
                        }
                    }
                }
            }
        }
    }
}

template <typename T, int K, int KC, bool GROUPED, bool PROJ, int NB, bool PF, bool PK = false>
__device__ __forceinline__ void af_item(const T *__restrict__ fT_in, double *__restrict__ QT_in,
                                        const Slice *__restrict__ slices, const int32_t *__restrict__ row_ptr,
                                        const int2 *__restrict__ zq, const T *__restrict__ zg,
                                        const double *__restrict__ R, const float *__restrict__ R32, int64_t B,
                                        int N_Q, int accumulate, int row, int64_t tile,
                                        const int32_t *__restrict__ row_mid = nullptr, int half = -1) {
    constexpr int K2 = K * K, K3 = K2 * K;
    const gptr<double> QT = (gptr<double>)QT_in;
    const cptr<Slice> slices_c = (cptr<Slice>)slices;
    const cptr<int32_t> row_ptr_c = (cptr<int32_t>)row_ptr;
    const cptr<int2> zq_c = (cptr<int2>)zq;
    const cptr<T> zg_c = (cptr<T>)zg;
    const cptr<double> R_c = (cptr<double>)R;
    const cptr<float> R32_c = (cptr<float>)R32;
    // Each lane owns NB distributions b_i = (tile*NB + i)*64 + lane, so every wave-uniform value
    // (table entry, Gaunt value, R row) fetched through the scalar path feeds NB FMAs.
    uint32_t lo[NB], lq[NB];  // byte offsets of the lane's distribution in an f row / a Q row (B * 8 < 2^32)
    bool act[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int64_t b = (tile * NB + i) * WAVE + (threadIdx.x % WAVE);
        act[i] = b < B;
        lo[i] = act[i] ? (uint32_t)b * (uint32_t)(PK ? pk_chunk_bytes<T>(KC) : sizeof(T)) : 0u;
        lq[i] = act[i] ? (uint32_t)b * 8u : 0u;
    }
    int s0 = row_ptr_c[row], s1 = row_ptr_c[row + 1];
    if (half >= 0) {  // row split between the two waves of a workgroup: half 0 [s0, mid), half 1 [mid, s1)
        const int mid = ((cptr<int32_t>)row_mid)[row];
        if (half == 0)
            s1 = mid;
        else
            s0 = mid;
    }
    const int64_t kstride = (int64_t)N_Q * B;  // [K][N_Q][B]
    static_assert(!PK || GROUPED, "packed rows: grouped tables");
    // PK: f as chunk-major packed rows -- chunk c of row (q, b) is KCQ values (KC rounded up to whole
    // 16-byte loads, zero padded) at element ((q NCH + c) B + b) KCQ, so one chunk is KCQ / VPL 16-byte
    // loads of the lane's own row (VPL = 4 floats or 2 doubles).  With KC = K in FP32 this is [N_Q][B][KP].
    constexpr int VPL = 16 / (int)sizeof(T), KCQ = (KC + VPL - 1) / VPL * VPL, NCH = (K + KC - 1) / KC;
    const uint32_t fbytes = PK ? (uint32_t)std::min<int64_t>((int64_t)N_Q * NCH * B * KCQ * (int64_t)sizeof(T), 0xFFFFFFFFLL)
                               : (uint32_t)(K * kstride * (int64_t)sizeof(T));
    const frsrc_t frs = make_frsrc(fT_in, fbytes);
    // byte offset of chunk ch of row q (unsigned: a buffer load's scalar offset is an unsigned 32-bit value)
    auto chunk_off = [&](int q, int ch) -> int {
        return (int)((uint32_t)(q * NCH + ch) * (uint32_t)B * (uint32_t)(KCQ * (int)sizeof(T)));
    };
    // the KC coefficients of chunk ch of row q of lane's distribution i (zero past K)
    auto load_chunk = [&](int q, int ch, int i, T (&dst)[KC]) {
        const int soff = chunk_off(q, ch);
#pragma unroll
        for (int c = 0; c < KCQ / VPL; ++c) {
            if constexpr (std::is_same<T, float>::value) {
                const float4 v = fload4(frs, lo[i] + 16u * c, soff);
                const float ve[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (4 * c + e < KC) dst[4 * c + e] = ve[e];
            } else {
                const f64x2_t d = fload2d(frs, lo[i] + 16u * c, soff);
                if (2 * c < KC) dst[2 * c] = d[0];
                if (2 * c + 1 < KC) dst[2 * c + 1] = d[1];
            }
        }
    };
    // all K coefficients of row q
    auto load_row = [&](int q, int i, T (&dst)[K]) {
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            T part[KC];
            load_chunk(q, ch, i, part);
#pragma unroll
            for (int j = 0; j < KC; ++j)
                if (ch * KC + j < K) dst[ch * KC + j] = part[j];
        }
    };
    double q[NB][K];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int k = 0; k < K; ++k) q[i][k] = 0.0;
    for (int s = s0; s < s1; ++s) {
        const Slice sl = slices_c[s];
        const cptr<double> Rt = R_c + (int64_t)sl.t * K3;
        const cptr<float> Rt32 = R32_c + (int64_t)sl.t * K3;
        const bool precise = !std::is_same<T, float>::value || (sl.q1 & SLICE_PRECISE) != 0;
        // (k2,k3)-symmetric block (l2 == l3): only the k2 <= k3 half of R_t is needed (FP32 stage, one chunk)
        const bool symm = KC == K && GROUPED && !precise && (sl.q1 & SLICE_SYMMETRIC) != 0;
#pragma unroll 1
        for (int c0 = 0; c0 < K; c0 += KC) {
            // Phi and the partial sums below are started by their first product (a multiply) rather
            // than zero-filled: 0 + x*y rounds exactly as x*y, and the fills were 8% of the
            // kernel's vector instructions.
            T phi[NB][K][KC];
            // Table entries are fetched ZB at a time (one scalar round trip per batch; the device arrays
            // carry ZB padding entries so the batch read never leaves them).
            if (GROUPED) {
                // A q2 group accumulates h = sum g f3 and is closed (phi += f2 (x) h) by the next
                // group's first entry or the end of the slice.  In FP32 with a full k3 chunk the slice's
                // first group (its length is in the slice flags) is peeled off and its close writes phi
                // with a multiply; elsewhere phi is zero-filled (in FP64 the peel costs an occupancy
                // step, and with half chunks at K = 14 it made the compiler serialize the f loads).
                constexpr bool PEEL = std::is_same<T, float>::value && KC == K;
                T f2[NB][K], h[NB][KC];
                int zmain = sl.z0;  // first entry of the main loop
                if constexpr (PEEL) {
                    const int zg2 = sl.z0 + (sl.q1 >> SLICE_G1_SHIFT);  // first entry of the second group
                    const int64_t u2 = (int64_t)(~zq_c[sl.z0].x) * B;
#pragma unroll
                    for (int i = 0; i < NB; ++i) {
                        if constexpr (PK)
                            load_row(~zq_c[sl.z0].x, i, f2[i]);
                        else
#pragma unroll
                            for (int k2 = 0; k2 < K; ++k2) f2[i][k2] = fload<T>(frs, lo[i], u2 + k2 * kstride);
#pragma unroll
                        for (int j = 0; j < KC; ++j) h[i][j] = T(0);
                    }
                    for (int zb = sl.z0; zb < zg2; zb += ZB) {
                        int2 qv[ZB];
                        T gv[ZB];
#pragma unroll
                        for (int e = 0; e < ZB; ++e) {
                            qv[e] = zq_c[zb + e];
                            gv[e] = zg_c[zb + e];
                        }
#pragma unroll
                        for (int e = 0; e < ZB; ++e) {
                            if (zb + e < zg2) {
                                const int64_t u3 = (int64_t)qv[e].y * B;
#pragma unroll
                                for (int i = 0; i < NB; ++i) {
                                    if constexpr (PK) {
                                        T f3[KC];
                                        load_chunk(qv[e].y, c0 / KC, i, f3);
                                        const T gg = to_vgpr(gv[e]);
#pragma unroll
                                        for (int j = 0; j < KC; ++j) h[i][j] = fma(gg, f3[j], h[i][j]);
                                    } else {
#pragma unroll
                                        for (int j = 0; j < KC; ++j)
                                            if (c0 + j < K) h[i][j] = fma(gv[e], fload<T>(frs, lo[i], u3 + (c0 + j) * kstride), h[i][j]);
                                    }
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int i = 0; i < NB; ++i)
#pragma unroll
                        for (int j = 0; j < KC; ++j) {
#pragma unroll
                            for (int k2 = 0; k2 < K; ++k2) phi[i][k2][j] = f2[i][k2] * h[i][j];
                            h[i][j] = T(0);
                        }
                    zmain = zg2;
                } else {
#pragma unroll
                    for (int i = 0; i < NB; ++i) {
#pragma unroll
                        for (int k = 0; k < K; ++k)
#pragma unroll
                            for (int j = 0; j < KC; ++j) phi[i][k][j] = T(0);
#pragma unroll
                        for (int j = 0; j < KC; ++j) h[i][j] = T(0);
                    }
                }
                bool have = false;  // wave-uniform: a group is open
                constexpr bool F2N = PK && std::is_same<T, float>::value && NB == 1;
                for (int zb = zmain; zb < sl.z1; zb += ZB) {
                    int2 qv[ZB];
                    T gv[ZB];
#pragma unroll
                    for (int e = 0; e < ZB; ++e) {
                        qv[e] = zq_c[zb + e];
                        gv[e] = zg_c[zb + e];
                    }
#pragma unroll
                    for (int e = 0; e < ZB; ++e) {
                        if (zb + e < sl.z1 && F2N) {
                            // Packed FP32 rows: the entry's f3 chunk and, for a new group, its f2 row
                            // are requested before the close of the previous group, whose products then
                            // cover their latency (the new row goes to f2n while the close reads f2).
                            const int2 qq = qv[e];
                            const T gg = to_vgpr(gv[e]);
                            T f3[NB][KC];
#pragma unroll
                            for (int i = 0; i < NB; ++i) load_chunk(qq.y, c0 / KC, i, f3[i]);
                            if (qq.x < 0) {
                                T f2n[NB][K];
#pragma unroll
                                for (int i = 0; i < NB; ++i) load_row(~qq.x, i, f2n[i]);
                                if (have) {
#pragma unroll
                                    for (int i = 0; i < NB; ++i)
#pragma unroll
                                        for (int j = 0; j < KC; ++j) {
#pragma unroll
                                            for (int k2 = 0; k2 < K; ++k2)
                                                phi[i][k2][j] = fma(f2[i][k2], h[i][j], phi[i][k2][j]);
                                            h[i][j] = T(0);
                                        }
                                }
                                have = true;
#pragma unroll
                                for (int i = 0; i < NB; ++i) {
#pragma unroll
                                    for (int j = 0; j < KC; ++j) h[i][j] = fma(gg, f3[i][j], h[i][j]);
#pragma unroll
                                    for (int k2 = 0; k2 < K; ++k2) f2[i][k2] = f2n[i][k2];
                                }
                            } else {
#pragma unroll
                                for (int i = 0; i < NB; ++i)
#pragma unroll
                                    for (int j = 0; j < KC; ++j) h[i][j] = fma(gg, f3[i][j], h[i][j]);
                            }
                        } else if (zb + e < sl.z1) {
                            const int2 qq = qv[e];
                            const T g = gv[e];
                            if (qq.x < 0) {  // new q2 group: close the open one, load f[q2,:]
                                if (have) {
#pragma unroll
                                    for (int i = 0; i < NB; ++i)
#pragma unroll
                                        for (int j = 0; j < KC; ++j) {
#pragma unroll
                                            for (int k2 = 0; k2 < K; ++k2)
                                                phi[i][k2][j] = fma(f2[i][k2], h[i][j], phi[i][k2][j]);
                                            h[i][j] = T(0);
                                        }
                                }
                                have = true;
                                const int64_t u2 = (int64_t)(~qq.x) * B;
#pragma unroll
                                for (int i = 0; i < NB; ++i) {
                                    if constexpr (PK)
                                        load_row(~qq.x, i, f2[i]);
                                    else
#pragma unroll
                                        for (int k2 = 0; k2 < K; ++k2) f2[i][k2] = fload<T>(frs, lo[i], u2 + k2 * kstride);
                                }
                            }
                            const int64_t u3 = (int64_t)qq.y * B;
#pragma unroll
                            for (int i = 0; i < NB; ++i) {
                                if constexpr (PK) {
                                    T f3[KC];
                                    load_chunk(qq.y, c0 / KC, i, f3);
                                    const T gg = to_vgpr(g);
#pragma unroll
                                    for (int j = 0; j < KC; ++j) h[i][j] = fma(gg, f3[j], h[i][j]);
                                } else {
#pragma unroll
                                    for (int j = 0; j < KC; ++j)
                                        if (c0 + j < K) h[i][j] = fma(g, fload<T>(frs, lo[i], u3 + (c0 + j) * kstride), h[i][j]);
                                }
                            }
                        }
                    }
                }
                if (have) {
#pragma unroll
                    for (int i = 0; i < NB; ++i)
#pragma unroll
                        for (int j = 0; j < KC; ++j)
#pragma unroll
                            for (int k2 = 0; k2 < K; ++k2) phi[i][k2][j] = fma(f2[i][k2], h[i][j], phi[i][k2][j]);
                }
            } else if (PF) {
                // one-entry-ahead prefetch of the next nonzero's f rows
                int2 qn = zq_c[sl.z0];
                T gn = zg_c[sl.z0];
                T f2n[NB][K], f3n[NB][KC];
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    const int64_t u2 = (int64_t)qn.x * B, u3 = (int64_t)qn.y * B;
#pragma unroll
                    for (int k = 0; k < K; ++k) f2n[i][k] = fload<T>(frs, lo[i], u2 + k * kstride);
#pragma unroll
                    for (int j = 0; j < KC; ++j)
                        f3n[i][j] = (c0 + j < K) ? fload<T>(frs, lo[i], u3 + (c0 + j) * kstride) : T(0);
                }
                for (int z = sl.z0; z < sl.z1; ++z) {
                    const T g = gn;
                    T gf2[NB][K], f3[NB][KC];
#pragma unroll
                    for (int i = 0; i < NB; ++i) {
#pragma unroll
                        for (int k2 = 0; k2 < K; ++k2) gf2[i][k2] = g * f2n[i][k2];
#pragma unroll
                        for (int j = 0; j < KC; ++j) f3[i][j] = f3n[i][j];
                    }
                    if (z + 1 < sl.z1) {
                        qn = zq_c[z + 1];
                        gn = zg_c[z + 1];
#pragma unroll
                        for (int i = 0; i < NB; ++i) {
                            const int64_t u2 = (int64_t)qn.x * B, u3 = (int64_t)qn.y * B;
#pragma unroll
                            for (int k = 0; k < K; ++k) f2n[i][k] = fload<T>(frs, lo[i], u2 + k * kstride);
#pragma unroll
                            for (int j = 0; j < KC; ++j)
                                f3n[i][j] = (c0 + j < K) ? fload<T>(frs, lo[i], u3 + (c0 + j) * kstride) : T(0);
                        }
                    }
                    if (z == sl.z0) {
#pragma unroll
                        for (int i = 0; i < NB; ++i)
#pragma unroll
                            for (int j = 0; j < KC; ++j)
#pragma unroll
                                for (int k2 = 0; k2 < K; ++k2) phi[i][k2][j] = gf2[i][k2] * f3[i][j];
                    } else {
#pragma unroll
                        for (int i = 0; i < NB; ++i)
#pragma unroll
                            for (int j = 0; j < KC; ++j)
#pragma unroll
                                for (int k2 = 0; k2 < K; ++k2) phi[i][k2][j] = fma(gf2[i][k2], f3[i][j], phi[i][k2][j]);
                    }
                }
            } else {
                for (int zb = sl.z0; zb < sl.z1; zb += ZB) {
                    int2 qv[ZB];
                    T gv[ZB];
#pragma unroll
                    for (int e = 0; e < ZB; ++e) {
                        qv[e] = zq_c[zb + e];
                        gv[e] = zg_c[zb + e];
                    }
#pragma unroll
                    for (int e = 0; e < ZB; ++e) {
                        if (zb + e < sl.z1) {
#pragma unroll
                            for (int i = 0; i < NB; ++i) {
                                const int64_t u2 = (int64_t)qv[e].x * B, u3 = (int64_t)qv[e].y * B;
                                T gf2[K];
#pragma unroll
                                for (int k2 = 0; k2 < K; ++k2) gf2[k2] = gv[e] * fload<T>(frs, lo[i], u2 + k2 * kstride);
                                if (zb + e == sl.z0) {
#pragma unroll
                                    for (int j = 0; j < KC; ++j) {
                                        const T f3 = (c0 + j < K) ? fload<T>(frs, lo[i], u3 + (c0 + j) * kstride) : T(0);
#pragma unroll
                                        for (int k2 = 0; k2 < K; ++k2) phi[i][k2][j] = gf2[k2] * f3;
                                    }
                                } else {
#pragma unroll
                                    for (int j = 0; j < KC; ++j) {
                                        if (c0 + j < K) {
                                            const T f3 = fload<T>(frs, lo[i], u3 + (c0 + j) * kstride);
#pragma unroll
                                            for (int k2 = 0; k2 < K; ++k2) phi[i][k2][j] = fma(gf2[k2], f3, phi[i][k2][j]);
                                        }
                                    }
                                }
                            }
                        }
                    }
                }
            }
            // R.Phi: one k3 plane of R_t (K*K contiguous values) per k3, fetched through the scalar path
            // in batches of RB values that are all in flight together (scalar loads return out of
            // order, so each batch costs one wait), then consumed by RB FMAs per distribution.
            if (precise) {
                constexpr int RB = (K * K + 1) / 2 < 25 ? (K * K + 1) / 2 : 25;  // doubles per batch
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
                            for (int e = 0; e < RB; ++e) {
                                if (e0 + e < K2) {
                                    const int k2 = (e0 + e) / K, k1 = (e0 + e) % K;
#pragma unroll
                                    for (int i = 0; i < NB; ++i) q[i][k1] = fma(rr[e], (double)phi[i][k2][j], q[i][k1]);
                                }
                            }
                        }
                    }
                }
            } else {
                // FP32 per-slice sum a[k1], started by the (k2,k3) = (0,c0) products.
                float a[NB][K];
                if (symm) {
                    // R_t is symmetric in (k2,k3) here, so R_t : Phi = sum_{k2<=k3} R_t[.,k2,k3] Phis[k2,k3]
                    // with Phis = Phi + Phi^T off the diagonal: the rows k2 <= k3 of each k3-plane are
                    // its first (k3+1)K values.
                    if constexpr (KC == K) {
#pragma unroll
                        for (int i = 0; i < NB; ++i)
#pragma unroll
                            for (int k3 = 1; k3 < K; ++k3)
#pragma unroll
                                for (int k2 = 0; k2 < k3; ++k2) phi[i][k2][k3] += phi[i][k3][k2];
                        r_phi_f32<K, KC, NB, true>(Rt32, phi, a, 0);
                    }
                } else {
                    r_phi_f32<K, KC, NB, false>(Rt32, phi, a, c0);
                }
#pragma unroll
                for (int i = 0; i < NB; ++i)
#pragma unroll
                    for (int k1 = 0; k1 < K; ++k1) q[i][k1] += (double)a[i][k1];
            }
        }
    }
    if (half >= 0) {
        // half 1 hands its partial sums to half 0 through LDS; half 0 adds them (q0 + q1, a fixed order)
        extern __shared__ double part_lds[];  // per tile of the workgroup: [NB][K][64]
        double *const part = part_lds + (threadIdx.x / (2 * WAVE)) * (NB * K * WAVE);
        const int lane = threadIdx.x % WAVE;
        if (half == 1) {
#pragma unroll
            for (int i = 0; i < NB; ++i)
#pragma unroll
                for (int k1 = 0; k1 < K; ++k1) part[(i * K + k1) * WAVE + lane] = q[i][k1];
        }
        __syncthreads();
        if (half == 1) return;
#pragma unroll
        for (int i = 0; i < NB; ++i)
#pragma unroll
            for (int k1 = 0; k1 < K; ++k1) q[i][k1] += part[(i * K + k1) * WAVE + lane];
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        if (!act[i]) continue;
#pragma unroll
        for (int k1 = 0; k1 < K; ++k1) {
            double v = q[i][k1];
            if (PROJ && ((row < 4 && k1 == 0) || (row == 0 && k1 == 1))) v = 0.0;
            const gptr<double> o = lane_ptr(QT, ((int64_t)k1 * N_Q + row) * B, lq[i]);
            *o = accumulate ? *o + v : v;
        }
    }
}

// A workgroup holds blockDim.x / 64 waves, each one batch tile of the same output row: the waves run
// the same slice sequence, so R_t and the table entries they fetch through the scalar path are
// shared in their CU's scalar cache instead of being fetched by waves of unrelated rows.
template <typename T, int K, int KC, bool GROUPED, bool PROJ, int NB, bool PF>
__global__ __launch_bounds__(4 * WAVE) void af_kernel(const T *__restrict__ fT, double *__restrict__ QT,
                                                  const Slice *__restrict__ slices, const int32_t *__restrict__ row_ptr,
                                                  const int32_t *__restrict__ row_order, const int2 *__restrict__ zq,
                                                  const T *__restrict__ zg, const double *__restrict__ R,
                                                  const float *__restrict__ R32, int64_t B, int N_Q, int accumulate,
                                                  const int32_t *__restrict__ row_mid, int nx, int gx) {
    // packed rows: the symmetrized FP64 kernels (any chunk) and the FP32 kernels with a full k3 chunk
    constexpr bool PK = GROUPED && (std::is_same<T, double>::value || KC == K);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    // Workgroup index -> (column workgroup xw, output row).  The nx column workgroups are cut into groups
    // of gx; the grid runs group by group, each group's rows heaviest first and, within a row, its column
    // workgroups in order.  Workgroups dispatch in index order, so the waves in flight read the f rows of
    // one group's distributions only, which stay in the XCDs' L2 (the whole batch's rows do not beyond
    // a few thousand distributions).
    const int64_t per = (int64_t)gx * N_Q, g = (int64_t)blockIdx.x / per, rem = (int64_t)blockIdx.x - g * per;
    const int64_t x0 = g * gx;
    const int gw = (int)(nx - x0 < gx ? nx - x0 : gx);
    const int rix = (int)(rem / gw);
    const int64_t xw = x0 + (rem - (int64_t)rix * gw);
    int64_t tile;
    int half = -1;
    if (row_mid) {
        // row-split launch: blockDim.x / 128 batch tiles per workgroup, two waves per tile, each wave half
        // of the row's slices.  Every wave reaches the barrier in af_item (a tile past the batch end runs
        // with all lanes inactive).
        tile = xw * (blockDim.x / (2 * WAVE)) + (w >> 1);
        half = w & 1;
    } else {
        tile = xw * (blockDim.x / WAVE) + w;
        if (tile * NB * WAVE >= B) return;  // wave-uniform; no barrier without a row split
    }
    af_item<T, K, KC, GROUPED, PROJ, NB, PF, PK>(fT, QT, slices, row_ptr, zq, zg, R, R32, B, N_Q, accumulate,
                                                 row_order[rix], tile, row_mid, half);
}

// ---------------------------------------------------------------------------------------------
// Batched uneven sizes: work items (group, output row, batch tile) of several geometry groups, one
// kernel per radial size K so that each is register-allocated for its own K (a single kernel that
// switches over K is allocated for the largest case and runs at one wave per SIMD).  The items of a
// K-class are ordered heaviest first by the host's cost model and workgroup i runs item i: the
// hardware dispatcher, which hands out workgroups in index order as slots free up, is the work
// queue.  (A persistent variant whose waves loop over a device counter measured slower: the loop
// around the item body raises register pressure enough that at K = 14 the compiler serializes the
// per-lane f loads.)  Items own disjoint outputs, so the result does not depend on which wave runs
// which item.
// ---------------------------------------------------------------------------------------------
template <int K, int KC>
__global__ __launch_bounds__(WAVE) void multi_af_kernel(const MultiGroupDev *__restrict__ groups,
                                                        const int4 *__restrict__ items, int i0, int i1) {
    const int it = i0 + (int)blockIdx.x;
    if (it >= i1) return;
    const int4 w = items[it];
    const MultiGroupDev &G = groups[w.x];
    af_item<float, K, KC, true, true, 1, false, KC == K>(G.fT, G.Q, G.slices, G.row_ptr, G.zq, G.zg, G.R, G.R32, G.B, G.N_Q,
                                                G.accumulate, w.y, (int64_t)w.z);
}

bool multi_supports_k(int K) { return K == 6 || K == 8 || K == 10 || K == 12 || K == 14; }

// Items [i0, i1) of one K-class; KC is K or ceil(K/2) (any other value selects ceil(K/2)).
int launch_multi(const MultiGroupDev *groups, const int4 *items, int K, int KC, int i0, int i1, hipStream_t st) {
    if (i1 <= i0) return CFK_OK;
    const dim3 grid((unsigned)(i1 - i0));
#define CFK_MULTI_LAUNCH(KK, KCC) \
    { hipLaunchKernelGGL((multi_af_kernel<KK, KCC>), grid, dim3(WAVE), 0, st, groups, items, i0, i1); }
    switch (K) {
        case 6: if (KC == 6) CFK_MULTI_LAUNCH(6, 6) else CFK_MULTI_LAUNCH(6, 3) break;
        case 8: if (KC == 8) CFK_MULTI_LAUNCH(8, 8) else CFK_MULTI_LAUNCH(8, 4) break;
        case 10: if (KC == 10) CFK_MULTI_LAUNCH(10, 10) else CFK_MULTI_LAUNCH(10, 5) break;
        case 12: if (KC == 12) CFK_MULTI_LAUNCH(12, 12) else CFK_MULTI_LAUNCH(12, 6) break;
        case 14: if (KC == 14) CFK_MULTI_LAUNCH(14, 14) else CFK_MULTI_LAUNCH(14, 7) break;
        default: return set_error(CFK_EUNSUP, "multi kernel: K=%d", K);
    }
#undef CFK_MULTI_LAUNCH
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

// ---------------------------------------------------------------------------------------------
// Naive (naive_collision_kernel_mex.cpp:23-46): K^3 per nonzero, loop order k2 -> k3 -> k1,
// Q[q1,k1] += g * R[k1,k2,k3,t] * (f2 f3).
// Radial-first (radial_first_collision_kernel_mex.cpp:41-72): Psi[k1] = sum R f2 f3 (k3 outer,
// k2 inner), then Q[q1,k1] += g * Psi[k1].
// Both use the same row ownership as the angular-first kernel.
// ---------------------------------------------------------------------------------------------
template <int K, bool RADIAL_FIRST>
__global__ __launch_bounds__(WAVE) void k3_fp64_kernel(const double *__restrict__ fT, double *__restrict__ QT,
                                                       const Slice *__restrict__ slices,
                                                       const int32_t *__restrict__ row_ptr,
                                                       const int32_t *__restrict__ row_order,
                                                       const int2 *__restrict__ zq, const double *__restrict__ zg,
                                                       const double *__restrict__ R, int64_t B, int N_Q,
                                                       int accumulate) {
    constexpr int K2 = K * K, K3 = K2 * K;
    const int64_t b = (int64_t)blockIdx.x * WAVE + threadIdx.x;
    const bool active = b < B;
    const int64_t bb = active ? b : 0;
    const int row = row_order[blockIdx.y];
    const int s0 = row_ptr[row], s1 = row_ptr[row + 1];
    const int64_t kstride = (int64_t)N_Q * B;  // [K][N_Q][B]
    double q[K];
#pragma unroll
    for (int k = 0; k < K; ++k) q[k] = 0.0;
    for (int s = s0; s < s1; ++s) {
        const Slice sl = slices[s];
        const double *Rt = R + (int64_t)sl.t * K3;
        for (int z = sl.z0; z < sl.z1; ++z) {
            const int2 qq = zq[z];
            const double g = zg[z];
            double f2[K], f3[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                f2[k] = fT[(int64_t)qq.x * B + k * kstride + bb];
                f3[k] = fT[(int64_t)qq.y * B + k * kstride + bb];
            }
            if (RADIAL_FIRST) {
#pragma unroll
                for (int k1 = 0; k1 < K; ++k1) {
                    double psi = 0.0;
#pragma unroll
                    for (int k3 = 0; k3 < K; ++k3)
#pragma unroll
                        for (int k2 = 0; k2 < K; ++k2) psi = fma(Rt[k1 + k2 * K + k3 * K2] * f2[k2], f3[k3], psi);
                    q[k1] = fma(g, psi, q[k1]);
                }
            } else {
#pragma unroll
                for (int k2 = 0; k2 < K; ++k2)
#pragma unroll
                    for (int k3 = 0; k3 < K; ++k3) {
                        const double f2f3 = f2[k2] * f3[k3];
#pragma unroll
                        for (int k1 = 0; k1 < K; ++k1) q[k1] = fma(g * Rt[k1 + k2 * K + k3 * K2], f2f3, q[k1]);
                    }
            }
        }
    }
    if (active) {
#pragma unroll
        for (int k1 = 0; k1 < K; ++k1) {
            double *o = QT + ((int64_t)k1 * N_Q + row) * B + b;
            *o = accumulate ? *o + q[k1] : q[k1];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Layout changes between [B][NK] (natural) and [NK][B] (batch fastest), 32x32 LDS tiles.
// ---------------------------------------------------------------------------------------------
template <typename TO, bool ADD>
__global__ __launch_bounds__(256) void transpose_kernel(const double *__restrict__ in, TO *__restrict__ out,
                                                         int64_t rows, int64_t cols) {
    // in: [rows][cols]; out: [cols][rows] (out += in^T when ADD); converts to TO on the way
    __shared__ double tile[32][33];
    const int64_t c0 = (int64_t)blockIdx.x * 32, r0 = (int64_t)blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int i = ty; i < 32; i += 8) {
        const int64_t r = r0 + i, c = c0 + tx;
        tile[i][tx] = (r < rows && c < cols) ? in[r * cols + c] : 0.0;
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int64_t c = c0 + i, r = r0 + tx;
        if (r < rows && c < cols) {
            TO *o = out + c * rows + r;
            *o = ADD ? (TO)(*o + tile[tx][i]) : (TO)tile[tx][i];
        }
    }
}

__global__ __launch_bounds__(256) void to_f32_kernel(const double *__restrict__ in, float *__restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

// FP64 f -> packed FP32 rows: for k3 chunks of width KC (NCH chunks, KCP = KC rounded up to 4), chunk c
// of row (q, b) is KCP floats at ((q NCH + c) B + b) KCP (coefficients c KC .. c KC + KC - 1, zero
// padding), so a wave's load of one chunk of one row covers a contiguous 64 KCP floats.  One thread
// per (q, distribution) row.  NATURAL: input [B][K][N_Q] (B MATLAB N_Q x K arrays), thread index q
// fastest (coalesced reads); otherwise batch-fastest [K][N_Q][B], thread index b fastest.
template <bool NATURAL>
__global__ __launch_bounds__(256) void pack_rows_kernel(const double *__restrict__ in, float *__restrict__ out,
                                                        int N_Q, int K, int KC, int64_t B) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)N_Q * B) return;
    int64_t q, b;
    if (NATURAL) {
        b = i / N_Q;
        q = i - b * N_Q;
    } else {
        q = i / B;
        b = i - q * B;
    }
    const int KCP = kcp_of(KC), NCH = (K + KC - 1) / KC;
    for (int c = 0; c < NCH; ++c) {
        float4 *o = (float4 *)(out + ((q * NCH + c) * B + b) * KCP);
        for (int v4 = 0; v4 < KCP / 4; ++v4) {
            float v[4];
            for (int r = 0; r < 4; ++r) {
                const int j = 4 * v4 + r, k = c * KC + j;
                v[r] = (j < KC && k < K) ? (float)(NATURAL ? in[(b * K + k) * N_Q + q] : in[((int64_t)k * N_Q + q) * B + b])
                                         : 0.0f;
            }
            o[v4] = make_float4(v[0], v[1], v[2], v[3]);
        }
    }
}

// Batch-fastest input with B even (and 16-byte aligned): one thread per (q, pair of distributions), the
// K 16-byte loads of the two distributions' coefficients issued together.  Each chunk of the
// workgroup's 512 rows goes through LDS, so that consecutive threads store consecutive 16-byte pieces of
// the packed rows (a thread writing its own two rows would store at a 2 KCP-float stride).
template <int K, int KC>
__global__ __launch_bounds__(256) void pack_rows_pair_kernel(const double *__restrict__ in, float *__restrict__ out,
                                                             int N_Q, int64_t B) {
    constexpr int KCP = kcp_of(KC), NCH = (K + KC - 1) / KC, V4 = KCP / 4;
    __shared__ float4 st[512 * V4];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t half = B / 2, n = (int64_t)N_Q * half;
    const bool act = i < n;
    double2 v[K];
    if (act) {
        const int64_t q = i / half, b = 2 * (i - q * half);
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = *reinterpret_cast<const double2 *>(in + ((int64_t)k * N_Q + q) * B + b);
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        if (act) {
#pragma unroll
            for (int v4 = 0; v4 < V4; ++v4) {
                float w0[4], w1[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int j = 4 * v4 + e, k = c * KC + j;
                    w0[e] = (j < KC && k < K) ? (float)v[j < KC && k < K ? k : 0].x : 0.0f;
                    w1[e] = (j < KC && k < K) ? (float)v[j < KC && k < K ? k : 0].y : 0.0f;
                }
                st[(2 * threadIdx.x) * V4 + v4] = make_float4(w0[0], w0[1], w0[2], w0[3]);
                st[(2 * threadIdx.x + 1) * V4 + v4] = make_float4(w1[0], w1[1], w1[2], w1[3]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 2 * V4; ++t) {
            const int o = (int)threadIdx.x + 256 * t;  // 16-byte piece o of the workgroup's 512 rows
            const int r = o / V4, v4 = o - r * V4;
            const int64_t ir = (int64_t)blockIdx.x * 256 + r / 2;
            if (ir < n) {
                const int64_t q = ir / half, b = 2 * (ir - q * half) + (r & 1);
                reinterpret_cast<float4 *>(out + ((q * NCH + c) * B + b) * KCP)[v4] = st[o];
            }
        }
        __syncthreads();
    }
}

// FP64 f -> packed FP64 rows for the symmetrized FP64 kernels: chunk c of row (q, b) is KCQ = KC rounded up
// to 2 doubles at element ((q NCH + c) B + b) KCQ (zero padded).  One thread per (q, distribution) row,
// 16-byte stores; NATURAL as in pack_rows_kernel.
template <bool NATURAL>
__global__ __launch_bounds__(256) void pack_rows_f64_kernel(const double *__restrict__ in, double *__restrict__ out,
                                                            int N_Q, int K, int KC, int64_t B) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)N_Q * B) return;
    int64_t q, b;
    if (NATURAL) {
        b = i / N_Q;
        q = i - b * N_Q;
    } else {
        q = i / B;
        b = i - q * B;
    }
    const int KCQ = (KC + 1) / 2 * 2, NCH = (K + KC - 1) / KC;
    for (int c = 0; c < NCH; ++c) {
        double2 *o = reinterpret_cast<double2 *>(out + ((q * NCH + c) * B + b) * KCQ);
        for (int v2 = 0; v2 < KCQ / 2; ++v2) {
            double v[2];
            for (int r = 0; r < 2; ++r) {
                const int j = 2 * v2 + r, k = c * KC + j;
                v[r] = (j < KC && k < K) ? (NATURAL ? in[(b * K + k) * N_Q + q] : in[((int64_t)k * N_Q + q) * B + b]) : 0.0;
            }
            o[v2] = make_double2(v[0], v[1]);
        }
    }
}

// Batch-fastest input with B even (and 16-byte aligned), FP64 rows: as pack_rows_pair_kernel, each
// chunk of the workgroup's 512 rows through LDS for stores of consecutive 16-byte pieces.
template <int K, int KC>
__global__ __launch_bounds__(256) void pack_rows_f64_pair_kernel(const double *__restrict__ in, double *__restrict__ out,
                                                                 int N_Q, int64_t B) {
    constexpr int KCQ = (KC + 1) / 2 * 2, NCH = (K + KC - 1) / KC, V2 = KCQ / 2;
    __shared__ double2 st[512 * V2];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t half = B / 2, n = (int64_t)N_Q * half;
    const bool act = i < n;
    double2 v[K];
    if (act) {
        const int64_t q = i / half, b = 2 * (i - q * half);
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = *reinterpret_cast<const double2 *>(in + ((int64_t)k * N_Q + q) * B + b);
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        if (act) {
#pragma unroll
            for (int v2 = 0; v2 < V2; ++v2) {
                double w0[2], w1[2];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int j = 2 * v2 + e, k = c * KC + j;
                    w0[e] = (j < KC && k < K) ? v[j < KC && k < K ? k : 0].x : 0.0;
                    w1[e] = (j < KC && k < K) ? v[j < KC && k < K ? k : 0].y : 0.0;
                }
                st[(2 * threadIdx.x) * V2 + v2] = make_double2(w0[0], w0[1]);
                st[(2 * threadIdx.x + 1) * V2 + v2] = make_double2(w1[0], w1[1]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 2 * V2; ++t) {
            const int o = (int)threadIdx.x + 256 * t;
            const int r = o / V2, v2 = o - r * V2;
            const int64_t ir = (int64_t)blockIdx.x * 256 + r / 2;
            if (ir < n) {
                const int64_t q = ir / half, b = 2 * (ir - q * half) + (r & 1);
                reinterpret_cast<double2 *>(out + ((q * NCH + c) * B + b) * KCQ)[v2] = st[o];
            }
        }
        __syncthreads();
    }
}

__global__ void conservation_kernel(double *Q, int64_t B, int N_Q, int K, int layout) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int slots_q[5] = {0, 1, 2, 3, 0};
    const int slots_k[5] = {0, 0, 0, 0, 1};
    for (int s = 0; s < 5; ++s) {
        if (slots_q[s] >= N_Q || slots_k[s] >= K) continue;
        const int64_t c = (int64_t)slots_k[s] * N_Q + slots_q[s];
        if (layout == CFK_LAYOUT_NATURAL)
            Q[b * (int64_t)N_Q * K + c] = 0.0;
        else
            Q[c * B + b] = 0.0;
    }
}

// ---------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------
namespace {

// Distributions per column group of a grouped-table launch (symmetrized FP64 and mixed).  With the whole
// batch as one group the waves in flight read the f rows of every distribution: at 16384 distributions
// the symmetrized FP64 kernel took 0.87 ms per 4096 against 0.57 ms at 4096.  Groups of 1024 (8 column
// workgroups of two batch tiles, one per XCD) measured best at L=8, K=10: mixed 0.257 ms and symmetrized
// FP64 0.52 ms per 4096 at 4096 distributions, 0.23 / 0.47-0.48 ms per 4096 at 65536 (groups of 4096:
// 0.266 / 0.57; 2048: 0.265 / 0.54; 512: 0.266 / 0.52).  The plain FP64 kernels run in groups of 4096
// (at 4096 distributions 1.39-1.44 ms against 1.45-1.52 ms in groups of 1024; at 16384 distributions in
// one group 1.78 ms per 4096 against 1.29 ms in groups of 4096).
#ifndef AF_GROUP_COLS
#define AF_GROUP_COLS 1024
#endif
#ifndef AF_GROUP_COLS_PLAIN
#define AF_GROUP_COLS_PLAIN 4096
#endif
template <typename T, int K, int KC, bool GROUPED, int NB, bool PF>
void launch_af_v(const SparseTables<T> &tb, int N_Q, int wpb, bool split, const T *fT, double *QT, int64_t B,
                 int acc, hipStream_t st) {
    wpb = wpb == 2 || wpb == 4 ? wpb : 1;
    const int64_t tiles = (B + WAVE * NB - 1) / (WAVE * NB);
    // double runs the plain operator (no projection), float the mixed path (always projected)
    constexpr bool PROJ = std::is_same<T, float>::value;
    // column groups of whole workgroups: AF_GROUP_COLS distributions for the grouped tables,
    // AF_GROUP_COLS_PLAIN for the plain FP64 kernels
    auto group_of = [&](int tpw) {
        return (int)std::max<int64_t>(1, (GROUPED ? AF_GROUP_COLS : AF_GROUP_COLS_PLAIN) / ((int64_t)WAVE * NB * tpw));
    };
    if (split && tb.row_mid) {
        const int tpw = wpb >= 2 ? 2 : 1;  // batch tiles per workgroup (two waves each, at most 4 waves)
        const int nx = (int)((tiles + tpw - 1) / tpw);
        const dim3 grid((unsigned)((int64_t)nx * N_Q));
        hipLaunchKernelGGL((af_kernel<T, K, KC, GROUPED, PROJ, NB, PF>), grid, dim3(2 * WAVE * tpw),
                           (size_t)tpw * NB * K * WAVE * sizeof(double), st, fT, QT, tb.slices, tb.row_ptr,
                           tb.row_order, tb.zq, tb.zg, tb.R, tb.R32, B, N_Q, acc, tb.row_mid, nx, group_of(tpw));
        return;
    }
    const int nx = (int)((tiles + wpb - 1) / wpb);
    const dim3 grid((unsigned)((int64_t)nx * N_Q));
    hipLaunchKernelGGL((af_kernel<T, K, KC, GROUPED, PROJ, NB, PF>), grid, dim3(WAVE * wpb), 0, st, fT, QT, tb.slices,
                       tb.row_ptr, tb.row_order, tb.zq, tb.zg, tb.R, tb.R32, B, N_Q, acc, (const int32_t *)nullptr, nx,
                       group_of(wpb));
}

// k3 chunk candidates: the full width and half of it.
constexpr int kc_half(int K) { return (K + 1) / 2; }
// The partial chunk of a tile: half the width, for FP64 from K = 12 a third (at K = 14 half a chunk of
// FP64 Phi alone takes 196 VGPRs and leaves one wave per SIMD).
template <typename T>
constexpr int kc_part(int K) { return (std::is_same<T, double>::value && K >= 12) ? (K + 2) / 3 : kc_half(K); }
// Default chunk: full width while the Phi accumulators stay within ~100 values.
constexpr int kc_default(int K) { return (K * K <= 100) ? K : kc_half(K); }

// K values that get the full set of tile variants (the workloads' sizes); other K get the default
// tile only, which keeps the build time bounded.
#ifndef CFK_TUNED_K
#define CFK_TUNED_K(K) ((K) == 6 || (K) == 8 || (K) == 10 || (K) == 12 || (K) == 14)
#endif

template <typename T, int K, bool GROUPED>
void launch_af_k(const SparseTables<T> &tb, int N_Q, const Tile &tl, const T *fT, double *QT, int64_t B, int acc,
                 hipStream_t st) {
    constexpr int KCD = kc_default(K);
    if constexpr (CFK_TUNED_K(K)) {
        const bool full = tl.kc >= K;
        const bool pf = tl.pf && !GROUPED;
#define CFK_AF_NB(KCV, PFV)                                                             \
    do {                                                                                \
        if (tl.nb >= 2)                                                                 \
            launch_af_v<T, K, KCV, GROUPED, 2, PFV>(tb, N_Q, tl.wpb, tl.split, fT, QT, B, acc, st);       \
        else                                                                            \
            launch_af_v<T, K, KCV, GROUPED, 1, PFV>(tb, N_Q, tl.wpb, tl.split, fT, QT, B, acc, st);       \
    } while (0)
        if constexpr (GROUPED) {
            (void)pf;
            if (full) CFK_AF_NB(K, false); else CFK_AF_NB(kc_part<T>(K), false);
        } else {
            if (full) {
                if (pf) CFK_AF_NB(K, true); else CFK_AF_NB(K, false);
            } else {
                if (pf) CFK_AF_NB(kc_part<T>(K), true); else CFK_AF_NB(kc_part<T>(K), false);
            }
        }
#undef CFK_AF_NB
    } else {
        launch_af_v<T, K, KCD, GROUPED, 1, false>(tb, N_Q, tl.wpb, tl.split, fT, QT, B, acc, st);
    }
}

template <int K>
void launch_k3(const cfk_plan_s *p, bool rf, const double *fT, double *QT, int64_t B, int acc, hipStream_t st) {
    dim3 grid((unsigned)((B + WAVE - 1) / WAVE), (unsigned)p->N_Q);
    if (rf)
        hipLaunchKernelGGL((k3_fp64_kernel<K, true>), grid, dim3(WAVE), 0, st, fT, QT, p->slices.p, p->row_ptr.p,
                           p->row_order.p, p->zq.p, p->zg.p, p->R.p, B, p->N_Q, acc);
    else
        hipLaunchKernelGGL((k3_fp64_kernel<K, false>), grid, dim3(WAVE), 0, st, fT, QT, p->slices.p, p->row_ptr.p,
                           p->row_order.p, p->zq.p, p->zg.p, p->R.p, B, p->N_Q, acc);
}

}  // namespace

int default_kc(int K) { return kc_default(K); }
int half_kc(int K) { return kc_half(K); }

int launch_sparse_fp64(const cfk_plan_s *p, int variant, const Tile &tl, const double *fT, double *QT, int64_t B,
                       int accumulate, hipStream_t st) {
    if (p->K < 1 || p->K > 16) return set_error(CFK_EUNSUP, "K=%d unsupported (1..16)", p->K);
    if (B <= 0) return CFK_OK;
    if ((B + WAVE - 1) / WAVE > 0x7fffffffLL) return set_error(CFK_EINVAL, "batch too large");
    const SparseTables<double> tb{p->slices.p, p->row_ptr.p, p->row_order.p, p->zq.p, p->zg.p, p->R.p, nullptr};
    switch (p->K) {
#define X(KK)                                                                                  \
    case KK:                                                                                   \
        if (variant == CFK_VARIANT_ANGULAR_FIRST)                                              \
            launch_af_k<double, KK, false>(tb, p->N_Q, tl, fT, QT, B, accumulate, st);         \
        else                                                                                   \
            launch_k3<KK>(p, variant == CFK_VARIANT_RADIAL_FIRST, fT, QT, B, accumulate, st);  \
        break;
        CFK_K_CASES(X)
#undef X
    }
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

// Angular-first over the symmetrized tables (q2 <= q3 entries only, merged R blocks).
template <typename T>
int launch_af_sym(const cfk_plan_s *p, const SparseTables<T> &tb, const Tile &tl, const T *fT, double *QT,
                  int64_t B, int accumulate, hipStream_t st) {
    if (p->K < 1 || p->K > 16) return set_error(CFK_EUNSUP, "K=%d unsupported (1..16)", p->K);
    if (B <= 0) return CFK_OK;
    switch (p->K) {
#define X(KK)                                                                   \
    case KK:                                                                    \
        launch_af_k<T, KK, true>(tb, p->N_Q, tl, fT, QT, B, accumulate, st);    \
        break;
        CFK_K_CASES(X)
#undef X
    }
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

int launch_af_sym_f64(const cfk_plan_s *p, const Tile &tl, const double *fT, double *QT, int64_t B, int accumulate,
                      hipStream_t st) {
    const SparseTables<double> tb{p->sym_slices.p, p->sym_row_ptr.p, p->sym_row_order.p, p->sym_zq.p, p->sym_zg.p,
                                  p->sym_R.p, nullptr, p->sym_row_mid.p};
    return launch_af_sym<double>(p, tb, tl, fT, QT, B, accumulate, st);
}

int launch_af_sym_f32(const cfk_plan_s *p, const Tile &tl, const float *fT, double *QT, int64_t B, int accumulate,
                      hipStream_t st) {
    const SparseTables<float> tb{p->sym_slices.p, p->sym_row_ptr.p, p->sym_row_order.p, p->sym_zq.p, p->sym_zg32.p,
                                 p->sym_R.p, p->sym_R32.p, p->sym_row_mid.p};
    return launch_af_sym<float>(p, tb, tl, fT, QT, B, accumulate, st);
}

int launch_transpose(const double *in, double *out, int64_t rows, int64_t cols, bool add, hipStream_t st) {
    if (rows <= 0 || cols <= 0) return CFK_OK;
    dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
    if ((rows + 31) / 32 > 65535) return set_error(CFK_EINVAL, "transpose: too many rows (%lld)", (long long)rows);
    if (add)
        hipLaunchKernelGGL((transpose_kernel<double, true>), grid, dim3(256), 0, st, in, out, rows, cols);
    else
        hipLaunchKernelGGL((transpose_kernel<double, false>), grid, dim3(256), 0, st, in, out, rows, cols);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

int launch_transpose_f32(const double *in, float *out, int64_t rows, int64_t cols, hipStream_t st) {
    if (rows <= 0 || cols <= 0) return CFK_OK;
    dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
    if ((rows + 31) / 32 > 65535) return set_error(CFK_EINVAL, "transpose: too many rows (%lld)", (long long)rows);
    hipLaunchKernelGGL((transpose_kernel<float, false>), grid, dim3(256), 0, st, in, out, rows, cols);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

int launch_pack_rows(const double *in, float *out, int N_Q, int K, int KC, int64_t B, bool natural, hipStream_t st) {
    const int64_t n = (int64_t)N_Q * B;
    if (n <= 0) return CFK_OK;
    if (KC < 1 || KC > K || rl_of(K, KC) > 16) return set_error(CFK_EUNSUP, "packed rows: K=%d KC=%d", K, KC);
    const dim3 grid((unsigned)((n + 255) / 256));
    if (natural)
        hipLaunchKernelGGL(pack_rows_kernel<true>, grid, dim3(256), 0, st, in, out, N_Q, K, KC, B);
    else if (B % 2 == 0 && reinterpret_cast<uintptr_t>(in) % 16 == 0 && (KC == K || KC == (K + 1) / 2)) {
        const dim3 g2((unsigned)((n / 2 + 255) / 256));
        switch (K * 2 + (KC == K ? 0 : 1)) {
#define CFK_PAIR(KK)                                                                                         \
    case 2 * KK:                                                                                            \
        hipLaunchKernelGGL((pack_rows_pair_kernel<KK, KK>), g2, dim3(256), 0, st, in, out, N_Q, B);         \
        break;                                                                                              \
    case 2 * KK + 1:                                                                                        \
        hipLaunchKernelGGL((pack_rows_pair_kernel<KK, (KK + 1) / 2>), g2, dim3(256), 0, st, in, out, N_Q, B); \
        break;
            CFK_PAIR(1) CFK_PAIR(2) CFK_PAIR(3) CFK_PAIR(4) CFK_PAIR(5) CFK_PAIR(6) CFK_PAIR(7) CFK_PAIR(8)
            CFK_PAIR(9) CFK_PAIR(10) CFK_PAIR(11) CFK_PAIR(12) CFK_PAIR(13) CFK_PAIR(14) CFK_PAIR(15) CFK_PAIR(16)
#undef CFK_PAIR
        }
    } else
        hipLaunchKernelGGL(pack_rows_kernel<false>, grid, dim3(256), 0, st, in, out, N_Q, K, KC, B);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

// Whether the FP32 angular-first launch for this tile reads packed rows (the kernel with a full k3
// chunk); mirrors the chunk selection of launch_af_k.
bool af_f32_packed(int K, const Tile &tl) {
    const int kc = CFK_TUNED_K(K) ? (tl.kc >= K ? K : kc_half(K)) : kc_default(K);
    return kc == K;
}

int launch_pack_rows_f64(const double *in, double *out, int N_Q, int K, int KC, int64_t B, bool natural,
                         hipStream_t st) {
    const int64_t n = (int64_t)N_Q * B;
    if (n <= 0) return CFK_OK;
    if (KC < 1 || KC > K) return set_error(CFK_EUNSUP, "packed FP64 rows: K=%d KC=%d", K, KC);
    const dim3 grid((unsigned)((n + 255) / 256));
    const int var = KC == K ? 0 : KC == (K + 1) / 2 ? 1 : KC == (K + 2) / 3 ? 2 : -1;
    if (!natural && B % 2 == 0 && reinterpret_cast<uintptr_t>(in) % 16 == 0 && var >= 0 && K <= 16) {
        const dim3 g2((unsigned)((n / 2 + 255) / 256));
        switch (K * 4 + var) {
#define CFK_PAIR64(KK)                                                                                             \
    case 4 * KK:                                                                                                  \
        hipLaunchKernelGGL((pack_rows_f64_pair_kernel<KK, KK>), g2, dim3(256), 0, st, in, out, N_Q, B);           \
        break;                                                                                                    \
    case 4 * KK + 1:                                                                                              \
        hipLaunchKernelGGL((pack_rows_f64_pair_kernel<KK, (KK + 1) / 2>), g2, dim3(256), 0, st, in, out, N_Q, B); \
        break;                                                                                                    \
    case 4 * KK + 2:                                                                                              \
        hipLaunchKernelGGL((pack_rows_f64_pair_kernel<KK, (KK + 2) / 3>), g2, dim3(256), 0, st, in, out, N_Q, B); \
        break;
            CFK_PAIR64(1) CFK_PAIR64(2) CFK_PAIR64(3) CFK_PAIR64(4) CFK_PAIR64(5) CFK_PAIR64(6) CFK_PAIR64(7)
            CFK_PAIR64(8) CFK_PAIR64(9) CFK_PAIR64(10) CFK_PAIR64(11) CFK_PAIR64(12) CFK_PAIR64(13) CFK_PAIR64(14)
            CFK_PAIR64(15) CFK_PAIR64(16)
#undef CFK_PAIR64
        }
    } else if (natural) {
        hipLaunchKernelGGL(pack_rows_f64_kernel<true>, grid, dim3(256), 0, st, in, out, N_Q, K, KC, B);
    } else {
        hipLaunchKernelGGL(pack_rows_f64_kernel<false>, grid, dim3(256), 0, st, in, out, N_Q, K, KC, B);
    }
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

// The k3 chunk the symmetrized FP64 launch for this tile runs with; mirrors the selection of launch_af_k.
int af_f64_sym_kc(int K, const Tile &tl) {
    if (!CFK_TUNED_K(K)) return kc_default(K);
    return tl.kc >= K ? K : kc_part<double>(K);
}

int launch_to_f32(const double *in, float *out, int64_t n, hipStream_t st) {
    if (n <= 0) return CFK_OK;
    hipLaunchKernelGGL(to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, out, n);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}


int launch_conservation(double *Q, int64_t B, int N_Q, int K, int layout, hipStream_t st) {
    if (B <= 0) return CFK_OK;
    hipLaunchKernelGGL(conservation_kernel, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, st, Q, B, N_Q, K,
                       layout);
    CFK_HIP(hipGetLastError());
    return CFK_OK;
}

}  // namespace cfk

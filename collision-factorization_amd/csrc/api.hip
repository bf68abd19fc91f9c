// C ABI of libcfk: argument validation, plan construction, MEX-equivalent one-shot entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <array>
#include <map>
#include <numeric>
#include <unordered_map>

#include "common.hpp"
#include "prep.hpp"

namespace cfk {

namespace {
thread_local std::string g_err;
}

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int check_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return set_error(CFK_ENODEV, "no usable HIP device (%s)", e != hipSuccess ? hipGetErrorString(e) : "0 devices");
    return CFK_OK;
}

int launch_sparse_fp64(const cfk_plan_s *p, int variant, const Tile &tl, const double *fT, double *QT, int64_t B,
                       int accumulate, hipStream_t st);
int launch_transpose(const double *in, double *out, int64_t rows, int64_t cols, bool add, hipStream_t st);
int launch_dense(double *Q, const double *f, const double *C, int64_t N, int64_t B, double *part, hipStream_t st);
int dense_part_doubles(int64_t N, int64_t *n);
int launch_dense_assemble(double *C, const int4 *zl, const double *zg, const double *R, int64_t N_G, int N_Q, int K,
                          hipStream_t st);
int launch_conservation(double *Q, int64_t B, int N_Q, int K, int layout, hipStream_t st);
int default_kc(int K);
int table_padding();
int half_kc(int K);
int launch_af_sym_f64(const cfk_plan_s *p, const Tile &tl, const double *fT, double *QT, int64_t B, int accumulate,
                      hipStream_t st);
int launch_af_sym_f32(const cfk_plan_s *p, const Tile &tl, const float *fT, double *QT, int64_t B, int accumulate,
                      hipStream_t st);
int launch_transpose_f32(const double *in, float *out, int64_t rows, int64_t cols, hipStream_t st);
int launch_to_f32(const double *in, float *out, int64_t n, hipStream_t st);
int launch_pack_rows(const double *in, float *out, int N_Q, int K, int KC, int64_t B, bool natural, hipStream_t st);
int launch_pack_rows_f64(const double *in, double *out, int N_Q, int K, int KC, int64_t B, bool natural,
                         hipStream_t st);
int af_f64_sym_kc(int K, const Tile &tl);
bool af_f32_packed(int K, const Tile &tl);
bool mx_supports_k(int K);
int mx_chunk(int K);
int build_mx_tables(cfk_plan_s *p, std::vector<float> &out, int64_t &n_groups);
int launch_af_mx(const cfk_plan_s *p, const float *Rmx, const Tile &tl, const float *fT, double *QT, int64_t B,
                 int accumulate, hipStream_t st);

// A operands of the matrix-core kernel, built on first use.
int ensure_mx_tables(cfk_plan_s *p) {
    if (p->sym_Rmx.p) return CFK_OK;
    std::vector<float> t;
    int64_t ng = 0;
    int rc = build_mx_tables(p, t, ng);
    if (rc) return rc;
    return p->sym_Rmx.upload(t.data(), t.size());
}

// 1-based integral index stored as double, checked against [1, hi].
static inline bool idx_ok(double v, int64_t hi) { return v >= 1.0 && v <= (double)hi && v == std::floor(v); }

// Symmetrized tables: keep entries with q2 <= q3 (g halved on the diagonal q2 == q3), merge each
// kept block t=(l1,l2,l3) with its mirror t'=(l1,l3,l2) into R'_t[k1,k2,k3] = R_t[k1,k2,k3] +
// R_t'[k1,k3,k2].  Valid only if every entry's mirror (q1,q3,q2) exists with the same value (the
// real Gaunt tensor is symmetric in its last two indices); otherwise the symmetrized paths are
// disabled for this plan and report why.
static int build_sym_tables(const double *labels, const double *vals, const double *ic_map, int64_t N_G,
                            const double *R, const std::vector<int32_t> &perm, cfk_plan_s *p) {
    const int N_Q = p->N_Q, K = p->K, n_l = p->n_l;
    const int64_t K3 = (int64_t)K * K * K;
    p->sym_ok = false;
    std::vector<std::array<int, 3>> trip(n_l, {-1, -1, -1});
    std::unordered_map<int64_t, double> gmap;
    gmap.reserve((size_t)N_G * 2);
    auto key = [&](int a, int b, int c) { return ((int64_t)a * N_Q + b) * N_Q + c; };
    for (int64_t z = 0; z < N_G; ++z) {
        const int q1 = (int)labels[z] - 1, q2 = (int)labels[z + N_G] - 1, q3 = (int)labels[z + 2 * N_G] - 1;
        const int t = (int)ic_map[z] - 1;
        const std::array<int, 3> lt = {l_of_q0(q1), l_of_q0(q2), l_of_q0(q3)};
        if (trip[t][0] < 0)
            trip[t] = lt;
        else if (trip[t] != lt) {
            p->sym_why = "ic_map groups entries of different (l1,l2,l3)";
            return CFK_OK;
        }
        gmap[key(q1, q2, q3)] = vals[z];
    }
    double gmax = 0.0;
    for (int64_t z = 0; z < N_G; ++z) gmax = std::max(gmax, std::fabs(vals[z]));
    for (int64_t z = 0; z < N_G; ++z) {
        const int q1 = (int)labels[z] - 1, q2 = (int)labels[z + N_G] - 1, q3 = (int)labels[z + 2 * N_G] - 1;
        auto it = gmap.find(key(q1, q3, q2));
        if (it == gmap.end() || std::fabs(it->second - vals[z]) > 1e-12 * gmax) {
            p->sym_why = "Gaunt COO is not symmetric in (q2,q3)";
            return CFK_OK;
        }
    }
    std::map<std::array<int, 3>, int> tof;
    for (int t = 0; t < n_l; ++t)
        if (trip[t][0] >= 0) tof[trip[t]] = t;
    std::vector<double> Rs((size_t)n_l * K3, 0.0);
    for (int t = 0; t < n_l; ++t) {
        if (trip[t][0] < 0 || trip[t][1] > trip[t][2]) continue;
        auto it = tof.find({trip[t][0], trip[t][2], trip[t][1]});
        if (it == tof.end()) {
            p->sym_why = "mirror l-triplet missing";
            return CFK_OK;
        }
        const double *Rt = R + (int64_t)t * K3, *Rm = R + (int64_t)it->second * K3;
        for (int k3 = 0; k3 < K; ++k3)
            for (int k2 = 0; k2 < K; ++k2)
                for (int k1 = 0; k1 < K; ++k1)
                    Rs[(size_t)t * K3 + k1 + k2 * K + k3 * K * K] =
                        Rt[k1 + k2 * K + k3 * K * K] + Rm[k1 + k3 * K + k2 * K * K];
    }
    // kept entries in (t,q1) order, then sorted by q2 inside each (t,q1) slice; an entry that
    // starts a new q2 group stores ~q2 (negative) so the kernel reloads f[q2,:] only then.
    std::vector<int32_t> kept;
    for (int64_t i = 0; i < N_G; ++i) {
        const int32_t z = perm[i];
        if (labels[z + N_G] <= labels[z + 2 * N_G]) kept.push_back(z);
    }
    std::vector<int2> zq;
    std::vector<double> zg;
    std::vector<Slice> tq;
    for (size_t i = 0; i < kept.size();) {
        const int32_t z0 = kept[i];
        const int t = (int)ic_map[z0] - 1, q1 = (int)labels[z0] - 1;
        size_t e = i;
        while (e < kept.size() && (int)ic_map[kept[e]] - 1 == t && (int)labels[kept[e]] - 1 == q1) ++e;
        std::stable_sort(kept.begin() + i, kept.begin() + e,
                         [&](int32_t a_, int32_t b_) { return labels[a_ + N_G] < labels[b_ + N_G]; });
        const int32_t j0 = (int32_t)zq.size();
        int prev_q2 = -1;
        for (size_t r = i; r < e; ++r) {
            const int32_t z = kept[r];
            const int q2 = (int)labels[z + N_G] - 1, q3 = (int)labels[z + 2 * N_G] - 1;
            zq.push_back(make_int2(q2 != prev_q2 ? ~q2 : q2, q3));
            zg.push_back(q2 == q3 ? 0.5 * vals[z] : vals[z]);
            prev_q2 = q2;
        }
        tq.push_back(Slice{t, j0, (int32_t)zq.size(), q1});
        i = e;
    }
    std::vector<int32_t> row_ptr(N_Q + 1, 0), order(N_Q);
    for (const Slice &s : tq) row_ptr[s.q1 + 1]++;
    for (int r = 0; r < N_Q; ++r) row_ptr[r + 1] += row_ptr[r];
    std::vector<Slice> sl(tq.size());
    {
        std::vector<int32_t> fill(row_ptr.begin(), row_ptr.end() - 1);
        for (const Slice &s : tq) sl[fill[s.q1]++] = s;
    }
    // within a row the slices whose partner block is l = 0 (one per row) come last: the matrix-core
    // kernel runs their FP64 radial stage after the row's other slices, so its FP64 sums are not held
    // across the row
    for (int r = 0; r < N_Q; ++r)
        std::stable_partition(sl.begin() + row_ptr[r], sl.begin() + row_ptr[r + 1],
                              [&](const Slice &x) { return trip[x.t][1] != 0; });
    std::vector<double> work(N_Q, 0.0);
    for (const Slice &s : tq) work[s.q1] += (double)(s.z1 - s.z0) * K * K + (double)K3;
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
    p->sym_row_work = work;
    // device copy: the last field holds the slice flags (partner block l = 0: kept with l2 <= l3, so l2 == 0;
    // (k2,k3)-symmetric merged block: l2 == l3) and the length of the slice's first q2 group
    for (Slice &s : sl) {
        int32_t g1 = 1;
        while (s.z0 + g1 < s.z1 && zq[s.z0 + g1].x >= 0) ++g1;
        s.q1 = (trip[s.t][1] == 0 ? SLICE_PRECISE : 0) | (trip[s.t][1] == trip[s.t][2] ? SLICE_SYMMETRIC : 0) |
               (g1 << SLICE_G1_SHIFT);
    }
    p->sym_N_G = (int64_t)zq.size();
    p->sym_S = (int64_t)tq.size();
    p->sym_groups = p->sym_symmetric = p->sym_precise = 0;
    for (const int2 &e : zq) p->sym_groups += e.x < 0;
    for (const Slice &s : sl) {
        p->sym_symmetric += (s.q1 & SLICE_SYMMETRIC) != 0;
        p->sym_precise += (s.q1 & SLICE_PRECISE) != 0;
    }
    for (int i = 0; i < table_padding(); ++i) {  // batch-fetch padding (never used as entries)
        zq.push_back(make_int2(0, 0));
        zg.push_back(0.0);
    }
    int rc;
    if ((rc = p->sym_slices.upload(sl.data(), sl.size()))) return rc;
    if ((rc = p->sym_row_ptr.upload(row_ptr.data(), row_ptr.size()))) return rc;
    {
        // split point of each row's slices for row-split launches: the halves' cost-model work balanced
        std::vector<int32_t> mid(N_Q);
        for (int r = 0; r < N_Q; ++r) {
            double tot = 0.0, acc = 0.0;
            for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) tot += (double)(sl[k].z1 - sl[k].z0) * K * K + (double)K3;
            int32_t m = row_ptr[r];
            while (m < row_ptr[r + 1] && acc + 0.5 * ((double)(sl[m].z1 - sl[m].z0) * K * K + (double)K3) < 0.5 * tot) {
                acc += (double)(sl[m].z1 - sl[m].z0) * K * K + (double)K3;
                ++m;
            }
            mid[r] = m;
        }
        if ((rc = p->sym_row_mid.upload(mid.data(), mid.size()))) return rc;
    }
    if ((rc = p->sym_row_order.upload(order.data(), order.size()))) return rc;
    if ((rc = p->sym_zq.upload(zq.data(), zq.size()))) return rc;
    if ((rc = p->sym_zg.upload(zg.data(), zg.size()))) return rc;
    if ((rc = p->sym_R.upload(Rs.data(), Rs.size()))) return rc;
    std::vector<float> g32(zg.begin(), zg.end()), r32(Rs.begin(), Rs.end());
    if ((rc = p->sym_zg32.upload(g32.data(), g32.size()))) return rc;
    if ((rc = p->sym_R32.upload(r32.data(), r32.size()))) return rc;
    p->sym_ok = true;
    return CFK_OK;
}

// Host-side validation of a COO + R (before any device work).
int validate_coo(const double *labels, const double *vals, const double *ic_map, int64_t N_G, const double *R,
                 int64_t R_numel, int32_t N_Q, int32_t K) {
    if (N_Q < 1) return set_error(CFK_EINVAL, "N_Q must be >= 1 (got %d)", N_Q);
    if (K < 1) return set_error(CFK_EINVAL, "K must be >= 1 (got %d)", K);
    if (K > 16) return set_error(CFK_EUNSUP, "K=%d unsupported (1..16)", K);
    if (N_G < 0) return set_error(CFK_EINVAL, "N_G must be >= 0");
    if (N_G > 0 && (!labels || !vals || !ic_map)) return set_error(CFK_EINVAL, "null COO pointer");
    const int64_t K3 = (int64_t)K * K * K;
    if (R_numel < 0 || R_numel % K3 != 0)
        return set_error(CFK_EINVAL, "R has %lld elements, not a multiple of K^3=%lld", (long long)R_numel, (long long)K3);
    if (R_numel > 0 && !R) return set_error(CFK_EINVAL, "null R pointer");
    const int64_t n_l = R_numel / K3;
    if (N_G > INT32_MAX / 2) return set_error(CFK_EINVAL, "N_G too large");
    for (int64_t z = 0; z < N_G; ++z) {
        for (int c = 0; c < 3; ++c)
            if (!idx_ok(labels[z + c * N_G], N_Q))
                return set_error(CFK_EINVAL, "labels(%lld,%d)=%g outside 1..N_Q=%d", (long long)z + 1, c + 1,
                                 labels[z + c * N_G], N_Q);
        if (!idx_ok(ic_map[z], n_l))
            return set_error(CFK_EINVAL, "ic_map(%lld)=%g outside 1..N_L=%lld", (long long)z + 1, ic_map[z],
                             (long long)n_l);
        if (!std::isfinite(vals[z])) return set_error(CFK_EINVAL, "vals(%lld) is not finite", (long long)z + 1);
    }
    return CFK_OK;
}

int build_plan(const double *labels, const double *vals, const double *ic_map, int64_t N_G, const double *R,
               int64_t R_numel, int32_t N_Q, int32_t K, cfk_plan_s *p) {
    int rc;
    if ((rc = validate_coo(labels, vals, ic_map, N_G, R, R_numel, N_Q, K))) return rc;
    const int64_t K3 = (int64_t)K * K * K;
    const int64_t n_l = R_numel / K3;
    p->N_Q = N_Q;
    p->K = K;
    p->n_l = (int)n_l;
    p->N_G = N_G;
    // Stable sort by (t, q1) as sortrows([ic_map, q1]) (benchmark_collision_contraction_scaling.m:82-85)
    std::vector<int32_t> perm(N_G);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) {
        if (ic_map[a] != ic_map[b]) return ic_map[a] < ic_map[b];
        return labels[a] < labels[b];
    });
    std::vector<int2> zq(N_G);
    std::vector<double> zg(N_G);
    std::vector<Slice> tq;  // slices in (t,q1) order
    for (int64_t i = 0; i < N_G; ++i) {
        const int32_t z = perm[i];
        zq[i] = make_int2((int)labels[z + N_G] - 1, (int)labels[z + 2 * N_G] - 1);
        zg[i] = vals[z];
        const int t = (int)ic_map[z] - 1, q1 = (int)labels[z] - 1;
        if (tq.empty() || tq.back().t != t || tq.back().q1 != q1)
            tq.push_back(Slice{t, (int32_t)i, (int32_t)i + 1, q1});
        else
            tq.back().z1 = (int32_t)i + 1;
    }
    p->S = (int64_t)tq.size();
    // group slices by output row q1 (stable: t ascending inside a row)
    p->h_row_ptr.assign(N_Q + 1, 0);
    for (const Slice &s : tq) p->h_row_ptr[s.q1 + 1]++;
    for (int r = 0; r < N_Q; ++r) p->h_row_ptr[r + 1] += p->h_row_ptr[r];
    p->h_slices.resize(tq.size());
    {
        std::vector<int32_t> fill(p->h_row_ptr.begin(), p->h_row_ptr.end() - 1);
        for (const Slice &s : tq) p->h_slices[fill[s.q1]++] = s;
    }
    // schedule heavy rows first
    std::vector<double> work(N_Q, 0.0);
    for (const Slice &s : tq) work[s.q1] += (double)(s.z1 - s.z0) * K * K + (double)K3;
    p->h_row_order.resize(N_Q);
    std::iota(p->h_row_order.begin(), p->h_row_order.end(), 0);
    std::stable_sort(p->h_row_order.begin(), p->h_row_order.end(), [&](int a, int b) { return work[a] > work[b]; });
    if ((rc = p->R.upload(R, (size_t)R_numel))) return rc;
    if ((rc = p->slices.upload(p->h_slices.data(), p->h_slices.size()))) return rc;
    if ((rc = p->row_ptr.upload(p->h_row_ptr.data(), p->h_row_ptr.size()))) return rc;
    if ((rc = p->row_order.upload(p->h_row_order.data(), p->h_row_order.size()))) return rc;
    // padding entries: the kernels fetch table entries in fixed-size batches
    for (int i = 0; i < table_padding(); ++i) {
        zq.push_back(make_int2(0, 0));
        zg.push_back(0.0);
    }
    if ((rc = p->zq.upload(zq.data(), zq.size()))) return rc;
    if ((rc = p->zg.upload(zg.data(), zg.size()))) return rc;
    return build_sym_tables(labels, vals, ic_map, N_G, R, perm, p);
}

// Untuned k3 chunk: the FP64 paths are register-bound and run fastest on half-width chunks from
// K = 8 up (measured at K = 10: 1.43 vs 1.98 ms plain, 0.69 vs 1.65 ms symmetrized, batch 4096); the
// mixed path keeps the full width while Phi fits (0.36 vs 0.42 ms).
static int default_tile_kc(int precision, int K) {
    return (precision != CFK_PREC_MIXED && K >= 8) ? half_kc(K) : default_kc(K);
}

// The tile a plan runs for a precision, defaults resolved.  The mixed path's default kernel follows
// the measured crossover (batch 1024-4096 on MI355X): the radial stage on the matrix cores from K = 12
// up, the VALU kernel below (K = 6, 8: the VALU kernel is faster; K = 10: the same).  The matrix-core
// kernel splits rows between two waves at K = 12 (L = 10, 1024 distributions: 0.32 vs 0.37 ms per call)
// but not at K = 14, where one wave per item already fills the GPU (L = 12, 1062 distributions: 0.91 vs
// 1.01 ms; 4096: 2.94 vs 2.99 ms) and the batched path's K = 14 class runs 1.05 instead of 1.18 ms.
//
// Small batches (B <= 1024 when the caller passes B; a surrogate's sampled residual check is 1% of the
// cells): the waves in flight are far fewer than the GPU's slots, so a call takes as long as its
// slowest wave.  Unless the plan fixes them, both symmetrized precisions then split rows between two
// waves, and up to 256 distributions the symmetrized FP64 kernel takes the whole k3 range in one chunk
// (K <= 10; one pass over the slice's entries instead of two).  Measured at L=8, K=10: 164
// distributions 0.098 ms in FP64 (0.197 ms with the large-batch tile, 0.113 ms split with two chunks)
// and 0.065 ms mixed (0.115 ms); 656 distributions 0.17 ms in FP64 split with two chunks (0.29 ms
// unsplit, 0.26 ms split in one chunk) and 0.081 ms mixed (0.137 ms).  The matrix-core kernel at L=12,
// K=14: 164 distributions 0.32 ms split against 0.52 ms, 656 distributions 0.72 against 0.70 ms.  (With
// the current kernels and launch order 1062 distributions measured 0.86 ms split against 0.92 ms, the
// reverse of the figures above; the large-batch default at K = 14 is left as measured there.)
Tile plan_tile(const cfk_plan_s *p, int precision, int64_t B = 0) {
    const bool small = B > 0 && B <= 1024;
    Tile tl{p->kc[precision] > 0 ? p->kc[precision] : default_tile_kc(precision, p->K),
            p->nb[precision] > 0 ? p->nb[precision] : 1, p->pf[precision],
            p->wpb[precision] > 0 ? p->wpb[precision] : 1, p->split[precision]};
    if (precision == CFK_PREC_FP64_SYM && small) {
        if (p->kc[precision] <= 0 && p->K <= 10 && B <= 256) tl.kc = p->K;
        if (tl.split < 0) tl.split = 1;
    }
    if (precision == CFK_PREC_MIXED) {
        if (tl.pf < 0) tl.pf = (p->K >= 12 && mx_supports_k(p->K)) ? 1 : 0;
        if (tl.pf == 1 && !mx_supports_k(p->K)) tl.pf = 0;
        if (tl.split < 0) tl.split = (tl.pf == 1 && p->K <= 12) || small ? 1 : 0;
        if (tl.pf == 1) {
            tl.kc = mx_chunk(p->K);
            tl.nb = 1;
        }
    }
    if (tl.split < 0) tl.split = 0;
    return tl;
}

// Device staging an execute needs, in elements: the FP32 copy of f (mixed precision: packed rows
// [N_Q][B][RL] for the VALU kernel with a full k3 chunk and for the matrix-core kernel (chunk-packed),
// else [K][N_Q][B]) and, for the natural layout, the
// batch-fastest FP64 copies of f (FP64 precisions) and of Q.
StagingNeed plan_staging_need(const cfk_plan_s *p, int precision, int layout, int64_t B) {
    StagingNeed n{0, 0, 0};
    const int64_t NK = (int64_t)p->N_Q * p->K;
    const bool natural = layout == CFK_LAYOUT_NATURAL;
    if (precision == CFK_PREC_MIXED) {
        const Tile tl = plan_tile(p, precision, B);
        const bool packed = tl.pf == 1 || af_f32_packed(p->K, tl);
        const int kc = tl.pf == 1 ? mx_chunk(p->K) : p->K;
        n.f32 = packed ? (int64_t)p->N_Q * rl_of(p->K, kc) * B : NK * B;
        if (natural) n.q64 = NK * B;
    } else if (precision == CFK_PREC_FP64_SYM) {
        // packed FP64 rows (chunk-major, from either layout)
        const int kc = af_f64_sym_kc(p->K, plan_tile(p, precision, B));
        n.f64 = (int64_t)p->N_Q * ((p->K + kc - 1) / kc) * ((kc + 1) / 2 * 2) * B;
        if (natural) n.q64 = NK * B;
    } else if (natural) {
        n.f64 = NK * B;
        n.q64 = NK * B;
    }
    return n;
}

int plan_execute(cfk_plan_s *p, int variant, int precision, int layout, double *Q, const double *f, int64_t B,
                 hipStream_t st, const Staging *stg) {
    if (B < 0) return set_error(CFK_EINVAL, "B must be >= 0");
    if (B == 0) return CFK_OK;
    if (!Q || !f) return set_error(CFK_EINVAL, "null Q or f");
    if (layout != CFK_LAYOUT_NATURAL && layout != CFK_LAYOUT_BATCH_FASTEST)
        return set_error(CFK_EINVAL, "unknown layout %d", layout);
    if (variant < 0 || variant > 2) return set_error(CFK_EINVAL, "unknown variant %d", variant);
    if (precision != CFK_PREC_FP64 && precision != CFK_PREC_MIXED && precision != CFK_PREC_FP64_SYM)
        return set_error(CFK_EINVAL, "unknown precision %d", precision);
    int dev = -1;
    CFK_HIP(hipGetDevice(&dev));
    if (dev != p->device) return set_error(CFK_EINVAL, "plan lives on device %d, current device is %d", p->device, dev);
    if (B > max_batch(p->N_Q, p->K))
        return set_error(CFK_EINVAL, "B=%lld exceeds %lld distributions per call (f must stay below 4 GiB)", (long long)B,
                         (long long)max_batch(p->N_Q, p->K));
    if (p->N_G == 0) return CFK_OK;  // no Gaunt entries: Q(f) = 0
    const int64_t NK = (int64_t)p->N_Q * p->K;
    const bool natural = layout == CFK_LAYOUT_NATURAL;
    const Tile tl = plan_tile(p, precision, B);
    int rc;
    // staging: the caller's (fixed area, no allocation) or the plan's own buffers, grown as needed
    const StagingNeed need = plan_staging_need(p, precision, layout, B);
    float *s_f32 = nullptr;
    double *s_f64 = nullptr, *s_q64 = nullptr;
    if (stg) {
        s_f32 = stg->f32;
        s_f64 = stg->f64;
        s_q64 = stg->q64;
    }
    if (precision != CFK_PREC_FP64) {
        if (variant != CFK_VARIANT_ANGULAR_FIRST)
            return set_error(CFK_EUNSUP, "symmetrized/mixed precision is implemented for the angular-first variant only");
        if (!p->sym_ok) return set_error(CFK_EUNSUP, "symmetrized tables unavailable: %s", p->sym_why.c_str());
    }
    if (precision == CFK_PREC_MIXED) {
        const bool mx = tl.pf == 1;  // radial stage on the matrix cores
        if (mx && !p->sym_Rmx.p) {
            if (stg) return set_error(CFK_EUNSUP, "matrix-core operand tables missing");
            if ((rc = ensure_mx_tables(p))) return rc;
        }
        // the matrix-core kernel and the VALU kernel with a full k3 chunk read packed rows [N_Q][B][RL]
        // (chunk-packed for the matrix-core kernel's k3 chunks); the others [K][N_Q][B]
        const bool packed = mx || af_f32_packed(p->K, tl);
        const int pkc = mx ? mx_chunk(p->K) : p->K;
        if (!stg) {
            if ((rc = p->fT32.ensure((size_t)need.f32))) return rc;
            s_f32 = p->fT32.p;
            if (natural) {
                if ((rc = p->QT.ensure((size_t)need.q64))) return rc;
                s_q64 = p->QT.p;
            }
        }
        if (natural) {
            if ((rc = packed ? launch_pack_rows(f, s_f32, p->N_Q, p->K, pkc, B, true, st)
                             : launch_transpose_f32(f, s_f32, B, NK, st)))
                return rc;
            KernelTimer tm(p, st, stg == nullptr);
            if ((rc = mx ? launch_af_mx(p, p->sym_Rmx.p, tl, s_f32, s_q64, B, 0, st)
                         : launch_af_sym_f32(p, tl, s_f32, s_q64, B, 0, st)))
                return rc;
            tm.stop();
            return launch_transpose(s_q64, Q, NK, B, p->accumulate != 0, st);
        }
        if ((rc = packed ? launch_pack_rows(f, s_f32, p->N_Q, p->K, pkc, B, false, st)
                         : launch_to_f32(f, s_f32, NK * B, st)))
            return rc;
        KernelTimer tm(p, st, stg == nullptr);
        if ((rc = mx ? launch_af_mx(p, p->sym_Rmx.p, tl, s_f32, Q, B, p->accumulate, st)
                     : launch_af_sym_f32(p, tl, s_f32, Q, B, p->accumulate, st)))
            return rc;
        tm.stop();
        return CFK_OK;
    }
    const bool sym = precision == CFK_PREC_FP64_SYM;
    if (sym) {
        // the symmetrized FP64 kernel reads f as chunk-major packed FP64 rows: one 16-byte load per two
        // coefficients at one scalar offset per row chunk, instead of one 8-byte load per coefficient
        if (!stg) {
            if ((rc = p->fT.ensure((size_t)need.f64))) return rc;
            s_f64 = p->fT.p;
            if (natural) {
                if ((rc = p->QT.ensure((size_t)need.q64))) return rc;
                s_q64 = p->QT.p;
            }
        }
        if ((rc = launch_pack_rows_f64(f, s_f64, p->N_Q, p->K, af_f64_sym_kc(p->K, tl), B, natural, st))) return rc;
        KernelTimer tm(p, st, stg == nullptr);
        if ((rc = launch_af_sym_f64(p, tl, s_f64, natural ? s_q64 : Q, B, natural ? 0 : p->accumulate, st))) return rc;
        tm.stop();
        return natural ? launch_transpose(s_q64, Q, NK, B, p->accumulate != 0, st) : CFK_OK;
    }
    if (!natural) {
        KernelTimer tm(p, st, stg == nullptr);
        rc = launch_sparse_fp64(p, variant, tl, f, Q, B, p->accumulate, st);
        if (rc) return rc;
        tm.stop();
        return CFK_OK;
    }
    if (!stg) {
        if ((rc = p->fT.ensure((size_t)need.f64))) return rc;
        if ((rc = p->QT.ensure((size_t)need.q64))) return rc;
        s_f64 = p->fT.p;
        s_q64 = p->QT.p;
    }
    if ((rc = launch_transpose(f, s_f64, B, NK, false, st))) return rc;
    {
        KernelTimer tm(p, st, stg == nullptr);
        rc = launch_sparse_fp64(p, variant, tl, s_f64, s_q64, B, 0, st);
        if (rc) return rc;
        tm.stop();
    }
    return launch_transpose(s_q64, Q, NK, B, p->accumulate != 0, st);
}

// One-shot host path shared by the three sparse MEX equivalents.
int oneshot_sparse(int variant, double *Q, const double *f, const double *labels, const double *vals,
                   const double *ic_map, const double *R, int64_t N_G, int64_t R_numel, int32_t N_Q, int32_t K) {
    int rc;
    if (!Q || !f) return set_error(CFK_EINVAL, "null Q or f");
    if ((rc = validate_coo(labels, vals, ic_map, N_G, R, R_numel, N_Q, K))) return rc;
    if (N_G == 0) return CFK_OK;  // empty sum: Q unchanged, as the reference's loops leave it
    if ((rc = check_device())) return rc;
    cfk_plan_s plan;
    CFK_HIP(hipGetDevice(&plan.device));
    if ((rc = build_plan(labels, vals, ic_map, N_G, R, R_numel, N_Q, K, &plan))) return rc;
    const size_t NK = (size_t)N_Q * K;
    DevBuf<double> dQ, df;
    if ((rc = dQ.upload(Q, NK))) return rc;
    if ((rc = df.upload(f, NK))) return rc;
    if ((rc = plan_execute(&plan, variant, CFK_PREC_FP64, CFK_LAYOUT_NATURAL, dQ.p, df.p, 1, nullptr))) return rc;
    CFK_HIP(hipMemcpy(Q, dQ.p, NK * sizeof(double), hipMemcpyDeviceToHost));
    return CFK_OK;
}

}  // namespace cfk

using namespace cfk;

extern "C" {

const char *cfk_last_error(void) { return g_err.c_str(); }

int cfk_version(void) { return 1; }

int cfk_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int cfk_angular_first_collision_kernel(double *Q, const double *f, const double *labels, const double *vals,
                                       const double *ic_map, const double *R, int64_t N_G, int64_t R_numel,
                                       int32_t N_Q, int32_t K) {
    return oneshot_sparse(CFK_VARIANT_ANGULAR_FIRST, Q, f, labels, vals, ic_map, R, N_G, R_numel, N_Q, K);
}

int cfk_radial_first_collision_kernel(double *Q, const double *f, const double *labels, const double *vals,
                                      const double *ic_map, const double *R, int64_t N_G, int64_t R_numel,
                                      int32_t N_Q, int32_t K) {
    return oneshot_sparse(CFK_VARIANT_RADIAL_FIRST, Q, f, labels, vals, ic_map, R, N_G, R_numel, N_Q, K);
}

int cfk_naive_collision_kernel(double *Q, const double *f, const double *labels, const double *vals,
                               const double *ic_map, const double *R, int64_t N_G, int64_t R_numel, int32_t N_Q,
                               int32_t K) {
    return oneshot_sparse(CFK_VARIANT_NAIVE, Q, f, labels, vals, ic_map, R, N_G, R_numel, N_Q, K);
}

int cfk_dense_execute(double *Q_dev, const double *f_dev, const double *C_dev, int64_t N, int64_t B, void *stream) {
    int rc;
    if ((rc = check_device())) return rc;
    if (N < 1 || B < 0) return set_error(CFK_EINVAL, "N must be >= 1 and B >= 0");
    if (N > 4096) return set_error(CFK_EINVAL, "N=%lld: dense tensor of N^3 doubles is too large", (long long)N);
    if (B == 0) return CFK_OK;
    if (!Q_dev || !f_dev || !C_dev) return set_error(CFK_EINVAL, "null pointer");
    int64_t n_part = 0;
    if ((rc = dense_part_doubles(N, &n_part))) return rc;
    // the per-workgroup partial sums live for the call only: stream-ordered allocation, so the call
    // neither waits for the GPU nor for the allocator
    hipStream_t st = (hipStream_t)stream;
    double *part = nullptr;
    CFK_HIP(hipMallocAsync((void **)&part, (size_t)n_part * sizeof(double), st));
    rc = launch_dense(Q_dev, f_dev, C_dev, N, B, part, st);
    const hipError_t fe = hipFreeAsync(part, st);
    if (rc) return rc;
    CFK_HIP(fe);
    return CFK_OK;
}

int cfk_dense_assemble(double *C_dev, const double *labels, const double *vals, const double *ic_map, int64_t N_G,
                       const double *R, int64_t R_numel, int32_t N_Q, int32_t K, void *stream) {
    int rc;
    if ((rc = validate_coo(labels, vals, ic_map, N_G, R, R_numel, N_Q, K))) return rc;
    if ((int64_t)N_Q * K > 4096)
        return set_error(CFK_EINVAL, "N = N_Q*K = %lld: dense tensor of N^3 doubles is too large", (long long)N_Q * K);
    if (!C_dev) return set_error(CFK_EINVAL, "null C pointer");
    if ((rc = check_device())) return rc;
    std::vector<int4> zl((size_t)N_G);
    for (int64_t z = 0; z < N_G; ++z)
        zl[z] = make_int4((int)labels[z] - 1, (int)labels[z + N_G] - 1, (int)labels[z + 2 * N_G] - 1, (int)ic_map[z] - 1);
    DevBuf<int4> dzl;
    DevBuf<double> dzg, dR;
    if ((rc = dzl.upload(zl.data(), zl.size()))) return rc;
    if ((rc = dzg.upload(vals, (size_t)N_G))) return rc;
    if ((rc = dR.upload(R, (size_t)R_numel))) return rc;
    hipStream_t st = (hipStream_t)stream;
    if ((rc = launch_dense_assemble(C_dev, dzl.p, dzg.p, dR.p, N_G, N_Q, K, st))) return rc;
    CFK_HIP(hipStreamSynchronize(st));  // the uploaded tables are freed on return
    return CFK_OK;
}

int cfk_dense_tensor_kernel(double *Q, const double *f, const double *C, int64_t N) {
    int rc;
    if (!Q || !f || !C) return set_error(CFK_EINVAL, "null pointer");
    if (N < 1) return set_error(CFK_EINVAL, "N must be >= 1");
    if (N > 4096) return set_error(CFK_EINVAL, "N=%lld: dense tensor of N^3 doubles is too large", (long long)N);
    if ((rc = check_device())) return rc;
    DevBuf<double> dQ, df, dC;
    if ((rc = dQ.upload(Q, (size_t)N))) return rc;
    if ((rc = df.upload(f, (size_t)N))) return rc;
    if ((rc = dC.upload(C, (size_t)N * N * N))) return rc;
    if ((rc = cfk_dense_execute(dQ.p, df.p, dC.p, N, 1, nullptr))) return rc;
    CFK_HIP(hipMemcpy(Q, dQ.p, (size_t)N * sizeof(double), hipMemcpyDeviceToHost));
    return CFK_OK;
}

int cfk_plan_create(const double *labels, const double *vals, const double *ic_map, int64_t N_G, const double *R,
                    int64_t R_numel, int32_t N_Q, int32_t K, cfk_plan *out) {
    int rc;
    if (!out) return set_error(CFK_EINVAL, "null plan output");
    *out = nullptr;
    if ((rc = validate_coo(labels, vals, ic_map, N_G, R, R_numel, N_Q, K))) return rc;
    if ((rc = check_device())) return rc;
    cfk_plan_s *p = new cfk_plan_s();
    if (hipGetDevice(&p->device) != hipSuccess) {
        delete p;
        return set_error(CFK_EHIP, "hipGetDevice failed");
    }
    if ((rc = build_plan(labels, vals, ic_map, N_G, R, R_numel, N_Q, K, p))) {
        delete p;
        return rc;
    }
    // the matrix-core kernel's operand tables are built here, so that no execute has to allocate them
    if (p->sym_ok && mx_supports_k(K) && (rc = ensure_mx_tables(p))) {
        delete p;
        return rc;
    }
    *out = p;
    return CFK_OK;
}

int cfk_plan_destroy(cfk_plan plan) {
    delete plan;
    return CFK_OK;
}

int cfk_plan_execute(cfk_plan plan, int32_t variant, int32_t precision, int32_t layout, double *Q_dev,
                     const double *f_dev, int64_t B, void *stream) {
    if (!plan) return set_error(CFK_EINVAL, "null plan");
    return plan_execute(plan, variant, precision, layout, Q_dev, f_dev, B, (hipStream_t)stream);
}

int cfk_plan_autotune(cfk_plan plan, int64_t B, void *stream, int32_t *kc_out) {  // kc_out = tiles_out
    if (!plan) return set_error(CFK_EINVAL, "null plan");
    if (B < 1) return set_error(CFK_EINVAL, "B must be >= 1");
    cfk_plan_s *p = plan;
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = (int64_t)p->N_Q * p->K * B;
    DevBuf<double> f, Q;
    int rc;
    if ((rc = f.alloc((size_t)n)) || (rc = Q.alloc((size_t)n))) return rc;
    CFK_HIP(hipMemsetAsync(f.p, 0, (size_t)n * sizeof(double), st));
    CFK_HIP(hipMemsetAsync(Q.p, 0, (size_t)n * sizeof(double), st));
    const int kcs[2] = {p->K, half_kc(p->K)};
    const bool was_timing = p->timing;
    for (int prec = 0; prec < 3; ++prec) {
        if (prec != CFK_PREC_FP64 && !p->sym_ok) continue;
        double best = 1e300;
        Tile bt = plan_tile(p, prec);
        for (int c = 0; c < 2; ++c) {
            if (c == 1 && kcs[1] == kcs[0]) continue;
            for (int nb = 1; nb <= 2; ++nb)
                for (int pf = 0; pf <= ((prec == CFK_PREC_FP64 || (prec == CFK_PREC_MIXED && mx_supports_k(p->K))) ? 1 : 0);
                     ++pf)
                for (int v = 0; v < 5; ++v) {
                    // v = 0..2: 1, 2, 4 waves per workgroup; v = 3, 4: row split (symmetrized tables, two
                    // waves per batch tile) with 1 or 2 tiles per workgroup
                    const int split = v >= 3 ? 1 : 0;
                    const int wpb = split ? v - 2 : 1 << v;
                    if (split && prec == CFK_PREC_FP64) continue;
                    // mixed: pf = 1 selects the matrix-core radial stage (its own k3 chunk, one distribution
                    // per lane), timed once
                    if (prec == CFK_PREC_MIXED && pf == 1 && (nb != 1 || c != 0)) continue;
                    p->kc[prec] = kcs[c];
                    p->nb[prec] = nb;
                    p->pf[prec] = pf;
                    p->wpb[prec] = wpb;
                    p->split[prec] = split;
                    p->timing = false;
                    if ((rc = plan_execute(p, CFK_VARIANT_ANGULAR_FIRST, prec, CFK_LAYOUT_BATCH_FASTEST, Q.p, f.p, B,
                                           st)))
                        return rc;  // warm-up
                    p->timing = true;
                    for (int rep = 0; rep < 3; ++rep)
                        if ((rc = plan_execute(p, CFK_VARIANT_ANGULAR_FIRST, prec, CFK_LAYOUT_BATCH_FASTEST, Q.p, f.p,
                                               B, st)))
                            return rc;
                    double ms = 0.0;
                    int64_t nl = 0;
                    if ((rc = cfk_plan_read_timing(p, &ms, &nl))) return rc;
                    if (std::getenv("CFK_AUTOTUNE_VERBOSE"))
                        fprintf(stderr, "cfk autotune: precision %d KC %d NB %d PF %d WPB %d SPLIT %d: %.4f ms\n", prec,
                                kcs[c], nb, pf, wpb, split, nl > 0 ? ms / nl : -1.0);
                    if (nl > 0 && ms / nl < best) {
                        best = ms / nl;
                        bt = plan_tile(p, prec);
                    }
                }
        }
        p->kc[prec] = bt.kc;
        p->nb[prec] = bt.nb;
        p->pf[prec] = bt.pf;
        p->wpb[prec] = bt.wpb;
        p->split[prec] = bt.split;
    }
    p->timing = was_timing;
    CFK_HIP(hipStreamSynchronize(st));
    if (kc_out)
        for (int i = 0; i < 3; ++i) {
            const Tile t = plan_tile(p, i);
            kc_out[3 * i] = t.kc;
            kc_out[3 * i + 1] = t.nb;
            kc_out[3 * i + 2] = t.pf;
        }
    return CFK_OK;
}

int cfk_plan_linearize(cfk_plan plan, const double *f0, double *J, void *stream) {
    if (!plan || !J) return set_error(CFK_EINVAL, "null argument");
    cfk_plan_s *p = plan;
    const int64_t NK = (int64_t)p->N_Q * p->K, B = 2 * NK;
    if (NK > 8192) return set_error(CFK_EINVAL, "linearization of %lld coefficients is too large", (long long)NK);
    // column j from Q(f0 + e_j) - Q(f0 - e_j) = 2 J e_j (Q is quadratic): 2 NK distributions, natural layout
    std::vector<double> f((size_t)(B * NK), 0.0);
    for (int64_t j = 0; j < NK; ++j)
        for (int64_t i = 0; i < NK; ++i) {
            const double base = f0 ? f0[i] : (i == 0 ? 1.0 : 0.0);
            f[(size_t)(j * NK + i)] = base;
            f[(size_t)((NK + j) * NK + i)] = base;
        }
    for (int64_t j = 0; j < NK; ++j) {
        f[(size_t)(j * NK + j)] += 1.0;
        f[(size_t)((NK + j) * NK + j)] -= 1.0;
    }
    hipStream_t st = (hipStream_t)stream;
    DevBuf<double> fd, Qd;
    int rc;
    if ((rc = fd.upload(f.data(), f.size())) || (rc = Qd.alloc(f.size()))) return rc;
    CFK_HIP(hipMemsetAsync(Qd.p, 0, f.size() * sizeof(double), st));
    if ((rc = plan_execute(p, CFK_VARIANT_ANGULAR_FIRST, CFK_PREC_FP64, CFK_LAYOUT_NATURAL, Qd.p, fd.p, B, st))) return rc;
    CFK_HIP(hipStreamSynchronize(st));
    CFK_HIP(hipMemcpy(f.data(), Qd.p, f.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t j = 0; j < NK; ++j)
        for (int64_t i = 0; i < NK; ++i) J[i + j * NK] = 0.5 * (f[(size_t)(j * NK + i)] - f[(size_t)((NK + j) * NK + i)]);
    return CFK_OK;
}

int cfk_plan_set_tile(cfk_plan plan, int32_t precision, int32_t kc, int32_t nb, int32_t pf) {
    if (!plan) return set_error(CFK_EINVAL, "null plan");
    if (precision < 0 || precision > 2) return set_error(CFK_EINVAL, "unknown precision %d", precision);
    if (kc < 0 || kc > plan->K || nb < 0 || nb > 2 || pf < -1 || pf > 1 || (pf == -1 && precision != CFK_PREC_MIXED))
        return set_error(CFK_EINVAL, "tile (%d, %d, %d) out of range", kc, nb, pf);
    if (precision == CFK_PREC_MIXED && pf == 1 && !mx_supports_k(plan->K))
        return set_error(CFK_EUNSUP, "matrix-core radial stage: K=%d not in {6,8,10,12,14}", plan->K);
    plan->kc[precision] = kc;
    plan->nb[precision] = nb;
    plan->pf[precision] = pf;
    return CFK_OK;
}

int cfk_plan_set_waves(cfk_plan plan, int32_t precision, int32_t wpb) {
    if (!plan) return set_error(CFK_EINVAL, "null plan");
    if (precision < 0 || precision > 2) return set_error(CFK_EINVAL, "unknown precision %d", precision);
    if (wpb != 0 && wpb != 1 && wpb != 2 && wpb != 4)
        return set_error(CFK_EINVAL, "waves per workgroup must be 0, 1, 2 or 4 (got %d)", wpb);
    plan->wpb[precision] = wpb;
    return CFK_OK;
}

int cfk_plan_set_accumulate(cfk_plan plan, int32_t accumulate) {
    if (!plan) return set_error(CFK_EINVAL, "null plan");
    if (accumulate != 0 && accumulate != 1) return set_error(CFK_EINVAL, "accumulate must be 0 or 1");
    plan->accumulate = accumulate;
    return CFK_OK;
}

int cfk_plan_set_split(cfk_plan plan, int32_t precision, int32_t split) {
    if (!plan) return set_error(CFK_EINVAL, "null plan");
    if (precision < 0 || precision > 2) return set_error(CFK_EINVAL, "unknown precision %d", precision);
    if (split < -1 || split > 1) return set_error(CFK_EINVAL, "split must be -1, 0 or 1 (got %d)", split);
    plan->split[precision] = split;
    return CFK_OK;
}

int cfk_plan_get_split(cfk_plan plan, int32_t precision, int32_t *split_out) {
    if (!plan || !split_out) return set_error(CFK_EINVAL, "null argument");
    if (precision < 0 || precision > 2) return set_error(CFK_EINVAL, "unknown precision %d", precision);
    *split_out = plan_tile(plan, precision).split;
    return CFK_OK;
}

int cfk_plan_get_waves(cfk_plan plan, int32_t precision, int32_t *wpb_out) {
    if (!plan || !wpb_out) return set_error(CFK_EINVAL, "null argument");
    if (precision < 0 || precision > 2) return set_error(CFK_EINVAL, "unknown precision %d", precision);
    *wpb_out = plan->wpb[precision] > 0 ? plan->wpb[precision] : 1;
    return CFK_OK;
}

int cfk_plan_get_stats(cfk_plan plan, int64_t stats[8]) {
    if (!plan || !stats) return set_error(CFK_EINVAL, "null plan or stats");
    stats[0] = plan->N_G;
    stats[1] = plan->S;
    stats[2] = plan->sym_ok ? plan->sym_N_G : 0;
    stats[3] = plan->sym_ok ? plan->sym_S : 0;
    stats[4] = plan->sym_ok ? plan->sym_groups : 0;
    stats[5] = plan->sym_ok ? plan->sym_symmetric : 0;
    stats[6] = plan->sym_ok ? plan->sym_precise : 0;
    stats[7] = plan->n_l;
    return CFK_OK;
}

int cfk_plan_set_timing(cfk_plan plan, int32_t enable) {
    if (!plan) return set_error(CFK_EINVAL, "null plan");
    plan->timing = enable != 0;
    return CFK_OK;
}

int cfk_plan_read_timing(cfk_plan plan, double *total_ms, int64_t *launches) {
    if (!plan || !total_ms || !launches) return set_error(CFK_EINVAL, "null argument");
    double tot = 0.0;
    for (auto &e : plan->events) {
        CFK_HIP(hipEventSynchronize(e.second));
        float ms = 0.0f;
        CFK_HIP(hipEventElapsedTime(&ms, e.first, e.second));
        tot += ms;
    }
    *total_ms = tot;
    *launches = (int64_t)plan->events.size();
    for (auto &e : plan->events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    plan->events.clear();
    return CFK_OK;
}

int cfk_conservation_project(double *Q_dev, int64_t B, int32_t N_Q, int32_t K, int32_t layout, void *stream) {
    int rc;
    if ((rc = check_device())) return rc;
    if (!Q_dev || N_Q < 1 || K < 1 || B < 0) return set_error(CFK_EINVAL, "bad arguments");
    if (layout != CFK_LAYOUT_NATURAL && layout != CFK_LAYOUT_BATCH_FASTEST)
        return set_error(CFK_EINVAL, "unknown layout %d", layout);
    return launch_conservation(Q_dev, B, N_Q, K, layout, (hipStream_t)stream);
}

int cfk_real_gaunt(int32_t L, int64_t *N_G, double *labels, double *vals, double *ic_map) {
    if (!N_G) return set_error(CFK_EINVAL, "null N_G");
    if (L < 0 || L > 40) return set_error(CFK_EINVAL, "L=%d outside 0..40", L);
    RealGaunt g;
    try {
        g = real_gaunt(L);
    } catch (const std::exception &e) {
        return set_error(CFK_EINVAL, "%s", e.what());
    }
    const int64_t n = (int64_t)g.val.size();
    if (!labels || !vals || !ic_map) {
        *N_G = n;
        return CFK_OK;
    }
    if (*N_G < n) return set_error(CFK_EINVAL, "output capacity %lld < N_G=%lld", (long long)*N_G, (long long)n);
    for (int64_t z = 0; z < n; ++z) {
        labels[z] = g.q1[z];
        labels[z + n] = g.q2[z];
        labels[z + 2 * n] = g.q3[z];
        vals[z] = g.val[z];
        ic_map[z] = g.t[z];
    }
    *N_G = n;
    return CFK_OK;
}

int cfk_gauss_legendre(int32_t n, double a, double b, double *x, double *w) {
    if (n < 1 || !x || !w) return set_error(CFK_EINVAL, "bad arguments");
    gauss_legendre(n, a, b, x, w);
    return CFK_OK;
}

int cfk_gauss_gen_laguerre(int32_t n, double alpha, double *x, double *w) {
    if (n < 1 || !x || !w || !(alpha > -1.0)) return set_error(CFK_EINVAL, "bad arguments");
    gauss_gen_laguerre(n, alpha, x, w);
    return CFK_OK;
}

int cfk_sumfac_grid_sizes(int32_t K_max, int32_t L, int32_t radial_pad, int32_t angular_pad, int32_t sizes[7]) {
    if (K_max < 0 || L < 0 || radial_pad < 0 || angular_pad < 0 || !sizes) return set_error(CFK_EINVAL, "bad arguments");
    int s[7];
    sumfac_grid_sizes(K_max, L, radial_pad, angular_pad, s);
    for (int i = 0; i < 7; ++i) sizes[i] = s[i];
    return CFK_OK;
}

}  // extern "C"

// Shared definitions for the libcfk sources: status handling, device buffers, plan layout.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/cfk.h"

namespace cfk {

int set_error(int code, const char *fmt, ...);
int check_device();

#define CFK_HIP(call)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return ::cfk::set_error(CFK_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                    __FILE__, __LINE__);                                            \
    } while (0)

// Owning device buffer (freed with hipFree).
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return set_error(CFK_ENOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
        }
        n = count;
        return CFK_OK;
    }
    int upload(const T *host, size_t count) {
        int rc = alloc(count);
        if (rc) return rc;
        if (count) CFK_HIP(hipMemcpy(p, host, count * sizeof(T), hipMemcpyHostToDevice));
        return CFK_OK;
    }
    int ensure(size_t count) { return (count <= n && p) ? CFK_OK : alloc(count); }
};

// Tile variant of the angular-first kernel (chosen by the autotuner): k3 chunk width, distributions
// per lane, one-entry-ahead prefetch of the next nonzero's f rows (plain tables only).
struct Tile {
    int kc, nb, pf;
    int wpb = 1;  // waves per workgroup: batch tiles of one output row that share a CU's scalar cache
    int split = 0;  // symmetrized tables: each row's slices split between two waves (one tile per workgroup)
};

// One group of the batched uneven-size kernel (mixed precision, symmetrized tables, batch-fastest).
struct MultiGroupDev {
    const float *fT;
    double *Q;
    const struct Slice *slices;
    const int32_t *row_ptr;
    const int2 *zq;
    const float *zg;
    const double *R;
    const float *R32;
    const float *Rmx;         // A operands of the matrix-core kernel (null for the VALU kernel)
    const int32_t *row_mid;   // row split points (matrix-core kernel with the split)
    int64_t B;
    int N_Q, K;
    int accumulate;  // the group's plan's output mode: Q += Q(f) (1) or Q = Q(f) (0)
};
// Packed FP32 rows of f.  A kernel with k3 chunks of width KC (NCH chunks) reads chunk c of row (q, b)
// -- the coefficients k = c KC .. c KC + KC - 1, zero padded to KCP = KC rounded up to 4 -- at float
// offset ((q NCH + c) B + b) KCP, so one chunk is KCP / 4 16-byte loads and a wave's 64 lanes read 64 KCP
// contiguous floats.  With KC = K this is [N_Q][B][KP].  RL = NCH KCP floats per (q, b) in all.
constexpr int kcp_of(int KC) { return (KC + 3) / 4 * 4; }
constexpr int rl_of(int K, int KC) { return (K + KC - 1) / KC * kcp_of(KC); }
int mx_chunk(int K);  // k3 chunk width of the matrix-core kernel

constexpr int MULTI_MAX_GROUPS = 8;
// Largest batch per call: the contraction kernels read f through a buffer resource, which spans
// at most 4 GiB (8-byte coefficients counted), and the mixed kernels' packed FP32 rows (RL floats per
// (q, distribution): K rounded up to 4, or the matrix-core kernel's chunk-packed rows) take a signed
// 32-bit offset.
// The symmetrized FP64 kernel's packed FP64 rows (chunks of K, K/2 or K/3 coefficients, each rounded up
// to 2) span at most 4 GiB.
inline int64_t max_batch(int N_Q, int K) {
    const int64_t a = (((int64_t)1 << 32) - 1) / (8 * (int64_t)N_Q * K);
    const int rl = rl_of(K, mx_chunk(K)) > (K + 3) / 4 * 4 ? rl_of(K, mx_chunk(K)) : (K + 3) / 4 * 4;
    const int64_t b = (((int64_t)1 << 31) - 1) / (4 * (int64_t)N_Q * rl);
    int rd = 0;
    for (const int kc : {K, (K + 1) / 2, (K + 2) / 3}) {
        const int r = (K + kc - 1) / kc * ((kc + 1) / 2 * 2);
        rd = r > rd ? r : rd;
    }
    const int64_t c = (((int64_t)1 << 32) - 1) / (8 * (int64_t)N_Q * rd);
    return a < b ? (a < c ? a : c) : (b < c ? b : c);
}

// One (t, q1) slice of the (t,q1)-sorted COO: entries [z0, z1) share t and q1.  In the device copy
// of the symmetrized tables the last field instead flags the precise slices (partner block l = 0)
// for the mixed-precision kernel; the kernels never need q1 (it is the row being computed).
struct Slice {
    int32_t t, z0, z1, q1;
};
// In the symmetrized device tables the q1 field of a Slice (rows are given by row_ptr there) holds
// flags instead: the partner block is l = 0 (precise: FP64 R.Phi on the mixed path), the block is
// (k2,k3)-symmetric (l2 == l3, so R'_t[k1,k2,k3] = R'_t[k1,k3,k2]), and, from bit SLICE_G1_SHIFT up,
// the number of entries of the slice's first q2 group.
constexpr int32_t SLICE_PRECISE = 1, SLICE_SYMMETRIC = 2, SLICE_G1_SHIFT = 8;

}  // namespace cfk

// Plan: geometry + R resident on the device.
struct cfk_plan_s {
    int device = 0;
    int accumulate = 1;  // execute: Q += Q(f) (1, the MEX semantics) or Q = Q(f) (0)
    int N_Q = 0, K = 0, n_l = 0;
    int64_t N_G = 0, S = 0;
    // host copies (for re-planning and validation)
    std::vector<cfk::Slice> h_slices;  // ordered by (q1, t)
    std::vector<int32_t> h_row_ptr;    // [N_Q+1] into slices
    std::vector<int32_t> h_row_order;  // rows by descending work
    // device tables
    cfk::DevBuf<double> R;          // [n_l][K^3], column-major inside each block (k1 fastest)
    cfk::DevBuf<cfk::Slice> slices;
    cfk::DevBuf<int32_t> row_ptr, row_order;
    cfk::DevBuf<int2> zq;           // (q2, q3) 0-based, (t,q1)-sorted
    cfk::DevBuf<double> zg;         // Gaunt values, (t,q1)-sorted
    // Symmetrized tables (used by CFK_PREC_FP64_SYM and CFK_PREC_MIXED): only entries with
    // q2 <= q3 are kept (g halved when q2 == q3) and each kept block t=(l1,l2,l3), l2 <= l3, carries
    // R'_t[k1,k2,k3] = R_t[k1,k2,k3] + R_t'[k1,k3,k2] with t' = (l1,l3,l2).  Exact because the real
    // Gaunt tensor is symmetric in (q2,q3) and only the (q2,k2)<->(q3,k3) symmetric part of the
    // quadratic form contributes to Q.
    bool sym_ok = false;
    std::string sym_why;
    int64_t sym_N_G = 0, sym_S = 0;
    int64_t sym_groups = 0, sym_symmetric = 0, sym_precise = 0;  // q2 groups; slices with l2 == l3; with l2 == 0
    cfk::DevBuf<cfk::Slice> sym_slices;
    cfk::DevBuf<int32_t> sym_row_ptr, sym_row_order, sym_row_mid;
    cfk::DevBuf<int2> sym_zq;
    cfk::DevBuf<double> sym_zg, sym_R;
    cfk::DevBuf<float> sym_zg32, sym_R32;
    cfk::DevBuf<float> sym_Rmx;  // R' as A operands of the matrix-core kernel, 4-step groups (built on first use)
    std::vector<double> sym_row_work;  // cost model per output row (symmetrized tables)
    // scratch for the natural layout (batch-fastest staging)
    cfk::DevBuf<double> fT, QT;
    cfk::DevBuf<float> fT32;
    // tile choice per precision (index = CFK_PREC_*), set by the autotuner; 0 = default
    int kc[3] = {0, 0, 0};
    int nb[3] = {0, 0, 0};   // distributions per lane
    int pf[3] = {0, -1, 0};  // FP64: prefetch next nonzero (plain tables); mixed: matrix-core radial stage (-1: by K)
    int wpb[3] = {0, 0, 0};  // waves per workgroup of the angular-first kernels (0: default)
    int split[3] = {0, -1, -1};  // row split between two waves (symmetrized tables; -1: default of the kernel)
    // kernel timing (cfk_plan_set_timing)
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    ~cfk_plan_s() {
        for (auto &e : events) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    }
};

namespace cfk {
// Device staging of one plan execute (element counts from plan_staging_need; any may be zero).
struct StagingNeed {
    int64_t f32, f64, q64;
};
struct Staging {
    float *f32;
    double *f64, *q64;
};
StagingNeed plan_staging_need(const cfk_plan_s *p, int precision, int layout, int64_t B);
int plan_execute(cfk_plan_s *p, int variant, int precision, int layout, double *Q, const double *f, int64_t B,
                 hipStream_t st, const Staging *stg = nullptr);

// Brackets one kernel launch with HIP events when the plan's timing is enabled (and `allowed`: the
// execute variant that allocates nothing creates no events).
struct KernelTimer {
    cfk_plan_s *p;
    hipStream_t st;
    bool on;
    hipEvent_t a = nullptr, b = nullptr;
    KernelTimer(cfk_plan_s *plan, hipStream_t stream, bool allowed = true)
        : p(plan), st(stream), on(allowed && plan->timing) {
        if (on && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
            (void)hipEventRecord(a, st);
    }
    void stop() {
        if (on && a && b) {
            (void)hipEventRecord(b, st);
            p->events.emplace_back(a, b);
            a = b = nullptr;
        }
    }
    ~KernelTimer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};
}  // namespace cfk

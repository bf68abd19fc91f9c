// cfk_plan_execute_static: cfk_plan_execute without any device allocation.
//
// The staging an execute needs (the FP32 copy of f of the mixed path; the batch-fastest FP64 copies
// of f and Q for the natural layout) comes from one fixed device array of CFK_STATIC_SCRATCH_BYTES
// per GPU, defined below.  Its device address is looked up once per device with
// hipGetSymbolAddress (on the host, the array's name is only a stand-in for that address) and kept;
// the kernels receive that address as they receive the plan-owned buffers on the allocating path.
// Every call on a GPU shares that GPU's array, so calls that overlap on one GPU (two streams, two
// threads) race on it; the array keeps the last call's contents and nothing in it needs to start at
// zero (each call writes what it reads).
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"

namespace cfk {

namespace {

constexpr size_t kScratchBytes = CFK_STATIC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's copy of `scratch` (the current device), looked up on first use.
int scratch_address(int dev, unsigned char **out) {
    if (dev < 0 || dev >= kMaxDevices) return set_error(CFK_EUNSUP, "device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(CFK_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = static_cast<unsigned char *>(g_scratch_ptr[dev]);
    return CFK_OK;
}

size_t round256(size_t n) { return (n + 255) / 256 * 256; }

}  // namespace

}  // namespace cfk

using namespace cfk;

extern "C" int cfk_plan_execute_static(cfk_plan plan, int32_t variant, int32_t precision, int32_t layout,
                                       double *Q_dev, const double *f_dev, int64_t B, void *stream) {
    if (!plan) return set_error(CFK_EINVAL, "null plan");
    if (B < 0) return set_error(CFK_EINVAL, "B must be >= 0");
    if (precision != CFK_PREC_FP64 && precision != CFK_PREC_MIXED && precision != CFK_PREC_FP64_SYM)
        return set_error(CFK_EINVAL, "unknown precision %d", precision);
    if (layout != CFK_LAYOUT_NATURAL && layout != CFK_LAYOUT_BATCH_FASTEST)
        return set_error(CFK_EINVAL, "unknown layout %d", layout);
    if (B > max_batch(plan->N_Q, plan->K))
        return set_error(CFK_EINVAL, "B=%lld exceeds %lld distributions per call", (long long)B,
                         (long long)max_batch(plan->N_Q, plan->K));
    const StagingNeed n = plan_staging_need(plan, precision, layout, B);
    const size_t b32 = round256((size_t)n.f32 * sizeof(float)), b64 = round256((size_t)n.f64 * sizeof(double));
    const size_t bq = round256((size_t)n.q64 * sizeof(double));
    if (b32 + b64 + bq > kScratchBytes) {
        const int64_t per = (n.f32 * 4 + n.f64 * 8 + n.q64 * 8) / B;  // bytes per distribution (B > 0 here)
        return set_error(CFK_EINVAL,
                         "B=%lld needs %zu bytes of staging; the fixed area holds %zu bytes (about %lld distributions "
                         "of this plan in this precision and layout)",
                         (long long)B, b32 + b64 + bq, kScratchBytes, (long long)((int64_t)(kScratchBytes - 768) / per));
    }
    int dev = -1;
    CFK_HIP(hipGetDevice(&dev));
    if (dev != plan->device)
        return set_error(CFK_EINVAL, "plan lives on device %d, current device is %d", plan->device, dev);
    unsigned char *base = nullptr;
    int rc;
    if ((rc = scratch_address(dev, &base))) return rc;
    const Staging stg{n.f32 ? reinterpret_cast<float *>(base) : nullptr,
                      n.f64 ? reinterpret_cast<double *>(base + b32) : nullptr,
                      n.q64 ? reinterpret_cast<double *>(base + b32 + b64) : nullptr};
    return plan_execute(plan, variant, precision, layout, Q_dev, f_dev, B, (hipStream_t)stream, &stg);
}

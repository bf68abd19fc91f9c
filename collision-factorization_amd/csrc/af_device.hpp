// Device-side helpers shared by the contraction kernels (contract.hip, contract_mf.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cfk {

constexpr int WAVE = 64;
// routing-table entries fetched per scalar round trip; the device tables carry ZB padding entries
constexpr int ZB = 8;

// The routing tables and R are read-only for the whole launch and read at wave-uniform addresses.
// They are read through the constant address space, which lets the compiler fetch them with scalar
// loads even where their pointers were themselves loaded from memory (the batched kernel's group
// descriptors) and it could not otherwise rule out aliasing with the output stores.
// Per-lane data (f in, Q out) is read and written through the global address space, so the
// compiler emits global rather than flat memory instructions for it even where the pointers were
// loaded from memory: flat loads also count against the scalar-load counter, and every wait for a
// batch of table or R values would then wait for the outstanding f loads too.
// (The host pass of this file only needs the types to exist.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
#else
template <typename T>
using cptr = const T *;
template <typename T>
using gptr = T *;
#endif

// Per-lane element u + b of a batch-fastest array, with u wave-uniform and lo = b * sizeof(T) the
// lane's 32-bit byte offset.
template <typename T>
__device__ __forceinline__ gptr<T> lane_ptr(gptr<T> base, int64_t u, uint32_t lo) {
    return (gptr<T>)((gptr<char>)(base + u) + lo);
}

// Reads of f go through a buffer resource: the address is resource base + lane offset (a VGPR set
// once) + uniform byte offset (an SGPR), so a load costs no per-lane address arithmetic.  (Plain
// pointer arithmetic gets the lane part hoisted out of the loops and then pays one 64-bit vector add
// per load, which at L=8, K=10 was 7% of the kernel's vector instructions.)  The host guarantees
// that the whole array is below 4 GiB.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __amdgpu_buffer_rsrc_t frsrc_t;
__device__ __forceinline__ frsrc_t make_frsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T fload(frsrc_t rs, uint32_t lo, int64_t u) {
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, lo, (int)(u * 4), 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, lo, (int)(u * 8), 0));
}
// Four consecutive FP32 values at lane offset lo + imm and uniform byte offset soff (one 16-byte load).
__device__ __forceinline__ float4 fload4(frsrc_t rs, uint32_t lo, int soff) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, lo, soff, 0));
}
// Two consecutive FP64 values, the same way.
typedef double f64x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ f64x2_t fload2d(frsrc_t rs, uint32_t lo, int soff) {
    return __builtin_bit_cast(f64x2_t, __builtin_amdgcn_raw_buffer_load_b128(rs, lo, soff, 0));
}
#else
typedef const void *frsrc_t;
typedef double f64x2_t __attribute__((ext_vector_type(2)));
__device__ inline f64x2_t fload2d(frsrc_t rs, uint32_t lo, int soff) {
    return *(const f64x2_t *)((const char *)rs + lo + soff);
}
__device__ inline float4 fload4(frsrc_t rs, uint32_t lo, int soff) {
    return *(const float4 *)((const char *)rs + lo + soff);
}
__device__ inline frsrc_t make_frsrc(const void *p, uint32_t) { return p; }
template <typename T>
__device__ inline T fload(frsrc_t rs, uint32_t lo, int64_t u) {
    return *(const T *)((const char *)rs + lo + u * (int64_t)sizeof(T));
}
#endif

// Packed FP32 rows: the single-plan mixed kernel with a full k3 chunk reads f as [N_Q][B][KP] with
// KP = K rounded up to 4 (zero padding), so the K coefficients of one (q, distribution) pair are one
// contiguous 16-byte-aligned row: three 16-byte loads per row at one scalar offset, instead of K
// 4-byte loads each with its own.
constexpr int kp_of(int K) { return (K + 3) / 4 * 4; }
// (Chunk-packed rows of the matrix-core kernel: kcp_of / rl_of in common.hpp.)

// A wave-uniform FP32 value (held in an SGPR) copied into a VGPR.  On gfx950 FP32 FMAs whose operands
// are all VGPRs issue from two waves of a SIMD together, FMAs that read an SGPR do not; a value that
// feeds several FMAs (a Gaunt value in the h sums: K of them) is moved once and read from the VGPR.
__device__ __forceinline__ float in_vgpr(float x) {
    float r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
}
// (FP64 FMAs issue one wave at a time whatever their operands: no copy.)
__device__ __forceinline__ float to_vgpr(float x) { return in_vgpr(x); }
__device__ __forceinline__ double to_vgpr(double x) { return x; }

// Bytes of one chunk of KC coefficients in the packed rows: KC rounded up to whole 16-byte loads.
template <typename T>
constexpr int pk_chunk_bytes(int KC) { return (KC * (int)sizeof(T) + 15) / 16 * 16; }

}  // namespace cfk

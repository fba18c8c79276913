// fzh_internal.h — shared host/device helpers of libfzh (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

#include "fzh.h"

namespace fzh {

// ---------------------------------------------------------------------------
// Host-side error plumbing: every C entry point returns a status code and
// leaves a thread-local message behind.
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
// Zeroes [p, p+n) with a fill kernel on stream s.
int zero_bytes(hipStream_t s, void* p, size_t n);

// Optional per-kernel device timing: while enabled (fzh_kernel_timing_enable)
// and the stream is not being captured, a KernelTimer records a HIP event
// pair on the launch stream around the one launch in its scope; the pairs
// are resolved into per-name totals by fzh_kernel_timing_get.
int ktime_begin(hipStream_t s, const char* name);  // slot, or -1 when off
void ktime_end(hipStream_t s, int slot);
struct KernelTimer {
    hipStream_t s;
    int slot;
    KernelTimer(hipStream_t st, const char* name) : s(st), slot(ktime_begin(st, name)) {}
    ~KernelTimer() {
        if (slot >= 0) ktime_end(s, slot);
    }
};

#define FZH_TRY_HIP(expr)                                                        \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess)                                                    \
            return ::fzh::fail(FZH_E_RUNTIME, std::string(#expr ": ") +          \
                                                  hipGetErrorString(e_));        \
    } while (0)

#define FZH_CHECK_LAUNCH(name)                                                   \
    do {                                                                         \
        hipError_t e_ = hipGetLastError();                                       \
        if (e_ != hipSuccess)                                                    \
            return ::fzh::fail(FZH_E_RUNTIME, std::string(name " launch: ") +    \
                                                  hipGetErrorString(e_));        \
    } while (0)

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }
inline size_t cdiv(size_t a, size_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t a, size_t b) { return cdiv(a, b) * b; }

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Wave64 primitives
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Orders LDS traffic between lanes of one wave (all earlier LDS writes of
// the wave are visible to its later LDS reads); also a compiler barrier.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive scan of a u32 across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int l = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off, 64);
        if (l >= off) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t t = __shfl_xor(v, off, 64);
        v = t > v ? t : v;
    }
    return v;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  This maps the hardware index to a logical one so that every XCD
// works through one contiguous range: neighbouring tiles, which share cache
// lines at their edges, then meet in the same L2.  A bijection on
// [0, nblocks).
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t b, uint32_t nblocks) {
    constexpr uint32_t X = 8;
    const uint32_t per = nblocks / X, rem = nblocks % X;
    const uint32_t xcd = b % X, k = b / X;
    const uint32_t base = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
    return base + k;
}

}  // namespace fzh

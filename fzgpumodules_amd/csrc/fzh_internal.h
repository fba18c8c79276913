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

// ---------------------------------------------------------------------------
// Cross-workgroup prefixes over a few hundred coarse units.
//
// Layout of the scan state inside a workspace (zeroed once):
//   ScanCtx  (ticket, done, epoch)          16 bytes
//   status[units]  u64: [63:34] epoch (30 bits) [33:32] flag [31:0] value
// Units are handed out in order through `ticket`, so every predecessor of a
// waiting unit is already running and every wait ends.  The last wave of a
// launch (counted through `done`) resets ticket/done and bumps `epoch`;
// entries left by older launches carry an older epoch and read as "not
// ready", so no per-launch clearing is needed (also under graph replay).
// Only the compaction kernels use this; per-tile work never waits on
// another workgroup.
// ---------------------------------------------------------------------------
struct ScanCtx {
    uint32_t ticket;
    uint32_t done;
    uint32_t epoch;
    uint32_t pad;
};

constexpr size_t scan_state_bytes(size_t max_tiles) {
    return sizeof(ScanCtx) + 8 * (max_tiles + 1);
}

__device__ __forceinline__ uint64_t ld_status(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t pack_status(uint32_t epoch, uint32_t flag, uint32_t v) {
    return ((uint64_t)(epoch & 0x3FFFFFFFu) << 34) | ((uint64_t)flag << 32) | v;
}

struct TileTicket {
    uint32_t tile;
    uint32_t epoch;
};

// Called by a whole wave; returns the same ticket in every lane.
__device__ __forceinline__ TileTicket take_ticket(ScanCtx* ctx) {
    uint32_t t = 0, e = 0;
    if (lane_id() == 0) {
        t = __hip_atomic_fetch_add(&ctx->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        e = __hip_atomic_load(&ctx->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    TileTicket r;
    r.tile = __shfl(t, 0, 64);
    r.epoch = __shfl(e, 0, 64);
    return r;
}

// Called by a whole wave once it will touch the scan state no more.
// `total_waves` = number of waves of the launch that take tickets.
__device__ __forceinline__ void release_ticket(ScanCtx* ctx, uint32_t epoch, uint32_t total_waves) {
    if (lane_id() == 0) {
        uint32_t d = __hip_atomic_fetch_add(&ctx->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (d == total_waves - 1) {
            __hip_atomic_store(&ctx->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctx->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ctx->epoch, epoch + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Returned by aggregate_prefix when a predecessor never published (only
// possible with a workspace that was not initialised or is shared by two
// concurrent launches).  Callers must then write nothing and flag an error.
constexpr uint32_t kLookbackFailed = 0xFFFFFFFFu;
constexpr uint32_t kSpinLimit = 1u << 22;

// Exclusive prefix for `unit` from the aggregates of ALL earlier units
// (published with flag 1, read 64 at a time).  There is no chain of
// inclusive prefixes (as in a decoupled look-back): every unit publishes at
// once and waits only for its predecessors' aggregates, so the latency is
// ceil(unit/64) rounds of device-scope loads.  Meant for a few hundred
// coarse units (O(units^2) loads overall).  Whole wave must call; bounded
// spin; returns kLookbackFailed if a predecessor never publishes.
__device__ __forceinline__ uint32_t aggregate_prefix(uint64_t* status, uint32_t unit, uint32_t epoch,
                                                     uint32_t aggregate) {
    const int l = lane_id();
    const uint32_t ep = epoch & 0x3FFFFFFFu;
    if (l == 0) st_status(&status[unit], pack_status(epoch, 1u, aggregate));
    uint32_t sum = 0;
    for (uint32_t b = 0; b < unit; b += 64) {
        const uint32_t j = b + (uint32_t)l;
        uint32_t v = 0;
        bool timed_out = false;
        if (j < unit) {
            uint32_t spins = 0;
            while (true) {
                const uint64_t s = ld_status(&status[j]);
                if ((uint32_t)(s >> 34) == ep && ((s >> 32) & 3u) != 0u) {
                    v = (uint32_t)s;
                    break;
                }
                if (++spins > kSpinLimit) { timed_out = true; break; }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (__ballot(timed_out) != 0ull) return kLookbackFailed;
        sum += wave_sum(v);
    }
    return sum;
}

}  // namespace fzh

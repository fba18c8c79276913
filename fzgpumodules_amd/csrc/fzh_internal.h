// fzh_internal.h — shared host/device helpers of libfzh (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

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
// Copies `bytes` device bytes to host memory and synchronises the stream.
// Reads of 4-byte multiples up to 64 KiB go through a host-mapped staging
// buffer (one small kernel; a shorter round trip than a device-to-host
// copy); others fall back to hipMemcpyAsync.
int read_small(hipStream_t s, const void* d_src, size_t bytes, void* h_dst);

// RZE encode without its pack launch: the encoder (and, past 4096 chunks,
// the offset scan) run, and the pack step is described for a later launch
// that also does the pipeline's concat and readback
// (rze_pack_concat_readback), saving one launch per compress.
struct RzePackDesc {
    const uint8_t* in;
    uint64_t nbytes;
    const uint8_t* scratch;
    const uint32_t* sizes;
    const uint32_t* cleans;
    const uint32_t* offs;  // null: each workgroup sums the clean sizes before its chunk
    uint8_t* out;
    uint64_t* out_size;
    uint32_t nc;
    uint64_t nshuf;
};
int rze_enc_prepare(hipStream_t s, const void* d_in, size_t nbytes, int sew, void* d_out, size_t out_cap,
                    uint64_t* d_out_size, void* d_ws, size_t ws_bytes, RzePackDesc* desc);
int rze_pack_concat_readback(hipStream_t s, const RzePackDesc& d, const fzh_concat_slot* slots, int n, void* d_out,
                             uint64_t out_cap, uint64_t* d_total, const fzh_readback_entry* entries, int n_entries,
                             void* d_mapped_dst);

// Optional per-kernel device timing.  launch_k launches like
// kernel<<<grid, block, shm, s>>>(args...); while timing is on
// (fzh_kernel_timing_enable) and the stream is not being captured, the
// launch carries a start/stop event pair that HIP fills with the dispatch's
// own begin/end timestamps, so the time is the kernel's execution alone.
// The pairs are resolved into per-name totals by fzh_kernel_timing_get.
struct KtEvents {
    hipEvent_t a = nullptr, b = nullptr;
};
KtEvents ktime_events(hipStream_t s, const char* name);  // null events when off

template <typename... P, typename... A>
inline void launch_k(const char* name, void (*kernel)(P...), dim3 grid, dim3 block, uint32_t shm, hipStream_t s,
                     A... args) {
    static_assert(sizeof...(P) == sizeof...(A), "launch_k: argument count");
    const KtEvents e = ktime_events(s, name);
    hipExtLaunchKernelGGL(kernel, grid, block, shm, s, e.a, e.b, 0u, static_cast<P>(args)...);
}

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

// This thread's wave within its workgroup, as a wave-uniform value: work
// indices derived from it (tiles, chunks) then live in scalar registers, so
// their divisions, bounds tests and base addresses are scalar work.
// The same pointer in the global address space.  Device code that reads a
// pointer out of memory (a table of fields, a plan) holds a generic pointer,
// and the compiler then emits FLAT instructions, which are counted by both
// the vector-memory and the LDS/scalar wait counters; accessing through this
// cast gives global instructions.  Only for pointers into device memory.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gp(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

// 8- and 16-byte global accesses through gp() (clang vector types: HIP's
// vector classes have no address-space-qualified members).
typedef uint32_t fzh_u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t fzh_u32x4 __attribute__((ext_vector_type(4)));
typedef float fzh_f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 gld_u4(const void* p) {
    const fzh_u32x4 v = *gp(reinterpret_cast<const fzh_u32x4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float4 gld_f4(const void* p) {
    const fzh_f32x4 v = *gp(reinterpret_cast<const fzh_f32x4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst_u4(void* p, uint4 v) {
    const fzh_u32x4 w = {v.x, v.y, v.z, v.w};
    *gp(reinterpret_cast<fzh_u32x4*>(p)) = w;
}
__device__ __forceinline__ void gst_u2(void* p, uint2 v) {
    const fzh_u32x2 w = {v.x, v.y};
    *gp(reinterpret_cast<fzh_u32x2*>(p)) = w;
}

__device__ __forceinline__ uint32_t wave_in_wg() {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

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

// Lane movement inside the VALU with DPP (data-parallel primitives) instead
// of the LDS crossbar (ds_bpermute, one LDS round trip per step).  Every
// caller runs them with all 64 lanes active.  dpp_ctrl encodings (GFX9):
// quad_perm 0x00-0xFF, row_shr:n 0x110+n, wave_shr:1 0x138, row_bcast:15
// 0x142, row_bcast:31 0x143; lanes without a source keep `old` (0 here).
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, false);
}

// Value of lane l - 1 (lane 0: 0).
__device__ __forceinline__ uint32_t wave_shr1(uint32_t v) { return dpp_u32<0x138>(v); }
__device__ __forceinline__ float wave_shr1(float v) { return __uint_as_float(dpp_u32<0x138>(__float_as_uint(v))); }
// Value of lane l ^ 1.
__device__ __forceinline__ uint32_t lane_xor1(uint32_t v) { return dpp_u32<0xB1>(v); }

// Inclusive scan of a u32 across the 64 lanes of a wave: Hillis-Steele
// inside each row of 16 lanes, then the row totals carried by broadcasts.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    v += dpp_u32<0x111>(v);
    v += dpp_u32<0x112>(v);
    v += dpp_u32<0x114>(v);
    v += dpp_u32<0x118>(v);
    v += dpp_u32<0x142, 0xA>(v);
    v += dpp_u32<0x143, 0xC>(v);
    return v;
}

// Inclusive running maximum of a u32 across the 64 lanes of a wave (the same
// DPP pattern as wave_incl_scan, lanes without a source contributing 0).
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t v) {
    v = max(v, dpp_u32<0x111>(v));
    v = max(v, dpp_u32<0x112>(v));
    v = max(v, dpp_u32<0x114>(v));
    v = max(v, dpp_u32<0x118>(v));
    v = max(v, dpp_u32<0x142, 0xA>(v));
    v = max(v, dpp_u32<0x143, 0xC>(v));
    return v;
}

// Lane 63's value, wave-uniform (the total of an inclusive scan).
__device__ __forceinline__ uint32_t lane63(uint32_t v) { return (uint32_t)__builtin_amdgcn_readlane((int)v, 63); }
__device__ __forceinline__ float lane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }

// Wave-uniform sum and maximum.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    return lane63(wave_incl_scan(v));
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

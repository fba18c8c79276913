// fzh_scan.h — device-wide exclusive scan of u32 counts in three launches
// (segment reduce, one-workgroup scan of the segment sums, apply).  Used
// where a stage needs offsets from per-unit sizes that are only known on the
// device; nothing spins on another workgroup.
#pragma once

#include "fzh_internal.h"

namespace fzh {
namespace {

constexpr uint32_t SCAN_SEG = 4096;   // elements per segment (256 threads x 16)
constexpr uint32_t SCAN_MAX_SEGS = 1024 * 64;

// seg_sum[s] = sum of in[s*SEG .. (s+1)*SEG), saturated to 0xFFFFFFFF (WRAP:
// modulo 2^32, for callers that only use differences of nearby prefixes).
// When d_n is non-null the element count is min(*d_n, n) (a count known only
// on the device; the grid is sized for n).
__device__ __forceinline__ uint64_t eff_count(uint64_t n, const uint32_t* d_n) {
    return d_n ? (*d_n < n ? (uint64_t)*d_n : n) : n;
}

template <bool WRAP>
__global__ void __launch_bounds__(256)
scan_reduce_kernel(const uint32_t* __restrict__ in, uint64_t n, const uint32_t* __restrict__ d_n,
                   uint32_t* __restrict__ seg_sum) {
    __shared__ uint64_t s_w[4];
    n = eff_count(n, d_n);
    const uint64_t b = (uint64_t)blockIdx.x * SCAN_SEG;
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const uint64_t i = b + threadIdx.x + 256ull * k;
        if (i < n) s += in[i];
    }
    // 64-bit wave sum
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane_id() == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t t = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        seg_sum[blockIdx.x] = (!WRAP && t > 0xFFFFFFFFull) ? 0xFFFFFFFFu : (uint32_t)t;
    }
}

// 64-bit inclusive scan across the 64 lanes of a wave.
__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v) {
    const int l = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t t = __shfl_up(v, off, 64);
        if (l >= off) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t sat_u32(uint64_t v, bool wrap) {
    return (!wrap && v > 0xFFFFFFFFull) ? 0xFFFFFFFFu : (uint32_t)v;
}

// Sums (mod 2^32) of NL arrays over [0, end), computed by one 256-thread
// workgroup and returned to every thread: a kernel whose workgroup needs
// the exclusive prefix at its own position reads the counts before it
// instead of waiting for a separate scan launch (worth it while the arrays
// are short: every workgroup reads up to `end` elements per array, from L2).
// Every thread of the workgroup must call it (it holds a barrier); s holds
// 4 * NL u32 of LDS.
// The same over NL u32 arrays p[k][0, end) read with 16-byte loads (the
// arrays need only 4-byte alignment): fewer, wider loads shorten the
// latency chain of a workgroup that sums thousands of counts.
template <int NL>
__device__ __forceinline__ void wg256_sums_before_u32(const uint32_t* const (&p)[NL], uint32_t end, uint32_t* s,
                                                      uint32_t (&sum)[NL]) {
    const uint32_t t = threadIdx.x;
    uint32_t a[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) {
        const uint32_t mis = (uint32_t)((reinterpret_cast<uintptr_t>(p[k]) >> 2) & 3u);
        uint32_t head = (4u - mis) & 3u;
        if (head > end) head = end;
        const uint32_t nq = (end - head) / 4;
        const uint32_t tail0 = head + 4 * nq;
        a[k] = 0;
        if (t < head) a[k] += p[k][t];
        if (tail0 + t < end) a[k] += p[k][tail0 + t];
        const uint4* q = reinterpret_cast<const uint4*>(p[k] + head);
#pragma unroll 4
        for (uint32_t i = t; i < nq; i += 256) {
            const uint4 v = q[i];
            a[k] += (v.x + v.y) + (v.z + v.w);
        }
    }
    const int w = (int)(threadIdx.x >> 6);
#pragma unroll
    for (int k = 0; k < NL; k++) {
        a[k] = wave_sum(a[k]);
        if (lane_id() == 0) s[4 * k + w] = a[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NL; k++) sum[k] = s[4 * k] + s[4 * k + 1] + s[4 * k + 2] + s[4 * k + 3];
}

template <int NL, typename Load>
__device__ __forceinline__ void wg256_sums_before(uint32_t end, Load load, uint32_t* s, uint32_t (&sum)[NL]) {
    uint32_t a[NL];
#pragma unroll
    for (int k = 0; k < NL; k++) a[k] = 0;
#pragma unroll 4
    for (uint32_t i = threadIdx.x; i < end; i += 256) {
#pragma unroll
        for (int k = 0; k < NL; k++) a[k] += load(k, i);
    }
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NL; k++) {
        a[k] = wave_sum(a[k]);
        if (lane_id() == 0) s[4 * k + w] = a[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NL; k++) sum[k] = s[4 * k] + s[4 * k + 1] + s[4 * k + 2] + s[4 * k + 3];
}

// Exclusive scan of in[0..n) into out (may alias in) by one 1024-thread
// workgroup in passes of 16384 elements (64 KiB of LDS; each pass costs one
// dependent HBM round trip, so short arrays take one): each pass is read and
// written with coalesced dword accesses through LDS, and thread t scans the
// pass's elements [16t, 16t+16) in registers.  Returns the sum (64-bit).
// CLEAR zeroes in[] as it is read (in and out distinct).
template <bool CLEAR>
__device__ __forceinline__ uint64_t block_scan_1024(uint32_t* in, uint32_t* out, uint32_t n, bool wrap) {
    constexpr uint32_t PER = 16, PASS = 1024 * PER;
    __shared__ __align__(16) uint32_t s_v[PASS];
    __shared__ uint64_t s_w[16];
    const int tid = threadIdx.x, w = tid >> 6, l = lane_id();
    uint64_t carry = 0;
    for (uint32_t base = 0; base < n; base += PASS) {
        const uint32_t m = n - base < PASS ? n - base : PASS;
#pragma unroll
        for (uint32_t r = 0; r < PER; r++) {
            const uint32_t i = 1024 * r + tid;
            uint32_t v = 0;
            if (i < m) {
                v = in[base + i];
                if constexpr (CLEAR) in[base + i] = 0u;
            }
            s_v[i] = v;
        }
        __syncthreads();
        uint32_t x[PER];
#pragma unroll
        for (uint32_t q = 0; q < PER / 4; q++) {
            const uint4 u = *reinterpret_cast<const uint4*>(s_v + PER * tid + 4 * q);
            x[4 * q] = u.x;
            x[4 * q + 1] = u.y;
            x[4 * q + 2] = u.z;
            x[4 * q + 3] = u.w;
        }
        uint64_t sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) sum += x[k];
        const uint64_t incl = wave_incl_scan64(sum);
        if (l == 63) s_w[w] = incl;
        __syncthreads();
        uint64_t run = carry + incl - sum, tot = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const uint64_t t = s_w[j];
            run += j < w ? t : 0ull;
            tot += t;
        }
#pragma unroll
        for (uint32_t q = 0; q < PER / 4; q++) {
            uint32_t y[4];
#pragma unroll
            for (uint32_t k = 0; k < 4; k++) {
                y[k] = sat_u32(run, wrap);
                run += x[4 * q + k];
            }
            *reinterpret_cast<uint4*>(s_v + PER * tid + 4 * q) = make_uint4(y[0], y[1], y[2], y[3]);
        }
        carry += tot;
        __syncthreads();
#pragma unroll
        for (uint32_t r = 0; r < PER; r++) {
            const uint32_t i = 1024 * r + tid;
            if (i < m) out[base + i] = s_v[i];
        }
        __syncthreads();  // s_v and s_w are rewritten by the next pass
    }
    return carry;
}

// In-place exclusive scan of seg[0..nseg) by one 1024-thread workgroup;
// *total = sum (saturated to 0xFFFFFFFF, or modulo 2^32 with WRAP).
template <bool WRAP>
__global__ void __launch_bounds__(1024)
scan_partials_kernel(uint32_t* __restrict__ seg, uint32_t nseg, uint32_t* __restrict__ total) {
    const uint64_t t = block_scan_1024<false>(seg, seg, nseg, WRAP);
    if (threadIdx.x == 0 && total) *total = sat_u32(t, WRAP);
}

// out[i] = seg_base[s] + exclusive prefix of in inside segment s; the
// segment's 4096 elements go through LDS (coalesced dword loads and stores)
// and thread t scans elements [16t, 16t+16) of it in registers.  CLEAR: in[i]
// is zeroed after it is read (in and out distinct), ready for counters
// accumulated next time.
template <bool WRAP, bool CLEAR>
__global__ void __launch_bounds__(256)
scan_apply_kernel(uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t n,
                  const uint32_t* __restrict__ d_n, const uint32_t* __restrict__ seg_base) {
    __shared__ __align__(16) uint32_t s_v[SCAN_SEG];
    __shared__ uint64_t s_w[4];
    n = eff_count(n, d_n);
    const int tid = threadIdx.x, w = tid >> 6;
    const uint64_t b = (uint64_t)blockIdx.x * SCAN_SEG;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint64_t i = b + 256u * r + tid;
        uint32_t v = 0;
        if (i < n) {
            v = in[i];
            if constexpr (CLEAR) in[i] = 0u;
        }
        s_v[256 * r + tid] = v;
    }
    __syncthreads();
    uint32_t x[16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint4 u = *reinterpret_cast<const uint4*>(s_v + 16 * tid + 4 * q);
        x[4 * q] = u.x;
        x[4 * q + 1] = u.y;
        x[4 * q + 2] = u.z;
        x[4 * q + 3] = u.w;
    }
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) sum += x[k];
    const uint64_t incl = wave_incl_scan64(sum);
    if (lane_id() == 63) s_w[w] = incl;
    __syncthreads();
    const uint32_t sb = seg_base[blockIdx.x];
    // a saturated segment base stays saturated
    uint64_t run = ((!WRAP && sb == 0xFFFFFFFFu) ? 0x100000000ull : (uint64_t)sb) + incl - sum;
    for (int j = 0; j < w; j++) run += s_w[j];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint32_t y[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            y[k] = sat_u32(run, WRAP);
            run += x[4 * q + k];
        }
        *reinterpret_cast<uint4*>(s_v + 16 * tid + 4 * q) = make_uint4(y[0], y[1], y[2], y[3]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const uint64_t i = b + 256u * r + tid;
        if (i < n) out[i] = s_v[256 * r + tid];
    }
}

// Whole scan in one 1024-thread workgroup for n <= SCAN_SMALL (saves two
// launches on the short arrays -- per-tile counts, chunk sizes of one field
// -- that most calls scan).
constexpr uint32_t SCAN_SMALL = 65536;

template <bool WRAP, bool CLEAR>
__global__ void __launch_bounds__(1024)
scan_small_kernel(uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n_max,
                  const uint32_t* __restrict__ d_n, uint32_t* __restrict__ total) {
    const uint32_t n = (uint32_t)eff_count(n_max, d_n);
    const uint64_t t = block_scan_1024<CLEAR>(in, out, n, WRAP);
    if (threadIdx.x == 0 && total) *total = sat_u32(t, WRAP);
}

inline size_t scan_ws_bytes(size_t n) { return align_up(4 * std::max<size_t>(cdiv(n, SCAN_SEG), 1), 256); }

// Exclusive scan of in[0..n) into out (may alias in); *d_total = sum.  With
// d_n, only min(*d_n, n) elements are scanned.  Workspace: scan_ws_bytes(n);
// n up to SCAN_SEG * SCAN_MAX_SEGS.  Sums saturate at 0xFFFFFFFF, or wrap
// modulo 2^32 with WRAP.  CLEAR zeroes in[] as it is consumed (in != out).
template <bool WRAP = false, bool CLEAR = false>
inline int scan_u32(hipStream_t s, const uint32_t* in, uint32_t* out, size_t n, const uint32_t* d_n,
                    uint32_t* d_total, void* ws) {
    if (n <= SCAN_SMALL) {
        launch_k("scan_small_kernel", scan_small_kernel<WRAP, CLEAR>, dim3(1), dim3(1024), 0, s,
                 const_cast<uint32_t*>(in), out, (uint32_t)n, d_n, d_total);
        return FZH_OK;
    }
    const size_t nseg = std::max<size_t>(cdiv(n, SCAN_SEG), 1);
    if (nseg > SCAN_MAX_SEGS) return fail(FZH_E_INVALID, "scan: more than 2^28 elements");
    uint32_t* seg = (uint32_t*)ws;
    launch_k("scan_reduce_kernel", scan_reduce_kernel<WRAP>, dim3((unsigned)nseg), dim3(256), 0, s, in, n, d_n, seg);
    launch_k("scan_partials_kernel", scan_partials_kernel<WRAP>, dim3(1), dim3(1024), 0, s, seg, (uint32_t)nseg, d_total);
    if (n) launch_k("scan_apply_kernel", scan_apply_kernel<WRAP, CLEAR>, dim3((unsigned)nseg), dim3(256), 0, s, const_cast<uint32_t*>(in), out, n, d_n, seg);
    return FZH_OK;
}

}  // namespace
}  // namespace fzh

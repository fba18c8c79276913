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

// In-place exclusive scan of seg[0..nseg) by one 1024-thread workgroup;
// *total = sum (saturated to 0xFFFFFFFF, or modulo 2^32 with WRAP).
template <bool WRAP>
__global__ void __launch_bounds__(1024)
scan_partials_kernel(uint32_t* __restrict__ seg, uint32_t nseg, uint32_t* __restrict__ total) {
    __shared__ uint64_t s_w[16];
    const int tid = threadIdx.x, w = tid >> 6, l = lane_id();
    const uint32_t per = (nseg + 1023) / 1024;
    const uint32_t a = tid * per;
    uint64_t sum = 0;
    for (uint32_t k = 0; k < per; k++)
        if (a + k < nseg) sum += seg[a + k];
    uint64_t incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t t = __shfl_up(incl, off, 64);
        if (l >= off) incl += t;
    }
    if (l == 63) s_w[w] = incl;
    __syncthreads();
    uint64_t run = incl - sum;
    for (int j = 0; j < w; j++) run += s_w[j];
    for (uint32_t k = 0; k < per; k++) {
        if (a + k >= nseg) break;
        const uint32_t v = seg[a + k];
        seg[a + k] = (!WRAP && run > 0xFFFFFFFFull) ? 0xFFFFFFFFu : (uint32_t)run;
        run += v;
    }
    if (tid == 1023 && total) *total = (!WRAP && run > 0xFFFFFFFFull) ? 0xFFFFFFFFu : (uint32_t)run;
}

// out[i] = seg_base[s] + exclusive prefix of in inside segment s.  Thread t
// owns elements [16t, 16t+16) of its segment.  CLEAR: in[i] is zeroed after
// it is read (in and out distinct), ready for counters accumulated next time.
template <bool CLEAR>
__global__ void __launch_bounds__(256)
scan_apply_kernel(uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t n,
                  const uint32_t* __restrict__ d_n, const uint32_t* __restrict__ seg_base) {
    __shared__ uint32_t s_w[4];
    n = eff_count(n, d_n);
    const int tid = threadIdx.x, w = tid >> 6;
    const uint64_t b = (uint64_t)blockIdx.x * SCAN_SEG + 16ull * tid;
    uint32_t v[16], sum = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        v[k] = (b + k < n) ? in[b + k] : 0u;
        sum += v[k];
    }
    if constexpr (CLEAR) {
#pragma unroll
        for (int k = 0; k < 16; k++)
            if (b + k < n) in[b + k] = 0u;
    }
    const uint32_t incl = wave_incl_scan(sum);
    if (lane_id() == 63) s_w[w] = incl;
    __syncthreads();
    uint32_t run = seg_base[blockIdx.x] + incl - sum;
    for (int j = 0; j < w; j++) run += s_w[j];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (b + k < n) out[b + k] = run;
        run += v[k];
    }
}

// Whole scan in one 1024-thread workgroup for n <= SCAN_SMALL: thread t owns
// elements [t*per, (t+1)*per).  Saves two launches on the short arrays
// (per-tile counts, chunk sizes of one field) that most calls scan.
constexpr uint32_t SCAN_SMALL = 65536;

template <bool WRAP, bool CLEAR>
__global__ void __launch_bounds__(1024)
scan_small_kernel(uint32_t* __restrict__ in, uint32_t* __restrict__ out, uint32_t n_max,
                  const uint32_t* __restrict__ d_n, uint32_t* __restrict__ total) {
    __shared__ uint64_t s_w[16];
    const uint32_t n = (uint32_t)eff_count(n_max, d_n);
    const int tid = threadIdx.x, w = tid >> 6, l = lane_id();
    const uint32_t per = (n_max + 1023) / 1024;
    const uint32_t a = tid * per;
    uint64_t sum = 0;
    for (uint32_t k = 0; k < per; k++)
        if (a + k < n) sum += in[a + k];
    uint64_t incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t t = __shfl_up(incl, off, 64);
        if (l >= off) incl += t;
    }
    if (l == 63) s_w[w] = incl;
    __syncthreads();
    uint64_t run = incl - sum;
    for (int j = 0; j < w; j++) run += s_w[j];
    for (uint32_t k = 0; k < per; k++) {
        if (a + k >= n) break;
        const uint32_t v = in[a + k];
        if constexpr (CLEAR) in[a + k] = 0u;
        out[a + k] = (!WRAP && run > 0xFFFFFFFFull) ? 0xFFFFFFFFu : (uint32_t)run;
        run += v;
    }
    if (tid == 1023 && total) *total = (!WRAP && run > 0xFFFFFFFFull) ? 0xFFFFFFFFu : (uint32_t)run;
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
        scan_small_kernel<WRAP, CLEAR><<<1, 1024, 0, s>>>(const_cast<uint32_t*>(in), out, (uint32_t)n, d_n, d_total);
        return FZH_OK;
    }
    const size_t nseg = std::max<size_t>(cdiv(n, SCAN_SEG), 1);
    if (nseg > SCAN_MAX_SEGS) return fail(FZH_E_INVALID, "scan: more than 2^28 elements");
    uint32_t* seg = (uint32_t*)ws;
    scan_reduce_kernel<WRAP><<<(unsigned)nseg, 256, 0, s>>>(in, n, d_n, seg);
    scan_partials_kernel<WRAP><<<1, 1024, 0, s>>>(seg, (uint32_t)nseg, d_total);
    if (n) scan_apply_kernel<CLEAR><<<(unsigned)nseg, 256, 0, s>>>(const_cast<uint32_t*>(in), out, n, d_n, seg);
    return FZH_OK;
}

}  // namespace
}  // namespace fzh

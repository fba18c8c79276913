// fzh_compact.h — ordered compaction of per-tile lists (device code).
//
// Forward kernels that emit variable-length lists (outliers, patches) write
// each tile's entries to a fixed per-tile scratch slot, in the tile's order,
// plus a per-tile count.  launch_compact then places every tile's list at
// its exclusive prefix with a place kernel in which a wave takes 32 tiles
// of each list, finds the non-empty ones by ballot and copies them with all
// 64 lanes.  Up to COMPACT_INLINE tiles each workgroup sums the counts
// before its 128 tiles itself (one launch in all); beyond that a
// device-wide scan of the counts (fzh_scan.h) runs first.  No workgroup
// waits on another.
#pragma once

#include "fzh_scan.h"

namespace fzh {
namespace {

// Up to two lists (e.g. outliers and patches) per launch; a list with a null
// s_val moves indices only.  For two lists the second list's counts must
// directly follow the first's (cnt1 = cnt0 + T), so one scan covers both.
// V: the value type (float, or double for FP64 outlier errors).
template <typename V>
struct CompactListT {
    const V* s_val;
    const uint32_t* s_idx;
    const uint32_t* cnt;
    V* val;
    uint32_t* idx;
    uint32_t* total;
    uint64_t cap;
};
using CompactList = CompactListT<float>;

constexpr uint32_t COMPACT_INLINE = 16384;

// excl: exclusive prefix (mod 2^32) of the NL*T counts, excl[NL*T] = sum
// (INL: unused, the prefixes are summed in the kernel).
template <int NL, bool INL, typename V>
__global__ void __launch_bounds__(256)
place_kernel(CompactListT<V> L0, CompactListT<V> L1, uint32_t T, uint32_t E, const uint32_t* __restrict__ excl) {
    const int w = (int)wave_in_wg();
    const uint32_t wave = blockIdx.x * 4 + (uint32_t)w;
    const int l = lane_id();
    const uint32_t tb = wave * 32;
    const uint32_t li = (uint32_t)l >> 5;
    const uint32_t t = tb + (uint32_t)(l & 31);
    uint32_t c = 0;
    if (t < T && li < (uint32_t)NL) c = L0.cnt[(uint64_t)li * T + t];
    uint32_t ex = 0;  // this lane's tile's exclusive prefix within its list
    const uint32_t inc = wave_incl_scan(c);
    if constexpr (INL) {
        __shared__ uint32_t s_sum[4 * NL];
        __shared__ uint32_t s_wave[4][2];
        uint32_t before[NL];
        const uint32_t* lists[NL];
#pragma unroll
        for (int k = 0; k < NL; k++) lists[k] = L0.cnt + (uint64_t)k * T;
        wg256_sums_before_u32<NL>(lists, min(blockIdx.x * 128u, T), s_sum, before);
        const uint32_t h0 = (uint32_t)__builtin_amdgcn_readlane((int)inc, 31);  // list 0 part of the wave
        const uint32_t h1 = lane63(inc) - h0;                                    // list 1 part
        if (l == 0) {
            s_wave[w][0] = h0;
            s_wave[w][1] = h1;
        }
        __syncthreads();
        uint32_t p0 = before[0], p1 = NL == 2 ? before[NL - 1] : 0u;
        for (int k = 0; k < w; k++) {
            p0 += s_wave[k][0];
            p1 += s_wave[k][1];
        }
        ex = inc - c - (li ? h0 : 0u) + (li ? p1 : p0);
        if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {  // the last workgroup ends at or past T
            uint32_t q0 = before[0], q1 = NL == 2 ? before[NL - 1] : 0u;
            for (int k = 0; k < 4; k++) {
                q0 += s_wave[k][0];
                q1 += s_wave[k][1];
            }
            *L0.total = q0;
            if constexpr (NL == 2) *L1.total = q1;
        }
    } else {
        // every lane's destination in one load (not one dependent load per
        // tile inside the placement loop)
        if (c) ex = excl[(uint64_t)li * T + t] - (li ? excl[T] : 0u);
        if (wave == 0 && l == 0) {
            *L0.total = excl[T];
            if constexpr (NL == 2) *L1.total = excl[2 * T] - excl[T];
        }
    }
    if (tb >= T) return;
    // The wave's entries (its lanes' tiles' lists, in lane order) spread
    // over the lanes, 64 at a time: entry g belongs to the first lane j
    // whose inclusive count exceeds g (a binary search over the lanes'
    // counts), so a round's 64 copies are independent.
    const uint32_t total = lane63(inc);
    for (uint32_t g0 = 0; g0 < total; g0 += 64) {
        const uint32_t g = g0 + (uint32_t)l;
        int j = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
            if ((uint32_t)__shfl((int)inc, j + step - 1, 64) <= g) j += step;
        const uint32_t cj = (uint32_t)__shfl((int)c, j, 64);
        const uint32_t ij = (uint32_t)__shfl((int)inc, j, 64);
        const uint32_t dst0 = (uint32_t)__shfl((int)ex, j, 64);
        if (g < total) {
            const uint32_t k = g - (ij - cj);
            const uint32_t lj = (uint32_t)j >> 5;
            const uint32_t tile = tb + (uint32_t)(j & 31);
            const uint64_t src = (uint64_t)tile * E + k;
            const V* sv = lj ? L1.s_val : L0.s_val;
            const uint32_t* si = lj ? L1.s_idx : L0.s_idx;
            V* dv = lj ? L1.val : L0.val;
            uint32_t* di = lj ? L1.idx : L0.idx;
            const uint64_t cap = lj ? L1.cap : L0.cap;
            const uint64_t pos = (uint64_t)dst0 + k;
            if (pos < cap) {
                if (sv) dv[pos] = sv[src];
                di[pos] = si[src];
            }
        }
    }
}

// Compaction state: exclusive prefixes (2T + 1 u32) and the scan workspace.
inline size_t compact_state_bytes(size_t T) {
    return align_up(4 * (2 * T + 1), 256) + scan_ws_bytes(2 * std::max<size_t>(T, 1));
}

template <typename V>
inline int launch_compact(hipStream_t s, void* state, size_t T, uint32_t E, CompactListT<V> L0, CompactListT<V>* L1) {
    if (T == 0) return FZH_OK;
    const int nl = L1 ? 2 : 1;
    if (L1 && L1->cnt != L0.cnt + T) return fail(FZH_E_INVALID, "compact: second count list must follow the first");
    uint32_t* excl = (uint32_t*)state;
    void* ws = (uint8_t*)state + align_up(4 * (2 * T + 1), 256);
    const bool inl = T <= COMPACT_INLINE;
    if (!inl)
        if (int rc = scan_u32<true>(s, L0.cnt, excl, nl * T, nullptr, excl + nl * T, ws)) return rc;
    const unsigned blocks = (unsigned)cdiv(cdiv(T, 32), 4);
    auto* k = inl ? (L1 ? place_kernel<2, true, V> : place_kernel<1, true, V>)
                  : (L1 ? place_kernel<2, false, V> : place_kernel<1, false, V>);
    launch_k("place_kernel", k, dim3(blocks), dim3(256), 0, s, L0, L1 ? *L1 : CompactListT<V>{}, (uint32_t)T, E,
             excl);
    FZH_CHECK_LAUNCH("compact");
    return FZH_OK;
}

}  // namespace
}  // namespace fzh

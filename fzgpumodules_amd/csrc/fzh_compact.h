// fzh_compact.h — ordered compaction of per-tile lists (device code).
//
// Forward kernels that emit variable-length lists (outliers, patches) write
// each tile's entries to a fixed per-tile scratch slot, in the tile's order,
// plus a per-tile count.  compact_kernel then places every tile's list at
// its exclusive prefix.  Each workgroup owns 1024 consecutive tiles; its
// base is the sum of the aggregates the earlier workgroups publish (no chain
// of inclusive prefixes, so no wave waits on a long dependency).
#pragma once

#include "fzh_internal.h"

namespace fzh {
namespace {

// Up to two lists (e.g. outliers and patches) per launch; a list with a null
// s_val moves indices only.
struct CompactList {
    const float* s_val;
    const uint32_t* s_idx;
    const uint32_t* cnt;
    float* val;
    uint32_t* idx;
    uint32_t* total;
    uint64_t cap;
    uint64_t* status;
};

template <int NL>
__global__ void __launch_bounds__(256)
compact_kernel(CompactList L0, CompactList L1, uint32_t T, uint32_t E, ScanCtx* ctx, uint32_t G) {
    __shared__ uint32_t s_off[NL][1024];
    __shared__ uint32_t s_cnt[NL][1024];
    __shared__ uint32_t s_wsum[NL][4];
    __shared__ uint32_t s_base[NL];
    __shared__ uint32_t s_unit, s_epoch;
    const int tid = threadIdx.x, w = tid >> 6, l = lane_id();
    if (w == 0) {
        const TileTicket tk = take_ticket(ctx);
        if (l == 0) {
            s_unit = tk.tile;
            s_epoch = tk.epoch;
        }
    }
    __syncthreads();
    const uint32_t g = s_unit, ep = s_epoch;
    if (g < G) {
#pragma unroll
        for (int li = 0; li < NL; li++) {
            const CompactList& L = li == 0 ? L0 : L1;
            uint32_t c[4], sum = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint64_t tile = (uint64_t)g * 1024 + 4 * tid + k;
                c[k] = tile < T ? L.cnt[tile] : 0u;
                sum += c[k];
            }
            const uint32_t incl = wave_incl_scan(sum);
            if (l == 63) s_wsum[li][w] = incl;
            __syncthreads();
            uint32_t wbase = 0;
            for (int j = 0; j < w; j++) wbase += s_wsum[li][j];
            uint32_t run = wbase + incl - sum;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                s_off[li][4 * tid + k] = run;
                s_cnt[li][4 * tid + k] = c[k];
                run += c[k];
            }
        }
        __syncthreads();
        if (w == 0) {
#pragma unroll
            for (int li = 0; li < NL; li++) {
                const CompactList& L = li == 0 ? L0 : L1;
                const uint32_t agg = s_wsum[li][0] + s_wsum[li][1] + s_wsum[li][2] + s_wsum[li][3];
                const uint32_t base = aggregate_prefix(L.status, g, ep, agg);
                if (l == 0) {
                    s_base[li] = base;
                    if (g == G - 1) *L.total = base == kLookbackFailed ? 0xFFFFFFFFu : base + agg;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int li = 0; li < NL; li++) {
            const CompactList& L = li == 0 ? L0 : L1;
            const uint32_t base = s_base[li];
            if (base == kLookbackFailed) continue;
            for (int t = w; t < 1024; t += 4) {
                const uint64_t tile = (uint64_t)g * 1024 + t;
                if (tile >= T) break;
                const uint32_t c = s_cnt[li][t];
                if (c == 0) continue;
                const uint64_t dst0 = (uint64_t)base + s_off[li][t];
                const uint64_t src0 = tile * E;
                for (uint32_t j = l; j < c; j += 64) {
                    const uint64_t pos = dst0 + j;
                    if (pos < L.cap) {
                        if (L.s_val) L.val[pos] = L.s_val[src0 + j];
                        L.idx[pos] = L.s_idx[src0 + j];
                    }
                }
            }
        }
    }
    if (w == 0) release_ticket(ctx, ep, G);
}


// Scan state for compact_kernel: ScanCtx + one status array per list.
inline size_t compact_state_bytes(size_t T) {
    const size_t G = std::max<size_t>(1, cdiv(T, 1024));
    return align_up(sizeof(ScanCtx), 256) + 2 * align_up(8 * (G + 1), 256);
}

inline int launch_compact(hipStream_t s, void* state, size_t T, uint32_t E, CompactList L0, CompactList* L1) {
    const size_t G = std::max<size_t>(1, cdiv(T, 1024));
    uint8_t* w = (uint8_t*)state;
    ScanCtx* ctx = (ScanCtx*)w;
    L0.status = (uint64_t*)(w + align_up(sizeof(ScanCtx), 256));
    if (L1) {
        L1->status = (uint64_t*)(w + align_up(sizeof(ScanCtx), 256) + align_up(8 * (G + 1), 256));
        compact_kernel<2><<<(unsigned)G, 256, 0, s>>>(L0, *L1, (uint32_t)T, E, ctx, (uint32_t)G);
    } else {
        compact_kernel<1><<<(unsigned)G, 256, 0, s>>>(L0, CompactList{}, (uint32_t)T, E, ctx, (uint32_t)G);
    }
    return FZH_OK;
}

}  // namespace
}  // namespace fzh

// fzh_lorenzo_common.h — geometry, workspace layouts and outlier binning
// shared by the FP32 (lorenzo.hip) and FP64 (lorenzo_f64.hip) Lorenzo
// predict+quantize kernels.  Tile geometry follows the reference: 1D tiles
// of 1024 elements (modules/fused/lorenzo_quant/lorenzo_quant.cu:262-364),
// 2D tiles of 32x32 (lorenzo_quant_nd.cu:29-105), 3D tiles of 8x8x8
// (lorenzo_quant_nd.cu:199-274).
#pragma once

#include "fzh_compact.h"

namespace fzh {

// A fresh odd tag per binning call (one counter for every caller, so two
// calls never share a tag, whichever element type they bin).
uint32_t next_bin_tag();

namespace {

int ndim_of(size_t ny, size_t nz) { return nz > 1 ? 3 : (ny > 1 ? 2 : 1); }

size_t tile_count(size_t n, size_t nx, size_t ny, size_t nz) {
    switch (ndim_of(ny, nz)) {
    case 3: return cdiv(nx, 8) * cdiv(ny, 8) * cdiv(nz, 8);
    case 2: return cdiv(nx, 32) * cdiv(ny, 32);
    default: return cdiv(n, 1024);
    }
}

int check_dims(size_t n, size_t nx, size_t ny, size_t nz, const char* who) {
    const int nd = ndim_of(ny, nz);
    if (n > 0xFFFFFFFFull) return fail(FZH_E_INVALID, std::string(who) + ": more than 2^32 elements");
    if (nd == 1 && nx > n) return fail(FZH_E_INVALID, std::string(who) + ": dims[0] exceeds the element count");
    if (nd >= 2) {
        if (nx == 0) return fail(FZH_E_INVALID, std::string(who) + ": dims[0] must be set for 2D/3D");
        const size_t cover = nx * ny * (nd == 3 ? nz : 1);
        if (cover > n) return fail(FZH_E_INVALID, std::string(who) + ": dims exceed the element count");
    }
    return FZH_OK;
}


// Forward workspace: compaction state (prefixes + scan workspace), per-tile
// counts, and the per-tile list scratch (T x E entries per list; written
// only where outliers/patches occur; vsize = bytes per value).
struct FwdWs {
    size_t T, E, G;
    size_t o_ctx, o_cnt, o_pcnt, o_err, o_idx, o_pval, o_pidx, total;
};

FwdWs fwd_layout(size_t n, size_t nx, size_t ny, size_t nz, bool strict, size_t vsize = 4) {
    FwdWs w;
    w.T = tile_count(n, nx, ny, nz);
    w.E = ndim_of(ny, nz) == 3 ? 512 : 1024;
    w.G = std::max<size_t>(1, cdiv(w.T, 1024));
    size_t o = 0;
    w.o_ctx = o;  o = align_up(o + compact_state_bytes(w.T), 256);
    // patch counts directly follow the outlier counts (one scan covers both)
    w.o_cnt = o;  w.o_pcnt = o + 4 * w.T;
    o = align_up(o + 4 * std::max<size_t>(w.T, 1) * (strict ? 2 : 1), 256);
    w.o_err = o;  o = align_up(o + vsize * w.T * w.E, 256);
    w.o_idx = o;  o = align_up(o + 4 * w.T * w.E, 256);
    w.o_pval = o; if (strict) o = align_up(o + 4 * w.T * w.E, 256);
    w.o_pidx = o; if (strict) o = align_up(o + 4 * w.T * w.E, 256);
    w.total = o;
    return w;
}


// ===========================================================================
// Inverse: outlier binning by tile (counting sort), then per-tile decode
// ===========================================================================
// Division of any 32-bit n by a divisor d fixed for a launch, as a multiply
// (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, fig. 4.1): with l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1, n / d = (t + ((n - t) >> s1)) >> s2,
// t = mulhi(m, n), s1 = min(l, 1), s2 = max(l - 1, 0).  A few instructions
// where a division by a run-time value takes several dozen.
struct FastDiv {
    uint32_t d, m, s1, s2;
};
inline FastDiv fast_div(uint32_t d) {
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < (uint64_t)d) l++;
    FastDiv f;
    f.d = d;
    f.m = (uint32_t)(((((1ull << l) - d) << 32) / d) + 1);
    f.s1 = l ? 1u : 0u;
    f.s2 = l ? l - 1u : 0u;
    return f;
}
__device__ __forceinline__ uint32_t div_by(uint32_t n, const FastDiv& f) {
    const uint32_t t = __umulhi(f.m, n);
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

struct TileMap {
    int ndim;
    uint32_t nx, ny, nz;
    uint32_t gx, gy;
    uint64_t total;  // elements covered by the tiles' valid region
    FastDiv dx, dy;  // division by nx and by ny
};

__device__ __forceinline__ bool tile_of(const TileMap& m, uint32_t idx, uint32_t& tile, uint32_t& local) {
    if ((uint64_t)idx >= m.total) return false;
    if (m.ndim == 1) {
        tile = idx >> 10;
        local = idx & 1023u;
    } else if (m.ndim == 2) {
        const uint32_t y = div_by(idx, m.dx), x = idx - y * m.nx;
        tile = (y >> 5) * m.gx + (x >> 5);
        local = ((y & 31u) << 5) | (x & 31u);
    } else {
        const uint32_t yz = div_by(idx, m.dx), x = idx - yz * m.nx;
        const uint32_t z = div_by(yz, m.dy), y = yz - z * m.ny;
        tile = ((z >> 3) * m.gy + (y >> 3)) * m.gx + (x >> 3);
        local = ((z & 7u) << 6) | ((y & 7u) << 3) | (x & 7u);
    }
    return true;
}

// Outlier binning.  A list in tile order (what the forward writes) is
// binned without atomics: bin_check_kernel finds out whether the list is in
// tile order with valid indices; if it is, the counting kernels return at
// once, the scan sees a count of zero, and bin_place_kernel takes each
// tile's range from the tile changes along the list.  Otherwise (the
// reference's atomic order) a counting sort runs: per-tile counts with
// atomics, their scan, and placement.  `flag` == `tag` marks an unordered
// list; `tag` is new for every call, so the flag needs no clearing, and
// `dn` (the scan's element count) is T for an unordered list and is reset
// to 0 by the ordered path.
__global__ void __launch_bounds__(256)
bin_check_kernel(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ count, uint64_t max_out,
                 TileMap m, uint32_t T, uint32_t* __restrict__ flag, uint32_t* __restrict__ dn, uint32_t tag) {
    const uint64_t k = min((uint64_t)*count, max_out);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    bool unordered = false;
    // four entries per step, their loads issued together
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < k; i0 += 4 * stride) {
        uint32_t q[4], qp[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint64_t i = i0 + u * stride;
            q[u] = i < k ? idx[i] : 0u;
            qp[u] = (i < k && i > 0) ? idx[i - 1] : 0u;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint64_t i = i0 + u * stride;
            if (i >= k) continue;
            uint32_t t, loc, tp, lp;
            if (!tile_of(m, q[u], t, loc)) unordered = true;
            else if (i > 0 && (!tile_of(m, qp[u], tp, lp) || tp > t)) unordered = true;
        }
    }
    if (__ballot(unordered) && lane_id() == 0) {
        *flag = tag;
        *dn = T;
    }
}

__global__ void __launch_bounds__(256)
bin_count_kernel(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ count, uint64_t max_out,
                 TileMap m, uint32_t* __restrict__ tile_cnt, uint32_t* __restrict__ slot,
                 const uint32_t* __restrict__ flag, uint32_t tag) {
    if (*flag != tag) return;  // in tile order: bin_place_kernel bins it
    const uint64_t k = min((uint64_t)*count, max_out);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t t, loc;
        if (tile_of(m, idx[i], t, loc)) slot[i] = atomicAdd(&tile_cnt[t], 1u);
    }
}

template <typename V>
__global__ void __launch_bounds__(256)
bin_place_kernel(const V* __restrict__ err, const uint32_t* __restrict__ idx,
                 const uint32_t* __restrict__ count, uint64_t max_out, TileMap m,
                 uint32_t* __restrict__ off, const uint32_t* __restrict__ slot,
                 V* __restrict__ serr, uint32_t* __restrict__ sloc, uint32_t T,
                 const uint32_t* __restrict__ flag, uint32_t* __restrict__ dn, uint32_t tag) {
    const uint64_t k = min((uint64_t)*count, max_out);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (*flag == tag) {
        for (uint64_t i = gid; i < k; i += stride) {
            uint32_t t, loc;
            if (tile_of(m, idx[i], t, loc)) {
                const uint32_t p = off[t] + slot[i];
                serr[p] = err[i];
                sloc[p] = loc;
            }
        }
        return;
    }
    // in tile order: entry i starts the range of every tile after its
    // predecessor's tile up to its own (entry k: every tile up to T); four
    // entries per step, their loads issued together
    for (uint64_t i0 = gid; i0 <= k; i0 += 4 * stride) {
        uint32_t q[4], qp[4];
        V e[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint64_t i = i0 + u * stride;
            q[u] = i < k ? idx[i] : 0u;
            qp[u] = (i <= k && i > 0) ? idx[i - 1] : 0u;
            e[u] = i < k ? err[i] : V(0);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint64_t i = i0 + u * stride;
            if (i > k) continue;
            uint32_t t = T, loc = 0, tp = 0, lp;
            if (i < k) (void)tile_of(m, q[u], t, loc);
            const bool head = i == 0 || (tile_of(m, qp[u], tp, lp), tp != t);
            if (head)
                for (uint32_t w = i == 0 ? 0u : tp + 1; w <= t; w++) off[w] = (uint32_t)i;
            if (i < k) {
                serr[i] = e[u];
                sloc[i] = loc;
            }
        }
    }
    if (gid == 0) *dn = 0;
}


// With few outliers (at most kDirectOutliers, known on the host from the
// outlier array's size) every tile's wave reads the whole list itself and
// applies the entries that fall in its tile: no binning launches.
constexpr uint32_t kDirectOutliers = 256;

// Inverse workspace: per-tile outlier counts (zero between calls: the scan
// clears them as it reads), their exclusive prefix, the scan workspace, and
// the outliers binned by tile (vsize = bytes per outlier value).
struct InvWs {
    size_t T, max_out;
    size_t o_scan, o_cnt, o_off, o_slot, o_err, o_loc, o_flag, total;
};

InvWs inv_layout(size_t n, size_t nx, size_t ny, size_t nz, size_t max_out, size_t vsize = 4) {
    InvWs w;
    w.T = tile_count(n, nx, ny, nz);
    w.max_out = max_out;
    size_t o = 0;
    w.o_scan = o; o = align_up(o + scan_ws_bytes(std::max<size_t>(w.T, 1)), 256);
    w.o_cnt = o;  o = align_up(o + 4 * w.T, 256);
    w.o_off = o;  o = align_up(o + 4 * (w.T + 1), 256);
    w.o_slot = o; o = align_up(o + 4 * std::max<size_t>(max_out, 1), 256);
    w.o_err = o;  o = align_up(o + vsize * std::max<size_t>(max_out, 1), 256);
    w.o_loc = o;  o = align_up(o + 4 * std::max<size_t>(max_out, 1), 256);
    w.o_flag = o; o = align_up(o + 8, 256);
    w.total = o;
    return w;
}

TileMap tile_map(size_t n, size_t nx, size_t ny, size_t nz) {
    const int nd = ndim_of(ny, nz);
    TileMap m;
    m.ndim = nd;
    m.nx = (uint32_t)nx; m.ny = (uint32_t)ny; m.nz = (uint32_t)nz;
    m.gx = (uint32_t)(nd == 2 ? cdiv(nx, 32) : cdiv(nx, 8));
    m.gy = (uint32_t)cdiv(ny, 8);
    m.total = nd == 3 ? (uint64_t)nx * ny * nz : nd == 2 ? (uint64_t)nx * ny : n;
    m.dx = fast_div((uint32_t)std::max<size_t>(nx, 1));
    m.dy = fast_div((uint32_t)std::max<size_t>(ny, 1));
    return m;
}

// Inverse, part 1: the outliers binned by tile into the workspace (needs
// only the outlier arrays, so it may run before the codes exist).
template <typename V>
int inv_bin(hipStream_t s, size_t n, size_t nx, size_t ny, size_t nz, const V* err, const uint32_t* idx,
            const uint32_t* count, size_t max_out, void* ws, size_t ws_bytes) {
    const InvWs L = inv_layout(n, nx, ny, nz, max_out, sizeof(V));
    if (ws_bytes < L.total) return fail(FZH_E_CAPACITY, "lorenzo inv: workspace too small");
    uint8_t* w = (uint8_t*)ws;
    uint32_t* cnt = (uint32_t*)(w + L.o_cnt);
    uint32_t* off = (uint32_t*)(w + L.o_off);
    uint32_t* slot = (uint32_t*)(w + L.o_slot);
    V* serr = (V*)(w + L.o_err);
    uint32_t* sloc = (uint32_t*)(w + L.o_loc);
    if (max_out <= kDirectOutliers) return FZH_OK;  // the inverse reads the list itself
    const TileMap m = tile_map(n, nx, ny, nz);
    const bool has_out = count != nullptr && max_out > 0;
    if (!has_out) return zero_bytes(s, off, 4 * (L.T + 1));
    uint32_t* flag = (uint32_t*)(w + L.o_flag);
    uint32_t* dn = flag + 1;
    const uint32_t tag = next_bin_tag();
    const unsigned ob = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(max_out, 256), 1024));
    launch_k("bin_check_kernel", bin_check_kernel, dim3(ob), dim3(256), 0, s, idx, count, max_out, m, (uint32_t)L.T,
             flag, dn, tag);
    launch_k("bin_count_kernel", bin_count_kernel, dim3(ob), dim3(256), 0, s, idx, count, max_out, m, cnt, slot,
             (const uint32_t*)flag, tag);
    if (int rc = scan_u32<false, true>(s, cnt, off, L.T, dn, off + L.T, w + L.o_scan)) return rc;
    launch_k("bin_place_kernel", bin_place_kernel<V>, dim3(ob), dim3(256), 0, s, err, idx, count, max_out, m, off,
             (const uint32_t*)slot, serr, sloc, (uint32_t)L.T, (const uint32_t*)flag, dn, tag);
    FZH_CHECK_LAUNCH("lorenzo inv binning");
    return FZH_OK;
}


}  // namespace
}  // namespace fzh

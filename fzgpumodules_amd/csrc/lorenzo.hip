// lorenzo.hip — fused Lorenzo predict+quantize (1D/2D/3D) and its inverse,
// plus the REL/NOA value-base reduction, for gfx950 (wave64).
//
// Format geometry follows the reference exactly: 1D tiles of 1024 elements
// (modules/fused/lorenzo_quant/lorenzo_quant.cu:262-364), 2D tiles of 32x32
// (lorenzo_quant_nd.cu:29-105), 3D tiles of 8x8x8 (lorenzo_quant_nd.cu:199-274).
// One wave64 owns one tile.  Lane layouts are chosen for wave64:
//   1D  lane l, step q: elements 4(l+64q) .. 4(l+64q)+3 of the tile
//   2D  lane l: column x = l&31, rows 16*(l>>5) .. +15 of the tile
//   3D  lane l: x = l&7, z = l>>3, rows y = 0..7 of the tile
// The arithmetic per element is the reference's (same FP32 operation order),
// and the inverse scans reproduce the reference's addition trees exactly.
//
// The per-tile forward work lives in fzh_lorenzo_tile.h (shared with the
// batched launch).  Each tile lists its outliers in scratch and a compaction
// kernel places the lists, so the outlier list is deterministic: tile-major,
// and sorted by element index inside each tile (the reference appends them
// with atomics in arbitrary order).
#include <atomic>
#include <mutex>

#include "fzh_lorenzo_common.h"
#include "fzh_lorenzo_tile.h"

namespace fzh {

uint32_t next_bin_tag() {
    static std::atomic<uint32_t> epoch{0};
    return 2u * (epoch.fetch_add(1, std::memory_order_relaxed) + 1u) + 1u;
}

namespace {

// Per-tile lists written by the forward kernels: tile t's outliers occupy
// err/idx[t*E, t*E + cnt[t]) in the tile's order (E = elements per tile), and
// likewise its strict-mode patches.  The compaction kernel then places every
// tile's list at its exclusive prefix, so no wave of the forward kernels
// waits on another (a single-pass ordered scan over thousands of tiles is
// limited by device-scope latency on this part).
struct TileLists {
    float* err;
    uint32_t* idx;
    uint32_t* cnt;
    float* pval;
    uint32_t* pidx;
    uint32_t* pcnt;
    const float* d_params;  // device eb params ([1] abs_eb, [3] ebx2) or null
    float eb, ebx2;         // used when d_params is null
};

__device__ __forceinline__ StrictEb strict_eb(const TileLists& tl) {
    return tl.d_params ? StrictEb{tl.d_params[1], tl.d_params[3]} : StrictEb{tl.eb, tl.ebx2};
}

template <bool STRICT>
__device__ __forceinline__ void store_counts(const TileLists& tl, uint32_t tile, TileCounts tc) {
    if (lane_id() == 0) {
        tl.cnt[tile] = tc.cnt;
        if constexpr (STRICT) tl.pcnt[tile] = tc.pcnt;
    }
}

// ===========================================================================
// Forward kernels: one wave64 per tile, four tiles per 256-thread workgroup
// ===========================================================================

template <typename TCode, bool ZZ, bool STRICT>
__global__ void __launch_bounds__(256)
lz1d_fwd_kernel(const float* __restrict__ in, uint64_t n, uint64_t nv, float r_host, const float* __restrict__ d_r,
                float rad, TCode* __restrict__ codes, TileLists tl, uint32_t T) {
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (tile >= T) return;
    const float r = d_r ? *d_r : r_host;
    const uint64_t b = (uint64_t)tile * 1024;
    const TileCounts tc = lz1d_tile<TCode, ZZ, STRICT>(in, n, nv, tile, r, rad, codes,
                                                       ListOut{tl.err + b, tl.idx + b, 1024u},
                                                       ListOut{tl.pval + b, tl.pidx + b, 1024u},
                                                       STRICT ? strict_eb(tl) : StrictEb{0.0f, 0.0f});
    store_counts<STRICT>(tl, tile, tc);
}

template <typename TCode, bool ZZ, bool STRICT>
__global__ void __launch_bounds__(256)
lz2d_fwd_kernel(const float* __restrict__ in, uint32_t nx, uint32_t ny, uint64_t n, float r_host,
                const float* __restrict__ d_r, float rad, TCode* __restrict__ codes, TileLists tl, uint32_t gx,
                uint32_t T) {
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (tile >= T) return;
    const float r = d_r ? *d_r : r_host;
    const uint64_t b = (uint64_t)tile * 1024;
    const TileCounts tc = lz2d_tile<TCode, ZZ, STRICT>(in, nx, ny, n, tile, gx, tile == T - 1, r, rad, codes,
                                                       ListOut{tl.err + b, tl.idx + b, 1024u},
                                                       ListOut{tl.pval + b, tl.pidx + b, 1024u},
                                                       STRICT ? strict_eb(tl) : StrictEb{0.0f, 0.0f});
    store_counts<STRICT>(tl, tile, tc);
}

// 3D: one wave per pair of x-adjacent tiles (lz3d_pair), four pairs per
// workgroup; P = ceil(gx / 2) gy gz pairs.
template <typename TCode, bool ZZ, bool STRICT>
__global__ void __launch_bounds__(256)
lz3d_fwd_kernel(const float* __restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz, uint64_t n,
                float r_host, const float* __restrict__ d_r, float rad, TCode* __restrict__ codes, TileLists tl,
                Tiles tg, uint32_t T, uint32_t P) {
    const uint32_t pair = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (pair >= P) return;
    const uint32_t gx2 = (tg.gx + 1) / 2;
    const uint32_t bx2 = pair % gx2, by = (pair / gx2) % tg.gy, bz = pair / (gx2 * tg.gy);
    const uint32_t ta = 2 * bx2 + tg.gx * (by + tg.gy * bz), tb = ta + 1;
    const bool has_b = 2 * bx2 + 1 < tg.gx;
    const float r = d_r ? *d_r : r_host;
    const uint64_t ba = (uint64_t)ta * 512, bb = (uint64_t)tb * 512;
    Lz3PairOut o;
    o.loA = ListOut{tl.err + ba, tl.idx + ba, 512u};
    o.lpA = ListOut{tl.pval + ba, tl.pidx + ba, 512u};
    o.loB = has_b ? ListOut{tl.err + bb, tl.idx + bb, 512u} : ListOut{nullptr, nullptr, 0u};
    o.lpB = has_b ? ListOut{tl.pval + bb, tl.pidx + bb, 512u} : ListOut{nullptr, nullptr, 0u};
    const bool last = ta == T - 1 || (has_b && tb == T - 1);
    lz3d_pair<TCode, ZZ, STRICT>(in, nx, ny, nz, n, bx2, by, bz, last, r, rad, codes,
                                 STRICT ? strict_eb(tl) : StrictEb{0.0f, 0.0f}, o);
    store_counts<STRICT>(tl, ta, o.A);
    if (has_b) store_counts<STRICT>(tl, tb, o.B);
}

// ===========================================================================
// REL / NOA value base (predictor_utils.cuh:27-117), resolved on the device
// ===========================================================================
// Pass 1: per-workgroup (min, max) of a grid-stride slice, the first element
// (and 0.0 when the input is zero-padded) folded in as the reference's
// reduction starts from them.  Pass 2: one workgroup reduces the partials
// and resolves the bound -- two launches instead of a last-workgroup atomic,
// whose device-scope fences are costly across the XCDs' separate L2s.
// Most workgroups of pass 1 (so at most this many partial pairs).
constexpr unsigned VB_MAX_PARTIALS = FZH_VB_NOALLOC_PARTIALS;

__global__ void __launch_bounds__(256)
minmax_partial_kernel(const float* __restrict__ in, uint64_t n, float2* __restrict__ partial, bool zero_pad) {
    __shared__ float smn[4], smx[4];
    const float first = in[0];
    float mn = first, mx = first;
    if (zero_pad) {  // the buffer continues with zeros up to the padded length
        if (0.0f < mn) mn = 0.0f;
        if (0.0f > mx) mx = 0.0f;
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n4 = n / 4;
    const float4* in4 = reinterpret_cast<const float4*>(in);
    auto fold = [&](const float4& v) {
        if (v.x < mn) mn = v.x;
        if (v.x > mx) mx = v.x;
        if (v.y < mn) mn = v.y;
        if (v.y > mx) mx = v.y;
        if (v.z < mn) mn = v.z;
        if (v.z > mx) mx = v.z;
        if (v.w < mn) mn = v.w;
        if (v.w > mx) mx = v.w;
    };
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // large fields: eight independent 16-byte loads in flight per thread
    // (the grid is at most 512 workgroups: 16 MiB in flight, what the HBM
    // latency needs); small ones keep four (fewer, shorter rounds)
    if (n4 >= 64 * stride)
    for (; i + 7 * stride < n4; i += 8 * stride) {
        float4 v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = in4[i + k * stride];
#pragma unroll
        for (int k = 0; k < 8; k++) fold(v[k]);
    }
    for (; i + 3 * stride < n4; i += 4 * stride) {
        const float4 a = in4[i], b = in4[i + stride], c = in4[i + 2 * stride], d = in4[i + 3 * stride];
        fold(a);
        fold(b);
        fold(c);
        fold(d);
    }
    for (; i < n4; i += stride) fold(in4[i]);
    for (uint64_t j = n4 * 4 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const float v = in[j];
        if (v < mn) mn = v;
        if (v > mx) mx = v;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float a = __shfl_xor(mn, off, 64), b = __shfl_xor(mx, off, 64);
        if (a < mn) mn = a;
        if (b > mx) mx = b;
    }
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) { smn[w] = mn; smx[w] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            if (smn[k] < mn) mn = smn[k];
            if (smx[k] > mx) mx = smx[k];
        }
        partial[blockIdx.x] = make_float2(mn, mx);
    }
}

__global__ void __launch_bounds__(256)
minmax_final_kernel(const float2* __restrict__ partial, uint32_t np, int mode, float user_eb,
                    float* __restrict__ params) {
    __shared__ float smn[4], smx[4];
    float gmn = INFINITY, gmx = -INFINITY;
    for (uint32_t b = threadIdx.x; b < np; b += 256) {
        const float2 v = partial[b];
        if (v.x < gmn) gmn = v.x;
        if (v.y > gmx) gmx = v.y;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float a = __shfl_xor(gmn, off, 64), b = __shfl_xor(gmx, off, 64);
        if (a < gmn) gmn = a;
        if (b > gmx) gmx = b;
    }
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) { smn[w] = gmn; smx[w] = gmx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 4; k++) {
            if (smn[k] < gmn) gmn = smn[k];
            if (smx[k] > gmx) gmx = smx[k];
        }
        resolve_eb(gmn, gmx, mode, user_eb, params);
    }
}

// Adds the tile's binned outliers into the per-lane registers: lane k
// loads entry o0 + j0 + k of the tile's range (64 independent loads per
// round, not a dependent chain), then the wave walks the round's entries
// through cross-lane reads.  LANE_OF / REG_OF map the local index `loc` to
// (lane, register); registers are updated with a select chain so no dynamic
// register indexing is needed.
#define FZH_APPLY_OUTLIERS(NREG, DVAR, LANE_OF, REG_OF)                      \
    do {                                                                     \
        const uint32_t o0 = off[tile], o1 = off[tile + 1];                   \
        for (uint32_t j0 = o0; j0 < o1; j0 += 64) {                          \
            const uint32_t jl = j0 + (uint32_t)l;                            \
            const uint32_t loc_ = jl < o1 ? sloc[jl] : 0u;                   \
            const float e_ = jl < o1 ? serr[jl] : 0.0f;                      \
            const uint32_t m_ = min(o1 - j0, 64u);                           \
            for (uint32_t k = 0; k < m_; k++) {                              \
                const uint32_t loc = (uint32_t)__builtin_amdgcn_readlane((int)loc_, (int)k); \
                const float e = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e_), (int)k)); \
                const int tl = (LANE_OF);                                    \
                const int tr = (REG_OF);                                     \
                if (l == tl) {                                               \
                    _Pragma("unroll") for (int rr = 0; rr < NREG; rr++)      \
                        if (rr == tr) DVAR[rr] = e + DVAR[rr];               \
                }                                                            \
            }                                                                \
        }                                                                    \
    } while (0)


// A short (value, index) list read whole by every tile's wave: the
// outliers (added to the decoded deltas) or, in strict mode, the patches
// (assigned to the reconstructed values).  n_out: patch indices past the
// tiles but below n_out are stored by tile 0's wave.
struct DirectOut {
    const float* err;
    const uint32_t* idx;
    const uint32_t* count;
    uint32_t max_out;
    TileMap m;
    uint64_t n_out;
};

#define FZH_LIST_DIRECT(LIST, ASSIGN, NREG, DVAR, LANE_OF, REG_OF)                            \
    do {                                                                                      \
        const uint32_t kk = LIST.count ? min(*LIST.count, LIST.max_out) : 0u;                 \
        for (uint32_t j0 = 0; j0 < kk; j0 += 64) {                                            \
            const uint32_t jj = j0 + (uint32_t)l;                                             \
            uint32_t t_ = 0xFFFFFFFFu, loc_ = 0;                                              \
            float e_ = 0.0f;                                                                  \
            if (jj < kk) {                                                                    \
                if (tile_of(LIST.m, LIST.idx[jj], t_, loc_)) e_ = LIST.err[jj];               \
                else t_ = 0xFFFFFFFFu;                                                        \
            }                                                                                 \
            uint64_t hit = __ballot(t_ == tile);                                              \
            while (hit) {                                                                     \
                const int src_ = __builtin_ctzll(hit);                                        \
                hit &= hit - 1;                                                               \
                const uint32_t loc = (uint32_t)__shfl((int)loc_, src_, 64);                   \
                const float e = __shfl(e_, src_, 64);                                         \
                const int tl = (LANE_OF);                                                     \
                const int tr = (REG_OF);                                                      \
                if (l == tl) {                                                                \
                    _Pragma("unroll") for (int rr = 0; rr < NREG; rr++)                       \
                        if (rr == tr) DVAR[rr] = (ASSIGN) ? e : e + DVAR[rr];                 \
                }                                                                             \
            }                                                                                 \
        }                                                                                     \
    } while (0)

#define FZH_APPLY_DIRECT(NREG, DVAR, LANE_OF, REG_OF) FZH_LIST_DIRECT(dout, false, NREG, DVAR, LANE_OF, REG_OF)

// Strict-mode patches inside the inverse kernel (PATCH), and tile 0's
// store of any patch index past the tiles.
#define FZH_PATCH_ANY(NREG, DVAR, LANE_OF, REG_OF)                                           \
    do {                                                                                     \
        if constexpr (PATCH) {                                                               \
            FZH_LIST_DIRECT(pin, true, NREG, DVAR, LANE_OF, REG_OF);                          \
            if (tile == 0) {                                                                 \
                const uint32_t kp = pin.count ? min(*pin.count, pin.max_out) : 0u;           \
                for (uint32_t j = (uint32_t)l; j < kp; j += 64) {                            \
                    const uint32_t q = pin.idx[j];                                           \
                    if ((uint64_t)q >= pin.m.total && (uint64_t)q < pin.n_out) out[q] = pin.err[j]; \
                }                                                                            \
            }                                                                                \
        }                                                                                    \
    } while (0)

#define FZH_APPLY_ANY(NREG, DVAR, LANE_OF, REG_OF)                 \
    do {                                                           \
        if constexpr (DIRECT) {                                    \
            FZH_APPLY_DIRECT(NREG, DVAR, LANE_OF, REG_OF);         \
        } else {                                                   \
            FZH_APPLY_OUTLIERS(NREG, DVAR, LANE_OF, REG_OF);       \
        }                                                          \
    } while (0)

// 1D inverse, scan tree of lorenzo_quant.cu:17-87 (see header comment).
template <typename TCode, bool ZZ, bool DIRECT, bool PATCH>
__global__ void __launch_bounds__(256)
lz1d_inv_kernel(const TCode* __restrict__ codes, uint64_t n, float ebx2, const uint32_t* __restrict__ off,
                const float* __restrict__ serr, const uint32_t* __restrict__ sloc, float* __restrict__ out,
                uint32_t T, DirectOut dout, DirectOut pin) {
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (tile >= T) return;
    const int l = lane_id();
    const int h = l >> 5, lam = l & 31;
    const uint64_t base = (uint64_t)tile * 1024;
    const bool full = base + 1024 <= n;
    float d[16];  // d[4q + k]
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint64_t e0 = base + 4 * (uint64_t)(l + 64 * q);
        if (full) {
            TCode c[4];
            if constexpr (sizeof(TCode) == 2) {
                const uint2 w = *reinterpret_cast<const uint2*>(codes + e0);
                c[0] = (TCode)w.x; c[1] = (TCode)(w.x >> 16); c[2] = (TCode)w.y; c[3] = (TCode)(w.y >> 16);
            } else {
                const uint32_t w = *reinterpret_cast<const uint32_t*>(codes + e0);
                c[0] = (TCode)w; c[1] = (TCode)(w >> 8); c[2] = (TCode)(w >> 16); c[3] = (TCode)(w >> 24);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) d[4 * q + k] = dec_code<TCode, ZZ>(c[k]);
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) d[4 * q + k] = (e0 + k < n) ? dec_code<TCode, ZZ>(codes[e0 + k]) : 0.0f;
        }
    }
    // local index li -> reference thread t = li/4 = l + 64q, k = li%4
    FZH_APPLY_ANY(16, d, (int)((loc >> 2) & 63u), (int)(((loc >> 8) << 2) | (loc & 3u)));
    // per reference thread: sequential prefix of 4
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int k = 1; k < 4; k++) d[4 * q + k] += d[4 * q + k - 1];
    // per 32-lane warp: Hillis-Steele over thread totals, add lane-1's inclusive
    float wt[8];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float ts = d[4 * q + 3];
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const float v = __shfl_up(ts, o, 32);
            if (lam >= o) ts += v;
        }
        const float pre = __shfl_up(ts, 1, 32);
        if (lam > 0) {
#pragma unroll
            for (int k = 0; k < 4; k++) d[4 * q + k] += pre;
        }
        wt[2 * q] = __shfl(d[4 * q + 3], 31, 64);
        wt[2 * q + 1] = __shfl(d[4 * q + 3], 63, 64);
    }
    // per block: Hillis-Steele over the 8 warp totals
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
        float nw[8];
#pragma unroll
        for (int i = 0; i < 8; i++) nw[i] = (i >= o) ? wt[i] + wt[i - o] : wt[i];
#pragma unroll
        for (int i = 0; i < 8; i++) wt[i] = nw[i];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int w = 2 * q + h;
        if (w > 0) {
            float add = 0.0f;
#pragma unroll
            for (int i = 0; i < 7; i++)
                if (i == w - 1) add = wt[i];
#pragma unroll
            for (int k = 0; k < 4; k++) d[4 * q + k] += add;
        }
    }
#pragma unroll
    for (int i = 0; i < 16; i++) d[i] = d[i] * ebx2;
    FZH_PATCH_ANY(16, d, (int)((loc >> 2) & 63u), (int)(((loc >> 8) << 2) | (loc & 3u)));
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint64_t e0 = base + 4 * (uint64_t)(l + 64 * q);
        float4 v = make_float4(d[4 * q], d[4 * q + 1], d[4 * q + 2], d[4 * q + 3]);
        if (full) {
            *reinterpret_cast<float4*>(out + e0) = v;
        } else {
            if (e0 < n) out[e0] = v.x;
            if (e0 + 1 < n) out[e0 + 1] = v.y;
            if (e0 + 2 < n) out[e0 + 2] = v.z;
            if (e0 + 3 < n) out[e0 + 3] = v.w;
        }
    }
}

// 2D inverse, scan tree of lorenzo_quant_nd.cu:112-190.  Interior tiles
// (inside the field, rows spanning less than 2^32 bytes) load codes and
// store values with 32-bit lane offsets from the tile's scalar base and no
// bounds tests.
template <typename TCode, bool ZZ, bool DIRECT, bool PATCH, bool INTERIOR>
__device__ __forceinline__ void lz2d_inv_tile(const TCode* __restrict__ codes, uint32_t nx, uint32_t ny, float ebx2,
                                              const uint32_t* __restrict__ off, const float* __restrict__ serr,
                                              const uint32_t* __restrict__ sloc, float* __restrict__ out,
                                              uint32_t gx, uint32_t tile, const DirectOut& dout,
                                              const DirectOut& pin) {
    const int l = lane_id();
    const int x = l & 31, h = l >> 5;
    const uint32_t bx = tile % gx, by = tile / gx;
    const uint32_t X = bx * 32 + x, Y0 = by * 32 + 16 * h;
    const bool xin = X < nx;
    const uint64_t tbase = (uint64_t)(by * 32) * nx + bx * 32;
    const uint32_t loff = (uint32_t)(16 * h) * nx + (uint32_t)x;
    float d[16];
    if constexpr (INTERIOR) {
        const char* cp = reinterpret_cast<const char*>(codes + tbase);
#pragma unroll
        for (int i = 0; i < 16; i++)
            d[i] = dec_code<TCode, ZZ>(
                *reinterpret_cast<const TCode*>(cp + (uint32_t)sizeof(TCode) * (loff + (uint32_t)i * nx)));
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++)
            d[i] = (xin && Y0 + i < ny) ? dec_code<TCode, ZZ>(codes[(uint64_t)(Y0 + i) * nx + X]) : 0.0f;
    }
    // local = (ty << 5) | tx ; lane = tx + 32*(ty >= 16), register = ty & 15
    FZH_APPLY_ANY(16, d, (int)((loc & 31u) | ((loc >> 4) & 32u)), (int)((loc >> 5) & 15u));
    // y: sequential inside each 8-row group
#pragma unroll
    for (int i = 1; i < 8; i++) d[i] += d[i - 1];
#pragma unroll
    for (int i = 9; i < 16; i++) d[i] += d[i - 1];
    // cross-group addends: A1 = T0, A2 = T0+T1, A3 = A2+T2
    const float tlo = d[7], thi = d[15];
    const float a2_own = thi + tlo;                 // meaningful in h = 0 lanes
    const float a2 = __shfl(a2_own, x, 64);         // lanes of h = 1 read A2 from h = 0
    if (h == 0) {
#pragma unroll
        for (int i = 8; i < 16; i++) d[i] += tlo;
    } else {
        const float a3 = tlo + a2;
#pragma unroll
        for (int i = 0; i < 8; i++) d[i] += a2;
#pragma unroll
        for (int i = 8; i < 16; i++) d[i] += a3;
    }
    // x: Hillis-Steele over the 32 columns, then scale
#pragma unroll
    for (int i = 0; i < 16; i++) {
        float v = d[i];
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const float u = __shfl_up(v, o, 32);
            if (x >= o) v += u;
        }
        d[i] = v * ebx2;
    }
    FZH_PATCH_ANY(16, d, (int)((loc & 31u) | ((loc >> 4) & 32u)), (int)((loc >> 5) & 15u));
    if constexpr (INTERIOR) {
        char* op = reinterpret_cast<char*>(out + tbase);
#pragma unroll
        for (int i = 0; i < 16; i++) *reinterpret_cast<float*>(op + 4u * (loff + (uint32_t)i * nx)) = d[i];
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (xin && Y0 + i < ny) out[(uint64_t)(Y0 + i) * nx + X] = d[i];
    }
}

template <typename TCode, bool ZZ, bool DIRECT, bool PATCH>
__global__ void __launch_bounds__(256)
lz2d_inv_kernel(const TCode* __restrict__ codes, uint32_t nx, uint32_t ny, float ebx2,
                const uint32_t* __restrict__ off, const float* __restrict__ serr,
                const uint32_t* __restrict__ sloc, float* __restrict__ out, uint32_t gx, uint32_t T,
                DirectOut dout, DirectOut pin) {
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (tile >= T) return;
    const uint32_t bx = tile % gx, by = tile / gx;
    if (bx * 32 + 32 <= nx && by * 32 + 32 <= ny && nx < (1u << 25))
        lz2d_inv_tile<TCode, ZZ, DIRECT, PATCH, true>(codes, nx, ny, ebx2, off, serr, sloc, out, gx, tile, dout, pin);
    else
        lz2d_inv_tile<TCode, ZZ, DIRECT, PATCH, false>(codes, nx, ny, ebx2, off, serr, sloc, out, gx, tile, dout, pin);
}

// 3D inverse, the addition tree of lorenzo_quant_nd.cu:282-360: outlier +
// decoded code, a sequential prefix along y, then Hillis-Steele prefixes
// (offsets 1, 2, 4) along x and then along z, times ebx2.  One wave per
// pair of x-adjacent tiles (as the forward): lane (x, zl) = (l & 15,
// l >> 4) holds column 16 bx2 + x, rows y = 0..7 of planes zl and 4 + zl,
// so loads and stores cover 4 rows of contiguous bytes.  The y prefix is
// register arithmetic, the x prefix DPP inside each tile's 8 lanes of a
// 16-lane row, the z prefix lane shuffles by 16 and 32 (offsets 1, 2) and
// the other register (offset 4).  Outliers are added in LDS tiles
// ([tile][z][y][x], 72 floats per z plane), one entry per lane;
// strict-mode patches of a short list are stored after the tiles' values.
template <typename TCode, bool ZZ, bool DIRECT, bool PATCH>
__global__ void __launch_bounds__(256)
lz3d_inv_kernel(const TCode* __restrict__ codes, uint32_t nx, uint32_t ny, uint32_t nz, float ebx2,
                const uint32_t* __restrict__ off, const float* __restrict__ serr,
                const uint32_t* __restrict__ sloc, float* __restrict__ out, Tiles tg, uint32_t T, uint32_t P,
                DirectOut dout, DirectOut pin) {
    constexpr uint32_t ZS = 72;  // LDS floats per z plane of a tile
    __shared__ float s_tile[4][2 * 8 * ZS];
    const uint32_t pair = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (pair >= P) return;
    float* st = s_tile[wave_in_wg()];
    const int l = lane_id();
    const int x = l & 15, zl = l >> 4, xt = x & 7;
    const uint32_t gx2 = (tg.gx + 1) / 2;
    const uint32_t bx2 = pair % gx2, by = (pair / gx2) % tg.gy, bz = pair / (gx2 * tg.gy);
    const uint32_t ta = 2 * bx2 + tg.gx * (by + tg.gy * bz), tb = ta + 1;
    const bool has_b = 2 * bx2 + 1 < tg.gx;
    const uint32_t X = bx2 * 16 + x, Y0 = by * 8, Z0 = bz * 8;
    const uint64_t plane = (uint64_t)nx * ny;
    const uint64_t tbase = (uint64_t)Z0 * plane + (uint64_t)Y0 * nx + bx2 * 16;
    const bool interior = bx2 * 16 + 16 <= nx && Y0 + 8 <= ny && Z0 + 8 <= nz && plane < (1ull << 26);
    const bool xin = X < nx;
    const uint32_t lo = (uint32_t)zl * (uint32_t)plane + (uint32_t)x;  // lane offset in plane Z0 + zl, row Y0
    // -- decoded codes: v[y][h] = element (X, Y0 + y, Z0 + 4 h + zl)
    float v[8][2];
    if (interior) {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int y = 0; y < 8; y++)
                v[y][h] = dec_code<TCode, ZZ>(*gp(reinterpret_cast<const TCode*>(
                    reinterpret_cast<const char*>(codes + tbase + (uint64_t)(4 * h) * plane + (uint64_t)y * nx) +
                    (uint32_t)sizeof(TCode) * lo)));
    } else {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int y = 0; y < 8; y++) {
                const uint32_t Z = Z0 + 4 * h + zl, Y = Y0 + y;
                v[y][h] = (xin && Y < ny && Z < nz)
                              ? dec_code<TCode, ZZ>(gp(codes)[(uint64_t)Z * plane + (uint64_t)Y * nx + X])
                              : 0.0f;
            }
    }
    // -- outliers: value + decoded code, in the LDS tiles [t][z][y][x]
    const uint32_t lbase = (uint32_t)(x >> 3) * (8 * ZS) + (uint32_t)zl * ZS + (uint32_t)xt;  // + 4 h ZS + 8 y
    bool any = false;
    uint32_t a0 = 0, a1 = 0, b0 = 0, b1 = 0, kk = 0;
    if constexpr (DIRECT) {
        kk = dout.count ? min(*dout.count, dout.max_out) : 0u;
        any = kk > 0;
    } else {
        a0 = off[ta];
        a1 = off[ta + 1];
        if (has_b) {
            b0 = a1;
            b1 = off[tb + 1];
        }
        any = a1 > a0 || b1 > b0;
    }
    if (any) {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int y = 0; y < 8; y++) st[lbase + 4 * h * ZS + 8 * y] = v[y][h];
        wave_lds_sync();
        if constexpr (DIRECT) {
            for (uint32_t j = (uint32_t)l; j < kk; j += 64) {
                uint32_t t, loc;
                if (tile_of(dout.m, dout.idx[j], t, loc) && (t == ta || (has_b && t == tb))) {
                    const uint32_t a = (t == ta ? 0u : 8 * ZS) + (loc >> 6) * ZS + (loc & 63u);
                    st[a] = dout.err[j] + st[a];
                }
            }
        } else {
            for (uint32_t j = a0 + (uint32_t)l; j < a1; j += 64) {
                const uint32_t loc = sloc[j];
                const uint32_t a = (loc >> 6) * ZS + (loc & 63u);
                st[a] = serr[j] + st[a];
            }
            for (uint32_t j = b0 + (uint32_t)l; j < b1; j += 64) {
                const uint32_t loc = sloc[j];
                const uint32_t a = 8 * ZS + (loc >> 6) * ZS + (loc & 63u);
                st[a] = serr[j] + st[a];
            }
        }
        wave_lds_sync();
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int y = 0; y < 8; y++) v[y][h] = st[lbase + 4 * h * ZS + 8 * y];
    }
    // -- y prefix (sequential, registers)
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int y = 1; y < 8; y++) v[y][h] += v[y - 1][h];
    // -- x prefix (offsets 1, 2, 4 inside each tile's 8-lane group of a 16-lane row)
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int y = 0; y < 8; y++) {
            float w = v[y][h];
            float u = __uint_as_float(dpp_u32<0x111>(__float_as_uint(w)));  // lane x - 1 (row_shr:1)
            if (xt >= 1) w += u;
            u = __uint_as_float(dpp_u32<0x112>(__float_as_uint(w)));        // x - 2
            if (xt >= 2) w += u;
            u = __uint_as_float(dpp_u32<0x114>(__float_as_uint(w)));        // x - 4
            if (xt >= 4) w += u;
            v[y][h] = w;
        }
    // -- z prefix, offsets 1, 2, 4 (each stage adds the previous stage's
    // values): z = 4 h + zl, lanes 16 apart per plane; offset 4 is h - 1
    const int src1 = zl == 0 ? l + 48 : l - 16;  // z - 1 for h = 1: (h 0, zl 3) or (h 1, zl - 1)
    const int src2 = zl < 2 ? l + 32 : l - 32;   // z - 2 for h = 1: (h 0, zl + 2) or (h 1, zl - 2)
#pragma unroll
    for (int y = 0; y < 8; y++) {
        float v0 = v[y][0], v1 = v[y][1];
        {
            const float u0 = __shfl_up(v0, 16, 64);
            const float u1 = __shfl(zl == 3 ? v0 : v1, src1, 64);
            v0 = zl >= 1 ? v0 + u0 : v0;
            v1 = v1 + u1;
        }
        {
            const float u0 = __shfl_up(v0, 32, 64);
            const float u1 = __shfl(zl >= 2 ? v0 : v1, src2, 64);
            v0 = zl >= 2 ? v0 + u0 : v0;
            v1 = v1 + u1;
        }
        v1 = v1 + v0;
        v[y][0] = v0;
        v[y][1] = v1;
    }
    if (interior) {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int y = 0; y < 8; y++)
                *gp(reinterpret_cast<float*>(reinterpret_cast<char*>(out + tbase + (uint64_t)(4 * h) * plane +
                                                                    (uint64_t)y * nx) + 4u * lo)) = v[y][h] * ebx2;
    } else {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int y = 0; y < 8; y++) {
                const uint32_t Z = Z0 + 4 * h + zl, Y = Y0 + y;
                if (xin && Y < ny && Z < nz) gp(out)[(uint64_t)Z * plane + (uint64_t)Y * nx + X] = v[y][h] * ebx2;
            }
    }
    // -- strict-mode patches of a short list: stored after the tiles' values
    // (a wave's stores to one address land in program order)
    if constexpr (PATCH) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // the tiles' stores first
        const uint32_t kp = pin.count ? min(*pin.count, pin.max_out) : 0u;
        for (uint32_t j = (uint32_t)l; j < kp; j += 64) {
            const uint32_t q = pin.idx[j];
            uint32_t t, loc;
            if (tile_of(pin.m, q, t, loc)) {
                if (t == ta || (has_b && t == tb)) out[q] = pin.err[j];
            } else if (ta == 0 && (uint64_t)q >= pin.m.total && (uint64_t)q < pin.n_out) {
                out[q] = pin.err[j];
            }
        }
    }
}

// Strict-mode patches: out[idx[i]] = val[i] for i < min(*count, cap).
__global__ void __launch_bounds__(256)
apply_patches_kernel(float* __restrict__ out, const float* __restrict__ val, const uint32_t* __restrict__ idx,
                     const uint32_t* __restrict__ count, uint64_t cap, uint64_t n_out) {
    const uint64_t k = min((uint64_t)*count, cap);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t j = idx[i];
        if (j < n_out) out[j] = val[i];
    }
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------
struct PatchOut {
    const float* d_params;
    float eb, ebx2;
    float* val;
    uint32_t* idx;
    uint32_t* count;
    size_t cap;
};

template <typename TCode, bool STRICT>
int launch_fwd(hipStream_t s, const float* in, size_t n, size_t nx, size_t ny, size_t nz, float r,
               const float* d_r, uint32_t radius, int zz, void* codes, float* oerr, uint32_t* oidx,
               uint32_t* ocount, size_t max_out, void* ws, size_t ws_bytes, const PatchOut& po) {
    const FwdWs L = fwd_layout(n, nx, ny, nz, STRICT);
    if (ws_bytes < L.total) return fail(FZH_E_CAPACITY, "lorenzo fwd: workspace too small");
    const size_t T = L.T;
    uint8_t* w = (uint8_t*)ws;
    TileLists tl{};
    tl.err = (float*)(w + L.o_err);
    tl.idx = (uint32_t*)(w + L.o_idx);
    tl.cnt = (uint32_t*)(w + L.o_cnt);
    if (STRICT) {
        tl.pval = (float*)(w + L.o_pval);
        tl.pidx = (uint32_t*)(w + L.o_pidx);
        tl.pcnt = (uint32_t*)(w + L.o_pcnt);
        tl.d_params = po.d_params;
        tl.eb = po.eb;
        tl.ebx2 = po.ebx2;
    }
    const float rad = (float)(TCode)radius;
    const unsigned blocks = (unsigned)cdiv(T, 4);
    TCode* c = (TCode*)codes;
    const int nd = ndim_of(ny, nz);
    const size_t nv = (nd == 1 && nx > 0 && nx < n) ? nx : n;
    const size_t P3 = nd == 3 ? cdiv(cdiv(nx, 8), 2) * cdiv(ny, 8) * cdiv(nz, 8) : 0;  // 3D tile pairs
#define FZH_LZ_FWD(ZZV)                                                                                 \
    if (nd == 1)                                                                                        \
        launch_k("lz1d_fwd_kernel", lz1d_fwd_kernel<TCode, ZZV, STRICT>, dim3(blocks), dim3(256), 0, s, in, n, \
                 nv, r, d_r, rad, c, tl, (uint32_t)T);                                                  \
    else if (nd == 2)                                                                                   \
        launch_k("lz2d_fwd_kernel", lz2d_fwd_kernel<TCode, ZZV, STRICT>, dim3(blocks), dim3(256), 0, s, in, \
                 (uint32_t)nx, (uint32_t)ny, n, r, d_r, rad, c, tl, (uint32_t)cdiv(nx, 32), (uint32_t)T);  \
    else                                                                                                \
        launch_k("lz3d_fwd_kernel", lz3d_fwd_kernel<TCode, ZZV, STRICT>, dim3((unsigned)cdiv(P3, 4)), dim3(256), 0, \
                 s, in, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz, n, r, d_r, rad, c, tl,               \
                 Tiles{(uint32_t)cdiv(nx, 8), (uint32_t)cdiv(ny, 8), 0}, (uint32_t)T, (uint32_t)P3);
    if (zz) {
        FZH_LZ_FWD(true)
    } else {
        FZH_LZ_FWD(false)
    }
#undef FZH_LZ_FWD
    FZH_CHECK_LAUNCH("lorenzo fwd");
    CompactList L0{tl.err, tl.idx, tl.cnt, oerr, oidx, ocount, max_out};
    CompactList L1{tl.pval, tl.pidx, tl.pcnt, po.val, po.idx, po.count, po.cap};
    if (int rc = launch_compact<float>(s, w + L.o_ctx, T, (uint32_t)L.E, L0, STRICT ? &L1 : nullptr)) return rc;
    return FZH_OK;
}

// Inverse, part 2: the tiles, with the binned outliers of part 1.
template <typename TCode>
int inv_apply(hipStream_t s, const void* codes, size_t n, size_t nx, size_t ny, size_t nz, const float* err,
              const uint32_t* idx, const uint32_t* count, size_t max_out, float ebx2, int zz, float* out, void* ws,
              size_t ws_bytes, const DirectOut* patches = nullptr) {
    const InvWs L = inv_layout(n, nx, ny, nz, max_out);
    if (ws_bytes < L.total) return fail(FZH_E_CAPACITY, "lorenzo inv: workspace too small");
    uint8_t* w = (uint8_t*)ws;
    const uint32_t* off = (const uint32_t*)(w + L.o_off);
    const float* serr = (const float*)(w + L.o_err);
    const uint32_t* sloc = (const uint32_t*)(w + L.o_loc);
    const int nd = ndim_of(ny, nz);
    const TileMap m = tile_map(n, nx, ny, nz);
    const unsigned blocks = (unsigned)cdiv(L.T, 4);
    const TCode* c = (const TCode*)codes;
    const bool direct = max_out <= kDirectOutliers;
    const size_t P3 = nd == 3 ? cdiv(cdiv(nx, 8), 2) * cdiv(ny, 8) * cdiv(nz, 8) : 0;  // 3D tile pairs
    const DirectOut dout{err, idx, (err && idx) ? count : nullptr, (uint32_t)max_out, m, 0};
    // short patch lists ride along with short outlier lists (one launch)
    const bool fold = patches && patches->max_out > 0 && direct && patches->max_out <= kDirectOutliers;
    DirectOut pin{};
    if (fold) pin = DirectOut{patches->err, patches->idx, patches->count, patches->max_out, m, patches->n_out};
#define FZH_LZ_INV(ZZV, DV, PV)                                                                         \
    if (nd == 1)                                                                                        \
        launch_k("lz1d_inv_kernel", lz1d_inv_kernel<TCode, ZZV, DV, PV>, dim3(blocks), dim3(256), 0, s, c, n, \
                 ebx2, off, serr, sloc, out, (uint32_t)L.T, dout, pin);                                 \
    else if (nd == 2)                                                                                   \
        launch_k("lz2d_inv_kernel", lz2d_inv_kernel<TCode, ZZV, DV, PV>, dim3(blocks), dim3(256), 0, s, c, \
                 (uint32_t)nx, (uint32_t)ny, ebx2, off, serr, sloc, out, m.gx, (uint32_t)L.T, dout, pin); \
    else                                                                                                \
        launch_k("lz3d_inv_kernel", lz3d_inv_kernel<TCode, ZZV, DV, PV>, dim3((unsigned)cdiv(P3, 4)), dim3(256), 0, s, \
                 c, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz, ebx2, off, serr, sloc, out, Tiles{m.gx, m.gy, 0}, \
                 (uint32_t)L.T, (uint32_t)P3, dout, pin);
    if (zz) {
        if (fold) { FZH_LZ_INV(true, true, true) }
        else if (direct) { FZH_LZ_INV(true, true, false) }
        else { FZH_LZ_INV(true, false, false) }
    } else {
        if (fold) { FZH_LZ_INV(false, true, true) }
        else if (direct) { FZH_LZ_INV(false, true, false) }
        else { FZH_LZ_INV(false, false, false) }
    }
#undef FZH_LZ_INV
    FZH_CHECK_LAUNCH("lorenzo inv");
    if (patches && patches->max_out > 0 && !fold) {
        const unsigned pb = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(patches->max_out, 256), 1024));
        launch_k("apply_patches_kernel", apply_patches_kernel, dim3(pb), dim3(256), 0, s, out, patches->err,
                 patches->idx, patches->count, (uint64_t)patches->max_out, patches->n_out);
        FZH_CHECK_LAUNCH("apply_patches_kernel");
    }
    return FZH_OK;
}

// Pass 1 writes one (min, max) pair per workgroup into `partial` (every
// pair pass 2 reads is written by the same call, so nothing needs clearing);
// about 16 float4 per thread, at most VB_MAX_PARTIALS workgroups.
int value_base_launch(hipStream_t s, const float* d_in, size_t n, int mode, float user_eb, float* d_params,
                      float2* partial) {
    const bool zero_pad = (mode & FZH_VB_ZERO_PADDED) != 0;
    mode &= ~FZH_VB_ZERO_PADDED;
    if (mode != 1 && mode != 2) return fail(FZH_E_INVALID, "value base: mode must be 1 (REL) or 2 (NOA)");
    if (n == 0) return fail(FZH_E_INVALID, "value base: empty input");
    if (!d_in || !d_params) return fail(FZH_E_INVALID, "value base: null pointer");
    if (reinterpret_cast<uintptr_t>(d_in) % 16 != 0) return fail(FZH_E_INVALID, "value base: input must be 16-byte aligned");
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(n, 256 * 64), VB_MAX_PARTIALS));
    launch_k("minmax_partial_kernel", minmax_partial_kernel, dim3(blocks), dim3(256), 0, s, d_in, n, partial,
             zero_pad);
    launch_k("minmax_final_kernel", minmax_final_kernel, dim3(1), dim3(256), 0, s, partial, blocks, mode, user_eb,
             d_params);
    FZH_CHECK_LAUNCH("minmax");
    return FZH_OK;
}

}  // namespace
}  // namespace fzh

// The partials of fzh_value_base_f32_noalloc: one fixed array per GPU, held
// until the process ends.  Host code never passes this name to a kernel (on
// the host it is only a handle); it resolves the array's device address per
// device with hipGetSymbolAddress and uses that.
static __device__ __align__(256) float2 scratch[fzh::VB_MAX_PARTIALS];

namespace fzh {
namespace {

constexpr int VB_MAX_DEVICES = 64;
std::mutex g_vb_mu;
float2* g_vb_scratch[VB_MAX_DEVICES] = {};

// Device address of `scratch` on the device that runs stream s (the current
// device for the null stream), resolved once per device with that device
// current.
int vb_scratch_address(hipStream_t s, float2** out) {
    int cur = 0;
    FZH_TRY_HIP(hipGetDevice(&cur));
    int dev = cur;
    if (s) FZH_TRY_HIP(hipStreamGetDevice(s, &dev));
    if (dev < 0 || dev >= VB_MAX_DEVICES) return fail(FZH_E_INVALID, "value base (no-alloc): device ordinal out of range");
    std::lock_guard<std::mutex> g(g_vb_mu);
    if (!g_vb_scratch[dev]) {
        if (dev != cur) FZH_TRY_HIP(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (dev != cur) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(FZH_E_RUNTIME, std::string("value base (no-alloc): hipGetSymbolAddress(scratch): ") +
                                           hipGetErrorString(e));
        g_vb_scratch[dev] = static_cast<float2*>(p);
    }
    *out = g_vb_scratch[dev];
    return FZH_OK;
}

}  // namespace
}  // namespace fzh

using namespace fzh;

extern "C" {

size_t fzh_lorenzo_quant_fwd_ws_bytes(size_t n, size_t nx, size_t ny, size_t nz) {
    return fwd_layout(n, nx, ny, nz, false).total;
}

size_t fzh_lorenzo_quant_fwd_strict_ws_bytes(size_t n, size_t nx, size_t ny, size_t nz) {
    return fwd_layout(n, nx, ny, nz, true).total;
}

int fzh_lorenzo_quant_fwd_f32(void* stream, const float* d_in, size_t n, size_t nx, size_t ny, size_t nz,
                              float ebx2_r, const float* d_ebx2_r, uint32_t radius, int zigzag,
                              int code_bytes, void* d_codes, float* d_outlier_err, uint32_t* d_outlier_idx,
                              uint32_t* d_outlier_count, size_t max_outliers, void* d_ws, size_t ws_bytes) {
    if (int rc = check_dims(n, nx, ny, nz, "lorenzo fwd")) return rc;
    if (n == 0) return FZH_OK;
    if (!d_in || !d_codes || !d_outlier_count) return fail(FZH_E_INVALID, "lorenzo fwd: null pointer");
    if (max_outliers > 0 && (!d_outlier_err || !d_outlier_idx))
        return fail(FZH_E_INVALID, "lorenzo fwd: null outlier buffers");
    if (ndim_of(ny, nz) == 1 && reinterpret_cast<uintptr_t>(d_in) % 16 != 0)
        return fail(FZH_E_INVALID, "lorenzo fwd: input must be 16-byte aligned");
    const PatchOut none{};
    switch (code_bytes) {
    case 1:
        return launch_fwd<uint8_t, false>(S(stream), d_in, n, nx, ny, nz, ebx2_r, d_ebx2_r, radius, zigzag, d_codes,
                                          d_outlier_err, d_outlier_idx, d_outlier_count, max_outliers, d_ws,
                                          ws_bytes, none);
    case 2:
        return launch_fwd<uint16_t, false>(S(stream), d_in, n, nx, ny, nz, ebx2_r, d_ebx2_r, radius, zigzag,
                                           d_codes, d_outlier_err, d_outlier_idx, d_outlier_count, max_outliers,
                                           d_ws, ws_bytes, none);
    default:
        return fail(FZH_E_INVALID, "lorenzo fwd: code_bytes must be 1 or 2");
    }
}

int fzh_lorenzo_quant_fwd_strict_f32(void* stream, const float* d_in, size_t n, size_t nx, size_t ny, size_t nz,
                                     float ebx2_r, const float* d_ebx2_r, uint32_t radius, int zigzag,
                                     int code_bytes, void* d_codes, float* d_outlier_err, uint32_t* d_outlier_idx,
                                     uint32_t* d_outlier_count, size_t max_outliers, float eb, float ebx2,
                                     const float* d_eb_params, float* d_patch_val, uint32_t* d_patch_idx,
                                     uint32_t* d_patch_count, size_t max_patches, void* d_ws, size_t ws_bytes) {
    if (int rc = check_dims(n, nx, ny, nz, "lorenzo fwd (strict)")) return rc;
    if (n == 0) return FZH_OK;
    if (!d_in || !d_codes || !d_outlier_count || !d_patch_count) return fail(FZH_E_INVALID, "lorenzo fwd (strict): null pointer");
    if (max_outliers > 0 && (!d_outlier_err || !d_outlier_idx))
        return fail(FZH_E_INVALID, "lorenzo fwd (strict): null outlier buffers");
    if (max_patches > 0 && (!d_patch_val || !d_patch_idx))
        return fail(FZH_E_INVALID, "lorenzo fwd (strict): null patch buffers");
    if (ndim_of(ny, nz) == 1 && reinterpret_cast<uintptr_t>(d_in) % 16 != 0)
        return fail(FZH_E_INVALID, "lorenzo fwd (strict): input must be 16-byte aligned");
    PatchOut po{};
    po.d_params = d_eb_params;
    po.eb = eb;
    po.ebx2 = ebx2;
    po.val = d_patch_val;
    po.idx = d_patch_idx;
    po.count = d_patch_count;
    po.cap = max_patches;
    switch (code_bytes) {
    case 1:
        return launch_fwd<uint8_t, true>(S(stream), d_in, n, nx, ny, nz, ebx2_r, d_ebx2_r, radius, zigzag, d_codes,
                                         d_outlier_err, d_outlier_idx, d_outlier_count, max_outliers, d_ws,
                                         ws_bytes, po);
    case 2:
        return launch_fwd<uint16_t, true>(S(stream), d_in, n, nx, ny, nz, ebx2_r, d_ebx2_r, radius, zigzag,
                                          d_codes, d_outlier_err, d_outlier_idx, d_outlier_count, max_outliers,
                                          d_ws, ws_bytes, po);
    default:
        return fail(FZH_E_INVALID, "lorenzo fwd (strict): code_bytes must be 1 or 2");
    }
}

int fzh_lorenzo_apply_patches(void* stream, float* d_out, const float* d_patch_val, const uint32_t* d_patch_idx,
                              const uint32_t* d_patch_count, size_t max_patches, size_t n_out) {
    if (max_patches == 0) return FZH_OK;
    if (!d_out || !d_patch_val || !d_patch_idx || !d_patch_count)
        return fail(FZH_E_INVALID, "lorenzo patches: null pointer");
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(max_patches, 256), 1024));
    launch_k("apply_patches_kernel", apply_patches_kernel, dim3(blocks), dim3(256), 0, S(stream), d_out, d_patch_val,
             d_patch_idx, d_patch_count, max_patches, n_out);
    FZH_CHECK_LAUNCH("apply_patches_kernel");
    return FZH_OK;
}

size_t fzh_lorenzo_quant_inv_ws_bytes(size_t n, size_t nx, size_t ny, size_t nz, size_t max_outliers) {
    return inv_layout(n, nx, ny, nz, max_outliers).total;
}

int fzh_lorenzo_quant_inv_bin(void* stream, size_t n, size_t nx, size_t ny, size_t nz, const float* d_outlier_err,
                              const uint32_t* d_outlier_idx, const uint32_t* d_outlier_count, size_t max_outliers,
                              void* d_ws, size_t ws_bytes) {
    if (int rc = check_dims(n, nx, ny, nz, "lorenzo inv")) return rc;
    if (n == 0) return FZH_OK;
    if (max_outliers > 0 && d_outlier_count && (!d_outlier_err || !d_outlier_idx))
        return fail(FZH_E_INVALID, "lorenzo inv: null outlier buffers");
    return inv_bin(S(stream), n, nx, ny, nz, d_outlier_err, d_outlier_idx, d_outlier_count, max_outliers, d_ws,
                   ws_bytes);
}

int fzh_lorenzo_quant_inv_apply_f32(void* stream, const void* d_codes, size_t n, size_t nx, size_t ny, size_t nz,
                                    const float* d_outlier_err, const uint32_t* d_outlier_idx,
                                    const uint32_t* d_outlier_count, size_t max_outliers, float ebx2, int zigzag,
                                    int code_bytes, float* d_out, void* d_ws, size_t ws_bytes) {
    if (int rc = check_dims(n, nx, ny, nz, "lorenzo inv")) return rc;
    if (n == 0) return FZH_OK;
    if (!d_codes || !d_out) return fail(FZH_E_INVALID, "lorenzo inv: null pointer");
    if (ndim_of(ny, nz) == 1 &&
        (reinterpret_cast<uintptr_t>(d_out) % 16 != 0 || reinterpret_cast<uintptr_t>(d_codes) % 8 != 0))
        return fail(FZH_E_INVALID, "lorenzo inv: 1D buffers must be 16/8-byte aligned");
    switch (code_bytes) {
    case 1:
        return inv_apply<uint8_t>(S(stream), d_codes, n, nx, ny, nz, d_outlier_err, d_outlier_idx, d_outlier_count,
                                  max_outliers, ebx2, zigzag, d_out, d_ws, ws_bytes);
    case 2:
        return inv_apply<uint16_t>(S(stream), d_codes, n, nx, ny, nz, d_outlier_err, d_outlier_idx,
                                   d_outlier_count, max_outliers, ebx2, zigzag, d_out, d_ws, ws_bytes);
    default:
        return fail(FZH_E_INVALID, "lorenzo inv: code_bytes must be 1 or 2");
    }
}

int fzh_lorenzo_quant_inv_f32(void* stream, const void* d_codes, size_t n, size_t nx, size_t ny, size_t nz,
                              const float* d_outlier_err, const uint32_t* d_outlier_idx,
                              const uint32_t* d_outlier_count, size_t max_outliers, float ebx2, int zigzag,
                              int code_bytes, float* d_out, void* d_ws, size_t ws_bytes) {
    if (code_bytes != 1 && code_bytes != 2) return fail(FZH_E_INVALID, "lorenzo inv: code_bytes must be 1 or 2");
    if (!d_codes || !d_out) {
        if (int rc = check_dims(n, nx, ny, nz, "lorenzo inv")) return rc;
        if (n == 0) return FZH_OK;
        return fail(FZH_E_INVALID, "lorenzo inv: null pointer");
    }
    if (int rc = fzh_lorenzo_quant_inv_bin(stream, n, nx, ny, nz, d_outlier_err, d_outlier_idx, d_outlier_count,
                                           max_outliers, d_ws, ws_bytes))
        return rc;
    return fzh_lorenzo_quant_inv_apply_f32(stream, d_codes, n, nx, ny, nz, d_outlier_err, d_outlier_idx,
                                           d_outlier_count, max_outliers, ebx2, zigzag, code_bytes, d_out, d_ws,
                                           ws_bytes);
}

int fzh_lorenzo_quant_inv_strict_f32(void* stream, const void* d_codes, size_t n, size_t nx, size_t ny, size_t nz,
                                     const float* d_outlier_err, const uint32_t* d_outlier_idx,
                                     const uint32_t* d_outlier_count, size_t max_outliers, float ebx2, int zigzag,
                                     int code_bytes, const float* d_patch_val, const uint32_t* d_patch_idx,
                                     const uint32_t* d_patch_count, size_t max_patches, float* d_out, void* d_ws,
                                     size_t ws_bytes) {
    if (code_bytes != 1 && code_bytes != 2) return fail(FZH_E_INVALID, "lorenzo inv: code_bytes must be 1 or 2");
    if (int rc = check_dims(n, nx, ny, nz, "lorenzo inv")) return rc;
    if (n == 0) return FZH_OK;
    if (!d_codes || !d_out) return fail(FZH_E_INVALID, "lorenzo inv: null pointer");
    if (max_patches > 0 && (!d_patch_val || !d_patch_idx || !d_patch_count))
        return fail(FZH_E_INVALID, "lorenzo patches: null pointer");
    if (max_patches > 0xFFFFFFFFull) return fail(FZH_E_INVALID, "lorenzo patches: more than 2^32 entries");
    if (ndim_of(ny, nz) == 1 &&
        (reinterpret_cast<uintptr_t>(d_out) % 16 != 0 || reinterpret_cast<uintptr_t>(d_codes) % 8 != 0))
        return fail(FZH_E_INVALID, "lorenzo inv: 1D buffers must be 16/8-byte aligned");
    if (int rc = fzh_lorenzo_quant_inv_bin(stream, n, nx, ny, nz, d_outlier_err, d_outlier_idx, d_outlier_count,
                                           max_outliers, d_ws, ws_bytes))
        return rc;
    const DirectOut pl{d_patch_val, d_patch_idx, d_patch_count, (uint32_t)max_patches, TileMap{}, (uint64_t)n};
    return code_bytes == 1
               ? inv_apply<uint8_t>(S(stream), d_codes, n, nx, ny, nz, d_outlier_err, d_outlier_idx, d_outlier_count,
                                    max_outliers, ebx2, zigzag, d_out, d_ws, ws_bytes, &pl)
               : inv_apply<uint16_t>(S(stream), d_codes, n, nx, ny, nz, d_outlier_err, d_outlier_idx,
                                     d_outlier_count, max_outliers, ebx2, zigzag, d_out, d_ws, ws_bytes, &pl);
}

size_t fzh_value_base_ws_bytes(size_t n) {
    (void)n;
    return 256 + 8 * 1024;
}

int fzh_value_base_f32(void* stream, const float* d_in, size_t n, int mode, float user_eb, float* d_params,
                       void* d_ws, size_t ws_bytes) {
    if (ws_bytes < fzh_value_base_ws_bytes(n)) return fail(FZH_E_CAPACITY, "value base: workspace too small");
    if (!d_ws) return fail(FZH_E_INVALID, "value base: null workspace");
    return value_base_launch(S(stream), d_in, n, mode, user_eb, d_params, (float2*)((uint8_t*)d_ws + 256));
}

int fzh_value_base_f32_noalloc(void* stream, const float* d_in, size_t n, int mode, float user_eb,
                               float* d_params) {
    float2* partial = nullptr;
    if (int rc = vb_scratch_address(S(stream), &partial)) return rc;
    return value_base_launch(S(stream), d_in, n, mode, user_eb, d_params, partial);
}

}  // extern "C"

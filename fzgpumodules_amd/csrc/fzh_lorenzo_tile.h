// fzh_lorenzo_tile.h — one Lorenzo predict+quantize tile on one wave64.
//
// Shared by the single-field kernels (lorenzo.hip) and the batched
// persistent kernels (batch.hip), so a field's codes and outlier lists are
// the same whichever launch produced them.
//
// Format geometry follows the reference exactly: 1D tiles of 1024 elements
// (modules/fused/lorenzo_quant/lorenzo_quant.cu:262-364), 2D tiles of 32x32
// (lorenzo_quant_nd.cu:29-105), 3D tiles of 8x8x8 (lorenzo_quant_nd.cu:199-274).
// Lane layouts are chosen for wave64:
//   1D  lane l, step q: elements 4(l+64q) .. 4(l+64q)+3 of the tile
//   2D  lane l: column x = l&31, rows 16*(l>>5) .. +15 of the tile
//   3D  lane l: x = l&7, z = l>>3, rows y = 0..7 of the tile
// The arithmetic per element is the reference's (same FP32 operation order).
// A tile's outliers (and strict-mode patches) are listed in element-index
// order inside the tile.
#pragma once

#include "fzh_internal.h"

namespace fzh {
namespace {

template <typename TCode> struct SignedOf;
template <> struct SignedOf<uint8_t> { using type = int8_t; };
template <> struct SignedOf<uint16_t> { using type = int16_t; };

// (TCode)(int)delta, or zigzag((SCode)(int)delta)   (lorenzo_quant.cu:328-334)
template <typename TCode, bool ZZ>
__device__ __forceinline__ TCode enc_code(float d) {
    using SC = typename SignedOf<TCode>::type;
    constexpr int W = 8 * sizeof(TCode);
    const int q = (int)d;
    if constexpr (ZZ) {
        const SC s = (SC)q;
        return (TCode)(((TCode)s << 1) ^ (TCode)(s >> (W - 1)));
    } else {
        return (TCode)q;
    }
}

// (TIn)(SCode)code, or zigzag-decode  (lorenzo_quant.cu:195-202)
template <typename TCode, bool ZZ>
__device__ __forceinline__ float dec_code(TCode c) {
    using SC = typename SignedOf<TCode>::type;
    if constexpr (ZZ) {
        return (float)(SC)((TCode)(c >> 1) ^ (TCode)(-(SC)(c & 1u)));
    } else {
        return (float)(SC)c;
    }
}

// round(x * r), halves away from zero (lorenzo_quant.cu:120-135 uses
// round()).  For every float v, round(v) == trunc(v + copysign(c, v)) with c
// the largest float below 1/2 under round-to-nearest-even addition: a
// fraction below 1/2 stays below the next integer (its distance to it is at
// least one ulp of v plus 2^-25), a fraction of 1/2 or more reaches it, and
// |v| >= 2^23 is unchanged -- three instructions where roundf takes six.
// Checked against round() over all 2^32 floats.  The product is rounded
// first (no fused multiply-add: the build disables contraction).
__device__ __forceinline__ float quant_round(float x, float r) {
    const float v = x * r;
    return __builtin_truncf(v + __builtin_copysignf(0.49999997f, v));
}

// Strict error bound (an extension; no counterpart in the reference).  The
// inverse reconstructs element i as fl(S_i * ebx2), S_i the exact integer
// prefix of the tile's deltas, whenever every partial sum of the inverse's
// addition tree is exact.  Partial sums are sums of deltas over sub-boxes of
// the tile, i.e. signed sums of at most 8 quantised values, so |s| < 2^21
// across the tile keeps them below 2^24 (exact in FP32).  An element is
// patched (its original value stored) when |fl(s*ebx2) - x| > eb, or when
// its tile breaks the exactness condition (then every element of the tile).
constexpr float kExactS = 2097152.0f;  // 2^21

__device__ __forceinline__ bool needs_patch(float x, float s, float eb, float ebx2) {
    const float xh = s * ebx2;
    const double e = fabs((double)xh - (double)x);
    return !(e <= (double)eb);  // NaN compares false: patched
}

// Destination of one tile's list: entry j of the tile's list goes to
// val/idx[j] when j < cap (entries past cap are counted, not stored).
struct ListOut {
    float* val;
    uint32_t* idx;
    uint32_t cap;
};

struct TileCounts {
    uint32_t cnt;   // outliers of the tile
    uint32_t pcnt;  // strict-mode patches of the tile
};

// Strict-mode bound of a tile (eb, 2*eb as the inverse uses them).
struct StrictEb {
    float eb, ebx2;
};

// Margins for deciding |fl(s*ebx2) - x| <= eb from the FP32 difference:
// the rounded difference is within a relative 2^-24 of the exact one, so
// below lo (<= eb (1 - 2^-23)) the exact difference is within eb and above
// hi (>= eb (1 + 2^-23)) it is not; between them (and for NaN) the exact
// FP64 test decides.
struct StrictBand {
    float lo, hi;
};

__device__ __forceinline__ StrictBand strict_band(StrictEb se) {
    return StrictBand{se.eb * 0.99999976f, se.eb * 1.00000024f};  // eb (1 -+ 2^-22), rounded
}

__device__ __forceinline__ bool needs_patch_fast(float x, float s, StrictEb se, StrictBand b) {
    const float ef = fabsf(s * se.ebx2 - x);
    if (ef < b.lo) return false;
    if (ef > b.hi) return true;
    return needs_patch(x, s, se.eb, se.ebx2);
}

// The same test split for wave-wide use: `sure` = patched by the FP32 test
// alone; `unsure` = the FP32 difference is within [lo, hi] or NaN, where the
// exact FP64 test decides.  needs_patch_fast == sure || (unsure && needs_patch).
__device__ __forceinline__ void patch_fast(float x, float s, StrictEb se, StrictBand b, bool& sure, bool& unsure) {
    const float ef = fabsf(s * se.ebx2 - x);
    sure = ef > b.hi;
    unsure = !(ef < b.lo) && !sure;
}

// Zero codes[from, to) (tail beyond nx*ny[*nz] when the input was padded).
template <typename TCode>
__device__ __forceinline__ void zero_code_tail(TCode* codes, uint64_t from, uint64_t to) {
    for (uint64_t i = from + lane_id(); i < to; i += 64) gp(codes)[i] = (TCode)0;
}

// ---------------------------------------------------------------------------
// 1D: tile `tile` of a field of n (padded) elements, nv of them in memory
// (elements in [nv, n) read as zero).  codes == nullptr: lists only.
// ---------------------------------------------------------------------------
template <typename TCode, bool ZZ, bool STRICT>
__device__ __forceinline__ TileCounts lz1d_tile(const float* __restrict__ in, uint64_t n, uint64_t nv,
                                                uint32_t tile, float r, float rad, TCode* __restrict__ codes,
                                                ListOut lo, ListOut lp, StrictEb se) {
    const int l = lane_id();
    const uint64_t base = (uint64_t)tile * 1024;
    const bool full = base + 1024 <= nv;
    float s[4][4], xv[4][4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint64_t e0 = base + 4 * (uint64_t)(l + 64 * q);
        if (full) {
            const float4 v = gld_f4(in + e0);
            xv[q][0] = v.x;
            xv[q][1] = v.y;
            xv[q][2] = v.z;
            xv[q][3] = v.w;
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) xv[q][k] = (e0 + k < nv) ? gp(in)[e0 + k] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) s[q][k] = quant_round(xv[q][k], r);
    }
    // delta_i = s_i - s_{i-1}, chain restarts at the tile start.
    float d[4][4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float prev = wave_shr1(s[q][3]);
        const float last_prev_q = lane63(q > 0 ? s[q > 0 ? q - 1 : 0][3] : 0.0f);
        if (l == 0) prev = (q > 0) ? last_prev_q : 0.0f;
        d[q][0] = s[q][0] - prev;
#pragma unroll
        for (int k = 1; k < 4; k++) d[q][k] = s[q][k] - s[q][k - 1];
    }
    TCode c[4][4];
    uint32_t mask = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint64_t e = base + 4 * (uint64_t)(l + 64 * q) + k;
            const bool ok = fabsf(d[q][k]) < rad;
            c[q][k] = ok ? enc_code<TCode, ZZ>(d[q][k]) : (TCode)0;
            mask |= (uint32_t)(!ok && e < n) << (q * 4 + k);
        }
    if (codes) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint64_t e0 = base + 4 * (uint64_t)(l + 64 * q);
            if (base + 1024 <= n) {
                if constexpr (sizeof(TCode) == 2) {
                    uint2 w;
                    w.x = (uint32_t)c[q][0] | ((uint32_t)c[q][1] << 16);
                    w.y = (uint32_t)c[q][2] | ((uint32_t)c[q][3] << 16);
                    gst_u2(codes + e0, w);
                } else {
                    *gp(reinterpret_cast<uint32_t*>(codes + e0)) =
                        (uint32_t)c[q][0] | ((uint32_t)c[q][1] << 8) | ((uint32_t)c[q][2] << 16) |
                        ((uint32_t)c[q][3] << 24);
                }
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (e0 + k < n) gp(codes)[e0 + k] = c[q][k];
            }
        }
    }
    // Tile-local lists in index order: key (q, lane, k).
    auto emit = [&](uint32_t m, bool patch, const ListOut& o) {
        uint32_t run = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t mq = (m >> (4 * q)) & 0xFu;
            const uint32_t cq = __popc(mq);
            const uint32_t incl = wave_incl_scan(cq);
            uint32_t pos = run + incl - cq;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (mq & (1u << k)) {
                    if (pos < o.cap) {
                        gp(o.val)[pos] = patch ? xv[q][k] : d[q][k];
                        gp(o.idx)[pos] = (uint32_t)(base + 4 * (uint64_t)(l + 64 * q) + k);
                    }
                    pos++;
                }
            run += lane63(incl);
        }
        return run;
    };
    TileCounts tc{emit(mask, false, lo), 0u};
    if constexpr (STRICT) {
        bool huge = false;
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int k = 0; k < 4; k++) huge |= !(fabsf(s[q][k]) < kExactS);
        const bool tile_huge = __ballot(huge) != 0ull;
        const StrictBand band = strict_band(se);
        uint32_t pm = 0;
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint64_t e = base + 4 * (uint64_t)(l + 64 * q) + k;
                if (e < nv && (tile_huge || needs_patch_fast(xv[q][k], s[q][k], se, band))) pm |= 1u << (q * 4 + k);
            }
        tc.pcnt = emit(pm, true, lp);
    }
    return tc;
}

// ---------------------------------------------------------------------------
// 2D: tile (tile % gx, tile / gx) of an nx x ny field; `last` zeroes the code
// tail [nx*ny, n).  codes == nullptr: lists only.
//
// Interior tiles (the tile lies inside the field and its rows span less
// than 2^32 bytes) run a variant with no bounds tests whose loads and code
// stores address from the tile's scalar base with 32-bit lane offsets.
// ---------------------------------------------------------------------------
// Whether a 2D tile lies inside the field with rows spanning less than 2^32
// bytes (32-bit lane offsets from the tile's scalar base).
__device__ __forceinline__ bool lz2d_interior(uint32_t nx, uint32_t ny, uint32_t bx, uint32_t by) {
    return bx * 32 + 32 <= nx && by * 32 + 32 <= ny && nx < (1u << 25);
}

// The 16 values of tile `tile` this lane holds (column l & 31, rows
// 16 (l >> 5) .. +15 of the tile; zero outside the field).
template <bool INTERIOR>
__device__ __forceinline__ void lz2d_load(const float* __restrict__ in, uint32_t nx, uint32_t ny, uint32_t tile,
                                          uint32_t gx, float (&xv)[16]) {
    const int l = lane_id();
    const int x = l & 31, h = l >> 5;
    const uint32_t bx = tile % gx, by = tile / gx;
    const uint32_t X = bx * 32 + x;
    const uint32_t Y0 = by * 32 + 16 * h;
    const uint64_t tbase = (uint64_t)(by * 32) * nx + bx * 32;
    if constexpr (INTERIOR) {
        // row i of both halves from the row's scalar base: no per-load address arithmetic
        const uint32_t lb = 4u * ((uint32_t)(16 * h) * nx + (uint32_t)x);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const char* rp = reinterpret_cast<const char*>(in + tbase + (uint64_t)i * nx);
            xv[i] = *gp(reinterpret_cast<const float*>(rp + lb));
        }
    } else {
        const bool xin = X < nx;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t Y = Y0 + i;
            xv[i] = (xin && Y < ny) ? gp(in)[(uint64_t)Y * nx + X] : 0.0f;
        }
    }
}

template <typename TCode, bool ZZ, bool STRICT, bool INTERIOR, bool CODES>
__device__ __forceinline__ TileCounts lz2d_tile_c(const float (&xin16)[16], uint32_t nx, uint32_t ny, uint64_t n,
                                                  uint32_t tile, uint32_t gx, bool last, float r, float rad,
                                                  TCode* __restrict__ codes, ListOut lo, ListOut lp, StrictEb se) {
    const int l = lane_id();
    const int x = l & 31, h = l >> 5;
    const uint32_t bx = tile % gx, by = tile / gx;
    const uint32_t X = bx * 32 + x;
    const uint32_t Y0 = by * 32 + 16 * h;
    const bool xin = X < nx;
    const uint64_t tbase = (uint64_t)(by * 32) * nx + bx * 32;  // the tile's first element (scalar)
    const uint32_t loff = (uint32_t)(16 * h) * nx + (uint32_t)x;  // this lane's first element in the tile
    float s[16], xv[16];
#pragma unroll
    for (int i = 0; i < 16; i++) xv[i] = xin16[i];
#pragma unroll
    for (int i = 0; i < 16; i++) s[i] = quant_round(xv[i], r);
    // t = s(y) - s(y-1) inside the tile (row 0 of the tile subtracts 0).
    const float north = __shfl(s[15], x, 64);  // row 15 of the tile, for h = 1
    float t[16];
    t[0] = s[0] - (h ? north : 0.0f);
#pragma unroll
    for (int i = 1; i < 16; i++) t[i] = s[i] - s[i - 1];
    uint64_t bal[16];
    const uint32_t cb = (uint32_t)sizeof(TCode) * loff;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float w = wave_shr1(t[i]);  // lane x - 1 (x > 0; lane 32's is unused)
        const float d = (x > 0) ? t[i] - w : t[i];
        const bool valid = INTERIOR || (xin && (Y0 + i) < ny);
        const bool ok = fabsf(d) < rad;
        if constexpr (CODES) {
            const TCode c = ok ? enc_code<TCode, ZZ>(d) : (TCode)0;
            if constexpr (INTERIOR) {
                char* cp = reinterpret_cast<char*>(codes + tbase + (uint64_t)i * nx);
                *gp(reinterpret_cast<TCode*>(cp + cb)) = c;
            } else if (valid) {
                gp(codes)[(uint64_t)(Y0 + i) * nx + X] = c;
            }
        }
        bal[i] = __ballot(valid && !ok);
        t[i] = d;  // keep the delta for the outlier pass
    }
    // Row-major order inside the tile: rows 0..15 (h=0) then 16..31 (h=1);
    // a lane's own flags are its bits of the ballots.
    auto emit = [&](const uint64_t (&bb)[16], bool patch, const ListOut& o) {
        uint32_t tot = 0, lw = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            tot += __popcll(bb[i]);
            lw += __popc((uint32_t)bb[i]);
        }
        if (tot) {
            uint32_t run = h ? lw : 0u;
            const uint32_t below = (1u << x) - 1u;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t half = h ? (uint32_t)(bb[i] >> 32) : (uint32_t)bb[i];
                if ((half >> x) & 1u) {
                    const uint32_t pos = run + __popc(half & below);
                    if (pos < o.cap) {
                        gp(o.val)[pos] = patch ? xv[i] : t[i];
                        gp(o.idx)[pos] = (Y0 + i) * nx + X;
                    }
                }
                run += __popc(half);
            }
        }
        return tot;
    };
    TileCounts tc{emit(bal, false, lo), 0u};
    if (CODES && last) zero_code_tail(codes, (uint64_t)nx * ny, n);
    if constexpr (STRICT) {
        bool huge = false;
#pragma unroll
        for (int i = 0; i < 16; i++) huge |= !(fabsf(s[i]) < kExactS);
        const bool tile_huge = __ballot(huge) != 0ull;
        uint64_t pbal[16];
        uint64_t unsure = 0;
        const StrictBand band = strict_band(se);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const bool valid = INTERIOR || (xin && (Y0 + i) < ny);
            bool sure, uns;
            patch_fast(xv[i], s[i], se, band, sure, uns);
            pbal[i] = __ballot(valid && (tile_huge || sure));
            unsure |= __ballot(valid && !tile_huge && uns);
        }
        if (unsure) {  // rare: elements within 2^-22 of the bound take the exact test
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const bool valid = INTERIOR || (xin && (Y0 + i) < ny);
                bool sure, uns;
                patch_fast(xv[i], s[i], se, band, sure, uns);
                pbal[i] = __ballot(valid && (tile_huge || sure || (uns && needs_patch(xv[i], s[i], se.eb, se.ebx2))));
            }
        }
        tc.pcnt = emit(pbal, true, lp);
    }
    return tc;
}


template <typename TCode, bool ZZ, bool STRICT, bool INTERIOR, bool CODES>
__device__ __forceinline__ TileCounts lz2d_tile_v(const float* __restrict__ in, uint32_t nx, uint32_t ny, uint64_t n,
                                                  uint32_t tile, uint32_t gx, bool last, float r, float rad,
                                                  TCode* __restrict__ codes, ListOut lo, ListOut lp, StrictEb se) {
    float xv[16];
    lz2d_load<INTERIOR>(in, nx, ny, tile, gx, xv);
    return lz2d_tile_c<TCode, ZZ, STRICT, INTERIOR, CODES>(xv, nx, ny, n, tile, gx, last, r, rad, codes, lo, lp, se);
}

template <typename TCode, bool ZZ, bool STRICT>
__device__ __forceinline__ TileCounts lz2d_tile(const float* __restrict__ in, uint32_t nx, uint32_t ny, uint64_t n,
                                                uint32_t tile, uint32_t gx, bool last, float r, float rad,
                                                TCode* __restrict__ codes, ListOut lo, ListOut lp, StrictEb se) {
    const uint32_t bx = tile % gx, by = tile / gx;
    const bool interior = lz2d_interior(nx, ny, bx, by);
    if (interior) {
        if (codes) return lz2d_tile_v<TCode, ZZ, STRICT, true, true>(in, nx, ny, n, tile, gx, last, r, rad, codes, lo, lp, se);
        return lz2d_tile_v<TCode, ZZ, STRICT, true, false>(in, nx, ny, n, tile, gx, last, r, rad, codes, lo, lp, se);
    }
    if (codes) return lz2d_tile_v<TCode, ZZ, STRICT, false, true>(in, nx, ny, n, tile, gx, last, r, rad, codes, lo, lp, se);
    return lz2d_tile_v<TCode, ZZ, STRICT, false, false>(in, nx, ny, n, tile, gx, last, r, rad, codes, lo, lp, se);
}

struct Tiles {
    uint32_t gx, gy, gz;  // tile grid
};

// ---------------------------------------------------------------------------
// 3D: tile (bx, by, bz) = (tile % gx, (tile / gx) % gy, tile / (gx gy)) of an
// nx x ny x nz field; `last` zeroes the code tail.  codes == nullptr: lists only.
// ---------------------------------------------------------------------------
// Tile geometry of this lane: column X, plane Z, first row Y0; the tile's
// first element tbase and the lane's 32-bit offset from it (interior tiles).
struct Lz3Geo {
    uint32_t X, Z, Y0;
    bool xz;
    uint64_t plane, tbase;
    uint32_t loff;
};
__device__ __forceinline__ Lz3Geo lz3d_geo(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t tile, Tiles tg) {
    const int l = lane_id();
    const int x = l & 7, z = l >> 3;
    const uint32_t bx = tile % tg.gx, by = (tile / tg.gx) % tg.gy, bz = tile / (tg.gx * tg.gy);
    Lz3Geo g;
    g.X = bx * 8 + x;
    g.Z = bz * 8 + z;
    g.Y0 = by * 8;
    g.xz = g.X < nx && g.Z < nz;
    g.plane = (uint64_t)nx * ny;
    g.tbase = (uint64_t)(bz * 8) * g.plane + (uint64_t)(by * 8) * nx + bx * 8;
    g.loff = (uint32_t)z * (uint32_t)g.plane + (uint32_t)x;
    return g;
}
// True when the tile lies inside the field and its 8 planes span less than
// 2^32 bytes (32-bit lane offsets from the tile base).
__device__ __forceinline__ bool lz3d_interior(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t tile, Tiles tg) {
    const uint32_t bx = tile % tg.gx, by = (tile / tg.gx) % tg.gy, bz = tile / (tg.gx * tg.gy);
    return bx * 8 + 8 <= nx && by * 8 + 8 <= ny && bz * 8 + 8 <= nz && (uint64_t)nx * ny < (1ull << 26);
}

// The tile's 8 values of this lane (rows Y0..Y0+7 of column X, plane Z;
// zero outside the field).
template <bool INTERIOR>
__device__ __forceinline__ void lz3d_load(const float* __restrict__ in, uint32_t nx, uint32_t ny, const Lz3Geo& g,
                                          float (&xv)[8]) {
    if constexpr (INTERIOR) {
        // row y of the tile from its scalar base plus a 32-bit lane offset
        const uint32_t lb = 4u * g.loff;
#pragma unroll
        for (int y = 0; y < 8; y++) {
            const char* rp = reinterpret_cast<const char*>(in + g.tbase + (uint64_t)y * nx);
            xv[y] = *gp(reinterpret_cast<const float*>(rp + lb));
        }
    } else {
#pragma unroll
        for (int y = 0; y < 8; y++)
            xv[y] = (g.xz && g.Y0 + y < ny) ? gp(in)[g.Z * g.plane + (uint64_t)(g.Y0 + y) * nx + g.X] : 0.0f;
    }
}

template <bool INTERIOR>
__device__ __forceinline__ void lz3d_load_tile(const float* __restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz,
                                               uint32_t tile, Tiles tg, float (&xv)[8]) {
    lz3d_load<INTERIOR>(in, nx, ny, lz3d_geo(nx, ny, nz, tile, tg), xv);
}

// The tile from its loaded values xv (lz3d_load).
template <typename TCode, bool ZZ, bool STRICT, bool INTERIOR, bool CODES>
__device__ __forceinline__ TileCounts lz3d_tile_c(const float (&xv)[8], uint32_t nx, uint32_t ny, uint32_t nz,
                                                  uint64_t n, uint32_t tile, Tiles tg, bool last, float r, float rad,
                                                  TCode* __restrict__ codes, ListOut lo, ListOut lp, StrictEb se) {
    const int l = lane_id();
    const int x = l & 7, z = l >> 3;
    const Lz3Geo g = lz3d_geo(nx, ny, nz, tile, tg);
    const uint32_t X = g.X, Z = g.Z, Y0 = g.Y0;
    const bool xz = g.xz;
    const uint64_t plane = g.plane, tbase = g.tbase;
    const uint32_t loff = g.loff;
    float s[8];
#pragma unroll
    for (int y = 0; y < 8; y++) s[y] = quant_round(xv[y], r);
    // a = s - s(z-1); b = a - a(x-1); d = b - b(y-1)   (lorenzo_quant_nd.cu:234-246)
    float b[8];
#pragma unroll
    for (int y = 0; y < 8; y++) {
        const float sz = __shfl_up(s[y], 8, 64);
        const float a = s[y] - (z > 0 ? sz : 0.0f);
        const float aw = wave_shr1(a);  // lane x - 1 of the group (x > 0)
        b[y] = (x > 0) ? a - aw : a;
    }
    float d[8];
    uint64_t bal[8];
    const uint32_t cb = (uint32_t)sizeof(TCode) * loff;
#pragma unroll
    for (int y = 0; y < 8; y++) {
        d[y] = (y > 0) ? b[y] - b[y - 1] : b[y];
        const bool valid = INTERIOR || (xz && (Y0 + y) < ny);
        const bool ok = fabsf(d[y]) < rad;
        if constexpr (CODES) {
            const TCode cv = ok ? enc_code<TCode, ZZ>(d[y]) : (TCode)0;
            if constexpr (INTERIOR) {
                char* cp = reinterpret_cast<char*>(codes + tbase + (uint64_t)y * nx);
                *gp(reinterpret_cast<TCode*>(cp + cb)) = cv;
            } else if (valid) {
                gp(codes)[Z * plane + (uint64_t)(Y0 + y) * nx + X] = cv;
            }
        }
        bal[y] = __ballot(valid && !ok);
    }
    // Order inside the tile: (z, y, x); a lane's own flags are its ballot bits.
    auto emit = [&](const uint64_t (&bb)[8], bool patch, const ListOut& o) {
        uint32_t tot = 0;
#pragma unroll
        for (int y = 0; y < 8; y++) tot += __popcll(bb[y]);
        if (tot) {
            const uint64_t zlow = (z == 0) ? 0ull : ((1ull << (8 * z)) - 1ull);
            uint32_t run = 0;
#pragma unroll
            for (int y = 0; y < 8; y++) run += __popcll(bb[y] & zlow);
            const uint32_t below = (1u << x) - 1u;
#pragma unroll
            for (int y = 0; y < 8; y++) {
                const uint32_t row = (uint32_t)(bb[y] >> (8 * z)) & 0xFFu;
                if ((row >> x) & 1u) {
                    const uint32_t pos = run + __popc(row & below);
                    if (pos < o.cap) {
                        gp(o.val)[pos] = patch ? xv[y] : d[y];
                        gp(o.idx)[pos] = (uint32_t)(Z * plane + (uint64_t)(Y0 + y) * nx + X);
                    }
                }
                run += __popc(row);
            }
        }
        return tot;
    };
    TileCounts tc{emit(bal, false, lo), 0u};
    if (CODES && last) zero_code_tail(codes, (uint64_t)nx * ny * nz, n);
    if constexpr (STRICT) {
        bool huge = false;
#pragma unroll
        for (int y = 0; y < 8; y++) huge |= !(fabsf(s[y]) < kExactS);
        const bool tile_huge = __ballot(huge) != 0ull;
        const StrictBand band = strict_band(se);
        uint64_t pbal[8];
        uint64_t unsure = 0;
#pragma unroll
        for (int y = 0; y < 8; y++) {
            const bool valid = INTERIOR || (xz && (Y0 + y) < ny);
            bool sure, uns;
            patch_fast(xv[y], s[y], se, band, sure, uns);
            pbal[y] = __ballot(valid && (tile_huge || sure));
            unsure |= __ballot(valid && !tile_huge && uns);
        }
        if (unsure) {  // rare: elements within 2^-22 of the bound take the exact test
#pragma unroll
            for (int y = 0; y < 8; y++) {
                const bool valid = INTERIOR || (xz && (Y0 + y) < ny);
                bool sure, uns;
                patch_fast(xv[y], s[y], se, band, sure, uns);
                pbal[y] = __ballot(valid && (tile_huge || sure || (uns && needs_patch(xv[y], s[y], se.eb, se.ebx2))));
            }
        }
        tc.pcnt = emit(pbal, true, lp);
    }
    return tc;
}

template <typename TCode, bool ZZ, bool STRICT, bool INTERIOR, bool CODES>
__device__ __forceinline__ TileCounts lz3d_tile_v(const float* __restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz,
                                                  uint64_t n, uint32_t tile, Tiles tg, bool last, float r, float rad,
                                                  TCode* __restrict__ codes, ListOut lo, ListOut lp, StrictEb se) {
    float xv[8];
    lz3d_load_tile<INTERIOR>(in, nx, ny, nz, tile, tg, xv);
    return lz3d_tile_c<TCode, ZZ, STRICT, INTERIOR, CODES>(xv, nx, ny, nz, n, tile, tg, last, r, rad, codes, lo, lp,
                                                           se);
}

template <typename TCode, bool ZZ, bool STRICT>
__device__ __forceinline__ TileCounts lz3d_tile(const float* __restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz,
                                                uint64_t n, uint32_t tile, Tiles tg, bool last, float r, float rad,
                                                TCode* __restrict__ codes, ListOut lo, ListOut lp, StrictEb se) {
    const bool interior = lz3d_interior(nx, ny, nz, tile, tg);
#define FZH_LZ3(I, C) lz3d_tile_v<TCode, ZZ, STRICT, I, C>(in, nx, ny, nz, n, tile, tg, last, r, rad, codes, lo, lp, se)
    if (interior) return codes ? FZH_LZ3(true, true) : FZH_LZ3(true, false);
    return codes ? FZH_LZ3(false, true) : FZH_LZ3(false, false);
#undef FZH_LZ3
}

// ---------------------------------------------------------------------------
// 3D, two x-adjacent tiles per wave: tiles A = (2 bx2, by, bz) and B =
// A + 1 of the tile grid (B absent when 2 bx2 + 1 == gx), for pair
// bx2 + gx2 (by + gy bz), gx2 = ceil(gx / 2).  Lane (x 0..15, yl 0..3)
// holds column X = 16 bx2 + x of rows Y0 + yl (h = 0) and Y0 + 4 + yl
// (h = 1) of the 8 planes: v[z][h] = element (X, Y0 + 4 h + yl, Z0 + z).
// A load instruction then covers 4 rows of 64 contiguous bytes; a lone
// tile's wave (lz3d_tile_v) covers 8 rows of 32 bytes, which the memory
// pipeline delivers at about 60% of the rate.  Codes, outlier and patch
// lists and counts are those of lz3d_tile_v for A and for B.
// ---------------------------------------------------------------------------
struct Lz3PairOut {
    ListOut loA, lpA, loB, lpB;
    TileCounts A, B;
};

__device__ __forceinline__ bool lz3d_pair_interior(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t bx2, uint32_t by,
                                                   uint32_t bz) {
    return bx2 * 16 + 16 <= nx && by * 8 + 8 <= ny && bz * 8 + 8 <= nz && (uint64_t)nx * ny < (1ull << 26);
}

template <typename TCode, bool ZZ, bool STRICT, bool INTERIOR>
__device__ __forceinline__ void lz3d_pair_v(const float* __restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz,
                                            uint64_t n, uint32_t bx2, uint32_t by, uint32_t bz, bool last, float r,
                                            float rad, TCode* __restrict__ codes, StrictEb se, Lz3PairOut& o) {
    const int l = lane_id();
    const int x = l & 15, yl = l >> 4;
    const uint32_t X = bx2 * 16 + x, Y0 = by * 8, Z0 = bz * 8;
    const bool xin = X < nx;
    const uint64_t plane = (uint64_t)nx * ny;
    const uint64_t tbase = (uint64_t)Z0 * plane + (uint64_t)Y0 * nx + bx2 * 16;
    const uint32_t loff = (uint32_t)yl * nx + (uint32_t)x;
    float xv[8][2];
#pragma unroll
    for (int z = 0; z < 8; z++)
#pragma unroll
        for (int h = 0; h < 2; h++) {
            if constexpr (INTERIOR) {
                const char* rp = reinterpret_cast<const char*>(in + tbase + (uint64_t)z * plane + (uint64_t)(4 * h) * nx);
                xv[z][h] = *gp(reinterpret_cast<const float*>(rp + 4u * loff));
            } else {
                const uint32_t Y = Y0 + 4 * h + yl, Z = Z0 + z;
                xv[z][h] = (xin && Y < ny && Z < nz) ? gp(in)[(uint64_t)Z * plane + (uint64_t)Y * nx + X] : 0.0f;
            }
        }
    // a = s - s(z-1); b = a - a(x-1); d = b - b(y-1)   (lorenzo_quant_nd.cu:234-246)
    float s[8][2], b[8][2];
#pragma unroll
    for (int z = 0; z < 8; z++)
#pragma unroll
        for (int h = 0; h < 2; h++) s[z][h] = quant_round(xv[z][h], r);
#pragma unroll
    for (int z = 0; z < 8; z++)
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const float a = s[z][h] - (z > 0 ? s[z > 0 ? z - 1 : 0][h] : 0.0f);
            const float aw = wave_shr1(a);  // lane x - 1 (x & 7 > 0: same tile)
            b[z][h] = (x & 7) ? a - aw : a;
        }
    float d[8][2];
    uint64_t bal[8][2];
    const uint32_t cb = (uint32_t)sizeof(TCode) * loff;
    const int src_h1 = yl == 0 ? l + 48 : l - 16;  // row y - 1 for h = 1: (h 0, yl 3) or (h 1, yl - 1)
#pragma unroll
    for (int z = 0; z < 8; z++) {
        const float up0 = __shfl_up(b[z][0], 16, 64);
        const float m = yl == 3 ? b[z][0] : b[z][1];
        const float up1 = __shfl(m, src_h1, 64);
        d[z][0] = yl > 0 ? b[z][0] - up0 : b[z][0];
        d[z][1] = b[z][1] - up1;
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint32_t Y = Y0 + 4 * h + yl, Z = Z0 + z;
            const bool valid = INTERIOR || (xin && Y < ny && Z < nz);
            const bool ok = fabsf(d[z][h]) < rad;
            const TCode cv = ok ? enc_code<TCode, ZZ>(d[z][h]) : (TCode)0;
            if constexpr (INTERIOR) {
                char* cp = reinterpret_cast<char*>(codes + tbase + (uint64_t)z * plane + (uint64_t)(4 * h) * nx);
                *gp(reinterpret_cast<TCode*>(cp + cb)) = cv;
            } else if (valid) {
                gp(codes)[(uint64_t)Z * plane + (uint64_t)Y * nx + X] = cv;
            }
            bal[z][h] = __ballot(valid && !ok);
        }
    }
    // Order inside a tile: (z, y, x); registers in (z, h) order and lanes in
    // (yl, x) order walk it, the tile's lanes being those with x & 8 equal.
    constexpr uint64_t kMaskA = 0x00FF00FF00FF00FFull;
    const uint64_t below = (1ull << l) - 1ull;
    auto emit = [&](const uint64_t (&bb)[8][2], bool patch, uint64_t mask, const ListOut& lo) {
        uint32_t tot = 0;
#pragma unroll
        for (int z = 0; z < 8; z++)
#pragma unroll
            for (int h = 0; h < 2; h++) tot += __popcll(bb[z][h] & mask);
        if (tot) {
            uint32_t run = 0;
#pragma unroll
            for (int z = 0; z < 8; z++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const uint64_t mm = bb[z][h] & mask;
                    if ((mm >> l) & 1ull) {
                        const uint32_t pos = run + __popcll(mm & below);
                        if (pos < lo.cap) {
                            gp(lo.val)[pos] = patch ? xv[z][h] : d[z][h];
                            gp(lo.idx)[pos] = (uint32_t)((uint64_t)(Z0 + z) * plane + (uint64_t)(Y0 + 4 * h + yl) * nx + X);
                        }
                    }
                    run += __popcll(mm);
                }
        }
        return tot;
    };
    o.A.cnt = emit(bal, false, kMaskA, o.loA);
    o.B.cnt = emit(bal, false, ~kMaskA, o.loB);
    if (last) zero_code_tail(codes, (uint64_t)nx * ny * nz, n);
    if constexpr (STRICT) {
        bool huge = false;
#pragma unroll
        for (int z = 0; z < 8; z++)
#pragma unroll
            for (int h = 0; h < 2; h++) huge |= !(fabsf(s[z][h]) < kExactS);
        const uint64_t hb = __ballot(huge);
        const bool tile_huge = (x & 8) ? (hb & ~kMaskA) != 0ull : (hb & kMaskA) != 0ull;
        const StrictBand band = strict_band(se);
        uint64_t pbal[8][2];
        uint64_t unsure = 0;
#pragma unroll
        for (int z = 0; z < 8; z++)
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const bool valid = INTERIOR || (xin && Y0 + 4 * h + yl < ny && Z0 + z < nz);
                bool sure, uns;
                patch_fast(xv[z][h], s[z][h], se, band, sure, uns);
                pbal[z][h] = __ballot(valid && (tile_huge || sure));
                unsure |= __ballot(valid && !tile_huge && uns);
            }
        if (unsure) {  // rare: elements within 2^-22 of the bound take the exact test
#pragma unroll
            for (int z = 0; z < 8; z++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const bool valid = INTERIOR || (xin && Y0 + 4 * h + yl < ny && Z0 + z < nz);
                    bool sure, uns;
                    patch_fast(xv[z][h], s[z][h], se, band, sure, uns);
                    pbal[z][h] = __ballot(valid && (tile_huge || sure ||
                                                    (uns && needs_patch(xv[z][h], s[z][h], se.eb, se.ebx2))));
                }
        }
        o.A.pcnt = emit(pbal, true, kMaskA, o.lpA);
        o.B.pcnt = emit(pbal, true, ~kMaskA, o.lpB);
    } else {
        o.A.pcnt = o.B.pcnt = 0;
    }
}

template <typename TCode, bool ZZ, bool STRICT>
__device__ __forceinline__ void lz3d_pair(const float* __restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz,
                                          uint64_t n, uint32_t bx2, uint32_t by, uint32_t bz, bool last, float r,
                                          float rad, TCode* __restrict__ codes, StrictEb se, Lz3PairOut& o) {
    if (lz3d_pair_interior(nx, ny, nz, bx2, by, bz))
        lz3d_pair_v<TCode, ZZ, STRICT, true>(in, nx, ny, nz, n, bx2, by, bz, last, r, rad, codes, se, o);
    else
        lz3d_pair_v<TCode, ZZ, STRICT, false>(in, nx, ny, nz, n, bx2, by, bz, last, r, rad, codes, se, o);
}

// Device-side resolution of the error bound from the value base, shared by
// the single-field and batched reductions (lorenzo_quant.cu:612-644):
//   params = [value_base, abs_eb, 1/(2 abs_eb), 2 abs_eb]
// with value_base = max(|min|, |max|) (REL, mode 1) or max - min (NOA, 2),
// and abs_eb = user_eb when value_base <= 0.
__device__ __forceinline__ void resolve_eb(float gmn, float gmx, int mode, float user_eb, float* params) {
    float vb;
    if (mode == 2) {
        vb = gmx - gmn;
    } else {
        const float amin = gmn < 0.0f ? -gmn : gmn;
        const float amax = gmx < 0.0f ? -gmx : gmx;
        vb = amin > amax ? amin : amax;
    }
    const float abs_eb = (vb <= 0.0f) ? user_eb : user_eb * vb;
    params[0] = vb;
    params[1] = abs_eb;
    params[2] = 1.0f / (2.0f * abs_eb);
    params[3] = 2.0f * abs_eb;
}

}  // namespace
}  // namespace fzh

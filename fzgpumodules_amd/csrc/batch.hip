// batch.hip — persistent, table-driven kernels of the batched compressor
// (LorenzoQuant -> Bitshuffle -> RZE over many fields at once), gfx950.
//
// No counterpart in the reference (one field per Pipeline::compress,
// src/pipeline/compressor_exec.cpp).  Every kernel here runs a fixed grid
// that walks a global index space (tiles, chunks or value-base units of all
// fields, fzh_batch.h) and does the per-tile / per-chunk work with the SAME
// device functions as the single-field kernels (fzh_lorenzo_tile.h,
// fzh_rze_chunk.h), so each field's bytes equal a single-field run's.
//
// Launch sequence of one compress (all on one stream, captured as a graph):
//   value_base  per-unit min/max, then one wave per field resolves eb
//   lorenzo     codes + per-tile lists (kListSlots entries) + per-tile counts,
//               one launch per tile shape (3D, 2D, 1D fields)
//   rze         fused Bitshuffle+RZE, one workgroup per chunk, into a chunk
//               scratch slot
//   scan        one exclusive scan (mod 2^32) over [tile outliers | tile
//               patches | chunk bytes]; per-field offsets are differences
//   finalize    one thread per field: result layout, concat header, pads
//   pack        chunks to their place in the field's RZE stream
//   scatter     tile lists to their place in the field's outlier / patch
//               arrays; a tile whose list outgrew kListSlots is recomputed
#include "fzh_batch.h"
#include "fzh_lorenzo_tile.h"
#include "fzh_rze_chunk.h"
#include "fzh_scan.h"

namespace fzh {
namespace batch {
namespace {

// Largest f with start[f] <= g (start strictly increasing, start[F] = total).
__device__ __forceinline__ uint32_t find_field(const uint32_t* __restrict__ start, uint32_t F, uint32_t g) {
    uint32_t lo = 0, hi = F;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (start[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// value base: per-unit min/max (first element and the zero padding folded in
// exactly as the single-field reduction does), then one wave per field
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) vb_partial_kernel(const Plan* __restrict__ pp) {
    const Plan& p = *pp;
    __shared__ float smn[4], smx[4];
    const int w = threadIdx.x >> 6;
    for (uint32_t u = blockIdx.x; u < p.U; u += gridDim.x) {
        const uint32_t f = p.ufield[u];
        const Field& fd = p.fields[f];
        const float* in = fd.in;
        const uint64_t b = (uint64_t)(u - fd.unit0) * kVbUnit;
        const uint64_t e = min(fd.nv, b + kVbUnit);
        const float first = gp(in)[0];
        float mn = first, mx = first;
        if (fd.nv < fd.n) {
            if (0.0f < mn) mn = 0.0f;
            if (0.0f > mx) mx = 0.0f;
        }
        const uint64_t e4 = b + ((e - b) & ~3ull);
        for (uint64_t i = b + 4 * (uint64_t)threadIdx.x; i < e4; i += 1024) {
            const float4 v = gld_f4(in + i);
            if (v.x < mn) mn = v.x;
            if (v.x > mx) mx = v.x;
            if (v.y < mn) mn = v.y;
            if (v.y > mx) mx = v.y;
            if (v.z < mn) mn = v.z;
            if (v.z > mx) mx = v.z;
            if (v.w < mn) mn = v.w;
            if (v.w > mx) mx = v.w;
        }
        for (uint64_t i = e4 + threadIdx.x; i < e; i += 256) {
            const float v = gp(in)[i];
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const float a = __shfl_xor(mn, off, 64), c = __shfl_xor(mx, off, 64);
            if (a < mn) mn = a;
            if (c > mx) mx = c;
        }
        if (lane_id() == 0) { smn[w] = mn; smx[w] = mx; }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = 1; i < 4; i++) {
                if (smn[i] < mn) mn = smn[i];
                if (smx[i] > mx) mx = smx[i];
            }
            p.vb_part[u] = make_float2(mn, mx);
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) vb_finalize_kernel(const Plan* __restrict__ pp) {
    const Plan& p = *pp;
    const uint32_t f = blockIdx.x * 4 + wave_in_wg();
    if (f >= p.F) return;
    float gmn = INFINITY, gmx = -INFINITY;
    if (p.mode != 0) {
        const Field& fd = p.fields[f];
        for (uint32_t u = lane_id(); u < fd.nunits; u += 64) {
            const float2 v = p.vb_part[fd.unit0 + u];
            if (v.x < gmn) gmn = v.x;
            if (v.y > gmx) gmx = v.y;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const float a = __shfl_xor(gmn, off, 64), c = __shfl_xor(gmx, off, 64);
            if (a < gmn) gmn = a;
            if (c > gmx) gmx = c;
        }
    } else {
        gmn = gmx = 0.0f;  // ABS: value base 0 -> abs_eb = user_eb
    }
    if (lane_id() == 0) resolve_eb(gmn, gmx, p.mode == 0 ? 1 : p.mode, p.user_eb, p.params + 4 * (size_t)f);
}

// ---------------------------------------------------------------------------
// Lorenzo forward over all tiles of all fields
// ---------------------------------------------------------------------------
template <bool ZZ, bool STRICT>
__device__ __forceinline__ TileCounts run_tile(const Plan& p, const Field& fd, uint32_t f, uint32_t lt,
                                               uint16_t* codes, ListOut lo, ListOut lp) {
    const float* prm = p.params + 4 * (size_t)f;
    const float r = prm[2];
    const StrictEb se{prm[1], prm[3]};
    const bool last = lt + 1 == fd.ntiles;
    if (fd.ndim == 1)
        return lz1d_tile<uint16_t, ZZ, STRICT>(fd.in, fd.n, fd.nv, lt, r, p.rad, codes, lo, lp, se);
    if (fd.ndim == 2)
        return lz2d_tile<uint16_t, ZZ, STRICT>(fd.in, fd.nx, fd.ny, fd.n, lt, fd.gx, last, r, p.rad, codes, lo, lp,
                                               se);
    return lz3d_tile<uint16_t, ZZ, STRICT>(fd.in, fd.nx, fd.ny, fd.nz, fd.n, lt, Tiles{fd.gx, fd.gy, 0}, last, r,
                                           p.rad, codes, lo, lp, se);
}

// Wave-uniform copies of table values (the tables are read with vector
// loads, so without this every field parameter and pointer would take
// vector registers for the whole tile).
__device__ __forceinline__ uint32_t rfl(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ float rflf(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
__device__ __forceinline__ Field uniform_field(const Field* fp) {
    static_assert(sizeof(Field) % 16 == 0, "Field is read as 16-byte words");
    Field f;
    const uint4* src = reinterpret_cast<const uint4*>(fp);
    uint4* dst = reinterpret_cast<uint4*>(&f);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(Field) / 16); i++) {
        const uint4 v = gld_u4(src + i);
        dst[i] = make_uint4(rfl(v.x), rfl(v.y), rfl(v.z), rfl(v.w));
    }
    return f;
}

// One launch per tile shape (ND = 1, 2, 3): a wave takes units of G
// consecutive tiles of its class and walks them field by field; tile
// indices, list slots and counts are the global ones.
template <int ND, bool ZZ, bool STRICT>
__device__ __forceinline__ TileCounts run_tile_nd(const Plan& p, const Field& fd, uint32_t f, uint32_t lt,
                                                  uint16_t* codes, ListOut lo, ListOut lp) {
    const float* prm = p.params + 4 * (size_t)f;
    const float r = rflf(gp(prm)[2]);
    const StrictEb se{rflf(gp(prm)[1]), rflf(gp(prm)[3])};
    const bool last = lt + 1 == fd.ntiles;
    // codes are always written here: only the code-writing tile variants
    if constexpr (ND == 1) {
        return lz1d_tile<uint16_t, ZZ, STRICT>(fd.in, fd.n, fd.nv, lt, r, p.rad, codes, lo, lp, se);
    } else if constexpr (ND == 2) {
        const uint32_t bx = lt % fd.gx, by = lt / fd.gx;
        if (lz2d_interior(fd.nx, fd.ny, bx, by))
            return lz2d_tile_v<uint16_t, ZZ, STRICT, true, true>(fd.in, fd.nx, fd.ny, fd.n, lt, fd.gx, last, r, p.rad,
                                                                 codes, lo, lp, se);
        return lz2d_tile_v<uint16_t, ZZ, STRICT, false, true>(fd.in, fd.nx, fd.ny, fd.n, lt, fd.gx, last, r, p.rad,
                                                              codes, lo, lp, se);
    } else {
        const Tiles tg{fd.gx, fd.gy, 0};
        const uint32_t bx = lt % tg.gx, by = (lt / tg.gx) % tg.gy, bz = lt / (tg.gx * tg.gy);
        if (bx * 8 + 8 <= fd.nx && by * 8 + 8 <= fd.ny && bz * 8 + 8 <= fd.nz && (uint64_t)fd.nx * fd.ny < (1ull << 26))
            return lz3d_tile_v<uint16_t, ZZ, STRICT, true, true>(fd.in, fd.nx, fd.ny, fd.nz, fd.n, lt, tg, last, r,
                                                                 p.rad, codes, lo, lp, se);
        return lz3d_tile_v<uint16_t, ZZ, STRICT, false, true>(fd.in, fd.nx, fd.ny, fd.nz, fd.n, lt, tg, last, r,
                                                              p.rad, codes, lo, lp, se);
    }
}

template <int ND, bool ZZ, bool STRICT>
__global__ void __launch_bounds__(256) lz_batch_kernel(const Plan* __restrict__ pp, uint32_t G) {
    const Plan& p = *pp;
    constexpr int d = ND - 1;
    const uint32_t* __restrict__ cf = p.cls_field[d];
    const uint32_t* __restrict__ ct0 = p.cls_tile0[d];
    const uint32_t T = p.cls_T[d];
    const uint32_t waves = gridDim.x * 4;
    const uint32_t nu = (T + G - 1) / G;
    for (uint32_t u = blockIdx.x * 4 + wave_in_wg(); u < nu; u += waves) {
        const uint32_t c0 = u * G, c1 = min(T, c0 + G);
        uint32_t j = rfl(gp(p.cls_tfield[d])[c0 >> 6]);
        while (c0 >= rfl(gp(ct0)[j + 1])) j++;
        for (uint32_t c = c0; c < c1; c++) {
            while (c >= rfl(gp(ct0)[j + 1])) j++;
            const uint32_t f = rfl(gp(cf)[j]);
            const Field fd = uniform_field(p.fields + f);
            const uint32_t lt = c - rfl(gp(ct0)[j]);
            const uint32_t g = fd.tile0 + lt;
            const uint64_t sl = (uint64_t)g * kListSlots;
            const ListOut lo{p.lerr + sl, p.lidx + sl, kListSlots};
            const ListOut lp = STRICT ? ListOut{p.lpval + sl, p.lpidx + sl, kListSlots} : ListOut{nullptr, nullptr, 0};
            const TileCounts tc = run_tile_nd<ND, ZZ, STRICT>(p, fd, f, lt, p.codes + fd.codes_off, lo, lp);
            if (lane_id() == 0) {
                gp(p.counts)[g] = tc.cnt;
                if constexpr (STRICT) gp(p.counts)[p.T + g] = tc.pcnt;
            }
        }
    }
}

// The same with one tile per wave and no loop (a grid covering the class):
// the registers of the single-field kernels, so more waves per SIMD.
template <int ND, bool ZZ, bool STRICT>
__global__ void __launch_bounds__(256) lz_batch_tile_kernel(const Plan* __restrict__ pp) {
    const Plan& p = *pp;
    constexpr int d = ND - 1;
    const uint32_t c = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (c >= p.cls_T[d]) return;
    const uint32_t* __restrict__ ct0 = p.cls_tile0[d];
    uint32_t j = rfl(gp(p.cls_tfield[d])[c >> 6]);
    while (c >= rfl(gp(ct0)[j + 1])) j++;
    const uint32_t f = rfl(gp(p.cls_field[d])[j]);
    const Field fd = uniform_field(p.fields + f);
    const uint32_t lt = c - rfl(gp(ct0)[j]);
    const uint32_t g = fd.tile0 + lt;
    const uint64_t sl = (uint64_t)g * kListSlots;
    const ListOut lo{p.lerr + sl, p.lidx + sl, kListSlots};
    const ListOut lp = STRICT ? ListOut{p.lpval + sl, p.lpidx + sl, kListSlots} : ListOut{nullptr, nullptr, 0};
    const TileCounts tc = run_tile_nd<ND, ZZ, STRICT>(p, fd, f, lt, p.codes + fd.codes_off, lo, lp);
    if (lane_id() == 0) {
        gp(p.counts)[g] = tc.cnt;
        if constexpr (STRICT) gp(p.counts)[p.T + g] = tc.pcnt;
    }
}

// 3D fields: one wave per pair of x-adjacent tiles of one field
// (lz3d_pair: a load instruction covers 4 rows of 64 bytes instead of 8 of
// 32), tile lists, counts and codes as lz3d_tile gives them.
template <bool ZZ, bool STRICT>
__global__ void __launch_bounds__(256) lz_batch_tile3_kernel(const Plan* __restrict__ pp) {
    const Plan& p = *pp;
    const uint32_t u = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (u >= p.P) return;
    const uint32_t* __restrict__ pr0 = p.pair0;
    uint32_t j = rfl(gp(p.pfield)[u >> 6]);
    while (u >= rfl(gp(pr0)[j + 1])) j++;
    const uint32_t f = rfl(gp(p.cls_field[2])[j]);
    const Field fd = uniform_field(p.fields + f);
    const uint32_t lp = u - rfl(gp(pr0)[j]);
    const uint32_t gx2 = (fd.gx + 1) / 2;
    const uint32_t bx2 = lp % gx2, by = (lp / gx2) % fd.gy, bz = lp / (gx2 * fd.gy);
    const uint32_t ta = 2 * bx2 + fd.gx * (by + fd.gy * bz), tb = ta + 1;
    const bool has_b = 2 * bx2 + 1 < fd.gx;
    const float* prm = p.params + 4 * (size_t)f;
    const float r = rflf(gp(prm)[2]);
    const StrictEb se{rflf(gp(prm)[1]), rflf(gp(prm)[3])};
    const uint32_t ga = fd.tile0 + ta, gb = fd.tile0 + tb;
    const uint64_t sa = (uint64_t)ga * kListSlots, sb = (uint64_t)gb * kListSlots;
    Lz3PairOut o;
    o.loA = ListOut{p.lerr + sa, p.lidx + sa, kListSlots};
    o.lpA = STRICT ? ListOut{p.lpval + sa, p.lpidx + sa, kListSlots} : ListOut{nullptr, nullptr, 0};
    o.loB = has_b ? ListOut{p.lerr + sb, p.lidx + sb, kListSlots} : ListOut{nullptr, nullptr, 0};
    o.lpB = (STRICT && has_b) ? ListOut{p.lpval + sb, p.lpidx + sb, kListSlots} : ListOut{nullptr, nullptr, 0};
    const bool last = ta + 1 == fd.ntiles || (has_b && tb + 1 == fd.ntiles);
    lz3d_pair<uint16_t, ZZ, STRICT>(fd.in, fd.nx, fd.ny, fd.nz, fd.n, bx2, by, bz, last, r, p.rad,
                                    p.codes + fd.codes_off, se, o);
    if (lane_id() == 0) {
        gp(p.counts)[ga] = o.A.cnt;
        if constexpr (STRICT) gp(p.counts)[p.T + ga] = o.A.pcnt;
        if (has_b) {
            gp(p.counts)[gb] = o.B.cnt;
            if constexpr (STRICT) gp(p.counts)[p.T + gb] = o.B.pcnt;
        }
    }
}

// ---------------------------------------------------------------------------
// Fused Bitshuffle + RZE over all chunks of all fields: one 256-thread
// workgroup per entry of the chunk table.  (Wrapped in a persistent loop the
// encoder's register allocation triples -- 74 -> 213 VGPRs -- and occupancy
// drops to 2 waves/SIMD, so this launch covers the table directly.)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) rze_batch_kernel(const Plan* __restrict__ pp) {
    const Plan& p = *pp;
    __shared__ RzeWgLds L;
    const uint32_t gc = xcd_swizzle(blockIdx.x, gridDim.x);
    if (gc >= p.NC) return;
    const uint32_t cbase = p.T * (p.strict ? 2u : 1u);
    const uint32_t f = rfl(gp(p.cfield)[gc]);
    const Field fd = uniform_field(p.fields + f);
    const uint32_t lc = gc - fd.chunk0;
    const uint64_t cb = 2 * fd.n;  // code bytes
    const uint64_t off = (uint64_t)lc * kChunk;
    const int in_size = (int)min((uint64_t)kChunk, cb - off);
    const uint8_t* src = reinterpret_cast<const uint8_t*>(p.codes + fd.codes_off) + off;
    const uint32_t size =
        rze_encode_chunk_wg<2>(src, in_size, off + kChunk <= cb, L, p.chunk_scratch + (uint64_t)gc * kChunk);
    if (threadIdx.x == 0) {
        p.sizes[gc] = size;
        p.counts[cbase + gc] = size & 0x7FFFFFFFu;
        uint32_t* rz = reinterpret_cast<uint32_t*>(p.out + fd.out_off + p.hdr);
        rz[2 + lc] = size;
        if (lc == 0) {
            rz[0] = (uint32_t)cb;
            rz[1] = fd.nchunks;
        }
    }
}

// ---------------------------------------------------------------------------
// Per-field layout of the result (fz::Pipeline's concat, fzh_core.hip)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t pad16(uint64_t v) { return (v + 15) & ~15ull; }

__device__ __forceinline__ void zero_range(uint8_t* b, uint64_t from, uint64_t to) {
    for (uint64_t i = from; i < to; i++) b[i] = 0;
}

__global__ void __launch_bounds__(256) finalize_kernel(const Plan* __restrict__ pp) {
    const Plan& p = *pp;
    const uint32_t f = blockIdx.x * 256 + threadIdx.x;
    if (f >= p.F) return;
    const Field& fd = p.fields[f];
    const uint32_t cbase = p.T * (p.strict ? 2u : 1u);
    const uint32_t ex_t = p.excl[fd.tile0];
    const uint32_t k = p.excl[fd.tile0 + fd.ntiles] - ex_t;
    uint32_t ex_p = 0, kp = 0;
    if (p.strict) {
        ex_p = p.excl[p.T + fd.tile0];
        kp = p.excl[p.T + fd.tile0 + fd.ntiles] - ex_p;
    }
    const uint32_t ex_c = p.excl[cbase + fd.chunk0];
    const uint64_t csum = (uint32_t)(p.excl[cbase + fd.chunk0 + fd.nchunks] - ex_c);
    const uint64_t rze_len = 8ull + 4ull * fd.nchunks + csum;
    const uint64_t rze_bytes = (rze_len + 3) & ~3ull;
    const uint32_t ko = min(k, fd.max_out), kpo = min(kp, fd.max_patch);
    uint64_t sz[7] = {rze_bytes, 4ull * ko, 4ull * ko, 4, 4ull * kpo, 4ull * kpo, 4};
    uint64_t off[7];
    uint64_t o = p.hdr;
    for (uint32_t i = 0; i < p.nout; i++) {
        off[i] = o;
        o += pad16(sz[i]);
    }
    Result res;
    res.bytes = o;
    res.k = k;
    res.kp = kp;
    res.flags = (k > fd.max_out ? kOverflowOutliers : 0u) | (kp > fd.max_patch ? kOverflowPatches : 0u);
    res.pad[0] = res.pad[1] = res.pad[2] = 0;
    Derived d;
    d.o_err = off[1];
    d.o_idx = off[2];
    d.o_pval = p.strict ? off[4] : 0;
    d.o_pidx = p.strict ? off[5] : 0;
    d.ko = ko;
    d.kpo = kpo;
    d.ex_t = ex_t;
    d.ex_p = ex_p;
    d.ex_c = ex_c;
    d.pad = 0;
    if (o > fd.out_cap) {
        res.flags |= kOverflowCapacity;
    } else {
        uint8_t* b = p.out + fd.out_off;
        uint32_t* b32 = reinterpret_cast<uint32_t*>(b);
        b32[0] = p.nout;
        for (uint32_t i = 0; i < p.nout; i++) {
            b32[1 + 2 * i] = (uint32_t)sz[i];
            b32[2 + 2 * i] = (uint32_t)(sz[i] >> 32);
        }
        zero_range(b, 4 + 8ull * p.nout, p.hdr);
        zero_range(b, off[0] + rze_len, off[0] + pad16(rze_bytes));
        for (uint32_t i = 1; i < p.nout; i++) zero_range(b, off[i] + sz[i], off[i] + pad16(sz[i]));
        b32[off[3] / 4] = k;
        if (p.strict) b32[off[6] / 4] = kp;
    }
    p.derived[f] = d;
    p.results[f] = res;
}

// ---------------------------------------------------------------------------
// Chunks to their place in the field's RZE stream
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) pack_batch_kernel(const Plan* __restrict__ pp) {
    const Plan& p = *pp;
    const uint32_t waves = gridDim.x * 4;
    const uint32_t cbase = p.T * (p.strict ? 2u : 1u);
    for (uint32_t gc = blockIdx.x * 4 + wave_in_wg(); gc < p.NC; gc += waves) {
        const uint32_t f = p.cfield[gc];
        if (p.results[f].flags & kOverflowCapacity) continue;
        const Field& fd = p.fields[f];
        const uint32_t lc = gc - fd.chunk0;
        const uint32_t size = p.sizes[gc];
        const uint32_t clean = size & 0x7FFFFFFFu;
        const bool raw = (size & 0x80000000u) != 0u;
        const uint64_t cb = 2 * fd.n;
        const uint64_t off = (uint64_t)lc * kChunk;
        const bool shuffled = off + kChunk <= cb;
        uint8_t* dst = p.out + fd.out_off + p.hdr + 8ull + 4ull * fd.nchunks +
                       (uint32_t)(p.excl[cbase + gc] - p.derived[f].ex_c);
        if (raw && !shuffled)
            copy_to_unaligned16(reinterpret_cast<const uint8_t*>(p.codes + fd.codes_off) + off, cb - off, dst, clean);
        else
            copy_to_unaligned16(p.chunk_scratch + (uint64_t)gc * kChunk, kChunk, dst, clean);
    }
}

// ---------------------------------------------------------------------------
// Tile lists to the field's outlier / patch arrays.  A wave takes 64 tiles,
// finds those with a non-empty list by ballot and places each; a list longer
// than kListSlots is produced again by recomputing the tile (lists only).
// ---------------------------------------------------------------------------
template <bool ZZ, bool STRICT>
__global__ void __launch_bounds__(256) scatter_batch_kernel(const Plan* __restrict__ pp) {
    const Plan& p = *pp;
    const uint32_t waves = gridDim.x * 4;
    const int l = lane_id();
    const uint32_t nblk = (p.T + 63) / 64;
    for (uint32_t blk = blockIdx.x * 4 + wave_in_wg(); blk < nblk; blk += waves) {
        const uint32_t gl = blk * 64 + (uint32_t)l;
        uint32_t c = 0, pc = 0;
        if (gl < p.T) {
            c = p.counts[gl];
            if constexpr (STRICT) pc = p.counts[p.T + gl];
        }
        uint64_t todo = __ballot((c | pc) != 0u);
        while (todo) {
            const int j = __builtin_ctzll(todo);
            todo &= todo - 1;
            const uint32_t g = blk * 64 + (uint32_t)j;
            const uint32_t tc = __shfl(c, j, 64), tp = __shfl(pc, j, 64);
            const uint32_t f = find_field(p.tile0, p.F, g);
            if (p.results[f].flags & kOverflowCapacity) continue;
            const Field& fd = p.fields[f];
            const Derived d = p.derived[f];
            uint8_t* b = p.out + fd.out_off;
            float* err = reinterpret_cast<float*>(b + d.o_err);
            uint32_t* idx = reinterpret_cast<uint32_t*>(b + d.o_idx);
            const uint64_t sl = (uint64_t)g * kListSlots;
            const uint32_t po = p.excl[g] - d.ex_t;
            bool redo_o = false, redo_p = false;
            if (tc <= kListSlots) {
                if ((uint32_t)l < tc && po + l < d.ko) {
                    err[po + l] = p.lerr[sl + l];
                    idx[po + l] = p.lidx[sl + l];
                }
            } else {
                redo_o = po < d.ko;
            }
            uint32_t pp = 0;
            float* pval = nullptr;
            uint32_t* pidx = nullptr;
            if constexpr (STRICT) {
                pval = reinterpret_cast<float*>(b + d.o_pval);
                pidx = reinterpret_cast<uint32_t*>(b + d.o_pidx);
                pp = p.excl[p.T + g] - d.ex_p;
                if (tp <= kListSlots) {
                    if ((uint32_t)l < tp && pp + l < d.kpo) {
                        pval[pp + l] = p.lpval[sl + l];
                        pidx[pp + l] = p.lpidx[sl + l];
                    }
                } else {
                    redo_p = pp < d.kpo;
                }
            }
            if (redo_o || redo_p) {
                const ListOut lo = redo_o ? ListOut{err + po, idx + po, d.ko - po} : ListOut{nullptr, nullptr, 0};
                const ListOut lp = redo_p ? ListOut{pval + pp, pidx + pp, d.kpo - pp} : ListOut{nullptr, nullptr, 0};
                (void)run_tile<ZZ, STRICT>(p, fd, f, g - fd.tile0, nullptr, lo, lp);
            }
        }
    }
}

__global__ void __launch_bounds__(256) readback_batch_kernel(const uint32_t* __restrict__ src, uint32_t* dst,
                                                             uint64_t words) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < words; i += (uint64_t)gridDim.x * 256)
        dst[i] = src[i];
}

}  // namespace

size_t scan_ws_bytes_for(size_t n_counts) { return fzh::scan_ws_bytes(n_counts); }

int launch_value_base(hipStream_t s, const Plan& p, const Plan* dp, uint32_t grid) {
    if (p.mode != 0 && p.U) vb_partial_kernel<<<std::max<uint32_t>(1, std::min(grid, p.U)), 256, 0, s>>>(dp);
    vb_finalize_kernel<<<(unsigned)cdiv(p.F, 4), 256, 0, s>>>(dp);
    FZH_CHECK_LAUNCH("batch value base");
    return FZH_OK;
}

int launch_lorenzo(hipStream_t s, const Plan& p, const Plan* dp, uint32_t grid, uint32_t G) {
    G = std::max<uint32_t>(1, G);
#define FZH_LZB(ND)                                                                                   \
    if (p.cls_T[ND - 1]) {                                                                            \
        const uint32_t nu = (uint32_t)cdiv(p.cls_T[ND - 1], G);                                       \
        const unsigned g = std::max<uint32_t>(1, std::min<uint64_t>(grid, cdiv(nu, 4)));             \
        if (G == 1 && (uint64_t)grid >= cdiv(nu, 4) && ND == 3) { /* two tiles per wave */          \
            const unsigned g3 = (unsigned)std::max<uint64_t>(1, cdiv(p.P, 4));                        \
            if (p.zigzag) {                                                                           \
                if (p.strict) lz_batch_tile3_kernel<true, true><<<g3, 256, 0, s>>>(dp);                \
                else lz_batch_tile3_kernel<true, false><<<g3, 256, 0, s>>>(dp);                        \
            } else {                                                                                  \
                if (p.strict) lz_batch_tile3_kernel<false, true><<<g3, 256, 0, s>>>(dp);               \
                else lz_batch_tile3_kernel<false, false><<<g3, 256, 0, s>>>(dp);                       \
            }                                                                                         \
        } else if (G == 1 && (uint64_t)grid >= cdiv(nu, 4)) { /* one tile per wave */               \
            if (p.zigzag) {                                                                           \
                if (p.strict) lz_batch_tile_kernel<ND, true, true><<<g, 256, 0, s>>>(dp);              \
                else lz_batch_tile_kernel<ND, true, false><<<g, 256, 0, s>>>(dp);                      \
            } else {                                                                                  \
                if (p.strict) lz_batch_tile_kernel<ND, false, true><<<g, 256, 0, s>>>(dp);             \
                else lz_batch_tile_kernel<ND, false, false><<<g, 256, 0, s>>>(dp);                     \
            }                                                                                         \
        } else if (p.zigzag) {                                                                        \
            if (p.strict) lz_batch_kernel<ND, true, true><<<g, 256, 0, s>>>(dp, G);                    \
            else lz_batch_kernel<ND, true, false><<<g, 256, 0, s>>>(dp, G);                            \
        } else {                                                                                      \
            if (p.strict) lz_batch_kernel<ND, false, true><<<g, 256, 0, s>>>(dp, G);                   \
            else lz_batch_kernel<ND, false, false><<<g, 256, 0, s>>>(dp, G);                           \
        }                                                                                             \
        FZH_CHECK_LAUNCH("batch lorenzo");                                                            \
    }
    FZH_LZB(3)
    FZH_LZB(2)
    FZH_LZB(1)
#undef FZH_LZB
    return FZH_OK;
}

int launch_rze(hipStream_t s, const Plan& p, const Plan* dp) {
    rze_batch_kernel<<<p.NC, 256, 0, s>>>(dp);
    FZH_CHECK_LAUNCH("batch rze");
    return FZH_OK;
}

int launch_scan(hipStream_t s, const Plan& p) {
    return fzh::scan_u32<true>(s, p.counts, p.excl, p.n_counts, nullptr, p.excl + p.n_counts, p.scan_ws);
}

int launch_finalize(hipStream_t s, const Plan& p, const Plan* dp) {
    finalize_kernel<<<(unsigned)cdiv(p.F, 256), 256, 0, s>>>(dp);
    FZH_CHECK_LAUNCH("batch finalize");
    return FZH_OK;
}

int launch_pack(hipStream_t s, const Plan& p, const Plan* dp, uint32_t grid) {
    const unsigned g = std::max<uint32_t>(1, std::min<uint32_t>(grid, (uint32_t)cdiv(p.NC, 4)));
    pack_batch_kernel<<<g, 256, 0, s>>>(dp);
    FZH_CHECK_LAUNCH("batch pack");
    return FZH_OK;
}

int launch_scatter(hipStream_t s, const Plan& p, const Plan* dp, uint32_t grid) {
    const unsigned g = std::max<uint32_t>(1, std::min<uint32_t>(grid, (uint32_t)cdiv(cdiv(p.T, 64), 4)));
    if (p.zigzag) {
        if (p.strict) scatter_batch_kernel<true, true><<<g, 256, 0, s>>>(dp);
        else scatter_batch_kernel<true, false><<<g, 256, 0, s>>>(dp);
    } else {
        if (p.strict) scatter_batch_kernel<false, true><<<g, 256, 0, s>>>(dp);
        else scatter_batch_kernel<false, false><<<g, 256, 0, s>>>(dp);
    }
    FZH_CHECK_LAUNCH("batch scatter");
    return FZH_OK;
}

int launch_readback(hipStream_t s, const Plan& p, void* dst) {
    const uint64_t words = (uint64_t)p.F * sizeof(Result) / 4;
    readback_batch_kernel<<<(unsigned)std::max<uint64_t>(1, std::min<uint64_t>(cdiv(words, 256), 64)), 256, 0, s>>>(
        reinterpret_cast<const uint32_t*>(p.results), reinterpret_cast<uint32_t*>(dst), words);
    FZH_CHECK_LAUNCH("batch readback");
    return FZH_OK;
}

}  // namespace batch
}  // namespace fzh

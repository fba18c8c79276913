// lorenzo_f64.hip — Lorenzo predict+quantize and its inverse for FP64 input
// (LorenzoQuantStage<double, uint16_t | uint32_t>, instantiated at
// modules/fused/lorenzo_quant/lorenzo_quant.h:335-339), and the FP64 value
// base (predictor_utils.cuh:75-116 with TInput = double).
//
// The tiles, their order and the outlier lists are those of the FP32 path
// (1D tiles of 1024, lorenzo_quant.cu:262-364; 2D 32x32, lorenzo_quant_nd.cu:
// 29-105; 3D 8x8x8, lorenzo_quant_nd.cu:199-274); every arithmetic step is
// in double, as the reference's kernels run with TInput = double:
// s = round(x * ebx2_r), the Lorenzo differences, |delta| < radius, and the
// outlier errors are stored as doubles.  The inverse runs the reference's
// addition trees (lorenzo_quant.cu:17-87,155-244; lorenzo_quant_nd.cu:
// 112-190, 281-364) in double.
//
// One wave64 per tile, four tiles per 256-thread workgroup.  Lane layouts:
//   1D  lane l, step q: elements 4(l+64q) .. 4(l+64q)+3 of the tile
//   2D  lane l: column x = l&31, rows 16*(l>>5) .. +15 of the tile
//   3D  lane l: x = l&7, z = l>>3, rows y = 0..7 of the tile
// A tile's outliers are listed in element-index order: the wave marks them
// in an LDS bitmap of the tile, and an element's rank is the number of
// marks before it.  The per-tile lists are then placed at their prefixes
// (fzh_compact.h), as in the FP32 path.
#include "fzh_lorenzo_common.h"
#include "fzh_lorenzo_tile.h"  // Tiles

namespace fzh {
namespace {

template <typename TCode> struct Signed64Of;
template <> struct Signed64Of<uint16_t> { using type = int16_t; };
template <> struct Signed64Of<uint32_t> { using type = int32_t; };

// (TCode)(int)delta or zigzag((SCode)(int)delta) (lorenzo_quant.cu:328-334).
// Only |delta| < radius reaches here; with u32 codes and a radius above 2^31
// the int conversion saturates (as the GPUs' float-to-int conversions do),
// which the clamp makes explicit.
template <typename TCode, bool ZZ>
__device__ __forceinline__ TCode enc_code64(double d) {
    using SC = typename Signed64Of<TCode>::type;
    constexpr int W = 8 * sizeof(TCode);
    const int q = (int)fmin(fmax(d, -2147483648.0), 2147483647.0);
    if constexpr (ZZ) {
        const SC s = (SC)q;
        return (TCode)(((TCode)s << 1) ^ (TCode)(s >> (W - 1)));
    } else {
        return (TCode)q;
    }
}

// (double)(SCode)code, or zigzag-decode (lorenzo_quant.cu:195-202).
template <typename TCode, bool ZZ>
__device__ __forceinline__ double dec_code64(TCode c) {
    using SC = typename Signed64Of<TCode>::type;
    if constexpr (ZZ) return (double)(SC)((TCode)(c >> 1) ^ (TCode)(-(SC)(c & 1u)));
    else return (double)(SC)c;
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Value of lane l - 1 (lane 0: 0.0), with DPP on both halves.
__device__ __forceinline__ double wave_shr1_d(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = wave_shr1((uint32_t)u), hi = wave_shr1((uint32_t)(u >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Per-tile list destinations and counts (as the FP32 path's TileLists).
struct TileLists64 {
    double* err;
    uint32_t* idx;
    uint32_t* cnt;
};

// Lists a tile's outliers in local-index order.  f[i]: register i holds an
// outlier with local index loc[i], global index gid[i] and error val[i].
// bits: this wave's 64 LDS words (bitmap words 0..31, word prefixes 32..63).
// Called by the whole wave; returns the tile's outlier count.
template <int NR>
__device__ __forceinline__ uint32_t emit_outliers(uint32_t* bits, const bool (&f)[NR], const uint32_t (&loc)[NR],
                                                  const double (&val)[NR], const uint32_t (&gid)[NR],
                                                  double* __restrict__ lerr, uint32_t* __restrict__ lidx) {
    bool any = false;
#pragma unroll
    for (int i = 0; i < NR; i++) any |= f[i];
    if (!__ballot(any)) return 0u;
    const int l = lane_id();
    if (l < 32) bits[l] = 0u;
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < NR; i++)
        if (f[i]) atomicOr(&bits[loc[i] >> 5], 1u << (loc[i] & 31u));
    wave_lds_sync();
    const uint32_t word = l < 32 ? bits[l] : 0u;
    const uint32_t c = (uint32_t)__popc(word);
    const uint32_t inc = wave_incl_scan(c);
    if (l < 32) bits[32 + l] = inc - c;
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < NR; i++)
        if (f[i]) {
            const uint32_t w = loc[i] >> 5;
            const uint32_t r = bits[32 + w] + (uint32_t)__popc(bits[w] & ((1u << (loc[i] & 31u)) - 1u));
            lerr[r] = val[i];
            lidx[r] = gid[i];
        }
    return lane63(inc);
}

// ===========================================================================
// Forward
// ===========================================================================

// 1D: chain restarts at every tile start (lorenzo_quant.cu:287-312).
// Elements [nv, n) read as 0.0 (the reference's zero-padded copy).
template <typename TCode, bool ZZ>
__global__ void __launch_bounds__(256)
lz1d_fwd64_kernel(const double* __restrict__ in, uint64_t n, uint64_t nv, double r_host,
                  const double* __restrict__ d_r, double rad, TCode* __restrict__ codes, TileLists64 tl,
                  uint32_t T) {
    __shared__ uint32_t s_bits[4][64];
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (tile >= T) return;
    const double r = d_r ? *d_r : r_host;
    const int l = lane_id();
    const uint64_t base = (uint64_t)tile * 1024;
    const bool full_in = base + 1024 <= nv, full_out = base + 1024 <= n;
    double s[16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint64_t e0 = base + 4 * (uint64_t)(l + 64 * q);
        double v[4];
        if (full_in) {
            const double2 a = *reinterpret_cast<const double2*>(in + e0);
            const double2 b = *reinterpret_cast<const double2*>(in + e0 + 2);
            v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) v[k] = e0 + k < nv ? in[e0 + k] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) s[4 * q + k] = __builtin_round(v[k] * r);
    }
    double dl[16];
    bool f[16];
    uint32_t loc[16], gid[16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        // the previous step's last element (lane 63), 0 at the tile start
        const double last = q > 0 ? readlane_d(s[q > 0 ? 4 * q - 1 : 0], 63) : 0.0;
        double p = wave_shr1_d(s[4 * q + 3]);
        if (l == 0) p = last;
        TCode c[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int i = 4 * q + k;
            const double delta = s[i] - (k ? s[i - 1] : p);
            loc[i] = 4u * (uint32_t)(l + 64 * q) + (uint32_t)k;
            gid[i] = (uint32_t)(base + loc[i]);
            const bool valid = base + loc[i] < n;
            const bool qz = fabs(delta) < rad;
            c[k] = qz ? enc_code64<TCode, ZZ>(delta) : (TCode)0;
            f[i] = valid && !qz;
            dl[i] = delta;
        }
        const uint64_t e0 = base + 4 * (uint64_t)(l + 64 * q);
        if (full_out) {
            if constexpr (sizeof(TCode) == 2) {
                *reinterpret_cast<uint2*>(codes + e0) =
                    make_uint2((uint32_t)c[0] | ((uint32_t)c[1] << 16), (uint32_t)c[2] | ((uint32_t)c[3] << 16));
            } else {
                *reinterpret_cast<uint4*>(codes + e0) = make_uint4(c[0], c[1], c[2], c[3]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (e0 + k < n) codes[e0 + k] = c[k];
        }
    }
    const uint64_t b = (uint64_t)tile * 1024;
    const uint32_t cnt = emit_outliers<16>(s_bits[wave_in_wg()], f, loc, dl, gid, tl.err + b, tl.idx + b);
    if (l == 0) tl.cnt[tile] = cnt;
}

// 2D (lorenzo_quant_nd.cu:29-105): t = s(x,y) - s(x,y-1) (0 above the
// tile), delta = t(x,y) - t(x-1,y) for x-1 inside the tile.
template <typename TCode, bool ZZ>
__global__ void __launch_bounds__(256)
lz2d_fwd64_kernel(const double* __restrict__ in, uint32_t nx, uint32_t ny, uint64_t n, double r_host,
                  const double* __restrict__ d_r, double rad, TCode* __restrict__ codes, TileLists64 tl, uint32_t gx,
                  uint32_t T) {
    __shared__ uint32_t s_bits[4][64];
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (tile >= T) return;
    const double r = d_r ? *d_r : r_host;
    const int l = lane_id();
    const int x = l & 31, h = l >> 5;
    const uint32_t bx = tile % gx, by = tile / gx;
    const uint32_t X = bx * 32 + (uint32_t)x, Y0 = by * 32 + 16 * (uint32_t)h;
    const bool xin = X < nx;
    double s[16];
#pragma unroll
    for (int i = 0; i < 16; i++)
        s[i] = (xin && Y0 + i < ny) ? __builtin_round(in[(uint64_t)(Y0 + i) * nx + X] * r) : 0.0;
    const double above = __shfl_xor(s[15], 32, 64);  // row 15 of the tile, for row 16
    double t[16];
#pragma unroll
    for (int i = 0; i < 16; i++) t[i] = s[i] - (i > 0 ? s[i - 1] : (h ? above : 0.0));
    double dl[16];
    bool f[16];
    uint32_t loc[16], gid[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const double left = __shfl_up(t[i], 1, 32);
        const double delta = x > 0 ? t[i] - left : t[i];
        const bool valid = xin && Y0 + i < ny;
        const bool qz = fabs(delta) < rad;
        if (valid) codes[(uint64_t)(Y0 + i) * nx + X] = qz ? enc_code64<TCode, ZZ>(delta) : (TCode)0;
        f[i] = valid && !qz;
        dl[i] = delta;
        loc[i] = (uint32_t)(16 * h + i) * 32u + (uint32_t)x;
        gid[i] = (uint32_t)((uint64_t)(Y0 + i) * nx + X);
    }
    if (tile == T - 1) {  // codes past the field (zero-padded input) are zero
        for (uint64_t e = (uint64_t)nx * ny + (uint64_t)l; e < n; e += 64) codes[e] = (TCode)0;
    }
    const uint64_t b = (uint64_t)tile * 1024;
    const uint32_t cnt = emit_outliers<16>(s_bits[wave_in_wg()], f, loc, dl, gid, tl.err + b, tl.idx + b);
    if (l == 0) tl.cnt[tile] = cnt;
}

// 3D (lorenzo_quant_nd.cu:199-274): a = s - s(z-1), b = a - a(x-1),
// delta = b - b(y-1), each neighbour inside the tile (else nothing is
// subtracted).
template <typename TCode, bool ZZ>
__global__ void __launch_bounds__(256)
lz3d_fwd64_kernel(const double* __restrict__ in, uint32_t nx, uint32_t ny, uint32_t nz, uint64_t n,
                  double r_host, const double* __restrict__ d_r, double rad, TCode* __restrict__ codes,
                  TileLists64 tl, Tiles tg, uint32_t T) {
    __shared__ uint32_t s_bits[4][64];
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (tile >= T) return;
    const double r = d_r ? *d_r : r_host;
    const int l = lane_id();
    const int x = l & 7, z = l >> 3;
    const uint32_t bx = tile % tg.gx, by = (tile / tg.gx) % tg.gy, bz = tile / (tg.gx * tg.gy);
    const uint32_t X = bx * 8 + (uint32_t)x, Z = bz * 8 + (uint32_t)z;
    const bool xz = X < nx && Z < nz;
    const uint64_t plane = (uint64_t)nx * ny;
    double s[8];
#pragma unroll
    for (int y = 0; y < 8; y++) {
        const uint32_t Y = by * 8 + (uint32_t)y;
        s[y] = (xz && Y < ny) ? __builtin_round(in[(uint64_t)Z * plane + (uint64_t)Y * nx + X] * r) : 0.0;
    }
    double a[8];
#pragma unroll
    for (int y = 0; y < 8; y++) {
        const double zm = __shfl_up(s[y], 8, 64);
        a[y] = s[y] - (z > 0 ? zm : 0.0);
    }
    double bb[8];
#pragma unroll
    for (int y = 0; y < 8; y++) {
        const double xm = __shfl_up(a[y], 1, 8);
        bb[y] = x > 0 ? a[y] - xm : a[y];
    }
    double dl[8];
    bool f[8];
    uint32_t loc[8], gid[8];
#pragma unroll
    for (int y = 0; y < 8; y++) {
        const uint32_t Y = by * 8 + (uint32_t)y;
        const double delta = y > 0 ? bb[y] - bb[y - 1] : bb[y];
        const bool valid = xz && Y < ny;
        const bool qz = fabs(delta) < rad;
        const uint64_t g = (uint64_t)Z * plane + (uint64_t)Y * nx + X;
        if (valid) codes[g] = qz ? enc_code64<TCode, ZZ>(delta) : (TCode)0;
        f[y] = valid && !qz;
        dl[y] = delta;
        loc[y] = (uint32_t)z * 64u + (uint32_t)y * 8u + (uint32_t)x;
        gid[y] = (uint32_t)g;
    }
    if (tile == T - 1) {
        for (uint64_t e = plane * nz + (uint64_t)l; e < n; e += 64) codes[e] = (TCode)0;
    }
    const uint64_t b = (uint64_t)tile * 512;
    const uint32_t cnt = emit_outliers<8>(s_bits[wave_in_wg()], f, loc, dl, gid, tl.err + b, tl.idx + b);
    if (l == 0) tl.cnt[tile] = cnt;
}

// ===========================================================================
// Inverse
// ===========================================================================

// The tile's outliers: the whole list (at most kDirectOutliers entries,
// every wave filtering its own), or the tile's binned range.
struct Outl64 {
    bool direct;
    const double* err;
    const uint32_t* idx;
    const uint32_t* count;
    uint32_t max_out;
    TileMap m;
    const uint32_t* off;
    const double* serr;
    const uint32_t* sloc;
};

// Adds the tile's outliers to the decoded deltas d[i] (local index loc_of(i))
// through this wave's LDS copy of the tile.  Outlier indices are distinct in
// any list the forward writes.
template <int NR, typename LOC>
__device__ __forceinline__ void add_outliers(double* st, const Outl64& o, uint32_t tile, double (&d)[NR], LOC loc_of) {
    const int l = lane_id();
    uint32_t kk = 0, o0 = 0, o1 = 0;
    bool any;
    if (o.direct) {
        kk = o.count ? min(*o.count, o.max_out) : 0u;
        any = kk > 0;
    } else {
        o0 = o.off[tile];
        o1 = o.off[tile + 1];
        any = o1 > o0;
    }
    if (!any) return;
#pragma unroll
    for (int i = 0; i < NR; i++) st[loc_of(i)] = d[i];
    wave_lds_sync();
    if (o.direct) {
        for (uint32_t j = (uint32_t)l; j < kk; j += 64) {
            uint32_t t, lc;
            if (tile_of(o.m, o.idx[j], t, lc) && t == tile) st[lc] = o.err[j] + st[lc];
        }
    } else {
        for (uint32_t j = o0 + (uint32_t)l; j < o1; j += 64) {
            const uint32_t lc = o.sloc[j];
            st[lc] = o.serr[j] + st[lc];
        }
    }
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < NR; i++) d[i] = st[loc_of(i)];
}

// 1D: per reference thread a sequential prefix of 4, per 32-lane warp a
// Hillis-Steele scan of the thread totals plus lane-1's inclusive, per
// block a Hillis-Steele scan of the 8 warp totals (lorenzo_quant.cu:17-87).
template <typename TCode, bool ZZ>
__global__ void __launch_bounds__(256)
lz1d_inv64_kernel(const TCode* __restrict__ codes, uint64_t n, double ebx2, Outl64 o, double* __restrict__ out,
                  uint32_t T) {
    __shared__ double s_tile[4][1024];
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (tile >= T) return;
    const int l = lane_id();
    const int h = l >> 5, lam = l & 31;
    const uint64_t base = (uint64_t)tile * 1024;
    const bool full = base + 1024 <= n;
    double d[16];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint64_t e0 = base + 4 * (uint64_t)(l + 64 * q);
        if (full) {
            TCode c[4];
            if constexpr (sizeof(TCode) == 2) {
                const uint2 w = *reinterpret_cast<const uint2*>(codes + e0);
                c[0] = (TCode)w.x; c[1] = (TCode)(w.x >> 16); c[2] = (TCode)w.y; c[3] = (TCode)(w.y >> 16);
            } else {
                const uint4 w = *reinterpret_cast<const uint4*>(codes + e0);
                c[0] = w.x; c[1] = w.y; c[2] = w.z; c[3] = w.w;
            }
#pragma unroll
            for (int k = 0; k < 4; k++) d[4 * q + k] = dec_code64<TCode, ZZ>(c[k]);
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) d[4 * q + k] = e0 + k < n ? dec_code64<TCode, ZZ>(codes[e0 + k]) : 0.0;
        }
    }
    add_outliers<16>(s_tile[wave_in_wg()], o, tile, d,
                     [&](int i) { return 4u * (uint32_t)(l + 64 * (i >> 2)) + (uint32_t)(i & 3); });
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int k = 1; k < 4; k++) d[4 * q + k] += d[4 * q + k - 1];
    double wt[8];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double ts = d[4 * q + 3];
#pragma unroll
        for (int off = 1; off < 32; off <<= 1) {
            const double v = __shfl_up(ts, off, 32);
            if (lam >= off) ts += v;
        }
        const double pre = __shfl_up(ts, 1, 32);
        if (lam > 0) {
#pragma unroll
            for (int k = 0; k < 4; k++) d[4 * q + k] += pre;
        }
        wt[2 * q] = readlane_d(d[4 * q + 3], 31);
        wt[2 * q + 1] = readlane_d(d[4 * q + 3], 63);
    }
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
        double nw[8];
#pragma unroll
        for (int i = 0; i < 8; i++) nw[i] = i >= off ? wt[i] + wt[i - off] : wt[i];
#pragma unroll
        for (int i = 0; i < 8; i++) wt[i] = nw[i];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int w = 2 * q + h;
        if (w > 0) {
            double add = 0.0;
#pragma unroll
            for (int i = 0; i < 7; i++)
                if (i == w - 1) add = wt[i];
#pragma unroll
            for (int k = 0; k < 4; k++) d[4 * q + k] += add;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint64_t e0 = base + 4 * (uint64_t)(l + 64 * q);
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = d[4 * q + k] * ebx2;
        if (full) {
            *reinterpret_cast<double2*>(out + e0) = make_double2(v[0], v[1]);
            *reinterpret_cast<double2*>(out + e0 + 2) = make_double2(v[2], v[3]);
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (e0 + k < n) out[e0 + k] = v[k];
        }
    }
}

// 2D (lorenzo_quant_nd.cu:112-190): y sequential inside each 8-row group,
// the group totals accumulated sequentially, then a Hillis-Steele scan over
// the 32 columns, times ebx2.
template <typename TCode, bool ZZ>
__global__ void __launch_bounds__(256)
lz2d_inv64_kernel(const TCode* __restrict__ codes, uint32_t nx, uint32_t ny, double ebx2, Outl64 o,
                  double* __restrict__ out, uint32_t gx, uint32_t T) {
    __shared__ double s_tile[4][1024];
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (tile >= T) return;
    const int l = lane_id();
    const int x = l & 31, h = l >> 5;
    const uint32_t bx = tile % gx, by = tile / gx;
    const uint32_t X = bx * 32 + (uint32_t)x, Y0 = by * 32 + 16 * (uint32_t)h;
    const bool xin = X < nx;
    double d[16];
#pragma unroll
    for (int i = 0; i < 16; i++)
        d[i] = (xin && Y0 + i < ny) ? dec_code64<TCode, ZZ>(codes[(uint64_t)(Y0 + i) * nx + X]) : 0.0;
    add_outliers<16>(s_tile[wave_in_wg()], o, tile, d,
                     [&](int i) { return (uint32_t)(16 * h + i) * 32u + (uint32_t)x; });
#pragma unroll
    for (int i = 1; i < 8; i++) d[i] += d[i - 1];
#pragma unroll
    for (int i = 9; i < 16; i++) d[i] += d[i - 1];
    // group addends: A1 = T0, A2 = T0 + T1, A3 = A2 + T2
    const double tlo = d[7], thi = d[15];
    const double a2 = __shfl(thi + tlo, x, 64);  // from the h = 0 lane of this column
    if (h == 0) {
#pragma unroll
        for (int i = 8; i < 16; i++) d[i] += tlo;
    } else {
        const double a3 = tlo + a2;
#pragma unroll
        for (int i = 0; i < 8; i++) d[i] += a2;
#pragma unroll
        for (int i = 8; i < 16; i++) d[i] += a3;
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        double v = d[i];
#pragma unroll
        for (int off = 1; off < 32; off <<= 1) {
            const double u = __shfl_up(v, off, 32);
            if (x >= off) v += u;
        }
        if (xin && Y0 + i < ny) out[(uint64_t)(Y0 + i) * nx + X] = v * ebx2;
    }
}

// 3D (lorenzo_quant_nd.cu:281-364): y sequential, then Hillis-Steele scans
// (offsets 1, 2, 4) along x and then along z, times ebx2.
template <typename TCode, bool ZZ>
__global__ void __launch_bounds__(256)
lz3d_inv64_kernel(const TCode* __restrict__ codes, uint32_t nx, uint32_t ny, uint32_t nz, double ebx2, Outl64 o,
                  double* __restrict__ out, Tiles tg, uint32_t T) {
    __shared__ double s_tile[4][512];
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + wave_in_wg();
    if (tile >= T) return;
    const int l = lane_id();
    const int x = l & 7, z = l >> 3;
    const uint32_t bx = tile % tg.gx, by = (tile / tg.gx) % tg.gy, bz = tile / (tg.gx * tg.gy);
    const uint32_t X = bx * 8 + (uint32_t)x, Z = bz * 8 + (uint32_t)z;
    const bool xz = X < nx && Z < nz;
    const uint64_t plane = (uint64_t)nx * ny;
    double d[8];
#pragma unroll
    for (int y = 0; y < 8; y++) {
        const uint32_t Y = by * 8 + (uint32_t)y;
        d[y] = (xz && Y < ny) ? dec_code64<TCode, ZZ>(codes[(uint64_t)Z * plane + (uint64_t)Y * nx + X]) : 0.0;
    }
    add_outliers<8>(s_tile[wave_in_wg()], o, tile, d,
                    [&](int y) { return (uint32_t)z * 64u + (uint32_t)y * 8u + (uint32_t)x; });
#pragma unroll
    for (int y = 1; y < 8; y++) d[y] += d[y - 1];
#pragma unroll
    for (int y = 0; y < 8; y++) {
        double v = d[y];
#pragma unroll
        for (int off = 1; off < 8; off <<= 1) {
            const double u = __shfl_up(v, off, 8);
            if (x >= off) v += u;
        }
#pragma unroll
        for (int off = 1; off < 8; off <<= 1) {
            const double u = __shfl_up(v, 8 * off, 64);
            if (z >= off) v += u;
        }
        const uint32_t Y = by * 8 + (uint32_t)y;
        if (xz && Y < ny) out[(uint64_t)Z * plane + (uint64_t)Y * nx + X] = v * ebx2;
    }
}

// ===========================================================================
// Value base (predictor_utils.cuh:27-117 with TInput = double): min/max in
// double with '<' / '>' from the first element, the base cast to float
// (predictor_utils.cuh:110-115), and the bound resolved in double
// (lorenzo_quant.cu:612-644: abs_eb = (double)user_eb * (double)base, or
// user_eb when base <= 0; ebx2_r = 1 / (2 abs_eb)).
// ===========================================================================
constexpr unsigned VB64_MAX_PARTIALS = 512;

__global__ void __launch_bounds__(256)
minmax64_partial_kernel(const double* __restrict__ in, uint64_t n, double2* __restrict__ partial, bool zero_pad) {
    __shared__ double smn[4], smx[4];
    const double first = in[0];
    double mn = first, mx = first;
    if (zero_pad) {
        if (0.0 < mn) mn = 0.0;
        if (0.0 > mx) mx = 0.0;
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n2 = n / 2;
    const double2* in2 = reinterpret_cast<const double2*>(in);
    auto fold = [&](double v) {
        if (v < mn) mn = v;
        if (v > mx) mx = v;
    };
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n2; i += 4 * stride) {
        const double2 a = in2[i], b = in2[i + stride], c = in2[i + 2 * stride], e = in2[i + 3 * stride];
        fold(a.x); fold(a.y); fold(b.x); fold(b.y); fold(c.x); fold(c.y); fold(e.x); fold(e.y);
    }
    for (; i < n2; i += stride) {
        const double2 a = in2[i];
        fold(a.x);
        fold(a.y);
    }
    for (uint64_t j = n2 * 2 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) fold(in[j]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double a = __shfl_xor(mn, off, 64), b = __shfl_xor(mx, off, 64);
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
        partial[blockIdx.x] = make_double2(mn, mx);
    }
}

__global__ void __launch_bounds__(64)
minmax64_final_kernel(const double2* __restrict__ partial, uint32_t np, int mode, float user_eb,
                      double* __restrict__ params) {
    double gmn = partial[0].x, gmx = partial[0].y;
    for (uint32_t b = threadIdx.x; b < np; b += 64) {
        const double2 v = partial[b];
        if (v.x < gmn) gmn = v.x;
        if (v.y > gmx) gmx = v.y;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double a = __shfl_xor(gmn, off, 64), b = __shfl_xor(gmx, off, 64);
        if (a < gmn) gmn = a;
        if (b > gmx) gmx = b;
    }
    if (threadIdx.x == 0) {
        float vb;
        if (mode == 2) {
            vb = (float)(gmx - gmn);
        } else {
            const double amin = gmn < 0.0 ? -gmn : gmn;
            const double amax = gmx < 0.0 ? -gmx : gmx;
            vb = (float)(amin > amax ? amin : amax);
        }
        const double abs_eb = vb <= 0.0f ? (double)user_eb : (double)user_eb * (double)vb;
        params[0] = (double)vb;
        params[1] = abs_eb;
        params[2] = 1.0 / (2.0 * abs_eb);
        params[3] = 2.0 * abs_eb;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <typename TCode>
int launch_fwd64(hipStream_t s, const double* in, size_t n, size_t nx, size_t ny, size_t nz, double r,
                 const double* d_r, uint32_t radius, int zz, void* codes, double* oerr, uint32_t* oidx,
                 uint32_t* ocount, size_t max_out, void* ws, size_t ws_bytes) {
    const FwdWs L = fwd_layout(n, nx, ny, nz, false, sizeof(double));
    if (ws_bytes < L.total) return fail(FZH_E_CAPACITY, "lorenzo fwd (f64): workspace too small");
    const size_t T = L.T;
    uint8_t* w = (uint8_t*)ws;
    const TileLists64 tl{(double*)(w + L.o_err), (uint32_t*)(w + L.o_idx), (uint32_t*)(w + L.o_cnt)};
    const double rad = (double)(TCode)radius;
    const unsigned blocks = (unsigned)cdiv(T, 4);
    TCode* c = (TCode*)codes;
    const int nd = ndim_of(ny, nz);
    const size_t nv = (nd == 1 && nx > 0 && nx < n) ? nx : n;
#define FZH_LZ64_FWD(ZZV)                                                                                     \
    if (nd == 1)                                                                                              \
        launch_k("lz1d_fwd64_kernel", lz1d_fwd64_kernel<TCode, ZZV>, dim3(blocks), dim3(256), 0, s, in, n, nv, r, \
                 d_r, rad, c, tl, (uint32_t)T);                                                               \
    else if (nd == 2)                                                                                         \
        launch_k("lz2d_fwd64_kernel", lz2d_fwd64_kernel<TCode, ZZV>, dim3(blocks), dim3(256), 0, s, in,           \
                 (uint32_t)nx, (uint32_t)ny, n, r, d_r, rad, c, tl, (uint32_t)cdiv(nx, 32), (uint32_t)T);        \
    else                                                                                                      \
        launch_k("lz3d_fwd64_kernel", lz3d_fwd64_kernel<TCode, ZZV>, dim3(blocks), dim3(256), 0, s, in,           \
                 (uint32_t)nx, (uint32_t)ny, (uint32_t)nz, n, r, d_r, rad, c, tl,                             \
                 Tiles{(uint32_t)cdiv(nx, 8), (uint32_t)cdiv(ny, 8), 0}, (uint32_t)T);
    if (zz) {
        FZH_LZ64_FWD(true)
    } else {
        FZH_LZ64_FWD(false)
    }
#undef FZH_LZ64_FWD
    FZH_CHECK_LAUNCH("lorenzo fwd (f64)");
    CompactListT<double> L0{tl.err, tl.idx, tl.cnt, oerr, oidx, ocount, max_out};
    return launch_compact<double>(s, w + L.o_ctx, T, (uint32_t)L.E, L0, nullptr);
}

template <typename TCode>
int launch_inv64(hipStream_t s, const void* codes, size_t n, size_t nx, size_t ny, size_t nz, const double* err,
                 const uint32_t* idx, const uint32_t* count, size_t max_out, double ebx2, int zz, double* out,
                 void* ws, size_t ws_bytes) {
    const InvWs L = inv_layout(n, nx, ny, nz, max_out, sizeof(double));
    if (ws_bytes < L.total) return fail(FZH_E_CAPACITY, "lorenzo inv (f64): workspace too small");
    if (int rc = inv_bin<double>(s, n, nx, ny, nz, err, idx, count, max_out, ws, ws_bytes)) return rc;
    uint8_t* w = (uint8_t*)ws;
    const int nd = ndim_of(ny, nz);
    const TileMap m = tile_map(n, nx, ny, nz);
    Outl64 o{};
    o.direct = max_out <= kDirectOutliers;
    o.err = err;
    o.idx = idx;
    o.count = (err && idx) ? count : nullptr;
    o.max_out = (uint32_t)max_out;
    o.m = m;
    o.off = (const uint32_t*)(w + L.o_off);
    o.serr = (const double*)(w + L.o_err);
    o.sloc = (const uint32_t*)(w + L.o_loc);
    const unsigned blocks = (unsigned)cdiv(L.T, 4);
    const TCode* c = (const TCode*)codes;
#define FZH_LZ64_INV(ZZV)                                                                                     \
    if (nd == 1)                                                                                              \
        launch_k("lz1d_inv64_kernel", lz1d_inv64_kernel<TCode, ZZV>, dim3(blocks), dim3(256), 0, s, c, n, ebx2, o, \
                 out, (uint32_t)L.T);                                                                         \
    else if (nd == 2)                                                                                         \
        launch_k("lz2d_inv64_kernel", lz2d_inv64_kernel<TCode, ZZV>, dim3(blocks), dim3(256), 0, s, c,            \
                 (uint32_t)nx, (uint32_t)ny, ebx2, o, out, m.gx, (uint32_t)L.T);                              \
    else                                                                                                      \
        launch_k("lz3d_inv64_kernel", lz3d_inv64_kernel<TCode, ZZV>, dim3(blocks), dim3(256), 0, s, c,            \
                 (uint32_t)nx, (uint32_t)ny, (uint32_t)nz, ebx2, o, out, Tiles{m.gx, m.gy, 0}, (uint32_t)L.T);
    if (zz) {
        FZH_LZ64_INV(true)
    } else {
        FZH_LZ64_INV(false)
    }
#undef FZH_LZ64_INV
    FZH_CHECK_LAUNCH("lorenzo inv (f64)");
    return FZH_OK;
}

}  // namespace
}  // namespace fzh

using namespace fzh;

extern "C" {

size_t fzh_lorenzo_quant_fwd_f64_ws_bytes(size_t n, size_t nx, size_t ny, size_t nz) {
    return fwd_layout(n, nx, ny, nz, false, sizeof(double)).total;
}

int fzh_lorenzo_quant_fwd_f64(void* stream, const double* d_in, size_t n, size_t nx, size_t ny, size_t nz,
                              double ebx2_r, const double* d_ebx2_r, uint32_t radius, int zigzag, int code_bytes,
                              void* d_codes, double* d_outlier_err, uint32_t* d_outlier_idx,
                              uint32_t* d_outlier_count, size_t max_outliers, void* d_ws, size_t ws_bytes) {
    if (code_bytes != 2 && code_bytes != 4) return fail(FZH_E_INVALID, "lorenzo fwd (f64): code_bytes must be 2 or 4");
    if (int rc = check_dims(n, nx, ny, nz, "lorenzo fwd (f64)")) return rc;
    if (n == 0) {
        if (d_outlier_count) return zero_bytes(S(stream), d_outlier_count, 4);
        return FZH_OK;
    }
    if (!d_in || !d_codes || !d_outlier_count || !d_ws || (max_outliers > 0 && (!d_outlier_err || !d_outlier_idx)))
        return fail(FZH_E_INVALID, "lorenzo fwd (f64): null pointer");
    if (!aligned16(d_in) || !aligned16(d_codes))
        return fail(FZH_E_INVALID, "lorenzo fwd (f64): input and codes must be 16-byte aligned");
    return code_bytes == 2
               ? launch_fwd64<uint16_t>(S(stream), d_in, n, nx, ny, nz, ebx2_r, d_ebx2_r, radius, zigzag, d_codes,
                                        d_outlier_err, d_outlier_idx, d_outlier_count, max_outliers, d_ws, ws_bytes)
               : launch_fwd64<uint32_t>(S(stream), d_in, n, nx, ny, nz, ebx2_r, d_ebx2_r, radius, zigzag, d_codes,
                                        d_outlier_err, d_outlier_idx, d_outlier_count, max_outliers, d_ws, ws_bytes);
}

size_t fzh_lorenzo_quant_inv_f64_ws_bytes(size_t n, size_t nx, size_t ny, size_t nz, size_t max_outliers) {
    return inv_layout(n, nx, ny, nz, max_outliers, sizeof(double)).total;
}

int fzh_lorenzo_quant_inv_f64(void* stream, const void* d_codes, size_t n, size_t nx, size_t ny, size_t nz,
                              const double* d_outlier_err, const uint32_t* d_outlier_idx,
                              const uint32_t* d_outlier_count, size_t max_outliers, double ebx2, int zigzag,
                              int code_bytes, double* d_out, void* d_ws, size_t ws_bytes) {
    if (code_bytes != 2 && code_bytes != 4) return fail(FZH_E_INVALID, "lorenzo inv (f64): code_bytes must be 2 or 4");
    if (int rc = check_dims(n, nx, ny, nz, "lorenzo inv (f64)")) return rc;
    if (n == 0) return FZH_OK;
    if (!d_codes || !d_out || !d_ws) return fail(FZH_E_INVALID, "lorenzo inv (f64): null pointer");
    if (!aligned16(d_codes) || !aligned16(d_out))
        return fail(FZH_E_INVALID, "lorenzo inv (f64): codes and output must be 16-byte aligned");
    return code_bytes == 2
               ? launch_inv64<uint16_t>(S(stream), d_codes, n, nx, ny, nz, d_outlier_err, d_outlier_idx,
                                        d_outlier_count, max_outliers, ebx2, zigzag, d_out, d_ws, ws_bytes)
               : launch_inv64<uint32_t>(S(stream), d_codes, n, nx, ny, nz, d_outlier_err, d_outlier_idx,
                                        d_outlier_count, max_outliers, ebx2, zigzag, d_out, d_ws, ws_bytes);
}

size_t fzh_value_base_f64_ws_bytes(size_t) { return VB64_MAX_PARTIALS * sizeof(double2); }

int fzh_value_base_f64(void* stream, const double* d_in, size_t n, int mode, float user_eb, double* d_params,
                       void* d_ws, size_t ws_bytes) {
    const bool zero_pad = (mode & FZH_VB_ZERO_PADDED) != 0;
    mode &= ~FZH_VB_ZERO_PADDED;
    if (mode != 1 && mode != 2) return fail(FZH_E_INVALID, "value base (f64): mode must be 1 (REL) or 2 (NOA)");
    if (n == 0) return fail(FZH_E_INVALID, "value base (f64): empty input");
    if (!d_in || !d_params || !d_ws) return fail(FZH_E_INVALID, "value base (f64): null pointer");
    if (!aligned16(d_in)) return fail(FZH_E_INVALID, "value base (f64): input must be 16-byte aligned");
    if (ws_bytes < fzh_value_base_f64_ws_bytes(n)) return fail(FZH_E_CAPACITY, "value base (f64): workspace too small");
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(n, 256 * 32), VB64_MAX_PARTIALS));
    double2* partial = (double2*)d_ws;
    launch_k("minmax64_partial_kernel", minmax64_partial_kernel, dim3(blocks), dim3(256), 0, S(stream), d_in,
             (uint64_t)n, partial, zero_pad);
    launch_k("minmax64_final_kernel", minmax64_final_kernel, dim3(1), dim3(64), 0, S(stream), (const double2*)partial,
             blocks, mode, user_eb, d_params);
    FZH_CHECK_LAUNCH("minmax (f64)");
    return FZH_OK;
}

}  // extern "C"

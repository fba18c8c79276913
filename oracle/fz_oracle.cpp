// fz_oracle.cpp — CPU restatement of the FZGPUModules stage-chain semantics.
//
// TEST INFRASTRUCTURE: this library is the checker for the HIP path.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Each routine below cites the reference file:line whose behaviour it
// restates.  Floating-point expressions keep the reference's operation order
// (compiled with -ffp-contract=off so no FMA contraction changes rounding).

#include "fz_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Code-type helpers.  Codes are stored as TCode (u8/u16/u32); the signed
// delta is cast through int and then to TCode (two's complement wrap), or
// zigzag-encoded in the signed type of the same width
// (modules/transforms/zigzag/zigzag.h:70-84).
// ---------------------------------------------------------------------------
template <typename S>
inline typename std::make_unsigned<S>::type zz_enc(S x) {
    using U = typename std::make_unsigned<S>::type;
    constexpr int W = sizeof(S) * 8;
    return (U)(((U)x << 1) ^ (U)(x >> (W - 1)));
}
template <typename S>
inline S zz_dec(typename std::make_unsigned<S>::type u) {
    using U = typename std::make_unsigned<S>::type;
    return (S)((U)(u >> 1) ^ (U)(-(S)(u & 1u)));
}

// (int)delta: exact for |delta| < 2^31; beyond that (reachable only with u32
// codes and a radius above 2^31) the conversion saturates, as the GPU's
// float-to-int conversion instructions do.
template <typename T>
inline int to_int_sat(T d) {
    if (d >= (T)2147483647.0) return 2147483647;
    if (d <= (T)-2147483648.0) return (int)(-2147483647 - 1);
    return (int)d;
}

template <typename T>
inline uint32_t encode_code(T delta, int zigzag, int code_bytes) {
    int q = to_int_sat(delta);
    switch (code_bytes) {
    case 1: return zigzag ? (uint32_t)zz_enc<int8_t>((int8_t)q) : (uint32_t)(uint8_t)q;
    case 2: return zigzag ? (uint32_t)zz_enc<int16_t>((int16_t)q) : (uint32_t)(uint16_t)q;
    default: return zigzag ? (uint32_t)zz_enc<int32_t>((int32_t)q) : (uint32_t)q;
    }
}
template <typename T>
inline T decode_code(uint32_t c, int zigzag, int code_bytes) {
    switch (code_bytes) {
    case 1: return zigzag ? (T)zz_dec<int8_t>((uint8_t)c) : (T)(int8_t)(uint8_t)c;
    case 2: return zigzag ? (T)zz_dec<int16_t>((uint16_t)c) : (T)(int16_t)(uint16_t)c;
    default: return zigzag ? (T)zz_dec<int32_t>(c) : (T)(int32_t)c;
    }
}
inline uint32_t load_code(const void* codes, size_t i, int code_bytes) {
    switch (code_bytes) {
    case 1: return ((const uint8_t*)codes)[i];
    case 2: return ((const uint16_t*)codes)[i];
    default: return ((const uint32_t*)codes)[i];
    }
}
inline void store_code(void* codes, size_t i, uint32_t c, int code_bytes) {
    switch (code_bytes) {
    case 1: ((uint8_t*)codes)[i] = (uint8_t)c; break;
    case 2: ((uint16_t*)codes)[i] = (uint16_t)c; break;
    default: ((uint32_t*)codes)[i] = c; break;
    }
}
// Radius is passed to the kernels as TCode (lorenzo_quant.cu:369-403), so it
// wraps for narrow code types.
template <typename T>
inline T radius_as_tcode(uint32_t radius, int code_bytes) {
    switch (code_bytes) {
    case 1: return (T)(uint8_t)radius;
    case 2: return (T)(uint16_t)radius;
    default: return (T)radius;
    }
}

// One outlier: the prediction error in the input type (TInput in
// lorenzo_quant.cu:270) and its element index.
template <typename T>
struct OutlierT { T err; uint32_t idx; };

// Quantize one predicted delta (lorenzo_quant.cu:321-350; _nd.cu:82-102).
template <typename T>
inline void quantize_delta(T delta, T rad, int zigzag, int code_bytes,
                           void* codes, size_t gid, std::vector<OutlierT<T>>& outl) {
    if (std::fabs(delta) < rad) {
        store_code(codes, gid, encode_code(delta, zigzag, code_bytes), code_bytes);
    } else {
        store_code(codes, gid, 0u, code_bytes);
        outl.push_back({delta, (uint32_t)gid});
    }
}

int g_threads = 0;
inline int nthreads() {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

// Collect per-tile outlier vectors in tile order into the output arrays.
template <typename T>
uint32_t flush_outliers(const std::vector<std::vector<OutlierT<T>>>& per_tile,
                        T* out_err, uint32_t* out_idx, size_t max_outliers) {
    size_t pos = 0;
    for (const auto& v : per_tile) {
        for (const auto& o : v) {
            if (pos < max_outliers) {
                out_err[pos] = o.err;
                out_idx[pos] = o.idx;
            }
            pos++;
        }
    }
    return (uint32_t)pos;
}

// ---------------------------------------------------------------------------
// Lorenzo forward
// ---------------------------------------------------------------------------

// 1D: 1024-element tiles, chain restarts at every tile start
// (lorenzo_quant.cu:287-312); s = round(x * r) with round() = half away
// from zero (lorenzo_quant.cu:292).
template <typename T>
void lorenzo_fwd_1d(const T* in, size_t n, T r, T rad, int zz,
                    int cb, void* codes, std::vector<std::vector<OutlierT<T>>>& per_tile) {
    const size_t nt = (n + 1023) / 1024;
    per_tile.assign(nt, {});
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long long t = 0; t < (long long)nt; t++) {
        const size_t base = (size_t)t * 1024;
        const size_t end = std::min(n, base + 1024);
        T prev = 0;
        for (size_t i = base; i < end; i++) {
            T cur = std::round(in[i] * r);
            T delta = cur - prev;
            quantize_delta(delta, rad, zz, cb, codes, i, per_tile[t]);
            prev = cur;
        }
    }
}

// 2D: 32x32 tiles (lorenzo_quant_nd.cu:29-105).  Operation order per element:
//   t(x,y) = s(x,y) - s(x,y-1)        (y-1 inside the tile, else 0)
//   delta  = t(x,y) - t(x-1,y)        (only if x-1 inside the tile)
template <typename T>
void lorenzo_fwd_2d(const T* in, size_t nx, size_t ny, T r, T rad,
                    int zz, int cb, void* codes,
                    std::vector<std::vector<OutlierT<T>>>& per_tile) {
    const size_t gx = (nx + 31) / 32, gy = (ny + 31) / 32;
    per_tile.assign(gx * gy, {});
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long long t = 0; t < (long long)(gx * gy); t++) {
        const size_t bx = (size_t)t % gx, by = (size_t)t / gx;
        T s[32][32];
        for (int ty = 0; ty < 32; ty++)
            for (int tx = 0; tx < 32; tx++) {
                size_t x = bx * 32 + tx, y = by * 32 + ty;
                s[ty][tx] = (x < nx && y < ny) ? std::round(in[y * nx + x] * r) : (T)0;
            }
        T tv[32][32];
        for (int ty = 0; ty < 32; ty++)
            for (int tx = 0; tx < 32; tx++)
                tv[ty][tx] = s[ty][tx] - (ty > 0 ? s[ty - 1][tx] : (T)0);
        for (int ty = 0; ty < 32; ty++)
            for (int tx = 0; tx < 32; tx++) {
                size_t x = bx * 32 + tx, y = by * 32 + ty;
                if (x >= nx || y >= ny) continue;
                T d = tv[ty][tx];
                if (tx > 0) d = d - tv[ty][tx - 1];
                quantize_delta(d, rad, zz, cb, codes, y * nx + x, per_tile[t]);
            }
    }
}

// 3D: 8x8x8 tiles (lorenzo_quant_nd.cu:199-274).  Order per element:
//   a = s(x,y,z) - s(x,y,z-1)       (z-1 in tile, else 0)
//   b = a(x,y,z) - a(x-1,y,z)       (x-1 in the 8-wide x segment)
//   d = b(x,y,z) - b(x,y-1,z)       (y-1 in tile; the reference multiplies
//                                    the ghost row by 0, i.e. subtracts 0)
// Canonical tile order: ((tz*GY + ty)*GX + tx) over 8-wide tiles.
template <typename T>
void lorenzo_fwd_3d(const T* in, size_t nx, size_t ny, size_t nz, T r,
                    T rad, int zz, int cb, void* codes,
                    std::vector<std::vector<OutlierT<T>>>& per_tile) {
    const size_t gx = (nx + 7) / 8, gy = (ny + 7) / 8, gz = (nz + 7) / 8;
    per_tile.assign(gx * gy * gz, {});
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long long t = 0; t < (long long)(gx * gy * gz); t++) {
        const size_t bx = (size_t)t % gx, by = ((size_t)t / gx) % gy, bz = (size_t)t / (gx * gy);
        T s[8][8][8], a[8][8][8], b[8][8][8];
        for (int z = 0; z < 8; z++)
            for (int y = 0; y < 8; y++)
                for (int x = 0; x < 8; x++) {
                    size_t X = bx * 8 + x, Y = by * 8 + y, Z = bz * 8 + z;
                    s[z][y][x] = (X < nx && Y < ny && Z < nz)
                                     ? std::round(in[(Z * ny + Y) * nx + X] * r) : (T)0;
                }
        for (int z = 0; z < 8; z++)
            for (int y = 0; y < 8; y++)
                for (int x = 0; x < 8; x++)
                    a[z][y][x] = s[z][y][x] - (z > 0 ? s[z - 1][y][x] : (T)0);
        for (int z = 0; z < 8; z++)
            for (int y = 0; y < 8; y++)
                for (int x = 0; x < 8; x++)
                    b[z][y][x] = x > 0 ? a[z][y][x] - a[z][y][x - 1] : a[z][y][x];
        for (int z = 0; z < 8; z++)
            for (int y = 0; y < 8; y++)
                for (int x = 0; x < 8; x++) {
                    size_t X = bx * 8 + x, Y = by * 8 + y, Z = bz * 8 + z;
                    if (X >= nx || Y >= ny || Z >= nz) continue;
                    T d = y > 0 ? b[z][y][x] - b[z][y - 1][x] : b[z][y][x];
                    quantize_delta(d, rad, zz, cb, codes, (Z * ny + Y) * nx + X, per_tile[t]);
                }
    }
}

// Hillis-Steele inclusive scan with shuffle-up semantics:
//   v'[i] = v[i] + v[i-off] for i >= off, offsets 1,2,4,... < len.
template <typename T>
inline void hs_scan(T* v, int len) {
    T tmp[64];
    for (int off = 1; off < len; off *= 2) {
        for (int i = 0; i < len; i++) tmp[i] = v[i];
        for (int i = off; i < len; i++) v[i] = tmp[i] + tmp[i - off];
    }
}

// ---------------------------------------------------------------------------
// Lorenzo inverse
// ---------------------------------------------------------------------------

// 1D tile scan tree of lorenzo_quant.cu:17-87 + 155-244 (TileDim 1024,
// Seq 4, 256 threads, 32-lane warps):
//   per thread: sequential sum of 4 consecutive deltas
//   per warp:   Hillis-Steele over 32 thread totals, add lane-1's inclusive
//   per block:  Hillis-Steele over 8 warp totals, add warp-1's inclusive
template <typename T>
void lorenzo_inv_1d_tile(T* d /*1024*/, T ebx2) {
    T v[256][4];
    for (int t = 0; t < 256; t++) {
        v[t][0] = d[4 * t];
        for (int i = 1; i < 4; i++) v[t][i] = v[t][i - 1] + d[4 * t + i];
    }
    for (int w = 0; w < 8; w++) {
        T ts[32];
        for (int l = 0; l < 32; l++) ts[l] = v[w * 32 + l][3];
        hs_scan(ts, 32);
        for (int l = 1; l < 32; l++)
            for (int i = 0; i < 4; i++) v[w * 32 + l][i] += ts[l - 1];
    }
    T wt[8];
    for (int w = 0; w < 8; w++) wt[w] = v[w * 32 + 31][3];
    hs_scan(wt, 8);
    for (int w = 1; w < 8; w++)
        for (int l = 0; l < 32; l++)
            for (int i = 0; i < 4; i++) v[w * 32 + l][i] += wt[w - 1];
    for (int t = 0; t < 256; t++)
        for (int i = 0; i < 4; i++) d[4 * t + i] = v[t][i] * ebx2;
}

// 2D tile scan of lorenzo_quant_nd.cu:112-190: y sequential within each
// 8-row warp group, warp totals accumulated sequentially, then an x
// Hillis-Steele scan over 32 lanes, then * ebx2.
template <typename T>
void lorenzo_inv_2d_tile(T d[32][32], T ebx2) {
    for (int x = 0; x < 32; x++) {
        T tot[4];
        for (int w = 0; w < 4; w++) {
            for (int r = 1; r < 8; r++) d[w * 8 + r][x] += d[w * 8 + r - 1][x];
            tot[w] = d[w * 8 + 7][x];
        }
        T acc[3] = {tot[0], tot[1], tot[2]};
        for (int i = 1; i < 3; i++) acc[i] += acc[i - 1];
        for (int w = 1; w < 4; w++)
            for (int r = 0; r < 8; r++) d[w * 8 + r][x] += acc[w - 1];
    }
    for (int y = 0; y < 32; y++) {
        hs_scan(d[y], 32);
        for (int x = 0; x < 32; x++) d[y][x] *= ebx2;
    }
}

// 3D tile scan of lorenzo_quant_nd.cu:281-364: y sequential, then per y an
// x Hillis-Steele scan over 8, then a z Hillis-Steele scan over 8, * ebx2.
template <typename T>
void lorenzo_inv_3d_tile(T d[8][8][8] /*[z][y][x]*/, T ebx2) {
    for (int z = 0; z < 8; z++)
        for (int x = 0; x < 8; x++)
            for (int y = 1; y < 8; y++) d[z][y][x] += d[z][y - 1][x];
    for (int y = 0; y < 8; y++) {
        for (int z = 0; z < 8; z++) hs_scan(d[z][y], 8);
        for (int x = 0; x < 8; x++) {
            T col[8];
            for (int z = 0; z < 8; z++) col[z] = d[z][y][x];
            hs_scan(col, 8);
            for (int z = 0; z < 8; z++) d[z][y][x] = col[z];
        }
    }
    for (int z = 0; z < 8; z++)
        for (int y = 0; y < 8; y++)
            for (int x = 0; x < 8; x++) d[z][y][x] *= ebx2;
}

template <typename T>
uint32_t lorenzo_fwd_any(const T* in, size_t n, size_t nx, size_t ny, size_t nz, T ebx2_r, uint32_t radius,
                         int zigzag, int code_bytes, void* codes, T* out_err, uint32_t* out_idx,
                         size_t max_outliers) {
    std::vector<std::vector<OutlierT<T>>> per_tile;
    const T rad = radius_as_tcode<T>(radius, code_bytes);
    if (n == 0) return 0;
    if (nz > 1)
        lorenzo_fwd_3d(in, nx, ny, nz, ebx2_r, rad, zigzag, code_bytes, codes, per_tile);
    else if (ny > 1)
        lorenzo_fwd_2d(in, nx, ny, ebx2_r, rad, zigzag, code_bytes, codes, per_tile);
    else
        lorenzo_fwd_1d(in, n, ebx2_r, rad, zigzag, code_bytes, codes, per_tile);
    return flush_outliers(per_tile, out_err, out_idx, max_outliers);
}

template <typename T>
void lorenzo_inv_any(const void* codes, size_t n, size_t nx, size_t ny, size_t nz, const T* err,
                     const uint32_t* idx, uint32_t count, size_t max_outliers, T ebx2, int zigzag, int cb,
                     T* out) {
    if (n == 0) return;
    // Scattered outlier deltas (lorenzo_quant.cu:426-450): zero + scatter of
    // min(count, capacity) entries.
    size_t total = (nz > 1) ? nx * ny * nz : (ny > 1) ? nx * ny : n;
    std::vector<T> sc(total, (T)0);
    const size_t k = std::min<size_t>(count, max_outliers);
    for (size_t t = 0; t < k; t++)
        if (idx[t] < total) sc[idx[t]] = err[t];

    if (nz > 1) {
        const size_t gx = (nx + 7) / 8, gy = (ny + 7) / 8, gz = (nz + 7) / 8;
#pragma omp parallel for schedule(static) num_threads(nthreads())
        for (long long t = 0; t < (long long)(gx * gy * gz); t++) {
            const size_t bx = (size_t)t % gx, by = ((size_t)t / gx) % gy, bz = (size_t)t / (gx * gy);
            T d[8][8][8];
            for (int z = 0; z < 8; z++)
                for (int y = 0; y < 8; y++)
                    for (int x = 0; x < 8; x++) {
                        size_t X = bx * 8 + x, Y = by * 8 + y, Z = bz * 8 + z;
                        if (X < nx && Y < ny && Z < nz) {
                            size_t g = (Z * ny + Y) * nx + X;
                            d[z][y][x] = sc[g] + decode_code<T>(load_code(codes, g, cb), zigzag, cb);
                        } else {
                            d[z][y][x] = (T)0;
                        }
                    }
            lorenzo_inv_3d_tile(d, ebx2);
            for (int z = 0; z < 8; z++)
                for (int y = 0; y < 8; y++)
                    for (int x = 0; x < 8; x++) {
                        size_t X = bx * 8 + x, Y = by * 8 + y, Z = bz * 8 + z;
                        if (X < nx && Y < ny && Z < nz) out[(Z * ny + Y) * nx + X] = d[z][y][x];
                    }
        }
    } else if (ny > 1) {
        const size_t gx = (nx + 31) / 32, gy = (ny + 31) / 32;
#pragma omp parallel for schedule(static) num_threads(nthreads())
        for (long long t = 0; t < (long long)(gx * gy); t++) {
            const size_t bx = (size_t)t % gx, by = (size_t)t / gx;
            T d[32][32];
            for (int y = 0; y < 32; y++)
                for (int x = 0; x < 32; x++) {
                    size_t X = bx * 32 + x, Y = by * 32 + y;
                    if (X < nx && Y < ny) {
                        size_t g = Y * nx + X;
                        d[y][x] = sc[g] + decode_code<T>(load_code(codes, g, cb), zigzag, cb);
                    } else {
                        d[y][x] = (T)0;
                    }
                }
            lorenzo_inv_2d_tile(d, ebx2);
            for (int y = 0; y < 32; y++)
                for (int x = 0; x < 32; x++) {
                    size_t X = bx * 32 + x, Y = by * 32 + y;
                    if (X < nx && Y < ny) out[Y * nx + X] = d[y][x];
                }
        }
    } else {
        const size_t nt = (n + 1023) / 1024;
#pragma omp parallel for schedule(static) num_threads(nthreads())
        for (long long t = 0; t < (long long)nt; t++) {
            T d[1024];
            const size_t base = (size_t)t * 1024;
            for (int i = 0; i < 1024; i++) {
                size_t g = base + i;
                d[i] = g < n ? decode_code<T>(load_code(codes, g, cb), zigzag, cb) + sc[g] : (T)0;
            }
            lorenzo_inv_1d_tile(d, ebx2);
            for (int i = 0; i < 1024; i++)
                if (base + i < n) out[base + i] = d[i];
        }
    }
}

// predictor_utils.cuh:27-117 — min/max with '<' / '>' comparisons in the
// input type, the value base returned as float (predictor_utils.cuh:110-115).
template <typename T>
float value_base_any(const T* in, size_t n, int mode) {
    if (n == 0) return 0.0f;
    T mn = in[0], mx = in[0];
    for (size_t i = 1; i < n; i++) {
        T v = in[i];
        if (v < mn) mn = v;
        if (v > mx) mx = v;
    }
    if (mode == 2) return (float)(mx - mn);
    T amin = mn < (T)0 ? -mn : mn;
    T amax = mx < (T)0 ? -mx : mx;
    return (float)(amin > amax ? amin : amax);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int orc_num_threads(void) { return nthreads(); }
void orc_set_num_threads(int n) { g_threads = n; }

uint32_t orc_lorenzo_fwd_f32(const float* in, size_t n, size_t nx, size_t ny,
                             size_t nz, float ebx2_r, uint32_t radius,
                             int zigzag, int code_bytes, void* codes,
                             float* out_err, uint32_t* out_idx,
                             size_t max_outliers) {
    return lorenzo_fwd_any<float>(in, n, nx, ny, nz, ebx2_r, radius, zigzag, code_bytes, codes, out_err, out_idx,
                                  max_outliers);
}

uint32_t orc_lorenzo_fwd_f64(const double* in, size_t n, size_t nx, size_t ny,
                             size_t nz, double ebx2_r, uint32_t radius,
                             int zigzag, int code_bytes, void* codes,
                             double* out_err, uint32_t* out_idx,
                             size_t max_outliers) {
    return lorenzo_fwd_any<double>(in, n, nx, ny, nz, ebx2_r, radius, zigzag, code_bytes, codes, out_err, out_idx,
                                   max_outliers);
}

void orc_lorenzo_inv_f32(const void* codes, size_t n, size_t nx, size_t ny,
                         size_t nz, const float* err, const uint32_t* idx,
                         uint32_t count, size_t max_outliers, float ebx2,
                         int zigzag, int cb, float* out) {
    lorenzo_inv_any<float>(codes, n, nx, ny, nz, err, idx, count, max_outliers, ebx2, zigzag, cb, out);
}

void orc_lorenzo_inv_f64(const void* codes, size_t n, size_t nx, size_t ny,
                         size_t nz, const double* err, const uint32_t* idx,
                         uint32_t count, size_t max_outliers, double ebx2,
                         int zigzag, int cb, double* out) {
    lorenzo_inv_any<double>(codes, n, nx, ny, nz, err, idx, count, max_outliers, ebx2, zigzag, cb, out);
}

float orc_value_base_f32(const float* in, size_t n, int mode) { return value_base_any<float>(in, n, mode); }
float orc_value_base_f64(const double* in, size_t n, int mode) { return value_base_any<double>(in, n, mode); }

// ---------------------------------------------------------------------------
// Bitshuffle (bitshuffle_stage.cu).  Element widths 1/2 use the ballot
// layout (:299-349): word (W-1-b)*npp + g holds bit b of elements
// 32g..32g+31, element 32g+j at bit j.  Widths 4/8 use the LC butterfly
// (:66-293), restated here by emulating the 32-lane xor shuffles.
// ---------------------------------------------------------------------------
static inline uint32_t byte_perm(uint32_t x, uint32_t y, uint32_t s) {
    uint8_t b[8];
    for (int i = 0; i < 4; i++) { b[i] = (uint8_t)(x >> (8 * i)); b[4 + i] = (uint8_t)(y >> (8 * i)); }
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) r |= (uint32_t)b[(s >> (4 * k)) & 7] << (8 * k);
    return r;
}

static void butterfly32(uint32_t a[32]) {
    uint32_t q[32];
    for (int l = 0; l < 32; l++) q[l] = a[l ^ 16];
    for (int l = 0; l < 32; l++)
        q[l] = ((l & 16) == 0) ? byte_perm(a[l], q[l], (3u << 12) | (2u << 8) | (7u << 4) | 6u)
                               : byte_perm(a[l], q[l], (5u << 12) | (4u << 8) | (1u << 4) | 0u);
    for (int l = 0; l < 32; l++) a[l] = q[l];
    for (int l = 0; l < 32; l++) q[l] = a[l ^ 8];
    for (int l = 0; l < 32; l++)
        q[l] = ((l & 8) == 0) ? byte_perm(a[l], q[l], (3u << 12) | (7u << 8) | (1u << 4) | 5u)
                              : byte_perm(a[l], q[l], (6u << 12) | (2u << 8) | (4u << 4) | 0u);
    for (int l = 0; l < 32; l++) a[l] = q[l];
    const uint32_t masks[3] = {0x0F0F0F0Fu, 0x33333333u, 0x55555555u};
    const int dist[3] = {4, 2, 1};
    for (int s = 0; s < 3; s++) {
        uint32_t m = masks[s];
        int dd = dist[s];
        for (int l = 0; l < 32; l++) q[l] = a[l ^ dd];
        for (int l = 0; l < 32; l++)
            q[l] = ((l & dd) == 0) ? ((a[l] & ~m) | ((q[l] >> dd) & m))
                                   : (((q[l] << dd) & ~m) | (a[l] & m));
        for (int l = 0; l < 32; l++) a[l] = q[l];
    }
}

static void butterfly64(uint64_t a0[32], uint64_t a1[32]) {
    for (int l = 0; l < 32; l++) {
        uint64_t b0 = a1[l], b1 = a0[l];
        uint64_t m = 0x00000000FFFFFFFFull;
        a0[l] = (a0[l] & ~m) | (b0 >> 32);
        a1[l] = (a1[l] & m) | (b1 << 32);
    }
    const uint64_t masks[5] = {0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull, 0x0F0F0F0F0F0F0F0Full,
                               0x3333333333333333ull, 0x5555555555555555ull};
    const int dist[5] = {16, 8, 4, 2, 1};
    uint64_t q0[32], q1[32];
    for (int s = 0; s < 5; s++) {
        const uint64_t m = masks[s];
        const int dd = dist[s];
        for (int l = 0; l < 32; l++) { q0[l] = a0[l ^ dd]; q1[l] = a1[l ^ dd]; }
        uint64_t n0[32], n1[32];
        for (int l = 0; l < 32; l++) {
            if ((l & dd) == 0) {
                n0[l] = (a0[l] & ~m) | ((q0[l] >> dd) & m);
                n1[l] = (a1[l] & ~m) | ((q1[l] >> dd) & m);
            } else {
                n0[l] = (a0[l] & m) | ((q0[l] << dd) & ~m);
                n1[l] = (a1[l] & m) | ((q1[l] << dd) & ~m);
            }
        }
        for (int l = 0; l < 32; l++) { a0[l] = n0[l]; a1[l] = n1[l]; }
    }
}

static void bitshuffle_block(const uint8_t* in, uint8_t* out, uint32_t bs, int ew, bool enc) {
    const uint32_t N = bs / ew;
    if (ew == 1 || ew == 2) {
        const int W = ew * 8;
        const uint32_t npp = N / 32;
        uint32_t* words = (uint32_t*)(enc ? out : (uint8_t*)in);
        if (enc) {
            for (uint32_t g = 0; g < npp; g++)
                for (int b = 0; b < W; b++) {
                    uint32_t w = 0;
                    for (int j = 0; j < 32; j++) {
                        uint32_t e = (ew == 1) ? in[32 * g + j] : ((const uint16_t*)in)[32 * g + j];
                        w |= ((e >> b) & 1u) << j;
                    }
                    std::memcpy(out + 4 * ((size_t)(W - 1 - b) * npp + g), &w, 4);
                }
        } else {
            for (uint32_t i = 0; i < N; i++) {
                uint32_t g = i / 32, j = i % 32, v = 0;
                for (int b = 0; b < W; b++) {
                    uint32_t w;
                    std::memcpy(&w, in + 4 * ((size_t)(W - 1 - b) * npp + g), 4);
                    v |= ((w >> j) & 1u) << b;
                }
                if (ew == 1) out[i] = (uint8_t)v; else ((uint16_t*)out)[i] = (uint16_t)v;
            }
        }
        (void)words;
    } else if (ew == 4) {
        const uint32_t npp = N / 32;
        for (uint32_t g = 0; g < npp; g++) {
            uint32_t a[32];
            for (int l = 0; l < 32; l++)
                std::memcpy(&a[l], in + 4 * (enc ? (32 * g + l) : (g + (size_t)l * npp)), 4);
            butterfly32(a);
            for (int l = 0; l < 32; l++)
                std::memcpy(out + 4 * (enc ? (g + (size_t)l * npp) : (32 * g + l)), &a[l], 4);
        }
    } else {
        const uint32_t npp = N / 64;
        for (uint32_t g = 0; g < npp; g++) {
            uint64_t a0[32], a1[32];
            for (int l = 0; l < 32; l++) {
                size_t i0 = enc ? (64 * (size_t)g + l) : (g + (size_t)l * npp);
                size_t i1 = enc ? (64 * (size_t)g + l + 32) : (g + (size_t)(l + 32) * npp);
                std::memcpy(&a0[l], in + 8 * i0, 8);
                std::memcpy(&a1[l], in + 8 * i1, 8);
            }
            butterfly64(a0, a1);
            for (int l = 0; l < 32; l++) {
                size_t i0 = enc ? (g + (size_t)l * npp) : (64 * (size_t)g + l);
                size_t i1 = enc ? (g + (size_t)(l + 32) * npp) : (64 * (size_t)g + l + 32);
                std::memcpy(out + 8 * i0, &a0[l], 8);
                std::memcpy(out + 8 * i1, &a1[l], 8);
            }
        }
    }
}

static void bitshuffle_all(const uint8_t* in, uint8_t* out, size_t nbytes, uint32_t bs, int ew, bool enc) {
    const size_t nblk = nbytes / bs;
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long long b = 0; b < (long long)nblk; b++)
        bitshuffle_block(in + (size_t)b * bs, out + (size_t)b * bs, bs, ew, enc);
    // Tail shorter than one block is copied verbatim (bitshuffle_stage.cu:460-469).
    std::memcpy(out + nblk * bs, in + nblk * bs, nbytes - nblk * bs);
}

void orc_bitshuffle_enc(const uint8_t* in, uint8_t* out, size_t nbytes, uint32_t bs, int ew) {
    bitshuffle_all(in, out, nbytes, bs, ew, true);
}
void orc_bitshuffle_dec(const uint8_t* in, uint8_t* out, size_t nbytes, uint32_t bs, int ew) {
    bitshuffle_all(in, out, nbytes, bs, ew, false);
}

// ---------------------------------------------------------------------------
// RZE (rze_stage.cu:135-445 encode, 457-551 decode, 774-994 host stream).
// ---------------------------------------------------------------------------
static constexpr int CS = 16384;

// Zero elimination: nonzero bytes in order + bitmap (bit i of byte i/8).
static int ze_encode(const uint8_t* in, int N, uint8_t* out, uint8_t* bm) {
    int w = 0;
    std::memset(bm, 0, (N + 7) / 8);
    for (int i = 0; i < N; i++)
        if (in[i]) { out[w++] = in[i]; bm[i >> 3] |= (uint8_t)(1u << (i & 7)); }
    return w;
}
// Repeat elimination: keep byte i if it differs from byte i-1 (byte -1 := 0).
static int re_encode(const uint8_t* in, int N, uint8_t* out, uint8_t* bm) {
    int w = 0;
    std::memset(bm, 0, (N + 7) / 8);
    for (int i = 0; i < N; i++) {
        uint8_t p = i > 0 ? in[i - 1] : 0;
        if (in[i] != p) { out[w++] = in[i]; bm[i >> 3] |= (uint8_t)(1u << (i & 7)); }
    }
    return w;
}

// Encodes one chunk into out (capacity CS); returns size (bit 31 = raw).
static uint32_t rze_encode_chunk(const uint8_t* in, int in_size, uint8_t* out) {
    uint8_t bm1[CS / 8], bm2[CS / 64], bm3[CS / 512], bm4[4];
    std::memset(bm4, 0, 4);
    auto raw = [&]() {
        std::memcpy(out, in, in_size);
        return (1u << 31) | (uint32_t)in_size;
    };
    int wpos = ze_encode(in, in_size, out, bm1);
    if (wpos == 0) {
        out[0] = (uint8_t)(in_size & 0xFF);
        out[1] = (uint8_t)((in_size >> 8) & 0xFF);
        return 2;
    }
    const int avail = CS - 2;
    const int bm1_bytes = (in_size + 7) / 8;
    if (wpos >= avail || wpos > avail - bm1_bytes) return raw();
    wpos += re_encode(bm1, bm1_bytes, out + wpos, bm2);
    if (wpos >= avail) return raw();
    const int bm2_bytes = (bm1_bytes + 7) / 8;
    if (wpos >= avail || wpos > avail - bm2_bytes) return raw();
    wpos += re_encode(bm2, bm2_bytes, out + wpos, bm3);
    if (wpos >= avail) return raw();
    const int bm3_bytes = (bm2_bytes + 7) / 8;
    if (wpos >= avail - 4 || wpos > (avail - 4) - bm3_bytes) return raw();
    wpos += re_encode(bm3, bm3_bytes, out + wpos, bm4);
    if (wpos >= avail - 4) return raw();
    std::memcpy(out + wpos, bm4, 4);
    wpos += 4;
    out[wpos] = (uint8_t)(in_size & 0xFF);
    out[wpos + 1] = (uint8_t)((in_size >> 8) & 0xFF);
    return (uint32_t)(wpos + 2);
}

size_t orc_rze_enc(const uint8_t* in, size_t nbytes, uint8_t* out) {
    if (nbytes == 0) return 0;
    const size_t nc = (nbytes + CS - 1) / CS;
    std::vector<uint8_t> scratch(nc * (size_t)CS);
    std::vector<uint32_t> sizes(nc);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads())
    for (long long c = 0; c < (long long)nc; c++) {
        int in_size = (int)std::min<size_t>(CS, nbytes - (size_t)c * CS);
        sizes[c] = rze_encode_chunk(in + (size_t)c * CS, in_size, scratch.data() + (size_t)c * CS);
    }
    const uint32_t hdr[2] = {(uint32_t)nbytes, (uint32_t)nc};
    std::memcpy(out, hdr, 8);
    std::memcpy(out + 8, sizes.data(), 4 * nc);
    size_t off = 8 + 4 * nc;
    std::vector<size_t> offs(nc);
    for (size_t c = 0; c < nc; c++) { offs[c] = off; off += sizes[c] & 0x7FFFFFFFu; }
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long long c = 0; c < (long long)nc; c++)
        std::memcpy(out + offs[c], scratch.data() + (size_t)c * CS, sizes[c] & 0x7FFFFFFFu);
    size_t total = (off + 3) & ~size_t(3);
    std::memset(out + off, 0, total - off);
    return total;
}

// Sequential RE decode (rze_stage.cu:283-298).
static bool re_decode(int decsize, const uint8_t* datain, int avail, const uint8_t* bm, uint8_t* out) {
    int pos = 0;
    uint8_t cur = 0;
    for (int i = 0; i < decsize; i++) {
        if ((bm[i >> 3] >> (i & 7)) & 1) {
            if (pos >= avail) return false;
            cur = datain[pos++];
        }
        out[i] = cur;
    }
    return true;
}
static int popcnt_bytes(const uint8_t* a, int n) {
    int c = 0;
    for (int i = 0; i < n; i++) c += __builtin_popcount(a[i]);
    return c;
}

static bool rze_decode_chunk(const uint8_t* in, uint32_t comp_size, uint32_t orig, uint8_t* out) {
    if (comp_size == 2) { std::memset(out, 0, orig); return true; }
    if (comp_size < 7) return false;
    const int bm1_bytes = ((int)orig + 7) / 8, bm2_bytes = (bm1_bytes + 7) / 8, bm3_bytes = (bm2_bytes + 7) / 8;
    uint8_t bm1[CS / 8], bm2[CS / 64], bm3[CS / 512], bm4[4];
    const int bm4_start = (int)comp_size - 2 - 4;
    std::memcpy(bm4, in + bm4_start, 4);
    const int bm3_start = bm4_start - popcnt_bytes(bm4, 4);
    if (bm3_start < 0 || !re_decode(bm3_bytes, in + bm3_start, bm4_start - bm3_start, bm4, bm3)) return false;
    const int bm2_start = bm3_start - popcnt_bytes(bm3, bm3_bytes);
    if (bm2_start < 0 || !re_decode(bm2_bytes, in + bm2_start, bm3_start - bm2_start, bm3, bm2)) return false;
    const int bm1_start = bm2_start - popcnt_bytes(bm2, bm2_bytes);
    if (bm1_start < 0 || !re_decode(bm1_bytes, in + bm1_start, bm2_start - bm1_start, bm2, bm1)) return false;
    int pos = 0;
    for (uint32_t i = 0; i < orig; i++) {
        if ((bm1[i >> 3] >> (i & 7)) & 1) {
            if (pos >= bm1_start) return false;
            out[i] = in[pos++];
        } else {
            out[i] = 0;
        }
    }
    return true;
}

size_t orc_rze_dec(const uint8_t* in, size_t in_bytes, uint8_t* out, size_t out_cap) {
    if (in_bytes < 8) return (size_t)-1;
    uint32_t orig, nc;
    std::memcpy(&orig, in, 4);
    std::memcpy(&nc, in + 4, 4);
    if (nc == 0 || orig == 0) return 0;
    if (orig > out_cap || 8 + 4 * (size_t)nc > in_bytes || nc != (orig + CS - 1) / CS) return (size_t)-1;
    std::vector<size_t> ioff(nc);
    size_t cur = 8 + 4 * (size_t)nc;
    for (uint32_t c = 0; c < nc; c++) {
        uint32_t e;
        std::memcpy(&e, in + 8 + 4 * c, 4);
        ioff[c] = cur;
        cur += e & 0x7FFFFFFFu;
    }
    if (cur > in_bytes) return (size_t)-1;
    bool ok = true;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads()) reduction(&& : ok)
    for (long long c = 0; c < (long long)nc; c++) {
        uint32_t e;
        std::memcpy(&e, in + 8 + 4 * c, 4);
        const uint32_t stored = e & 0x7FFFFFFFu;
        const uint32_t o = (c + 1 < (long long)nc) ? (uint32_t)CS : orig - (uint32_t)c * CS;
        if (e & 0x80000000u) {
            if (stored < o) { ok = false; continue; }
            std::memcpy(out + (size_t)c * CS, in + ioff[c], o);
        } else {
            ok = ok && rze_decode_chunk(in + ioff[c], stored, o, out + (size_t)c * CS);
        }
    }
    return ok ? orig : (size_t)-1;
}

// ---------------------------------------------------------------------------
// RLE (rle.cu): [num_runs u32][values T x R, zero-padded to 4][lengths u32 x R]
// ---------------------------------------------------------------------------
static inline uint32_t ld(const void* p, size_t i, int eb) { return load_code(p, i, eb); }

size_t orc_rle_enc(const void* in, size_t n, int eb, uint8_t* out) {
    if (n == 0) { uint32_t z = 0; std::memcpy(out, &z, 4); return 4; }
    std::vector<uint32_t> vals, lens;
    size_t start = 0;
    for (size_t i = 1; i <= n; i++) {
        if (i == n || ld(in, i, eb) != ld(in, i - 1, eb)) {
            vals.push_back(ld(in, start, eb));
            lens.push_back((uint32_t)(i - start));
            start = i;
        }
    }
    const uint32_t R = (uint32_t)vals.size();
    const size_t vb = (size_t)R * eb, va = (vb + 3) & ~size_t(3);
    std::memcpy(out, &R, 4);
    for (uint32_t r = 0; r < R; r++) store_code(out + 4, r, vals[r], eb);
    std::memset(out + 4 + vb, 0, va - vb);
    std::memcpy(out + 4 + va, lens.data(), 4 * (size_t)R);
    return 4 + va + 4 * (size_t)R;
}

size_t orc_rle_dec(const uint8_t* in, size_t in_bytes, int eb, void* out, size_t cap) {
    if (in_bytes < 4) return (size_t)-1;
    uint32_t R;
    std::memcpy(&R, in, 4);
    const size_t vb = (size_t)R * eb, va = (vb + 3) & ~size_t(3);
    if (4 + va + 4 * (size_t)R > in_bytes) return (size_t)-1;
    size_t pos = 0;
    for (uint32_t r = 0; r < R; r++) {
        uint32_t len;
        std::memcpy(&len, in + 4 + va + 4 * (size_t)r, 4);
        uint32_t v = ld(in + 4, r, eb);
        if (pos + len > cap) return (size_t)-1;
        for (uint32_t j = 0; j < len; j++) store_code(out, pos + j, v, eb);
        pos += len;
    }
    return pos;
}

// ---------------------------------------------------------------------------
// Histogram (histogram.cu:26-60): symbols >= bklen are skipped.
// ---------------------------------------------------------------------------
void orc_histogram(const void* in, size_t n, int eb, uint32_t bklen, uint32_t* freq) {
    std::memset(freq, 0, 4 * (size_t)bklen);
    for (size_t i = 0; i < n; i++) {
        uint32_t s = ld(in, i, eb);
        if (s < bklen) freq[s]++;
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Huffman codebook (phf/hf_bk_impl1.cc, hf_bk_internal.cc, hf_bk.cc,
// hf_canon.cc), restated.
// ---------------------------------------------------------------------------
namespace {

struct HNode {
    HNode* left = nullptr;
    HNode* right = nullptr;
    uint64_t freq = 0;
    uint32_t symbol = 0;
};

struct HHeap {  // 1-indexed binary heap, hf_bk_impl1.cc:73-106
    std::vector<HNode*> q;
    int qend = 1;
    explicit HHeap(size_t cap) : q(cap + 2, nullptr) {}
    void insert(HNode* n) {
        int j, i = qend++;
        while ((j = (i >> 1))) {
            if (q[j]->freq <= n->freq) break;
            q[i] = q[j];
            i = j;
        }
        q[i] = n;
    }
    HNode* remove() {
        int i = 1, l;
        HNode* n = q[1];
        if (qend < 2) return nullptr;
        qend--;
        q[i] = q[qend];
        while ((l = (i << 1)) < qend) {
            if (l + 1 < qend && q[l + 1]->freq < q[l]->freq) l++;
            if (q[i]->freq > q[l]->freq) {
                std::swap(q[i], q[l]);
                i = l;
            } else {
                break;
            }
        }
        return n;
    }
};

// 64-bit codebook word: prefix code in bits 0..57, bitcount in bits 58..63.
void build_codebook_u64(const uint32_t* freq_in, uint32_t bklen, uint64_t* book) {
    const size_t all_nodes = 4 * (size_t)bklen;
    std::vector<HNode> pool(2 * all_nodes);
    size_t used = 0;
    HHeap heap(2 * all_nodes);
    for (size_t i = 0; i < bklen; i++) {  // symbols >= bklen have zero frequency
        if (freq_in[i]) {
            HNode* n = &pool[used++];
            n->freq = freq_in[i];
            n->symbol = (uint32_t)i;
            heap.insert(n);
        }
    }
    while (heap.qend > 2) {
        // new_node(tree, 0, 0, qremove(tree), qremove(tree)): the host
        // compiler evaluates the last argument first, so the first node
        // removed becomes the right child.
        HNode* b = heap.remove();
        HNode* a = heap.remove();
        HNode* n = &pool[used++];
        n->left = a;
        n->right = b;
        n->freq = a->freq + b->freq;
        heap.insert(n);
    }
    for (uint32_t i = 0; i < bklen; i++) book[i] = ~0ull;
    HNode* root = heap.qend >= 2 ? heap.q[1] : nullptr;
    if (!root) return;
    // In-order traversal with an explicit stack (hf_bk_internal.cc:58-101),
    // left edge = 0, right edge = 1, depth limit 58.
    struct Fr { HNode* n; uint64_t path, len; };
    std::vector<Fr> st;
    HNode* p = root;
    uint64_t out1 = 0, len = 0;
    while (p) {
        if (p->left || p->right) {
            if (st.size() + 1 > 58) throw std::runtime_error("huffman: stack overflow");
            st.push_back({p, out1, len});
            p = p->left;
            out1 = (out1 << 1) | 0u;
            len += 1;
            if (len > 58) throw std::runtime_error("huffman: exceeding max len");
        } else {
            book[p->symbol] = (out1 & ((1ull << 58) - 1)) | (len << 58);
            if (!st.empty()) {
                Fr f = st.back();
                st.pop_back();
                out1 = f.path;
                len = f.len;
                p = f.n->right;
                out1 = (out1 << 1) | 1u;
                len += 1;
                if (len > 58) throw std::runtime_error("huffman: exceeding max len");
            } else {
                p = nullptr;
            }
        }
    }
}

}  // namespace

extern "C" {

size_t orc_huffman_revbook_bytes(uint32_t bklen, int eb) {
    return 4 * 64 + (size_t)eb * bklen;
}

void orc_huffman_book(const uint32_t* freq, uint32_t bklen, int eb, uint32_t* bk4, uint8_t* revbook) {
    std::vector<uint64_t> bk8(bklen);
    build_codebook_u64(freq, bklen, bk8.data());
    // Truncate to the 32-bit word (hf_bk.cc:104-122).
    for (uint32_t i = 0; i < bklen; i++) {
        if (bk8[i] == ~0ull) { bk4[i] = 0xFFFFFFFFu; continue; }
        uint32_t bc = (uint32_t)(bk8[i] >> 58);
        uint32_t pc = (uint32_t)(bk8[i] & ((1ull << 58) - 1));
        if (bc > 27) bk4[i] = 28u << 27;
        else bk4[i] = (bc << 27) | (pc & 0x07FFFFFFu);
    }
    // Canonize (hf_canon.cc:86-131), TYPE_BITS = 32.
    int numl[32] = {0}, iterby[32] = {0}, first[32] = {0}, entry[32] = {0};
    int max_l = 0;
    for (uint32_t i = 0; i < bklen; i++) {
        uint32_t c = bk4[i];
        int l = (int)(c >> 27);
        if (c != 0xFFFFFFFFu) { max_l = l > max_l ? l : max_l; numl[l] += 1; }
    }
    for (int i = 1; i < 32; i++) entry[i] = numl[i - 1];
    for (int i = 1; i < 32; i++) entry[i] += entry[i - 1];
    for (int i = 0; i < 32; i++) iterby[i] = entry[i];
    first[max_l] = 0;
    for (int l = max_l - 1; l >= 1; l--) first[l] = (int)((first[l + 1] + numl[l + 1]) / 2.0 + 0.5);
    first[0] = 0xff;
    std::vector<uint32_t> canon(bklen, 0xFFFFFFFFu), obk(bklen, 0xFFFFFFFFu);
    std::vector<uint32_t> keys(bklen, 0);
    for (uint32_t i = 0; i < bklen; i++) {
        uint32_t c = bk4[i];
        uint8_t l = (uint8_t)(c >> 27);
        if (c != 0xFFFFFFFFu) {
            uint32_t v = (uint32_t)(first[l] + iterby[l] - entry[l]);
            canon[iterby[l]] = (v & 0x07FFFFFFu) | ((uint32_t)l << 27);
            keys[iterby[l]] = i;
            iterby[l]++;
        }
    }
    for (uint32_t i = 0; i < bklen; i++)
        if (canon[i] != 0xFFFFFFFFu) obk[keys[i]] = canon[i];
    std::memcpy(bk4, obk.data(), 4 * (size_t)bklen);
    std::memcpy(revbook, first, 128);
    std::memcpy(revbook + 128, entry, 128);
    for (uint32_t i = 0; i < bklen; i++) store_code(revbook + 256, i, keys[i], eb);
}

size_t orc_huffman_encode_bound(size_t n, uint32_t bklen, int eb) {
    const size_t pardeg = n ? (n - 1) / 768 + 1 : 0;
    return 128 + orc_huffman_revbook_bytes(bklen, eb) + 8 * pardeg + 4 * (n + pardeg) + 64;
}

size_t orc_huffman_encode(const void* in, size_t n, int eb, uint32_t bklen, uint8_t* out, size_t cap) {
    return orc_huffman_encode_mode(in, n, eb, bklen, 0, out, cap);
}

// fine = 1: the Fine mode of HuffmanStage (huffman_stage.cu:41): 1024-symbol
// partitions (hf_buf.cc:48-49), taken only when no codeword is longer than
// 8 bits, else the coarse 768 (hf_hl.cc:69-80).  The ReVISIT-lite kernel
// (hf_kernels.cu:145-250) builds each partition's cells by the same MSB-first
// concatenation as the coarse deflate below.
size_t orc_huffman_encode_mode(const void* in, size_t n, int eb, uint32_t bklen, int fine, uint8_t* out,
                               size_t cap) {
    if (n == 0) return 0;
    std::vector<uint32_t> freq(bklen);
    orc_histogram(in, n, eb, bklen, freq.data());
    uint64_t fsum = 0;
    for (uint32_t i = 0; i < bklen; i++) fsum += freq[i];
    if (fsum != n) return 0;
    std::vector<uint32_t> bk4(bklen);
    const size_t rbytes = orc_huffman_revbook_bytes(bklen, eb);
    std::vector<uint8_t> rev(rbytes);
    orc_huffman_book(freq.data(), bklen, eb, bk4.data(), rev.data());

    uint32_t max_cl = 0;
    for (uint32_t i = 0; i < bklen; i++) max_cl = std::max(max_cl, bk4[i] >> 27);
    const size_t sublen = (fine && max_cl <= 8) ? 1024 : 768, pardeg = (n - 1) / sublen + 1;
    std::vector<uint32_t> nbit(pardeg), ncell(pardeg), entry(pardeg);
    std::vector<uint32_t> gapped(n + 1, 0);
    // Phase 2 deflate (hf_kernels.cu:73-122), one partition at a time.
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long long p = 0; p < (long long)pardeg; p++) {
        int residue = 32;
        uint32_t total = 0, bufr = 0;
        size_t ptr = (size_t)p * sublen;
        for (size_t i = 0; i < sublen; i++) {
            size_t did = (size_t)p * sublen + i;
            if (did == n) break;
            uint32_t w = bk4[ld(in, did, eb)];
            uint32_t width = w >> 27;
            w &= 0x07FFFFFFu;
            if (residue == 32) bufr = 0;
            if ((int)width <= residue) {
                residue -= (int)width;
                bufr |= residue >= 32 ? 0u : (w << residue);
                if (residue == 0) { residue = 32; gapped[ptr++] = bufr; }
            } else {
                int lb = (int)width - residue, rb = 32 - lb;
                bufr |= w >> lb;
                gapped[ptr++] = bufr;
                bufr = w << rb;
                residue = rb;
            }
            total += width;
        }
        if (ptr < n + 1) gapped[ptr] = bufr;
        nbit[p] = total;
        ncell[p] = (total + 31) / 32;
    }
    uint64_t total_nbit = 0, total_ncell = 0;
    for (size_t p = 0; p < pardeg; p++) {
        entry[p] = (uint32_t)total_ncell;
        total_ncell += ncell[p];
        total_nbit += nbit[p];
    }
    uint32_t ent[6];
    const uint32_t nb[5] = {128, (uint32_t)rbytes, (uint32_t)(4 * pardeg), (uint32_t)(4 * pardeg),
                            (uint32_t)(4 * total_ncell)};
    ent[0] = 0;
    for (int i = 1; i < 6; i++) ent[i] = ent[i - 1] + nb[i - 1];
    const size_t total = ent[5];
    if (total > cap) return 0;
    std::memset(out, 0, 128);
    int32_t bk16 = (int32_t)(bklen & 0xFFFF);
    int32_t sl = (int32_t)sublen, pd = (int32_t)pardeg;
    uint64_t ol = n;
    std::memcpy(out + 0, &bk16, 2);
    std::memcpy(out + 4, &sl, 4);
    std::memcpy(out + 8, &pd, 4);
    std::memcpy(out + 16, &ol, 8);
    std::memcpy(out + 24, &total_nbit, 8);
    std::memcpy(out + 32, &total_ncell, 8);
    std::memcpy(out + 40, ent, 24);
    std::memcpy(out + ent[1], rev.data(), rbytes);
    std::memcpy(out + ent[2], nbit.data(), 4 * pardeg);
    std::memcpy(out + ent[3], entry.data(), 4 * pardeg);
    for (size_t p = 0; p < pardeg; p++)
        std::memcpy(out + ent[4] + 4 * (size_t)entry[p], gapped.data() + p * sublen, 4 * (size_t)ncell[p]);
    return total;
}

// Canonical decode, one partition at a time (hf_kernels.cu:258-312).
int orc_huffman_decode(const uint8_t* in, size_t in_bytes, int eb, void* out, size_t n) {
    if (in_bytes < 128) return -1;
    int32_t sublen, pardeg;
    uint32_t ent[6];
    std::memcpy(&sublen, in + 4, 4);
    std::memcpy(&pardeg, in + 8, 4);
    std::memcpy(ent, in + 40, 24);
    if (ent[5] > in_bytes || sublen <= 0 || pardeg < 0) return -1;
    const uint32_t* first = (const uint32_t*)(in + ent[1]);
    const uint32_t* entry = first + 32;
    const uint8_t* keys = in + ent[1] + 256;
    const uint32_t* nbit = (const uint32_t*)(in + ent[2]);
    const uint32_t* pent = (const uint32_t*)(in + ent[3]);
    const uint32_t* cells = (const uint32_t*)(in + ent[4]);
    const size_t ncells = (ent[5] - ent[4]) / 4;
    const uint32_t nkeys = (ent[2] - ent[1] - 256) / eb;
    int bad = 0;
#pragma omp parallel for schedule(static) num_threads(nthreads()) reduction(| : bad)
    for (long long p = 0; p < pardeg; p++) {
        const size_t obase = (size_t)p * sublen;
        const size_t ocap = obase < n ? std::min<size_t>(sublen, n - obase) : 0;
        const uint32_t total_bw = nbit[p];
        const size_t cbase = pent[p];
        auto cell = [&](size_t k) -> uint32_t { return cbase + k < ncells ? cells[cbase + k] : 0u; };
        if (entry[1] == 1 && total_bw == 0) {  // single zero-length code
            for (size_t j = 0; j < ocap; j++) store_code(out, obase + j, ld(keys, 0, eb), eb);
            continue;
        }
        size_t idx_out = 0;
        uint32_t bufr = cell(0);
        uint32_t v = (bufr >> 31) & 1u;
        uint32_t l = 1, i = 0;
        while (i < total_bw) {
            while (v < first[l]) {
                ++i;
                if (i % 32 == 0) bufr = cell(i / 32);
                v = (v << 1) | ((bufr >> (31 - i % 32)) & 1u);
                ++l;
                if (l > 31) { bad = 1; break; }
            }
            if (l > 31) break;
            uint32_t k = entry[l] + v - first[l];
            if (idx_out >= ocap || k >= nkeys) { bad = 1; break; }
            store_code(out, obase + idx_out++, ld(keys, k, eb), eb);
            ++i;
            if (i % 32 == 0) bufr = cell(i / 32);
            v = (bufr >> (31 - i % 32)) & 1u;
            l = 1;
        }
    }
    return bad ? -1 : 0;
}

// ---------------------------------------------------------------------------
// Quantizer ABS (quantizer.cu:30-67, 116-137) + Bitpack (bitpack_stage.cu)
// ---------------------------------------------------------------------------
static inline int f2i_rn(float f) {  // __float2int_rn: nearest-even, saturating, NaN -> 0
    if (std::isnan(f)) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int)0x80000000;
    return (int)std::nearbyint(f);
}

uint32_t orc_quant_abs_fwd_f32(const float* in, size_t n, float r, float threshold, uint32_t radius,
                               int zigzag, int cb, void* codes, float* out_val, uint32_t* out_idx,
                               size_t max_outliers) {
    uint32_t cnt = 0;
    const int rad = (int)(uint32_t)(cb == 1 ? (uint8_t)radius : cb == 2 ? (uint16_t)radius : radius);
    for (size_t i = 0; i < n; i++) {
        float x = in[i];
        int q = f2i_rn(x * r);
        if (q > -rad && q < rad && std::fabs(x) < threshold) {
            uint32_t c;
            if (cb == 1) c = zigzag ? zz_enc<int8_t>((int8_t)q) : (uint8_t)q;
            else if (cb == 2) c = zigzag ? zz_enc<int16_t>((int16_t)q) : (uint16_t)q;
            else c = zigzag ? zz_enc<int32_t>(q) : (uint32_t)q;
            store_code(codes, i, c, cb);
        } else {
            store_code(codes, i, 0, cb);
            if (cnt < max_outliers) { out_val[cnt] = x; out_idx[cnt] = (uint32_t)i; }
            cnt++;
        }
    }
    return cnt;
}

void orc_quant_abs_inv_f32(const void* codes, size_t n, float ebx2, int zigzag, int cb, const float* vals,
                           const uint32_t* idx, uint32_t count, size_t max_outliers, float* out) {
    for (size_t i = 0; i < n; i++) {
        uint32_t c = ld(codes, i, cb);
        int q;
        if (cb == 1) q = zigzag ? zz_dec<int8_t>((uint8_t)c) : (int8_t)c;
        else if (cb == 2) q = zigzag ? zz_dec<int16_t>((uint16_t)c) : (int16_t)c;
        else q = zigzag ? zz_dec<int32_t>(c) : (int32_t)c;
        out[i] = (float)q * ebx2;
    }
    const size_t k = std::min<size_t>(count, max_outliers);
    for (size_t t = 0; t < k; t++)
        if (idx[t] < n) out[idx[t]] = vals[t];
}

// ---------------------------------------------------------------------------
// Quantizer REL (quantizer.cu:171-352, 426-436, 563-580): log2-space bins.
// The reference evaluates log2 with __log2f and 2^t with exp2f (CUDA,
// within 2 ulp); this restatement uses the correctly rounded float results
// (the double-precision functions, rounded once to float), as the HIP
// kernels do.  The host constants are the reference's float expressions.
// ---------------------------------------------------------------------------
void orc_quant_rel_params(float eps, float* p /*[4]*/) {
    const float log2eb = 2.0f * std::log2(1.0f + eps);  // quantizer.cu:565
    p[0] = log2eb;
    p[1] = 1.0f / log2eb;
    p[2] = 1.0f + eps;
    p[3] = 1.0f / (1.0f + eps);
}

static inline float log2_cr(float x) { return (float)std::log2((double)x); }
static inline float exp2_cr(float t) { return (float)std::exp2((double)t); }

uint32_t orc_quant_rel_fwd_f32(const float* in, size_t n, float eps, float threshold, uint32_t radius, int cb,
                               void* codes, float* out_val, uint32_t* out_idx, size_t max_outliers) {
    float p[4];
    orc_quant_rel_params(eps, p);
    const float log2eb = p[0], log2eb_r = p[1], opp = p[2], oopp = p[3];
    const uint32_t rad = cb == 1 ? (uint8_t)radius : cb == 2 ? (uint16_t)radius : radius;
    const uint32_t cmax = cb == 1 ? 0xFFu : cb == 2 ? 0xFFFFu : 0xFFFFFFFFu;
    uint32_t cnt = 0;
    for (size_t i = 0; i < n; i++) {
        const float x = in[i];
        uint32_t bits;
        std::memcpy(&bits, &x, 4);
        const uint32_t expo = (bits >> 23) & 0xFFu;
        bool ok = false;
        uint32_t stored = 0;
        if (!(std::fabs(x) >= threshold) && expo != 0 && expo != 255) {  // quantizer.cu:249-275
            const bool neg = x < 0.0f;
            const float ax = neg ? -x : x;
            const int lb = f2i_rn(log2_cr(ax) * log2eb_r);
            const uint32_t alb = lb < 0 ? 0u - (uint32_t)lb : (uint32_t)lb;
            stored = ((alb << 2) | ((lb < 0 ? 1u : 0u) << 1) | (neg ? 1u : 0u)) + 1u;  // quantizer.cu:224-230
            if (alb < rad && stored <= cmax) {
                const float rec = exp2_cr((float)lb * log2eb);
                ok = rec >= ax * oopp && rec <= ax * opp;  // quantizer.cu:300-311
            }
        }
        if (ok) {
            store_code(codes, i, stored, cb);
        } else {
            store_code(codes, i, 0, cb);
            if (cnt < max_outliers) { out_val[cnt] = x; out_idx[cnt] = (uint32_t)i; }
            cnt++;
        }
    }
    return cnt;
}

void orc_quant_rel_inv_f32(const void* codes, size_t n, float eps, int cb, const float* vals, const uint32_t* idx,
                           uint32_t count, size_t max_outliers, float* out) {
    float p[4];
    orc_quant_rel_params(eps, p);
    for (size_t i = 0; i < n; i++) {
        const uint32_t st = ld(codes, i, cb);
        if (st == 0) { out[i] = 0.0f; continue; }  // quantizer.cu:340-342
        const uint32_t cv = st - 1u;
        const int alb = (int)(cv >> 2);
        const int lb = ((cv >> 1) & 1u) ? -alb : alb;
        const float rec = exp2_cr((float)lb * p[0]);
        out[i] = (cv & 1u) ? -rec : rec;
    }
    const size_t k = std::min<size_t>(count, max_outliers);
    for (size_t t = 0; t < k; t++)
        if (idx[t] < n) out[idx[t]] = vals[t];
}

int orc_bitpack_nbits_for_max(uint32_t max_val, int eb) {
    const int full = 8 * eb;
    for (int b = 1; b < full; b *= 2)
        if ((uint64_t)max_val < (1ull << b)) return b;
    return full;
}

size_t orc_bitpack_enc(const void* in, size_t n, int eb, int nbits, uint8_t* out) {
    const size_t nout = (n * nbits + 7) / 8;
    if (nbits < 8) {
        const int epb = 8 / nbits;
        const uint32_t mask = (1u << nbits) - 1u;
        for (size_t b = 0; b < nout; b++) {
            uint8_t packed = 0;
            for (int k = 0; k < epb; k++) {
                size_t e = b * epb + k;
                if (e < n) packed |= (uint8_t)((ld(in, e, eb) & mask) << (k * nbits));
            }
            out[b] = packed;
        }
    } else {
        const int bytes = nbits / 8;
        const uint64_t mask = nbits >= 32 ? 0xFFFFFFFFull : ((1ull << nbits) - 1);
        for (size_t i = 0; i < n; i++) {
            uint64_t v = ld(in, i, eb) & mask;
            for (int b = 0; b < bytes; b++) { out[i * bytes + b] = (uint8_t)(v & 0xFF); v >>= 8; }
        }
    }
    return nout;
}

void orc_bitpack_dec(const uint8_t* in, size_t n, int eb, int nbits, void* out) {
    if (nbits < 8) {
        const int epb = 8 / nbits;
        const uint32_t mask = (1u << nbits) - 1u;
        for (size_t i = 0; i < n; i++)
            store_code(out, i, (in[i / epb] >> ((i % epb) * nbits)) & mask, eb);
    } else {
        const int bytes = nbits / 8;
        for (size_t i = 0; i < n; i++) {
            uint32_t v = 0;
            for (int b = 0; b < bytes; b++) v |= (uint32_t)in[i * bytes + b] << (8 * b);
            store_code(out, i, v, eb);
        }
    }
}

// IEEE 802.3 CRC32, reflected polynomial 0xEDB88320 (compressor_io.cpp:22-68).
uint32_t orc_crc32(const uint8_t* data, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; i++) c = table[(c ^ data[i]) & 0xFF] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

size_t orc_chain_l2d_bs_rze(const float* in, size_t n_padded, size_t nx, size_t ny, float ebx2_r,
                            uint32_t radius, int zigzag, uint16_t* codes_tmp, uint8_t* shuf_tmp,
                            uint8_t* out, float* out_err, uint32_t* out_idx, size_t max_outliers,
                            uint32_t* n_outliers) {
    std::memset(codes_tmp + nx * ny, 0, 2 * (n_padded - nx * ny));
    *n_outliers = orc_lorenzo_fwd_f32(in, n_padded, nx, ny, 1, ebx2_r, radius, zigzag, 2, codes_tmp,
                                      out_err, out_idx, max_outliers);
    orc_bitshuffle_enc((const uint8_t*)codes_tmp, shuf_tmp, 2 * n_padded, 16384, 2);
    return orc_rze_enc(shuf_tmp, 2 * n_padded, out);
}

}  // extern "C"

// transform.hip — the element-wise and prefix-sum stages: Zigzag, Negabinary
// (modules/transforms/zigzag/zigzag.h:70-80, modules/transforms/negabinary/
// negabinary.h), Difference with optional chunking and fused negabinary
// output (modules/predictors/diff/diff.cu:22-240), and the integer Lorenzo
// predictor (modules/predictors/lorenzo/lorenzo_stage.cu).
//
// Element-wise kernels move 16-byte vectors.  Prefix sums run in three
// launches over 512-element tiles (tile sums; per-chunk left fold of the
// tile sums; tiles with their prefix), with the FP32/FP64 association of the
// reference's inverse difference, which runs cub::BlockScan<T, 512> (raking:
// 32 segments of 16 summed left to right, a Kogge-Stone scan over the 32
// segment sums, then each segment's elements added left to right onto its
// exclusive prefix) tile by tile and adds the running sum of the earlier
// tiles' aggregates.  Integer sums wrap, so their association is free.
#include "fzh_internal.h"

namespace fzh {
namespace {

template <int EB> struct UIntOf;
template <> struct UIntOf<1> { using type = uint8_t; };
template <> struct UIntOf<2> { using type = uint16_t; };
template <> struct UIntOf<4> { using type = uint32_t; };
template <> struct UIntOf<8> { using type = uint64_t; };
template <int EB> struct SIntOf;
template <> struct SIntOf<1> { using type = int8_t; };
template <> struct SIntOf<2> { using type = int16_t; };
template <> struct SIntOf<4> { using type = int32_t; };
template <> struct SIntOf<8> { using type = int64_t; };

// zigzag.h:70-80
template <typename S>
__device__ __forceinline__ typename std::make_unsigned<S>::type zz_enc(S x) {
    using U = typename std::make_unsigned<S>::type;
    constexpr int W = 8 * sizeof(S);
    return (U)((U)((U)x << 1) ^ (U)(x >> (W - 1)));
}
template <typename S>
__device__ __forceinline__ S zz_dec(typename std::make_unsigned<S>::type u) {
    using U = typename std::make_unsigned<S>::type;
    return (S)(U)((U)(u >> 1) ^ (U)(-(S)(u & 1u)));
}
// negabinary.h: (n + M) ^ M and (u ^ M) - M with M = 0xAA..AA
template <typename S>
__device__ __forceinline__ typename std::make_unsigned<S>::type nb_enc(S n) {
    using U = typename std::make_unsigned<S>::type;
    constexpr U M = (U)0xAAAAAAAAAAAAAAAAull;
    return (U)((U)((U)n + M) ^ M);
}
template <typename S>
__device__ __forceinline__ S nb_dec(typename std::make_unsigned<S>::type u) {
    using U = typename std::make_unsigned<S>::type;
    constexpr U M = (U)0xAAAAAAAAAAAAAAAAull;
    return (S)(U)((U)(u ^ M) - M);
}

// ---------------------------------------------------------------------------
// Element-wise codes: OP 0 zigzag encode, 1 zigzag decode, 2 negabinary
// encode, 3 negabinary decode.  16-byte vectors, grid-stride; the < 16-byte
// tail by the first thread.
// ---------------------------------------------------------------------------
template <int EB, int OP>
__device__ __forceinline__ typename UIntOf<EB>::type code_op(typename UIntOf<EB>::type v) {
    using U = typename UIntOf<EB>::type;
    using S = typename SIntOf<EB>::type;
    if constexpr (OP == 0) return zz_enc<S>((S)v);
    else if constexpr (OP == 1) return (U)zz_dec<S>(v);
    else if constexpr (OP == 2) return nb_enc<S>((S)v);
    else return (U)nb_dec<S>(v);
}

template <int EB, int OP>
__global__ void __launch_bounds__(256) code_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                                   uint64_t n) {
    using U = typename UIntOf<EB>::type;
    constexpr int PER = 16 / EB;
    const uint64_t nv = n / PER;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
        uint4 v = reinterpret_cast<const uint4*>(in)[i];
        U e[PER];
        __builtin_memcpy(e, &v, 16);
#pragma unroll
        for (int k = 0; k < PER; k++) e[k] = code_op<EB, OP>(e[k]);
        __builtin_memcpy(&v, e, 16);
        reinterpret_cast<uint4*>(out)[i] = v;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (uint64_t k = nv * PER; k < n; k++)
            reinterpret_cast<U*>(out)[k] = code_op<EB, OP>(reinterpret_cast<const U*>(in)[k]);
}

template <int OP>
int launch_code(hipStream_t s, const void* in, void* out, size_t n, int eb) {
    if (n == 0) return FZH_OK;
    if (!in || !out) return fail(FZH_E_INVALID, "transform: null pointer");
    if (((uintptr_t)in | (uintptr_t)out) % 16 != 0) return fail(FZH_E_INVALID, "transform: buffers must be 16-byte aligned");
    const unsigned grid = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(n * eb, 16 * 256), 8192));
    switch (eb) {
    case 1: launch_k("code_kernel", code_kernel<1, OP>, dim3(grid), dim3(256), 0, s, (const uint8_t*)in, (uint8_t*)out, (uint64_t)n); break;
    case 2: launch_k("code_kernel", code_kernel<2, OP>, dim3(grid), dim3(256), 0, s, (const uint8_t*)in, (uint8_t*)out, (uint64_t)n); break;
    case 4: launch_k("code_kernel", code_kernel<4, OP>, dim3(grid), dim3(256), 0, s, (const uint8_t*)in, (uint8_t*)out, (uint64_t)n); break;
    case 8: launch_k("code_kernel", code_kernel<8, OP>, dim3(grid), dim3(256), 0, s, (const uint8_t*)in, (uint8_t*)out, (uint64_t)n); break;
    default: return fail(FZH_E_INVALID, "transform: element size must be 1, 2, 4 or 8");
    }
    FZH_CHECK_LAUNCH("code_kernel");
    return FZH_OK;
}

// ---------------------------------------------------------------------------
// Difference forward (diff.cu:22-45): out[i] = in[i] - in[i-1], the first
// element of every chunk (and of the array) as is; NB: negabinary output.
// T arithmetic as the reference's (wrapping for integers).
// ---------------------------------------------------------------------------
template <typename T, bool NB>
__global__ void __launch_bounds__(256) diff_fwd_kernel(const T* __restrict__ in, void* __restrict__ out, uint64_t n,
                                                       uint64_t chunk_mask) {
    constexpr int PER = 16 / sizeof(T);
    const uint64_t nv = n / PER;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    auto one = [&](uint64_t i, T cur, T prev) -> T {
        const bool boundary = i == 0 || (chunk_mask != ~0ull && (i & chunk_mask) == 0);
        return boundary ? cur : (T)(cur - prev);
    };
    auto store = [&](uint64_t i, T d) {
        if constexpr (NB) {
            using S = T;
            reinterpret_cast<typename std::make_unsigned<S>::type*>(out)[i] = nb_enc<S>(d);
        } else {
            reinterpret_cast<T*>(out)[i] = d;
        }
    };
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += stride) {
        const uint64_t i0 = v * PER;
        T e[PER];
        const uint4 x = reinterpret_cast<const uint4*>(in)[v];
        __builtin_memcpy(e, &x, 16);
        T prev = i0 ? in[i0 - 1] : (T)0;
        T d[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            d[k] = one(i0 + k, e[k], prev);
            prev = e[k];
        }
        if constexpr (NB) {
            using U = typename std::make_unsigned<T>::type;
            U u[PER];
#pragma unroll
            for (int k = 0; k < PER; k++) u[k] = nb_enc<T>(d[k]);
            uint4 y;
            __builtin_memcpy(&y, u, 16);
            reinterpret_cast<uint4*>(out)[v] = y;
        } else {
            uint4 y;
            __builtin_memcpy(&y, d, 16);
            reinterpret_cast<uint4*>(out)[v] = y;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (uint64_t i = nv * PER; i < n; i++) store(i, one(i, in[i], i ? in[i - 1] : (T)0));
}

// ---------------------------------------------------------------------------
// Prefix sums in the reference's tile association (see the file comment).
// A tile is 512 elements of one chunk (the chunk's last tile may be short:
// its missing elements are +0, as the reference's padded threads).  Lane
// s < 32 of a wave owns segment s (elements 16s .. 16s+15) of one tile;
// lanes 32..63 the next tile.
// ---------------------------------------------------------------------------
constexpr int kTile = 512;

struct TileGeom {
    uint64_t n, chunk;  // elements; chunk length (n when unchunked)
    uint64_t tpc;       // tiles per chunk
    uint64_t ntiles;    // nchunks * tpc
};

__device__ __forceinline__ bool tile_origin(const TileGeom& g, uint64_t tile, uint64_t& base, uint32_t& len) {
    const uint64_t c = tile / g.tpc, k = tile % g.tpc;
    const uint64_t cb = c * g.chunk;
    const uint64_t clen = min(g.chunk, g.n - cb);
    const uint64_t off = k * kTile;
    if (off >= clen) return false;
    base = cb + off;
    len = (uint32_t)min((uint64_t)kTile, clen - off);
    return true;
}

// Element loads of a segment (NB: negabinary decode first, diff.cu:205-214).
template <typename T, bool NB>
__device__ __forceinline__ T load_elem(const void* in, uint64_t i) {
    if constexpr (NB) return nb_dec<T>(reinterpret_cast<const typename std::make_unsigned<T>::type*>(in)[i]);
    else return reinterpret_cast<const T*>(in)[i];
}

// Kogge-Stone inclusive scan over the 32 lanes of each half wave (lane
// order: the lower lane's value is the left operand, as shfl-up scans add).
template <typename T>
__device__ __forceinline__ T half_wave_incl(T v, int s) {
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const T u = __shfl_up(v, o, 32);
        if (s >= o) v = u + v;
    }
    return v;
}

// Pass 1: aggregate of every tile (the Kogge-Stone total of its 32 segment sums).
template <typename T, bool NB>
__global__ void __launch_bounds__(256) tile_sum_kernel(const void* __restrict__ in, TileGeom g, T* __restrict__ agg) {
    const int l = lane_id(), s = l & 31;
    const uint64_t tile = ((uint64_t)blockIdx.x * 4 + wave_in_wg()) * 2 + (l >> 5);
    uint64_t base = 0;
    uint32_t len = 0;
    const bool live = tile < g.ntiles && tile_origin(g, tile, base, len);
    T seg = (T)0;
    if (live) {
        const uint32_t j0 = 16 * s;
        seg = j0 < len ? load_elem<T, NB>(in, base + j0) : (T)0;
#pragma unroll
        for (uint32_t k = 1; k < 16; k++) seg = seg + (j0 + k < len ? load_elem<T, NB>(in, base + j0 + k) : (T)0);
    }
    const T inc = half_wave_incl(seg, s);
    if (live && s == 31) agg[tile] = inc;
}

// Pass 2: per chunk, the running sum of its tiles' aggregates, folded left
// to right from 0 (the reference's `prefix += agg`); tprefix[tile] = the sum
// of the chunk's earlier tiles.
template <typename T>
__global__ void __launch_bounds__(256) tile_fold_kernel(const T* __restrict__ agg, TileGeom g,
                                                        T* __restrict__ tprefix, uint64_t nchunks) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    T p = (T)0;
    for (uint64_t k = 0; k < g.tpc; k++) {
        const uint64_t t = c * g.tpc + k;
        uint64_t b;
        uint32_t len;
        if (!tile_origin(g, t, b, len)) break;
        tprefix[t] = p;
        p = p + agg[t];
    }
}

// Pass 3: every element = (its tile's inclusive scan, in the raking order) + tile prefix.
template <typename T, bool NB>
__global__ void __launch_bounds__(256) tile_apply_kernel(const void* __restrict__ in, TileGeom g,
                                                         const T* __restrict__ tprefix, T* __restrict__ out) {
    const int l = lane_id(), s = l & 31;
    const uint64_t tile = ((uint64_t)blockIdx.x * 4 + wave_in_wg()) * 2 + (l >> 5);
    uint64_t base = 0;
    uint32_t len = 0;
    const bool live = tile < g.ntiles && tile_origin(g, tile, base, len);
    T x[16];
    const uint32_t j0 = 16 * s;
#pragma unroll
    for (uint32_t k = 0; k < 16; k++) x[k] = (live && j0 + k < len) ? load_elem<T, NB>(in, base + j0 + k) : (T)0;
    T seg = x[0];
#pragma unroll
    for (int k = 1; k < 16; k++) seg = seg + x[k];
    const T inc = half_wave_incl(seg, s);
    T excl = __shfl_up(inc, 1, 32);
    const T pre = live ? tprefix[tile] : (T)0;
    // inclusive downsweep of the segment onto its exclusive prefix (none for segment 0)
    T run = s ? excl + x[0] : x[0];
#pragma unroll
    for (uint32_t k = 0; k < 16; k++) {
        if (k) run = run + x[k];
        if (live && j0 + k < len) out[base + j0 + k] = run + pre;
    }
}

template <typename T, bool NB>
int launch_prefix(hipStream_t s, const void* in, void* out, size_t n, size_t chunk, void* ws, size_t ws_bytes) {
    TileGeom g;
    g.n = n;
    g.chunk = chunk ? chunk : n;
    g.tpc = cdiv(g.chunk, kTile);
    const uint64_t nchunks = cdiv(n, g.chunk);
    g.ntiles = nchunks * g.tpc;
    const size_t need = 2 * align_up(sizeof(T) * g.ntiles, 256);
    if (ws_bytes < need) return fail(FZH_E_CAPACITY, "difference inverse: workspace too small");
    T* agg = (T*)ws;
    T* tpre = (T*)((uint8_t*)ws + align_up(sizeof(T) * g.ntiles, 256));
    const unsigned grid = (unsigned)cdiv(g.ntiles, 8);  // 8 tiles per 256-thread workgroup
    launch_k("tile_sum_kernel", tile_sum_kernel<T, NB>, dim3(grid), dim3(256), 0, s, in, g, agg);
    launch_k("tile_fold_kernel", tile_fold_kernel<T>, dim3((unsigned)cdiv(nchunks, 256)), dim3(256), 0, s,
             (const T*)agg, g, tpre, nchunks);
    launch_k("tile_apply_kernel", tile_apply_kernel<T, NB>, dim3(grid), dim3(256), 0, s, in, g, (const T*)tpre,
             (T*)out);
    FZH_CHECK_LAUNCH("difference inverse");
    return FZH_OK;
}

size_t prefix_ws_bytes(size_t n, size_t eb, size_t chunk) {
    const size_t c = chunk ? chunk : std::max<size_t>(n, 1);
    const size_t nt = cdiv(n, c) * cdiv(c, kTile);
    return std::max<size_t>(2 * align_up(eb * nt, 256), 256);
}

// ---------------------------------------------------------------------------
// Integer Lorenzo (lorenzo_stage.cu).  1D: out[i] = in[i] - in[i-1] inside
// 256-element blocks (the predecessor of a block's first element is 0),
// inverse an inclusive scan per block.  2D/3D: the inclusion-exclusion
// delta over the whole field with zeros outside it, inverse inclusive
// prefix sums along x, then y, then z.  Wrapping integer arithmetic (done
// in 64 bits, stored truncated: equal modulo 2^W).
// ---------------------------------------------------------------------------
template <int EB>
__device__ __forceinline__ uint64_t ld_w(const void* p, uint64_t i) {
    return (uint64_t)(int64_t)reinterpret_cast<const typename SIntOf<EB>::type*>(p)[i];
}
template <int EB>
__device__ __forceinline__ void st_w(void* p, uint64_t i, uint64_t v) {
    reinterpret_cast<typename UIntOf<EB>::type*>(p)[i] = (typename UIntOf<EB>::type)v;
}

template <int EB>
__global__ void __launch_bounds__(256) ilz_fwd_kernel(const void* __restrict__ in, void* __restrict__ out,
                                                      uint64_t nx, uint64_t ny, uint64_t nz, int nd) {
    const uint64_t n = nx * ny * nz;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t v;
        if (nd == 1) {
            v = ld_w<EB>(in, i) - ((i % 256) ? ld_w<EB>(in, i - 1) : 0ull);
        } else {
            const uint64_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
            auto at = [&](uint64_t dx, uint64_t dy, uint64_t dz) -> uint64_t {
                if (x < dx || y < dy || z < dz) return 0ull;
                return ld_w<EB>(in, i - dx - dy * nx - dz * nx * ny);
            };
            if (nd == 2) {
                v = at(0, 0, 0) - at(1, 0, 0) - at(0, 1, 0) + at(1, 1, 0);
            } else {
                v = at(0, 0, 0) - at(1, 0, 0) - at(0, 1, 0) - at(0, 0, 1) + at(1, 1, 0) + at(1, 0, 1) +
                    at(0, 1, 1) - at(1, 1, 1);
            }
        }
        st_w<EB>(out, i, v);
    }
}

// Inclusive scan along one axis: `lines` lines of `len` elements, element k
// of line q at q_base(q) + k*step.  Thread per line and block of 64 elements:
// pass 1 block sums, pass 2 per line the running sum over blocks, pass 3 the
// block's elements onto the block's prefix.
struct AxisGeom {
    uint64_t lines, len, step;  // elements
    uint64_t inner;             // lines are q = outer*inner + r: base = outer*outer_stride + r*inner_stride
    uint64_t outer_stride, inner_stride;
    uint64_t nblk;              // blocks of 64 per line
    uint64_t limit;             // elements at index >= limit do not exist (the short last 1D block)
};

__device__ __forceinline__ uint64_t line_base(const AxisGeom& a, uint64_t q) {
    return (q / a.inner) * a.outer_stride + (q % a.inner) * a.inner_stride;
}

template <int EB>
__global__ void __launch_bounds__(256) axis_blk_kernel(const void* __restrict__ in, AxisGeom a,
                                                       uint64_t* __restrict__ bsum) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.lines * a.nblk) return;
    const uint64_t q = t % a.lines, b = t / a.lines;  // neighbouring threads: neighbouring lines
    const uint64_t base = line_base(a, q);
    uint64_t sum = 0;
    for (uint64_t k = 64 * b; k < min(a.len, 64 * b + 64); k++) {
        const uint64_t i = base + k * a.step;
        if (i >= a.limit) break;
        sum += ld_w<EB>(in, i);
    }
    bsum[t] = sum;
}

__global__ void __launch_bounds__(256) axis_line_kernel(uint64_t* __restrict__ bsum, AxisGeom a) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.lines) return;
    uint64_t run = 0;
    for (uint64_t b = 0; b < a.nblk; b++) {
        const uint64_t v = bsum[b * a.lines + q];
        bsum[b * a.lines + q] = run;
        run += v;
    }
}

template <int EB>
__global__ void __launch_bounds__(256) axis_apply_kernel(const void* __restrict__ in, void* __restrict__ out,
                                                         AxisGeom a, const uint64_t* __restrict__ bsum) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.lines * a.nblk) return;
    const uint64_t q = t % a.lines, b = t / a.lines;
    const uint64_t base = line_base(a, q);
    uint64_t run = bsum[t];
    for (uint64_t k = 64 * b; k < min(a.len, 64 * b + 64); k++) {
        const uint64_t i = base + k * a.step;
        if (i >= a.limit) break;
        run += ld_w<EB>(in, i);
        st_w<EB>(out, i, run);
    }
}

template <int EB>
int axis_scan(hipStream_t s, const void* in, void* out, const AxisGeom& a, void* ws) {
    const uint64_t nt = a.lines * a.nblk;
    const unsigned grid = (unsigned)cdiv(nt, 256);
    launch_k("axis_blk_kernel", axis_blk_kernel<EB>, dim3(grid), dim3(256), 0, s, in, a, (uint64_t*)ws);
    launch_k("axis_line_kernel", axis_line_kernel, dim3((unsigned)cdiv(a.lines, 256)), dim3(256), 0, s,
             (uint64_t*)ws, a);
    launch_k("axis_apply_kernel", axis_apply_kernel<EB>, dim3(grid), dim3(256), 0, s, in, out, a,
             (const uint64_t*)ws);
    FZH_CHECK_LAUNCH("integer lorenzo inverse");
    return FZH_OK;
}

AxisGeom axis_geom(uint64_t nx, uint64_t ny, uint64_t nz, int axis, int nd) {
    AxisGeom a{};
    if (nd == 1) {  // 1D: lines are the 256-element blocks
        a.len = 256;
        a.lines = cdiv(nx, 256);
        a.step = 1;
        a.inner = 1;
        a.outer_stride = 256;
        a.inner_stride = 0;
    } else if (axis == 0) {  // rows
        a.len = nx; a.lines = ny * nz; a.step = 1;
        a.inner = 1; a.outer_stride = nx; a.inner_stride = 0;
    } else if (axis == 1) {  // columns of each plane
        a.len = ny; a.lines = nx * nz; a.step = nx;
        a.inner = nx; a.outer_stride = nx * ny; a.inner_stride = 1;
    } else {  // z lines
        a.len = nz; a.lines = nx * ny; a.step = nx * ny;
        a.inner = 1; a.outer_stride = 1; a.inner_stride = 0;
    }
    a.nblk = cdiv(a.len, 64);
    a.limit = nx * ny * nz;
    return a;
}

template <int EB>
int launch_ilz_inv(hipStream_t s, const void* in, void* out, uint64_t nx, uint64_t ny, uint64_t nz, int nd,
                   void* ws) {
    if (nd == 1) return axis_scan<EB>(s, in, out, axis_geom(nx, 1, 1, 0, 1), ws);
    if (int rc = axis_scan<EB>(s, in, out, axis_geom(nx, ny, nz, 0, nd), ws)) return rc;
    if (int rc = axis_scan<EB>(s, out, out, axis_geom(nx, ny, nz, 1, nd), ws)) return rc;
    if (nd == 3)
        if (int rc = axis_scan<EB>(s, out, out, axis_geom(nx, ny, nz, 2, nd), ws)) return rc;
    return FZH_OK;
}

size_t ilz_ws_bytes(uint64_t nx, uint64_t ny, uint64_t nz) {
    const int nd = nz > 1 ? 3 : ny > 1 ? 2 : 1;
    uint64_t m = 0;
    for (int ax = 0; ax < (nd == 1 ? 1 : nd); ax++) {
        const AxisGeom a = axis_geom(nx, ny, nz, ax, nd);
        m = std::max<uint64_t>(m, a.lines * a.nblk);
    }
    return std::max<size_t>(align_up(8 * m, 256), 256);
}

}  // namespace
}  // namespace fzh

using namespace fzh;

extern "C" {

int fzh_zigzag(void* stream, const void* d_in, void* d_out, size_t n, int elem_bytes, int inverse) {
    return inverse ? launch_code<1>(S(stream), d_in, d_out, n, elem_bytes)
                   : launch_code<0>(S(stream), d_in, d_out, n, elem_bytes);
}

int fzh_negabinary(void* stream, const void* d_in, void* d_out, size_t n, int elem_bytes, int inverse) {
    return inverse ? launch_code<3>(S(stream), d_in, d_out, n, elem_bytes)
                   : launch_code<2>(S(stream), d_in, d_out, n, elem_bytes);
}

static int diff_check(const void* in, const void* out, size_t n, int dtype, size_t chunk, int nb, int& eb) {
    if (n && (!in || !out)) return fail(FZH_E_INVALID, "difference: null pointer");
    if (((uintptr_t)in | (uintptr_t)out) % 16 != 0) return fail(FZH_E_INVALID, "difference: buffers must be 16-byte aligned");
    if (chunk && (chunk & (chunk - 1)) != 0) return fail(FZH_E_INVALID, "difference: chunk_elems must be a power of 2");
    switch (dtype) {
    case 0: case 4: eb = 1; break;
    case 1: case 5: eb = 2; break;
    case 2: case 6: case 8: eb = 4; break;
    case 3: case 7: case 9: eb = 8; break;
    default: return fail(FZH_E_INVALID, "difference: unsupported data type");
    }
    if (nb && (dtype < 4 || dtype > 7)) return fail(FZH_E_INVALID, "difference: negabinary needs a signed integer type");
    return FZH_OK;
}

int fzh_diff_fwd(void* stream, const void* d_in, void* d_out, size_t n, int dtype, size_t chunk_elems,
                 int negabinary) {
    int eb = 0;
    if (int rc = diff_check(d_in, d_out, n, dtype, chunk_elems, negabinary, eb)) return rc;
    if (n == 0) return FZH_OK;
    hipStream_t s = S(stream);
    const uint64_t mask = chunk_elems ? (uint64_t)chunk_elems - 1 : ~0ull;
    const unsigned grid = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(n * eb, 16 * 256), 8192));
#define FZH_DIFF(T, NB) launch_k("diff_fwd_kernel", diff_fwd_kernel<T, NB>, dim3(grid), dim3(256), 0, s, (const T*)d_in, d_out, (uint64_t)n, mask)
    switch (dtype) {
    case 0: FZH_DIFF(uint8_t, false); break;
    case 1: FZH_DIFF(uint16_t, false); break;
    case 2: FZH_DIFF(uint32_t, false); break;
    case 3: FZH_DIFF(uint64_t, false); break;
    case 4: if (negabinary) FZH_DIFF(int8_t, true); else FZH_DIFF(int8_t, false); break;
    case 5: if (negabinary) FZH_DIFF(int16_t, true); else FZH_DIFF(int16_t, false); break;
    case 6: if (negabinary) FZH_DIFF(int32_t, true); else FZH_DIFF(int32_t, false); break;
    case 7: if (negabinary) FZH_DIFF(int64_t, true); else FZH_DIFF(int64_t, false); break;
    case 8: FZH_DIFF(float, false); break;
    default: FZH_DIFF(double, false); break;
    }
#undef FZH_DIFF
    FZH_CHECK_LAUNCH("diff_fwd_kernel");
    return FZH_OK;
}

size_t fzh_diff_inv_ws_bytes(size_t n, int dtype, size_t chunk_elems) {
    const size_t eb = dtype == 0 || dtype == 4 ? 1 : dtype == 1 || dtype == 5 ? 2
                    : dtype == 2 || dtype == 6 || dtype == 8 ? 4 : 8;
    return prefix_ws_bytes(n, eb, chunk_elems);
}

int fzh_diff_inv(void* stream, const void* d_in, void* d_out, size_t n, int dtype, size_t chunk_elems,
                 int negabinary, void* d_ws, size_t ws_bytes) {
    int eb = 0;
    if (int rc = diff_check(d_in, d_out, n, dtype, chunk_elems, negabinary, eb)) return rc;
    if (n == 0) return FZH_OK;
    if (d_in == d_out) return fail(FZH_E_INVALID, "difference inverse: input and output must differ");
    hipStream_t s = S(stream);
    switch (dtype) {
    case 0: return launch_prefix<uint8_t, false>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes);
    case 1: return launch_prefix<uint16_t, false>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes);
    case 2: return launch_prefix<uint32_t, false>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes);
    case 3: return launch_prefix<uint64_t, false>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes);
    case 4: return negabinary ? launch_prefix<int8_t, true>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes)
                              : launch_prefix<int8_t, false>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes);
    case 5: return negabinary ? launch_prefix<int16_t, true>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes)
                              : launch_prefix<int16_t, false>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes);
    case 6: return negabinary ? launch_prefix<int32_t, true>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes)
                              : launch_prefix<int32_t, false>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes);
    case 7: return negabinary ? launch_prefix<int64_t, true>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes)
                              : launch_prefix<int64_t, false>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes);
    case 8: return launch_prefix<float, false>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes);
    default: return launch_prefix<double, false>(s, d_in, d_out, n, chunk_elems, d_ws, ws_bytes);
    }
}

int fzh_lorenzo_int_fwd(void* stream, const void* d_in, void* d_out, size_t nx, size_t ny, size_t nz,
                        int elem_bytes) {
    const uint64_t n = (uint64_t)nx * ny * nz;
    if (n == 0) return FZH_OK;
    if (!d_in || !d_out) return fail(FZH_E_INVALID, "integer lorenzo: null pointer");
    if (d_in == d_out) return fail(FZH_E_INVALID, "integer lorenzo: input and output must differ");
    const int nd = nz > 1 ? 3 : ny > 1 ? 2 : 1;
    const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(cdiv(n, 256), 16384));
    hipStream_t s = S(stream);
    switch (elem_bytes) {
    case 1: launch_k("ilz_fwd_kernel", ilz_fwd_kernel<1>, dim3(grid), dim3(256), 0, s, d_in, d_out, (uint64_t)nx, (uint64_t)ny, (uint64_t)nz, nd); break;
    case 2: launch_k("ilz_fwd_kernel", ilz_fwd_kernel<2>, dim3(grid), dim3(256), 0, s, d_in, d_out, (uint64_t)nx, (uint64_t)ny, (uint64_t)nz, nd); break;
    case 4: launch_k("ilz_fwd_kernel", ilz_fwd_kernel<4>, dim3(grid), dim3(256), 0, s, d_in, d_out, (uint64_t)nx, (uint64_t)ny, (uint64_t)nz, nd); break;
    case 8: launch_k("ilz_fwd_kernel", ilz_fwd_kernel<8>, dim3(grid), dim3(256), 0, s, d_in, d_out, (uint64_t)nx, (uint64_t)ny, (uint64_t)nz, nd); break;
    default: return fail(FZH_E_INVALID, "integer lorenzo: element size must be 1, 2, 4 or 8");
    }
    FZH_CHECK_LAUNCH("ilz_fwd_kernel");
    return FZH_OK;
}

size_t fzh_lorenzo_int_inv_ws_bytes(size_t nx, size_t ny, size_t nz) { return ilz_ws_bytes(nx, ny, nz); }

int fzh_lorenzo_int_inv(void* stream, const void* d_in, void* d_out, size_t nx, size_t ny, size_t nz,
                        int elem_bytes, void* d_ws, size_t ws_bytes) {
    const uint64_t n = (uint64_t)nx * ny * nz;
    if (n == 0) return FZH_OK;
    if (!d_in || !d_out || !d_ws) return fail(FZH_E_INVALID, "integer lorenzo: null pointer");
    if (ws_bytes < ilz_ws_bytes(nx, ny, nz)) return fail(FZH_E_CAPACITY, "integer lorenzo: workspace too small");
    const int nd = nz > 1 ? 3 : ny > 1 ? 2 : 1;
    hipStream_t s = S(stream);
    switch (elem_bytes) {
    case 1: return launch_ilz_inv<1>(s, d_in, d_out, nx, ny, nz, nd, d_ws);
    case 2: return launch_ilz_inv<2>(s, d_in, d_out, nx, ny, nz, nd, d_ws);
    case 4: return launch_ilz_inv<4>(s, d_in, d_out, nx, ny, nz, nd, d_ws);
    case 8: return launch_ilz_inv<8>(s, d_in, d_out, nx, ny, nz, nd, d_ws);
    default: return fail(FZH_E_INVALID, "integer lorenzo: element size must be 1, 2, 4 or 8");
    }
}

}  // extern "C"

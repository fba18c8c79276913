// rle.hip — global run-length encoding (gfx950).
//
// Wire format (modules/coders/rle/rle.h:22, rle.cu:119-148):
//   [num_runs u32][values T x R, zero-padded to 4 bytes][run lengths u32 x R]
// Runs are global (not chunked): element i starts a run iff i == 0 or
// x[i] != x[i-1] (rle.cu:45-55).
//
// MI355X design
//   encode: K1 — one wave per 8 x 64 x (16/sizeof T) elements; boundary
//     flags from 16-byte loads, run offsets from the ordered tile scan; it
//     writes run values straight into the output and run start positions into
//     scratch.  K2 — run lengths = next start - start.  (The reference does
//     5 passes plus a CUB scan through ~15 B/element of scratch.)
//   decode: K1 — ordered scan of the run lengths into run starts, plus a map
//     "output tile -> first run overlapping it" (every run writes the tile
//     boundaries it covers).  K2 — one wave per 4096-element output tile
//     stages that tile's runs in LDS; each lane binary-searches its first
//     run and fills 64 consecutive outputs.  No host round trip (the
//     reference syncs twice and fills with one thread per run).
#include "fzh_internal.h"

namespace fzh {
namespace {

constexpr int STEPS = 8;
constexpr uint32_t DTILE = 4096;  // decode output tile (elements)

template <typename T>
__global__ void __launch_bounds__(256)
rle_enc_runs_kernel(const T* __restrict__ in, uint64_t n, uint8_t* __restrict__ out, uint32_t* __restrict__ pos,
                    uint32_t* __restrict__ d_R, ScanCtx* ctx, uint64_t* status, uint32_t ntiles,
                    uint32_t total_waves) {
    constexpr int E = 16 / sizeof(T);       // elements per 16-byte load
    constexpr uint32_t TILE = 64 * E * STEPS;
    const TileTicket tk = take_ticket(ctx);
    if (tk.tile < ntiles) {
        const int l = lane_id();
        const uint64_t base = (uint64_t)tk.tile * TILE;
        T v[STEPS][E];
#pragma unroll
        for (int j = 0; j < STEPS; j++) {
            const uint64_t e0 = base + (uint64_t)j * 64 * E + (uint64_t)l * E;
            if (e0 + E <= n) {
                const uint4 w = *reinterpret_cast<const uint4*>(in + e0);
                const T* p = reinterpret_cast<const T*>(&w);
#pragma unroll
                for (int k = 0; k < E; k++) v[j][k] = p[k];
            } else {
#pragma unroll
                for (int k = 0; k < E; k++) v[j][k] = (e0 + k < n) ? in[e0 + k] : (T)0;
            }
        }
        const T before = (base > 0) ? in[base - 1] : (T)0;
        uint32_t mask[STEPS];
        uint32_t cnt = 0;
#pragma unroll
        for (int j = 0; j < STEPS; j++) {
            T prev = __shfl_up(v[j][E - 1], 1, 64);
            const T prevj = __shfl(v[j > 0 ? j - 1 : 0][E - 1], 63, 64);
            if (l == 0) prev = (j == 0) ? before : prevj;
            uint32_t m = 0;
            const uint64_t e0 = base + (uint64_t)j * 64 * E + (uint64_t)l * E;
#pragma unroll
            for (int k = 0; k < E; k++) {
                const T p = (k == 0) ? prev : v[j][k - 1];
                const uint64_t e = e0 + k;
                const bool b = e < n && (e == 0 || v[j][k] != p);
                m |= (b ? 1u : 0u) << k;
            }
            mask[j] = m;
            cnt += __popc(m);
        }
        const uint32_t agg = wave_sum(cnt);
        const uint32_t roff = lookback_exclusive(status, tk.tile, tk.epoch, agg);
        if (roff != kLookbackFailed) {
            T* vals = reinterpret_cast<T*>(out + 4);
            uint32_t run = roff;
#pragma unroll
            for (int j = 0; j < STEPS; j++) {
                const uint32_t c = __popc(mask[j]);
                const uint32_t incl = wave_incl_scan(c);
                uint32_t r = run + incl - c;
                const uint64_t e0 = base + (uint64_t)j * 64 * E + (uint64_t)l * E;
#pragma unroll
                for (int k = 0; k < E; k++)
                    if (mask[j] & (1u << k)) {
                        vals[r] = v[j][k];
                        pos[r] = (uint32_t)(e0 + k);
                        r++;
                    }
                run += __shfl(incl, 63, 64);
            }
        }
        if (tk.tile == ntiles - 1 && l == 0) *d_R = (roff == kLookbackFailed) ? 0xFFFFFFFFu : roff + agg;
    }
    release_ticket(ctx, tk.epoch, total_waves);
}

template <typename T>
__global__ void __launch_bounds__(256)
rle_enc_lengths_kernel(uint8_t* __restrict__ out, const uint32_t* __restrict__ pos, const uint32_t* __restrict__ d_R,
                       uint64_t n, uint64_t* __restrict__ out_size) {
    const uint32_t R = *d_R;
    if (R == 0xFFFFFFFFu) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *out_size = 0;
        return;
    }
    const uint64_t vb = (uint64_t)R * sizeof(T);
    const uint64_t va = (vb + 3) & ~3ull;
    uint32_t* lens = reinterpret_cast<uint32_t*>(out + 4 + va);
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = pos[r];
        const uint32_t e = (r + 1 < R) ? pos[r + 1] : (uint32_t)n;
        lens[r] = e - s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *reinterpret_cast<uint32_t*>(out) = R;
        for (uint64_t b = vb; b < va; b++) out[4 + b] = 0;
        *out_size = 4 + va + 4ull * R;
    }
}

__global__ void rle_enc_empty_kernel(uint8_t* out, uint64_t* out_size) {
    if (threadIdx.x == 0) {
        *reinterpret_cast<uint32_t*>(out) = 0u;
        *out_size = 4;
    }
}

// ---------------------------------------------------------------------------
// Decode
// ---------------------------------------------------------------------------
struct RleDecState {
    uint32_t R;
    uint32_t total;
};

// Scan of run lengths -> start[r] (exclusive), start[R] = total.  One wave per
// 1024 runs; waves past R only take/release tickets.  Each run also records
// itself as the first run of every output tile whose first element it covers.
template <typename T>
__global__ void __launch_bounds__(256)
rle_dec_scan_kernel(const uint8_t* __restrict__ in, uint64_t in_bytes, uint64_t n, uint32_t* __restrict__ start,
                    uint32_t* __restrict__ tile_first, int* __restrict__ d_status, ScanCtx* ctx, uint64_t* status,
                    uint32_t nseg_max, uint32_t total_waves) {
    const TileTicket tk = take_ticket(ctx);
    const int l = lane_id();
    uint32_t R = 0;
    bool hdr_ok = in_bytes >= 4;
    if (hdr_ok) R = *reinterpret_cast<const uint32_t*>(in);
    const uint64_t va = ((uint64_t)R * sizeof(T) + 3) & ~3ull;
    if (hdr_ok) hdr_ok = R <= n && 4 + va + 4ull * R <= in_bytes;
    const uint32_t nseg = hdr_ok ? (uint32_t)((R + 1023) / 1024) : 0u;
    if (tk.tile == 0 && l == 0 && !hdr_ok) *d_status = 1;
    if (tk.tile < nseg && tk.tile < nseg_max) {
        const uint32_t* lens = reinterpret_cast<const uint32_t*>(in + 4 + va);
        const uint32_t b = tk.tile * 1024 + 16 * l;
        uint32_t v[16];
        uint64_t sum = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            v[i] = (b + i < R) ? lens[b + i] : 0u;
            sum += v[i];
        }
        const uint32_t s32 = sum > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)sum;
        const uint32_t incl = wave_incl_scan(s32);
        const uint32_t agg = __shfl(incl, 63, 64);
        uint32_t pre = lookback_exclusive(status, tk.tile, tk.epoch, agg);
        const bool ok = pre != kLookbackFailed;
        if (!ok) pre = 0;
        uint64_t run = (uint64_t)pre + incl - s32;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t r = b + i;
            if (r < R) {
                const uint64_t s = run, e = run + v[i];
                start[r] = (uint32_t)(s < n ? s : n);
                // tiles whose first element lies in [s, e)
                if (s < n && v[i] > 0) {
                    const uint64_t e2 = e < n ? e : n;
                    for (uint64_t t = (s + DTILE - 1) / DTILE; t * DTILE < e2; t++) tile_first[t] = r;
                }
            }
            run += v[i];
        }
        if (tk.tile == nseg - 1 && l == 0) {
            const uint64_t total = (uint64_t)pre + agg;
            start[R] = (uint32_t)(total < n ? total : n);
            if (!ok || total != n) *d_status = 1;
        }
    }
    if (tk.tile == 0 && l == 0 && hdr_ok && R == 0 && n != 0) *d_status = 1;
    release_ticket(ctx, tk.epoch, total_waves);
}

template <typename T>
__global__ void __launch_bounds__(128)
rle_dec_expand_kernel(const uint8_t* __restrict__ in, uint64_t n, const uint32_t* __restrict__ start,
                      const uint32_t* __restrict__ tile_first, const int* __restrict__ d_status, T* __restrict__ out,
                      uint32_t ntiles) {
    __shared__ uint32_t s_start[2][DTILE + 2];
    __shared__ T s_val[2][DTILE + 2];
    const int wv = threadIdx.x >> 6;
    const uint32_t t = blockIdx.x * 2 + wv;
    if (t >= ntiles || *d_status != 0) return;
    const int l = lane_id();
    const uint32_t R = *reinterpret_cast<const uint32_t*>(in);
    const T* vals = reinterpret_cast<const T*>(in + 4);
    const uint64_t o0 = (uint64_t)t * DTILE;
    const uint64_t o1 = (o0 + DTILE < n) ? o0 + DTILE : n;
    const uint32_t r0 = tile_first[t];
    const uint32_t r1 = (t + 1 < ntiles) ? tile_first[t + 1] + 1 : R;  // runs [r0, r1) touch the tile
    const uint32_t m = (r1 > r0) ? (r1 - r0) : 1u;
    const uint32_t mm = m < DTILE + 1 ? m : DTILE + 1;
    for (uint32_t i = l; i < mm; i += 64) {
        s_start[wv][i] = start[r0 + i];
        s_val[wv][i] = vals[r0 + i];
    }
    wave_lds_sync();
    // lane l fills [o0 + 64 l, o0 + 64 l + 64)
    const uint64_t my0 = o0 + 64ull * l;
    if (my0 < o1) {
        // last run with start <= my0
        uint32_t lo = 0, hi = mm;  // invariant: s_start[lo] <= my0 < s_start[hi] (hi may be mm)
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if ((uint64_t)s_start[wv][mid] <= my0) lo = mid; else hi = mid;
        }
        uint32_t r = lo;
        uint64_t next = (r + 1 < mm) ? s_start[wv][r + 1] : (r0 + r + 1 <= R ? (uint64_t)start[r0 + r + 1] : n);
        T cur = s_val[wv][r];
        const uint64_t my1 = (my0 + 64 < o1) ? my0 + 64 : o1;
        constexpr int E = 16 / sizeof(T);
        if (my1 - my0 == 64) {
            for (int q = 0; q < 64 / E; q++) {
                T buf[E];
#pragma unroll
                for (int k = 0; k < E; k++) {
                    const uint64_t i = my0 + (uint64_t)q * E + k;
                    while (i >= next) {
                        r++;
                        cur = s_val[wv][r < mm ? r : mm - 1];
                        next = (r + 1 < mm) ? s_start[wv][r + 1] : n;
                    }
                    buf[k] = cur;
                }
                *reinterpret_cast<uint4*>(out + my0 + (uint64_t)q * E) = *reinterpret_cast<const uint4*>(buf);
            }
        } else {
            for (uint64_t i = my0; i < my1; i++) {
                while (i >= next) {
                    r++;
                    cur = s_val[wv][r < mm ? r : mm - 1];
                    next = (r + 1 < mm) ? s_start[wv][r + 1] : n;
                }
                out[i] = cur;
            }
        }
    }
}

struct EncWs {
    size_t o_ctx, o_pos, o_R, total;
};
template <typename T>
EncWs enc_layout(size_t n) {
    constexpr size_t TILE = 64 * (16 / sizeof(T)) * STEPS;
    EncWs w;
    size_t o = 0;
    w.o_ctx = o; o = align_up(o + scan_state_bytes(cdiv(n, TILE)), 256);
    w.o_pos = o; o = align_up(o + 4 * std::max<size_t>(n, 1), 256);
    w.o_R = o;   o = align_up(o + 16, 256);
    w.total = o;
    return w;
}

template <typename T>
int enc_impl(hipStream_t s, const void* in, size_t n, void* out, size_t cap, uint64_t* out_size, void* ws,
             size_t ws_bytes) {
    const EncWs L = enc_layout<T>(n);
    if (ws_bytes < L.total) return fail(FZH_E_CAPACITY, "RLE enc: workspace too small");
    if (cap < fzh_rle_enc_bound(n, sizeof(T))) return fail(FZH_E_CAPACITY, "RLE enc: output below fzh_rle_enc_bound()");
    if (n == 0) {
        rle_enc_empty_kernel<<<1, 64, 0, s>>>((uint8_t*)out, out_size);
        FZH_CHECK_LAUNCH("rle_enc_empty_kernel");
        return FZH_OK;
    }
    if (((uintptr_t)in % 16) || ((uintptr_t)out % 4)) return fail(FZH_E_INVALID, "RLE enc: input must be 16-byte, output 4-byte aligned");
    constexpr size_t TILE = 64 * (16 / sizeof(T)) * STEPS;
    const uint32_t ntiles = (uint32_t)cdiv(n, TILE);
    uint8_t* w = (uint8_t*)ws;
    ScanCtx* ctx = (ScanCtx*)(w + L.o_ctx);
    uint64_t* status = (uint64_t*)(w + L.o_ctx + sizeof(ScanCtx));
    uint32_t* pos = (uint32_t*)(w + L.o_pos);
    uint32_t* dR = (uint32_t*)(w + L.o_R);
    const unsigned blocks = (unsigned)cdiv(ntiles, 4);
    rle_enc_runs_kernel<T><<<blocks, 256, 0, s>>>((const T*)in, n, (uint8_t*)out, pos, dR, ctx, status, ntiles, blocks * 4);
    const unsigned lb = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(n, 256), 2048));
    rle_enc_lengths_kernel<T><<<lb, 256, 0, s>>>((uint8_t*)out, pos, dR, n, out_size);
    FZH_CHECK_LAUNCH("rle encode");
    return FZH_OK;
}

struct DecWs {
    size_t nseg, ntiles, o_ctx, o_start, o_tf, total;
};
DecWs dec_layout(size_t n) {
    DecWs w;
    w.nseg = cdiv(std::max<size_t>(n, 1), 1024);
    w.ntiles = cdiv(std::max<size_t>(n, 1), DTILE);
    size_t o = 0;
    w.o_ctx = o;   o = align_up(o + scan_state_bytes(w.nseg), 256);
    w.o_start = o; o = align_up(o + 4 * (n + 2), 256);
    w.o_tf = o;    o = align_up(o + 4 * (w.ntiles + 1), 256);
    w.total = o;
    return w;
}

template <typename T>
int dec_impl(hipStream_t s, const void* in, size_t in_bytes, void* out, size_t n, int* d_status, void* ws,
             size_t ws_bytes) {
    const DecWs L = dec_layout(n);
    if (ws_bytes < L.total) return fail(FZH_E_CAPACITY, "RLE dec: workspace too small");
    if (!d_status) return fail(FZH_E_INVALID, "RLE dec: null status pointer");
    if (n == 0) return FZH_OK;
    if (((uintptr_t)in % 4) || ((uintptr_t)out % 16)) return fail(FZH_E_INVALID, "RLE dec: input must be 4-byte, output 16-byte aligned");
    if (n > 0xFFFFFFF0ull) return fail(FZH_E_INVALID, "RLE dec: more than 2^32 elements");
    uint8_t* w = (uint8_t*)ws;
    ScanCtx* ctx = (ScanCtx*)(w + L.o_ctx);
    uint64_t* status = (uint64_t*)(w + L.o_ctx + sizeof(ScanCtx));
    uint32_t* start = (uint32_t*)(w + L.o_start);
    uint32_t* tf = (uint32_t*)(w + L.o_tf);
    const unsigned sb = (unsigned)cdiv(L.nseg, 4);
    rle_dec_scan_kernel<T><<<sb, 256, 0, s>>>((const uint8_t*)in, in_bytes, n, start, tf, d_status, ctx, status,
                                              (uint32_t)L.nseg, sb * 4);
    const unsigned eb = (unsigned)cdiv(L.ntiles, 2);
    rle_dec_expand_kernel<T><<<eb, 128, 0, s>>>((const uint8_t*)in, n, start, tf, d_status, (T*)out, (uint32_t)L.ntiles);
    FZH_CHECK_LAUNCH("rle decode");
    return FZH_OK;
}

}  // namespace
}  // namespace fzh

using namespace fzh;

extern "C" {

size_t fzh_rle_enc_bound(size_t n, int elem_bytes) {
    return 4 + align_up(n * (size_t)elem_bytes, 4) + 4 * n + 4;
}

// Enough for every element width (4-byte elements have the most tiles).
size_t fzh_rle_enc_ws_bytes(size_t n) {
    return std::max(enc_layout<uint8_t>(n).total, std::max(enc_layout<uint16_t>(n).total, enc_layout<uint32_t>(n).total));
}

int fzh_rle_enc(void* stream, const void* d_in, size_t n, int elem_bytes, void* d_out, size_t out_cap,
                uint64_t* d_out_size, void* d_ws, size_t ws_bytes) {
    if (n > 0xFFFFFFF0ull) return fail(FZH_E_INVALID, "RLE enc: more than 2^32 elements");
    switch (elem_bytes) {
    case 1: return enc_impl<uint8_t>(S(stream), d_in, n, d_out, out_cap, d_out_size, d_ws, ws_bytes);
    case 2: return enc_impl<uint16_t>(S(stream), d_in, n, d_out, out_cap, d_out_size, d_ws, ws_bytes);
    case 4: return enc_impl<uint32_t>(S(stream), d_in, n, d_out, out_cap, d_out_size, d_ws, ws_bytes);
    default: return fail(FZH_E_INVALID, "RLE: elem_bytes must be 1, 2 or 4");
    }
}

size_t fzh_rle_dec_ws_bytes(size_t n) { return dec_layout(n).total; }

int fzh_rle_dec(void* stream, const void* d_in, size_t in_bytes, int elem_bytes, void* d_out, size_t n,
                int* d_status, void* d_ws, size_t ws_bytes) {
    switch (elem_bytes) {
    case 1: return dec_impl<uint8_t>(S(stream), d_in, in_bytes, d_out, n, d_status, d_ws, ws_bytes);
    case 2: return dec_impl<uint16_t>(S(stream), d_in, in_bytes, d_out, n, d_status, d_ws, ws_bytes);
    case 4: return dec_impl<uint32_t>(S(stream), d_in, in_bytes, d_out, n, d_status, d_ws, ws_bytes);
    default: return fail(FZH_E_INVALID, "RLE: elem_bytes must be 1, 2 or 4");
    }
}

}  // extern "C"

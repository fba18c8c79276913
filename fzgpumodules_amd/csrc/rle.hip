// rle.hip — global run-length encoding (gfx950).
//
// Wire format (modules/coders/rle/rle.h:22, rle.cu:119-148):
//   [num_runs u32][values T x R, zero-padded to 4 bytes][run lengths u32 x R]
// Runs are global (not chunked): element i starts a run iff i == 0 or
// x[i] != x[i-1] (rle.cu:45-55).
//
// MI355X design
//   encode: one wave per tile of 8 x 64 x (16/sizeof T) elements.  K1
//     counts the tile's run starts (boundary flags from 16-byte loads); a
//     scan of the counts gives each tile's first run index and R; K2 finds
//     the starts again and writes every run's value and every length whose
//     run ends inside the tile straight to the output (the previous start
//     from a running maximum across lanes), leaving the start of the tile's
//     last run in its length slot; K3, one thread per tile, turns those into
//     lengths with the next tile's first start and writes the header.  The
//     input is read twice and the output written once, no position scratch.
//     (The reference does 5 passes plus a CUB scan through ~15 B/element of
//     scratch.)
//   decode: a three-launch segmented scan of the run lengths gives the run
//     starts and a map "output tile -> first run overlapping it" (every run
//     marks the tile boundaries it covers; the lengths and starts pass
//     through LDS so global accesses stay coalesced); then one wave per
//     1024-element output tile stages that tile's runs in LDS (33 KiB per
//     four-wave workgroup), each lane binary-searches its first run and
//     fills 16 consecutive outputs.  No host round trip
//     (the reference syncs twice and fills with one thread per run).
#include "fzh_scan.h"

namespace fzh {
namespace {

constexpr int STEPS = 8;
constexpr uint32_t DTILE = 1024;  // decode output tile (elements), one wave each

// Run-start flags of one wave's tile: mask[j] bit k <=> element
// base + 64 E j + E l + k starts a run (k < E, E elements per 16-byte load).
template <typename T>
__device__ __forceinline__ void rle_tile_starts(const T* __restrict__ in, uint64_t n, uint64_t base,
                                                T (&v)[STEPS][16 / sizeof(T)], uint32_t (&mask)[STEPS]) {
    constexpr int E = 16 / sizeof(T);
    const int l = lane_id();
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
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
rle_enc_count_kernel(const T* __restrict__ in, uint64_t n, uint32_t* __restrict__ cnt, uint32_t ntiles) {
    constexpr int E = 16 / sizeof(T);
    constexpr uint32_t TILE = 64 * E * STEPS;
    const uint32_t tile = blockIdx.x * 4 + wave_in_wg();
    if (tile >= ntiles) return;
    T v[STEPS][E];
    uint32_t mask[STEPS];
    rle_tile_starts<T>(in, n, (uint64_t)tile * TILE, v, mask);
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < STEPS; j++) c += __popc(mask[j]);
    c = wave_sum(c);
    if (lane_id() == 0) cnt[tile] = c;
}

// Layout of the stream for R runs: values at 4, lengths at 4 + va.
__device__ __forceinline__ uint64_t rle_values_bytes(uint32_t R, int eb) { return ((uint64_t)R * eb + 3) & ~3ull; }

// pre[tile] = index of the tile's first run (exclusive scan of the counts),
// *d_R = the run count.  Writes values[r] for every run of the tile and
// lengths[r] for every run that ends inside it; the tile's last run gets its
// start position in its length slot and first[tile] the tile's first start,
// for rle_enc_fix_kernel.
template <typename T>
__global__ void __launch_bounds__(256)
rle_enc_write_kernel(const T* __restrict__ in, uint64_t n, const uint32_t* __restrict__ pre,
                     const uint32_t* __restrict__ d_R, uint32_t* __restrict__ first, uint8_t* __restrict__ out,
                     uint32_t ntiles) {
    constexpr int E = 16 / sizeof(T);
    constexpr uint32_t TILE = 64 * E * STEPS;
    constexpr uint32_t SR = 64 * E;  // most run starts of one step
    // a step's runs staged in LDS, then stored contiguously (coalesced):
    // value of run r + i, length of run r - 1 + i
    __shared__ T s_v[4][SR];
    __shared__ uint32_t s_l[4][SR];
    const uint32_t tile = blockIdx.x * 4 + wave_in_wg();
    if (tile >= ntiles) return;
    const int l = lane_id();
    const int wv = (int)wave_in_wg();
    const uint64_t base = (uint64_t)tile * TILE;
    T v[STEPS][E];
    uint32_t mask[STEPS];
    rle_tile_starts<T>(in, n, base, v, mask);
    const uint32_t R = *d_R;
    T* vals = reinterpret_cast<T*>(out + 4);
    uint32_t* lens = reinterpret_cast<uint32_t*>(out + 4 + rle_values_bytes(R, sizeof(T)));
    uint32_t r = pre[tile];     // wave-uniform: index of the next run start in tile order
    uint32_t last1 = 0;         // wave-uniform: (last start so far in this tile) + 1, 0 if none
    uint32_t first1 = 0;        // wave-uniform: (first start of the tile) + 1
#pragma unroll
    for (int j = 0; j < STEPS; j++) {
        const uint32_t m = mask[j];
        const uint32_t c = __popc(m);
        const uint32_t incl = wave_incl_scan(c);
        const uint32_t e0 = (uint32_t)(base + (uint64_t)j * 64 * E + (uint64_t)l * E);
        const uint32_t lp1 = m ? e0 + (31u - (uint32_t)__clz(m)) + 1u : 0u;  // last start in this lane + 1
        const uint32_t pmax = wave_incl_max(lp1);
        uint32_t prev1 = max(wave_shr1(pmax), last1);  // previous start in tile order + 1
        uint32_t q = incl - c;  // this lane's first run within the step
#pragma unroll
        for (int k = 0; k < E; k++) {
            if (m & (1u << k)) {
                const uint32_t e = e0 + (uint32_t)k;
                s_v[wv][q] = v[j][k];
                s_l[wv][q] = prev1 ? e - (prev1 - 1u) : 0u;
                prev1 = e + 1u;
                q++;
            }
        }
        const uint32_t ctot = lane63(incl);
        wave_lds_sync();
        // the step's first run closes the run before it only when that run
        // started in this tile (else the fix kernel closes it)
        for (uint32_t i = (uint32_t)l; i < ctot; i += 64) {
            vals[r + i] = s_v[wv][i];
            if (i > 0 || last1) lens[r - 1 + i] = s_l[wv][i];
        }
        wave_lds_sync();
        const uint64_t any = __ballot(m != 0u);
        if (!first1 && any) {
            const int fl = __builtin_ctzll(any);
            first1 = (uint32_t)__builtin_amdgcn_readlane((int)(m ? e0 + (uint32_t)__builtin_ctz(m) + 1u : 0u), fl);
        }
        last1 = max(last1, lane63(pmax));
        r += lane63(incl);
    }
    if (l == 0) {
        first[tile] = first1;
        if (last1) lens[r - 1] = last1 - 1u;
    }
}

// One thread per tile: the length of the run before the tile's first start
// (the previous tile's last run, whose slot holds its start), and of the
// last run; thread 0 writes the header and the values' zero padding.
template <typename T>
__global__ void __launch_bounds__(256)
rle_enc_fix_kernel(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ pre,
                   const uint32_t* __restrict__ first, const uint32_t* __restrict__ d_R, uint64_t n,
                   uint8_t* __restrict__ out, uint64_t* __restrict__ out_size, uint32_t ntiles) {
    const uint32_t R = *d_R;
    const uint64_t vb = (uint64_t)R * sizeof(T);
    const uint64_t va = rle_values_bytes(R, sizeof(T));
    uint32_t* lens = reinterpret_cast<uint32_t*>(out + 4 + va);
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ntiles) {
        const uint32_t c = cnt[k], p = pre[k];
        if (c) {
            if (p) lens[p - 1] = (first[k] - 1u) - lens[p - 1];
            if (p + c == R) lens[R - 1] = (uint32_t)n - lens[R - 1];
        }
    }
    if (k == 0) {
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

// Validated run count (0 and status = 1 if the header is inconsistent).
template <typename T>
__global__ void rle_dec_prep_kernel(const uint8_t* __restrict__ in, uint64_t in_bytes, uint64_t n,
                                    uint32_t* __restrict__ Rv, int* __restrict__ d_status) {
    if (threadIdx.x != 0) return;
    uint32_t R = 0;
    bool ok = in_bytes >= 4;
    if (ok) R = *reinterpret_cast<const uint32_t*>(in);
    const uint64_t va = ((uint64_t)R * sizeof(T) + 3) & ~3ull;
    if (ok) ok = R <= n && 4 + va + 4ull * R <= in_bytes && !(R == 0 && n != 0);
    *Rv = ok ? R : 0u;
    if (!ok) *d_status = 1;
}

__device__ __forceinline__ const uint32_t* rle_lengths(const uint8_t* in, uint32_t R, int eb) {
    return reinterpret_cast<const uint32_t*>(in + 4 + (((uint64_t)R * eb + 3) & ~3ull));
}

template <typename T>
__global__ void __launch_bounds__(256)
rle_dec_reduce_kernel(const uint8_t* __restrict__ in, const uint32_t* __restrict__ Rv,
                      uint32_t* __restrict__ seg_sum) {
    __shared__ uint64_t s_w[4];
    const uint32_t R = *Rv;
    const uint32_t* lens = rle_lengths(in, R, sizeof(T));
    const uint64_t b = (uint64_t)blockIdx.x * SCAN_SEG;
    uint64_t sm = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const uint64_t i = b + threadIdx.x + 256ull * k;
        if (i < R) sm += lens[i];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sm += __shfl_xor(sm, off, 64);
    if (lane_id() == 0) s_w[threadIdx.x >> 6] = sm;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t t = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        seg_sum[blockIdx.x] = t > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)t;
    }
}

// start[r] = exclusive prefix of the lengths (clamped to n), start[R] = the
// total; each run marks the output tiles whose first element it covers.
template <typename T>
__global__ void __launch_bounds__(256)
rle_dec_apply_kernel(const uint8_t* __restrict__ in, uint64_t n, const uint32_t* __restrict__ Rv,
                     const uint32_t* __restrict__ seg_base, const uint32_t* __restrict__ total,
                     uint32_t* __restrict__ start, uint32_t* __restrict__ tile_first, int* __restrict__ d_status) {
    __shared__ uint32_t s_w[4];
    // the segment's lengths, then its starts, in LDS: loaded and stored
    // coalesced, read by thread t as its 16 consecutive runs (one padding
    // word per 16 keeps those reads on distinct banks)
    __shared__ uint32_t s_x[SCAN_SEG + SCAN_SEG / 16];
    auto px = [](uint32_t i) { return i + (i >> 4); };
    const uint32_t R = *Rv;
    const uint32_t* lens = rle_lengths(in, R, sizeof(T));
    const int tid = threadIdx.x, w = tid >> 6;
    const uint64_t seg0 = (uint64_t)blockIdx.x * SCAN_SEG;
    const uint64_t b = seg0 + 16ull * tid;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const uint32_t i = (uint32_t)(256 * k + tid);
        s_x[px(i)] = (seg0 + i < R) ? lens[seg0 + i] : 0u;
    }
    __syncthreads();
    uint32_t v[16], sum = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        v[k] = s_x[px(16 * tid + k)];
        sum += v[k];
    }
    const uint32_t incl = wave_incl_scan(sum);
    if (lane_id() == 63) s_w[w] = incl;
    __syncthreads();
    uint64_t run = (uint64_t)seg_base[blockIdx.x] + incl - sum;
    for (int j = 0; j < w; j++) run += s_w[j];
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const uint64_t r = b + k;
        const uint64_t s0 = run, e = run + v[k];
        s_x[px(16 * tid + k)] = (uint32_t)(s0 < n ? s0 : n);  // after this thread's own reads
        if (r < R && s0 < n && v[k] > 0) {
            const uint64_t e2 = e < n ? e : n;
            for (uint64_t t = (s0 + DTILE - 1) / DTILE; t * DTILE < e2; t++) tile_first[t] = (uint32_t)r;
        }
        run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) {
        const uint32_t i = (uint32_t)(256 * k + tid);
        if (seg0 + i < R) start[seg0 + i] = s_x[px(i)];
    }
    if (blockIdx.x == 0 && tid == 0) {
        const uint32_t tot = *total;
        start[R] = tot < n ? tot : (uint32_t)n;
        if (R != 0 && tot != n) *d_status = 1;
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
rle_dec_expand_kernel(const uint8_t* __restrict__ in, uint64_t n, const uint32_t* __restrict__ start,
                      const uint32_t* __restrict__ tile_first, const int* __restrict__ d_status, T* __restrict__ out,
                      uint32_t ntiles) {
    __shared__ uint32_t s_start[4][DTILE + 2];
    __shared__ T s_val[4][DTILE + 2];
    const int wv = threadIdx.x >> 6;
    const uint32_t t = blockIdx.x * 4 + wv;
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
    // lane l fills [o0 + LSPAN l, o0 + LSPAN l + LSPAN)
    constexpr uint32_t LSPAN = DTILE / 64;
    const uint64_t my0 = o0 + (uint64_t)LSPAN * l;
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
        const uint64_t my1 = (my0 + LSPAN < o1) ? my0 + LSPAN : o1;
        constexpr int E = 16 / sizeof(T);
        if (my1 - my0 == LSPAN) {
            for (int q = 0; q < (int)LSPAN / E; q++) {
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
    size_t ntiles, o_cnt, o_pre, o_first, o_R, o_scan, total;
};
template <typename T>
EncWs enc_layout(size_t n) {
    constexpr size_t TILE = 64 * (16 / sizeof(T)) * STEPS;
    EncWs w;
    w.ntiles = std::max<size_t>(cdiv(n, TILE), 1);
    size_t o = 0;
    w.o_cnt = o;   o = align_up(o + 4 * w.ntiles, 256);
    w.o_pre = o;   o = align_up(o + 4 * w.ntiles, 256);
    w.o_first = o; o = align_up(o + 4 * w.ntiles, 256);
    w.o_R = o;     o = align_up(o + 16, 256);
    w.o_scan = o;  o = align_up(o + scan_ws_bytes(w.ntiles), 256);
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
        launch_k("rle_enc_empty_kernel", rle_enc_empty_kernel, dim3(1), dim3(64), 0, s, (uint8_t*)out, out_size);
        FZH_CHECK_LAUNCH("rle_enc_empty_kernel");
        return FZH_OK;
    }
    if (((uintptr_t)in % 16) || ((uintptr_t)out % 4)) return fail(FZH_E_INVALID, "RLE enc: input must be 16-byte, output 4-byte aligned");
    constexpr size_t TILE = 64 * (16 / sizeof(T)) * STEPS;
    const uint32_t ntiles = (uint32_t)cdiv(n, TILE);
    uint8_t* w = (uint8_t*)ws;
    uint32_t* cnt = (uint32_t*)(w + L.o_cnt);
    uint32_t* pre = (uint32_t*)(w + L.o_pre);
    uint32_t* first = (uint32_t*)(w + L.o_first);
    uint32_t* dR = (uint32_t*)(w + L.o_R);
    const unsigned blocks = (unsigned)cdiv(ntiles, 4);
    launch_k("rle_enc_count_kernel", rle_enc_count_kernel<T>, dim3(blocks), dim3(256), 0, s, (const T*)in, n, cnt, ntiles);
    if (int rc = scan_u32(s, cnt, pre, ntiles, nullptr, dR, w + L.o_scan)) return rc;
    launch_k("rle_enc_write_kernel", rle_enc_write_kernel<T>, dim3(blocks), dim3(256), 0, s, (const T*)in, n,
             (const uint32_t*)pre, (const uint32_t*)dR, first, (uint8_t*)out, ntiles);
    launch_k("rle_enc_fix_kernel", rle_enc_fix_kernel<T>, dim3((unsigned)cdiv(ntiles, 256)), dim3(256), 0, s,
             (const uint32_t*)cnt, (const uint32_t*)pre, (const uint32_t*)first, (const uint32_t*)dR, (uint64_t)n,
             (uint8_t*)out, out_size, ntiles);
    FZH_CHECK_LAUNCH("rle encode");
    return FZH_OK;
}

struct DecWs {
    size_t nseg, ntiles, o_R, o_seg, o_start, o_tf, total;
};
DecWs dec_layout(size_t n) {
    DecWs w;
    w.nseg = std::max<size_t>(cdiv(std::max<size_t>(n, 1), SCAN_SEG), 1);
    w.ntiles = cdiv(std::max<size_t>(n, 1), DTILE);
    size_t o = 0;
    w.o_R = o;     o = align_up(o + 16, 256);
    w.o_seg = o;   o = align_up(o + 4 * w.nseg, 256);
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
    if (L.nseg > SCAN_MAX_SEGS) return fail(FZH_E_INVALID, "RLE dec: too many elements");
    uint8_t* w = (uint8_t*)ws;
    uint32_t* Rv = (uint32_t*)(w + L.o_R);
    uint32_t* tot = Rv + 1;
    uint32_t* seg = (uint32_t*)(w + L.o_seg);
    uint32_t* start = (uint32_t*)(w + L.o_start);
    uint32_t* tf = (uint32_t*)(w + L.o_tf);
    const uint8_t* bin = (const uint8_t*)in;
    launch_k("rle_dec_prep_kernel", rle_dec_prep_kernel<T>, dim3(1), dim3(64), 0, s, bin, in_bytes, n, Rv, d_status);
    launch_k("rle_dec_reduce_kernel", rle_dec_reduce_kernel<T>, dim3((unsigned)L.nseg), dim3(256), 0, s, bin, Rv, seg);
    launch_k("scan_partials_kernel", scan_partials_kernel<false>, dim3(1), dim3(1024), 0, s, seg, (uint32_t)L.nseg, tot);
    launch_k("rle_dec_apply_kernel", rle_dec_apply_kernel<T>, dim3((unsigned)L.nseg), dim3(256), 0, s, bin, n, Rv, seg, tot, start, tf, d_status);
    const unsigned eb = (unsigned)cdiv(L.ntiles, 4);
    launch_k("rle_dec_expand_kernel", rle_dec_expand_kernel<T>, dim3(eb), dim3(256), 0, s, (const uint8_t*)in, n, start, tf, d_status, (T*)out, (uint32_t)L.ntiles);
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

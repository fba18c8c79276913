// label.hip — exhaustive labeller of the chain selector (gfx950): the exact
// size of fz::Pipeline::compress's result for every chunk of a batch under
// each of the four candidate chains (fzs.h), without producing the streams.
// No counterpart in the reference.
//
// Per chunk the kernels gather exactly what each stream's size depends on:
//   LorenzoQuant     codes (kept for the coders) and the outlier count
//   RLE              the number of runs of the codes
//   Bitshuffle+RZE   the size of every 16 KiB block (same device function
//                    as the encoder, size only)
//   Huffman          the code histogram; the host builds each chunk's
//                    codebook; then the cells of every 768-symbol partition
//   Quantizer        the largest code and the outlier count (-> Bitpack nbits)
// The host then adds the container sizes (fzs_label_sizes below).
#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "fzh_lorenzo_tile.h"
#include "fzh_rze_chunk.h"
#include "fzs.h"

namespace fzh {
namespace label {
namespace {

constexpr uint32_t kSeg = 4096;      // codes per workgroup in the per-code kernels
constexpr uint32_t kBklen = FZS_HUF_BKLEN;
constexpr uint32_t kSub = 768;       // Huffman partition (hf_bk.cc:28-31)

struct Stats {
    uint32_t* k;      // Lorenzo outliers
    uint32_t* runs;   // RLE runs
    uint32_t* rze;    // sum of clean RZE block sizes
    uint32_t* ncell;  // Huffman cells
    uint32_t* qmax;   // largest Quantizer code
    uint32_t* qout;   // Quantizer outliers
};

template <int ND>
__global__ void __launch_bounds__(256)
lorenzo_kernel(const float* __restrict__ chunks, uint32_t nx, uint32_t ny, uint32_t nz, uint64_t T,
               uint32_t tpc, const float* __restrict__ d_eb, uint16_t* __restrict__ codes, Stats st) {
    const uint64_t g = (uint64_t)blockIdx.x * 4 + wave_in_wg();
    if (g >= T) return;
    const uint64_t E = (uint64_t)nx * ny * nz;
    const uint32_t c = (uint32_t)(g / tpc), lt = (uint32_t)(g % tpc);
    const float eb = d_eb[c];
    const float r = 1.0f / (2.0f * eb);  // as LorenzoQuantStage computes it for ABS
    const float rad = (float)(uint16_t)FZS_LZ_RADIUS;
    const float* in = chunks + c * E;
    uint16_t* cd = codes + c * E;
    const ListOut none{nullptr, nullptr, 0};
    TileCounts tc;
    if constexpr (ND == 1) {
        tc = lz1d_tile<uint16_t, true, false>(in, E, E, lt, r, rad, cd, none, none, StrictEb{0, 0});
    } else if constexpr (ND == 2) {
        tc = lz2d_tile<uint16_t, true, false>(in, nx, ny, E, lt, (nx + 31) / 32, lt + 1 == tpc, r, rad, cd, none,
                                              none, StrictEb{0, 0});
    } else {
        tc = lz3d_tile<uint16_t, true, false>(in, nx, ny, nz, E, lt, Tiles{(nx + 7) / 8, (ny + 7) / 8, 0},
                                              lt + 1 == tpc, r, rad, cd, none, none, StrictEb{0, 0});
    }
    if (lane_id() == 0 && tc.cnt) atomicAdd(&st.k[c], tc.cnt);
}

// Runs and histogram of kSeg codes (segments never straddle chunks).  Each
// thread walks 16 consecutive codes and adds runs of equal codes to the LDS
// histogram with one atomic per run (smooth data has long runs of one code).
__global__ void __launch_bounds__(256)
codes_kernel(const uint16_t* __restrict__ codes, uint64_t E, uint32_t* __restrict__ hist, Stats st) {
    __shared__ uint32_t s_h[kBklen];
    __shared__ uint32_t s_r[4];
    for (uint32_t i = threadIdx.x; i < kBklen; i += 256) s_h[i] = 0;
    __syncthreads();
    const uint64_t b = (uint64_t)blockIdx.x * kSeg;
    const uint32_t c = (uint32_t)(b / E);
    const uint64_t c0 = (uint64_t)c * E;
    const uint64_t i0 = b + 16ull * threadIdx.x;
    uint32_t runs = 0;
    uint16_t prev = i0 == c0 ? (uint16_t)0 : codes[i0 - 1];
    uint16_t cur = 0;
    uint32_t len = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const uint16_t v = codes[i0 + j];
        runs += (i0 + j == c0 || v != prev) ? 1u : 0u;
        prev = v;
        if (len && v == cur) {
            len++;
        } else {
            if (len && cur < kBklen) atomicAdd(&s_h[cur], len);
            cur = v;
            len = 1;
        }
    }
    if (cur < kBklen) atomicAdd(&s_h[cur], len);
    runs = wave_sum(runs);
    if (lane_id() == 0) s_r[threadIdx.x >> 6] = runs;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&st.runs[c], s_r[0] + s_r[1] + s_r[2] + s_r[3]);
    uint32_t* h = hist + (uint64_t)c * kBklen;
    for (uint32_t i = threadIdx.x; i < kBklen; i += 256)
        if (s_h[i]) atomicAdd(&h[i], s_h[i]);
}

// Bitshuffle + RZE size of every 16 KiB block of codes (one workgroup per
// block, the encoder's own device function, size only).
__global__ void __launch_bounds__(256)
rze_kernel(const uint16_t* __restrict__ codes, uint64_t nblocks, uint32_t bpc, Stats st) {
    __shared__ RzeWgLds L;
    const uint64_t b = blockIdx.x;
    if (b >= nblocks) return;
    const uint8_t* src = reinterpret_cast<const uint8_t*>(codes) + b * (uint64_t)CS;
    const uint32_t size = rze_encode_chunk_wg<2>(src, CS, true, L, nullptr);
    if (threadIdx.x == 0) atomicAdd(&st.rze[b / bpc], size & 0x7FFFFFFFu);
}

__device__ __forceinline__ int f2i_rn(float f) {
    if (f != f) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int)0x80000000;
    return (int)__builtin_rintf(f);
}

// QuantizerStage (ABS, zigzag, radius FZS_Q_RADIUS, uint16 codes): largest
// code and outlier count of kSeg values.
__global__ void __launch_bounds__(256)
quant_kernel(const float* __restrict__ chunks, uint64_t E, const float* __restrict__ d_eb, Stats st) {
    __shared__ uint32_t s_m[4], s_o[4];
    const uint64_t b = (uint64_t)blockIdx.x * kSeg;
    const uint32_t c = (uint32_t)(b / E);
    const float r = 1.0f / (2.0f * d_eb[c]);
    const int rad = (int)(uint16_t)FZS_Q_RADIUS;
    uint32_t mx = 0, out = 0;
    for (uint32_t j = threadIdx.x; j < kSeg; j += 256) {
        const float v = chunks[b + j];
        const int q = f2i_rn(v * r);
        if (q > -rad && q < rad && fabsf(v) < INFINITY) {
            const int16_t s = (int16_t)q;
            const uint32_t code = (uint16_t)(((uint16_t)s << 1) ^ (uint16_t)(s >> 15));
            mx = code > mx ? code : mx;
        } else {
            out++;
        }
    }
    mx = wave_max(mx);
    out = wave_sum(out);
    if (lane_id() == 0) {
        s_m[threadIdx.x >> 6] = mx;
        s_o[threadIdx.x >> 6] = out;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = s_m[0];
        for (int i = 1; i < 4; i++) m = s_m[i] > m ? s_m[i] : m;
        atomicMax(&st.qmax[c], m);
        const uint32_t o = s_o[0] + s_o[1] + s_o[2] + s_o[3];
        if (o) atomicAdd(&st.qout[c], o);
    }
}

// Huffman cells: one wave per 768-symbol partition of a chunk.
__global__ void __launch_bounds__(256)
huff_kernel(const uint16_t* __restrict__ codes, uint64_t E, uint32_t pardeg, uint64_t nparts,
            const uint8_t* __restrict__ lens, Stats st) {
    const uint64_t g = (uint64_t)blockIdx.x * 4 + wave_in_wg();
    if (g >= nparts) return;
    const uint32_t c = (uint32_t)(g / pardeg), p = (uint32_t)(g % pardeg);
    const uint64_t s0 = (uint64_t)p * kSub;
    const uint32_t cnt = (uint32_t)min((uint64_t)kSub, E - s0);
    const uint16_t* cd = codes + (uint64_t)c * E + s0;
    const uint8_t* L = lens + (uint64_t)c * kBklen;
    uint32_t bits = 0;
    for (uint32_t j = lane_id(); j < cnt; j += 64) {
        const uint16_t v = cd[j];
        bits += v < kBklen ? L[v] : 0u;
    }
    bits = wave_sum(bits);
    if (lane_id() == 0) atomicAdd(&st.ncell[c], (bits + 31) / 32);
}

}  // namespace

// Layout of the labeller workspace.
struct LabelWs {
    size_t o_codes, o_stats, o_hist, o_lens, total;
};

LabelWs label_layout(size_t n, size_t E) {
    LabelWs w;
    size_t o = 0;
    w.o_codes = o; o = align_up(o + 2 * n * E, 256);
    w.o_stats = o; o = align_up(o + 6 * 4 * n, 256);
    w.o_hist = o;  o = align_up(o + 4 * kBklen * n, 256);
    w.o_lens = o;  o = align_up(o + kBklen * n, 256);
    w.total = o;
    return w;
}

static Stats stats_of(void* ws, const LabelWs& L, size_t n) {
    uint32_t* s = (uint32_t*)((uint8_t*)ws + L.o_stats);
    return Stats{s, s + n, s + 2 * n, s + 3 * n, s + 4 * n, s + 5 * n};
}

// Stage 1: everything up to the histograms.
int label_launch_pass1(hipStream_t s, const float* chunks, size_t n, uint32_t nx, uint32_t ny, uint32_t nz,
                       const float* d_eb, void* ws) {
    const size_t E = (size_t)nx * ny * nz;
    const LabelWs L = label_layout(n, E);
    uint8_t* w = (uint8_t*)ws;
    if (int rc = zero_bytes(s, w + L.o_stats, L.o_lens - L.o_stats)) return rc;
    const Stats st = stats_of(ws, L, n);
    uint16_t* codes = (uint16_t*)(w + L.o_codes);
    const int nd = nz > 1 ? 3 : (ny > 1 ? 2 : 1);
    uint32_t tpc;
    if (nd == 3) tpc = (uint32_t)(cdiv(nx, 8) * cdiv(ny, 8) * cdiv(nz, 8));
    else if (nd == 2) tpc = (uint32_t)(cdiv(nx, 32) * cdiv(ny, 32));
    else tpc = (uint32_t)cdiv(E, 1024);
    const uint64_t T = (uint64_t)tpc * n;
    const unsigned tb = (unsigned)cdiv(T, 4);
    if (nd == 3) lorenzo_kernel<3><<<tb, 256, 0, s>>>(chunks, nx, ny, nz, T, tpc, d_eb, codes, st);
    else if (nd == 2) lorenzo_kernel<2><<<tb, 256, 0, s>>>(chunks, nx, ny, nz, T, tpc, d_eb, codes, st);
    else lorenzo_kernel<1><<<tb, 256, 0, s>>>(chunks, nx, ny, nz, T, tpc, d_eb, codes, st);
    const uint64_t segs = (uint64_t)n * E / kSeg;
    codes_kernel<<<(unsigned)segs, 256, 0, s>>>(codes, E, (uint32_t*)(w + L.o_hist), st);
    const uint32_t bpc = (uint32_t)(2 * E / CS);
    const uint64_t nb = (uint64_t)bpc * n;
    rze_kernel<<<(unsigned)nb, 256, 0, s>>>(codes, nb, bpc, st);
    quant_kernel<<<(unsigned)segs, 256, 0, s>>>(chunks, E, d_eb, st);
    FZH_CHECK_LAUNCH("labeller pass 1");
    return FZH_OK;
}

// Stage 2: Huffman cells from the per-chunk code lengths (lens in ws).
int label_launch_pass2(hipStream_t s, size_t n, size_t E, void* ws) {
    const LabelWs L = label_layout(n, E);
    uint8_t* w = (uint8_t*)ws;
    const Stats st = stats_of(ws, L, n);
    const uint32_t pardeg = (uint32_t)cdiv(E, kSub);
    const uint64_t np = (uint64_t)pardeg * n;
    huff_kernel<<<(unsigned)cdiv(np, 4), 256, 0, s>>>((const uint16_t*)(w + L.o_codes), E, pardeg, np,
                                                     w + L.o_lens, st);
    FZH_CHECK_LAUNCH("labeller pass 2");
    return FZH_OK;
}

}  // namespace label
}  // namespace fzh

using namespace fzh;
using namespace fzh::label;

namespace {

uint64_t pad16(uint64_t v) { return (v + 15) & ~15ull; }
uint64_t pad4(uint64_t v) { return (v + 3) & ~3ull; }

int check_shape(uint32_t nx, uint32_t ny, uint32_t nz) {
    const uint64_t E = (uint64_t)nx * ny * nz;
    if (E == 0) return fail(FZH_E_INVALID, "labeller: empty chunk shape");
    if ((4 * E) % CS != 0) return fail(FZH_E_INVALID, "labeller: chunk bytes must be a multiple of 16384");
    if (E > (1ull << 24)) return fail(FZH_E_INVALID, "labeller: chunks hold at most 2^24 values");
    return FZH_OK;
}

}  // namespace

extern "C" {

size_t fzs_label_ws_bytes(size_t n_chunks, uint32_t nx, uint32_t ny, uint32_t nz) {
    return label_layout(n_chunks, (size_t)nx * ny * nz).total;
}

int fzs_label_sizes(void* stream, const float* d_chunks, size_t n, uint32_t nx, uint32_t ny, uint32_t nz,
                    const float* d_abs_eb, uint64_t* h_sizes, void* d_ws, size_t ws_bytes, int host_threads) {
    if (n == 0) return FZH_OK;
    if (int rc = check_shape(nx, ny, nz)) return rc;
    if (!d_chunks || !d_abs_eb || !h_sizes || !d_ws) return fail(FZH_E_INVALID, "labeller: null pointer");
    if ((uintptr_t)d_chunks % 16) return fail(FZH_E_INVALID, "labeller: chunks must be 16-byte aligned");
    const size_t E = (size_t)nx * ny * nz;
    const LabelWs L = label_layout(n, E);
    if (ws_bytes < L.total) return fail(FZH_E_CAPACITY, "labeller: workspace too small");
    hipStream_t s = S(stream);
    if (int rc = label_launch_pass1(s, d_chunks, n, nx, ny, nz, d_abs_eb, d_ws)) return rc;
    uint8_t* w = (uint8_t*)d_ws;
    std::vector<uint32_t> hist(kBklen * n);
    FZH_TRY_HIP(hipMemcpyAsync(hist.data(), w + L.o_hist, 4 * hist.size(), hipMemcpyDeviceToHost, s));
    FZH_TRY_HIP(hipStreamSynchronize(s));
    // per-chunk canonical codebooks (the HuffmanStage's host build), code lengths only
    std::vector<uint8_t> lens(kBklen * n);
    std::atomic<size_t> next{0};
    std::atomic<int> bad{0};
    unsigned nt = host_threads > 0 ? (unsigned)host_threads : std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 16u));
    auto work = [&] {
        std::vector<uint32_t> bk4(kBklen);
        std::vector<uint8_t> rev(fzh_huffman_revbook_bytes(kBklen, 2));
        for (size_t c; (c = next.fetch_add(1)) < n;) {
            if (fzh_huffman_build_book(hist.data() + c * kBklen, kBklen, 2, bk4.data(), rev.data()) != FZH_OK) {
                bad = 1;
                continue;
            }
            for (uint32_t i = 0; i < kBklen; i++) lens[c * kBklen + i] = (uint8_t)(bk4[i] >> 27);
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    if (bad) return fail(FZH_E_RUNTIME, "labeller: codebook build failed");
    FZH_TRY_HIP(hipMemcpyAsync(w + L.o_lens, lens.data(), lens.size(), hipMemcpyHostToDevice, s));
    if (int rc = label_launch_pass2(s, n, E, d_ws)) return rc;
    std::vector<uint32_t> st(6 * n);
    FZH_TRY_HIP(hipMemcpyAsync(st.data(), w + L.o_stats, 4 * st.size(), hipMemcpyDeviceToHost, s));
    FZH_TRY_HIP(hipStreamSynchronize(s));
    // container sizes: [u32 4][u64 x 4] -> 48 B, then each output padded to 16
    const uint64_t max_out = (uint64_t)std::ceil((double)E * FZS_OUTLIER_CAP);
    const uint64_t nb = 2 * E / CS;
    const uint64_t pardeg = (E + kSub - 1) / kSub;
    const uint64_t rev = fzh_huffman_revbook_bytes(kBklen, 2);
    for (size_t c = 0; c < n; c++) {
        const uint64_t k = std::min<uint64_t>(st[c], max_out);
        const uint64_t R = st[n + c], rz = st[2 * n + c], cells = st[3 * n + c];
        const uint64_t qmax = st[4 * n + c], qo = std::min<uint64_t>(st[5 * n + c], max_out);
        const uint64_t lists = 2 * pad16(4 * k) + 16;
        uint64_t nbits = 16;
        for (uint64_t b = 1; b < 16; b *= 2)
            if (qmax < (1ull << b)) { nbits = b; break; }
        uint64_t* o = h_sizes + c * FZS_NCHAINS;
        o[0] = 48 + pad16(4 + pad4(2 * R) + 4 * R) + lists;
        o[1] = 48 + pad16(pad4(8 + 4 * nb + rz)) + lists;
        o[2] = 48 + pad16(128 + rev + 8 * pardeg + 4 * cells) + lists;
        o[3] = 48 + pad16((E * nbits + 7) / 8) + 2 * pad16(4 * qo) + 16;
    }
    return FZH_OK;
}

}  // extern "C"

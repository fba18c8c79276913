// huffman.hip — histogram and coarse Huffman encode/decode (gfx950).
//
// PHF stream (modules/coders/huffman/phf/hf_hl.cc:43-120, hf_buf.cc:151-176):
//   [phf_header 64 B][64 B (zeros here)][revbook][par_nbit u32 x pardeg]
//   [par_entry u32 x pardeg][bitstream u32 cells]
// phf_header: bklen (int:16) @0, sublen @4, pardeg @8, original_len u64 @16,
// total_nbit u64 @24, total_ncell u64 @32, entry[6] u32 @40 (entry[5] = total
// stream bytes).  Partitions of sublen = 768 symbols (coarse,
// hf_bk.cc:28-31) or 1024 (fine, the ReVISIT-lite path, hf_buf.cc:48-49);
// each partition's codewords are packed MSB-first into 32-bit cells
// (hf_kernels.cu:73-122, :145-250 build the same cells); par_entry =
// exclusive prefix of the per-partition cell counts.
//
// MI355X design
//   histogram: privatized per-wave LDS histograms, consecutive equal symbols
//     merged in-lane before the LDS atomic (the dominant code otherwise
//     serialises a 64-lane conflict).
//   encode: one wave per partition packs its codewords in parallel (bit
//     offsets from a wave prefix sum, LDS OR into cells) and writes them to
//     the partition's scratch slot plus par_nbit; a segmented device scan of
//     the cell counts writes par_entry into the stream; a pack kernel moves
//     the cells and a one-workgroup kernel writes the header: no host prefix
//     sum, no host sync (reference phases 1-4 + two syncs).
//   decode: one lane per partition with a 1024-entry prefix lookup table in
//     LDS (codes up to 10 bits resolve in one probe), canonical bit loop only
//     for longer codes; outputs leave in 16-byte stores.
#include "fzh_scan.h"

namespace fzh {
namespace {

constexpr int SUBLEN = 768;       // coarse partitions
constexpr int SUBLEN_FINE = 1024; // fine partitions
constexpr int HDR_BYTES = 128;
constexpr int LUT_BITS = 10;

template <typename T>
__device__ __forceinline__ uint32_t ld_sym(const T* p, uint64_t i) { return (uint32_t)p[i]; }

// ---------------------------------------------------------------------------
// Histogram (histogram.cu:26-60 semantics: symbols >= bklen are skipped)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
hist_kernel(const T* __restrict__ in, uint64_t n, uint32_t bklen, uint32_t* __restrict__ freq, int replicas) {
    extern __shared__ uint32_t sh[];
    const int wv = threadIdx.x >> 6;
    const int rep = replicas == 4 ? wv : 0;
    for (uint32_t i = threadIdx.x; i < bklen * (uint32_t)replicas; i += blockDim.x) sh[i] = 0;
    __syncthreads();
    uint32_t* h = sh + (size_t)rep * bklen;
    constexpr int E = 16 / sizeof(T);
    const uint64_t nv = n / E;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += stride) {
        const uint4 w = reinterpret_cast<const uint4*>(in)[v];
        const T* p = reinterpret_cast<const T*>(&w);
        uint32_t cur = (uint32_t)p[0], cnt = 1;
#pragma unroll
        for (int k = 1; k < E; k++) {
            const uint32_t s = (uint32_t)p[k];
            if (s == cur) {
                cnt++;
            } else {
                if (cur < bklen) atomicAdd(&h[cur], cnt);
                cur = s;
                cnt = 1;
            }
        }
        if (cur < bklen) atomicAdd(&h[cur], cnt);
    }
    for (uint64_t i = nv * E + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t s = (uint32_t)in[i];
        if (s < bklen) atomicAdd(&h[s], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < bklen; b += blockDim.x) {
        uint32_t s = 0;
        for (int r = 0; r < replicas; r++) s += sh[(size_t)r * bklen + b];
        if (s) atomicAdd(&freq[b], s);
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
hist_global_kernel(const T* __restrict__ in, uint64_t n, uint32_t bklen, uint32_t* __restrict__ freq) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = (uint32_t)in[i];
        if (s < bklen) atomicAdd(&freq[s], 1u);
    }
}

// ---------------------------------------------------------------------------
// Encode
// ---------------------------------------------------------------------------
// cells per partition slot (every codeword is at most 28 bits)
constexpr int cells_of(int sl) { return (sl * 28 + 31) / 32 + 2; }

template <typename T, int SL>
__global__ void __launch_bounds__(256)
huf_enc_kernel(const T* __restrict__ in, uint64_t n, uint32_t bklen, const uint32_t* __restrict__ bk4,
               const uint8_t* __restrict__ revbook, uint32_t rev_bytes, uint8_t* __restrict__ out,
               uint32_t* __restrict__ scells, uint32_t* __restrict__ ncell_out, uint32_t pardeg, int book_in_lds) {
    constexpr int CELLS = cells_of(SL);
    constexpr int K = SL / 64;  // symbols per lane
    extern __shared__ uint32_t smem[];
    const int wv = threadIdx.x >> 6;
    uint32_t* cells = smem + wv * CELLS;
    uint32_t* s_bk = smem + 4 * CELLS;
    if (book_in_lds) {
        for (uint32_t i = threadIdx.x; i < bklen; i += blockDim.x) s_bk[i] = bk4[i];
        __syncthreads();
    }
    const uint32_t* book = book_in_lds ? s_bk : bk4;
    const uint32_t ent1 = HDR_BYTES, ent2 = ent1 + rev_bytes;
    const uint32_t p = blockIdx.x * 4 + wv;
    const int l = lane_id();
    if (p >= pardeg) return;
    const uint64_t s0 = (uint64_t)p * SL;
    const uint32_t cnt = (uint32_t)((n - s0) < (uint64_t)SL ? (n - s0) : SL);
    for (int i = l; i < CELLS; i += 64) cells[i] = 0u;
    // symbol s0 + l + 64k, bit offsets in (k, lane) order = symbol order
    uint32_t code[K], len[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const uint32_t j = l + 64 * k;
        uint32_t w = 0;
        if (j < cnt) w = book[ld_sym(in, s0 + j)];
        len[k] = j < cnt ? (w >> 27) : 0u;
        code[k] = w & 0x07FFFFFFu;
    }
    wave_lds_sync();
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const uint32_t incl = wave_incl_scan(len[k]);
        const uint32_t b = run + incl - len[k];
        const uint32_t L = len[k];
        if (L) {
            const uint32_t c = b >> 5, sh = b & 31u;
            if (sh + L <= 32) {
                atomicOr(&cells[c], code[k] << (32 - sh - L));
            } else {
                atomicOr(&cells[c], code[k] >> (sh + L - 32));
                atomicOr(&cells[c + 1], code[k] << (64 - sh - L));
            }
        }
        run += lane63(incl);
    }
    const uint32_t nbit = run;
    const uint32_t ncell = (nbit + 31) / 32;
    wave_lds_sync();
    uint32_t* dst = scells + (uint64_t)p * CELLS;
    for (uint32_t i = l; i < ncell; i += 64) dst[i] = cells[i];
    if (l == 0) {
        reinterpret_cast<uint32_t*>(out + ent2)[p] = nbit;
        ncell_out[p] = ncell;
    }
    if (p == 0) {
        for (uint32_t i = l; i < rev_bytes; i += 64) out[ent1 + i] = revbook[i];
        for (uint32_t i = 64 + l; i < HDR_BYTES; i += 64) out[i] = 0;
    }
}

// Cells of partition p -> stream cell par_entry[p] (one wave per partition).
__global__ void __launch_bounds__(256)
huf_pack_kernel(const uint32_t* __restrict__ scells, const uint32_t* __restrict__ ncell,
                uint8_t* __restrict__ out, uint32_t ent3, uint32_t ent4, uint32_t pardeg, uint32_t cells) {
    const uint32_t p = blockIdx.x * 4 + wave_in_wg();
    if (p >= pardeg) return;
    const int l = lane_id();
    const uint32_t e = reinterpret_cast<const uint32_t*>(out + ent3)[p];
    const uint32_t c = ncell[p];
    const uint32_t* src = scells + (uint64_t)p * cells;
    uint32_t* dst = reinterpret_cast<uint32_t*>(out + ent4) + e;
    for (uint32_t i = l; i < c; i += 64) dst[i] = src[i];
}

// phf_header (hf.h:38-44) once every partition is placed.
__global__ void __launch_bounds__(1024)
huf_header_kernel(uint8_t* __restrict__ out, uint64_t n, uint32_t bklen, uint32_t sublen, uint32_t ent1, uint32_t ent2,
                  uint32_t ent3, uint32_t ent4, uint32_t pardeg, const uint32_t* __restrict__ total_ncell,
                  uint64_t* __restrict__ out_size) {
    __shared__ unsigned long long s_w[16];
    const uint32_t* nb = reinterpret_cast<const uint32_t*>(out + ent2);
    unsigned long long t = 0;
    for (uint32_t p = threadIdx.x; p < pardeg; p += 1024) t += nb[p];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane_id() == 0) s_w[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tnb = 0;
        for (int i = 0; i < 16; i++) tnb += s_w[i];
        const uint32_t tc = *total_ncell;
        const uint32_t ent5 = ent4 + 4 * tc;
        uint32_t* h32 = reinterpret_cast<uint32_t*>(out);
        h32[0] = bklen & 0xFFFFu;
        h32[1] = sublen;
        h32[2] = pardeg;
        h32[3] = 0;
        reinterpret_cast<uint64_t*>(out)[2] = n;
        reinterpret_cast<uint64_t*>(out)[3] = tnb;
        reinterpret_cast<uint64_t*>(out)[4] = tc;
        h32[10] = 0;
        h32[11] = ent1;
        h32[12] = ent2;
        h32[13] = ent3;
        h32[14] = ent4;
        h32[15] = ent5;
        *out_size = ent5;
    }
}

// ---------------------------------------------------------------------------
// Decode
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
huf_dec_kernel(const uint8_t* __restrict__ in, uint64_t in_bytes, T* __restrict__ out, uint64_t n,
               int* __restrict__ d_status) {
    __shared__ uint32_t s_first[32], s_entry[32];
    __shared__ uint32_t s_lut[1 << LUT_BITS];  // (len << 24) | key index ; 0 = long code
    if (in_bytes < HDR_BYTES) {
        if (threadIdx.x == 0) *d_status = 1;
        return;
    }
    const int32_t sublen = *reinterpret_cast<const int32_t*>(in + 4);
    const int32_t pardeg = *reinterpret_cast<const int32_t*>(in + 8);
    const uint32_t* ent = reinterpret_cast<const uint32_t*>(in + 40);
    const uint32_t e1 = ent[1], e2 = ent[2], e3 = ent[3], e4 = ent[4], e5 = ent[5];
    const bool hdr_ok = sublen > 0 && pardeg >= 0 && (uint64_t)pardeg == (n + sublen - 1) / sublen &&
                        e1 <= e2 && e2 - e1 >= 256 && e2 <= e3 && e3 <= e4 && e4 <= e5 && e5 <= in_bytes &&
                        e3 - e2 == 4ull * pardeg && e4 - e3 == 4ull * pardeg && (e4 % 4) == 0 && (e1 % 4) == 0;
    if (!hdr_ok) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *d_status = 1;
        return;
    }
    const uint32_t nkeys = (e2 - e1 - 256) / sizeof(T);
    const T* keys = reinterpret_cast<const T*>(in + e1 + 256);
    if (threadIdx.x < 32) {
        s_first[threadIdx.x] = reinterpret_cast<const uint32_t*>(in + e1)[threadIdx.x];
        s_entry[threadIdx.x] = reinterpret_cast<const uint32_t*>(in + e1 + 128)[threadIdx.x];
    }
    __syncthreads();
    // prefix table: simulate the canonical loop on every LUT_BITS-bit prefix
    for (uint32_t pat = threadIdx.x; pat < (1u << LUT_BITS); pat += blockDim.x) {
        uint32_t v = pat >> (LUT_BITS - 1), lv = 1, res = 0;
        for (int i = 1;; i++) {
            if (!(v < s_first[lv])) {
                const uint32_t k = s_entry[lv] + v - s_first[lv];
                res = (k < nkeys && k < (1u << 24)) ? ((lv << 24) | k) : 0u;
                break;
            }
            if (i == LUT_BITS) break;
            v = (v << 1) | ((pat >> (LUT_BITS - 1 - i)) & 1u);
            lv++;
        }
        s_lut[pat] = res;
    }
    __syncthreads();
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)pardeg) return;
    const uint32_t nbit = reinterpret_cast<const uint32_t*>(in + e2)[p];
    const uint32_t pe = reinterpret_cast<const uint32_t*>(in + e3)[p];
    const uint32_t ncells_total = (e5 - e4) / 4;
    const uint32_t ncell = (nbit + 31) / 32;
    const uint64_t o0 = (uint64_t)p * sublen;
    const uint32_t cnt = (uint32_t)((n - o0) < (uint64_t)sublen ? (n - o0) : (uint64_t)sublen);
    T* dst = out + o0;
    if ((uint64_t)pe + ncell > ncells_total) {
        *d_status = 1;
        return;
    }
    if (s_entry[1] == 1 && nbit == 0) {  // one symbol with a zero-length code
        const T k0 = nkeys ? keys[0] : (T)0;
        for (uint32_t j = 0; j < cnt; j++) dst[j] = k0;
        return;
    }
    const uint32_t* cells = reinterpret_cast<const uint32_t*>(in + e4) + pe;
    // 64-bit window: `buf` holds the next `avail` bits MSB-aligned
    uint64_t buf = 0;
    int avail = 0;
    uint32_t next_cell = 0;
    auto refill = [&]() {
        while (avail <= 32) {
            const uint32_t c = next_cell < ncell ? cells[next_cell] : 0u;
            next_cell++;
            buf |= (uint64_t)c << (32 - avail);
            avail += 32;
        }
    };
    uint32_t used = 0;
    bool bad = false;
    auto decode_one = [&]() -> T {
        refill();
        const uint32_t pat = (uint32_t)(buf >> (64 - LUT_BITS));
        const uint32_t e = s_lut[pat];
        uint32_t key, L;
        if (e) {
            L = e >> 24;
            key = e & 0xFFFFFFu;
        } else {
            // long code: continue the canonical loop from LUT_BITS bits
            uint32_t v = pat, lv = LUT_BITS;
            int i = LUT_BITS;
            bool done = false;
            while (lv < 31) {
                v = (v << 1) | (uint32_t)((buf >> (63 - i)) & 1u);
                lv++;
                i++;
                if (!(v < s_first[lv])) { done = true; break; }
            }
            const uint32_t k = s_entry[lv] + v - s_first[lv];
            if (!done || k >= nkeys) {
                bad = true;
                return (T)0;
            }
            L = lv;
            key = k;
        }
        buf <<= L;
        avail -= (int)L;
        used += L;
        return keys[key];
    };
    constexpr int E = 16 / sizeof(T);
    const bool vec = ((uint64_t)sublen * sizeof(T)) % 16 == 0;
    for (uint32_t g = 0; g < cnt && !bad; g += E) {
        T pack[E];
#pragma unroll
        for (int q = 0; q < E; q++) pack[q] = (g + q < cnt && !bad) ? decode_one() : (T)0;
        if (vec && g + E <= cnt) {
            *reinterpret_cast<uint4*>(dst + g) = *reinterpret_cast<const uint4*>(pack);
        } else {
#pragma unroll
            for (int q = 0; q < E; q++)
                if (g + q < cnt) dst[g + q] = pack[q];
        }
    }
    if (bad || used != nbit) *d_status = 1;
}

struct EncWs {
    size_t pardeg, o_ncell, o_tot, o_scan, o_cells, total;
};
EncWs enc_layout(size_t n, int sl) {
    EncWs w;
    w.pardeg = cdiv(std::max<size_t>(n, 1), (size_t)sl);
    size_t o = 0;
    w.o_ncell = o; o = align_up(o + 4 * w.pardeg, 256);
    w.o_tot = o;   o = align_up(o + 16, 256);
    w.o_scan = o;  o = align_up(o + scan_ws_bytes(w.pardeg), 256);
    w.o_cells = o; o = align_up(o + 4 * (size_t)cells_of(sl) * w.pardeg, 256);
    w.total = o;
    return w;
}

template <typename T, int SL>
int enc_impl(hipStream_t s, const void* in, size_t n, uint32_t bklen, const uint32_t* bk4, const uint8_t* rev,
             void* out, size_t cap, uint64_t* out_size, void* ws, size_t ws_bytes) {
    constexpr int CELLS = cells_of(SL);
    const EncWs L = enc_layout(n, SL);
    if (ws_bytes < L.total) return fail(FZH_E_CAPACITY, "huffman enc: workspace too small");
    if (n == 0) return fail(FZH_E_INVALID, "huffman enc: empty input");
    if (n > 0x7FFFFFFFull * SL) return fail(FZH_E_INVALID, "huffman enc: input too large");
    if (cap < fzh_huffman_enc_bound(n, bklen, sizeof(T))) return fail(FZH_E_CAPACITY, "huffman enc: output below bound");
    const uint32_t rev_bytes = (uint32_t)fzh_huffman_revbook_bytes(bklen, sizeof(T));
    if (rev_bytes % 4) return fail(FZH_E_INVALID, "huffman enc: bklen * sizeof(T) must be a multiple of 4");
    if ((uintptr_t)out % 16) return fail(FZH_E_INVALID, "huffman enc: output must be 16-byte aligned");
    const uint32_t pardeg = (uint32_t)L.pardeg;
    const uint32_t ent1 = HDR_BYTES, ent2 = ent1 + rev_bytes, ent3 = ent2 + 4 * pardeg, ent4 = ent3 + 4 * pardeg;
    uint8_t* w = (uint8_t*)ws;
    uint32_t* ncell = (uint32_t*)(w + L.o_ncell);
    uint32_t* tot = (uint32_t*)(w + L.o_tot);
    uint32_t* scells = (uint32_t*)(w + L.o_cells);
    uint8_t* o8 = (uint8_t*)out;
    const int book_in_lds = bklen <= 8192 ? 1 : 0;
    const size_t lds = (size_t)4 * CELLS * 4 + (book_in_lds ? 4 * (size_t)bklen : 0);
    const unsigned blocks = (unsigned)cdiv(pardeg, 4);
    launch_k("huf_enc_kernel", huf_enc_kernel<T, SL>, dim3(blocks), dim3(256), (uint32_t)lds, s, (const T*)in, n, bklen, bk4, rev,
             rev_bytes, o8, scells, ncell, pardeg, book_in_lds);
    if (int rc = scan_u32(s, ncell, reinterpret_cast<uint32_t*>(o8 + ent3), pardeg, nullptr, tot, w + L.o_scan))
        return rc;
    launch_k("huf_pack_kernel", huf_pack_kernel, dim3(blocks), dim3(256), 0, s, scells, ncell, o8, ent3, ent4, pardeg, (uint32_t)CELLS);
    launch_k("huf_header_kernel", huf_header_kernel, dim3(1), dim3(1024), 0, s, o8, n, bklen, (uint32_t)SL, ent1, ent2, ent3, ent4, pardeg, tot, out_size);
    FZH_CHECK_LAUNCH("huffman encode");
    return FZH_OK;
}

template <int SL>
int encode_sl(void* stream, const void* d_in, size_t n, int elem_bytes, uint32_t bklen, const uint32_t* d_bk4,
                     const uint8_t* d_revbook, void* d_out, size_t out_cap, uint64_t* d_out_size, void* d_ws,
                     size_t ws_bytes) {
    if (ws_bytes < std::max(enc_layout(n, SUBLEN).total, enc_layout(n, SUBLEN_FINE).total)) return fail(FZH_E_CAPACITY, "huffman enc: workspace too small");
    switch (elem_bytes) {
    case 1: return enc_impl<uint8_t, SL>(S(stream), d_in, n, bklen, d_bk4, d_revbook, d_out, out_cap, d_out_size, d_ws, ws_bytes);
    case 2: return enc_impl<uint16_t, SL>(S(stream), d_in, n, bklen, d_bk4, d_revbook, d_out, out_cap, d_out_size, d_ws, ws_bytes);
    case 4: return enc_impl<uint32_t, SL>(S(stream), d_in, n, bklen, d_bk4, d_revbook, d_out, out_cap, d_out_size, d_ws, ws_bytes);
    default: return fail(FZH_E_INVALID, "huffman enc: elem_bytes must be 1, 2 or 4");
    }
}

}  // namespace
}  // namespace fzh

using namespace fzh;

extern "C" {

int fzh_histogram(void* stream, const void* d_in, size_t n, int elem_bytes, uint32_t bklen, uint32_t* d_freq) {
    if (bklen == 0) return fail(FZH_E_INVALID, "histogram: bklen must be > 0");
    if (int rc = zero_bytes(S(stream), d_freq, 4 * (size_t)bklen)) return rc;
    if (n == 0) return FZH_OK;
    if ((uintptr_t)d_in % 16) return fail(FZH_E_INVALID, "histogram: input must be 16-byte aligned");
    const int replicas = bklen <= 4096 ? 4 : (bklen <= 16384 ? 1 : 0);
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(n, 256 * 64), 1024));
    hipStream_t s = S(stream);
#define FZH_HIST(T)                                                                                       \
    if (replicas)                                                                                         \
        launch_k("hist_kernel", hist_kernel<T>, dim3(blocks), dim3(256), (uint32_t)(4 * bklen * replicas), s, (const T*)d_in, n, bklen, d_freq, replicas); \
    else                                                                                                  \
        launch_k("hist_global_kernel", hist_global_kernel<T>, dim3(blocks), dim3(256), 0, s, (const T*)d_in, n, bklen, d_freq);
    switch (elem_bytes) {
    case 1: FZH_HIST(uint8_t) break;
    case 2: FZH_HIST(uint16_t) break;
    case 4: FZH_HIST(uint32_t) break;
    default: return fail(FZH_E_INVALID, "histogram: elem_bytes must be 1, 2 or 4");
    }
#undef FZH_HIST
    FZH_CHECK_LAUNCH("hist_kernel");
    return FZH_OK;
}

size_t fzh_huffman_enc_bound(size_t n, uint32_t bklen, int elem_bytes) {
    const size_t pardeg = cdiv(std::max<size_t>(n, 1), SUBLEN);
    // every codeword is at most 28 bits
    return HDR_BYTES + fzh_huffman_revbook_bytes(bklen, elem_bytes) + 8 * pardeg + 4 * cdiv(n * 28, 32) + 4 * pardeg + 64;
}

size_t fzh_huffman_enc_ws_bytes(size_t n) {
    return std::max(enc_layout(n, SUBLEN).total, enc_layout(n, SUBLEN_FINE).total);
}

int fzh_huffman_encode(void* stream, const void* d_in, size_t n, int elem_bytes, uint32_t bklen,
                       const uint32_t* d_bk4, const uint8_t* d_revbook, void* d_out, size_t out_cap,
                       uint64_t* d_out_size, void* d_ws, size_t ws_bytes) {
    return encode_sl<SUBLEN>(stream, d_in, n, elem_bytes, bklen, d_bk4, d_revbook, d_out, out_cap, d_out_size, d_ws,
                             ws_bytes);
}

int fzh_huffman_encode_fine(void* stream, const void* d_in, size_t n, int elem_bytes, uint32_t bklen,
                            const uint32_t* d_bk4, const uint8_t* d_revbook, void* d_out, size_t out_cap,
                            uint64_t* d_out_size, void* d_ws, size_t ws_bytes) {
    return encode_sl<SUBLEN_FINE>(stream, d_in, n, elem_bytes, bklen, d_bk4, d_revbook, d_out, out_cap, d_out_size,
                                  d_ws, ws_bytes);
}

int fzh_huffman_decode(void* stream, const void* d_in, size_t in_bytes, int elem_bytes, void* d_out, size_t n,
                       int* d_status) {
    if (n == 0) return FZH_OK;
    if (!d_status) return fail(FZH_E_INVALID, "huffman dec: null status");
    if ((uintptr_t)d_in % 4 || (uintptr_t)d_out % 16) return fail(FZH_E_INVALID, "huffman dec: input 4-byte, output 16-byte alignment required");
    const size_t pardeg = cdiv(n, SUBLEN);
    const unsigned blocks = (unsigned)cdiv(pardeg, 256);
    hipStream_t s = S(stream);
    switch (elem_bytes) {
    case 1: launch_k("huf_dec_kernel", huf_dec_kernel<uint8_t>, dim3(blocks), dim3(256), 0, s, (const uint8_t*)d_in, in_bytes, (uint8_t*)d_out, n, d_status); break;
    case 2: launch_k("huf_dec_kernel", huf_dec_kernel<uint16_t>, dim3(blocks), dim3(256), 0, s, (const uint8_t*)d_in, in_bytes, (uint16_t*)d_out, n, d_status); break;
    case 4: launch_k("huf_dec_kernel", huf_dec_kernel<uint32_t>, dim3(blocks), dim3(256), 0, s, (const uint8_t*)d_in, in_bytes, (uint32_t*)d_out, n, d_status); break;
    default: return fail(FZH_E_INVALID, "huffman dec: elem_bytes must be 1, 2 or 4");
    }
    FZH_CHECK_LAUNCH("huf_dec_kernel");
    return FZH_OK;
}

}  // extern "C"

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
//   decode: one wave per partition: 64 lanes decode equal spans of its bits
//     speculatively and settle the span starts in rounds (a prefix code
//     resynchronises within a few codewords), a 4096-entry prefix table
//     yields the symbol of any code of up to 12 bits in one LDS probe and
//     longer codes' lengths come from per-length limit compares; symbols are
//     staged in LDS and leave in 16-byte stores.
#include "fzh_scan.h"

namespace fzh {
namespace {

constexpr int SUBLEN = 768;       // coarse partitions
constexpr int SUBLEN_FINE = 1024; // fine partitions
constexpr int HDR_BYTES = 128;

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
    // 8 replicas: two per wave, by lane parity (half the same-bin lanes of
    // one LDS atomic); replica stride bklen + 1 words, so one bin of two
    // replicas sits in different banks
    const uint32_t rs = replicas == 8 ? bklen + 1 : bklen;
    const int rep = replicas == 8 ? 2 * wv + (threadIdx.x & 1) : (replicas == 4 ? wv : 0);
    for (uint32_t i = threadIdx.x; i < rs * (uint32_t)replicas; i += blockDim.x) sh[i] = 0;
    __syncthreads();
    uint32_t* h = sh + (size_t)rep * rs;
    constexpr int E = 16 / sizeof(T);
    const uint64_t nv = n / E;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // four 16-byte loads per thread and step, issued before the atomics
    for (uint64_t v0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v0 < nv; v0 += 4 * stride) {
        uint4 w[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint64_t v = v0 + (uint64_t)u * stride;
            w[u] = v < nv ? reinterpret_cast<const uint4*>(in)[v] : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (v0 + (uint64_t)u * stride >= nv) continue;
            const T* p = reinterpret_cast<const T*>(&w[u]);
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
    }
    for (uint64_t i = nv * E + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t s = (uint32_t)in[i];
        if (s < bklen) atomicAdd(&h[s], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < bklen; b += blockDim.x) {
        uint32_t s = 0;
        for (int r = 0; r < replicas; r++) s += sh[(size_t)r * rs + b];
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

// BOOK_LDS: the codebook is staged in LDS (bklen <= 8192), else read from
// global memory.  (A template parameter, not a pointer chosen at run time:
// a pointer that may be either LDS or global makes every lookup a FLAT
// access.)
template <typename T, int SL, bool BOOK_LDS>
__global__ void __launch_bounds__(256)
huf_enc_kernel(const T* __restrict__ in, uint64_t n, uint32_t bklen, const uint32_t* __restrict__ bk4,
               const uint8_t* __restrict__ revbook, uint32_t rev_bytes, uint8_t* __restrict__ out,
               uint32_t* __restrict__ scells, uint32_t* __restrict__ ncell_out, uint32_t pardeg) {
    constexpr int CELLS = cells_of(SL);
    constexpr int K = SL / 64;  // symbols per lane
    extern __shared__ uint32_t smem[];
    const int wv = threadIdx.x >> 6;
    uint32_t* cells = smem + wv * CELLS;
    uint32_t* s_bk = smem + 4 * CELLS;
    if constexpr (BOOK_LDS) {
        for (uint32_t i = threadIdx.x; i < bklen; i += blockDim.x) s_bk[i] = bk4[i];
        __syncthreads();
    }
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
        if (j < cnt) {
            const uint32_t sym = ld_sym(in, s0 + j);
            if (sym < bklen) w = BOOK_LDS ? s_bk[sym] : bk4[sym];  // an out-of-range symbol (reported by its caller) gets no code
        }
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

// Cells of partition p -> stream cell par_entry[p] (one wave per partition);
// each workgroup adds its partitions' bit counts to one of NBIT_SLOTS
// accumulators (one atomic per workgroup) for the header.
constexpr uint32_t NBIT_SLOTS = 256;

__global__ void __launch_bounds__(256)
huf_pack_kernel(const uint32_t* __restrict__ scells, const uint32_t* __restrict__ ncell,
                uint8_t* __restrict__ out, uint32_t ent2, uint32_t ent3, uint32_t ent4, uint32_t pardeg,
                uint32_t cells, unsigned long long* __restrict__ nbit_acc) {
    __shared__ uint32_t s_nb[4];
    const int w = (int)wave_in_wg();
    const uint32_t p = blockIdx.x * 4 + w;
    const int l = lane_id();
    uint32_t nb = 0;
    if (p < pardeg) {
        const uint32_t e = reinterpret_cast<const uint32_t*>(out + ent3)[p];
        const uint32_t c = ncell[p];
        const uint32_t* src = scells + (uint64_t)p * cells;
        uint32_t* dst = reinterpret_cast<uint32_t*>(out + ent4) + e;
        // up to four loads in flight per lane before their stores
        for (uint32_t i = l; i < c; i += 256) {
            uint32_t v[4];
#pragma unroll
            for (int k = 0; k < 4; k++) v[k] = i + 64 * k < c ? src[i + 64 * k] : 0u;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (i + 64 * k < c) dst[i + 64 * k] = v[k];
        }
        nb = reinterpret_cast<const uint32_t*>(out + ent2)[p];
    }
    if (l == 0) s_nb[w] = nb;
    __syncthreads();
    if (threadIdx.x == 0)  // spread over NBIT_SLOTS words: one address serialises every atomic
        atomicAdd(nbit_acc + (blockIdx.x % NBIT_SLOTS), (unsigned long long)s_nb[0] + s_nb[1] + s_nb[2] + s_nb[3]);
}

// phf_header (hf.h:38-44) once every partition is placed; sums the bit
// total from the NBIT_SLOTS accumulators and clears them for the next call.
__global__ void __launch_bounds__(256)
huf_header_kernel(uint8_t* __restrict__ out, uint64_t n, uint32_t bklen, uint32_t sublen, uint32_t ent1, uint32_t ent2,
                  uint32_t ent3, uint32_t ent4, uint32_t pardeg, const uint32_t* __restrict__ total_ncell,
                  uint64_t* __restrict__ out_size, unsigned long long* __restrict__ nbit_acc) {
    __shared__ unsigned long long s_w[4];
    unsigned long long a = 0;
    for (uint32_t i = threadIdx.x; i < NBIT_SLOTS; i += 256) {
        a += nbit_acc[i];
        nbit_acc[i] = 0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane_id() == 0) s_w[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long tnb = s_w[0] + s_w[1] + s_w[2] + s_w[3];
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
// One wave per partition.  Codeword boundaries inside a partition are only
// known by decoding from its first bit, so the wave cuts the partition's
// bits into 64 equal spans (of at least 32 bits, longer than any codeword)
// and decodes them speculatively, all lanes at once: lane j starts at the
// first bit of span j and decodes the codewords that start inside its span,
// which leaves it at the end of its last codeword -- the first codeword
// boundary at or past span j+1's first bit, i.e. where lane j+1 really
// starts.  Lanes whose start moves decode again from the corrected start,
// and stop as soon as they reach a boundary of their previous decode (the
// first 64 bits of each span keep a boundary mask): from there on both
// decodes see the same codewords, so the count follows from the mask and
// the end is unchanged.  A prefix code falls back into step within a few
// codewords, so one short round nearly always settles the wave; lane 0 is
// always right and a lane whose predecessor is right becomes right in the
// next round, so 64 rounds bound it.  A prefix sum of the per-lane counts
// gives each lane its output offset; a last decode of each span writes its
// symbols to LDS and the wave stores the partition with 16-byte stores.
//
// Per codeword: a probe of a 2^12-entry table (length and symbol -- the
// symbol itself for 1- and 2-byte keys -- of every code of up to 12 bits);
// the 12-bit prefixes of longer codes point into a second table indexed by
// the next B2 bits (B2 chosen so the table fits; all long codes of typical
// books); codes longer still take the canonical comparison against the
// per-length limits first[l] << (32 - l), which do not increase with l
// (first[l] = round((first[l+1] + numl[l+1]) / 2)), so the length is one
// more than the number of limits above the window (the bit loop of
// hf_kernels.cu:258-312 in a few compares).  A window with no code of the
// book decodes as a 31-bit code with a flag that fails the partition, so
// the loops never branch on it.  Bits come from a 64-bit register buffer
// refilled from the partition's cells in LDS.
//
// A partition's cells are staged in LDS when they fit (DEC_CELLS cells and
// partitions of at most DEC_SUBLEN_MAX symbols, which covers the coarse 768
// and fine 1024 layouts); any other partition is decoded by the same steps
// reading its cells from memory and writing its symbols there.
constexpr int DEC_LUT_BITS = 12;
constexpr uint32_t DEC_LUT2 = 1024;  // second-level entries
constexpr uint32_t DEC_CELLS = 448;
constexpr uint32_t DEC_SUBLEN_MAX = 1024;
constexpr uint32_t DEC_INVALID = 0xFFFFFFFFu;
// table entry: value (symbol, or key index with DEC_KEY) in bits 0-23,
// code length in bits 24-28
constexpr uint32_t DEC_KEY = 1u << 29;
constexpr uint32_t DEC_BAD = (1u << 30) | (31u << 24);  // no code: skip 31 bits, fail
constexpr uint32_t DEC_LONG = 1u << 31;                 // first level: prefix of a longer code

// waves per workgroup: the tables are built once per workgroup, so
// workgroups are large; 4-byte keys halve it to keep the LDS under 64 KiB
template <typename T>
constexpr int dec_waves() { return sizeof(T) == 4 ? 4 : 8; }

struct DecTables {
    const uint32_t* lut;   // DEC_LUT_BITS-bit prefix -> entry, or DEC_LONG | prefix
    const uint32_t* lut2;  // long prefix, next b2 bits -> entry; 0 = longer code or none
    const uint64_t* lim;   // first[l] << (32 - l)
    const int32_t* diff;   // entry[l] - first[l]
    uint32_t maxl, nkeys, b2;
};

// Entry of a code longer than DEC_LUT_BITS bits at the head of W, given its
// first-level entry e.
__device__ __forceinline__ uint32_t dec_long(const DecTables& t, uint32_t W, uint32_t e) {
    if (t.b2) {
        const uint32_t e2 = t.lut2[((e & 0xFFFFu) << t.b2) | ((W << DEC_LUT_BITS) >> (32 - t.b2))];
        if (e2) return e2;
    }
    uint32_t lv = DEC_LUT_BITS + 1;
    for (uint32_t l = DEC_LUT_BITS + 1; l < t.maxl; l++) lv += (uint64_t)W < t.lim[l] ? 1u : 0u;
    const uint32_t k = (W >> (32 - lv)) + (uint32_t)t.diff[lv];
    if (lv > t.maxl || (uint64_t)W < t.lim[lv] || k >= t.nkeys) return DEC_BAD;
    return (lv << 24) | DEC_KEY | k;
}

__device__ __forceinline__ uint32_t dec_entry(const DecTables& t, uint32_t W) {
    uint32_t e = t.lut[W >> (32 - DEC_LUT_BITS)];
    if ((int32_t)e < 0) e = dec_long(t, W, e);
    return e;
}

// MSB-first reader of a partition's cells: staged in LDS (with zero cells
// past the end) or read from memory with the end checked.  `buf` holds bits
// [p, 32 next) of the partition at its top; it is refilled whenever fewer
// than 32 remain, i.e. when p > thr = 32 (next - 1).
template <bool STAGED>
struct DecBits {
    const uint32_t* c;
    uint32_t ncell;
    uint64_t buf;
    uint32_t next, thr;
    __device__ __forceinline__ uint32_t at(uint32_t k) const { return STAGED ? c[k] : (k < ncell ? c[k] : 0u); }
    __device__ __forceinline__ void seek(uint32_t p) {
        const uint32_t k = p >> 5;
        buf = (((uint64_t)at(k) << 32) | at(k + 1)) << (p & 31u);
        next = k + 2;
        thr = 32 * k + 32;
    }
    __device__ __forceinline__ uint32_t window() const { return (uint32_t)(buf >> 32); }
    // advance from p to p + L (L < 32)
    __device__ __forceinline__ void consume(uint32_t p, uint32_t L) {
        buf <<= L;
        p += L;
        if (p > thr) {
            buf |= (uint64_t)at(next) << (p - thr);
            next++;
            thr += 32;
        }
    }
};

// One partition of nbit bits and cnt symbols, decoded by one wave into
// `ost` (LDS staging, STAGED) or straight into `dst`.  Returns false if the
// bits do not decode to exactly cnt symbols ending at bit nbit.
template <typename T, bool STAGED>
__device__ __forceinline__ bool dec_partition(const uint32_t* cells, uint32_t ncell, const DecTables& t, uint32_t nbit,
                                              uint32_t cnt, const T* __restrict__ keys, T* ost, T* __restrict__ dst) {
    const uint32_t l = (uint32_t)lane_id();
    // spans of at least 32 bits: longer than any codeword, so every span
    // holds the start of a codeword and a correction moves on by whole spans
    const uint32_t seg = max((nbit + 63) / 64, 32u);
    const uint32_t act = (nbit + seg - 1) / seg;  // lanes with a span (wave-uniform, <= 64)
    if (act == 0) return cnt == 0;
    const bool on = l < act;
    const uint32_t lo = min(l * seg, nbit), lim = min(l * seg + seg, nbit);
    const uint32_t lim64 = min(lim, lo + 64);
    DecBits<STAGED> bits{cells, ncell, 0, 0, 0};
    // speculative decode from lo: count, end, boundaries in [lo, lo + 64)
    uint32_t b = lo, c = 0, e = nbit;
    uint64_t m = 0;
    if (on) {
        uint32_t p = lo, acc = 0;
        bits.seek(p);
        while (p < lim64) {
            const uint32_t en = dec_entry(t, bits.window());
            const uint32_t L = (en >> 24) & 31u;
            acc |= en;
            m |= 1ull << (p - lo);
            bits.consume(p, L);
            p += L;
            c++;
        }
        while (p < lim) {
            const uint32_t en = dec_entry(t, bits.window());
            const uint32_t L = (en >> 24) & 31u;
            acc |= en;
            bits.consume(p, L);
            p += L;
            c++;
        }
        e = (acc & (1u << 30)) ? DEC_INVALID : p;
    }
    for (int round = 0; round < 64; round++) {
        const uint32_t prev = wave_shr1(e);  // every lane takes part in the lane shift
        const uint32_t nb = l == 0 ? 0u : prev;
        const bool moved = on && nb != b;
        if (!__ballot(moved)) break;
        if (moved) {
            // decode again from nb, up to the first boundary of the previous decode
            b = nb;
            uint32_t p = nb, steps = 0, acc = 0;
            uint64_t mn = 0;
            bool synced = false;
            if (nb != DEC_INVALID) {
                bits.seek(p);
                while (p < lim64) {
                    if ((m >> (p - lo)) & 1ull) {
                        synced = true;
                        break;
                    }
                    const uint32_t en = dec_entry(t, bits.window());
                    const uint32_t L = (en >> 24) & 31u;
                    acc |= en;
                    mn |= 1ull << (p - lo);
                    bits.consume(p, L);
                    p += L;
                    steps++;
                }
            }
            if (nb == DEC_INVALID) {
                e = DEC_INVALID;
                c = 0;
                m = 0;
            } else if (synced && !(acc & (1u << 30))) {
                // from p on both decodes see the same codewords: e stays
                const uint64_t below = (1ull << (p - lo)) - 1ull;
                c = steps + c - (uint32_t)__popcll(m & below);
                m = mn | (m & ~below);
            } else {
                while (p < lim) {
                    const uint32_t en = dec_entry(t, bits.window());
                    const uint32_t L = (en >> 24) & 31u;
                    acc |= en;
                    bits.consume(p, L);
                    p += L;
                    steps++;
                }
                e = (acc & (1u << 30)) ? DEC_INVALID : p;
                c = steps;
                m = mn;
            }
        }
    }
    const uint32_t incl = wave_incl_scan(c);
    const uint32_t last_e = (uint32_t)__builtin_amdgcn_readlane((int)e, (int)(act - 1));
    if (last_e != nbit || lane63(incl) != cnt) return false;
    // every span settled: decode it once more, writing its symbols
    uint32_t o = incl - c;
    if (b < lim) {
        uint32_t p = b;
        bits.seek(p);
        while (p < lim) {
            const uint32_t en = dec_entry(t, bits.window());
            const uint32_t L = (en >> 24) & 31u;
            uint32_t v = en & 0xFFFFFFu;
            if (en & DEC_KEY) v = (uint32_t)keys[v];
            if (STAGED)
                ost[o] = (T)v;
            else
                dst[o] = (T)v;
            o++;
            bits.consume(p, L);
            p += L;
        }
    }
    return true;
}

template <typename T>
__global__ void __launch_bounds__(64 * dec_waves<T>())
huf_dec_kernel(const uint8_t* __restrict__ in, uint64_t in_bytes, T* __restrict__ out, uint64_t n,
               int* __restrict__ d_status) {
    constexpr bool FOLD = sizeof(T) <= 2;
    constexpr int DEC_WAVES = dec_waves<T>();
    __shared__ uint32_t s_first[32], s_entry[32];
    __shared__ uint64_t s_lim[32];
    __shared__ int32_t s_diff[32];
    __shared__ uint32_t s_lut[1 << DEC_LUT_BITS];
    __shared__ uint32_t s_lut2[DEC_LUT2];
    __shared__ uint32_t s_cells[DEC_WAVES][DEC_CELLS + 4];
    __shared__ uint4 s_out[DEC_WAVES][DEC_SUBLEN_MAX * sizeof(T) / 16];
    if (in_bytes < HDR_BYTES) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *d_status = 1;
        return;
    }
    const int32_t sublen = *reinterpret_cast<const int32_t*>(in + 4);
    const int32_t pardeg = *reinterpret_cast<const int32_t*>(in + 8);
    const uint32_t* ent = reinterpret_cast<const uint32_t*>(in + 40);
    const uint32_t e1 = ent[1], e2 = ent[2], e3 = ent[3], e4 = ent[4], e5 = ent[5];
    const bool hdr_ok = sublen > 0 && pardeg >= 0 && (uint64_t)pardeg == (n + sublen - 1) / sublen &&
                        e1 <= e2 && e2 - e1 >= 256 && (e2 - e1 - 256) / sizeof(T) <= 65536 && e2 <= e3 &&
                        e3 <= e4 && e4 <= e5 && e5 <= in_bytes && e3 - e2 == 4ull * pardeg &&
                        e4 - e3 == 4ull * pardeg && (e4 % 4) == 0 && (e1 % 4) == 0;
    if (!hdr_ok) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *d_status = 1;
        return;
    }
    const uint32_t nkeys = (e2 - e1 - 256) / sizeof(T);
    const T* keys = reinterpret_cast<const T*>(in + e1 + 256);
    if (threadIdx.x < 32) {
        const uint32_t f = reinterpret_cast<const uint32_t*>(in + e1)[threadIdx.x];
        const uint32_t en = reinterpret_cast<const uint32_t*>(in + e1 + 128)[threadIdx.x];
        s_first[threadIdx.x] = f;
        s_entry[threadIdx.x] = en;
        s_lim[threadIdx.x] = threadIdx.x ? (uint64_t)f << (32 - threadIdx.x) : ~0ull;
        s_diff[threadIdx.x] = (int32_t)(en - f);
    }
    __syncthreads();
    // longest code of the book: the last length with a symbol (entry[l+1] > entry[l])
    uint32_t maxl = 1;
    for (uint32_t l = 1; l < 31; l++)
        if (s_entry[l + 1] > s_entry[l]) maxl = l;
    if (s_entry[31] < nkeys) maxl = 31;
    // The canonical loop on the window W from length l0 up to length lmax:
    // the entry of the code, DEC_BAD when the window starts no code of the
    // book, or (first level) DEC_LONG | (W's 12-bit prefix) for a code longer
    // than 12 bits; 0 (second level) for a code longer than lmax.
    auto canon = [&](uint32_t W, uint32_t l0, uint32_t lmax, bool first_level) -> uint32_t {
        for (uint32_t lv = l0; lv <= lmax; lv++) {
            const uint32_t v = W >> (32 - lv);
            if (!(v < s_first[lv])) {
                const uint32_t k = s_entry[lv] + v - s_first[lv];
                if (k >= nkeys) return DEC_BAD;
                return (lv << 24) | (FOLD ? (uint32_t)keys[k] : (DEC_KEY | k));
            }
        }
        return first_level ? (DEC_LONG | (W >> (32 - DEC_LUT_BITS))) : 0u;
    };
    for (uint32_t pat = threadIdx.x; pat < (1u << DEC_LUT_BITS); pat += blockDim.x)
        s_lut[pat] = canon(pat << (32 - DEC_LUT_BITS), 1, DEC_LUT_BITS, true);
    // second level: the long codes' 12-bit prefixes are [0, first[12]) (the
    // smallest values of their lengths); B2 more bits index it
    const uint32_t n_long = maxl > (uint32_t)DEC_LUT_BITS ? s_first[DEC_LUT_BITS] : 0u;
    uint32_t b2 = 0;
    if (n_long > 0)
        while (b2 < maxl - DEC_LUT_BITS && (n_long << (b2 + 1)) <= DEC_LUT2) b2++;
    for (uint32_t i = threadIdx.x; i < (b2 ? n_long << b2 : 0u); i += blockDim.x) {
        const uint32_t W = ((i >> b2) << (32 - DEC_LUT_BITS)) | ((i & ((1u << b2) - 1u)) << (32 - DEC_LUT_BITS - b2));
        s_lut2[i] = canon(W, DEC_LUT_BITS + 1, DEC_LUT_BITS + b2, false);
    }
    __syncthreads();
    const DecTables tb{s_lut, s_lut2, s_lim, s_diff, maxl, nkeys, b2};
    const uint32_t w = wave_in_wg();
    const uint32_t l = (uint32_t)lane_id();
    const uint32_t ncells_total = (e5 - e4) / 4;
    const uint32_t* par_nbit = reinterpret_cast<const uint32_t*>(in + e2);
    const uint32_t* par_entry = reinterpret_cast<const uint32_t*>(in + e3);
    const uint32_t* cells = reinterpret_cast<const uint32_t*>(in + e4);
    uint32_t* cst = s_cells[w];
    T* ost = reinterpret_cast<T*>(s_out[w]);
    for (uint32_t p = blockIdx.x * DEC_WAVES + w; p < (uint32_t)pardeg; p += gridDim.x * DEC_WAVES) {
        const uint32_t nbit = par_nbit[p];
        const uint32_t pe = par_entry[p];
        const uint32_t ncell = (uint32_t)(((uint64_t)nbit + 31) / 32);
        const uint64_t o0 = (uint64_t)p * (uint32_t)sublen;
        const uint32_t cnt = (uint32_t)((n - o0) < (uint64_t)sublen ? (n - o0) : (uint64_t)sublen);
        T* dst = out + o0;
        if ((uint64_t)pe + ncell > ncells_total) {
            if (l == 0) *d_status = 1;
            continue;
        }
        if (s_entry[1] == 1 && nbit == 0) {  // one symbol with a zero-length code
            const T k0 = nkeys ? keys[0] : (T)0;
            for (uint32_t j = l; j < cnt; j += 64) dst[j] = k0;
            continue;
        }
        bool ok;
        if (ncell <= DEC_CELLS && (uint32_t)sublen <= DEC_SUBLEN_MAX) {
            wave_lds_sync();  // the previous partition's staging reads are done
            for (uint32_t i = l; i < ncell + 4; i += 64) cst[i] = i < ncell ? cells[pe + i] : 0u;
            wave_lds_sync();
            ok = dec_partition<T, true>(cst, ncell, tb, nbit, cnt, keys, ost, dst);
            wave_lds_sync();
            if (ok) {
                const uint32_t bytes = cnt * (uint32_t)sizeof(T);
                uint32_t done = 0;
                if (((uintptr_t)dst & 15u) == 0) {
                    const uint32_t nv = bytes / 16;
                    for (uint32_t i = l; i < nv; i += 64) reinterpret_cast<uint4*>(dst)[i] = s_out[w][i];
                    done = nv * 16 / (uint32_t)sizeof(T);
                }
                for (uint32_t i = done + l; i < cnt; i += 64) dst[i] = ost[i];
            }
        } else {
            ok = dec_partition<T, false>(cells + pe, ncell, tb, nbit, cnt, keys, ost, dst);
        }
        if (!ok && l == 0) *d_status = 1;
    }
}

struct EncWs {
    size_t pardeg, o_ncell, o_tot, o_nbit, o_scan, o_cells, total;
};
EncWs enc_layout(size_t n, int sl) {
    EncWs w;
    w.pardeg = cdiv(std::max<size_t>(n, 1), (size_t)sl);
    size_t o = 0;
    w.o_ncell = o; o = align_up(o + 4 * w.pardeg, 256);
    w.o_tot = o;   o = align_up(o + 256, 256);  // total cells (u32)
    w.o_nbit = o;  o = align_up(o + 8 * NBIT_SLOTS, 256);  // bit-total accumulators (u64)
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
    const bool book_in_lds = bklen <= 8192;
    const size_t lds = (size_t)4 * CELLS * 4 + (book_in_lds ? 4 * (size_t)bklen : 0);
    const unsigned blocks = (unsigned)cdiv(pardeg, 4);
    if (book_in_lds)
        launch_k("huf_enc_kernel", huf_enc_kernel<T, SL, true>, dim3(blocks), dim3(256), (uint32_t)lds, s, (const T*)in, n,
                 bklen, bk4, rev, rev_bytes, o8, scells, ncell, pardeg);
    else
        launch_k("huf_enc_kernel", huf_enc_kernel<T, SL, false>, dim3(blocks), dim3(256), (uint32_t)lds, s, (const T*)in,
                 n, bklen, bk4, rev, rev_bytes, o8, scells, ncell, pardeg);
    if (int rc = scan_u32(s, ncell, reinterpret_cast<uint32_t*>(o8 + ent3), pardeg, nullptr, tot, w + L.o_scan))
        return rc;
    unsigned long long* nbit_acc = reinterpret_cast<unsigned long long*>(w + L.o_nbit);
    launch_k("huf_pack_kernel", huf_pack_kernel, dim3(blocks), dim3(256), 0, s, scells, ncell, o8, ent2, ent3, ent4,
             pardeg, (uint32_t)CELLS, nbit_acc);
    launch_k("huf_header_kernel", huf_header_kernel, dim3(1), dim3(256), 0, s, o8, n, bklen, (uint32_t)SL, ent1, ent2,
             ent3, ent4, pardeg, tot, out_size, nbit_acc);
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
    const int replicas = bklen <= 1024 ? 8 : (bklen <= 4096 ? 4 : (bklen <= 16384 ? 1 : 0));
    const uint32_t rs = replicas == 8 ? bklen + 1 : bklen;  // replica stride in words (LDS <= 64 KiB)
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(n, 256 * 64), 1024));
    hipStream_t s = S(stream);
#define FZH_HIST(T)                                                                                       \
    if (replicas)                                                                                         \
        launch_k("hist_kernel", hist_kernel<T>, dim3(blocks), dim3(256), (uint32_t)(4 * rs * replicas), s, (const T*)d_in, n, bklen, d_freq, replicas); \
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
    // persistent workgroups: each builds the decode tables once and then
    // takes partitions in turn (the partition count is in the stream; the
    // coarse 768-symbol partitions bound it)
    const size_t pardeg = cdiv(n, SUBLEN);
    hipStream_t s = S(stream);
#define FZH_HUF_DEC(T)                                                                                         \
    {                                                                                                          \
        const unsigned blocks = (unsigned)std::min<size_t>(cdiv(pardeg, dec_waves<T>()), 768);                \
        launch_k("huf_dec_kernel", huf_dec_kernel<T>, dim3(blocks), dim3(64 * dec_waves<T>()), 0, s,          \
                 (const uint8_t*)d_in, in_bytes, (T*)d_out, n, d_status);                                      \
    }
    switch (elem_bytes) {
    case 1: FZH_HUF_DEC(uint8_t) break;
    case 2: FZH_HUF_DEC(uint16_t) break;
    case 4: FZH_HUF_DEC(uint32_t) break;
    default: return fail(FZH_E_INVALID, "huffman dec: elem_bytes must be 1, 2 or 4");
    }
#undef FZH_HUF_DEC
    FZH_CHECK_LAUNCH("huf_dec_kernel");
    return FZH_OK;
}

}  // extern "C"

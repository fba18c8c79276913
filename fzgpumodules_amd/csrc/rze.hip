// rze.hip — Recursive Zero-byte Elimination, 16 KiB chunks, 4 levels (gfx950).
//
// Stream format (modules/coders/rze/rze_stage.cu:774-843, 324-445):
//   [orig_bytes u32][n_chunks u32][size_with_flag u32 x n_chunks][chunks...]
//   zero-padded to a multiple of 4 bytes.  size bit 31 = chunk stored raw.
// Chunk: [nonzero bytes][bm1 non-repeated][bm2 non-repeated][bm3 non-repeated]
//        [bm4 raw 4 B][in_size u16 LE]; an all-zero chunk is the 2-byte tag.
//   bm1 bit i of byte i/8 = (data byte i != 0)            (zero elimination)
//   bm(k+1) bit i = (bm(k) byte i != bm(k) byte i-1)      (repeat elimination,
//                                                          byte -1 := 0)
// The raw-fallback thresholds are the reference's, verbatim (:369-431).
//
// MI355X design
//   encode: one 256-thread workgroup per chunk (fzh_rze_chunk.h).  The
//   chunk is held in registers (16 dwords per thread), bitmaps are built in
//   LDS and the compressed chunk is assembled in LDS, then written to its
//   slot of a chunk-sized scratch (raw chunks are taken from the input by
//   the pack kernel, except bit-shuffled blocks).  A pack kernel
//   places every chunk at its offset, which each workgroup sums directly from
//   the sizes the encode kernel wrote (beyond 4096 chunks a device-wide scan
//   runs first) -- no host round trip and no wave waiting on another (the
//   reference stages per-chunk scratch too, then scans on the host stream
//   with CUB and copies).
//   decode: one 256-thread workgroup per chunk sums the header's sizes
//   before its chunk (beyond 4096 chunks a one-workgroup kernel scans them
//   first) and decodes, with the
//   repeat-elimination decode done as a parallel prefix-popcount gather
//   instead of a single-thread loop.
#include "fzh_concat.h"
#include "fzh_rze_chunk.h"
#include "fzh_scan.h"

namespace fzh {
namespace {

// Streams of up to PACK_INLINE chunks (64 MiB) take their chunk offsets
// inside the pack and decode kernels (each workgroup reads the sizes before
// its chunk, at most 16 KiB, from L2); longer ones scan them in separate
// launches first.
constexpr uint32_t PACK_INLINE = 4096;

// ---------------------------------------------------------------------------
// Encode
// ---------------------------------------------------------------------------
// SEW = 0: plain RZE; SEW = 1/2: the input is first bit-shuffled (16 KiB
// blocks of SEW-byte elements; the first nfull chunks are full blocks, the
// tail is taken verbatim, as BitshuffleStage leaves it).
template <int SEW>
__global__ void __launch_bounds__(256)
rze_enc_kernel(const uint8_t* __restrict__ in, uint64_t nbytes, uint8_t* __restrict__ out,
               uint8_t* __restrict__ scratch, uint32_t* __restrict__ sizes, uint32_t* __restrict__ clean,
               uint32_t nc, uint64_t nfull) {
    __shared__ RzeWgLds L;
    const uint32_t c = xcd_swizzle(blockIdx.x, gridDim.x);
    if (c >= nc) return;
    const uint64_t in_off = (uint64_t)c * CS;
    const int in_size = (int)((nbytes - in_off) < (uint64_t)CS ? (nbytes - in_off) : CS);
    const uint32_t size =
        rze_encode_chunk_wg<SEW>(in + in_off, in_size, SEW != 0 && c < nfull, L, scratch + (uint64_t)c * CS);
    if (threadIdx.x == 0) {
        sizes[c] = size;
        clean[c] = size & 0x7FFFFFFFu;
        reinterpret_cast<uint32_t*>(out + 8)[c] = size;
        if (c == 0) {
            reinterpret_cast<uint32_t*>(out)[0] = (uint32_t)nbytes;
            reinterpret_cast<uint32_t*>(out)[1] = nc;
        }
    }
}

// Places chunk c at [8 + 4 nc + offs[c]) of the stream, one workgroup per
// chunk (offs = exclusive scan of the clean sizes, offs[nc] = their sum;
// with offs null, up to PACK_INLINE chunks, each workgroup sums the clean
// sizes before its chunk itself).  Raw chunks below `nshuf` (bit-shuffled
// blocks) come from the scratch slot, other raw chunks straight from the
// input.
__device__ __forceinline__ void rze_pack_chunk(const RzePackDesc& d, uint32_t c, uint32_t* s_sum) {
    const uint8_t* __restrict__ in = d.in;
    const uint64_t nbytes = d.nbytes;
    const uint8_t* __restrict__ scratch = d.scratch;
    const uint32_t* __restrict__ sizes = d.sizes;
    const uint32_t* __restrict__ cleans = d.cleans;
    const uint32_t* __restrict__ offs = d.offs;
    uint8_t* __restrict__ out = d.out;
    const uint32_t nc = d.nc;
    const uint64_t nshuf = d.nshuf;
    uint32_t off;
    if (offs) {
        off = offs[c];
    } else {
        uint32_t before[1];
        wg256_sums_before<1>(c, [&](int, uint32_t i) { return cleans[i]; }, s_sum, before);
        off = before[0];
    }
    const uint64_t hdr = 8ull + 4ull * nc;
    const uint32_t size = sizes[c];
    const uint32_t clean = size & 0x7FFFFFFFu;
    const bool raw = (size & 0x80000000u) != 0u;
    const uint64_t dst = hdr + off;
    if (raw && c >= nshuf) {
        const uint64_t src = (uint64_t)c * CS;
        copy_to_unaligned16<256>(in + src, nbytes - src, out + dst, clean);
    } else {
        copy_to_unaligned16<256>(scratch + (uint64_t)c * CS, CS, out + dst, clean);
    }
    if (c == nc - 1) {
        const uint64_t end = hdr + (uint32_t)(off + clean);
        const uint64_t padded = (end + 3) & ~3ull;
        if ((uint64_t)threadIdx.x < padded - end) out[end + threadIdx.x] = 0;
        if (threadIdx.x == 0) *d.out_size = padded;
    }
}

__global__ void __launch_bounds__(256) rze_pack_kernel(RzePackDesc d) {
    __shared__ uint32_t s_sum[4];
    const uint32_t c = xcd_swizzle(blockIdx.x, gridDim.x);
    if (c >= d.nc) return;  // uniform over the workgroup
    rze_pack_chunk(d, c, s_sum);
}

// The pack step and the pipeline's concat + readback in one launch:
// workgroups [0, nc) place the chunks (slot 0 of the concat, already in
// place), the rest lay out the other slots and serve the readback, with the
// RZE stream's size summed from the clean sizes rather than read from the
// size word the pack workgroups are still writing.
__global__ void __launch_bounds__(256)
rze_pack_concat_kernel(RzePackDesc d, ConcatArgs a, int n, uint8_t* __restrict__ out, uint64_t* __restrict__ total,
                       ReadbackArgs r, int nr, uint8_t* __restrict__ mapped) {
    __shared__ uint32_t s_sum[4];
    const uint32_t b = blockIdx.x;
    if (b < d.nc) {
        rze_pack_chunk(d, b, s_sum);
        return;
    }
    uint32_t sum;
    if (d.offs) {
        sum = d.offs[d.nc];
    } else {
        uint32_t before[1];
        wg256_sums_before<1>(d.nc, [&](int, uint32_t i) { return d.cleans[i]; }, s_sum, before);
        sum = before[0];
    }
    const uint64_t end = 8ull + 4ull * d.nc + sum;
    const uint64_t size0 = (end + 3) & ~3ull;
    concat_readback_body(a, n, out, total, r, nr, mapped, b - d.nc, gridDim.x - d.nc,
                         SizeOverride{true, size0, reinterpret_cast<uintptr_t>(d.out_size), size0});
}

// ---------------------------------------------------------------------------
// Decode
// ---------------------------------------------------------------------------
// Repeat-elimination decode of `n` bytes: out[i] = data[rank(i)-1], or 0 when
// no bit of bm[0..i] is set (rze_stage.cu:283-298, done with prefix popcounts).

// One workgroup: exclusive scan of the header's chunk sizes into offsets.
__global__ void __launch_bounds__(1024)
rze_offsets_kernel(const uint8_t* __restrict__ in, uint64_t in_bytes, uint32_t* __restrict__ offs, uint32_t nc) {
    __shared__ uint32_t s_wsum[16];
    const int tid = threadIdx.x, w = tid >> 6, l = lane_id();
    const uint64_t hdr = 8ull + 4ull * nc;
    const uint32_t per = (nc + 1023) / 1024;
    const uint32_t a = tid * per;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < per; k++) {
        const uint32_t c = a + k;
        if (c < nc && hdr <= in_bytes) sum += reinterpret_cast<const uint32_t*>(in + 8)[c] & 0x7FFFFFFFu;
    }
    const uint32_t incl = wave_incl_scan(sum);
    if (l == 63) s_wsum[w] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
    for (int j = 0; j < w; j++) run += s_wsum[j];
    for (uint32_t k = 0; k < per; k++) {
        const uint32_t c = a + k;
        if (c >= nc) break;
        offs[c] = run;
        if (hdr <= in_bytes) run += reinterpret_cast<const uint32_t*>(in + 8)[c] & 0x7FFFFFFFu;
    }
}

// Decoder, one 256-thread workgroup per chunk, in the encoder's layout:
// thread t rebuilds bm1 bytes 8t..8t+7 and produces the chunk image's bytes
// [64 t, 64 t + 64).  bm3 is rebuilt by every wave on its own (lanes 0..31,
// no barrier), bm2 byte t by thread t, bm1 bytes 8t..8t+7 by thread t from
// its run of stored bytes with one repeat-expansion v_perm, and the data
// bytes by one block scan and an expansion v_perm per 8 bytes.  Bit-shuffled
// blocks go back through an LDS exchange to the element-group layout and are
// un-shuffled in registers (thread t owns element group t, or groups t and
// 256 + t for 1-byte elements).
template <int SEW>
struct RzeDecLds {
    uint32_t stw[(CS + 64) / 4 + 8];      // the chunk's stored bytes, dword aligned, zero tail
    uint32_t xch[SEW != 0 ? CS / 4 : 4];  // decoded image of a shuffled block (img_swz order)
    alignas(8) uint32_t rep[256][2];      // kRepeat8
    alignas(8) uint32_t exp[256][2];      // kExpand8
    uint32_t wsum[4][2];                  // per-wave sums: stored bm1 bytes, data bytes
    uint32_t osum[4];                     // per-wave sums of the sizes before the chunk
};
constexpr uint32_t STW_LAST = (CS + 64) / 4 + 8 - 3;  // staged reads start at or below this word

// The 8 stream bytes starting at staged byte q (any alignment).
__device__ __forceinline__ void staged_pair(const uint32_t* stw, uint32_t q, uint32_t& lo, uint32_t& hi) {
    uint32_t i = q >> 2;
    i = i < STW_LAST ? i : STW_LAST;
    const uint32_t s = q & 3u;
    const uint32_t a = stw[i], b = stw[i + 1], c = stw[i + 2];
    lo = __builtin_amdgcn_alignbyte(b, a, s);
    hi = __builtin_amdgcn_alignbyte(c, b, s);
}

// stream dword starting at chunk byte q (any alignment) from the staged words
__device__ __forceinline__ uint32_t staged_dword(const uint32_t* stw, uint32_t q) {
    const uint32_t lo = stw[q >> 2], hi = stw[(q >> 2) + 1];
    return (q & 3) ? __builtin_amdgcn_alignbyte(hi, lo, q & 3) : lo;
}

template <int SEW>
__global__ void __launch_bounds__(256)
rze_dec_kernel(const uint8_t* __restrict__ in, uint64_t in_bytes, uint8_t* __restrict__ out,
               uint64_t orig_bytes, int* __restrict__ d_status, const uint32_t* __restrict__ offs, uint32_t nc,
               uint64_t nfull) {
    constexpr int SE = SEW == 0 ? 2 : SEW;
    __shared__ RzeDecLds<SEW> L;
    const uint32_t c = xcd_swizzle(blockIdx.x, gridDim.x);
    if (c >= nc) return;
    const int t = threadIdx.x, w = (int)wave_in_wg(), l = lane_id();
    const uint32_t hdr = 8u + 4u * nc;
    bool bad = in_bytes < hdr;
    uint32_t off_c = 0;
    if (offs) {
        off_c = offs[c];
    } else {  // short header: the workgroup sums the sizes before chunk c itself
        const uint32_t* ent = reinterpret_cast<const uint32_t*>(in + 8);
        uint32_t before[1];
        wg256_sums_before<1>(bad ? 0u : c, [&](int, uint32_t i) { return ent[i] & 0x7FFFFFFFu; }, L.osum, before);
        off_c = before[0];
    }
    uint32_t entry = 0;
    if (!bad) {
        const uint32_t h0 = reinterpret_cast<const uint32_t*>(in)[0];
        const uint32_t h1 = reinterpret_cast<const uint32_t*>(in)[1];
        bad = (uint64_t)h0 != orig_bytes || h1 != nc;
        entry = reinterpret_cast<const uint32_t*>(in + 8)[c];
    }
    const uint32_t stored = bad ? 0u : (entry & 0x7FFFFFFFu);
    const uint64_t src = (uint64_t)hdr + off_c;
    const uint32_t o = (c + 1 < nc) ? (uint32_t)CS : (uint32_t)(orig_bytes - (uint64_t)c * CS);
    const bool israw = (entry & 0x80000000u) != 0u;
    if (!bad) bad = src + stored > in_bytes || stored > (uint32_t)CS || (israw && stored < o) ||
                    (!israw && stored != 2 && stored < 7);
    uint8_t* dst = out + (uint64_t)c * CS;
    if (bad) {  // uniform over the workgroup
        if (t == 0) *d_status = 1;
        return;
    }
    if (!israw && stored == 2) {
        for (uint32_t i = 16 * t; i < o; i += 4096) {
            if (i + 16 <= o) *reinterpret_cast<uint4*>(dst + i) = make_uint4(0, 0, 0, 0);
            else for (uint32_t k = i; k < o; k++) dst[k] = 0;
        }
        return;
    }
    if (!israw) {
        const uint2 a = *reinterpret_cast<const uint2*>(kRepeat8.v[t]);
        const uint2 b = *reinterpret_cast<const uint2*>(kExpand8.v[t]);
        L.rep[t][0] = a.x;
        L.rep[t][1] = a.y;
        L.exp[t][0] = b.x;
        L.exp[t][1] = b.y;
    }
    // -- stage the chunk's bytes into LDS (dword aligned start)
    const uint64_t a0 = src & ~3ull;
    const uint32_t shift = (uint32_t)(src - a0);
    const uint32_t nwd = (uint32_t)((((src + stored + 3) & ~3ull) - a0) / 4);
    for (uint32_t i = t; i < nwd; i += 256) {
        const uint64_t ga = a0 + 4ull * i;
        uint32_t v;
        if (ga + 4 <= in_bytes) {
            v = *reinterpret_cast<const uint32_t*>(in + ga);
        } else {
            v = 0;
            for (uint32_t k = 0; k < 4; k++)
                if (ga + k < in_bytes) v |= (uint32_t)in[ga + k] << (8 * k);
        }
        L.stw[i] = v;
    }
    if (t < 8) L.stw[nwd + t] = 0u;  // the words after the last one the staged reads may touch
    __syncthreads();
    const uint8_t* cs = reinterpret_cast<const uint8_t*>(L.stw) + shift;  // chunk byte 0
    const bool unshuf = SEW != 0 && c < nfull;
    if (israw) {
        if (unshuf) {
            uint32_t d[16];
#pragma unroll
            for (int p = 0; p < 16; p++) d[p] = staged_dword(L.stw, 1024 * p + 4 * t + shift);
            unshuffled_store<SE>(d, dst);
        } else {
            for (uint32_t i = 16 * t; i < o; i += 4096) {
                if (i + 16 <= o) {
                    *reinterpret_cast<uint4*>(dst + i) =
                        make_uint4(staged_dword(L.stw, i + shift), staged_dword(L.stw, i + 4 + shift),
                                   staged_dword(L.stw, i + 8 + shift), staged_dword(L.stw, i + 12 + shift));
                } else {
                    for (uint32_t k = i; k < o; k++) dst[k] = cs[k];
                }
            }
        }
        return;
    }
    const int bm1_bytes = ((int)o + 7) / 8;
    const int bm2_bytes = (bm1_bytes + 7) / 8;
    const int bm3_bytes = (bm2_bytes + 7) / 8;
    const int bm4_start = (int)stored - 6;
    const uint32_t bm4 = (uint32_t)cs[bm4_start] | ((uint32_t)cs[bm4_start + 1] << 8) |
                         ((uint32_t)cs[bm4_start + 2] << 16) | ((uint32_t)cs[bm4_start + 3] << 24);
    const int bm3_start = bm4_start - __popc(bm4);
    // -- bm3, in every wave: lane j < 32 holds byte j = the last stored byte at or before j
    uint32_t b3 = 0;
    if (l < 32) {
        const uint32_t rank = __popc(bm4 & (l >= 31 ? 0xFFFFFFFFu : ((2u << l) - 1u)));
        const int p = bm3_start + (int)rank - 1;
        if (l < bm3_bytes && rank > 0) {
            if (p < 0) bad = true;
            else b3 = cs[p];
        }
    }
    const uint32_t b3_cnt = __popc(b3);
    const uint32_t b3_incl = wave_incl_scan(b3_cnt);
    const int bm2_start = bm3_start - (int)lane63(b3_incl);
    // -- bm2 byte t: its flag is bm3 bit t
    uint32_t b2 = 0;
    {
        const int j = t >> 3;  // 8 w + (l >> 3): a lane below 32 of this wave
        const uint32_t byte_j = (uint32_t)__shfl((int)b3, j & 63, 64);
        const uint32_t before_j = (uint32_t)__shfl((int)(b3_incl - b3_cnt), j & 63, 64);
        const uint32_t rank = before_j + __popc(byte_j & ((2u << (t & 7)) - 1u));
        const int p = bm2_start + (int)rank - 1;
        if (t < bm2_bytes && rank > 0) {
            if (p < 0) bad = true;
            else b2 = cs[p];
        }
    }
    const uint32_t b2_cnt = __popc(b2);
    const uint32_t b2_incl = wave_incl_scan(b2_cnt);
    if (l == 63) L.wsum[w][0] = b2_incl;
    __syncthreads();
    const uint32_t u0 = L.wsum[0][0], u1 = L.wsum[1][0], u2 = L.wsum[2][0], u3 = L.wsum[3][0];
    const int bm1_start = bm2_start - (int)(u0 + u1 + u2 + u3);
    const uint32_t pre2 = (w > 0 ? u0 : 0u) + (w > 1 ? u1 : 0u) + (w > 2 ? u2 : 0u) + b2_incl - b2_cnt;
    // -- bm1 bytes 8t..8t+7: byte k = stored bm1 byte (pre2 + r_k - 1), r_k = set bits of b2 at
    //    positions <= k; the bytes before b2's first set bit repeat the byte before the run
    uint32_t Xlo, Xhi;
    {
        const int q = bm1_start + (int)pre2;
        if ((b2 && q < 0) || (pre2 > 0 && q - 1 < 0)) bad = true;
        const uint32_t qq = q < 0 ? 0u : (uint32_t)q;
        uint32_t plo, phi;
        staged_pair(L.stw, qq + shift, plo, phi);
        const uint32_t prev = (pre2 > 0 && q > 0) ? (uint32_t)cs[q - 1] : 0u;
        const uint2 rs = *reinterpret_cast<const uint2*>(L.rep[b2]);
        Xlo = __builtin_amdgcn_perm(phi, plo, rs.x);
        Xhi = __builtin_amdgcn_perm(phi, plo, rs.y);
        const uint32_t tz = b2 ? (uint32_t)__builtin_ctz(b2) : 8u;  // leading bytes without a stored byte
        const uint64_t lead = tz >= 8 ? ~0ull : ((1ull << (8 * tz)) - 1ull);
        const uint32_t pr = prev * 0x01010101u;
        Xlo |= pr & (uint32_t)lead;
        Xhi |= pr & (uint32_t)(lead >> 32);
        const int nb = bm1_bytes - 8 * t;  // bytes of bm1 this thread covers
        if (nb < 8) {
            const uint64_t keep = nb <= 0 ? 0ull : ((1ull << (8 * nb)) - 1ull);
            Xlo &= (uint32_t)keep;
            Xhi &= (uint32_t)(keep >> 32);
        }
    }
    // -- data: one scan places this thread's nonzero bytes
    const uint32_t cnt = __popc(Xlo) + __popc(Xhi);
    const uint32_t cnt_incl = wave_incl_scan(cnt);
    if (l == 63) L.wsum[w][1] = cnt_incl;
    __syncthreads();
    const uint32_t v0 = L.wsum[0][1], v1 = L.wsum[1][1], v2 = L.wsum[2][1], v3 = L.wsum[3][1];
    if ((int)(v0 + v1 + v2 + v3) > bm1_start) bad = true;
    uint32_t D[16];
    {
        uint32_t p = (w > 0 ? v0 : 0u) + (w > 1 ? v1 : 0u) + (w > 2 ? v2 : 0u) + cnt_incl - cnt;
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint32_t m = ((b < 4 ? Xlo : Xhi) >> (8 * (b & 3))) & 0xFFu;
            if (m) {
                uint32_t lo, hi;
                staged_pair(L.stw, p + shift, lo, hi);
                const uint2 es = *reinterpret_cast<const uint2*>(L.exp[m]);
                D[2 * b] = __builtin_amdgcn_perm(hi, lo, es.x);
                D[2 * b + 1] = __builtin_amdgcn_perm(hi, lo, es.y);
            } else {
                D[2 * b] = 0u;
                D[2 * b + 1] = 0u;
            }
            p += __popc(m);
        }
    }
    if (unshuf) {
#pragma unroll
        for (int q = 0; q < 4; q++)
            *reinterpret_cast<uint4*>(&L.xch[img_swz(16 * t + 4 * q)]) =
                make_uint4(D[4 * q], D[4 * q + 1], D[4 * q + 2], D[4 * q + 3]);
        __syncthreads();
        uint32_t d[16];
#pragma unroll
        for (int p = 0; p < 16; p++) d[p] = L.xch[img_swz(256 * p + t)];
        unshuffled_store<SE>(d, dst);
    } else {
        const uint32_t b0 = 64u * t;
        if (b0 + 64 <= o) {
#pragma unroll
            for (int q = 0; q < 4; q++)
                st16(dst + b0 + 16 * q, make_uint4(D[4 * q], D[4 * q + 1], D[4 * q + 2], D[4 * q + 3]));
        } else {
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const uint32_t b = b0 + 4 * j;
                if (b + 4 <= o) *reinterpret_cast<uint32_t*>(dst + b) = D[j];
                else for (uint32_t k = b; k < o; k++) dst[k] = (uint8_t)(D[j] >> (8 * (k - b)));
            }
        }
    }
    if (bad) *d_status = 1;
}

// Encode workspace: flagged sizes, clean sizes, offsets (nc + 1), the scan
// workspace, then one CS-byte scratch slot per chunk.
struct EncWs {
    size_t o_sizes, o_clean, o_offs, o_scan, o_scratch, total;
};

EncWs enc_layout(size_t nbytes) {
    const size_t nc = std::max<size_t>(cdiv(nbytes, CS), 1);
    EncWs w;
    size_t o = 0;
    w.o_sizes = o;   o = align_up(o + 4 * nc, 256);
    w.o_clean = o;   o = align_up(o + 4 * nc, 256);
    w.o_offs = o;    o = align_up(o + 4 * (nc + 1), 256);
    w.o_scan = o;    o = align_up(o + scan_ws_bytes(nc), 256);
    w.o_scratch = o; o += nc * CS;
    w.total = o;
    return w;
}

}  // namespace
}  // namespace fzh

using namespace fzh;

extern "C" {

size_t fzh_rze_enc_bound(size_t nbytes) {
    const size_t nc = cdiv(nbytes, CS);
    // A partial last chunk may grow by up to ~1/7 (the thresholds compare
    // against the full 16 KiB chunk), plus the header and 4-byte padding.
    return nbytes + 8 + 4 * nc + (nbytes % CS) / 7 + 64;
}

size_t fzh_rze_enc_ws_bytes(size_t nbytes) { return enc_layout(nbytes).total; }

}  // extern "C"

namespace fzh {
namespace {

}  // namespace

int rze_enc_prepare(hipStream_t s, const void* d_in, size_t nbytes, int sew, void* d_out, size_t out_cap,
                    uint64_t* d_out_size, void* d_ws, size_t ws_bytes, RzePackDesc* desc) {
    if (nbytes == 0) return fail(FZH_E_INVALID, "RZE: empty input");
    if (nbytes > 0xFFFFFFFFull) return fail(FZH_E_INVALID, "RZE: input larger than 4 GiB");
    if (out_cap < fzh_rze_enc_bound(nbytes)) return fail(FZH_E_CAPACITY, "RZE: output buffer below fzh_rze_enc_bound()");
    if (ws_bytes < fzh_rze_enc_ws_bytes(nbytes)) return fail(FZH_E_CAPACITY, "RZE: workspace too small");
    if (((uintptr_t)d_in | (uintptr_t)d_out) % 16 != 0) return fail(FZH_E_INVALID, "RZE: buffers must be 16-byte aligned");
    const uint32_t nc = (uint32_t)cdiv(nbytes, CS);
    const uint64_t nfull = nbytes / CS;
    const EncWs L = enc_layout(nbytes);
    uint8_t* w = (uint8_t*)d_ws;
    uint32_t* sizes = (uint32_t*)(w + L.o_sizes);
    uint32_t* clean = (uint32_t*)(w + L.o_clean);
    uint32_t* offs = (uint32_t*)(w + L.o_offs);
    uint8_t* scratch = w + L.o_scratch;
    const uint8_t* in = (const uint8_t*)d_in;
    uint8_t* out = (uint8_t*)d_out;
    auto* ek = sew == 1 ? rze_enc_kernel<1> : sew == 2 ? rze_enc_kernel<2> : rze_enc_kernel<0>;
    launch_k("rze_enc_kernel", ek, dim3(nc), dim3(256), 0, s, in, nbytes, out, scratch, sizes, clean, nc,
             sew ? nfull : (uint64_t)0);
    FZH_CHECK_LAUNCH("rze_enc_kernel");
    const bool inline_offs = nc <= PACK_INLINE;
    if (!inline_offs)
        if (int rc = scan_u32(s, clean, offs, nc, nullptr, offs + nc, w + L.o_scan)) return rc;
    *desc = RzePackDesc{in,    nbytes, scratch, sizes, clean, inline_offs ? nullptr : offs,
                        out,   d_out_size, nc, sew ? nfull : (uint64_t)0};
    return FZH_OK;
}

int rze_pack_concat_readback(hipStream_t s, const RzePackDesc& d, const fzh_concat_slot* slots, int n, void* d_out,
                             uint64_t out_cap, uint64_t* d_total, const fzh_readback_entry* entries, int n_entries,
                             void* d_mapped_dst) {
    ConcatArgs a;
    ReadbackArgs r;
    unsigned grid = 1;
    if (int rc = build_concat_args(slots, n, d_out, out_cap, d_total, entries, n_entries, d_mapped_dst, a, r, grid))
        return rc;
    if (n < 1 || slots[0].src != d.out)
        return fail(FZH_E_INVALID, "RZE pack + concat: slot 0 must be the RZE stream");
    launch_k("rze_pack_concat_kernel", rze_pack_concat_kernel, dim3(d.nc + grid), dim3(256), 0, s, d, a, n,
             (uint8_t*)d_out, d_total, r, n_entries, (uint8_t*)d_mapped_dst);
    FZH_CHECK_LAUNCH("rze_pack_concat_kernel");
    return FZH_OK;
}

namespace {

int rze_enc_impl(hipStream_t s, const void* d_in, size_t nbytes, int sew, void* d_out, size_t out_cap,
                 uint64_t* d_out_size, void* d_ws, size_t ws_bytes) {
    RzePackDesc d;
    if (int rc = rze_enc_prepare(s, d_in, nbytes, sew, d_out, out_cap, d_out_size, d_ws, ws_bytes, &d)) return rc;
    launch_k("rze_pack_kernel", rze_pack_kernel, dim3(d.nc), dim3(256), 0, s, d);
    FZH_CHECK_LAUNCH("rze_pack_kernel");
    return FZH_OK;
}

int rze_dec_impl(hipStream_t s, const void* d_in, size_t in_bytes, void* d_out, size_t orig_bytes, int sew,
                 int* d_status, void* d_ws, size_t ws_bytes) {
    if (orig_bytes == 0) return FZH_OK;
    if (orig_bytes > 0xFFFFFFFFull) return fail(FZH_E_INVALID, "RZE: output larger than 4 GiB");
    if (ws_bytes < fzh_rze_dec_ws_bytes(orig_bytes)) return fail(FZH_E_CAPACITY, "RZE: workspace too small");
    if (((uintptr_t)d_in % 4) != 0 || ((uintptr_t)d_out % 16) != 0)
        return fail(FZH_E_INVALID, "RZE: input must be 4-byte and output 16-byte aligned");
    if (!d_status) return fail(FZH_E_INVALID, "RZE: null status pointer");
    const uint32_t nc = (uint32_t)cdiv(orig_bytes, CS);
    const uint64_t nfull = orig_bytes / CS;
    uint32_t* offs = nc <= PACK_INLINE ? nullptr : (uint32_t*)d_ws;
    const uint8_t* in = (const uint8_t*)d_in;
    uint8_t* out = (uint8_t*)d_out;
    if (offs) launch_k("rze_offsets_kernel", rze_offsets_kernel, dim3(1), dim3(1024), 0, s, in, in_bytes, offs, nc);
    auto* dk = sew == 1 ? rze_dec_kernel<1> : sew == 2 ? rze_dec_kernel<2> : rze_dec_kernel<0>;
    launch_k("rze_dec_kernel", dk, dim3(nc), dim3(256), 0, s, in, in_bytes, out, orig_bytes, d_status, offs, nc,
             sew ? nfull : (uint64_t)0);
    FZH_CHECK_LAUNCH("rze_dec_kernel");
    return FZH_OK;
}

}  // namespace
}  // namespace fzh

extern "C" {

int fzh_rze_enc(void* stream, const void* d_in, size_t nbytes, void* d_out, size_t out_cap, uint64_t* d_out_size,
                void* d_ws, size_t ws_bytes) {
    return rze_enc_impl(S(stream), d_in, nbytes, 0, d_out, out_cap, d_out_size, d_ws, ws_bytes);
}

int fzh_rze_enc_shuffled(void* stream, const void* d_in, size_t nbytes, int elem_width, void* d_out,
                         size_t out_cap, uint64_t* d_out_size, void* d_ws, size_t ws_bytes) {
    if (elem_width != 1 && elem_width != 2)
        return fail(FZH_E_INVALID, "RZE (shuffled): element width must be 1 or 2");
    return rze_enc_impl(S(stream), d_in, nbytes, elem_width, d_out, out_cap, d_out_size, d_ws, ws_bytes);
}

size_t fzh_rze_dec_ws_bytes(size_t orig_bytes) { return align_up(4 * std::max<size_t>(cdiv(orig_bytes, CS), 1), 256); }

int fzh_rze_dec(void* stream, const void* d_in, size_t in_bytes, void* d_out, size_t orig_bytes, int* d_status,
                void* d_ws, size_t ws_bytes) {
    return rze_dec_impl(S(stream), d_in, in_bytes, d_out, orig_bytes, 0, d_status, d_ws, ws_bytes);
}

int fzh_rze_dec_shuffled(void* stream, const void* d_in, size_t in_bytes, void* d_out, size_t orig_bytes,
                         int elem_width, int* d_status, void* d_ws, size_t ws_bytes) {
    if (elem_width != 1 && elem_width != 2)
        return fail(FZH_E_INVALID, "RZE (shuffled): element width must be 1 or 2");
    return rze_dec_impl(S(stream), d_in, in_bytes, d_out, orig_bytes, elem_width, d_status, d_ws, ws_bytes);
}

}  // extern "C"

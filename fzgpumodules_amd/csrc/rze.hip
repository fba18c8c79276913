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
//   the sizes the encode kernel wrote -- no host round trip and no wave
//   waiting on another (the reference stages per-chunk scratch too, then
//   scans on the host stream with CUB and copies).
//   decode: a one-workgroup kernel scans the header's sizes into chunk
//   offsets; then one wave64 per chunk decodes, with the repeat-elimination
//   decode done as a parallel prefix-popcount gather instead of a
//   single-thread loop.
#include "fzh_rze_chunk.h"
#include "fzh_scan.h"

namespace fzh {
namespace {

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

// Places chunk c at [8 + 4 nc + offs[c]) of the stream, one wave per chunk
// (offs = exclusive scan of the clean sizes, offs[nc] = their sum).  Raw
// chunks below `nshuf` (bit-shuffled blocks) come from the scratch slot,
// other raw chunks straight from the input.
__global__ void __launch_bounds__(256)
rze_pack_kernel(const uint8_t* __restrict__ in, uint64_t nbytes, const uint8_t* __restrict__ scratch,
                const uint32_t* __restrict__ sizes, const uint32_t* __restrict__ offs, uint8_t* __restrict__ out,
                uint64_t* __restrict__ out_size, uint32_t nc, uint64_t nshuf) {
    const uint32_t c = xcd_swizzle(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (c >= nc) return;
    const int l = lane_id();
    const uint64_t hdr = 8ull + 4ull * nc;
    const uint32_t size = sizes[c];
    const uint32_t clean = size & 0x7FFFFFFFu;
    const bool raw = (size & 0x80000000u) != 0u;
    const uint64_t dst = hdr + offs[c];
    if (raw && c >= nshuf) {
        const uint64_t src = (uint64_t)c * CS;
        copy_to_unaligned16(in + src, nbytes - src, out + dst, clean);
    } else {
        copy_to_unaligned16(scratch + (uint64_t)c * CS, CS, out + dst, clean);
    }
    if (c == nc - 1) {
        const uint64_t end = hdr + offs[nc];
        const uint64_t padded = (end + 3) & ~3ull;
        if ((uint64_t)l < padded - end) out[end + l] = 0;
        if (l == 0) *out_size = padded;
    }
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

template <int SEW>
__global__ void __launch_bounds__(64 * WAVES)
rze_dec_kernel(const uint8_t* __restrict__ in, uint64_t in_bytes, uint8_t* __restrict__ out,
               uint64_t orig_bytes, int* __restrict__ d_status, const uint32_t* __restrict__ offs, uint32_t nc,
               uint64_t nfull) {
    constexpr int SE = SEW == 0 ? 2 : SEW;
    extern __shared__ __align__(16) uint8_t smem[];
    const int l = lane_id();
    const int wv = threadIdx.x >> 6;
    uint8_t* st = smem + wv * LDS_PER_WAVE;
    uint8_t* bm1 = st + LDS_STREAM;

    const uint32_t c = blockIdx.x * WAVES + wv;
    if (c < nc) {
        const uint32_t hdr = 8u + 4u * nc;
        bool bad = in_bytes < hdr;
        uint32_t entry = 0;
        if (!bad) {
            const uint32_t h0 = reinterpret_cast<const uint32_t*>(in)[0];
            const uint32_t h1 = reinterpret_cast<const uint32_t*>(in)[1];
            bad = (uint64_t)h0 != orig_bytes || h1 != nc;
            entry = reinterpret_cast<const uint32_t*>(in + 8)[c];
        }
        const uint32_t stored = bad ? 0u : (entry & 0x7FFFFFFFu);
        const uint32_t off = offs[c];
        const uint64_t src = (uint64_t)hdr + off;
        const uint32_t o = (c + 1 < nc) ? (uint32_t)CS : (uint32_t)(orig_bytes - (uint64_t)c * CS);
        const bool israw = (entry & 0x80000000u) != 0u;
        if (!bad) bad = src + stored > in_bytes || stored > (uint32_t)CS || (israw && stored < o) ||
                        (!israw && stored != 2 && stored < 7);
        uint8_t* dst = out + (uint64_t)c * CS;
        if (bad) {
            if (l == 0) *d_status = 1;
        } else if (!israw && stored == 2) {
            for (uint32_t i = 16 * l; i < o; i += 1024) {
                if (i + 16 <= o) *reinterpret_cast<uint4*>(dst + i) = make_uint4(0, 0, 0, 0);
                else for (uint32_t k = i; k < o; k++) dst[k] = 0;
            }
        } else {
            // -- stage the chunk's bytes into LDS (dword aligned start)
            const uint64_t a0 = src & ~3ull;
            const uint32_t shift = (uint32_t)(src - a0);
            const uint32_t span = (uint32_t)(((src + stored + 3) & ~3ull) - a0);  // bytes to fetch
            const uint32_t nwd = span / 4;
            uint32_t* stw = reinterpret_cast<uint32_t*>(st);
            for (uint32_t i = l; i < nwd; i += 64) {
                const uint64_t ga = a0 + 4ull * i;
                uint32_t w;
                if (ga + 4 <= in_bytes) {
                    w = *reinterpret_cast<const uint32_t*>(in + ga);
                } else {
                    w = 0;
                    for (uint32_t k = 0; k < 4; k++)
                        if (ga + k < in_bytes) w |= (uint32_t)in[ga + k] << (8 * k);
                }
                stw[i] = w;
            }
            wave_lds_sync();
            const uint8_t* cs = st + shift;  // chunk byte 0
            if (israw && SEW != 0 && c < nfull) {
                uint32_t vv[16][4];
#pragma unroll
                for (int j = 0; j < 16; j++)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const uint32_t p = 1024 * j + 16 * l + 4 * q + shift;
                        const uint32_t lo = stw[p >> 2], hi = stw[(p >> 2) + 1];
                        vv[j][q] = (p & 3) ? __builtin_amdgcn_alignbyte(hi, lo, p & 3) : lo;
                    }
                unshuffled_store<SE>(vv, dst);
            } else if (israw) {
                // raw copy: cs[0 .. o) -> dst (dst 16-byte aligned)
                for (uint32_t i = 16 * l; i < o; i += 1024) {
                    uint32_t w4[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const uint32_t p = i + 4 * q + shift;
                        const uint32_t lo = stw[p >> 2], hi = stw[(p >> 2) + 1];
                        w4[q] = (p & 3) ? __builtin_amdgcn_alignbyte(hi, lo, p & 3) : lo;
                    }
                    if (i + 16 <= o) *reinterpret_cast<uint4*>(dst + i) = make_uint4(w4[0], w4[1], w4[2], w4[3]);
                    else for (uint32_t k = i; k < o; k++) dst[k] = cs[k];
                }
            } else {
                const int bm1_bytes = ((int)o + 7) / 8;
                const int bm2_bytes = (bm1_bytes + 7) / 8;
                const int bm3_bytes = (bm2_bytes + 7) / 8;
                const int bm4_start = (int)stored - 6;
                const uint32_t bm4 = (uint32_t)cs[bm4_start] | ((uint32_t)cs[bm4_start + 1] << 8) |
                                     ((uint32_t)cs[bm4_start + 2] << 16) | ((uint32_t)cs[bm4_start + 3] << 24);
                const int bm3_start = bm4_start - __popc(bm4);
                // bm3: lanes 0..31, byte l
                uint32_t b3 = 0;
                {
                    const uint32_t rank = __popc(bm4 & (l >= 31 ? 0xFFFFFFFFu : ((2u << l) - 1u)));
                    const int p = bm3_start + (int)rank - 1;
                    b3 = (l < bm3_bytes && rank > 0 && p >= 0) ? cs[p] : 0u;
                    if (l < bm3_bytes && rank > 0 && p < 0) bad = true;
                }
                const uint32_t bm3_nr = wave_sum(l < 32 ? __popc(b3) : 0u);
                const int bm2_start = bm3_start - (int)bm3_nr;
                // bm2: lane l owns bytes 4l..4l+3; bits 4l..4l+3 of bm3 are its nibble
                uint32_t b2w = 0;
                {
                    const uint32_t b3byte = __shfl(b3, (l >> 1) & 63, 64);
                    const uint32_t nib = (b3byte >> (4 * (l & 1))) & 0xFu;
                    const uint32_t cnt = __popc(nib);
                    const uint32_t excl = wave_incl_scan(cnt) - cnt;
                    uint32_t r = excl;
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        if (nib & (1u << k)) r++;
                        const int p = bm2_start + (int)r - 1;
                        uint32_t bv = 0;
                        if (4 * l + k < bm2_bytes && r > 0) {
                            if (p < 0) bad = true; else bv = cs[p];
                        }
                        b2w |= bv << (8 * k);
                    }
                }
                const uint32_t bm2_nr = wave_sum(__popc(b2w));
                const int bm1_start = bm2_start - (int)bm2_nr;
                // bm1: lane l owns bytes 32l..32l+31; bits = bm2 bytes 4l..4l+3 = b2w
                {
                    const uint32_t cnt = __popc(b2w);
                    const uint32_t excl = wave_incl_scan(cnt) - cnt;
                    uint32_t r = excl;
                    uint32_t wds[8];
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        uint32_t wq = 0;
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            const int bi = 4 * q + k;
                            if (b2w & (1u << bi)) r++;
                            const int p = bm1_start + (int)r - 1;
                            uint32_t bv = 0;
                            if (32 * l + bi < bm1_bytes && r > 0) {
                                if (p < 0) bad = true; else bv = cs[p];
                            }
                            wq |= bv << (8 * k);
                        }
                        wds[q] = wq;
                    }
                    *reinterpret_cast<uint4*>(bm1 + 32 * l) = make_uint4(wds[0], wds[1], wds[2], wds[3]);
                    *reinterpret_cast<uint4*>(bm1 + 32 * l + 16) = make_uint4(wds[4], wds[5], wds[6], wds[7]);
                }
                const uint32_t data_nz = wave_sum(__popc(*reinterpret_cast<const uint32_t*>(bm1 + 32 * l)) +
                                                  __popc(*reinterpret_cast<const uint32_t*>(bm1 + 32 * l + 4)) +
                                                  __popc(*reinterpret_cast<const uint32_t*>(bm1 + 32 * l + 8)) +
                                                  __popc(*reinterpret_cast<const uint32_t*>(bm1 + 32 * l + 12)) +
                                                  __popc(*reinterpret_cast<const uint32_t*>(bm1 + 32 * l + 16)) +
                                                  __popc(*reinterpret_cast<const uint32_t*>(bm1 + 32 * l + 20)) +
                                                  __popc(*reinterpret_cast<const uint32_t*>(bm1 + 32 * l + 24)) +
                                                  __popc(*reinterpret_cast<const uint32_t*>(bm1 + 32 * l + 28)));
                if ((int)data_nz > bm1_start) bad = true;
                wave_lds_sync();
                // zero-elimination decode: lane l, step j -> output bytes [1024 j + 16 l, +16)
                const bool unshuf = SEW != 0 && c < nfull;
                uint32_t vv[16][4];
                uint32_t run = 0;
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    const uint32_t mm = *reinterpret_cast<const uint16_t*>(bm1 + 128 * j + 2 * l);
                    const uint32_t cnt = __popc(mm);
                    const uint32_t incl = wave_incl_scan(cnt);
                    uint32_t r = run + incl - cnt;
                    uint32_t w4[4] = {0, 0, 0, 0};
#pragma unroll
                    for (int k = 0; k < 16; k++) {
                        if (mm & (1u << k)) {
                            const uint32_t bv = (r < data_nz && (int)r < bm1_start) ? cs[r] : 0u;
                            w4[k >> 2] |= bv << (8 * (k & 3));
                            r++;
                        }
                    }
                    run += __shfl(incl, 63, 64);
                    const uint32_t p = 1024 * j + 16 * l;
                    if (unshuf) {
#pragma unroll
                        for (int q = 0; q < 4; q++) vv[j][q] = w4[q];
                    } else if (p + 16 <= o) {
                        *reinterpret_cast<uint4*>(dst + p) = make_uint4(w4[0], w4[1], w4[2], w4[3]);
                    } else {
                        for (uint32_t k = p; k < o; k++) dst[k] = (uint8_t)(w4[(k - p) >> 2] >> (8 * ((k - p) & 3)));
                    }
                }
                if (unshuf) unshuffled_store<SE>(vv, dst);
                if (__ballot(bad) != 0ull && l == 0) *d_status = 1;
            }
        }
    }
}

constexpr size_t kLds = (size_t)WAVES * LDS_PER_WAVE;

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

int rze_enc_impl(hipStream_t s, const void* d_in, size_t nbytes, int sew, void* d_out, size_t out_cap,
                 uint64_t* d_out_size, void* d_ws, size_t ws_bytes) {
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
    if (sew == 1)
        rze_enc_kernel<1><<<nc, 256, 0, s>>>(in, nbytes, out, scratch, sizes, clean, nc, nfull);
    else if (sew == 2)
        rze_enc_kernel<2><<<nc, 256, 0, s>>>(in, nbytes, out, scratch, sizes, clean, nc, nfull);
    else
        rze_enc_kernel<0><<<nc, 256, 0, s>>>(in, nbytes, out, scratch, sizes, clean, nc, 0);
    FZH_CHECK_LAUNCH("rze_enc_kernel");
    if (int rc = scan_u32(s, clean, offs, nc, nullptr, offs + nc, w + L.o_scan)) return rc;
    rze_pack_kernel<<<(unsigned)cdiv(nc, 4), 256, 0, s>>>(in, nbytes, scratch, sizes, offs, out, d_out_size, nc,
                                                          sew ? nfull : 0);
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
    const unsigned blocks = (unsigned)cdiv(nc, WAVES);
    uint32_t* offs = (uint32_t*)d_ws;
    const uint8_t* in = (const uint8_t*)d_in;
    uint8_t* out = (uint8_t*)d_out;
    rze_offsets_kernel<<<1, 1024, 0, s>>>(in, in_bytes, offs, nc);
    if (sew == 1)
        rze_dec_kernel<1><<<blocks, 64 * WAVES, kLds, s>>>(in, in_bytes, out, orig_bytes, d_status, offs, nc, nfull);
    else if (sew == 2)
        rze_dec_kernel<2><<<blocks, 64 * WAVES, kLds, s>>>(in, in_bytes, out, orig_bytes, d_status, offs, nc, nfull);
    else
        rze_dec_kernel<0><<<blocks, 64 * WAVES, kLds, s>>>(in, in_bytes, out, orig_bytes, d_status, offs, nc, 0);
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

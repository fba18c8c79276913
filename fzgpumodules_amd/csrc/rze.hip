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
//   encode: one wave64 per chunk.  The chunk is held in registers (16 x 16 B
//   per lane), bitmaps are built with per-lane masks in LDS and the
//   compressed chunk is assembled in LDS, then written to its slot of a
//   chunk-sized scratch (raw chunks are not copied at all).  A pack kernel
//   places every chunk at its offset, which each workgroup sums directly from
//   the sizes the encode kernel wrote -- no host round trip and no wave
//   waiting on another (the reference stages per-chunk scratch too, then
//   scans on the host stream with CUB and copies).
//   decode: a one-workgroup kernel scans the header's sizes into chunk
//   offsets; then one wave64 per chunk decodes, with the repeat-elimination
//   decode done as a parallel prefix-popcount gather instead of a
//   single-thread loop.
#include "fzh_bitplane.h"

namespace fzh {
namespace {

constexpr int CS = 16384;
constexpr int WAVES = 2;
// LDS per wave: stream/staging buffer (+ slack for the aligned dword reads
// that straddle its end) and bm1.
constexpr int LDS_STREAM = CS + 64;
constexpr int LDS_BM1 = CS / 8;
constexpr int LDS_PER_WAVE = LDS_STREAM + LDS_BM1;

// Fused Bitshuffle (16 KiB blocks, element width SEW = 1 or 2, ballot
// layout): a full chunk's bit-plane image in the register layout the RZE
// kernels use (lane l, step j: bytes [1024 j + 16 l, +16) of the shuffled
// block).  Shuffled word index = plane * npp + group with npp = 8192/SEW/32;
// lane l step j holds words 256 j + 4 l + i, i = 0..3, so each lane
// transposes its own groups: SEW 2 -> groups 4l+i (plane j);
// SEW 1 -> groups 256 (j&1) + 4l + i (plane j/2).
template <int SEW>
__device__ __forceinline__ void shuffled_load(const uint8_t* blk, uint4 (&v)[16]) {
    const int l = lane_id();
    uint32_t vv[16][4];
    if constexpr (SEW == 2) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint8_t* g = blk + 64 * (4 * l + i);
            uint32_t w[16];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint4 t = ld16(g + 16 * q);
                w[4 * q] = t.x; w[4 * q + 1] = t.y; w[4 * q + 2] = t.z; w[4 * q + 3] = t.w;
            }
            uint32_t A[16];
            planes16(w, A);
#pragma unroll
            for (int j = 0; j < 16; j++) vv[j][i] = A[15 - j];
        }
    } else {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint8_t* g = blk + 32 * (256 * h + 4 * l + i);
                const uint4 t0 = ld16(g), t1 = ld16(g + 16);
                const uint32_t w[8] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w};
                uint32_t A[8];
                planes8(w, A);
#pragma unroll
                for (int P = 0; P < 8; P++) vv[2 * P + h][i] = A[7 - P];
            }
    }
#pragma unroll
    for (int j = 0; j < 16; j++) v[j] = make_uint4(vv[j][0], vv[j][1], vv[j][2], vv[j][3]);
}

// Inverse of shuffled_load: vv[j][i] = shuffled words in the register layout;
// writes the chunk's unshuffled 16 KiB to blk.
template <int SEW>
__device__ __forceinline__ void unshuffled_store(const uint32_t (&vv)[16][4], uint8_t* blk) {
    const int l = lane_id();
    if constexpr (SEW == 2) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            uint32_t A[16], w[16];
#pragma unroll
            for (int b = 0; b < 16; b++) A[b] = vv[15 - b][i];
            unplanes16(A, w);
            uint8_t* g = blk + 64 * (4 * l + i);
#pragma unroll
            for (int q = 0; q < 4; q++) st16(g + 16 * q, make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]));
        }
    } else {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                uint32_t A[8], w[8];
#pragma unroll
                for (int P = 0; P < 8; P++) A[7 - P] = vv[2 * P + h][i];
                unplanes8(A, w);
                uint8_t* g = blk + 32 * (256 * h + 4 * l + i);
                st16(g, make_uint4(w[0], w[1], w[2], w[3]));
                st16(g + 16, make_uint4(w[4], w[5], w[6], w[7]));
            }
    }
}

__device__ __forceinline__ uint32_t byte_of(const uint4& v, int k) {
    const uint32_t w = k < 4 ? v.x : k < 8 ? v.y : k < 12 ? v.z : v.w;
    return (w >> (8 * (k & 3))) & 0xFFu;
}

// Copy `size` bytes from LDS buffer `lds` (dword aligned) to global `dst`
// (any alignment), touching no byte outside [dst, dst+size).
__device__ __forceinline__ void lds_to_global(const uint8_t* lds, uint8_t* dst, uint32_t size) {
    if (size == 0) return;
    const int l = lane_id();
    const uintptr_t a0 = (uintptr_t)dst;
    const uint32_t head = (uint32_t)((4 - (a0 & 3)) & 3);  // bytes before the first aligned dword
    const uint32_t hb = head < size ? head : size;
    if ((uint32_t)l < hb) dst[l] = lds[l];
    if (size <= hb) return;
    uint32_t* dw = reinterpret_cast<uint32_t*>(dst + hb);
    const uint32_t nd = (size - hb) / 4;
    const uint32_t sh = hb & 3;  // stream position hb is at lds byte offset hb
    const uint32_t* lw = reinterpret_cast<const uint32_t*>(lds);
    for (uint32_t i = l; i < nd; i += 64) {
        const uint32_t p = hb + 4 * i;  // stream byte position of this dword
        const uint32_t q = p >> 2;
        uint32_t v;
        if (sh == 0) {
            v = lw[q];
        } else {
            const uint32_t lo = lw[q], hi = lw[q + 1];
            v = __builtin_amdgcn_alignbyte(hi, lo, sh);
        }
        dw[i] = v;
    }
    const uint32_t done = hb + 4 * nd;
    if ((uint32_t)l < size - done) dst[done + l] = lds[done + l];
}

// ---------------------------------------------------------------------------
// Encode
// ---------------------------------------------------------------------------
// SEW = 0: plain RZE; SEW = 1/2: the input is first bit-shuffled (16 KiB
// blocks of SEW-byte elements; the first nfull chunks are full blocks, the
// tail is taken verbatim, as BitshuffleStage leaves it).
template <int SEW>
__global__ void __launch_bounds__(64 * WAVES)
rze_enc_kernel(const uint8_t* __restrict__ in, uint64_t nbytes, uint8_t* __restrict__ out,
               uint8_t* __restrict__ scratch, uint32_t* __restrict__ sizes, uint32_t nc, uint64_t nfull) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int l = lane_id();
    const int wv = threadIdx.x >> 6;
    uint8_t* st = smem + wv * LDS_PER_WAVE;  // assembled chunk stream
    uint8_t* bm1 = st + LDS_STREAM;

    const uint32_t c = blockIdx.x * WAVES + wv;
    if (c < nc) {
        const uint64_t in_off = (uint64_t)c * CS;
        const int in_size = (int)((nbytes - in_off) < (uint64_t)CS ? (nbytes - in_off) : CS);
        const bool full = in_size == CS;

        // -- load the chunk: lane l, step j holds bytes [1024 j + 16 l, +16)
        uint4 v[16];
        if (SEW != 0 && c < nfull) {
            shuffled_load<SEW == 0 ? 2 : SEW>(in + in_off, v);
        } else {
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int p = 1024 * j + 16 * l;
            if (full || p + 16 <= in_size) {
                v[j] = *reinterpret_cast<const uint4*>(in + in_off + p);
            } else {
                uint32_t wds[4] = {0, 0, 0, 0};
                for (int k = 0; k < 16; k++)
                    if (p + k < in_size) wds[k >> 2] |= (uint32_t)in[in_off + p + k] << (8 * (k & 3));
                v[j] = make_uint4(wds[0], wds[1], wds[2], wds[3]);
            }
        }
        }
        // -- level 1: nonzero masks (16 bits per lane per step) -> bm1 in LDS
        uint32_t m[16];
        uint32_t my_nz = 0;
#pragma unroll
        for (int j = 0; j < 16; j++) {
            uint32_t mm = 0;
#pragma unroll
            for (int k = 0; k < 16; k++) mm |= (byte_of(v[j], k) != 0u ? 1u : 0u) << k;
            m[j] = mm;
            my_nz += __popc(mm);
            *reinterpret_cast<uint16_t*>(bm1 + 128 * j + 2 * l) = (uint16_t)mm;
        }
        const uint32_t data_nz = wave_sum(my_nz);
        wave_lds_sync();

        const int avail = CS - 2;
        const int bm1_bytes = (in_size + 7) / 8;
        const int bm2_bytes = (bm1_bytes + 7) / 8;
        const int bm3_bytes = (bm2_bytes + 7) / 8;
        uint32_t size = 0;
        bool raw = false;
        uint32_t bm2_nr = 0, bm3_nr = 0, bm4_nr = 0;
        uint32_t rmask1 = 0;   // bm1 bytes 32l..32l+31 non-repeated
        uint32_t rmask2 = 0;   // bm2 bytes 4l..4l+3 non-repeated (4 bits)
        uint32_t bm3_byte = 0; // lanes 0..31: bm3 byte l
        uint32_t rmask3 = 0;   // ballot over bm3 bytes (= bm4 as u32)
        uint32_t bm2w = 0;     // bm2 bytes 4l..4l+3 as a word
        if (data_nz == 0) {
            size = 2;
        } else {
            int wpos = (int)data_nz;
            if (wpos >= avail || wpos > avail - bm1_bytes) raw = true;
            if (!raw) {
                // -- level 2: RE over bm1 (N = bm1_bytes); lane l owns bytes 32l..32l+31
                const uint4 a = *reinterpret_cast<const uint4*>(bm1 + 32 * l);
                const uint4 b = *reinterpret_cast<const uint4*>(bm1 + 32 * l + 16);
                const uint32_t last_prev = __shfl_up(b.w >> 24, 1, 64);
                uint32_t prev = (l == 0) ? 0u : last_prev;
                const uint32_t wds[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
                for (int k = 0; k < 32; k++) {
                    const uint32_t cur = (wds[k >> 2] >> (8 * (k & 3))) & 0xFFu;
                    if (cur != prev && 32 * l + k < bm1_bytes) rmask1 |= 1u << k;
                    prev = cur;
                }
                bm2w = rmask1;  // bm2 bytes 4l..4l+3 (bit k <-> bm1 byte 32l+k)
                bm2_nr = wave_sum(__popc(rmask1));
                wpos += (int)bm2_nr;
                if (wpos >= avail) raw = true;
                if (!raw && (wpos >= avail || wpos > avail - bm2_bytes)) raw = true;
            }
            if (!raw) {
                // -- level 3: RE over bm2 (N = bm2_bytes); lane l owns bytes 4l..4l+3
                uint32_t prev = __shfl_up(bm2w >> 24, 1, 64);
                if (l == 0) prev = 0u;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t cur = (bm2w >> (8 * k)) & 0xFFu;
                    if (cur != prev && 4 * l + k < bm2_bytes) rmask2 |= 1u << k;
                    prev = cur;
                }
                bm3_nr = wave_sum(__popc(rmask2));
                // bm3 byte m = nibble(2m) | nibble(2m+1) << 4
                const uint32_t nlo = __shfl(rmask2, (2 * l) & 63, 64);
                const uint32_t nhi = __shfl(rmask2, (2 * l + 1) & 63, 64);
                bm3_byte = (l < 32) ? (nlo | (nhi << 4)) : 0u;
                wpos += (int)bm3_nr;
                if (wpos >= avail) raw = true;
                if (!raw && (wpos >= avail - 4 || wpos > (avail - 4) - bm3_bytes)) raw = true;
            }
            if (!raw) {
                // -- level 4: RE over bm3 (N = bm3_bytes); lanes 0..31 own bm3 bytes
                uint32_t prev = __shfl_up(bm3_byte, 1, 64);
                if (l == 0) prev = 0u;
                const bool nr = (l < bm3_bytes) && bm3_byte != prev;
                rmask3 = (uint32_t)__ballot(nr);
                bm4_nr = __popc(rmask3);
                wpos += (int)bm4_nr;
                if (wpos >= avail - 4) raw = true;
            }
            size = raw ? ((1u << 31) | (uint32_t)in_size) : (uint32_t)(wpos + 4 + 2);
        }
        const uint32_t clean = size & 0x7FFFFFFFu;

        // -- assemble the chunk stream in LDS
        if (size == 2) {
            if (l == 0) { st[0] = (uint8_t)(in_size & 0xFF); st[1] = (uint8_t)((in_size >> 8) & 0xFF); }
        } else if (raw) {
#pragma unroll
            for (int j = 0; j < 16; j++) *reinterpret_cast<uint4*>(st + 1024 * j + 16 * l) = v[j];
        } else {
            // nonzero bytes in order (j, lane, k)
            uint32_t run = 0;
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const uint32_t cnt = __popc(m[j]);
                const uint32_t incl = wave_incl_scan(cnt);
                uint32_t pos = run + incl - cnt;
#pragma unroll
                for (int k = 0; k < 16; k++)
                    if (m[j] & (1u << k)) st[pos++] = (uint8_t)byte_of(v[j], k);
                run += __shfl(incl, 63, 64);
            }
            // bm1 non-repeated bytes
            {
                const uint32_t cnt = __popc(rmask1);
                const uint32_t incl = wave_incl_scan(cnt);
                uint32_t pos = data_nz + incl - cnt;
                for (int k = 0; k < 32; k++)
                    if (rmask1 & (1u << k)) st[pos++] = bm1[32 * l + k];
            }
            // bm2 non-repeated bytes
            {
                const uint32_t cnt = __popc(rmask2);
                const uint32_t incl = wave_incl_scan(cnt);
                uint32_t pos = data_nz + bm2_nr + incl - cnt;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (rmask2 & (1u << k)) st[pos++] = (uint8_t)((bm2w >> (8 * k)) & 0xFFu);
            }
            // bm3 non-repeated bytes, bm4, tag
            const uint32_t base3 = data_nz + bm2_nr + bm3_nr;
            if (rmask3 & (1u << (l & 31)) && l < 32) st[base3 + __popc(rmask3 & ((1u << l) - 1u))] = (uint8_t)bm3_byte;
            const uint32_t base4 = base3 + bm4_nr;
            if (l < 4) st[base4 + l] = (uint8_t)((rmask3 >> (8 * l)) & 0xFFu);
            if (l == 0) {
                st[base4 + 4] = (uint8_t)(in_size & 0xFF);
                st[base4 + 5] = (uint8_t)((in_size >> 8) & 0xFF);
            }
        }
        wave_lds_sync();
        if (!raw && clean > 0) {
            // compressed bytes -> this chunk's scratch slot (dword stores; the
            // slot is CS bytes and a non-raw chunk is shorter)
            const uint32_t nd = (clean + 3) / 4;
            const uint32_t* sw = reinterpret_cast<const uint32_t*>(st);
            uint32_t* dw = reinterpret_cast<uint32_t*>(scratch + (uint64_t)c * CS);
            for (uint32_t i = l; i < nd; i += 64) dw[i] = sw[i];
        }
        if (l == 0) {
            sizes[c] = size;
            reinterpret_cast<uint32_t*>(out + 8)[c] = size;
            if (c == 0) {
                reinterpret_cast<uint32_t*>(out)[0] = (uint32_t)nbytes;
                reinterpret_cast<uint32_t*>(out)[1] = nc;
            }
        }
    }
}

// Copy `size` bytes from `src` (4-byte aligned; `limit` bytes readable) to
// `dst` (any alignment), touching no byte outside [dst, dst+size).  Whole
// wave calls.
__device__ __forceinline__ void copy_to_unaligned(const uint8_t* src, uint64_t limit, uint8_t* dst, uint32_t size) {
    if (size == 0) return;
    const int l = lane_id();
    const uint32_t head = (uint32_t)((4 - ((uintptr_t)dst & 3)) & 3);
    const uint32_t hb = head < size ? head : size;
    if ((uint32_t)l < hb) dst[l] = src[l];
    if (size <= hb) return;
    uint32_t* dw = reinterpret_cast<uint32_t*>(dst + hb);
    const uint32_t nd = (size - hb) / 4;
    const uint32_t* sw = reinterpret_cast<const uint32_t*>(src);
    auto word = [&](uint32_t q) -> uint32_t {
        if (4ull * q + 4 <= limit) return sw[q];
        uint32_t v = 0;
        for (uint32_t k = 0; k < 4; k++)
            if (4ull * q + k < limit) v |= (uint32_t)src[4 * q + k] << (8 * k);
        return v;
    };
    for (uint32_t i = l; i < nd; i += 64) {
        const uint32_t p = hb + 4 * i;  // source byte of this dword
        const uint32_t q = p >> 2;
        dw[i] = hb == 0 ? sw[q] : __builtin_amdgcn_alignbyte(word(q + 1), sw[q], hb);
    }
    const uint32_t done = hb + 4 * nd;
    if ((uint32_t)l < size - done) dst[done + l] = src[done + l];
}

// Places every chunk at [8 + 4 nc + offset): workgroup g owns chunks
// [PACK * g, PACK * g + PACK) and sums all earlier sizes itself (the encode
// kernel has finished, so they are all in memory).
constexpr uint32_t PACK = 16;

__global__ void __launch_bounds__(256)
rze_pack_kernel(const uint8_t* __restrict__ in, uint64_t nbytes, const uint8_t* __restrict__ scratch,
                const uint32_t* __restrict__ sizes, uint8_t* __restrict__ out, uint64_t* __restrict__ out_size,
                uint32_t nc) {
    __shared__ uint32_t s_red[4];
    __shared__ uint32_t s_off[PACK];
    const int tid = threadIdx.x, w = tid >> 6, l = lane_id();
    const uint32_t c0 = blockIdx.x * PACK;
    uint32_t part = 0;
    for (uint32_t i = tid; i < c0; i += 256) part += sizes[i] & 0x7FFFFFFFu;
    part = wave_sum(part);
    if (l == 0) s_red[w] = part;
    if (tid == 0) {
        uint32_t run = 0;
        for (uint32_t k = 0; k < PACK; k++) {
            s_off[k] = run;
            if (c0 + k < nc) run += sizes[c0 + k] & 0x7FFFFFFFu;
        }
    }
    __syncthreads();
    const uint64_t base = (uint64_t)s_red[0] + s_red[1] + s_red[2] + s_red[3];
    const uint64_t hdr = 8ull + 4ull * nc;
    for (uint32_t k = w; k < PACK; k += 4) {
        const uint32_t c = c0 + k;
        if (c >= nc) break;
        const uint32_t size = sizes[c];
        const uint32_t clean = size & 0x7FFFFFFFu;
        const bool raw = (size & 0x80000000u) != 0u;
        const uint64_t dst = hdr + base + s_off[k];
        if (raw) {
            const uint64_t src = (uint64_t)c * CS;
            copy_to_unaligned(in + src, nbytes - src, out + dst, clean);
        } else {
            copy_to_unaligned(scratch + (uint64_t)c * CS, CS, out + dst, clean);
        }
        if (c == nc - 1) {
            const uint64_t end = dst + clean;
            const uint64_t padded = (end + 3) & ~3ull;
            if ((uint64_t)l < padded - end) out[end + l] = 0;
            if (l == 0) *out_size = padded;
        }
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

// [sizes u32 x nc][chunk scratch, CS bytes per chunk]
size_t fzh_rze_enc_ws_bytes(size_t nbytes) {
    const size_t nc = cdiv(nbytes, CS);
    return align_up(4 * std::max<size_t>(nc, 1), 256) + nc * CS;
}

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
    const unsigned blocks = (unsigned)cdiv(nc, WAVES);
    uint32_t* sizes = (uint32_t*)d_ws;
    uint8_t* scratch = (uint8_t*)d_ws + align_up(4 * (size_t)nc, 256);
    const uint8_t* in = (const uint8_t*)d_in;
    uint8_t* out = (uint8_t*)d_out;
    if (sew == 1) rze_enc_kernel<1><<<blocks, 64 * WAVES, kLds, s>>>(in, nbytes, out, scratch, sizes, nc, nfull);
    else if (sew == 2) rze_enc_kernel<2><<<blocks, 64 * WAVES, kLds, s>>>(in, nbytes, out, scratch, sizes, nc, nfull);
    else rze_enc_kernel<0><<<blocks, 64 * WAVES, kLds, s>>>(in, nbytes, out, scratch, sizes, nc, 0);
    rze_pack_kernel<<<(unsigned)cdiv(nc, PACK), 256, 0, s>>>(in, nbytes, scratch, sizes, out, d_out_size, nc);
    FZH_CHECK_LAUNCH("rze_enc_kernel");
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

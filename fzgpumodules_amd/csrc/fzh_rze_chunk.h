// fzh_rze_chunk.h — encoding of one 16 KiB RZE chunk on one wave64, with
// the optional fused Bitshuffle of its block (gfx950).
//
// Shared by the single-field kernels (rze.hip) and the batched persistent
// kernels (batch.hip), so a chunk's bytes are the same whichever launch
// produced them.  Chunk format (modules/coders/rze/rze_stage.cu:324-445):
// [nonzero bytes][bm1 non-repeated][bm2 non-repeated][bm3 non-repeated]
// [bm4 raw 4 B][in_size u16 LE]; an all-zero chunk is the 2-byte tag; size
// bit 31 = chunk stored raw.  The raw-fallback thresholds are the
// reference's, verbatim (:369-431).
#pragma once

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

// Encodes the chunk at `src` (in_size <= CS bytes; a full block that is
// bit-shuffled first when SEW != 0 and `shuffled`).  Returns the size word
// (bit 31 = raw).  The chunk's stored bytes are written to `slot` (CS bytes,
// dword aligned) when they cannot be taken from the input as they are: a
// compressed chunk, or a raw bit-shuffled block (slot == nullptr: size
// only).  `st` and `bm1` are this wave's LDS (LDS_STREAM and LDS_BM1
// bytes).  Whole wave calls.
template <int SEW>
__device__ __forceinline__ uint32_t rze_encode_chunk(const uint8_t* __restrict__ src, int in_size, bool shuffled,
                                                     uint8_t* st, uint8_t* bm1, uint8_t* __restrict__ slot) {
    const int l = lane_id();
    const bool full = in_size == CS;

    // -- load the chunk: lane l, step j holds bytes [1024 j + 16 l, +16)
    uint4 v[16];
    if (SEW != 0 && shuffled) {
        shuffled_load<SEW == 0 ? 2 : SEW>(src, v);
    } else {
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int p = 1024 * j + 16 * l;
            if (full || p + 16 <= in_size) {
                v[j] = *reinterpret_cast<const uint4*>(src + p);
            } else {
                uint32_t wds[4] = {0, 0, 0, 0};
                for (int k = 0; k < 16; k++)
                    if (p + k < in_size) wds[k >> 2] |= (uint32_t)src[p + k] << (8 * (k & 3));
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
    // A raw chunk is copied from the input by the pack kernel, except a
    // bit-shuffled block, whose stored bytes are the shuffled image.
    const bool from_scratch = !raw || (SEW != 0 && shuffled);
    if (slot && from_scratch && clean > 0) {
        // chunk bytes -> this chunk's scratch slot (dword stores; the slot
        // is CS bytes)
        const uint32_t nd = (clean + 3) / 4;
        const uint32_t* sw = reinterpret_cast<const uint32_t*>(st);
        uint32_t* dw = reinterpret_cast<uint32_t*>(slot);
        for (uint32_t i = l; i < nd; i += 64) dw[i] = sw[i];
    }
    wave_lds_sync();  // LDS free for the wave's next chunk
    return size;
}

// Copy `size` bytes from `src` (4-byte aligned; `limit` bytes readable) to
// `dst` (any alignment), touching no byte outside [dst, dst+size): bytes up
// to the first 16-byte boundary of dst, then 16-byte stores (each lane
// assembles its word from five source dwords with alignbyte), then the
// tail.  Whole wave calls.
__device__ __forceinline__ void copy_to_unaligned16(const uint8_t* src, uint64_t limit, uint8_t* dst,
                                                    uint32_t size) {
    if (size == 0) return;
    const int l = lane_id();
    const uint32_t head = (uint32_t)((16 - ((uintptr_t)dst & 15)) & 15);
    const uint32_t hb = head < size ? head : size;
    if ((uint32_t)l < hb) dst[l] = src[l];
    if (size <= hb) return;
    const uint32_t sh = hb & 3;  // source byte hb sits at this offset in its dword
    const uint32_t* sw = reinterpret_cast<const uint32_t*>(src);
    auto word = [&](uint32_t q) -> uint32_t {
        if (4ull * q + 4 <= limit) return sw[q];
        uint32_t v = 0;
        for (uint32_t k = 0; k < 4; k++)
            if (4ull * q + k < limit) v |= (uint32_t)src[4 * q + k] << (8 * k);
        return v;
    };
    uint4* dq = reinterpret_cast<uint4*>(dst + hb);
    const uint32_t n16 = (size - hb) / 16;
    for (uint32_t i = l; i < n16; i += 64) {
        const uint32_t q = (hb + 16 * i) >> 2;
        const uint32_t w0 = sw[q], w1 = sw[q + 1], w2 = sw[q + 2], w3 = sw[q + 3];
        uint4 v;
        if (sh == 0) {
            v = make_uint4(w0, w1, w2, w3);
        } else {
            const uint32_t w4 = word(q + 4);
            v = make_uint4(__builtin_amdgcn_alignbyte(w1, w0, sh), __builtin_amdgcn_alignbyte(w2, w1, sh),
                           __builtin_amdgcn_alignbyte(w3, w2, sh), __builtin_amdgcn_alignbyte(w4, w3, sh));
        }
        dq[i] = v;
    }
    const uint32_t done = hb + 16 * n16;
    if ((uint32_t)l < size - done) dst[done + l] = src[done + l];
}

}  // namespace
}  // namespace fzh

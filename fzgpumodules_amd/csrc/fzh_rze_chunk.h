// fzh_rze_chunk.h — encoding of one 16 KiB RZE chunk by one 256-thread
// workgroup, with the optional fused Bitshuffle of its block, plus the
// decode-side un-shuffle and the unaligned copy helpers (gfx950).
//
// Shared by the single-field kernels (rze.hip), the batched kernels
// (batch.hip) and the labeller (label.hip), so a chunk's bytes are the same
// whichever launch produced them.  Chunk format (modules/coders/rze/rze_stage.cu:324-445):
// [nonzero bytes][bm1 non-repeated][bm2 non-repeated][bm3 non-repeated]
// [bm4 raw 4 B][in_size u16 LE]; an all-zero chunk is the 2-byte tag; size
// bit 31 = chunk stored raw.  The raw-fallback thresholds are the
// reference's, verbatim (:369-431).
#pragma once

#include "fzh_bitplane.h"

namespace fzh {
namespace {

constexpr int CS = 16384;
// Stream/staging buffer (+ slack for the aligned dword reads that straddle
// its end) and level-1 bitmap sizes in LDS.
constexpr int LDS_STREAM = CS + 64;
constexpr int LDS_BM1 = CS / 8;

// Fused Bitshuffle (16 KiB blocks, element width SEW = 1 or 2) on the decode
// side, in the workgroup layout: d[p] is the shuffled block's dword at byte
// 1024 p + 4 t (thread t).  Shuffled word index = plane * npp + group with
// npp = 8192 / SEW / 32, so thread t holds every plane of element group t
// (SEW 2: plane 15 - p in d[p]) or of groups t and 256 + t (SEW 1: plane
// 7 - P of group 256 h + t in d[2P + h]) and writes them back transposed.
template <int SEW>
__device__ __forceinline__ void unshuffled_store(const uint32_t (&d)[16], uint8_t* blk) {
    const int t = threadIdx.x;
    if constexpr (SEW == 2) {
        uint32_t A[16], w[16];
#pragma unroll
        for (int b = 0; b < 16; b++) A[b] = d[15 - b];
        unplanes16(A, w);
        uint8_t* g = blk + 64 * t;
#pragma unroll
        for (int q = 0; q < 4; q++) st16(g + 16 * q, make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]));
    } else {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint32_t A[8], w[8];
#pragma unroll
            for (int P = 0; P < 8; P++) A[7 - P] = d[2 * P + h];
            unplanes8(A, w);
            uint8_t* g = blk + 32 * (256 * h + t);
            st16(g, make_uint4(w[0], w[1], w[2], w[3]));
            st16(g + 16, make_uint4(w[4], w[5], w[6], w[7]));
        }
    }
}

// ---------------------------------------------------------------------------
// The encoder: one 256-thread workgroup (four waves) per chunk; thread t,
// step p holds the chunk's dword at byte 1024 p + 4 t, so every
// load and store of a step covers 1 KiB contiguously and a thread keeps 16
// dwords instead of 64.  For bit-shuffled blocks thread t transposes the
// element group(s) whose plane words land at exactly those positions.
// Level 2 gives thread t bm1 bytes 8t..8t+7 (so bm2 byte t is its 8-bit
// repeat mask), level 3 one bm2 byte per thread (bm3 bytes from per-wave
// ballots), level 4 runs in wave 0.
// ---------------------------------------------------------------------------
struct RzeWgLds {
    uint8_t st[LDS_STREAM];   // assembled stream (dword aligned: offset 0)
    uint32_t sel[16];         // v_perm selectors packing the nonzero bytes of a dword, by nibble
    uint8_t bm1[LDS_BM1];     // level-1 bitmap
    uint32_t wsum[4][16];     // per-wave partial sums (scratch for block scans)
    uint32_t tot[16];         // broadcast totals
    uint8_t bm2[256];
    uint64_t bal[4];          // level-3 ballots (bm3 bytes 8w..8w+7)
};

// Block-wide exclusive scan of one u32 per thread (256 threads); returns
// the thread's exclusive prefix, *total gets the sum.  Uses L.wsum[..][slot].
__device__ __forceinline__ uint32_t wg_excl_scan(RzeWgLds& L, uint32_t v, int slot, uint32_t& total) {
    const int w = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_scan(v);
    if (lane_id() == 63) L.wsum[w][slot] = incl;
    __syncthreads();
    uint32_t base = 0;
    for (int j = 0; j < w; j++) base += L.wsum[j][slot];
    total = L.wsum[0][slot] + L.wsum[1][slot] + L.wsum[2][slot] + L.wsum[3][slot];
    return base + incl - v;
}

template <int SEW>
__device__ __forceinline__ uint32_t rze_encode_chunk_wg(const uint8_t* __restrict__ src, int in_size, bool shuffled,
                                                        RzeWgLds& L, uint8_t* __restrict__ slot) {
    const int t = threadIdx.x;
    const int w = t >> 6;
    const int l = lane_id();
    const bool full = in_size == CS;
    if (t < 16) {  // selector for nibble t: byte slot c <- c-th nonzero byte, zeros after
        uint32_t sel = 0x0C0C0C0Cu;
        int c = 0;
        for (int k = 0; k < 4; k++)
            if (t & (1 << k)) {
                sel = (sel & ~(0xFFu << (8 * c))) | ((uint32_t)k << (8 * c));
                c++;
            }
        L.sel[t] = sel;
    }
    // -- load: dword d[p] = chunk bytes [1024 p + 4 t, +4)
    uint32_t d[16];
    if (SEW == 2 && shuffled) {
        const uint8_t* g = src + 64 * t;  // element group t (32 x u16)
        uint32_t wd[16];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 x = ld16(g + 16 * q);
            wd[4 * q] = x.x; wd[4 * q + 1] = x.y; wd[4 * q + 2] = x.z; wd[4 * q + 3] = x.w;
        }
        uint32_t A[16];
        planes16(wd, A);
#pragma unroll
        for (int p = 0; p < 16; p++) d[p] = A[15 - p];
    } else if (SEW == 1 && shuffled) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const uint8_t* g = src + 32 * (256 * h + t);  // element group 256 h + t (32 x u8)
            const uint4 x0 = ld16(g), x1 = ld16(g + 16);
            const uint32_t wd[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
            uint32_t A[8];
            planes8(wd, A);
#pragma unroll
            for (int P = 0; P < 8; P++) d[2 * P + h] = A[7 - P];
        }
    } else {
#pragma unroll
        for (int p = 0; p < 16; p++) {
            const int b = 1024 * p + 4 * t;
            if (full || b + 4 <= in_size) {
                d[p] = *reinterpret_cast<const uint32_t*>(src + b);
            } else {
                uint32_t x = 0;
                for (int k = 0; k < 4; k++)
                    if (b + k < in_size) x |= (uint32_t)src[b + k] << (8 * k);
                d[p] = x;
            }
        }
    }
    // -- level 1: nibble of nonzero bytes per step; bm1 byte 128 p + t/2
    uint64_t nib = 0;  // 4 bits per step
    uint32_t my_nz = 0;
#pragma unroll
    for (int p = 0; p < 16; p++) {
        uint32_t m = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) m |= (((d[p] >> (8 * k)) & 0xFFu) != 0u ? 1u : 0u) << k;
        nib |= (uint64_t)m << (4 * p);
        my_nz += __popc(m);
        const uint32_t other = lane_xor1(m);
        if ((t & 1) == 0) L.bm1[128 * p + (t >> 1)] = (uint8_t)(m | (other << 4));
    }
    uint32_t data_nz;
    (void)wg_excl_scan(L, my_nz, 0, data_nz);  // barrier: bm1 complete
    // zero the data part of the stream: the data bytes are ORed in below
    for (uint32_t i = t; i < (data_nz + 3) / 4; i += 256) reinterpret_cast<uint32_t*>(L.st)[i] = 0u;

    const int avail = CS - 2;
    const int bm1_bytes = (in_size + 7) / 8;
    const int bm2_bytes = (bm1_bytes + 7) / 8;
    const int bm3_bytes = (bm2_bytes + 7) / 8;
    uint32_t size = 0;
    bool raw = false;
    uint32_t bm2_nr = 0, bm3_nr = 0, bm4_nr = 0;
    uint32_t rmask1 = 0;  // bm1 bytes 8t..8t+7 non-repeated (= bm2 byte t)
    uint32_t nr2 = 0;     // bm2 byte t non-repeated
    uint32_t rmask3 = 0;  // bm4 (wave 0 computes it; broadcast through LDS)
    if (data_nz == 0) {
        size = 2;
    } else {
        int wpos = (int)data_nz;
        if (wpos >= avail || wpos > avail - bm1_bytes) raw = true;
        if (!raw) {
            // -- level 2: RE over bm1; thread t owns bytes 8t..8t+7
            const uint32_t lo = *reinterpret_cast<const uint32_t*>(L.bm1 + 8 * t);
            const uint32_t hi = *reinterpret_cast<const uint32_t*>(L.bm1 + 8 * t + 4);
            uint32_t prev = t > 0 ? L.bm1[8 * t - 1] : 0u;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const uint32_t cur = ((k < 4 ? lo : hi) >> (8 * (k & 3))) & 0xFFu;
                if (cur != prev && 8 * t + k < bm1_bytes) rmask1 |= 1u << k;
                prev = cur;
            }
            L.bm2[t] = (uint8_t)rmask1;
            (void)wg_excl_scan(L, __popc(rmask1), 1, bm2_nr);  // barrier: bm2 complete
            wpos += (int)bm2_nr;
            if (wpos >= avail) raw = true;
            if (!raw && (wpos >= avail || wpos > avail - bm2_bytes)) raw = true;
        }
        if (!raw) {
            // -- level 3: RE over bm2; thread t owns bm2 byte t
            const uint32_t prev = t > 0 ? L.bm2[t - 1] : 0u;
            nr2 = (t < bm2_bytes && rmask1 != prev) ? 1u : 0u;
            const uint64_t b = __ballot(nr2 != 0u);
            if (l == 0) L.bal[w] = b;
            (void)wg_excl_scan(L, nr2, 2, bm3_nr);  // barrier: ballots visible
            wpos += (int)bm3_nr;
            if (wpos >= avail) raw = true;
            if (!raw && (wpos >= avail - 4 || wpos > (avail - 4) - bm3_bytes)) raw = true;
        }
        if (!raw) {
            // -- level 4: RE over bm3 (32 bytes: ballot bytes), in wave 0
            if (w == 0) {
                const uint32_t cur = l < 32 ? (uint32_t)((L.bal[l >> 3] >> (8 * (l & 7))) & 0xFFu) : 0u;
                const uint32_t prev = wave_shr1(cur);  // lane 0: 0
                const uint32_t m3 = (uint32_t)__ballot(l < bm3_bytes && cur != prev);
                if (l == 0) L.tot[0] = m3;
            }
            __syncthreads();
            rmask3 = L.tot[0];
            bm4_nr = __popc(rmask3);
            wpos += (int)bm4_nr;
            if (wpos >= avail - 4) raw = true;
        }
        size = raw ? ((1u << 31) | (uint32_t)in_size) : (uint32_t)(wpos + 4 + 2);
    }
    const uint32_t clean = size & 0x7FFFFFFFu;
    const bool from_scratch = !raw || (SEW != 0 && shuffled);
    if (raw) {
        // stored bytes = the (shuffled) chunk image, straight from registers
        if (slot && from_scratch) {
#pragma unroll
            for (int p = 0; p < 16; p++) reinterpret_cast<uint32_t*>(slot)[256 * p + t] = d[p];
        }
        __syncthreads();
        return size;
    }
    if (size == 2) {
        if (t == 0) { L.st[0] = (uint8_t)(in_size & 0xFF); L.st[1] = (uint8_t)((in_size >> 8) & 0xFF); }
    } else {
        // nonzero bytes in chunk order: step-major, then thread, then byte
        uint32_t incl[16];
#pragma unroll
        for (int p = 0; p < 16; p++) {
            incl[p] = wave_incl_scan((uint32_t)__popcll((nib >> (4 * p)) & 0xFull));
            if (l == 63) L.wsum[w][p] = incl[p];
        }
        __syncthreads();
        // lane q < 16 of each wave: where this wave's bytes of step q start =
        // bytes of all steps before q + bytes of step q in earlier waves
        uint32_t qb = 0;
        {
            uint32_t before = 0, all = 0;
            if (l < 16) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint32_t v = L.wsum[j][l];
                    before += j < w ? v : 0u;
                    all += v;
                }
            }
            qb = wave_incl_scan(all) - all + before;  // lanes 0..15 form one DPP row
        }
#pragma unroll
        for (int p = 0; p < 16; p++) {
            const uint32_t m = (uint32_t)((nib >> (4 * p)) & 0xFull);
            const uint32_t base = (uint32_t)__builtin_amdgcn_readlane((int)qb, p);
            const uint32_t pos = base + incl[p] - __popc(m);
            if (m) {
                // the dword's nonzero bytes packed to its low end, ORed into
                // the zeroed stream at [pos, pos + popc(m)) with two aligned
                // dword ORs (neighbouring threads' bytes share the boundary
                // dwords, never a byte)
                const uint32_t packed = __builtin_amdgcn_perm(0u, d[p], L.sel[m]);
                const uint64_t sh = (uint64_t)packed << (8 * (pos & 3));
                uint32_t* sw = reinterpret_cast<uint32_t*>(L.st) + (pos >> 2);
                atomicOr(sw, (uint32_t)sh);
                if ((uint32_t)(sh >> 32)) atomicOr(sw + 1, (uint32_t)(sh >> 32));
            }
        }
        __syncthreads();  // wsum reused below
        // bm1 non-repeated bytes (thread t: bm1 bytes 8t..8t+7)
        uint32_t dummy;
        uint32_t pos1 = data_nz + wg_excl_scan(L, __popc(rmask1), 3, dummy);
#pragma unroll
        for (int k = 0; k < 8; k++)
            if (rmask1 & (1u << k)) L.st[pos1++] = L.bm1[8 * t + k];
        // bm2 non-repeated bytes (one per thread at most)
        const uint32_t pos2 = data_nz + bm2_nr + wg_excl_scan(L, nr2, 4, dummy);
        if (nr2) L.st[pos2] = (uint8_t)rmask1;
        // bm3 non-repeated bytes, bm4, tag
        const uint32_t base3 = data_nz + bm2_nr + bm3_nr;
        if (w == 0 && l < 32 && (rmask3 & (1u << l))) {
            const uint32_t b3 = (uint32_t)((L.bal[l >> 3] >> (8 * (l & 7))) & 0xFFu);
            L.st[base3 + __popc(rmask3 & ((1u << l) - 1u))] = (uint8_t)b3;
        }
        const uint32_t base4 = base3 + bm4_nr;
        if (t < 4) L.st[base4 + t] = (uint8_t)((rmask3 >> (8 * t)) & 0xFFu);
        if (t == 0) {
            L.st[base4 + 4] = (uint8_t)(in_size & 0xFF);
            L.st[base4 + 5] = (uint8_t)((in_size >> 8) & 0xFF);
        }
    }
    __syncthreads();
    if (slot && clean > 0) {
        const uint32_t nd = (clean + 3) / 4;
        const uint32_t* sw = reinterpret_cast<const uint32_t*>(L.st);
        uint32_t* dw = reinterpret_cast<uint32_t*>(slot);
        for (uint32_t i = t; i < nd; i += 256) dw[i] = sw[i];
    }
    __syncthreads();  // LDS free for the workgroup's next chunk
    return size;
}

// Copy `size` bytes from `src` (4-byte aligned; `limit` bytes readable) to
// `dst` (any alignment), touching no byte outside [dst, dst+size): bytes up
// to the first 16-byte boundary of dst, then 16-byte stores (each lane
// assembles its word from five source dwords with alignbyte), then the
// tail.  Called by a whole wave (NT = 64) or a whole workgroup of NT threads.
template <int NT = 64>
__device__ __forceinline__ void copy_to_unaligned16(const uint8_t* src, uint64_t limit, uint8_t* dst,
                                                    uint32_t size) {
    if (size == 0) return;
    const int l = NT == 64 ? lane_id() : (int)threadIdx.x;
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
    for (uint32_t i = l; i < n16; i += NT) {
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

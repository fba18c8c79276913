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
// Byte-mask helpers (v_dot4_u32_u8 gathers one flag bit per byte)
// ---------------------------------------------------------------------------
// 0x80 in every nonzero byte of v, 0 in every zero byte.
__device__ __forceinline__ uint32_t nz_bytes(uint32_t v) {
    return (((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u;
}
// Bit i set <=> byte i of the 8-byte little-endian pair (lo, hi) is nonzero.
__device__ __forceinline__ uint32_t nz_mask8(uint32_t lo, uint32_t hi) {
    return __builtin_amdgcn_udot4(nz_bytes(hi), 0x80402010u,
                                  __builtin_amdgcn_udot4(nz_bytes(lo), 0x08040201u, 0u, false), false) >> 7;
}
// Repeat-elimination flags of 8 consecutive bitmap bytes (lo, hi): bit k set
// <=> byte k differs from byte k-1, with byte -1 = prevb.
__device__ __forceinline__ uint32_t repeat_flags8(uint32_t lo, uint32_t hi, uint32_t prevb) {
    const uint32_t ylo = (lo << 8) | prevb;
    const uint32_t yhi = __builtin_amdgcn_alignbyte(hi, lo, 3);
    return nz_mask8(lo ^ ylo, hi ^ yhi);
}
// Low min(max(k, 0), 8) bits set.
__device__ __forceinline__ uint32_t low_bits8(int k) { return (1u << (k < 0 ? 0 : k > 8 ? 8 : k)) - 1u; }

// v_perm selectors that pack the bytes of an 8-byte pair (lo = bytes 0-3,
// hi = 4-7, v_perm(hi, lo, sel)) named by the set bits of a mask byte to the
// low end in order: [m][0] gives packed bytes 0-3, [m][1] bytes 4-7; 0x0C
// selects a zero byte.  kExpand8 is the inverse: output byte k takes packed
// byte (number of set bits below k) when bit k is set, else zero.
struct alignas(16) Sel8Table {
    uint32_t v[256][2];
};
constexpr Sel8Table make_pack8() {
    Sel8Table T{};
    for (int m = 0; m < 256; m++) {
        uint32_t s[2] = {0x0C0C0C0Cu, 0x0C0C0C0Cu};
        int c = 0;
        for (int i = 0; i < 8; i++)
            if (m & (1 << i)) {
                s[c >> 2] = (s[c >> 2] & ~(0xFFu << (8 * (c & 3)))) | ((uint32_t)i << (8 * (c & 3)));
                c++;
            }
        T.v[m][0] = s[0];
        T.v[m][1] = s[1];
    }
    return T;
}
constexpr Sel8Table make_expand8() {
    Sel8Table T{};
    for (int m = 0; m < 256; m++) {
        uint32_t s[2] = {0x0C0C0C0Cu, 0x0C0C0C0Cu};
        int c = 0;
        for (int k = 0; k < 8; k++)
            if (m & (1 << k)) {
                s[k >> 2] = (s[k >> 2] & ~(0xFFu << (8 * (k & 3)))) | ((uint32_t)c << (8 * (k & 3)));
                c++;
            }
        T.v[m][0] = s[0];
        T.v[m][1] = s[1];
    }
    return T;
}
// Repeat expansion: output byte k takes packed byte r - 1, r = set bits of
// the mask at positions <= k, or zero when r = 0 (the caller fills those
// leading bytes with the byte before the run).
constexpr Sel8Table make_repeat8() {
    Sel8Table T{};
    for (int m = 0; m < 256; m++) {
        uint32_t s[2] = {0x0C0C0C0Cu, 0x0C0C0C0Cu};
        int r = 0;
        for (int k = 0; k < 8; k++) {
            if (m & (1 << k)) r++;
            if (r > 0) s[k >> 2] = (s[k >> 2] & ~(0xFFu << (8 * (k & 3)))) | ((uint32_t)(r - 1) << (8 * (k & 3)));
        }
        T.v[m][0] = s[0];
        T.v[m][1] = s[1];
    }
    return T;
}
__constant__ Sel8Table kPack8 = make_pack8();
__constant__ Sel8Table kExpand8 = make_expand8();
__constant__ Sel8Table kRepeat8 = make_repeat8();

// ORs the 8 bytes (lo, hi) into the byte stream `st` at byte position pos
// (bytes past the packed count are zero, so neighbours' bytes that share a
// dword are untouched).
__device__ __forceinline__ void or_bytes8(uint32_t* st, uint32_t pos, uint32_t lo, uint32_t hi) {
    const uint32_t sh = 8u * (pos & 3u);
    const uint64_t r = ((uint64_t)hi << 32 | lo) << sh;
    const uint32_t r2 = (uint32_t)(((uint64_t)hi << sh) >> 32);
    uint32_t* d = st + (pos >> 2);
    atomicOr(d, (uint32_t)r);
    if ((uint32_t)(r >> 32)) atomicOr(d + 1, (uint32_t)(r >> 32));
    if (r2) atomicOr(d + 2, r2);
}

// Word w of a chunk image staged in LDS sits at img_swz(w): the 16-byte
// groups of every 64-byte row are rotated by the row's bits 4-5, so a wave
// reading one 16-byte group per lane at a 64-byte stride spreads over the
// banks.  A bijection that keeps 16-byte groups and 64-word blocks intact.
__device__ __forceinline__ uint32_t img_swz(uint32_t w) { return w ^ (((w >> 4) & 3u) << 2); }

// ---------------------------------------------------------------------------
// The encoder: one 256-thread workgroup (four waves) per chunk.  Thread t
// owns the chunk image's bytes [64 t, 64 t + 64) (16 dwords), so its
// level-1 bitmap is bm1 bytes 8t..8t+7 (one 64-bit word in registers), its
// level-2 repeat mask is bm2 byte t, and one block scan places its nonzero
// bytes.  Bit-shuffled blocks are transposed per element group (thread t:
// group t, or groups t and 256 + t for 1-byte elements) and exchanged through
// LDS into that layout first.  Four barriers: the exchange; the data-byte
// totals; the level-2/3 totals and ballots; the assembled stream.  The
// compressed chunk is assembled in LDS with byte-exact ORs and copied to
// `slot` (a raw bit-shuffled block is written from registers; other raw
// chunks are left to the caller, which copies the input).  With slot ==
// nullptr only the size is computed.  Returns the size (bit 31 = raw).  One
// call per workgroup launch: LDS is still in use by other waves on return.
// ---------------------------------------------------------------------------
struct RzeWgLds {
    uint32_t st[LDS_STREAM / 4];  // the assembled stream; first the shuffled image (img_swz order)
    alignas(8) uint32_t sel[256][2];  // kPack8
    uint32_t wsum[4][2];          // per-wave totals: data bytes, non-repeated bm1 bytes
    uint32_t edge[4][3];          // per-wave lane 63: bm1 words, the bm1 byte before them
    uint64_t bal[4];              // level-3 flags: bm3 bytes 8w..8w+7
};

template <int SEW>
__device__ __forceinline__ uint32_t rze_encode_chunk_wg(const uint8_t* __restrict__ src, int in_size, bool shuffled,
                                                        RzeWgLds& L, uint8_t* __restrict__ slot) {
    const int t = threadIdx.x;
    const int w = (int)wave_in_wg();
    const int l = lane_id();
    if (slot) {
        const uint2 s = *reinterpret_cast<const uint2*>(kPack8.v[t]);
        L.sel[t][0] = s.x;
        L.sel[t][1] = s.y;
    }
    // -- the chunk image, thread t: D[j] = bytes [64 t + 4 j, +4)
    uint32_t D[16];
    if (SEW != 0 && shuffled) {
        uint32_t d[16];  // d[p] = image word 256 p + t
        if constexpr (SEW == 2) {
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
        } else {
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
        }
#pragma unroll
        for (int p = 0; p < 16; p++) L.st[img_swz(256 * p + t)] = d[p];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 v = *reinterpret_cast<const uint4*>(&L.st[img_swz(16 * t + 4 * q)]);
            D[4 * q] = v.x; D[4 * q + 1] = v.y; D[4 * q + 2] = v.z; D[4 * q + 3] = v.w;
        }
    } else if (in_size == CS) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint4 v = ld16(src + 64 * t + 16 * q);
            D[4 * q] = v.x; D[4 * q + 1] = v.y; D[4 * q + 2] = v.z; D[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int b = 64 * t + 4 * j;
            if (b + 4 <= in_size) {
                D[j] = *gp(reinterpret_cast<const uint32_t*>(src + b));
            } else {
                uint32_t x = 0;
                for (int k = 0; k < 4; k++)
                    if (b + k < in_size) x |= (uint32_t)gp(src)[b + k] << (8 * k);
                D[j] = x;
            }
        }
    }
    // -- level 1: bm1 bytes 8t..8t+7 and the data byte count
    uint32_t Xlo = 0, Xhi = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        Xlo |= nz_mask8(D[2 * b], D[2 * b + 1]) << (8 * b);
        Xhi |= nz_mask8(D[8 + 2 * b], D[9 + 2 * b]) << (8 * b);
    }
    const uint32_t cnt = __popc(Xlo) + __popc(Xhi);
    const uint32_t cnt_incl = wave_incl_scan(cnt);
    const uint32_t prevb_lane = wave_shr1(Xhi >> 24);  // bm1 byte 8t - 1 (lane 0: not yet known)
    if (l == 63) {
        L.wsum[w][0] = cnt_incl;
        L.edge[w][0] = Xlo;
        L.edge[w][1] = Xhi;
        L.edge[w][2] = prevb_lane;
    }
    __syncthreads();
    const uint32_t s0 = L.wsum[0][0], s1 = L.wsum[1][0], s2 = L.wsum[2][0], s3 = L.wsum[3][0];
    const uint32_t data_nz = s0 + s1 + s2 + s3;
    const uint32_t data_pos = (w > 0 ? s0 : 0u) + (w > 1 ? s1 : 0u) + (w > 2 ? s2 : 0u) + cnt_incl - cnt;
    const int bm1_bytes = (in_size + 7) / 8;
    const int bm2_bytes = (bm1_bytes + 7) / 8;
    const int bm3_bytes = (bm2_bytes + 7) / 8;
    // -- level 2: bm2 byte t = repeat flags of bm1 bytes 8t..8t+7 (byte -1 := 0)
    uint32_t e_lo = 0, e_hi = 0, e_prev = 0;  // the previous wave's lane 63
    if (w > 0) {
        e_lo = L.edge[w - 1][0];
        e_hi = L.edge[w - 1][1];
        e_prev = L.edge[w - 1][2];
    }
    const uint32_t prevb = l == 0 ? (e_hi >> 24) : prevb_lane;
    const uint32_t rm = repeat_flags8(Xlo, Xhi, prevb) & low_bits8(bm1_bytes - 8 * t);
    // -- level 3: bm3 bit t = bm2 byte t differs from bm2 byte t-1 (byte -1 := 0)
    const uint32_t rm_edge = w > 0 ? (repeat_flags8(e_lo, e_hi, e_prev) & low_bits8(bm1_bytes - 8 * (64 * w - 1))) : 0u;
    const uint32_t rm_shifted = wave_shr1(rm);  // every lane takes part in the lane shift
    const uint32_t rm_prev = l == 0 ? rm_edge : rm_shifted;
    const bool nr2 = t < bm2_bytes && rm != rm_prev;
    const uint64_t bal = __ballot(nr2);
    const uint32_t rm_cnt = __popc(rm);
    const uint32_t rm_incl = wave_incl_scan(rm_cnt);
    if (l == 63) L.wsum[w][1] = rm_incl;
    if (l == 0) L.bal[w] = bal;
    if (slot) {  // clear the stream for the ORs (the staged image is in registers by now)
        uint4* z = reinterpret_cast<uint4*>(L.st);
#pragma unroll
        for (int q = 0; q < 4; q++) z[4 * t + q] = make_uint4(0, 0, 0, 0);
        if (t < (LDS_STREAM - CS) / 16) z[4 * 256 + t] = make_uint4(0, 0, 0, 0);
    }
    __syncthreads();
    const uint32_t r0 = L.wsum[0][1], r1 = L.wsum[1][1], r2 = L.wsum[2][1], r3 = L.wsum[3][1];
    const uint32_t bm2_nr = r0 + r1 + r2 + r3;
    const uint64_t b0 = L.bal[0], b1 = L.bal[1], b2 = L.bal[2], b3 = L.bal[3];
    const uint32_t bm3_nr = __popcll(b0) + __popcll(b1) + __popcll(b2) + __popcll(b3);
    // -- level 4: bm4 bit j = bm3 byte j differs from bm3 byte j-1 (byte -1 := 0)
    uint32_t bm4;
    {
        const uint64_t d0 = b0 ^ (b0 << 8), d1 = b1 ^ (b1 << 8 | b0 >> 56), d2 = b2 ^ (b2 << 8 | b1 >> 56),
                       d3 = b3 ^ (b3 << 8 | b2 >> 56);
        bm4 = nz_mask8((uint32_t)d0, (uint32_t)(d0 >> 32)) | nz_mask8((uint32_t)d1, (uint32_t)(d1 >> 32)) << 8 |
              nz_mask8((uint32_t)d2, (uint32_t)(d2 >> 32)) << 16 | nz_mask8((uint32_t)d3, (uint32_t)(d3 >> 32)) << 24;
        if (bm3_bytes < 32) bm4 &= (1u << bm3_bytes) - 1u;
    }
    const uint32_t bm4_nr = __popc(bm4);
    // -- size and the raw fallback, with the reference's thresholds
    const int avail = CS - 2;
    uint32_t size;
    bool raw = false;
    if (data_nz == 0) {
        size = 2;
    } else {
        int wpos = (int)data_nz;
        if (wpos >= avail || wpos > avail - bm1_bytes) raw = true;
        if (!raw) {
            wpos += (int)bm2_nr;
            if (wpos >= avail || wpos > avail - bm2_bytes) raw = true;
        }
        if (!raw) {
            wpos += (int)bm3_nr;
            if (wpos >= avail || wpos >= avail - 4 || wpos > (avail - 4) - bm3_bytes) raw = true;
        }
        if (!raw) {
            wpos += (int)bm4_nr;
            if (wpos >= avail - 4) raw = true;
        }
        size = raw ? ((1u << 31) | (uint32_t)in_size) : (uint32_t)(wpos + 4 + 2);
    }
    if (!slot) return size;
    if (raw) {
        // stored bytes = the (shuffled) chunk image, straight from registers
        if (SEW != 0 && shuffled) {
#pragma unroll
            for (int q = 0; q < 4; q++)
                st16(slot + 64 * t + 16 * q, make_uint4(D[4 * q], D[4 * q + 1], D[4 * q + 2], D[4 * q + 3]));
        }
        return size;
    }
    uint8_t* stb = reinterpret_cast<uint8_t*>(L.st);
    if (size == 2) {
        if (t == 0) { stb[0] = (uint8_t)(in_size & 0xFF); stb[1] = (uint8_t)((in_size >> 8) & 0xFF); }
    } else {
        // nonzero data bytes in chunk order
        uint32_t pos = data_pos;
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint32_t m = ((b < 4 ? Xlo : Xhi) >> (8 * (b & 3))) & 0xFFu;
            if (m) {
                const uint2 sel = *reinterpret_cast<const uint2*>(L.sel[m]);
                or_bytes8(L.st, pos, __builtin_amdgcn_perm(D[2 * b + 1], D[2 * b], sel.x),
                          __builtin_amdgcn_perm(D[2 * b + 1], D[2 * b], sel.y));
            }
            pos += __popc(m);
        }
        // non-repeated bm1 bytes (this thread's: the set bits of rm)
        if (rm) {
            const uint32_t pos1 = data_nz + (w > 0 ? r0 : 0u) + (w > 1 ? r1 : 0u) + (w > 2 ? r2 : 0u) + rm_incl - rm_cnt;
            const uint2 sel = *reinterpret_cast<const uint2*>(L.sel[rm]);
            or_bytes8(L.st, pos1, __builtin_amdgcn_perm(Xhi, Xlo, sel.x), __builtin_amdgcn_perm(Xhi, Xlo, sel.y));
        }
        // non-repeated bm2 bytes (this thread's byte rm when nr2)
        if (nr2) {
            const uint32_t before = (w > 0 ? (uint32_t)__popcll(b0) : 0u) + (w > 1 ? (uint32_t)__popcll(b1) : 0u) +
                                    (w > 2 ? (uint32_t)__popcll(b2) : 0u) + lanes_below(bal);
            stb[data_nz + bm2_nr + before] = (uint8_t)rm;
        }
        // non-repeated bm3 bytes, bm4, the size tag
        const uint32_t base3 = data_nz + bm2_nr + bm3_nr;
        if (t < 32 && (bm4 & (1u << t))) {
            const uint64_t bw = t < 8 ? b0 : t < 16 ? b1 : t < 24 ? b2 : b3;
            stb[base3 + __popc(bm4 & ((1u << t) - 1u))] = (uint8_t)((bw >> (8 * (t & 7))) & 0xFFu);
        }
        const uint32_t base4 = base3 + bm4_nr;
        if (t < 4) stb[base4 + t] = (uint8_t)((bm4 >> (8 * t)) & 0xFFu);
        if (t == 0) {
            stb[base4 + 4] = (uint8_t)(in_size & 0xFF);
            stb[base4 + 5] = (uint8_t)((in_size >> 8) & 0xFF);
        }
    }
    __syncthreads();
    const uint32_t nq = (size + 15) / 16;  // size < CS here
    const uint4* sq = reinterpret_cast<const uint4*>(L.st);
    uint4* dq = reinterpret_cast<uint4*>(slot);
    for (uint32_t i = t; i < nq; i += 256) gst_u4(dq + i, sq[i]);
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
    if ((uint32_t)l < hb) gp(dst)[l] = gp(src)[l];
    if (size <= hb) return;
    const uint32_t sh = hb & 3;  // source byte hb sits at this offset in its dword
    const uint32_t* sw = reinterpret_cast<const uint32_t*>(src);
    auto word = [&](uint32_t q) -> uint32_t {
        if (4ull * q + 4 <= limit) return gp(sw)[q];
        uint32_t v = 0;
        for (uint32_t k = 0; k < 4; k++)
            if (4ull * q + k < limit) v |= (uint32_t)gp(src)[4 * q + k] << (8 * k);
        return v;
    };
    uint4* dq = reinterpret_cast<uint4*>(dst + hb);
    const uint32_t n16 = (size - hb) / 16;
    for (uint32_t i = l; i < n16; i += NT) {
        const uint32_t q = (hb + 16 * i) >> 2;
        const uint32_t w0 = gp(sw)[q], w1 = gp(sw)[q + 1], w2 = gp(sw)[q + 2], w3 = gp(sw)[q + 3];
        uint4 v;
        if (sh == 0) {
            v = make_uint4(w0, w1, w2, w3);
        } else {
            const uint32_t w4 = word(q + 4);
            v = make_uint4(__builtin_amdgcn_alignbyte(w1, w0, sh), __builtin_amdgcn_alignbyte(w2, w1, sh),
                           __builtin_amdgcn_alignbyte(w3, w2, sh), __builtin_amdgcn_alignbyte(w4, w3, sh));
        }
        gst_u4(dq + i, v);
    }
    const uint32_t done = hb + 16 * n16;
    if ((uint32_t)l < size - done) gp(dst)[done + l] = gp(src)[done + l];
}

}  // namespace
}  // namespace fzh

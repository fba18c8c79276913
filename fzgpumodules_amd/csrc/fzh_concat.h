// fzh_concat.h — device-side concat of a pipeline's outputs and the host
// readback of small values, shared by the plain concat launch (fzh_core.hip)
// and the one fused with the RZE pack step (rze.hip).
#pragma once

#include <cstring>

#include "fzh_internal.h"

namespace fzh {
namespace {

struct ConcatArgs {
    fzh_concat_slot slot[FZH_CONCAT_MAX_SLOTS];
};
struct ReadbackArgs {
    fzh_readback_entry e[FZH_READBACK_MAX];
};

__device__ __forceinline__ uint64_t slot_size(const fzh_concat_slot& s) {
    uint64_t v = s.size;
    if (s.kind == 1) v = *reinterpret_cast<const uint64_t*>(s.size_ptr);
    else if (s.kind == 2) v = (uint64_t)*reinterpret_cast<const uint32_t*>(s.size_ptr) * s.elem;
    return v < s.cap ? v : s.cap;
}

// The fused RZE launch knows slot 0's size before its own pack workgroups
// have stored it: size0 replaces slot 0's device size, and readback words
// at ptr (a u64) read val.
struct SizeOverride {
    bool has_size0;
    uint64_t size0;
    uintptr_t ptr;
    uint64_t val;
};

// Concat and readback in one launch.  Every workgroup's first wave reads
// the slot sizes and forms the offsets (n <= 32 values); all workgroups copy
// the slots' 16-byte words grid-stride; workgroup 0 writes the header and
// the total and then serves the readback entries, taking any word that lies
// in the header or the total from the values it just computed (so the
// readback never depends on another workgroup's stores).
__device__ __forceinline__ void concat_readback_body(const ConcatArgs& a, int n, uint8_t* __restrict__ out,
                                                    uint64_t* __restrict__ total, const ReadbackArgs& r, int nr,
                                                    uint8_t* __restrict__ mapped, uint32_t blk, uint32_t nblk,
                                                    const SizeOverride& ov) {
    __shared__ uint64_t s_off[FZH_CONCAT_MAX_SLOTS], s_sz[FZH_CONCAT_MAX_SLOTS];
    __shared__ uint64_t s_total;
    const int t = threadIdx.x;
    const uint64_t hdr = (4 + 8 * (uint64_t)n + 15) & ~uint64_t(15);
    if (t < 64) {
        const uint64_t sz = t < n ? (t == 0 && ov.has_size0 ? ov.size0 : slot_size(a.slot[t])) : 0;
        const uint64_t padded = (sz + 15) & ~uint64_t(15);
        uint64_t incl = padded;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t v = __shfl_up(incl, off, 64);
            if (t >= off) incl += v;
        }
        if (t < n) {
            s_off[t] = hdr + incl - padded;
            s_sz[t] = sz;
        }
        if (t == 63) s_total = hdr + incl;
    }
    __syncthreads();
    const uint64_t G = (uint64_t)nblk * 256, g = (uint64_t)blk * 256 + t;
    for (int si = 0; si < n; si++) {
        const uint64_t sz = s_sz[si], nw = (sz + 15) / 16;
        const uint8_t* src = reinterpret_cast<const uint8_t*>(a.slot[si].src);
        uint8_t* d8 = out + s_off[si];
        if (reinterpret_cast<const void*>(d8) == reinterpret_cast<const void*>(src)) {
            if (g < nw * 16 - sz) d8[sz + g] = 0;  // produced in place: the pad only
            continue;
        }
        uint4* dst = reinterpret_cast<uint4*>(d8);
        const uint4* src4 = reinterpret_cast<const uint4*>(src);
        const uint64_t full = sz / 16;
        uint64_t w = g;
        // four words per step, loads issued before the stores
        for (; w + 3 * G < full; w += 4 * G) {
            const uint4 v0 = src4[w], v1 = src4[w + G], v2 = src4[w + 2 * G], v3 = src4[w + 3 * G];
            dst[w] = v0;
            dst[w + G] = v1;
            dst[w + 2 * G] = v2;
            dst[w + 3 * G] = v3;
        }
        for (; w < nw; w += G) {
            if (w < full) {
                dst[w] = src4[w];
            } else {
                uint32_t v[4] = {0, 0, 0, 0};
                for (uint64_t b = 16 * w; b < sz; b++) v[(b & 15) >> 2] |= (uint32_t)src[b] << (8 * (b & 3));
                dst[w] = make_uint4(v[0], v[1], v[2], v[3]);
            }
        }
    }
    if (blk != 0) return;
    // header word k: 0 = n, 1 + 2j / 2 + 2j = size j (lo / hi), then zero pad
    auto hword = [&](uint64_t k) -> uint32_t {
        if (k == 0) return (uint32_t)n;
        const uint64_t j = (k - 1) >> 1;
        if (j >= (uint64_t)n) return 0u;
        return (k & 1) ? (uint32_t)s_sz[j] : (uint32_t)(s_sz[j] >> 32);
    };
    for (uint64_t k = t; k < hdr / 4; k += 256) reinterpret_cast<uint32_t*>(out)[k] = hword(k);
    if (t == 0) *total = s_total;
    // the entries' words flattened over the workgroup's threads: one load
    // round trip for all of them instead of one per entry
    const uintptr_t o0 = reinterpret_cast<uintptr_t>(out), t0 = reinterpret_cast<uintptr_t>(total);
    uint32_t all = 0;
    for (int i = 0; i < nr; i++) all += r.e[i].bytes / 4;
    for (uint32_t g = t; g < all; g += 256) {
        int i = 0;
        uint32_t base = 0;
        while (base + r.e[i].bytes / 4 <= g) base += r.e[i++].bytes / 4;
        const uint32_t w = g - base;
        const uintptr_t at = reinterpret_cast<uintptr_t>(r.e[i].src) + 4ull * w;
        uint32_t v;
        if (at >= o0 && at < o0 + hdr) v = hword((at - o0) / 4);
        else if (at >= t0 && at < t0 + 8) v = at == t0 ? (uint32_t)s_total : (uint32_t)(s_total >> 32);
        else if (ov.ptr && at >= ov.ptr && at < ov.ptr + 8)
            v = at == ov.ptr ? (uint32_t)ov.val : (uint32_t)(ov.val >> 32);
        else v = *reinterpret_cast<const uint32_t*>(at);
        reinterpret_cast<uint32_t*>(mapped + r.e[i].dst_offset)[w] = v;
    }
}


// Validates and packs the concat slots and readback entries for a launch;
// grid: workgroups for the slots' copies.
inline int build_concat_args(const fzh_concat_slot* slots, int n, void* d_out, uint64_t out_cap, uint64_t* d_total,
                             const fzh_readback_entry* entries, int n_entries, void* d_mapped_dst, ConcatArgs& a,
                             ReadbackArgs& r, unsigned& grid) {
    if (n < 0 || n > FZH_CONCAT_MAX_SLOTS) return fail(FZH_E_INVALID, "concat: 0..32 slots");
    if (n_entries < 0 || n_entries > FZH_READBACK_MAX) return fail(FZH_E_INVALID, "readback: at most 32 entries");
    if (!d_out || !d_total) return fail(FZH_E_INVALID, "concat: null pointer");
    if (reinterpret_cast<uintptr_t>(d_out) % 16 != 0 || reinterpret_cast<uintptr_t>(d_total) % 8 != 0)
        return fail(FZH_E_INVALID, "concat: output must be 16-byte and the total 8-byte aligned");
    if (n_entries > 0 && (!d_mapped_dst || reinterpret_cast<uintptr_t>(d_mapped_dst) % 4 != 0))
        return fail(FZH_E_INVALID, "readback: destination must be 4-byte aligned");
    std::memset(&a, 0, sizeof(a));
    const uint64_t hdr = (4 + 8 * (uint64_t)n + 15) & ~uint64_t(15);
    uint64_t copy_words = 0, bound = hdr;
    for (int i = 0; i < n; i++) {
        const fzh_concat_slot& s = slots[i];
        if (s.kind > 2) return fail(FZH_E_INVALID, "concat: slot kind must be 0, 1 or 2");
        if (s.kind != 0 && !s.size_ptr) return fail(FZH_E_INVALID, "concat: null size pointer");
        if (s.cap > 0 && (!s.src || reinterpret_cast<uintptr_t>(s.src) % 16 != 0))
            return fail(FZH_E_INVALID, "concat: slot sources must be 16-byte aligned");
        if (s.kind == 0 && s.size > s.cap) return fail(FZH_E_INVALID, "concat: slot size exceeds its capacity");
        a.slot[i] = s;
        bound += (s.cap + 15) & ~uint64_t(15);
        const bool first_in_place = i == 0 && s.src == static_cast<const uint8_t*>(d_out) + hdr;
        if (!first_in_place) copy_words += (s.cap + 15) / 16;
    }
    if (bound > out_cap) return fail(FZH_E_CAPACITY, "concat: output buffer too small");
    std::memset(&r, 0, sizeof(r));
    for (int i = 0; i < n_entries; i++) {
        const fzh_readback_entry& e = entries[i];
        if (!e.src || e.bytes % 4 != 0 || e.dst_offset % 4 != 0 || reinterpret_cast<uintptr_t>(e.src) % 4 != 0)
            return fail(FZH_E_INVALID, "readback: entries must be non-null 4-byte multiples");
        r.e[i] = e;
    }
    grid = (unsigned)std::min<uint64_t>(std::max<uint64_t>(cdiv(copy_words, 1024), 1), 256);
    return FZH_OK;
}

}  // namespace
}  // namespace fzh

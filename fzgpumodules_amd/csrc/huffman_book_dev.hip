// huffman_book_dev.hip — the canonical Huffman codebook built on the device,
// byte-identical to the host builder (huffman_book.cpp), so a Huffman stage
// needs no host round trip between its histogram and its encoder and can be
// captured in a HIP graph.
//
// The book depends on the code lengths the reference's CPU builder assigns
// (modules/coders/huffman/phf/hf_bk_impl1.cc:55-163, hf_bk_internal.cc:58-101)
// and those depend on the tie order of its binary min-heap, so the heap is
// replayed exactly, one operation after another, by one lane with the heap
// in LDS (a wave first compacts the nonzero counts, so the lane's chain of
// dependent LDS accesses is about 3 log2(m) steps per merge for m symbols
// present): symbols with nonzero frequency pushed in symbol order (sift-up
// stops at a parent with frequency <=), then while two or more nodes
// remain: pop (last node to the root, sifted down past strictly smaller
// children, the right child taken only when strictly smaller than the left)
// twice, the first node popped becoming the right child, and push their
// sum.  A node's code length is its depth; the workgroup finds every leaf's
// depth by walking parent links in parallel, truncates lengths above 27 to
// 28 (hf_bk.cc:84-140 keeps 27-bit codes in a 32-bit word) and canonises as
// hf_canon.cc:86-131 does: per-length counts, first[l] =
// (first[l+1] + numl[l+1] + 1) >> 1 downwards from the longest length
// (the reference's (int)((a) / 2.0 + 0.5) for a >= 0), entry[] their
// prefix, and symbols of one length ranked in symbol order.
#include "fzh_internal.h"

namespace fzh {
namespace {

constexpr uint32_t DBOOK_MAX = 1024;  // bklen limit of the device build
constexpr int NODE_BITS = 20;         // heap entry = freq << 20 | node
constexpr uint64_t NODE_MASK = (1ull << NODE_BITS) - 1;

__device__ __forceinline__ uint64_t efreq(uint64_t e) { return e >> NODE_BITS; }

// Status codes (0 = ok): 1 = the histogram does not sum to n (symbols out of
// range), 2 = a code longer than 58 bits (the host builder's error; a tree
// that deep needs a total count above fib(61), about 2.5e12).
__global__ void __launch_bounds__(256)
huf_book_kernel(const uint32_t* __restrict__ freq, uint32_t bklen, int eb, uint64_t n_expected,
                uint32_t* __restrict__ bk4, uint8_t* __restrict__ revbook, int* __restrict__ status) {
    __shared__ __align__(16) uint64_t q[2 * DBOOK_MAX + 2];  // the heap, 1-indexed
    __shared__ uint32_t fq[DBOOK_MAX];           // the histogram
    __shared__ uint32_t lfq[DBOOK_MAX];          // leaf node -> frequency (nonzero counts in symbol order)
    __shared__ uint32_t parent[2 * DBOOK_MAX];   // node -> parent node
    __shared__ uint16_t sym_of[DBOOK_MAX];       // leaf node -> symbol
    __shared__ uint8_t len[DBOOK_MAX];           // symbol -> code length (255: no code)
    __shared__ uint32_t rank[DBOOK_MAX];         // symbol -> rank inside its length
    __shared__ int numl[32], first[32], entry[32];
    __shared__ uint32_t base[32];
    __shared__ uint64_t s_sum[4];
    __shared__ uint32_t s_nl, s_bad;
    const int t = threadIdx.x;
    const int l = lane_id();
    // -- histogram total (the out-of-range check of the host path)
    uint64_t sum = 0;
    for (uint32_t i = t; i < bklen; i += 256) {
        const uint32_t f = freq[i];
        fq[i] = f;
        sum += f;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if (l == 0) s_sum[t >> 6] = sum;
    if (t < 32) { numl[t] = 0; first[t] = 0; entry[t] = 0; base[t] = 0; }
    if (t == 0) s_bad = 0;
    __syncthreads();
    // -- leaves: the nonzero counts compacted in symbol order (one wave)
    if (t < 64) {
        uint32_t nz = 0;
        for (uint32_t c0 = 0; c0 < bklen; c0 += 64) {
            const uint32_t s = c0 + (uint32_t)l;
            const uint32_t f = s < bklen ? fq[s] : 0u;
            const uint64_t m = __ballot(f != 0);
            if (f) {
                const uint32_t k = nz + lanes_below(m);
                sym_of[k] = (uint16_t)s;
                lfq[k] = f;
            }
            nz += (uint32_t)__popcll(m);
        }
        if (l == 0) s_nl = nz;
    }
    __syncthreads();
    // -- the heap replay (one lane)
    if (t == 0) {
        const uint32_t nl = s_nl;
        uint32_t end = 1;
        for (uint32_t k = 0; k < nl; k++) {
            const uint32_t f = lfq[k];
            const uint64_t e = ((uint64_t)f << NODE_BITS) | k;
            uint32_t i = end++, j;
            while ((j = i >> 1) != 0) {
                const uint64_t p = q[j];
                if (efreq(p) <= (uint64_t)f) break;
                q[i] = p;
                i = j;
            }
            q[i] = e;
        }
        uint32_t next = nl;  // internal nodes follow the leaves
        auto pop = [&]() -> uint64_t {
            const uint64_t top = q[1];
            end--;
            const uint64_t e = q[end];
            const uint64_t fe = efreq(e);
            uint32_t i = 1, c;
            while ((c = i << 1) < end) {
                uint64_t ch;
                if (c + 1 < end) {  // both children in one 16-byte LDS read (c is even)
                    const ulonglong2 pr = *reinterpret_cast<const ulonglong2*>(&q[c]);
                    ch = pr.x;
                    if (efreq(pr.y) < efreq(pr.x)) { ch = pr.y; c++; }
                } else {
                    ch = q[c];
                }
                if (fe > efreq(ch)) {
                    q[i] = ch;
                    i = c;
                } else {
                    break;
                }
            }
            q[i] = e;
            return top;
        };
        while (end > 2) {
            const uint64_t right = pop();
            const uint64_t left = pop();
            const uint32_t node = next++;
            parent[(uint32_t)(left & NODE_MASK)] = node;
            parent[(uint32_t)(right & NODE_MASK)] = node;
            const uint64_t f = efreq(left) + efreq(right);
            const uint64_t e = (f << NODE_BITS) | node;
            uint32_t i = end++, j;
            while ((j = i >> 1) != 0) {
                const uint64_t p = q[j];
                if (efreq(p) <= f) break;
                q[i] = p;
                i = j;
            }
            q[i] = e;
        }
        if (s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3] != n_expected) s_bad = 1;
    }
    __syncthreads();
    const uint32_t nl = s_nl;
    const uint32_t root = nl > 0 ? 2 * nl - 2 : 0;  // the last node made (the single leaf when nl == 1)
    // -- code lengths: depth of every leaf
    for (uint32_t s = t; s < bklen; s += 256) len[s] = 255;
    __syncthreads();
    for (uint32_t leaf = t; leaf < nl; leaf += 256) {
        uint32_t d = 0, v = leaf;
        while (v != root && d <= 58) {
            v = parent[v];
            d++;
        }
        if (d > 58) s_bad = 2;
        len[sym_of[leaf]] = (uint8_t)(d > 27 ? 28 : d);
    }
    __syncthreads();
    // -- per-length counts, ranks in symbol order (one wave, 64 symbols at a time)
    if (t < 64) {
        for (uint32_t c0 = 0; c0 < bklen; c0 += 64) {
            const uint32_t s = c0 + (uint32_t)l;
            const int ln = s < bklen ? (int)len[s] : 255;
            uint64_t pend = __ballot(ln != 255);
            while (pend) {
                const int src = __builtin_ctzll(pend);
                const int lv = __builtin_amdgcn_readlane(ln, src);
                const uint64_t m = __ballot(ln == lv);
                const uint32_t b0 = base[lv];
                if (ln == lv) rank[s] = b0 + lanes_below(m);
                wave_lds_sync();
                if (l == 0) base[lv] = b0 + (uint32_t)__popcll(m);
                wave_lds_sync();
                pend &= ~m;
            }
        }
        wave_lds_sync();
        if (l < 32) numl[l] = (int)base[l];
        wave_lds_sync();
        if (l == 0) {
            int max_l = 0;
            for (int i = 0; i < 32; i++)
                if (numl[i] && i > max_l) max_l = i;
            for (int i = 1; i < 32; i++) entry[i] = entry[i - 1] + numl[i - 1];
            first[max_l] = 0;
            for (int i = max_l - 1; i >= 1; i--) first[i] = (first[i + 1] + numl[i + 1] + 1) >> 1;
            first[0] = 0xff;
        }
    }
    __syncthreads();
    // -- the book, the reverse book (first[32], entry[32], keys[bklen])
    uint8_t* keys = revbook + 256;
    for (uint32_t s = t; s < bklen; s += 256) {
        // keys past the last code stay 0
        if (eb == 1) keys[s] = 0;
        else if (eb == 2) reinterpret_cast<uint16_t*>(keys)[s] = 0;
        else reinterpret_cast<uint32_t*>(keys)[s] = 0;
    }
    if (t < 32) {
        reinterpret_cast<int*>(revbook)[t] = first[t];
        reinterpret_cast<int*>(revbook + 128)[t] = entry[t];
    }
    __syncthreads();
    for (uint32_t s = t; s < bklen; s += 256) {
        const int ln = len[s];
        if (ln == 255) {
            bk4[s] = 0xFFFFFFFFu;
            continue;
        }
        const uint32_t v = (uint32_t)first[ln] + rank[s];
        bk4[s] = (v & 0x07FFFFFFu) | ((uint32_t)ln << 27);
        const uint32_t p = (uint32_t)entry[ln] + rank[s];
        if (eb == 1) keys[p] = (uint8_t)s;
        else if (eb == 2) reinterpret_cast<uint16_t*>(keys)[p] = (uint16_t)s;
        else reinterpret_cast<uint32_t*>(keys)[p] = s;
    }
    if (t == 0) *status = (int)s_bad;
}

}  // namespace
}  // namespace fzh

using namespace fzh;

extern "C" {

int fzh_huffman_build_book_device(void* stream, const uint32_t* d_freq, uint32_t bklen, int elem_bytes,
                                  uint64_t n_expected, uint32_t* d_bk4, uint8_t* d_revbook, int* d_status) {
    if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4)
        return fail(FZH_E_INVALID, "huffman device book: bad elem_bytes");
    if (bklen == 0 || bklen > DBOOK_MAX)
        return fail(FZH_E_INVALID, "huffman device book: bklen must be in [1, 1024]");
    if (elem_bytes == 1 && bklen > 256)
        return fail(FZH_E_INVALID, "huffman device book: bklen above 256 with 1-byte symbols");
    if (!d_freq || !d_bk4 || !d_revbook || !d_status) return fail(FZH_E_INVALID, "huffman device book: null pointer");
    launch_k("huf_book_kernel", huf_book_kernel, dim3(1), dim3(256), 0, S(stream), d_freq, bklen, elem_bytes,
             n_expected, d_bk4, d_revbook, d_status);
    FZH_CHECK_LAUNCH("huf_book_kernel");
    return FZH_OK;
}

}  // extern "C"

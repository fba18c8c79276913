// huffman_book.cpp — canonical Huffman codebook on the host.
//
// Same algorithm and tie-breaking as the reference's CPU builder
// (modules/coders/huffman/phf/hf_bk.cc:84-140 with hf_bk_impl1.cc:55-163,
// hf_bk_internal.cc:58-101, hf_canon.cc:86-131):
//   1. binary min-heap (1-indexed) of symbols with nonzero frequency, inserted
//      in symbol order; insert breaks ties with <=, remove picks the smaller
//      child with <;
//   2. repeatedly remove two nodes; the first removed becomes the right child
//      and the second the left (the reference passes both removals as
//      arguments of one call and its host compiler evaluates the last
//      argument first);
//   3. in-order traversal, left edge 0, right edge 1, codes of at most 58
//      bits in a 64-bit word (bitcount in bits 58..63);
//   4. truncation to a 32-bit word: bitcount in bits 27..31, codes longer than
//      27 bits become bitcount 28 / prefix 0;
//   5. canonization: longest codes numbered from 0, first[l] =
//      (int)((first[l+1] + numl[l+1]) / 2.0 + 0.5), symbols of equal length in
//      symbol order.
// Output: bk4[bklen] and revbook = first[32] i32, entry[32] i32, keys[bklen].
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "fzh_internal.h"

namespace fzh {
namespace {

struct Node {
    Node* left = nullptr;
    Node* right = nullptr;
    uint64_t freq = 0;
    uint32_t symbol = 0;
};

class MinHeap {
  public:
    explicit MinHeap(size_t cap) : q_(cap + 2, nullptr) {}
    int size_plus_one() const { return end_; }
    Node* top() const { return q_[1]; }
    void push(Node* n) {
        int i = end_++, j;
        while ((j = i >> 1) != 0) {
            if (q_[j]->freq <= n->freq) break;
            q_[i] = q_[j];
            i = j;
        }
        q_[i] = n;
    }
    Node* pop() {
        if (end_ < 2) return nullptr;
        Node* n = q_[1];
        end_--;
        q_[1] = q_[end_];
        int i = 1, c;
        while ((c = i << 1) < end_) {
            if (c + 1 < end_ && q_[c + 1]->freq < q_[c]->freq) c++;
            if (q_[i]->freq > q_[c]->freq) {
                std::swap(q_[i], q_[c]);
                i = c;
            } else {
                break;
            }
        }
        return n;
    }

  private:
    std::vector<Node*> q_;
    int end_ = 1;
};

constexpr uint64_t kNoCode = ~0ull;

void tree_codes(const uint32_t* freq, uint32_t bklen, std::vector<uint64_t>& book) {
    book.assign(bklen, kNoCode);
    std::vector<Node> pool(2 * (size_t)bklen + 2);
    size_t used = 0;
    MinHeap heap(2 * (size_t)bklen + 2);
    for (uint32_t s = 0; s < bklen; s++) {
        if (!freq[s]) continue;
        Node* n = &pool[used++];
        n->freq = freq[s];
        n->symbol = s;
        heap.push(n);
    }
    while (heap.size_plus_one() > 2) {
        Node* right = heap.pop();
        Node* left = heap.pop();
        Node* n = &pool[used++];
        n->left = left;
        n->right = right;
        n->freq = left->freq + right->freq;
        heap.push(n);
    }
    if (heap.size_plus_one() < 2) return;
    struct Frame {
        Node* node;
        uint64_t code, len;
    };
    std::vector<Frame> stack;
    Node* p = heap.top();
    uint64_t code = 0, len = 0;
    while (p) {
        if (p->left || p->right) {
            if (stack.size() + 1 > 58) throw std::runtime_error("huffman: code tree deeper than 58");
            stack.push_back({p, code, len});
            p = p->left;
            code <<= 1;
            if (++len > 58) throw std::runtime_error("huffman: code longer than 58 bits");
        } else {
            book[p->symbol] = (code & ((1ull << 58) - 1)) | (len << 58);
            if (stack.empty()) break;
            Frame f = stack.back();
            stack.pop_back();
            p = f.node->right;
            code = (f.code << 1) | 1u;
            len = f.len + 1;
            if (len > 58) throw std::runtime_error("huffman: code longer than 58 bits");
        }
    }
}

void store_key(uint8_t* keys, uint32_t i, uint32_t v, int eb) {
    switch (eb) {
    case 1: keys[i] = (uint8_t)v; break;
    case 2: reinterpret_cast<uint16_t*>(keys)[i] = (uint16_t)v; break;
    default: reinterpret_cast<uint32_t*>(keys)[i] = v; break;
    }
}

}  // namespace
}  // namespace fzh

extern "C" {

size_t fzh_huffman_revbook_bytes(uint32_t bklen, int elem_bytes) { return 256 + (size_t)elem_bytes * bklen; }

int fzh_huffman_build_book(const uint32_t* h_freq, uint32_t bklen, int elem_bytes, uint32_t* h_bk4,
                           uint8_t* h_revbook) {
    using namespace fzh;
    if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4) return fail(FZH_E_INVALID, "huffman: bad elem_bytes");
    if (bklen == 0 || bklen > 65536) return fail(FZH_E_INVALID, "huffman: bklen must be in [1, 65536]");
    std::vector<uint64_t> bk8;
    try {
        tree_codes(h_freq, bklen, bk8);
    } catch (const std::exception& e) {
        return fail(FZH_E_INVALID, e.what());
    }
    std::vector<uint32_t> bk4(bklen);
    for (uint32_t i = 0; i < bklen; i++) {
        if (bk8[i] == kNoCode) { bk4[i] = 0xFFFFFFFFu; continue; }
        const uint32_t bc = (uint32_t)(bk8[i] >> 58);
        const uint32_t pc = (uint32_t)(bk8[i] & ((1ull << 58) - 1));
        bk4[i] = bc > 27 ? (28u << 27) : ((bc << 27) | (pc & 0x07FFFFFFu));
    }
    int numl[32] = {0}, iterby[32] = {0}, first[32] = {0}, entry[32] = {0};
    int max_l = 0;
    for (uint32_t i = 0; i < bklen; i++) {
        if (bk4[i] == 0xFFFFFFFFu) continue;
        const int l = (int)(bk4[i] >> 27);
        max_l = l > max_l ? l : max_l;
        numl[l]++;
    }
    for (int i = 1; i < 32; i++) entry[i] = entry[i - 1] + numl[i - 1];
    for (int i = 0; i < 32; i++) iterby[i] = entry[i];
    first[max_l] = 0;
    for (int l = max_l - 1; l >= 1; l--) first[l] = (int)((first[l + 1] + numl[l + 1]) / 2.0 + 0.5);
    first[0] = 0xff;
    std::vector<uint32_t> canon(bklen, 0xFFFFFFFFu), keys(bklen, 0);
    for (uint32_t i = 0; i < bklen; i++) {
        if (bk4[i] == 0xFFFFFFFFu) continue;
        const int l = (int)(bk4[i] >> 27);
        const uint32_t v = (uint32_t)(first[l] + iterby[l] - entry[l]);
        canon[iterby[l]] = (v & 0x07FFFFFFu) | ((uint32_t)l << 27);
        keys[iterby[l]] = i;
        iterby[l]++;
    }
    for (uint32_t i = 0; i < bklen; i++) h_bk4[i] = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < bklen; i++)
        if (canon[i] != 0xFFFFFFFFu) h_bk4[keys[i]] = canon[i];
    std::memcpy(h_revbook, first, 128);
    std::memcpy(h_revbook + 128, entry, 128);
    for (uint32_t i = 0; i < bklen; i++) store_key(h_revbook + 256, i, keys[i], elem_bytes);
    return FZH_OK;
}

}  // extern "C"

// C entry points onto the reference's host-only Huffman codebook builder
// (modules/coders/huffman/phf/hf_bk.cc:84-140), so ctypes can call the
// reference's own template instantiations.  Test infrastructure only.
#include <cstdint>
#include "hf_impl.hh"

extern "C" void ref_phf_book(const uint32_t* freq, int bklen, int elem_bytes,
                             uint32_t* bk4, uint8_t* revbook) {
    uint32_t* f = const_cast<uint32_t*>(freq);
    switch (elem_bytes) {
    case 1: phf_CPU_build_canonized_codebook_v2<uint8_t, uint32_t>(f, bklen, bk4, revbook, nullptr); break;
    case 2: phf_CPU_build_canonized_codebook_v2<uint16_t, uint32_t>(f, bklen, bk4, revbook, nullptr); break;
    default: phf_CPU_build_canonized_codebook_v2<uint32_t, uint32_t>(f, bklen, bk4, revbook, nullptr); break;
    }
}

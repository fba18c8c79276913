// fzh_bitplane.h — in-register bit-plane transposes for 32-element groups
// (the Bitshuffle ballot layout for 1- and 2-byte elements,
// modules/shufflers/bitshuffle/bitshuffle_stage.cu:299-349): for a group of
// 32 elements E_0..E_31, plane word W_b holds bit b of E_j at bit j.
// Shared by the Bitshuffle kernels and the fused Bitshuffle+RZE kernels.
#pragma once

#include "fzh_internal.h"

namespace fzh {
namespace {

// In-register transpose step on rows A[k], A[k+s] (LSB-first columns):
// swaps bits (row k, col i+s) <-> (row k+s, col i) for i with bit s clear.
template <int S, int N>
__device__ __forceinline__ void transpose_step(uint32_t (&A)[N]) {
    constexpr uint32_t M = S == 16 ? 0x0000FFFFu : S == 8 ? 0x00FF00FFu : S == 4 ? 0x0F0F0F0Fu
                         : S == 2 ? 0x33333333u : 0x55555555u;
#pragma unroll
    for (int k = 0; k < N; k++) {
        if ((k & S) == 0 && k + S < N) {
            const uint32_t t = ((A[k] >> S) ^ A[k + S]) & M;
            A[k] ^= t << S;
            A[k + S] ^= t;
        }
    }
}

// 2-byte elements: w[16] = the group's 64 bytes as little-endian words
// (word k holds E_2k | E_2k+1 << 16) -> A[b] = plane word for bit b.
// Packed rows P[k] = E_k | E_{k+16}<<16 are the 16-step of the 32x32
// transpose with the upper 16 columns zero; steps 8..1 finish it.
__device__ __forceinline__ void planes16(const uint32_t (&w)[16], uint32_t (&A)[16]) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
        A[2 * i] = (w[i] & 0xFFFFu) | (w[8 + i] << 16);
        A[2 * i + 1] = (w[i] >> 16) | (w[8 + i] & 0xFFFF0000u);
    }
    transpose_step<8>(A);
    transpose_step<4>(A);
    transpose_step<2>(A);
    transpose_step<1>(A);
}

__device__ __forceinline__ void unplanes16(uint32_t (&A)[16], uint32_t (&w)[16]) {
    transpose_step<8>(A);
    transpose_step<4>(A);
    transpose_step<2>(A);
    transpose_step<1>(A);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        w[i] = (A[2 * i] & 0xFFFFu) | (A[2 * i + 1] << 16);
        w[8 + i] = (A[2 * i] >> 16) | (A[2 * i + 1] & 0xFFFF0000u);
    }
}

// 1-byte elements: w[8] = the group's 32 bytes (byte j in w[j/4], bits
// 8*(j%4)); P[k] = E_k | E_{k+8}<<8 | E_{k+16}<<16 | E_{k+24}<<24.
__device__ __forceinline__ void planes8(const uint32_t (&w)[8], uint32_t (&A)[8]) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int sh = 8 * (k & 3);
        const int wi = k >> 2;
        A[k] = ((w[wi] >> sh) & 0xFFu) | (((w[wi + 2] >> sh) & 0xFFu) << 8) |
               (((w[wi + 4] >> sh) & 0xFFu) << 16) | (((w[wi + 6] >> sh) & 0xFFu) << 24);
    }
    transpose_step<4>(A);
    transpose_step<2>(A);
    transpose_step<1>(A);
}

__device__ __forceinline__ void unplanes8(uint32_t (&A)[8], uint32_t (&w)[8]) {
    transpose_step<4>(A);
    transpose_step<2>(A);
    transpose_step<1>(A);
#pragma unroll
    for (int k = 0; k < 8; k++) w[k] = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int sh = 8 * (k & 3);
        const int wi = k >> 2;
        w[wi] |= (A[k] & 0xFFu) << sh;
        w[wi + 2] |= ((A[k] >> 8) & 0xFFu) << sh;
        w[wi + 4] |= ((A[k] >> 16) & 0xFFu) << sh;
        w[wi + 6] |= ((A[k] >> 24) & 0xFFu) << sh;
    }
}

// 16-byte global-memory load / store
__device__ __forceinline__ uint4 ld16(const uint8_t* p) { return gld_u4(p); }
__device__ __forceinline__ void st16(uint8_t* p, uint4 v) { gst_u4(p, v); }

}  // namespace
}  // namespace fzh

// fzh_synth.h -- seeded synthetic fields on the device (benchmark and
// training inputs; no counterpart in the reference, whose tests generate
// host data).
#pragma once

#include "fzh_internal.h"

namespace fzh {
namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Uniform in [0, 1) and in (0, 1) from 24 bits of a hash.
__device__ __forceinline__ float u01(uint64_t h) { return (float)(uint32_t)(h & 0xFFFFFFu) * (1.0f / 16777216.0f); }
__device__ __forceinline__ float u01_open(uint64_t h) {
    return ((float)(uint32_t)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
}

// Standard normal from one 64-bit hash (Box-Muller on its two 24-bit halves).
__device__ __forceinline__ float hash_normal(uint64_t h) {
    return sqrtf(-2.0f * __logf(u01_open(h))) * __cosf(6.28318531f * u01(h));
}

__device__ __forceinline__ float lattice_normal(uint64_t seed, uint32_t o, uint64_t i, uint64_t j, uint64_t k) {
    return hash_normal(mix64(seed ^ mix64(((uint64_t)o << 56) ^ (i * 0xD1B54A32D192ED03ull) ^
                                          (j * 0xABC98388FB8FAC03ull) ^ (k * 0x8CB92BA72F3D8DD7ull))));
}

__device__ __forceinline__ float quintic(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }

// Gaussian random field with spectrum ~ k^expo at point (x, y, z): a sum of
// octaves o = 0 .. O-1 of lattice noise -- iid standard-normal values on a
// lattice of spacing 2^o, blended between the 2^d surrounding lattice points
// with the quintic weight t^3 (t (6t - 15) + 10) -- with octave amplitude
// 2^(ea o), ea = -(expo + d) / 2 (a power law P(k) ~ k^expo puts variance
// ~ k^(expo + d) in the octave of wavenumbers around 2^-o), the largest
// amplitude scaled to 1.  A linear combination of normal variables, so the
// field is Gaussian.
__device__ __forceinline__ float power_law_value(uint64_t x, uint64_t y, uint64_t z, int d, int octaves, float ea,
                                                 uint64_t seed) {
    float acc = 0.0f;
    for (int o = 0; o < octaves; o++) {
        const float amp = exp2f(ea * (float)(ea >= 0.0f ? o - (octaves - 1) : o));
        const uint64_t ix = x >> o, iy = y >> o, iz = z >> o;
        const float inv = 1.0f / (float)(1u << o);
        const float tx = quintic((float)(x & ((1u << o) - 1u)) * inv);
        const float ty = quintic((float)(y & ((1u << o) - 1u)) * inv);
        const float tz = quintic((float)(z & ((1u << o) - 1u)) * inv);
        float v;
        if (d == 1) {
            v = lattice_normal(seed, o, ix, 0, 0) * (1.0f - tx) + lattice_normal(seed, o, ix + 1, 0, 0) * tx;
        } else if (d == 2) {
            const float a = lattice_normal(seed, o, ix, iy, 0) * (1.0f - tx) + lattice_normal(seed, o, ix + 1, iy, 0) * tx;
            const float b = lattice_normal(seed, o, ix, iy + 1, 0) * (1.0f - tx) + lattice_normal(seed, o, ix + 1, iy + 1, 0) * tx;
            v = a * (1.0f - ty) + b * ty;
        } else {
            float c[2];
            for (int dz = 0; dz < 2; dz++) {
                const float a = lattice_normal(seed, o, ix, iy, iz + dz) * (1.0f - tx) +
                                lattice_normal(seed, o, ix + 1, iy, iz + dz) * tx;
                const float b = lattice_normal(seed, o, ix, iy + 1, iz + dz) * (1.0f - tx) +
                                lattice_normal(seed, o, ix + 1, iy + 1, iz + dz) * tx;
                c[dz] = a * (1.0f - ty) + b * ty;
            }
            v = c[0] * (1.0f - tz) + c[1] * tz;
        }
        acc += amp * v;
    }
    return acc;
}

// Octaves for a field whose longest side is m.
__host__ __device__ inline int power_law_octaves(uint64_t m) {
    int o = 1;
    while ((1ull << o) < m && o < 24) o++;
    return o;
}

}  // namespace
}  // namespace fzh

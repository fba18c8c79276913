// select.hip — learned chain selector: per-chunk features and a small MLP
// trained and run on the device (gfx950).  No counterpart in the reference.
//
// Features (chunk of E = nx*ny*nz values, absolute bound eb; pairs_x =
// (nx-1) ny nz etc.; "res" is the Lorenzo residual of the raw values over
// the whole chunk, zero outside it; l = |res| / (2 eb)):
//    0 log2(1 + std / eb)              16..23  (continued)
//    1 log2(1 + (max - min) / eb)       9..23 fraction of l in magnitude bins
//    2 fraction of exact zeros                bin 0: l < 0.5, bin k >= 1:
//    3 fraction of integer values             2^(k-1) <= l + 0.5 < 2^k (k <= 14)
//    4 fraction x == west neighbour    24 fraction |x - west| / (2 eb) < 0.5
//    5 log2(1 + mean |dx| / eb)        25 log2(1 + mean |x| / eb)
//    6 log2(1 + mean |dy| / eb)        26 log2(1 + max |x| / eb)
//    7 log2(1 + mean |dz| / eb) (0 2D) 27 fraction l >= 511.5
//    8 log2(1 + mean l)                28 fraction |x| / (2 eb) >= 32767.5
//                                      29 1 for 3D chunks, else 0
//                                      30 entropy (bits) of the 15 bins
//                                      31 fraction x > 0
// One workgroup per chunk; per-thread partials in a fixed order, then a
// fixed reduction tree, so the features are deterministic.
//
// MLP: 32 -> 64 -> 64 -> 4, ReLU.  Each layer reads BF16 inputs (features
// after normalisation, hidden activations, and the weights, all rounded to
// nearest-even BF16) and accumulates in FP32 in index order; biases and the
// master weights are FP32.  Backward rounds the propagated deltas to BF16
// before they enter a product, and weight gradients are FP32 sums over the
// batch in sample order.
#include <algorithm>
#include <cmath>

#include "fzh_internal.h"
#include "fzh_synth.h"
#include "fzs.h"

namespace fzh {
namespace sel {
namespace {

constexpr int NF = FZS_NFEAT, H1 = FZS_H1, H2 = FZS_H2, NC = FZS_NCHAINS;
constexpr int OW1 = 0, OB1 = OW1 + H1 * NF, OW2 = OB1 + H1, OB2 = OW2 + H2 * H1, OW3 = OB2 + H2, OB3 = OW3 + NC * H2;
constexpr int NP = OB3 + NC;  // 6532

__device__ __forceinline__ float bf16r(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (u & 0x007FFFFFu) ? __uint_as_float(u | 0x00400000u) : f;
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xFFFF0000u);
}

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
struct Acc {
    double sx, sx2, sdx, sdy, sdz, sl, sabs;
    float mn, mx, mxabs;
    uint32_t c[7];  // zero, int, rep, smooth, out512, qout, pos
    uint32_t h[15];
};

__device__ __forceinline__ double wsum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ float wmin_f(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ float wmax_f(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

__global__ void __launch_bounds__(256)
features_kernel(const float* __restrict__ chunks, uint32_t nx, uint32_t ny, uint32_t nz,
                const float* __restrict__ d_eb, float* __restrict__ feat) {
    __shared__ double s_d[4][7];
    __shared__ float s_f[4][3];
    __shared__ uint32_t s_c[4][22];
    const uint32_t c = blockIdx.x;
    const uint64_t E = (uint64_t)nx * ny * nz;
    const float* x = chunks + c * E;
    const float eb = d_eb[c];
    const float inv2eb = 0.5f / eb;
    const uint64_t sxy = (uint64_t)nx * ny;
    Acc a;
    a.sx = a.sx2 = a.sdx = a.sdy = a.sdz = a.sl = a.sabs = 0.0;
    a.mn = INFINITY;
    a.mx = -INFINITY;
    a.mxabs = 0.0f;
#pragma unroll
    for (int k = 0; k < 7; k++) a.c[k] = 0;
#pragma unroll
    for (int k = 0; k < 15; k++) a.h[k] = 0;
    for (uint64_t i = threadIdx.x; i < E; i += 256) {
        const uint32_t ix = (uint32_t)(i % nx);
        const uint32_t iy = (uint32_t)((i / nx) % ny);
        const uint32_t iz = (uint32_t)(i / sxy);
        const float v = x[i];
        const float w = ix ? x[i - 1] : 0.0f;
        const float nn = iy ? x[i - nx] : 0.0f;
        const float wn = (ix && iy) ? x[i - 1 - nx] : 0.0f;
        float res;
        if (nz > 1) {
            const float b = iz ? x[i - sxy] : 0.0f;
            const float wb = (ix && iz) ? x[i - 1 - sxy] : 0.0f;
            const float nb = (iy && iz) ? x[i - nx - sxy] : 0.0f;
            const float wnb = (ix && iy && iz) ? x[i - 1 - nx - sxy] : 0.0f;
            res = v - w - nn - b + wn + wb + nb - wnb;
            if (iz) a.sdz += fabs((double)v - (double)b);
        } else {
            res = v - w - nn + wn;
        }
        a.sx += v;
        a.sx2 += (double)v * v;
        a.mn = fminf(a.mn, v);
        a.mx = fmaxf(a.mx, v);
        a.mxabs = fmaxf(a.mxabs, fabsf(v));
        a.sabs += fabsf(v);
        a.c[0] += v == 0.0f;
        a.c[1] += v == rintf(v);
        a.c[6] += v > 0.0f;
        if (ix) {
            a.c[2] += v == w;
            const double d = fabs((double)v - (double)w);
            a.sdx += d;
            a.c[3] += d * inv2eb < 0.5;
        }
        if (iy) a.sdy += fabs((double)v - (double)nn);
        const float l = fabsf(res) * inv2eb;
        a.sl += l;
        const float lp = l + 0.5f;
        int bin;
        if (!(lp >= 1.0f)) bin = 0;
        else if (lp >= 16384.0f) bin = 14;
        else bin = 32 - __clz((uint32_t)lp);
        bin = bin > 14 ? 14 : bin;
#pragma unroll
        for (int k = 0; k < 15; k++) a.h[k] += bin == k;  // no dynamic register indexing
        a.c[4] += l >= 511.5f;
        a.c[5] += fabsf(v) * inv2eb >= 32767.5f;
    }
    const int wv = threadIdx.x >> 6, l = lane_id();
    double d[7] = {a.sx, a.sx2, a.sdx, a.sdy, a.sdz, a.sl, a.sabs};
#pragma unroll
    for (int k = 0; k < 7; k++) d[k] = wsum_d(d[k]);
    const float fmn = wmin_f(a.mn), fmx = wmax_f(a.mx), fma = wmax_f(a.mxabs);
    uint32_t cc[22];
#pragma unroll
    for (int k = 0; k < 7; k++) cc[k] = wave_sum(a.c[k]);
#pragma unroll
    for (int k = 0; k < 15; k++) cc[7 + k] = wave_sum(a.h[k]);
    if (l == 0) {
        for (int k = 0; k < 7; k++) s_d[wv][k] = d[k];
        s_f[wv][0] = fmn;
        s_f[wv][1] = fmx;
        s_f[wv][2] = fma;
        for (int k = 0; k < 22; k++) s_c[wv][k] = cc[k];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double D[7];
    for (int k = 0; k < 7; k++) D[k] = ((s_d[0][k] + s_d[1][k]) + s_d[2][k]) + s_d[3][k];
    float mn = s_f[0][0], mx = s_f[0][1], mxa = s_f[0][2];
    for (int j = 1; j < 4; j++) {
        mn = fminf(mn, s_f[j][0]);
        mx = fmaxf(mx, s_f[j][1]);
        mxa = fmaxf(mxa, s_f[j][2]);
    }
    uint32_t C[22];
    for (int k = 0; k < 22; k++) C[k] = s_c[0][k] + s_c[1][k] + s_c[2][k] + s_c[3][k];
    const double En = (double)E;
    const double px = (double)(nx - 1) * ny * nz, py = (double)nx * (ny - 1) * nz, pz = (double)nx * ny * (nz - 1);
    const double e = (double)eb;
    const double mean = D[0] / En;
    const double var = fmax(D[1] / En - mean * mean, 0.0);
    float* f = feat + (uint64_t)c * NF;
    f[0] = (float)log2(1.0 + sqrt(var) / e);
    f[1] = (float)log2(1.0 + ((double)mx - (double)mn) / e);
    f[2] = (float)(C[0] / En);
    f[3] = (float)(C[1] / En);
    f[4] = px > 0 ? (float)(C[2] / px) : 0.0f;
    f[5] = px > 0 ? (float)log2(1.0 + D[2] / px / e) : 0.0f;
    f[6] = py > 0 ? (float)log2(1.0 + D[3] / py / e) : 0.0f;
    f[7] = (nz > 1 && pz > 0) ? (float)log2(1.0 + D[4] / pz / e) : 0.0f;
    f[8] = (float)log2(1.0 + D[5] / En);
    double ent = 0.0;
    for (int k = 0; k < 15; k++) {
        const double p = C[7 + k] / En;
        f[9 + k] = (float)p;
        if (p > 0) ent -= p * log2(p);
    }
    f[24] = px > 0 ? (float)(C[3] / px) : 0.0f;
    f[25] = (float)log2(1.0 + D[6] / En / e);
    f[26] = (float)log2(1.0 + (double)mxa / e);
    f[27] = (float)(C[4] / En);
    f[28] = (float)(C[5] / En);
    f[29] = nz > 1 ? 1.0f : 0.0f;
    f[30] = (float)ent;
    f[31] = (float)(C[6] / En);
}

// ---------------------------------------------------------------------------
// MLP
// ---------------------------------------------------------------------------
struct Lds {
    float w1[H1 * NF];
    float b1[H1];
    float w2[H2 * H1];
    float b2[H2];
    float w3[NC * H2];
    float b3[NC];
    float shift[NF];
    float scale[NF];
};

__device__ void load_params(Lds& s, const float* __restrict__ p, const float* __restrict__ norm) {
    for (int i = threadIdx.x; i < H1 * NF; i += blockDim.x) s.w1[i] = bf16r(p[OW1 + i]);
    for (int i = threadIdx.x; i < H2 * H1; i += blockDim.x) s.w2[i] = bf16r(p[OW2 + i]);
    for (int i = threadIdx.x; i < NC * H2; i += blockDim.x) s.w3[i] = bf16r(p[OW3 + i]);
    for (int i = threadIdx.x; i < H1; i += blockDim.x) s.b1[i] = p[OB1 + i];
    for (int i = threadIdx.x; i < H2; i += blockDim.x) s.b2[i] = p[OB2 + i];
    for (int i = threadIdx.x; i < NC; i += blockDim.x) s.b3[i] = p[OB3 + i];
    for (int i = threadIdx.x; i < NF; i += blockDim.x) {
        s.shift[i] = norm[i];
        s.scale[i] = norm[NF + i];
    }
}

// Forward of one sample: xb (BF16 inputs), h1b/h2b (BF16 activations), z.
__device__ __forceinline__ void forward_one(const Lds& s, const float* __restrict__ f, float (&xb)[NF],
                                            float (&h1b)[H1], float (&h2b)[H2], float (&z)[NC]) {
#pragma unroll
    for (int i = 0; i < NF; i++) xb[i] = bf16r((f[i] - s.shift[i]) * s.scale[i]);
#pragma unroll
    for (int o = 0; o < H1; o++) {
        float acc = s.b1[o];
#pragma unroll
        for (int i = 0; i < NF; i++) acc = acc + s.w1[o * NF + i] * xb[i];
        h1b[o] = bf16r(acc > 0.0f ? acc : 0.0f);
    }
#pragma unroll
    for (int o = 0; o < H2; o++) {
        float acc = s.b2[o];
#pragma unroll
        for (int i = 0; i < H1; i++) acc = acc + s.w2[o * H1 + i] * h1b[i];
        h2b[o] = bf16r(acc > 0.0f ? acc : 0.0f);
    }
#pragma unroll
    for (int o = 0; o < NC; o++) {
        float acc = s.b3[o];
#pragma unroll
        for (int i = 0; i < H2; i++) acc = acc + s.w3[o * H2 + i] * h2b[i];
        z[o] = acc;
    }
}

__global__ void __launch_bounds__(256)
mlp_forward_kernel(const float* __restrict__ p, const float* __restrict__ norm, const float* __restrict__ feat,
                   uint64_t n, float* __restrict__ logits) {
    __shared__ Lds s;
    load_params(s, p, norm);
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float xb[NF], h1b[H1], h2b[H2], z[NC];
    forward_one(s, feat + i * NF, xb, h1b, h2b, z);
#pragma unroll
    for (int o = 0; o < NC; o++) logits[i * NC + o] = z[o];
}

// Per-sample record kept for the gradient kernel (BF16 values stored as float).
struct Rec {
    float* xb;   // [n][NF]
    float* h1b;  // [n][H1]
    float* h2b;  // [n][H2]
    float* d1b;  // [n][H1]
    float* d2b;  // [n][H2]
    float* d3b;  // [n][NC]
    float* loss; // [n]
};

Rec rec_layout(void* ws, size_t n) {
    float* w = (float*)ws;
    Rec r;
    r.xb = w;
    r.h1b = r.xb + n * NF;
    r.h2b = r.h1b + n * H1;
    r.d1b = r.h2b + n * H2;
    r.d2b = r.d1b + n * H1;
    r.d3b = r.d2b + n * H2;
    r.loss = r.d3b + n * NC;
    return r;
}

__global__ void __launch_bounds__(256)
mlp_backward_kernel(const float* __restrict__ p, const float* __restrict__ norm, const float* __restrict__ feat,
                    const float* __restrict__ cost, uint64_t n, float tau, Rec r) {
    __shared__ Lds s;
    load_params(s, p, norm);
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float xb[NF], h1b[H1], h2b[H2], z[NC];
    forward_one(s, feat + i * NF, xb, h1b, h2b, z);
    // softmax and the soft target softmax(-cost / tau)
    float zm = z[0], cm = -cost[i * NC] / tau;
    float nc[NC];
#pragma unroll
    for (int o = 0; o < NC; o++) {
        zm = fmaxf(zm, z[o]);
        nc[o] = -cost[i * NC + o] / tau;
        cm = fmaxf(cm, nc[o]);
    }
    float pe[NC], te[NC], ps = 0.0f, ts = 0.0f;
#pragma unroll
    for (int o = 0; o < NC; o++) {
        pe[o] = expf(z[o] - zm);
        te[o] = expf(nc[o] - cm);
        ps += pe[o];
        ts += te[o];
    }
    float loss = 0.0f, d3b[NC];
#pragma unroll
    for (int o = 0; o < NC; o++) {
        const float pr = pe[o] / ps, t = te[o] / ts;
        loss -= t * (z[o] - zm - logf(ps));
        d3b[o] = bf16r(pr - t);
    }
    float d2b[H2];
#pragma unroll
    for (int k = 0; k < H2; k++) {
        float g = 0.0f;
#pragma unroll
        for (int o = 0; o < NC; o++) g = g + s.w3[o * H2 + k] * d3b[o];
        d2b[k] = bf16r(h2b[k] > 0.0f ? g : 0.0f);
    }
    float d1b[H1];
#pragma unroll
    for (int j = 0; j < H1; j++) {
        float g = 0.0f;
#pragma unroll
        for (int k = 0; k < H2; k++) g = g + s.w2[k * H1 + j] * d2b[k];
        d1b[j] = bf16r(h1b[j] > 0.0f ? g : 0.0f);
    }
#pragma unroll
    for (int k = 0; k < NF; k++) r.xb[i * NF + k] = xb[k];
#pragma unroll
    for (int k = 0; k < H1; k++) {
        r.h1b[i * H1 + k] = h1b[k];
        r.d1b[i * H1 + k] = d1b[k];
    }
#pragma unroll
    for (int k = 0; k < H2; k++) {
        r.h2b[i * H2 + k] = h2b[k];
        r.d2b[i * H2 + k] = d2b[k];
    }
#pragma unroll
    for (int k = 0; k < NC; k++) r.d3b[i * NC + k] = d3b[k];
    r.loss[i] = loss;
}

// Gradient: the batch is cut into segments of GRAD_SEG samples; thread
// (parameter q, segment g) sums its segment's samples in order
// (mlp_grad_part_kernel), then one thread per parameter adds the segment
// sums in order and divides by n (mlp_grad_final_kernel) -- a fixed order,
// so the result is deterministic, and the work spreads over the GPU for
// large batches.  Index NP carries the loss.
constexpr uint32_t GRAD_SEG = 256;

__global__ void __launch_bounds__(256)
mlp_grad_part_kernel(Rec r, uint64_t n, float* __restrict__ part) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q > NP) return;
    const uint64_t i0 = (uint64_t)blockIdx.y * GRAD_SEG, i1 = min(n, i0 + GRAD_SEG);
    float acc = 0.0f;
    if (q == NP) {
        for (uint64_t i = i0; i < i1; i++) acc = acc + r.loss[i];
    } else if (q < OB1) {
        const int o = q / NF, k = q % NF;
        for (uint64_t i = i0; i < i1; i++) acc = acc + r.d1b[i * H1 + o] * r.xb[i * NF + k];
    } else if (q < OW2) {
        const int o = q - OB1;
        for (uint64_t i = i0; i < i1; i++) acc = acc + r.d1b[i * H1 + o];
    } else if (q < OB2) {
        const int o = (q - OW2) / H1, k = (q - OW2) % H1;
        for (uint64_t i = i0; i < i1; i++) acc = acc + r.d2b[i * H2 + o] * r.h1b[i * H1 + k];
    } else if (q < OW3) {
        const int o = q - OB2;
        for (uint64_t i = i0; i < i1; i++) acc = acc + r.d2b[i * H2 + o];
    } else if (q < OB3) {
        const int o = (q - OW3) / H2, k = (q - OW3) % H2;
        for (uint64_t i = i0; i < i1; i++) acc = acc + r.d3b[i * NC + o] * r.h2b[i * H2 + k];
    } else {
        const int o = q - OB3;
        for (uint64_t i = i0; i < i1; i++) acc = acc + r.d3b[i * NC + o];
    }
    part[(uint64_t)blockIdx.y * (NP + 1) + q] = acc;
}

__global__ void __launch_bounds__(256)
mlp_grad_final_kernel(const float* __restrict__ part, uint32_t nseg, uint64_t n, float* __restrict__ grad,
                      float* __restrict__ loss_out) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q > NP) return;
    float acc = 0.0f;
    for (uint32_t g = 0; g < nseg; g++) acc = acc + part[(uint64_t)g * (NP + 1) + q];
    if (q == NP)
        *loss_out = acc / (float)n;
    else
        grad[q] = acc / (float)n;
}

__global__ void __launch_bounds__(256)
adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
            float lr, float b1, float b2, float eps, float wd, float c1, float c2) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= NP) return;
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.0f - b1) * gq;
    const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    const float mh = mq / c1, vh = vq / c2;
    p[q] = p[q] - lr * (mh / (sqrtf(vh) + eps) + wd * p[q]);
}

// Chunk-major copy of a field: chunk c = (bz, by, bx) in row-major order of
// the chunk grid, element (z, y, x) of the chunk at field (bz cz + z, by cy +
// y, bx cx + x).  TO_CHUNKS = false scatters back.  One thread per element.
template <bool TO_CHUNKS>
__global__ void __launch_bounds__(256)
chunk_copy_kernel(const float* __restrict__ src, float* __restrict__ dst, uint64_t total, uint32_t nx, uint32_t ny,
                  uint32_t cx, uint32_t cy, uint32_t cz, uint32_t gx, uint32_t gy) {
    const uint64_t E = (uint64_t)cx * cy * cz;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (uint64_t)gridDim.x * 256) {
        const uint64_t c = i / E, e = i % E;
        const uint32_t x = (uint32_t)(e % cx), y = (uint32_t)((e / cx) % cy), z = (uint32_t)(e / ((uint64_t)cx * cy));
        const uint32_t bx = (uint32_t)(c % gx), by = (uint32_t)((c / gx) % gy), bz = (uint32_t)(c / ((uint64_t)gx * gy));
        const uint64_t f = ((uint64_t)(bz * cz + z) * ny + (by * cy + y)) * nx + (bx * cx + x);
        if constexpr (TO_CHUNKS) dst[i] = src[f];
        else dst[f] = src[i];
    }
}

__global__ void __launch_bounds__(256) select_kernel(const float* __restrict__ logits, uint64_t n,
                                                     uint8_t* __restrict__ choice) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int best = 0;
    float bv = logits[i * NC];
#pragma unroll
    for (int o = 1; o < NC; o++)
        if (logits[i * NC + o] > bv) {
            bv = logits[i * NC + o];
            best = o;
        }
    choice[i] = (uint8_t)best;
}


// ---------------------------------------------------------------------------
// Selector training/evaluation chunks generated on the device.  Chunk c
// belongs to parent field f = c / CPP (CPP = 4 chunks of 256 x 256 in a
// 512 x 512 field, or 8 chunks of 16 x 64 x 64 in a 32 x 128 x 128 field, in
// row-major order of the chunk grid) of regime f % 5, seeded by (seed, nd, f):
//   0  Gaussian random field, power ~ k^-4: a window of the caller's base
//      field (fzh_synth_power_law_f32 output) at a seeded offset, wrapping
//      periodically
//   1  white noise, N(0, 1)
//   2  80% exact zeros, the rest N(0, 1)
//   3  piecewise constant: 48 random axis-aligned boxes painted in order,
//      each a N(0, 1) value, on zeros
//   4  Poisson(20) integers (inverse CDF)
// and its bound is float(1e-3f * max |parent|) (1e-3 for an all-zero
// parent).  Deterministic in (seed, nd, chunk index).
// ---------------------------------------------------------------------------
constexpr int SYN_BOXES = 48;
constexpr int POIS_K = 64;

struct PoissonCdf {
    double c[POIS_K];
};
constexpr PoissonCdf make_poisson_cdf() {
    PoissonCdf t{};
    double p = 2.061153622438557828e-09;  // e^-20
    double acc = 0.0;
    for (int k = 0; k < POIS_K; k++) {
        acc += p;
        t.c[k] = acc;
        p = p * 20.0 / (double)(k + 1);
    }
    return t;
}
__constant__ PoissonCdf kPoisCdf = make_poisson_cdf();

struct SynGeom {
    uint32_t cx, cy, cz, gx, gy, cpp, px, py, pz;
};

__device__ __forceinline__ float wave_max_f(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

__host__ __device__ inline SynGeom syn_geom(int nd) {
    // chunk (cx, cy, cz), chunk grid of a parent (gx, gy, gz = cpp / (gx gy))
    if (nd == 2) return SynGeom{256, 256, 1, 2, 2, 4, 512, 512, 1};
    return SynGeom{64, 64, 16, 2, 2, 8, 128, 128, 32};
}

__device__ __forceinline__ uint64_t parent_seed(uint64_t seed, int nd, uint64_t f) {
    return mix64(seed ^ mix64(0x5E1EC7ull * (uint64_t)nd + f * 0x9E3779B97F4A7C15ull));
}

__global__ void __launch_bounds__(256)
synth_chunks_kernel(float* __restrict__ out, float* __restrict__ cmax, uint8_t* __restrict__ regime, int nd,
                    uint64_t seed, const float* __restrict__ base, uint32_t bsx, uint32_t bsy, uint32_t bsz) {
    const SynGeom g = syn_geom(nd);
    const uint64_t c = blockIdx.x;
    const uint64_t f = c / g.cpp;
    const uint32_t q = (uint32_t)(c % g.cpp);
    const uint32_t bx = q % g.gx, by = (q / g.gx) % g.gy, bz = q / (g.gx * g.gy);
    const int r = (int)(f % 5);
    const uint64_t ps = parent_seed(seed, nd, f);
    __shared__ int s_box[SYN_BOXES][6];
    __shared__ float s_val[SYN_BOXES];
    __shared__ float s_max[4];
    if (r == 3 && threadIdx.x < SYN_BOXES) {
        const uint32_t b = threadIdx.x;
        const uint32_t dims[3] = {g.px, g.py, g.pz};
        for (int a = 0; a < 3; a++) {
            const uint64_t h = mix64(ps ^ mix64(0xB0C5ull + 8 * b + a));
            uint32_t lo = (uint32_t)((h & 0xFFFFFFFFull) % (dims[a] + 1)), hi = (uint32_t)((h >> 32) % (dims[a] + 1));
            if (lo > hi) { const uint32_t t = lo; lo = hi; hi = t; }
            if (hi == lo) hi = lo + 1;
            s_box[b][2 * a] = (int)lo;
            s_box[b][2 * a + 1] = (int)hi;
        }
        s_val[b] = hash_normal(mix64(ps ^ mix64(0xBA1ull + b)));
    }
    __syncthreads();
    const uint32_t E = g.cx * g.cy * g.cz;
    const uint64_t wo = mix64(ps ^ 0x3141592653589793ull);  // the window's offset in the base
    const uint32_t ox = (uint32_t)(wo % bsx), oy = (uint32_t)((wo >> 21) % bsy), oz = (uint32_t)((wo >> 42) % bsz);
    float m = 0.0f;
    float* dst = out + c * E;
    for (uint32_t e = threadIdx.x; e < E; e += 256) {
        const uint32_t x = bx * g.cx + e % g.cx, y = by * g.cy + (e / g.cx) % g.cy, z = bz * g.cz + e / (g.cx * g.cy);
        const uint64_t p = ((uint64_t)z * g.py + y) * g.px + x;
        float v;
        if (r == 0) {
            v = base[((uint64_t)((oz + z) % bsz) * bsy + (oy + y) % bsy) * bsx + (ox + x) % bsx];
        } else if (r == 1) {
            v = hash_normal(mix64(ps ^ mix64(p)));
        } else if (r == 2) {
            const uint64_t h = mix64(ps ^ mix64(p));
            v = u01(h >> 24) >= 0.8f ? hash_normal(mix64(h)) : 0.0f;
        } else if (r == 3) {
            v = 0.0f;
            for (int b = 0; b < SYN_BOXES; b++)
                if ((int)x >= s_box[b][0] && (int)x < s_box[b][1] && (int)y >= s_box[b][2] && (int)y < s_box[b][3] &&
                    (int)z >= s_box[b][4] && (int)z < s_box[b][5])
                    v = s_val[b];
        } else {
            const double u = (double)(mix64(ps ^ mix64(p)) >> 11) * (1.0 / 9007199254740992.0);
            int lo = 0, hi = POIS_K - 1;  // smallest k with CDF(k) > u
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (kPoisCdf.c[mid] > u) hi = mid;
                else lo = mid + 1;
            }
            v = (float)lo;
        }
        dst[e] = v;
        m = fmaxf(m, fabsf(v));
    }
    m = wave_max_f(m);
    if (lane_id() == 0) s_max[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        cmax[c] = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
        if (regime) regime[c] = (uint8_t)r;
    }
}

// One thread per parent: its chunks' bound from their maxima (in place).
__global__ void __launch_bounds__(256) synth_eb_kernel(float* __restrict__ eb, uint64_t n_parents, uint32_t cpp) {
    const uint64_t f = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (f >= n_parents) return;
    float vb = 0.0f;
    for (uint32_t k = 0; k < cpp; k++) vb = fmaxf(vb, eb[f * cpp + k]);
    const float e = vb > 0.0f ? 1e-3f * vb : 1e-3f;
    for (uint32_t k = 0; k < cpp; k++) eb[f * cpp + k] = e;
}

}  // namespace
}  // namespace sel
}  // namespace fzh

using namespace fzh;
using namespace fzh::sel;

extern "C" {

int fzs_features_f32(void* stream, const float* d_chunks, size_t n_chunks, uint32_t nx, uint32_t ny, uint32_t nz,
                     const float* d_abs_eb, float* d_feat) {
    if (n_chunks == 0) return FZH_OK;
    if (!d_chunks || !d_abs_eb || !d_feat) return fail(FZH_E_INVALID, "features: null pointer");
    if (nx == 0 || ny == 0 || nz == 0) return fail(FZH_E_INVALID, "features: empty chunk shape");
    if (n_chunks > 0x7FFFFFFFull) return fail(FZH_E_INVALID, "features: too many chunks");
    features_kernel<<<(unsigned)n_chunks, 256, 0, S(stream)>>>(d_chunks, nx, ny, nz, d_abs_eb, d_feat);
    FZH_CHECK_LAUNCH("features_kernel");
    return FZH_OK;
}

size_t fzs_mlp_param_count(void) { return NP; }

int fzs_mlp_forward(void* stream, const float* d_params, const float* d_norm, const float* d_feat, size_t n,
                    float* d_logits) {
    if (n == 0) return FZH_OK;
    if (!d_params || !d_norm || !d_feat || !d_logits) return fail(FZH_E_INVALID, "mlp forward: null pointer");
    mlp_forward_kernel<<<(unsigned)cdiv(n, 256), 256, 0, S(stream)>>>(d_params, d_norm, d_feat, n, d_logits);
    FZH_CHECK_LAUNCH("mlp_forward_kernel");
    return FZH_OK;
}

size_t fzs_mlp_train_ws_bytes(size_t n) {
    return 4 * n * (NF + 2 * H1 + 2 * H2 + NC + 1) + 256 + 4 * cdiv(n, GRAD_SEG) * (NP + 1);
}

int fzs_mlp_train_step(void* stream, const float* d_params, const float* d_norm, const float* d_feat,
                       const float* d_cost, size_t n, float tau, float* d_grad, float* d_loss, void* d_ws,
                       size_t ws_bytes) {
    if (n == 0) return fail(FZH_E_INVALID, "mlp train: empty batch");
    if (!(tau > 0.0f)) return fail(FZH_E_INVALID, "mlp train: tau must be positive");
    if (!d_params || !d_norm || !d_feat || !d_cost || !d_grad || !d_loss || !d_ws)
        return fail(FZH_E_INVALID, "mlp train: null pointer");
    if (ws_bytes < fzs_mlp_train_ws_bytes(n)) return fail(FZH_E_CAPACITY, "mlp train: workspace too small");
    const Rec r = rec_layout(d_ws, n);
    mlp_backward_kernel<<<(unsigned)cdiv(n, 256), 256, 0, S(stream)>>>(d_params, d_norm, d_feat, d_cost, n, tau, r);
    const uint32_t nseg = (uint32_t)cdiv(n, GRAD_SEG);
    float* part = reinterpret_cast<float*>(reinterpret_cast<uint8_t*>(d_ws) + 4 * n * (NF + 2 * H1 + 2 * H2 + NC + 1) + 256);
    mlp_grad_part_kernel<<<dim3((unsigned)cdiv(NP + 1, 256), nseg), 256, 0, S(stream)>>>(r, n, part);
    mlp_grad_final_kernel<<<(unsigned)cdiv(NP + 1, 256), 256, 0, S(stream)>>>(part, nseg, n, d_grad, d_loss);
    FZH_CHECK_LAUNCH("mlp train");
    return FZH_OK;
}

int fzs_mlp_adam(void* stream, float* d_params, const float* d_grad, float* d_m, float* d_v, uint32_t step,
                 float lr, float beta1, float beta2, float eps, float weight_decay) {
    if (step == 0) return fail(FZH_E_INVALID, "adam: step counts from 1");
    if (!d_params || !d_grad || !d_m || !d_v) return fail(FZH_E_INVALID, "adam: null pointer");
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    adam_kernel<<<(unsigned)cdiv(NP, 256), 256, 0, S(stream)>>>(d_params, d_grad, d_m, d_v, lr, beta1, beta2, eps,
                                                               weight_decay, c1, c2);
    FZH_CHECK_LAUNCH("adam_kernel");
    return FZH_OK;
}

static int chunk_copy(void* stream, const float* src, float* dst, uint64_t nx, uint64_t ny, uint64_t nz, uint32_t cx,
                      uint32_t cy, uint32_t cz, bool to_chunks) {
    if (!src || !dst) return fail(FZH_E_INVALID, "chunk copy: null pointer");
    if (!cx || !cy || !cz || nx % cx || ny % cy || nz % cz)
        return fail(FZH_E_INVALID, "chunk copy: the field must be a whole number of chunks");
    const uint64_t total = nx * ny * nz;
    if (total == 0) return FZH_OK;
    const unsigned blocks = (unsigned)std::min<uint64_t>(cdiv(total, 256), 65536);
    const uint32_t gx = (uint32_t)(nx / cx), gy = (uint32_t)(ny / cy);
    if (to_chunks)
        chunk_copy_kernel<true><<<blocks, 256, 0, S(stream)>>>(src, dst, total, (uint32_t)nx, (uint32_t)ny, cx, cy, cz,
                                                               gx, gy);
    else
        chunk_copy_kernel<false><<<blocks, 256, 0, S(stream)>>>(src, dst, total, (uint32_t)nx, (uint32_t)ny, cx, cy,
                                                                cz, gx, gy);
    FZH_CHECK_LAUNCH("chunk_copy_kernel");
    return FZH_OK;
}

int fzs_gather_chunks(void* stream, const float* d_field, uint64_t nx, uint64_t ny, uint64_t nz, uint32_t cx,
                      uint32_t cy, uint32_t cz, float* d_chunks) {
    return chunk_copy(stream, d_field, d_chunks, nx, ny, nz, cx, cy, cz, true);
}

int fzs_scatter_chunks(void* stream, const float* d_chunks, uint64_t nx, uint64_t ny, uint64_t nz, uint32_t cx,
                       uint32_t cy, uint32_t cz, float* d_field) {
    return chunk_copy(stream, d_chunks, d_field, nx, ny, nz, cx, cy, cz, false);
}

int fzs_synth_chunks(void* stream, int nd, uint64_t n_chunks, uint64_t seed, const float* d_base, uint32_t bx,
                     uint32_t by, uint32_t bz, float* d_chunks, float* d_eb, uint8_t* d_regime) {
    if (nd != 2 && nd != 3) return fail(FZH_E_INVALID, "synth chunks: nd must be 2 or 3");
    if (n_chunks == 0) return FZH_OK;
    if (!d_chunks || !d_eb || !d_base) return fail(FZH_E_INVALID, "synth chunks: null pointer");
    if (!bx || !by || !bz || (nd == 2 && bz != 1)) return fail(FZH_E_INVALID, "synth chunks: bad base field shape");
    const SynGeom g = syn_geom(nd);
    if (n_chunks % g.cpp) return fail(FZH_E_INVALID, "synth chunks: a whole number of parent fields is required");
    if (n_chunks > 0x7FFFFFFFull) return fail(FZH_E_INVALID, "synth chunks: too many chunks");
    synth_chunks_kernel<<<(unsigned)n_chunks, 256, 0, S(stream)>>>(d_chunks, d_eb, d_regime, nd, seed, d_base, bx, by,
                                                                   bz);
    const uint64_t np = n_chunks / g.cpp;
    synth_eb_kernel<<<(unsigned)cdiv(np, 256), 256, 0, S(stream)>>>(d_eb, np, g.cpp);
    FZH_CHECK_LAUNCH("synth chunks");
    return FZH_OK;
}

int fzs_select(void* stream, const float* d_logits, size_t n, uint8_t* d_choice) {
    if (n == 0) return FZH_OK;
    if (!d_logits || !d_choice) return fail(FZH_E_INVALID, "select: null pointer");
    select_kernel<<<(unsigned)cdiv(n, 256), 256, 0, S(stream)>>>(d_logits, n, d_choice);
    FZH_CHECK_LAUNCH("select_kernel");
    return FZH_OK;
}

}  // extern "C"

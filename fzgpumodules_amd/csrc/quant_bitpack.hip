// quant_bitpack.hip — direct-value quantizer (ABS/NOA) and power-of-two bit
// packing (gfx950).
//
// Quantizer ABS/NOA (modules/quantizers/quantizer/quantizer.cu:30-67, 116-137):
//   q = __float2int_rn(x * ebx2_r)  (round half to even, saturating)
//   |q| < radius and |x| < threshold -> code = q (two's complement in TCode)
//   or zigzag(q); otherwise code 0 and an outlier (x, i).
//   inverse: out = (float)q * ebx2, then outliers assign their value.
// Outliers come out in index order: each 1024-element tile writes its list
// to a scratch slot and a compaction kernel places the lists (the reference
// appends with atomics).
// In-place outliers (quantizer.cu:83-104, 147-168): u32 codes, valid ones
// zigzag(q) in [0, 2 radius), an outlier's code is the raw bits of x; the
// inverse takes (code >> 1) >= radius as an outlier.
//
// Bitpack (modules/coders/bitpack/bitpack_stage.cu:30-124): the low `nbits`
// bits of each element, nbits in {1,2,4,8,16,32}; sub-byte: element k of a
// byte at bit k*nbits; multi-byte: little endian.
#include "fzh_compact.h"

namespace fzh {
namespace {

template <typename TCode> struct SignedOf;
template <> struct SignedOf<uint8_t> { using type = int8_t; };
template <> struct SignedOf<uint16_t> { using type = int16_t; };
template <> struct SignedOf<uint32_t> { using type = int32_t; };

// __float2int_rn: nearest-even; out of range saturates; NaN -> 0.
__device__ __forceinline__ int f2i_rn(float f) {
    if (f != f) return 0;
    if (f >= 2147483647.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int)0x80000000;
    return (int)__builtin_rintf(f);
}

constexpr int QTILE = 1024;  // elements per wave (16 per lane)

template <typename TCode, bool ZZ>
__global__ void __launch_bounds__(256)
quant_fwd_kernel(const float* __restrict__ in, uint64_t n, float r, float threshold, int rad,
                 TCode* __restrict__ codes, float* __restrict__ sval, uint32_t* __restrict__ sidx,
                 uint32_t* __restrict__ scnt, uint32_t T) {
    using SC = typename SignedOf<TCode>::type;
    constexpr int W = 8 * sizeof(TCode);
    const uint32_t tile = blockIdx.x * 4 + wave_in_wg();
    if (tile < T) {
        const int l = lane_id();
        const uint64_t base = (uint64_t)tile * QTILE;
        float x[16];
        uint32_t mask = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint64_t e0 = base + 4 * (uint64_t)(l + 64 * q);
            if (e0 + 4 <= n) {
                const float4 v = *reinterpret_cast<const float4*>(in + e0);
                x[4 * q] = v.x; x[4 * q + 1] = v.y; x[4 * q + 2] = v.z; x[4 * q + 3] = v.w;
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++) x[4 * q + k] = (e0 + k < n) ? in[e0 + k] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint64_t e = e0 + k;
                const int qv = f2i_rn(x[4 * q + k] * r);
                const bool ok = qv > -rad && qv < rad && fabsf(x[4 * q + k]) < threshold;
                TCode c = 0;
                if (ok) {
                    if constexpr (ZZ) {
                        const SC s = (SC)qv;
                        c = (TCode)(((TCode)s << 1) ^ (TCode)(s >> (W - 1)));
                    } else {
                        c = (TCode)qv;
                    }
                }
                if (e < n) {
                    codes[e] = c;
                    if (!ok) mask |= 1u << (4 * q + k);
                }
            }
        }
        const uint32_t agg = wave_sum(__popc(mask));
        if (agg) {
            uint32_t run = 0;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t mq = (mask >> (4 * q)) & 0xFu;
                const uint32_t c = __popc(mq);
                const uint32_t incl = wave_incl_scan(c);
                uint32_t pos = run + incl - c;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (mq & (1u << k)) {
                        sval[base + pos] = x[4 * q + k];
                        sidx[base + pos] = (uint32_t)(base + 4 * (uint64_t)(l + 64 * q) + k);
                        pos++;
                    }
                run += lane63(incl);
            }
        }
        if (l == 0) scnt[tile] = agg;
    }
}

template <typename TCode, bool ZZ>
__global__ void __launch_bounds__(256)
quant_inv_kernel(const TCode* __restrict__ codes, uint64_t n, float ebx2, float* __restrict__ out) {
    using SC = typename SignedOf<TCode>::type;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const TCode c = codes[i];
        int q;
        if constexpr (ZZ) q = (int)(SC)((TCode)(c >> 1) ^ (TCode)(-(SC)(c & 1u)));
        else q = (int)(SC)c;
        out[i] = (float)q * ebx2;
    }
}

// ---------------------------------------------------------------------------
// REL: log2-space bins (quantizer.cu:171-352).  For x normal, finite and
// below the threshold: log_bin = __float2int_rn(log2|x| * log2eb_r), the
// code packs (|log_bin|, sign of log_bin, sign of x) + 1 (0 marks an
// outlier), and it is kept when |log_bin| < radius, it fits TCode and
// 2^(log_bin * log2eb) lies in [|x| / (1+eps), |x| (1+eps)]; everything else
// is an outlier.  The reference takes log2 and 2^t from __log2f / exp2f
// (within 2 ulp); here both are the correctly rounded float results (the
// double-precision functions rounded once), which is what the bin and the
// bound test are defined on.
// ---------------------------------------------------------------------------
struct RelParams {
    float log2eb, log2eb_r, opp, oopp;
};

__device__ __forceinline__ float log2_cr(float x) { return (float)log2((double)x); }
__device__ __forceinline__ float exp2_cr(float t) { return (float)exp2((double)t); }

// Code of one element (0: outlier).
template <typename TCode>
__device__ __forceinline__ uint32_t rel_code(float x, const RelParams& p, float threshold, uint32_t rad) {
    const uint32_t bits = __float_as_uint(x);
    const uint32_t expo = (bits >> 23) & 0xFFu;
    if (fabsf(x) >= threshold || expo == 0u || expo == 255u) return 0u;
    const bool neg = x < 0.0f;
    const float ax = neg ? -x : x;
    const int lb = f2i_rn(log2_cr(ax) * p.log2eb_r);
    const uint32_t alb = lb < 0 ? 0u - (uint32_t)lb : (uint32_t)lb;
    const uint32_t stored = ((alb << 2) | ((lb < 0 ? 1u : 0u) << 1) | (neg ? 1u : 0u)) + 1u;
    constexpr uint32_t cmax = sizeof(TCode) == 1 ? 0xFFu : sizeof(TCode) == 2 ? 0xFFFFu : 0xFFFFFFFFu;
    if (!(alb < rad && stored <= cmax)) return 0u;
    const float rec = exp2_cr((float)lb * p.log2eb);
    return (rec >= ax * p.oopp && rec <= ax * p.opp) ? stored : 0u;
}

template <typename TCode>
__global__ void __launch_bounds__(256)
quant_rel_fwd_kernel(const float* __restrict__ in, uint64_t n, RelParams p, float threshold, uint32_t rad,
                     TCode* __restrict__ codes, float* __restrict__ sval, uint32_t* __restrict__ sidx,
                     uint32_t* __restrict__ scnt, uint32_t T) {
    const uint32_t tile = blockIdx.x * 4 + wave_in_wg();
    if (tile >= T) return;
    const int l = lane_id();
    const uint64_t base = (uint64_t)tile * QTILE;
    float x[16];
    uint32_t mask = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint64_t e0 = base + 4 * (uint64_t)(l + 64 * q);
        if (e0 + 4 <= n) {
            const float4 v = *reinterpret_cast<const float4*>(in + e0);
            x[4 * q] = v.x; x[4 * q + 1] = v.y; x[4 * q + 2] = v.z; x[4 * q + 3] = v.w;
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) x[4 * q + k] = (e0 + k < n) ? in[e0 + k] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint64_t e = e0 + k;
            const uint32_t c = rel_code<TCode>(x[4 * q + k], p, threshold, rad);
            if (e < n) {
                codes[e] = (TCode)c;
                if (c == 0u) mask |= 1u << (4 * q + k);
            }
        }
    }
    const uint32_t agg = wave_sum(__popc(mask));
    if (agg) {
        uint32_t run = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t mq = (mask >> (4 * q)) & 0xFu;
            const uint32_t c = __popc(mq);
            const uint32_t incl = wave_incl_scan(c);
            uint32_t pos = run + incl - c;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (mq & (1u << k)) {
                    sval[base + pos] = x[4 * q + k];
                    sidx[base + pos] = (uint32_t)(base + 4 * (uint64_t)(l + 64 * q) + k);
                    pos++;
                }
            run += lane63(incl);
        }
    }
    if (l == 0) scnt[tile] = agg;
}

// Inverse (quantizer.cu:328-352): code 0 -> 0 (an outlier's value is
// assigned afterwards), else +-2^(log_bin * log2eb).
template <typename TCode>
__global__ void __launch_bounds__(256)
quant_rel_inv_kernel(const TCode* __restrict__ codes, uint64_t n, float log2eb, float* __restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t st = (uint32_t)codes[i];
        float v = 0.0f;
        if (st != 0u) {
            const uint32_t cv = st - 1u;
            const int alb = (int)(cv >> 2);
            const int lb = ((cv >> 1) & 1u) ? -alb : alb;
            const float rec = exp2_cr((float)lb * log2eb);
            v = (cv & 1u) ? -rec : rec;
        }
        out[i] = v;
    }
}

// Grid-stride, one element per thread per step.
__global__ void __launch_bounds__(256)
quant_inplace_fwd_kernel(const float* __restrict__ in, uint64_t n, float r, float threshold, int rad,
                         uint32_t* __restrict__ codes) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        const float x = in[i];
        const int q = f2i_rn(x * r);
        const bool ok = q > -rad && q < rad && fabsf(x) < threshold;
        codes[i] = ok ? (uint32_t)((q << 1) ^ (q >> 31)) : __float_as_uint(x);
    }
}

__global__ void __launch_bounds__(256)
quant_inplace_inv_kernel(const uint32_t* __restrict__ codes, uint64_t n, float ebx2, uint32_t rad,
                         float* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        const uint32_t c = codes[i];
        if ((c >> 1) >= rad) {
            out[i] = __uint_as_float(c);
        } else {
            const int ci = (int)c;
            const int q = (ci & 1) ? -((ci + 1) >> 1) : (ci >> 1);
            out[i] = (float)q * ebx2;
        }
    }
}

__global__ void __launch_bounds__(256)
scatter_assign_kernel(const float* __restrict__ val, const uint32_t* __restrict__ idx,
                      const uint32_t* __restrict__ count, uint64_t max_out, uint64_t n, float* __restrict__ out) {
    const uint64_t k = min((uint64_t)*count, max_out);
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < k; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = idx[t];
        if (i < n) out[i] = val[t];
    }
}

// ---------------------------------------------------------------------------
// Bitpack
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ uint32_t ldv(const T* p, uint64_t i) { return (uint32_t)p[i]; }

// one thread per output byte (sub-byte), per element (>= 8 bits)
template <typename T>
__global__ void __launch_bounds__(256)
bitpack_enc_kernel(const T* __restrict__ in, uint64_t n, int nbits, uint8_t* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    if (nbits < 8) {
        const int epb = 8 / nbits;
        const uint32_t mask = (1u << nbits) - 1u;
        const uint64_t nout = (n * nbits + 7) / 8;
        for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nout; b += stride) {
            uint32_t packed = 0;
            for (int k = 0; k < epb; k++) {
                const uint64_t e = b * epb + k;
                if (e < n) packed |= (ldv(in, e) & mask) << (k * nbits);
            }
            out[b] = (uint8_t)packed;
        }
    } else {
        const int bytes = nbits / 8;
        const uint32_t mask = nbits >= 32 ? 0xFFFFFFFFu : ((1u << nbits) - 1u);
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
            uint32_t v = ldv(in, i) & mask;
            for (int b = 0; b < bytes; b++) {
                out[i * bytes + b] = (uint8_t)(v & 0xFFu);
                v >>= 8;
            }
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
bitpack_dec_kernel(const uint8_t* __restrict__ in, uint64_t n, int nbits, T* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (nbits < 8) {
            const int epb = 8 / nbits;
            const uint32_t mask = (1u << nbits) - 1u;
            out[i] = (T)((in[i / epb] >> ((i % epb) * nbits)) & mask);
        } else {
            const int bytes = nbits / 8;
            uint32_t v = 0;
            for (int b = 0; b < bytes; b++) v |= (uint32_t)in[i * bytes + b] << (8 * b);
            out[i] = (T)v;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
max_kernel(const T* __restrict__ in, uint64_t n, uint32_t* __restrict__ partial, uint32_t* __restrict__ counter,
           uint32_t* __restrict__ result) {
    __shared__ uint32_t sm[4];
    uint32_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = (uint32_t)in[i];
        m = v > m ? v : m;
    }
    m = wave_max(m);
    if (lane_id() == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 4; i++) m = sm[i] > m ? sm[i] : m;
        m = sm[0] > m ? sm[0] : m;
        partial[blockIdx.x] = m;
        __threadfence();
        if (atomicAdd(counter, 1u) == gridDim.x - 1) {
            __threadfence();
            uint32_t g = 0;
            for (uint32_t b = 0; b < gridDim.x; b++) {
                const uint32_t v = __hip_atomic_load(&partial[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                g = v > g ? v : g;
            }
            *result = g;
            *counter = 0u;
        }
    }
}

}  // namespace
}  // namespace fzh

using namespace fzh;

extern "C" {

// [compaction state][counts u32 x T][scratch values f32 x T*QTILE][scratch indices u32 x T*QTILE]
size_t fzh_quant_abs_fwd_ws_bytes(size_t n) {
    const size_t T = cdiv(std::max<size_t>(n, 1), QTILE);
    return compact_state_bytes(T) + align_up(4 * T, 256) + 2 * align_up(4 * T * QTILE, 256);
}

int fzh_quant_abs_fwd_f32(void* stream, const float* d_in, size_t n, float ebx2_r, float threshold, uint32_t radius,
                          int zigzag, int code_bytes, void* d_codes, float* d_oval, uint32_t* d_oidx,
                          uint32_t* d_ocount, size_t max_outliers, void* d_ws, size_t ws_bytes) {
    if (ws_bytes < fzh_quant_abs_fwd_ws_bytes(n)) return fail(FZH_E_CAPACITY, "quantizer: workspace too small");
    if (n == 0) return FZH_OK;
    if (n > 0xFFFFFFFFull) return fail(FZH_E_INVALID, "quantizer: more than 2^32 elements");
    if ((uintptr_t)d_in % 16) return fail(FZH_E_INVALID, "quantizer: input must be 16-byte aligned");
    if (max_outliers > 0 && (!d_oval || !d_oidx)) return fail(FZH_E_INVALID, "quantizer: null outlier buffers");
    const uint32_t T = (uint32_t)cdiv(n, QTILE);
    const unsigned blocks = (unsigned)cdiv(T, 4);
    uint8_t* w = (uint8_t*)d_ws;
    const size_t o_cnt = compact_state_bytes(T);
    const size_t o_val = o_cnt + align_up(4 * (size_t)T, 256);
    const size_t o_idx = o_val + align_up(4 * (size_t)T * QTILE, 256);
    uint32_t* scnt = (uint32_t*)(w + o_cnt);
    float* sval = (float*)(w + o_val);
    uint32_t* sidx = (uint32_t*)(w + o_idx);
    hipStream_t s = S(stream);
#define FZH_Q(TC, ZZV)                                                                                              \
    quant_fwd_kernel<TC, ZZV><<<blocks, 256, 0, s>>>(d_in, n, ebx2_r, threshold, (int)(TC)radius, (TC*)d_codes,   \
                                                     sval, sidx, scnt, T)
    switch (code_bytes) {
    case 1: if (zigzag) FZH_Q(uint8_t, true); else FZH_Q(uint8_t, false); break;
    case 2: if (zigzag) FZH_Q(uint16_t, true); else FZH_Q(uint16_t, false); break;
    case 4: if (zigzag) FZH_Q(uint32_t, true); else FZH_Q(uint32_t, false); break;
    default: return fail(FZH_E_INVALID, "quantizer: code_bytes must be 1, 2 or 4");
    }
#undef FZH_Q
    CompactList L0{sval, sidx, scnt, d_oval, d_oidx, d_ocount, max_outliers};
    FZH_CHECK_LAUNCH("quant_fwd_kernel");
    if (int rc = launch_compact<float>(s, w, T, QTILE, L0, nullptr)) return rc;
    return FZH_OK;
}

int fzh_quant_abs_inv_f32(void* stream, const void* d_codes, size_t n, float ebx2, int zigzag, int code_bytes,
                          const float* d_oval, const uint32_t* d_oidx, const uint32_t* d_ocount, size_t max_outliers,
                          float* d_out) {
    if (n == 0) return FZH_OK;
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(n, 256), 4096));
    hipStream_t s = S(stream);
#define FZH_QI(TC, ZZV) quant_inv_kernel<TC, ZZV><<<blocks, 256, 0, s>>>((const TC*)d_codes, n, ebx2, d_out)
    switch (code_bytes) {
    case 1: if (zigzag) FZH_QI(uint8_t, true); else FZH_QI(uint8_t, false); break;
    case 2: if (zigzag) FZH_QI(uint16_t, true); else FZH_QI(uint16_t, false); break;
    case 4: if (zigzag) FZH_QI(uint32_t, true); else FZH_QI(uint32_t, false); break;
    default: return fail(FZH_E_INVALID, "quantizer: code_bytes must be 1, 2 or 4");
    }
#undef FZH_QI
    if (d_ocount && max_outliers > 0) {
        const unsigned ob = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(max_outliers, 256), 1024));
        scatter_assign_kernel<<<ob, 256, 0, s>>>(d_oval, d_oidx, d_ocount, max_outliers, n, d_out);
    }
    FZH_CHECK_LAUNCH("quantizer inverse");
    return FZH_OK;
}

void fzh_quant_rel_params(float eps, float* params) {
    const float log2eb = 2.0f * std::log2(1.0f + eps);  // quantizer.cu:563-568
    params[0] = log2eb;
    params[1] = 1.0f / log2eb;
    params[2] = 1.0f + eps;
    params[3] = 1.0f / (1.0f + eps);
}

int fzh_quant_rel_fwd_f32(void* stream, const float* d_in, size_t n, float eps, float threshold, uint32_t radius,
                          int code_bytes, void* d_codes, float* d_oval, uint32_t* d_oidx, uint32_t* d_ocount,
                          size_t max_outliers, void* d_ws, size_t ws_bytes) {
    if (ws_bytes < fzh_quant_abs_fwd_ws_bytes(n)) return fail(FZH_E_CAPACITY, "quantizer (REL): workspace too small");
    if (n == 0) return FZH_OK;
    if (n > 0xFFFFFFFFull) return fail(FZH_E_INVALID, "quantizer (REL): more than 2^32 elements");
    if (!d_in || !d_codes || !d_ocount) return fail(FZH_E_INVALID, "quantizer (REL): null pointer");
    if ((uintptr_t)d_in % 16) return fail(FZH_E_INVALID, "quantizer (REL): input must be 16-byte aligned");
    if (max_outliers > 0 && (!d_oval || !d_oidx)) return fail(FZH_E_INVALID, "quantizer (REL): null outlier buffers");
    float pf[4];
    fzh_quant_rel_params(eps, pf);
    const RelParams p{pf[0], pf[1], pf[2], pf[3]};
    const uint32_t T = (uint32_t)cdiv(n, QTILE);
    const unsigned blocks = (unsigned)cdiv(T, 4);
    uint8_t* w = (uint8_t*)d_ws;
    const size_t o_cnt = compact_state_bytes(T);
    const size_t o_val = o_cnt + align_up(4 * (size_t)T, 256);
    const size_t o_idx = o_val + align_up(4 * (size_t)T * QTILE, 256);
    uint32_t* scnt = (uint32_t*)(w + o_cnt);
    float* sval = (float*)(w + o_val);
    uint32_t* sidx = (uint32_t*)(w + o_idx);
    hipStream_t s = S(stream);
    switch (code_bytes) {
    case 1:
        launch_k("quant_rel_fwd_kernel", quant_rel_fwd_kernel<uint8_t>, dim3(blocks), dim3(256), 0, s, d_in, n, p,
                 threshold, (uint32_t)(uint8_t)radius, (uint8_t*)d_codes, sval, sidx, scnt, T);
        break;
    case 2:
        launch_k("quant_rel_fwd_kernel", quant_rel_fwd_kernel<uint16_t>, dim3(blocks), dim3(256), 0, s, d_in, n, p,
                 threshold, (uint32_t)(uint16_t)radius, (uint16_t*)d_codes, sval, sidx, scnt, T);
        break;
    case 4:
        launch_k("quant_rel_fwd_kernel", quant_rel_fwd_kernel<uint32_t>, dim3(blocks), dim3(256), 0, s, d_in, n, p,
                 threshold, radius, (uint32_t*)d_codes, sval, sidx, scnt, T);
        break;
    default: return fail(FZH_E_INVALID, "quantizer (REL): code_bytes must be 1, 2 or 4");
    }
    FZH_CHECK_LAUNCH("quant_rel_fwd_kernel");
    CompactList L0{sval, sidx, scnt, d_oval, d_oidx, d_ocount, max_outliers};
    return launch_compact<float>(s, w, T, QTILE, L0, nullptr);
}

int fzh_quant_rel_inv_f32(void* stream, const void* d_codes, size_t n, float eps, int code_bytes,
                          const float* d_oval, const uint32_t* d_oidx, const uint32_t* d_ocount, size_t max_outliers,
                          float* d_out) {
    if (n == 0) return FZH_OK;
    if (!d_codes || !d_out) return fail(FZH_E_INVALID, "quantizer (REL) inverse: null pointer");
    float pf[4];
    fzh_quant_rel_params(eps, pf);
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(n, 256), 4096));
    hipStream_t s = S(stream);
    switch (code_bytes) {
    case 1:
        launch_k("quant_rel_inv_kernel", quant_rel_inv_kernel<uint8_t>, dim3(blocks), dim3(256), 0, s,
                 (const uint8_t*)d_codes, n, pf[0], d_out);
        break;
    case 2:
        launch_k("quant_rel_inv_kernel", quant_rel_inv_kernel<uint16_t>, dim3(blocks), dim3(256), 0, s,
                 (const uint16_t*)d_codes, n, pf[0], d_out);
        break;
    case 4:
        launch_k("quant_rel_inv_kernel", quant_rel_inv_kernel<uint32_t>, dim3(blocks), dim3(256), 0, s,
                 (const uint32_t*)d_codes, n, pf[0], d_out);
        break;
    default: return fail(FZH_E_INVALID, "quantizer (REL): code_bytes must be 1, 2 or 4");
    }
    if (d_ocount && max_outliers > 0) {
        const unsigned ob = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(max_outliers, 256), 1024));
        launch_k("scatter_assign_kernel", scatter_assign_kernel, dim3(ob), dim3(256), 0, s, d_oval, d_oidx, d_ocount,
                 (uint64_t)max_outliers, (uint64_t)n, d_out);
    }
    FZH_CHECK_LAUNCH("quantizer (REL) inverse");
    return FZH_OK;
}

int fzh_quant_abs_fwd_inplace_f32(void* stream, const float* d_in, size_t n, float ebx2_r, float threshold,
                                  uint32_t radius, uint32_t* d_codes) {
    if (n == 0) return FZH_OK;
    if (!d_in || !d_codes) return fail(FZH_E_INVALID, "quantizer (in place): null pointer");
    if (radius == 0 || radius > (1u << 30)) return fail(FZH_E_INVALID, "quantizer (in place): radius must be in [1, 2^30]");
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(n, 256), 8192));
    quant_inplace_fwd_kernel<<<blocks, 256, 0, S(stream)>>>(d_in, n, ebx2_r, threshold, (int)radius, d_codes);
    FZH_CHECK_LAUNCH("quant_inplace_fwd_kernel");
    return FZH_OK;
}

int fzh_quant_abs_inv_inplace_f32(void* stream, const uint32_t* d_codes, size_t n, float ebx2, uint32_t radius,
                                  float* d_out) {
    if (n == 0) return FZH_OK;
    if (!d_codes || !d_out) return fail(FZH_E_INVALID, "quantizer (in place): null pointer");
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(n, 256), 8192));
    quant_inplace_inv_kernel<<<blocks, 256, 0, S(stream)>>>(d_codes, n, ebx2, radius, d_out);
    FZH_CHECK_LAUNCH("quant_inplace_inv_kernel");
    return FZH_OK;
}

size_t fzh_max_u32_ws_bytes(void) { return 256 + 4 * 1024; }

int fzh_max_u32(void* stream, const void* d_in, size_t n, int elem_bytes, uint32_t* d_max, void* d_ws,
                   size_t ws_bytes) {
    if (ws_bytes < fzh_max_u32_ws_bytes()) return fail(FZH_E_CAPACITY, "max: workspace too small");
    uint32_t* counter = (uint32_t*)d_ws;
    uint32_t* partial = (uint32_t*)((uint8_t*)d_ws + 256);
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(std::max<size_t>(n, 1), 256 * 8), 1024));
    hipStream_t s = S(stream);
    switch (elem_bytes) {
    case 1: max_kernel<uint8_t><<<blocks, 256, 0, s>>>((const uint8_t*)d_in, n, partial, counter, d_max); break;
    case 2: max_kernel<uint16_t><<<blocks, 256, 0, s>>>((const uint16_t*)d_in, n, partial, counter, d_max); break;
    case 4: max_kernel<uint32_t><<<blocks, 256, 0, s>>>((const uint32_t*)d_in, n, partial, counter, d_max); break;
    default: return fail(FZH_E_INVALID, "max: elem_bytes must be 1, 2 or 4");
    }
    FZH_CHECK_LAUNCH("max_kernel");
    return FZH_OK;
}

int fzh_bitpack_enc(void* stream, const void* d_in, size_t n, int elem_bytes, int nbits, void* d_out) {
    if (nbits <= 0 || nbits > 8 * elem_bytes || (nbits & (nbits - 1)) != 0)
        return fail(FZH_E_INVALID, "bitpack: nbits must be a power of two <= 8*sizeof(T)");
    if (n == 0) return FZH_OK;
    const size_t work = nbits < 8 ? (n * nbits + 7) / 8 : n;
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(work, 256), 4096));
    hipStream_t s = S(stream);
    switch (elem_bytes) {
    case 1: bitpack_enc_kernel<uint8_t><<<blocks, 256, 0, s>>>((const uint8_t*)d_in, n, nbits, (uint8_t*)d_out); break;
    case 2: bitpack_enc_kernel<uint16_t><<<blocks, 256, 0, s>>>((const uint16_t*)d_in, n, nbits, (uint8_t*)d_out); break;
    case 4: bitpack_enc_kernel<uint32_t><<<blocks, 256, 0, s>>>((const uint32_t*)d_in, n, nbits, (uint8_t*)d_out); break;
    default: return fail(FZH_E_INVALID, "bitpack: elem_bytes must be 1, 2 or 4");
    }
    FZH_CHECK_LAUNCH("bitpack_enc_kernel");
    return FZH_OK;
}

int fzh_bitpack_dec(void* stream, const void* d_in, size_t n, int elem_bytes, int nbits, void* d_out) {
    if (nbits <= 0 || nbits > 8 * elem_bytes || (nbits & (nbits - 1)) != 0)
        return fail(FZH_E_INVALID, "bitpack: nbits must be a power of two <= 8*sizeof(T)");
    if (n == 0) return FZH_OK;
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(cdiv(n, 256), 4096));
    hipStream_t s = S(stream);
    switch (elem_bytes) {
    case 1: bitpack_dec_kernel<uint8_t><<<blocks, 256, 0, s>>>((const uint8_t*)d_in, n, nbits, (uint8_t*)d_out); break;
    case 2: bitpack_dec_kernel<uint16_t><<<blocks, 256, 0, s>>>((const uint8_t*)d_in, n, nbits, (uint16_t*)d_out); break;
    case 4: bitpack_dec_kernel<uint32_t><<<blocks, 256, 0, s>>>((const uint8_t*)d_in, n, nbits, (uint32_t*)d_out); break;
    default: return fail(FZH_E_INVALID, "bitpack: elem_bytes must be 1, 2 or 4");
    }
    FZH_CHECK_LAUNCH("bitpack_dec_kernel");
    return FZH_OK;
}

}  // extern "C"

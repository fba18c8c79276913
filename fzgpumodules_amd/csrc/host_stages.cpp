// host_stages.cpp — fz:: stage classes on top of the fzh_* kernels.
//
// Each execute() enqueues kernels only; values the host needs afterwards
// (outlier counts, stream sizes, error status, resolved REL/NOA bounds) are
// read in postStreamSync(), after the pipeline's one stream synchronisation.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "fz/stages.h"
#include "fzh_internal.h"

namespace fz {

void throwIfFzh(int rc, const char* what) {
    if (rc != FZH_OK) throw std::runtime_error(std::string(what) + ": " + fzh_last_error());
}

#define FZ_HIP(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

void DeviceScratch::ensure(size_t need, MemoryPool* p, hipStream_t s) {
    need = std::max<size_t>(need, 256);
    if (ptr && bytes >= need) {
        // Layouts inside a workspace depend on the problem size: clear the
        // parts a new size may read as zero.
        if (need != last_need) {
            throwIfFzh(fzh_workspace_init(s, ptr, bytes), "workspace init");
            last_need = need;
        }
        return;
    }
    release();
    pool = p;
    if (pool) ptr = pool->allocatePersistentDevice(need, "stage scratch");
    else FZ_HIP(hipMalloc(&ptr, need));
    if (!ptr) throw std::runtime_error("stage scratch allocation failed");
    bytes = need;
    last_need = need;
    throwIfFzh(fzh_workspace_init(s, ptr, bytes), "workspace init");
}

void DeviceScratch::release() {
    if (!ptr) return;
    if (pool) pool->freePersistentDevice(ptr);
    else (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

// ===========================================================================
// LorenzoQuantStage
// ===========================================================================
template <typename TI, typename TC>
LorenzoQuantStage<TI, TC>::LorenzoQuantStage(const Config& c)
    : config_(c), computed_abs_eb_(c.error_bound), computed_value_base_(c.precomputed_value_base) {}

template <typename TI, typename TC>
void LorenzoQuantStage<TI, TC>::resolveDims(size_t n, size_t& nx, size_t& ny, size_t& nz) const {
    const int nd = ndim();
    ny = nd >= 2 ? config_.dims[1] : 1;
    nz = nd >= 3 ? config_.dims[2] : 1;
    nx = config_.dims[0] > 0 ? config_.dims[0] : n / (ny * nz);
    if (nd == 1) nx = n;
}

template <typename TI, typename TC>
std::vector<size_t> LorenzoQuantStage<TI, TC>::estimateOutputSizes(const std::vector<size_t>& in) const {
    if (is_inverse_) return {num_elements_ * sizeof(TI)};
    const size_t n = in.empty() ? 0 : in[0] / sizeof(TI);
    const size_t m = maxOutliers(n);
    std::vector<size_t> r = {n * sizeof(TC), m * sizeof(TI), m * sizeof(uint32_t), sizeof(uint32_t)};
    if (config_.strict_error_bound) {
        const size_t pm = maxPatches(n);
        r.insert(r.end(), {pm * sizeof(TI), pm * sizeof(uint32_t), sizeof(uint32_t)});
    }
    return r;
}

template <typename TI, typename TC>
std::unordered_map<std::string, size_t> LorenzoQuantStage<TI, TC>::getActualOutputSizesByName() const {
    std::unordered_map<std::string, size_t> r;
    const auto names = getOutputNames();
    for (size_t i = 0; i < names.size() && i < actual_output_sizes_.size(); i++) r[names[i]] = actual_output_sizes_[i];
    return r;
}

template <typename TI, typename TC>
void LorenzoQuantStage<TI, TC>::execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                                        const std::vector<void*>& outputs, const std::vector<size_t>& sizes) {
    if (is_inverse_) {
        if (inputs.size() < numPorts() || outputs.empty() || sizes.size() < numPorts())
            throw std::runtime_error("LorenzoQuantStage (inverse): Requires " + std::to_string(numPorts()) +
                                     " inputs and 1 output");
        const size_t n = sizes[0] / sizeof(TC);
        actual_output_sizes_.assign(1, n * sizeof(TI));
        if (n == 0) return;
        size_t nx, ny, nz;
        resolveDims(n, nx, ny, nz);
        const size_t max_out = sizes[1] / sizeof(TI);
        if constexpr (std::is_same_v<TI, double>) {
            // ebx2 = 2 * abs_eb in double (lorenzo_quant.cu:509)
            inv_ws_.ensure(fzh_lorenzo_quant_inv_f64_ws_bytes(n, nx, ny, nz, max_out), pool, stream);
            throwIfFzh(fzh_lorenzo_quant_inv_f64(stream, inputs[0], n, nx, ny, nz, (const double*)inputs[1],
                                                 (const uint32_t*)inputs[2], (const uint32_t*)inputs[3], max_out,
                                                 2.0 * computed_abs_eb_, config_.zigzag_codes ? 1 : 0, (int)sizeof(TC),
                                                 (double*)outputs[0], inv_ws_.ptr, inv_ws_.bytes),
                       "LorenzoQuantStage (inverse)");
            return;
        }
        const float ebx2 = 2.0f * (float)computed_abs_eb_;
        inv_ws_.ensure(fzh_lorenzo_quant_inv_ws_bytes(n, nx, ny, nz, max_out), pool, stream);
        if (config_.strict_error_bound) {
            // the patches go in with the inverse (one launch for short lists)
            throwIfFzh(fzh_lorenzo_quant_inv_strict_f32(
                           stream, inputs[0], n, nx, ny, nz, (const float*)inputs[1], (const uint32_t*)inputs[2],
                           (const uint32_t*)inputs[3], max_out, ebx2, config_.zigzag_codes ? 1 : 0, (int)sizeof(TC),
                           (const float*)inputs[4], (const uint32_t*)inputs[5], (const uint32_t*)inputs[6],
                           sizes[4] / sizeof(TI), (float*)outputs[0], inv_ws_.ptr, inv_ws_.bytes),
                       "LorenzoQuantStage (inverse)");
            return;
        }
        throwIfFzh(fzh_lorenzo_quant_inv_f32(stream, inputs[0], n, nx, ny, nz, (const float*)inputs[1],
                                             (const uint32_t*)inputs[2], (const uint32_t*)inputs[3], max_out, ebx2,
                                             config_.zigzag_codes ? 1 : 0, (int)sizeof(TC), (float*)outputs[0],
                                             inv_ws_.ptr, inv_ws_.bytes),
                   "LorenzoQuantStage (inverse)");
        return;
    }
    if (inputs.empty() || outputs.size() < numPorts() || sizes.empty())
        throw std::runtime_error("LorenzoQuantStage: Requires 1 input and " + std::to_string(numPorts()) + " outputs");
    const size_t n = sizes[0] / sizeof(TI);
    const size_t valid = (valid_elems_ > 0 && valid_elems_ < n) ? valid_elems_ : n;
    valid_elems_ = 0;
    num_elements_ = n;
    const size_t max_out = maxOutliers(n);
    actual_outlier_count_ = 0;
    clearReadbacks();
    d_count_ptr_ = nullptr;
    d_pcount_ptr_ = nullptr;
    h_pcount_ = 0;
    if (n == 0) {
        actual_output_sizes_.assign(numPorts(), 0);
        return;
    }
    size_t nx, ny, nz;
    resolveDims(n, nx, ny, nz);
    if (ndim() == 1) nx = valid;
    else if (nx * ny * nz > valid)
        throw std::runtime_error("LorenzoQuantStage: dims exceed the unpadded input");
    if constexpr (std::is_same_v<TI, double>) {
        executeF64(stream, pool, inputs, outputs, n, valid, nx, ny, nz, max_out);
    } else {
        // Resolve the absolute bound (lorenzo_quant.cu:612-644).
        const float* d_r = nullptr;
        float r = 0.0f;
        params_pending_ = false;
        if (config_.eb_mode == ErrorBoundMode::ABS) {
            computed_abs_eb_ = config_.error_bound;
            computed_value_base_ = 0.0f;
            r = 1.0f / (2.0f * computed_abs_eb_);
        } else if (config_.precomputed_value_base > 0.0f) {
            computed_value_base_ = config_.precomputed_value_base;
            computed_abs_eb_ = config_.error_bound * computed_value_base_;
            r = 1.0f / (2.0f * computed_abs_eb_);
        } else {
            // value base reduced on the device; no host round trip before the kernel
            params_.ensure(256 + fzh_value_base_ws_bytes(n), pool, stream);
            float* d_params = (float*)params_.ptr;
            const int mode = (config_.eb_mode == ErrorBoundMode::NOA ? 2 : 1) | (valid < n ? FZH_VB_ZERO_PADDED : 0);
            throwIfFzh(fzh_value_base_f32(stream, (const float*)inputs[0], valid, mode, config_.error_bound, d_params,
                                          (uint8_t*)params_.ptr + 256, params_.bytes - 256),
                       "LorenzoQuantStage value base");
            d_r = d_params + 2;
            params_pending_ = true;
            addReadback(d_params, h_params_, sizeof(h_params_));
        }
        ws_.ensure(config_.strict_error_bound ? fzh_lorenzo_quant_fwd_strict_ws_bytes(n, nx, ny, nz)
                                              : fzh_lorenzo_quant_fwd_ws_bytes(n, nx, ny, nz),
                   pool, stream);
        if (!config_.strict_error_bound) {
            throwIfFzh(fzh_lorenzo_quant_fwd_f32(stream, (const float*)inputs[0], n, nx, ny, nz, r, d_r,
                                                 (uint32_t)config_.quant_radius, config_.zigzag_codes ? 1 : 0,
                                                 (int)sizeof(TC), outputs[0], (float*)outputs[1], (uint32_t*)outputs[2],
                                                 (uint32_t*)outputs[3], max_out, ws_.ptr, ws_.bytes),
                       "LorenzoQuantStage");
        } else {
            const size_t pm = maxPatches(n);
            const float eb = computed_abs_eb_;
            throwIfFzh(fzh_lorenzo_quant_fwd_strict_f32(
                           stream, (const float*)inputs[0], n, nx, ny, nz, r, d_r, (uint32_t)config_.quant_radius,
                           config_.zigzag_codes ? 1 : 0, (int)sizeof(TC), outputs[0], (float*)outputs[1],
                           (uint32_t*)outputs[2], (uint32_t*)outputs[3], max_out, eb, 2.0f * eb,
                           params_pending_ ? (const float*)params_.ptr : nullptr, (float*)outputs[4],
                           (uint32_t*)outputs[5], (uint32_t*)outputs[6], pm, ws_.ptr, ws_.bytes),
                       "LorenzoQuantStage (strict)");
            d_pcount_ptr_ = outputs[6];
            addReadback(d_pcount_ptr_, &h_pcount_, 4);
        }
        d_count_ptr_ = outputs[3];
        addReadback(d_count_ptr_, &h_count_, 4);
        actual_output_sizes_ = {n * sizeof(TC), max_out * sizeof(TI), max_out * sizeof(uint32_t), sizeof(uint32_t)};
        if (config_.strict_error_bound) {
            const size_t pm = maxPatches(n);
            actual_output_sizes_.insert(actual_output_sizes_.end(), {pm * sizeof(TI), pm * sizeof(uint32_t), sizeof(uint32_t)});
        }
    }
}

// FP64 forward: the bound resolved in double (lorenzo_quant.cu:612-644 with
// TInput = double: abs_eb = (double)user_eb * (double)value_base, value_base
// a float; ebx2_r = 1 / (2 abs_eb)), then the FP64 kernels.
template <typename TI, typename TC>
void LorenzoQuantStage<TI, TC>::executeF64(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                                           const std::vector<void*>& outputs, size_t n, size_t valid, size_t nx,
                                           size_t ny, size_t nz, size_t max_out) {
    if constexpr (std::is_same_v<TI, double>) {
        if (config_.strict_error_bound)
            throw std::runtime_error("LorenzoQuantStage: the strict error bound is available for float input only");
        const double* d_r = nullptr;
        double r = 0.0;
        params_pending_ = false;
        if (config_.eb_mode == ErrorBoundMode::ABS) {
            computed_abs_eb_ = (TI)config_.error_bound;
            computed_value_base_ = 0.0f;
            r = 1.0 / (2.0 * (double)computed_abs_eb_);
        } else if (config_.precomputed_value_base > 0.0f) {
            computed_value_base_ = config_.precomputed_value_base;
            computed_abs_eb_ = (TI)config_.error_bound * (TI)computed_value_base_;
            r = 1.0 / (2.0 * (double)computed_abs_eb_);
        } else {
            params_.ensure(256 + fzh_value_base_f64_ws_bytes(n), pool, stream);
            double* d_params = (double*)params_.ptr;
            const int mode = (config_.eb_mode == ErrorBoundMode::NOA ? 2 : 1) | (valid < n ? FZH_VB_ZERO_PADDED : 0);
            throwIfFzh(fzh_value_base_f64(stream, (const double*)inputs[0], valid, mode, config_.error_bound, d_params,
                                          (uint8_t*)params_.ptr + 256, params_.bytes - 256),
                       "LorenzoQuantStage value base");
            d_r = d_params + 2;
            params_pending_ = true;
            addReadback(d_params, h_params_, sizeof(h_params_));
        }
        ws_.ensure(fzh_lorenzo_quant_fwd_f64_ws_bytes(n, nx, ny, nz), pool, stream);
        throwIfFzh(fzh_lorenzo_quant_fwd_f64(stream, (const double*)inputs[0], n, nx, ny, nz, r, d_r,
                                             (uint32_t)config_.quant_radius, config_.zigzag_codes ? 1 : 0, (int)sizeof(TC),
                                             outputs[0], (double*)outputs[1], (uint32_t*)outputs[2], (uint32_t*)outputs[3],
                                             max_out, ws_.ptr, ws_.bytes),
                   "LorenzoQuantStage");
        d_count_ptr_ = outputs[3];
        addReadback(d_count_ptr_, &h_count_, 4);
        actual_output_sizes_ = {n * sizeof(TC), max_out * sizeof(TI), max_out * sizeof(uint32_t), sizeof(uint32_t)};
    }
}

template <typename TI, typename TC>
void LorenzoQuantStage<TI, TC>::postStreamSync(hipStream_t) {
    if (is_inverse_ || !d_count_ptr_) return;
    completeReadbacks();
    if (params_pending_) {
        computed_value_base_ = (float)h_params_[0];
        computed_abs_eb_ = h_params_[1];
    }
    uint32_t cnt = h_count_;
    if (cnt == 0xFFFFFFFFu) throw std::runtime_error("LorenzoQuantStage: device prefix failed (workspace state)");
    const size_t max_out = maxOutliers(num_elements_);
    if (cnt > max_out) {
        if (config_.outlier_capacity != 0.0f)
            std::fprintf(stderr,
                         "[fz] LorenzoQuant outlier overflow: %u outliers, capacity %zu; the excess was dropped\n",
                         cnt, max_out);
        cnt = (uint32_t)max_out;
    }
    actual_outlier_count_ = cnt;
    actual_output_sizes_[1] = (size_t)cnt * sizeof(TI);
    actual_output_sizes_[2] = (size_t)cnt * sizeof(uint32_t);
    if (config_.strict_error_bound && d_pcount_ptr_) {
        const uint32_t pc = h_pcount_;
        if (pc == 0xFFFFFFFFu) throw std::runtime_error("LorenzoQuantStage: device prefix failed (workspace state)");
        if (pc > maxPatches(num_elements_))
            throw std::runtime_error("LorenzoQuantStage: strict error bound needs " + std::to_string(pc) +
                                     " patched values but has room for " + std::to_string(maxPatches(num_elements_)) +
                                     "; raise setPatchCapacity()");
        if (cnt < h_count_ && h_count_ != 0xFFFFFFFFu)
            throw std::runtime_error("LorenzoQuantStage: strict error bound cannot hold with " +
                                     std::to_string(h_count_) + " outliers for " + std::to_string(cnt) +
                                     " slots; raise setOutlierCapacity()");
        actual_output_sizes_[4] = (size_t)pc * sizeof(TI);
        actual_output_sizes_[5] = (size_t)pc * sizeof(uint32_t);
    }
}

template <typename TI, typename TC>
size_t LorenzoQuantStage<TI, TC>::serializeHeader(size_t, uint8_t* buf, size_t max_size) const {
    if (max_size < sizeof(LorenzoQuantConfig)) throw std::runtime_error("Insufficient buffer for Lorenzo config");
    LorenzoQuantConfig c;
    c.error_bound = (float)computed_abs_eb_;  // lorenzo_quant.h:218
    c.quant_radius = (uint32_t)config_.quant_radius;
    c.num_elements = (uint32_t)num_elements_;
    c.outlier_count = actual_outlier_count_;
    c.input_type = dataTypeOf<TI>();
    c.code_type = dataTypeOf<TC>();
    c.ndim = (uint8_t)ndim();
    c.eb_mode = (uint8_t)config_.eb_mode;
    c.dim_x = (uint32_t)config_.dims[0];
    c.dim_y = (uint32_t)config_.dims[1];
    c.dim_z = (uint32_t)config_.dims[2];
    c.user_eb = config_.error_bound;
    c.value_base = computed_value_base_;
    c.zigzag_codes = config_.zigzag_codes ? 1 : 0;
    c.reserved[0] = config_.strict_error_bound ? 1 : 0;
    std::memcpy(buf, &c, sizeof(c));
    return sizeof(c);
}

// Accepts the 32/40/44-byte layouts (lorenzo_quant.h:243-298).
template <typename TI, typename TC>
void LorenzoQuantStage<TI, TC>::deserializeHeader(const uint8_t* buf, size_t size) {
    if (size < 32) throw std::runtime_error("Invalid Lorenzo config size");
    LorenzoQuantConfig c;
    std::memcpy(&c, buf, std::min(size, sizeof(LorenzoQuantConfig)));
    config_.error_bound = c.error_bound;
    computed_abs_eb_ = (TI)c.error_bound;  // lorenzo_quant.h:255
    config_.quant_radius = (TC)c.quant_radius;
    num_elements_ = c.num_elements;
    actual_outlier_count_ = c.outlier_count;
    if (size >= 40) {
        config_.eb_mode = (ErrorBoundMode)c.eb_mode;
        config_.precomputed_value_base = c.value_base;
        computed_value_base_ = c.value_base;
    } else {
        config_.eb_mode = ErrorBoundMode::ABS;
        config_.precomputed_value_base = 0.0f;
        computed_value_base_ = 0.0f;
    }
    config_.zigzag_codes = size >= sizeof(LorenzoQuantConfig) ? c.zigzag_codes != 0 : false;
    config_.strict_error_bound = size >= sizeof(LorenzoQuantConfig) && (c.reserved[0] & 1u) != 0;
    const int nd = c.ndim == 0 ? 1 : c.ndim;
    if (c.dim_x > 0) config_.dims[0] = c.dim_x;
    else if (c.num_elements > 0) config_.dims[0] = c.num_elements / (std::max<size_t>(1, c.dim_y) * std::max<size_t>(1, c.dim_z));
    else config_.dims[0] = 0;
    config_.dims[1] = nd >= 2 ? (c.dim_y > 0 ? c.dim_y : 1) : 1;
    config_.dims[2] = nd >= 3 ? (c.dim_z > 0 ? c.dim_z : 1) : 1;
}

template class LorenzoQuantStage<float, uint16_t>;
template class LorenzoQuantStage<float, uint8_t>;
template class LorenzoQuantStage<double, uint16_t>;
template class LorenzoQuantStage<double, uint32_t>;

// ===========================================================================
// Bitshuffle
// ===========================================================================
void BitshuffleStage::execute(hipStream_t stream, MemoryPool*, const std::vector<void*>& inputs,
                              const std::vector<void*>& outputs, const std::vector<size_t>& sizes) {
    if (inputs.empty() || outputs.empty() || sizes.empty()) throw std::runtime_error("BitshuffleStage: invalid inputs/outputs");
    if (element_width_ != 1 && element_width_ != 2 && element_width_ != 4 && element_width_ != 8)
        throw std::invalid_argument("BitshuffleStage: element_width must be 1, 2, 4, or 8");
    if (block_size_ == 0 || block_size_ % (1024u * element_width_) != 0)
        throw std::invalid_argument(
            "BitshuffleStage: block_size must be a positive multiple of 1024 * element_width");
    actual_output_size_ = sizes[0];
    if (sizes[0] == 0) return;
    const int rc = is_inverse_ ? fzh_bitshuffle_dec(stream, inputs[0], outputs[0], sizes[0], block_size_, element_width_)
                               : fzh_bitshuffle_enc(stream, inputs[0], outputs[0], sizes[0], block_size_, element_width_);
    throwIfFzh(rc, "BitshuffleStage");
}

// ===========================================================================
// RZE
// ===========================================================================
std::vector<size_t> RZEStage::estimateOutputSizes(const std::vector<size_t>& in) const {
    if (is_inverse_) return {cached_orig_bytes_ ? (size_t)cached_orig_bytes_ : (in.empty() ? 0 : in[0])};
    return {fzh_rze_enc_bound(in.empty() ? 0 : in[0])};
}

void RZEStage::execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                       const std::vector<void*>& outputs, const std::vector<size_t>& sizes) {
    if (inputs.empty() || outputs.empty() || sizes.empty()) throw std::runtime_error("RZEStage: invalid inputs/outputs");
    if (chunk_size_ != 16384)
        throw std::runtime_error("RZEStage: only chunk_size == 16384 is currently supported; got " + std::to_string(chunk_size_));
    if (levels_ != 4)
        throw std::runtime_error("RZEStage: only levels == 4 is currently supported; got " + std::to_string((int)levels_));
    small_.ensure(256, pool, stream);
    clearReadbacks();
    size_pending_ = status_pending_ = false;
    pack_deferred_ = false;
    if (!is_inverse_) {
        const size_t in_bytes = sizes[0];
        if (in_bytes == 0) { actual_output_size_ = 0; return; }
        cached_orig_bytes_ = (uint32_t)in_bytes;
        ws_.ensure(fzh_rze_enc_ws_bytes(in_bytes), pool, stream);
        static_assert(sizeof(fzh::RzePackDesc) <= sizeof(pack_desc_), "pack description storage");
        if (defer_pack_) {
            throwIfFzh(fzh::rze_enc_prepare(stream, inputs[0], in_bytes, pre_shuffle_, outputs[0],
                                            fzh_rze_enc_bound(in_bytes), (uint64_t*)small_.ptr, ws_.ptr, ws_.bytes,
                                            reinterpret_cast<fzh::RzePackDesc*>(pack_desc_)),
                       "RZEStage");
            pack_deferred_ = true;
        } else if (pre_shuffle_)
            throwIfFzh(fzh_rze_enc_shuffled(stream, inputs[0], in_bytes, pre_shuffle_, outputs[0],
                                            fzh_rze_enc_bound(in_bytes), (uint64_t*)small_.ptr, ws_.ptr, ws_.bytes),
                       "RZEStage (fused Bitshuffle)");
        else
            throwIfFzh(fzh_rze_enc(stream, inputs[0], in_bytes, outputs[0], fzh_rze_enc_bound(in_bytes),
                                   (uint64_t*)small_.ptr, ws_.ptr, ws_.bytes),
                       "RZEStage");
        actual_output_size_ = fzh_rze_enc_bound(in_bytes);
        size_pending_ = true;
        addReadback(small_.ptr, &h_size_, 8);
    } else {
        const size_t orig = cached_orig_bytes_;
        if (orig == 0) { actual_output_size_ = 0; return; }
        inv_ws_.ensure(fzh_rze_dec_ws_bytes(orig), pool, stream);
        if (pre_shuffle_)
            throwIfFzh(fzh_rze_dec_shuffled(stream, inputs[0], sizes[0], outputs[0], orig, pre_shuffle_,
                                            (int*)((uint8_t*)small_.ptr + 8), inv_ws_.ptr, inv_ws_.bytes),
                       "RZEStage (inverse, fused Bitshuffle)");
        else
            throwIfFzh(fzh_rze_dec(stream, inputs[0], sizes[0], outputs[0], orig, (int*)((uint8_t*)small_.ptr + 8),
                                   inv_ws_.ptr, inv_ws_.bytes),
                       "RZEStage (inverse)");
        actual_output_size_ = orig;
        status_pending_ = true;
        addReadback((uint8_t*)small_.ptr + 8, &h_status_, 4);
    }
}

void RZEStage::postStreamSync(hipStream_t) {
    if (!size_pending_ && !status_pending_) return;
    completeReadbacks();
    if (size_pending_) {
        if (h_size_ == 0) throw std::runtime_error("RZEStage: device prefix failed (workspace state)");
        actual_output_size_ = h_size_;
    }
    if (status_pending_) {
        const int st = h_status_;
        if (st != 0) {
            const int zero = 0;
            FZ_HIP(hipMemcpy((uint8_t*)small_.ptr + 8, &zero, 4, hipMemcpyHostToDevice));
            throw std::runtime_error("RZEStage (inverse): malformed or mismatching RZE stream");
        }
    }
}

// ===========================================================================
// RLE
// ===========================================================================
template <typename T>
void RLEStage<T>::execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                          const std::vector<void*>& outputs, const std::vector<size_t>& sizes) {
    if (inputs.empty() || outputs.empty() || sizes.empty()) throw std::runtime_error("RLEStage: Invalid inputs/outputs");
    small_.ensure(256, pool, stream);
    clearReadbacks();
    size_pending_ = status_pending_ = false;
    if (!is_inverse_) {
        const size_t n = sizes[0] / sizeof(T);
        cached_num_elements_ = (uint32_t)n;
        ws_.ensure(fzh_rle_enc_ws_bytes(n), pool, stream);
        throwIfFzh(fzh_rle_enc(stream, inputs[0], n, (int)sizeof(T), outputs[0], fzh_rle_enc_bound(n, sizeof(T)),
                               (uint64_t*)small_.ptr, ws_.ptr, ws_.bytes),
                   "RLEStage");
        actual_output_size_ = fzh_rle_enc_bound(n, sizeof(T));
        size_pending_ = true;
        addReadback(small_.ptr, &h_size_, 8);
    } else {
        const size_t n = cached_num_elements_;
        actual_output_size_ = n * sizeof(T);
        if (n == 0) return;
        inv_ws_.ensure(fzh_rle_dec_ws_bytes(n), pool, stream);
        throwIfFzh(fzh_rle_dec(stream, inputs[0], sizes[0], (int)sizeof(T), outputs[0], n,
                               (int*)((uint8_t*)small_.ptr + 8), inv_ws_.ptr, inv_ws_.bytes),
                   "RLEStage (inverse)");
        status_pending_ = true;
        addReadback((uint8_t*)small_.ptr + 8, &h_status_, 4);
    }
}

template <typename T>
void RLEStage<T>::postStreamSync(hipStream_t) {
    if (!size_pending_ && !status_pending_) return;
    completeReadbacks();
    if (size_pending_) {
        if (h_size_ == 0) throw std::runtime_error("RLEStage: device prefix failed (workspace state)");
        actual_output_size_ = h_size_;
    }
    if (status_pending_) {
        const int st = h_status_;
        if (st != 0) {
            const int zero = 0;
            FZ_HIP(hipMemcpy((uint8_t*)small_.ptr + 8, &zero, 4, hipMemcpyHostToDevice));
            throw std::runtime_error("RLEStage (inverse): malformed RLE stream or wrong element count");
        }
    }
}

template class RLEStage<uint8_t>;
template class RLEStage<uint16_t>;
template class RLEStage<uint32_t>;
template class RLEStage<int32_t>;

// ===========================================================================
// Huffman
// ===========================================================================
template <typename T>
std::vector<size_t> HuffmanStage<T>::estimateOutputSizes(const std::vector<size_t>& in) const {
    if (is_inverse_) return {(size_t)original_len_ * sizeof(T)};
    const size_t n = in.empty() ? 0 : in[0] / sizeof(T);
    return {fzh_huffman_enc_bound(n, bklen_, sizeof(T))};
}

template <typename T>
void HuffmanStage<T>::execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                              const std::vector<void*>& outputs, const std::vector<size_t>& sizes) {
    if (inputs.empty() || outputs.empty() || sizes.empty())
        throw std::runtime_error("HuffmanStage: inputs, outputs, and sizes must be non-empty");
    clearReadbacks();
    size_pending_ = status_pending_ = book_pending_ = false;
    if (sizes[0] == 0) { actual_output_size_ = 0; return; }
    small_.ensure(256, pool, stream);
    if (!is_inverse_ && device_book_) {
        // histogram -> codebook -> encode, all on the stream; the codebook
        // status (the host path's out-of-range check) is read after the sync
        if (encode_mode_ == HuffmanEncodeMode::Fine)
            throw std::runtime_error("HuffmanStage: the device codebook serves the coarse encoding only "
                                     "(the fine/coarse choice reads the codebook on the host)");
        const size_t n = sizes[0] / sizeof(T);
        const size_t rev_bytes = fzh_huffman_revbook_bytes(bklen_, sizeof(T));
        freq_.ensure(4 * (size_t)bklen_, pool, stream);
        book_.ensure(4 * (size_t)bklen_ + rev_bytes, pool, stream);
        ws_.ensure(fzh_huffman_enc_ws_bytes(n), pool, stream);
        throwIfFzh(fzh_histogram(stream, inputs[0], n, (int)sizeof(T), bklen_, (uint32_t*)freq_.ptr), "HuffmanStage histogram");
        throwIfFzh(fzh_huffman_build_book_device(stream, (const uint32_t*)freq_.ptr, bklen_, (int)sizeof(T), n,
                                                 (uint32_t*)book_.ptr, (uint8_t*)book_.ptr + 4 * (size_t)bklen_,
                                                 (int*)((uint8_t*)small_.ptr + 16)),
                   "HuffmanStage device codebook");
        throwIfFzh(fzh_huffman_encode(stream, inputs[0], n, (int)sizeof(T), bklen_, (const uint32_t*)book_.ptr,
                                      (const uint8_t*)book_.ptr + 4 * (size_t)bklen_, outputs[0],
                                      fzh_huffman_enc_bound(n, bklen_, sizeof(T)), (uint64_t*)small_.ptr, ws_.ptr,
                                      ws_.bytes),
                   "HuffmanStage encode");
        original_len_ = n;
        actual_output_size_ = fzh_huffman_enc_bound(n, bklen_, sizeof(T));
        size_pending_ = book_pending_ = true;
        addReadback(small_.ptr, &h_size_, 8);
        addReadback((uint8_t*)small_.ptr + 16, &h_book_status_, 4);
    } else if (!is_inverse_) {
        const size_t n = sizes[0] / sizeof(T);
        freq_.ensure(4 * (size_t)bklen_, pool, stream);
        throwIfFzh(fzh_histogram(stream, inputs[0], n, (int)sizeof(T), bklen_, (uint32_t*)freq_.ptr), "HuffmanStage histogram");
        std::vector<uint32_t> h(bklen_);
        throwIfFzh(fzh::read_small(stream, freq_.ptr, 4 * (size_t)bklen_, h.data()), "HuffmanStage histogram readback");
        uint64_t sum = 0;
        for (uint32_t v : h) sum += v;
        if (sum != (uint64_t)n)
            throw std::runtime_error("HuffmanStage: " + std::to_string(n - sum) +
                                     " out-of-range symbol(s) detected; all symbols must be in [0, " +
                                     std::to_string(bklen_) +
                                     "). Increase bklen or use setZigzagCodes(true) with LorenzoQuantStage.");
        const size_t rev_bytes = fzh_huffman_revbook_bytes(bklen_, sizeof(T));
        std::vector<uint8_t> book(4 * (size_t)bklen_ + rev_bytes);
        throwIfFzh(fzh_huffman_build_book(h.data(), bklen_, (int)sizeof(T), (uint32_t*)book.data(),
                                          book.data() + 4 * (size_t)bklen_),
                   "HuffmanStage codebook");
        book_.ensure(book.size(), pool, stream);
        FZ_HIP(hipMemcpyAsync(book_.ptr, book.data(), book.size(), hipMemcpyHostToDevice, stream));
        ws_.ensure(fzh_huffman_enc_ws_bytes(n), pool, stream);
        // Fine mode: 1024-symbol partitions when no codeword exceeds 8 bits
        // (hf_hl.cc:69-80), else the coarse stream
        uint32_t max_cl = 0;
        for (uint32_t i = 0; i < bklen_; i++) max_cl = std::max(max_cl, ((const uint32_t*)book.data())[i] >> 27);
        const bool fine = encode_mode_ == HuffmanEncodeMode::Fine && max_cl <= 8;
        throwIfFzh((fine ? fzh_huffman_encode_fine : fzh_huffman_encode)(
                       stream, inputs[0], n, (int)sizeof(T), bklen_, (const uint32_t*)book_.ptr,
                       (const uint8_t*)book_.ptr + 4 * (size_t)bklen_, outputs[0],
                       fzh_huffman_enc_bound(n, bklen_, sizeof(T)), (uint64_t*)small_.ptr, ws_.ptr, ws_.bytes),
                   "HuffmanStage encode");
        // keep the host copy alive until the copy has been consumed
        FZ_HIP(hipStreamSynchronize(stream));
        original_len_ = n;
        actual_output_size_ = fzh_huffman_enc_bound(n, bklen_, sizeof(T));
        size_pending_ = true;
        addReadback(small_.ptr, &h_size_, 8);
    } else {
        if (original_len_ == 0)
            throw std::runtime_error("HuffmanStage: inverse called with original_len_=0; call deserializeHeader() before decompressing");
        throwIfFzh(fzh_huffman_decode(stream, inputs[0], sizes[0], (int)sizeof(T), outputs[0], original_len_,
                                      (int*)((uint8_t*)small_.ptr + 8)),
                   "HuffmanStage decode");
        actual_output_size_ = (size_t)original_len_ * sizeof(T);
        status_pending_ = true;
        addReadback((uint8_t*)small_.ptr + 8, &h_status_, 4);
    }
}

template <typename T>
void HuffmanStage<T>::postStreamSync(hipStream_t) {
    if (!size_pending_ && !status_pending_) return;
    completeReadbacks();
    if (book_pending_ && h_book_status_ != 0) {
        size_pending_ = book_pending_ = false;
        if (h_book_status_ == 1)
            throw std::runtime_error("HuffmanStage: out-of-range symbol(s) detected; all symbols must be in [0, " +
                                     std::to_string(bklen_) +
                                     "). Increase bklen or use setZigzagCodes(true) with LorenzoQuantStage.");
        throw std::runtime_error("HuffmanStage: device codebook: a code would exceed 58 bits");
    }
    if (size_pending_) {
        if (h_size_ == 0) throw std::runtime_error("HuffmanStage: device prefix failed (workspace state)");
        actual_output_size_ = h_size_;
    }
    if (status_pending_) {
        const int st = h_status_;
        if (st != 0) {
            const int zero = 0;
            FZ_HIP(hipMemcpy((uint8_t*)small_.ptr + 8, &zero, 4, hipMemcpyHostToDevice));
            throw std::runtime_error("HuffmanStage (inverse): malformed Huffman stream");
        }
    }
}

template class HuffmanStage<uint8_t>;
template class HuffmanStage<uint16_t>;
template class HuffmanStage<uint32_t>;

// ===========================================================================
// Quantizer (ABS / NOA)
// ===========================================================================
template <typename TI, typename TC>
std::vector<size_t> QuantizerStage<TI, TC>::estimateOutputSizes(const std::vector<size_t>& in) const {
    if (is_inverse_) return {num_elements_ * sizeof(TI)};
    const size_t n = in.empty() ? 0 : in[0] / sizeof(TI);
    if (isInplaceMode()) return {n * sizeof(TC)};
    const size_t m = maxOutliers(n);
    return {n * sizeof(TC), m * sizeof(TI), m * 4, 4};
}

template <typename TI, typename TC>
std::unordered_map<std::string, size_t> QuantizerStage<TI, TC>::getActualOutputSizesByName() const {
    std::unordered_map<std::string, size_t> r;
    const auto names = getOutputNames();
    for (size_t i = 0; i < names.size() && i < actual_output_sizes_.size(); i++) r[names[i]] = actual_output_sizes_[i];
    return r;
}

template <typename TI, typename TC>
void QuantizerStage<TI, TC>::execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                                     const std::vector<void*>& outputs, const std::vector<size_t>& sizes) {
    if (isInplaceMode()) {  // quantizer.cu:396-452 / 492-615, one port each way
        if (!config_.zigzag_codes)
            throw std::runtime_error("QuantizerStage: setInplaceOutliers(true) requires setZigzagCodes(true)");
        if (sizeof(TC) != sizeof(TI))
            throw std::runtime_error("QuantizerStage: setInplaceOutliers(true) requires sizeof(TCode)==sizeof(TInput) "
                                     "(use QuantizerStage<float, uint32_t>)");
    }
    if (is_inverse_ && isInplaceMode()) {
        if (inputs.empty() || outputs.empty() || sizes.empty())
            throw std::runtime_error("QuantizerStage (inverse, inplace): requires 1 input and 1 output");
        const size_t n = sizes[0] / sizeof(TC);
        actual_output_sizes_.assign(1, n * sizeof(TI));
        throwIfFzh(fzh_quant_abs_inv_inplace_f32(stream, (const uint32_t*)inputs[0], n, 2.0f * computed_abs_eb_,
                                                 (uint32_t)config_.quant_radius, (float*)outputs[0]),
                   "QuantizerStage (inverse, inplace)");
        return;
    }
    if (is_inverse_) {
        if (inputs.size() < 4 || outputs.empty() || sizes.size() < 4)
            throw std::runtime_error("QuantizerStage (inverse): requires 4 inputs");
        const size_t n = sizes[0] / sizeof(TC);
        actual_output_sizes_.assign(1, n * sizeof(TI));
        if (config_.eb_mode == ErrorBoundMode::REL) {  // quantizer.cu:426-436: log2eb from the user bound
            throwIfFzh(fzh_quant_rel_inv_f32(stream, inputs[0], n, config_.error_bound, (int)sizeof(TC),
                                             (const float*)inputs[1], (const uint32_t*)inputs[2],
                                             (const uint32_t*)inputs[3], sizes[1] / sizeof(TI), (float*)outputs[0]),
                       "QuantizerStage (inverse, REL)");
            return;
        }
        throwIfFzh(fzh_quant_abs_inv_f32(stream, inputs[0], n, 2.0f * computed_abs_eb_, config_.zigzag_codes ? 1 : 0,
                                         (int)sizeof(TC), (const float*)inputs[1], (const uint32_t*)inputs[2],
                                         (const uint32_t*)inputs[3], sizes[1] / sizeof(TI), (float*)outputs[0]),
                   "QuantizerStage (inverse)");
        return;
    }
    if (inputs.empty() || outputs.size() < (isInplaceMode() ? 1u : 4u) || sizes.empty())
        throw std::runtime_error(isInplaceMode() ? "QuantizerStage (inplace): requires 1 input and 1 output"
                                                 : "QuantizerStage: requires 1 input and 4 outputs");
    const size_t n = sizes[0] / sizeof(TI);
    num_elements_ = n;
    const size_t m = maxOutliers(n);
    clearReadbacks();
    d_count_ptr_ = nullptr;
    if (n == 0) { actual_output_sizes_.assign(isInplaceMode() ? 1 : 4, 0); return; }
    if (config_.eb_mode == ErrorBoundMode::REL) {  // quantizer.cu:530-533, 563-580
        computed_abs_eb_ = 0.0f;
        computed_value_base_ = 0.0f;
        ws_.ensure(256 + fzh_quant_abs_fwd_ws_bytes(n), pool, stream);
        throwIfFzh(fzh_quant_rel_fwd_f32(stream, (const float*)inputs[0], n, config_.error_bound,
                                         config_.outlier_threshold, (uint32_t)config_.quant_radius, (int)sizeof(TC),
                                         outputs[0], (float*)outputs[1], (uint32_t*)outputs[2], (uint32_t*)outputs[3],
                                         m, (uint8_t*)ws_.ptr + 256, ws_.bytes - 256),
                   "QuantizerStage (REL)");
        d_count_ptr_ = outputs[3];
        addReadback(d_count_ptr_, &h_count_, 4);
        actual_output_sizes_ = {n * sizeof(TC), m * sizeof(TI), m * 4, 4};
        return;
    }
    if (config_.eb_mode == ErrorBoundMode::ABS) {
        computed_abs_eb_ = config_.error_bound;
        computed_value_base_ = 0.0f;
    } else {
        float vb = config_.precomputed_value_base;
        if (vb <= 0.0f) {
            vb_.ensure(256 + fzh_value_base_ws_bytes(n), pool, stream);
            throwIfFzh(fzh_value_base_f32(stream, (const float*)inputs[0], n, 2, config_.error_bound, (float*)vb_.ptr,
                                          (uint8_t*)vb_.ptr + 256, vb_.bytes - 256),
                       "QuantizerStage value range");
            float p[4];
            throwIfFzh(fzh::read_small(stream, vb_.ptr, 16, p), "QuantizerStage value range readback");
            vb = p[0];
        }
        computed_value_base_ = vb;
        computed_abs_eb_ = vb > 0.0f ? config_.error_bound * vb : config_.error_bound;
    }
    const float r = 1.0f / (2.0f * computed_abs_eb_);
    if (isInplaceMode()) {
        throwIfFzh(fzh_quant_abs_fwd_inplace_f32(stream, (const float*)inputs[0], n, r, config_.outlier_threshold,
                                                 (uint32_t)config_.quant_radius, (uint32_t*)outputs[0]),
                   "QuantizerStage (inplace)");
        outlier_count_ = 0;
        actual_output_sizes_ = {n * sizeof(TC)};
        return;
    }
    ws_.ensure(256 + fzh_quant_abs_fwd_ws_bytes(n), pool, stream);
    throwIfFzh(fzh_quant_abs_fwd_f32(stream, (const float*)inputs[0], n, r, config_.outlier_threshold,
                                     (uint32_t)config_.quant_radius, config_.zigzag_codes ? 1 : 0, (int)sizeof(TC),
                                     outputs[0], (float*)outputs[1], (uint32_t*)outputs[2], (uint32_t*)outputs[3], m,
                                     (uint8_t*)ws_.ptr + 256, ws_.bytes - 256),
               "QuantizerStage");
    d_count_ptr_ = outputs[3];
    addReadback(d_count_ptr_, &h_count_, 4);
    actual_output_sizes_ = {n * sizeof(TC), m * sizeof(TI), m * 4, 4};
}

template <typename TI, typename TC>
void QuantizerStage<TI, TC>::postStreamSync(hipStream_t) {
    if (is_inverse_ || !d_count_ptr_) return;
    completeReadbacks();
    uint32_t cnt = h_count_;
    if (cnt == 0xFFFFFFFFu) throw std::runtime_error("QuantizerStage: device prefix failed (workspace state)");
    cnt = (uint32_t)std::min<size_t>(cnt, maxOutliers(num_elements_));
    outlier_count_ = cnt;
    actual_output_sizes_[1] = (size_t)cnt * sizeof(TI);
    actual_output_sizes_[2] = (size_t)cnt * 4;
}

template <typename TI, typename TC>
size_t QuantizerStage<TI, TC>::serializeHeader(size_t, uint8_t* buf, size_t max_size) const {
    if (max_size < sizeof(QuantizerConfig)) throw std::runtime_error("Insufficient buffer for Quantizer config");
    QuantizerConfig c;
    c.abs_error_bound = computed_abs_eb_;
    c.user_error_bound = config_.error_bound;
    c.value_base = computed_value_base_;
    c.quant_radius = (uint32_t)config_.quant_radius;
    c.num_elements = (uint32_t)num_elements_;
    c.outlier_count = outlier_count_;
    c.input_type = DataType::FLOAT32;
    c.code_type = dataTypeOf<TC>();
    c.eb_mode = (uint8_t)config_.eb_mode;
    c.zigzag_codes = config_.zigzag_codes ? 1 : 0;
    c.outlier_threshold = config_.outlier_threshold;
    c.inplace_outliers = config_.inplace_outliers ? 1 : 0;
    std::memcpy(buf, &c, sizeof(c));
    return sizeof(c);
}

template <typename TI, typename TC>
void QuantizerStage<TI, TC>::deserializeHeader(const uint8_t* buf, size_t size) {
    QuantizerConfig c;
    std::memcpy(&c, buf, std::min(size, sizeof(c)));
    computed_abs_eb_ = c.abs_error_bound;
    config_.error_bound = c.user_error_bound;
    computed_value_base_ = c.value_base;
    config_.quant_radius = (int)c.quant_radius;
    num_elements_ = c.num_elements;
    outlier_count_ = c.outlier_count;
    config_.eb_mode = (ErrorBoundMode)c.eb_mode;
    config_.zigzag_codes = c.zigzag_codes != 0;
    // quantizer.cu:714-745: headers shorter than 36 bytes predate the last two fields
    if (size >= sizeof(QuantizerConfig)) {
        config_.outlier_threshold = c.outlier_threshold;
        config_.inplace_outliers = c.inplace_outliers != 0;
    } else {
        config_.outlier_threshold = INFINITY;
        config_.inplace_outliers = false;
    }
}

template class QuantizerStage<float, uint8_t>;
template class QuantizerStage<float, uint16_t>;
template class QuantizerStage<float, uint32_t>;

// ===========================================================================
// Bitpack
// ===========================================================================
template <typename T>
void BitpackStage<T>::execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                              const std::vector<void*>& outputs, const std::vector<size_t>& sizes) {
    if (inputs.empty() || outputs.empty() || sizes.empty())
        throw std::runtime_error("BitpackStage: inputs, outputs, and sizes must be non-empty");
    if (!is_inverse_) {
        const size_t n = sizes[0] / sizeof(T);
        num_elements_ = n;
        if (n == 0) { actual_output_size_ = 0; return; }
        if (auto_detect_) {
            ws_.ensure(256 + fzh_max_u32_ws_bytes(), pool, stream);
            throwIfFzh(fzh_max_u32(stream, inputs[0], n, (int)sizeof(T), (uint32_t*)ws_.ptr, (uint8_t*)ws_.ptr + 256,
                                   ws_.bytes - 256),
                       "BitpackStage max");
            uint32_t mx = 0;
            throwIfFzh(fzh::read_small(stream, ws_.ptr, 4, &mx), "BitpackStage max readback");
            uint8_t nb = 8 * sizeof(T);
            for (uint8_t b = 1; b < 8 * sizeof(T); b = (uint8_t)(b * 2))
                if ((uint64_t)mx < (1ull << b)) { nb = b; break; }
            nbits_ = nb;
        }
        throwIfFzh(fzh_bitpack_enc(stream, inputs[0], n, (int)sizeof(T), nbits_, outputs[0]), "BitpackStage");
        actual_output_size_ = (n * nbits_ + 7) / 8;
    } else {
        const size_t n = num_elements_;
        actual_output_size_ = n * sizeof(T);
        if (n == 0) return;
        throwIfFzh(fzh_bitpack_dec(stream, inputs[0], n, (int)sizeof(T), nbits_, outputs[0]), "BitpackStage (inverse)");
    }
}

template class BitpackStage<uint8_t>;
template class BitpackStage<uint16_t>;
template class BitpackStage<uint32_t>;

// ===========================================================================
// Factory (stage_factory.h:36-248): types supported by this build
// ===========================================================================
Stage* createStage(StageType type, const uint8_t* config, size_t config_size) {
    Stage* s = nullptr;
    switch (type) {
    case StageType::LORENZO_QUANT: {
        if (config_size < sizeof(LorenzoQuantConfig))
            throw std::runtime_error("Lorenzo config too small: " + std::to_string(config_size));
        LorenzoQuantConfig lc;
        std::memcpy(&lc, config, sizeof(lc));
        if (lc.input_type == DataType::FLOAT32 && lc.code_type == DataType::UINT16) s = new LorenzoQuantStage<float, uint16_t>();
        else if (lc.input_type == DataType::FLOAT32 && lc.code_type == DataType::UINT8) s = new LorenzoQuantStage<float, uint8_t>();
        else if (lc.input_type == DataType::FLOAT64 && lc.code_type == DataType::UINT16) s = new LorenzoQuantStage<double, uint16_t>();
        else if (lc.input_type == DataType::FLOAT64 && lc.code_type == DataType::UINT32) s = new LorenzoQuantStage<double, uint32_t>();
        else throw std::runtime_error("Unsupported Lorenzo template instantiation: input_type=" +
                                      std::to_string((int)lc.input_type) + " code_type=" + std::to_string((int)lc.code_type));
        break;
    }
    case StageType::BITSHUFFLE: s = new BitshuffleStage(); break;
    case StageType::RZE: s = new RZEStage(); break;
    case StageType::RLE: {
        const DataType dt = config_size >= 1 ? (DataType)config[0] : DataType::UINT16;
        switch (dt) {
        case DataType::UINT8: s = new RLEStage<uint8_t>(); break;
        case DataType::UINT16: s = new RLEStage<uint16_t>(); break;
        case DataType::UINT32: s = new RLEStage<uint32_t>(); break;
        case DataType::INT32: s = new RLEStage<int32_t>(); break;
        default: throw std::runtime_error("Unsupported RLE data type " + std::to_string((int)dt));
        }
        break;
    }
    case StageType::HUFFMAN: {
        const DataType dt = config_size >= 1 ? (DataType)config[0] : DataType::UINT16;
        if (dt == DataType::UINT8) s = new HuffmanStage<uint8_t>();
        else if (dt == DataType::UINT32) s = new HuffmanStage<uint32_t>();
        else s = new HuffmanStage<uint16_t>();
        break;
    }
    case StageType::QUANTIZER: {
        QuantizerConfig qc;
        std::memcpy(&qc, config, std::min(config_size, sizeof(qc)));
        if (qc.code_type == DataType::UINT8) s = new QuantizerStage<float, uint8_t>();
        else if (qc.code_type == DataType::UINT32) s = new QuantizerStage<float, uint32_t>();
        else s = new QuantizerStage<float, uint16_t>();
        break;
    }
    case StageType::BITPACK: {
        const DataType dt = config_size >= 1 ? (DataType)config[0] : DataType::UINT16;
        if (dt == DataType::UINT8) s = new BitpackStage<uint8_t>();
        else if (dt == DataType::UINT32) s = new BitpackStage<uint32_t>();
        else s = new BitpackStage<uint16_t>();
        break;
    }
    case StageType::DIFFERENCE: {
        // [0] T, [1] TOut, [2..5] chunk size; a 1-byte legacy header names T only
        const DataType ti = config_size >= 1 ? (DataType)config[0] : DataType::FLOAT32;
        const DataType to = config_size >= 2 ? (DataType)config[1] : ti;
        if (ti == DataType::INT8 && to == DataType::UINT8) s = new DifferenceStage<int8_t, uint8_t>();
        else if (ti == DataType::INT16 && to == DataType::UINT16) s = new DifferenceStage<int16_t, uint16_t>();
        else if (ti == DataType::INT32 && to == DataType::UINT32) s = new DifferenceStage<int32_t, uint32_t>();
        else if (ti == DataType::INT64 && to == DataType::UINT64) s = new DifferenceStage<int64_t, uint64_t>();
        else if (ti == DataType::FLOAT32) s = new DifferenceStage<float>();
        else if (ti == DataType::FLOAT64) s = new DifferenceStage<double>();
        else if (ti == DataType::UINT8) s = new DifferenceStage<uint8_t>();
        else if (ti == DataType::UINT16) s = new DifferenceStage<uint16_t>();
        else if (ti == DataType::UINT32) s = new DifferenceStage<uint32_t>();
        else if (ti == DataType::INT32) s = new DifferenceStage<int32_t>();
        else if (ti == DataType::INT64) s = new DifferenceStage<int64_t>();
        else throw std::runtime_error("Unsupported Difference data type: " + std::to_string((int)ti));
        break;
    }
    case StageType::ZIGZAG:
    case StageType::NEGABINARY: {
        const bool nb = type == StageType::NEGABINARY;
        const DataType ti = config_size >= 2 ? (DataType)config[0] : DataType::INT32;
        const DataType to = config_size >= 2 ? (DataType)config[1] : DataType::UINT32;
        if (ti == DataType::INT8 && to == DataType::UINT8) s = nb ? (Stage*)new NegabinaryStage<int8_t, uint8_t>() : new ZigzagStage<int8_t, uint8_t>();
        else if (ti == DataType::INT16 && to == DataType::UINT16) s = nb ? (Stage*)new NegabinaryStage<int16_t, uint16_t>() : new ZigzagStage<int16_t, uint16_t>();
        else if (ti == DataType::INT32 && to == DataType::UINT32) s = nb ? (Stage*)new NegabinaryStage<int32_t, uint32_t>() : new ZigzagStage<int32_t, uint32_t>();
        else if (ti == DataType::INT64 && to == DataType::UINT64) s = nb ? (Stage*)new NegabinaryStage<int64_t, uint64_t>() : new ZigzagStage<int64_t, uint64_t>();
        else throw std::runtime_error(std::string(nb ? "Unsupported NegabinaryStage" : "Unsupported ZigzagStage") +
                                      " type pair: TIn=" + std::to_string((int)ti) + " TOut=" + std::to_string((int)to));
        break;
    }
    case StageType::LORENZO: {
        const DataType dt = config_size >= sizeof(LorenzoConfig) ? (DataType)config[0] : DataType::INT32;
        if (dt == DataType::INT8) s = new LorenzoStage<int8_t>();
        else if (dt == DataType::INT16) s = new LorenzoStage<int16_t>();
        else if (dt == DataType::INT32) s = new LorenzoStage<int32_t>();
        else if (dt == DataType::INT64) s = new LorenzoStage<int64_t>();
        else throw std::runtime_error("Unsupported LorenzoStage DataType: " + std::to_string((int)dt));
        if (config_size < sizeof(LorenzoConfig)) return s;  // no header: default dims
        break;
    }
    default:
        throw std::runtime_error("createStage: stage type " + stageTypeToString(type) + " (" +
                                 std::to_string((int)type) + ") is not part of this build");
    }
    s->deserializeHeader(config, config_size);
    return s;
}

}  // namespace fz

// host_transforms.cpp — DifferenceStage, ZigzagStage / NegabinaryStage and
// the integer LorenzoStage on top of the fzh_* transform kernels
// (transform.hip).  Semantics and headers: modules/predictors/diff/diff.cu,
// modules/transforms/{zigzag,negabinary}/*_stage.cu,
// modules/predictors/lorenzo/lorenzo_stage.cu.
#include <stdexcept>
#include <string>

#include "fz/stages.h"
#include "fzh_internal.h"

namespace fz {

void throwIfFzh(int rc, const char* what);

namespace {
template <typename T> constexpr int dtypeCode() { return (int)dataTypeOf<T>(); }
}  // namespace

// ---------------------------------------------------------------------------
// Difference
// ---------------------------------------------------------------------------
template <typename T, typename TOut>
void DifferenceStage<T, TOut>::execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                                       const std::vector<void*>& outputs, const std::vector<size_t>& sizes) {
    if (inputs.empty() || outputs.empty() || sizes.empty())
        throw std::runtime_error("DifferenceStage: inputs, outputs, and sizes vectors must all be non-empty");
    const size_t bytes = sizes[0];
    const size_t n = bytes / sizeof(T);
    if (chunk_size_ % sizeof(T) != 0)
        throw std::runtime_error("DifferenceStage: chunk size must be a multiple of the element size");
    const size_t chunk = chunk_size_ / sizeof(T);
    actual_output_size_ = n == 0 ? 0 : bytes;
    if (n == 0) return;
    const int nb = std::is_same_v<T, TOut> ? 0 : 1;
    if (!is_inverse_) {
        throwIfFzh(fzh_diff_fwd(stream, inputs[0], outputs[0], n, dtypeCode<T>(), chunk, nb), "DifferenceStage");
    } else {
        ws_.ensure(fzh_diff_inv_ws_bytes(n, dtypeCode<T>(), chunk), pool, stream);
        throwIfFzh(fzh_diff_inv(stream, inputs[0], outputs[0], n, dtypeCode<T>(), chunk, nb, ws_.ptr, ws_.bytes),
                   "DifferenceStage (inverse)");
    }
}

template class DifferenceStage<float>;
template class DifferenceStage<double>;
template class DifferenceStage<int32_t>;
template class DifferenceStage<int64_t>;
template class DifferenceStage<uint16_t>;
template class DifferenceStage<uint8_t>;
template class DifferenceStage<uint32_t>;
template class DifferenceStage<int8_t, uint8_t>;
template class DifferenceStage<int16_t, uint16_t>;
template class DifferenceStage<int32_t, uint32_t>;
template class DifferenceStage<int64_t, uint64_t>;

// ---------------------------------------------------------------------------
// Zigzag / Negabinary
// ---------------------------------------------------------------------------
template <bool NEGABINARY, typename TIn, typename TOut>
void IntCodeStage<NEGABINARY, TIn, TOut>::execute(hipStream_t stream, MemoryPool*, const std::vector<void*>& inputs,
                                                  const std::vector<void*>& outputs,
                                                  const std::vector<size_t>& sizes) {
    if (inputs.empty() || outputs.empty() || sizes.empty())
        throw std::runtime_error(getName() + "Stage: invalid inputs/outputs");
    const size_t bytes = sizes[0];
    actual_output_size_ = bytes;
    const size_t n = bytes / sizeof(TIn);
    if (n == 0) return;
    const int rc = NEGABINARY ? fzh_negabinary(stream, inputs[0], outputs[0], n, (int)sizeof(TIn), is_inverse_ ? 1 : 0)
                              : fzh_zigzag(stream, inputs[0], outputs[0], n, (int)sizeof(TIn), is_inverse_ ? 1 : 0);
    throwIfFzh(rc, NEGABINARY ? "NegabinaryStage" : "ZigzagStage");
}

template class IntCodeStage<false, int8_t, uint8_t>;
template class IntCodeStage<false, int16_t, uint16_t>;
template class IntCodeStage<false, int32_t, uint32_t>;
template class IntCodeStage<false, int64_t, uint64_t>;
template class IntCodeStage<true, int8_t, uint8_t>;
template class IntCodeStage<true, int16_t, uint16_t>;
template class IntCodeStage<true, int32_t, uint32_t>;
template class IntCodeStage<true, int64_t, uint64_t>;

// ---------------------------------------------------------------------------
// Integer Lorenzo
// ---------------------------------------------------------------------------
template <typename T>
void LorenzoStage<T>::execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                              const std::vector<void*>& outputs, const std::vector<size_t>& sizes) {
    if (inputs.empty() || outputs.empty() || sizes.empty())
        throw std::runtime_error("LorenzoStage: inputs, outputs, and sizes must be non-empty");
    const size_t bytes = sizes[0];
    actual_output_size_ = bytes;
    if (bytes == 0) return;
    const size_t n = bytes / sizeof(T);
    // lorenzo_stage.cu:384-388: dim_x 0 means a flat 1-D array of n elements
    const int nd = ndim();
    const size_t nx = nd == 1 ? n : (dims_[0] > 0 ? dims_[0] : n);
    const size_t ny = nd >= 2 ? dims_[1] : 1, nz = nd >= 3 ? dims_[2] : 1;
    if (nx * ny * nz > n) throw std::runtime_error("LorenzoStage: dims exceed the input");
    if (!is_inverse_) {
        throwIfFzh(fzh_lorenzo_int_fwd(stream, inputs[0], outputs[0], nx, ny, nz, (int)sizeof(T)), "LorenzoStage");
    } else {
        ws_.ensure(fzh_lorenzo_int_inv_ws_bytes(nx, ny, nz), pool, stream);
        throwIfFzh(fzh_lorenzo_int_inv(stream, inputs[0], outputs[0], nx, ny, nz, (int)sizeof(T), ws_.ptr, ws_.bytes),
                   "LorenzoStage (inverse)");
    }
    // elements past the field (a padded input) are written as zero
    const size_t used = nx * ny * nz * sizeof(T);
    if (used < bytes) throwIfFzh(fzh::zero_bytes(stream, (uint8_t*)outputs[0] + used, bytes - used), "LorenzoStage pad");
}

template class LorenzoStage<int8_t>;
template class LorenzoStage<int16_t>;
template class LorenzoStage<int32_t>;
template class LorenzoStage<int64_t>;

}  // namespace fz

// fz/stages.h — the hot-path stages, backed by the fzh_* HIP kernels.
//
// Class names, configuration fields, setters/getters, port names and header
// layouts follow the reference:
//   LorenzoQuantStage   modules/fused/lorenzo_quant/lorenzo_quant.h:43-334
//   BitshuffleStage     modules/shufflers/bitshuffle/bitshuffle_stage.h:38-151
//   RZEStage            modules/coders/rze/rze_stage.h:47-249
//   RLEStage            modules/coders/rle/rle.h:31-207
//   HuffmanStage        modules/coders/huffman/huffman_stage.h:61-256
//   QuantizerStage      modules/quantizers/quantizer/quantizer.h:25-260 (ABS/NOA)
//   BitpackStage        modules/coders/bitpack/bitpack_stage.h (nbits / auto)
//   DifferenceStage     modules/predictors/diff/diff.h:16-200
//   ZigzagStage         modules/transforms/zigzag/zigzag_stage.h
//   NegabinaryStage     modules/transforms/negabinary/negabinary_stage.h
//   LorenzoStage        modules/predictors/lorenzo/lorenzo_stage.h (integer)
// Differences in behaviour are confined to where the reference is
// nondeterministic or waits on the host: outlier lists come out in a fixed
// order, REL/NOA error bounds and all inverse paths resolve on the device
// without a host round trip, and padding bytes are always zero.
#pragma once

#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>

#include "fz/mempool.h"
#include "fz/stage.h"

namespace fz {

enum class ErrorBoundMode : uint8_t { ABS = 0, REL = 1, NOA = 2 };

// Device scratch owned by a stage, zero-initialised once on allocation.
struct DeviceScratch {
    void* ptr = nullptr;
    size_t bytes = 0;
    size_t last_need = 0;
    MemoryPool* pool = nullptr;
    // Holds at least `need` bytes; zeroed whenever `need` differs from the
    // previous call's (a fill kernel on `s`, so steady-state calls add nothing).
    void ensure(size_t need, MemoryPool* p, hipStream_t s);
    void release();
    ~DeviceScratch() { release(); }
};

template <typename T> constexpr DataType dataTypeOf() {
    if constexpr (std::is_same_v<T, uint8_t>) return DataType::UINT8;
    else if constexpr (std::is_same_v<T, uint16_t>) return DataType::UINT16;
    else if constexpr (std::is_same_v<T, uint32_t>) return DataType::UINT32;
    else if constexpr (std::is_same_v<T, uint64_t>) return DataType::UINT64;
    else if constexpr (std::is_same_v<T, int8_t>) return DataType::INT8;
    else if constexpr (std::is_same_v<T, int16_t>) return DataType::INT16;
    else if constexpr (std::is_same_v<T, int32_t>) return DataType::INT32;
    else if constexpr (std::is_same_v<T, int64_t>) return DataType::INT64;
    else if constexpr (std::is_same_v<T, float>) return DataType::FLOAT32;
    else return DataType::FLOAT64;
}

// ---------------------------------------------------------------------------
// LorenzoQuant
// ---------------------------------------------------------------------------
// 44-byte header (lorenzo_quant.h:43-67).  reserved[0] bit 0 is set when the
// stage carries the strict-error-bound patch list (an extension; zero
// otherwise, so headers without it are byte-identical to the reference's).
struct LorenzoQuantConfig {
    float error_bound;
    uint32_t quant_radius;
    uint32_t num_elements;
    uint32_t outlier_count;
    DataType input_type;
    DataType code_type;
    uint8_t ndim;
    uint8_t eb_mode;
    uint32_t dim_x;
    uint32_t dim_y;
    uint32_t dim_z;
    float user_eb;
    float value_base;
    uint8_t zigzag_codes;
    uint8_t reserved[3];
    LorenzoQuantConfig() { std::memset(this, 0, sizeof(*this)); input_type = DataType::FLOAT32; code_type = DataType::UINT16; ndim = 1; dim_y = 1; dim_z = 1; }
};
static_assert(sizeof(LorenzoQuantConfig) == 44, "LorenzoQuantConfig is 44 bytes");

// Instantiations (lorenzo_quant.h:335-339): <float, uint16_t>,
// <float, uint8_t>, <double, uint16_t>, <double, uint32_t>.  With double
// input every step runs in double and outlier errors are doubles; the
// strict-error-bound extension is FP32 only.
template <typename TInput = float, typename TCode = uint16_t>
class LorenzoQuantStage : public Stage {
    static_assert((std::is_same_v<TInput, float> && (std::is_same_v<TCode, uint16_t> || std::is_same_v<TCode, uint8_t>)) ||
                      (std::is_same_v<TInput, double> &&
                       (std::is_same_v<TCode, uint16_t> || std::is_same_v<TCode, uint32_t>)),
                  "LorenzoQuantStage: <float, uint8_t|uint16_t> or <double, uint16_t|uint32_t>");

  public:
    struct Config {
        float error_bound = 1e-3f;
        int quant_radius = 32768;
        float outlier_capacity = 0.2f;
        std::array<size_t, 3> dims = {0, 1, 1};
        ErrorBoundMode eb_mode = ErrorBoundMode::ABS;
        float precomputed_value_base = 0.0f;
        bool zigzag_codes = false;
        bool strict_error_bound = false;  // extension: patch list, |x - x_hat| <= eb everywhere
        float patch_capacity = 0.01f;     // patch slots = max(256, ceil(n * patch_capacity))
        Config() = default;
        Config(TInput eb, TCode radius = (TCode)32768, float outlier_cap = 0.2f, std::array<size_t, 3> d = {0, 1, 1})
            : error_bound((float)eb), quant_radius(radius), outlier_capacity(outlier_cap), dims(d) {}
    };

    explicit LorenzoQuantStage(const Config& config = Config());

    void execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                 const std::vector<void*>& outputs, const std::vector<size_t>& sizes) override;
    void postStreamSync(hipStream_t stream) override;

    std::string getName() const override { return "LorenzoQuant"; }
    size_t getNumInputs() const override { return is_inverse_ ? numPorts() : 1; }
    size_t getNumOutputs() const override { return is_inverse_ ? 1 : numPorts(); }
    std::vector<std::string> getOutputNames() const override {
        if (config_.strict_error_bound)
            return {"codes", "outlier_errors", "outlier_indices", "outlier_count",
                    "patch_values", "patch_indices", "patch_count"};
        return {"codes", "outlier_errors", "outlier_indices", "outlier_count"};
    }
    std::vector<size_t> estimateOutputSizes(const std::vector<size_t>& input_sizes) const override;
    std::unordered_map<std::string, size_t> getActualOutputSizesByName() const override;
    size_t getActualOutputSize(int index) const override {
        return (index >= 0 && index < (int)actual_output_sizes_.size()) ? actual_output_sizes_[index] : 0;
    }
    void saveState() override { saved_output_sizes_ = actual_output_sizes_; }
    void restoreState() override { actual_output_sizes_ = saved_output_sizes_; }

    void setErrorBound(TInput eb) { config_.error_bound = (float)eb; }
    void setQuantRadius(TCode r) { config_.quant_radius = r; }
    void setOutlierCapacity(float c) { config_.outlier_capacity = c; }
    void setDims(const std::array<size_t, 3>& dims) override { config_.dims = dims; }
    void setDims(size_t x, size_t y = 1, size_t z = 1) { config_.dims = {x, y, z}; }
    void setErrorBoundMode(ErrorBoundMode m) { config_.eb_mode = m; }
    void setValueBase(float vb) { config_.precomputed_value_base = vb; }
    void setZigzagCodes(bool on) { config_.zigzag_codes = on; }
    // Extension: every decompressed value within eb (the reference's FP32
    // arithmetic overshoots by a few ULPs on a few elements); codes and
    // outliers are unchanged, three more outputs carry the patched values.
    void setStrictErrorBound(bool on) { config_.strict_error_bound = on; }
    void setPatchCapacity(float c) { config_.patch_capacity = c; }
    bool getStrictErrorBound() const { return config_.strict_error_bound; }
    uint32_t getPatchCount() const { return h_pcount_; }

    TInput getErrorBound() const { return (TInput)config_.error_bound; }
    TCode getQuantRadius() const { return (TCode)config_.quant_radius; }
    float getOutlierCapacity() const { return config_.outlier_capacity; }
    std::array<size_t, 3> getDims() const { return config_.dims; }
    ErrorBoundMode getErrorBoundMode() const { return config_.eb_mode; }
    float getValueBase() const { return config_.precomputed_value_base; }
    bool getZigzagCodes() const { return config_.zigzag_codes; }
    // Absolute bound of the last compression (REL/NOA resolved) and its base.
    TInput getComputedErrorBound() const { return computed_abs_eb_; }
    float getComputedValueBase() const { return computed_value_base_; }
    uint32_t getOutlierCount() const { return actual_outlier_count_; }

    int ndim() const { return config_.dims[2] > 1 ? 3 : (config_.dims[1] > 1 ? 2 : 1); }
    bool acceptZeroPaddedInput(size_t valid_bytes) override {
        valid_elems_ = valid_bytes / sizeof(TInput);
        return !is_inverse_ && valid_bytes % sizeof(TInput) == 0;
    }
    DeviceSize getDeviceOutputSize(int index) const override {
        if (is_inverse_) return {};
        if (index == 1 && d_count_ptr_) return {d_count_ptr_, 2, (uint32_t)sizeof(TInput)};
        if (index == 2 && d_count_ptr_) return {d_count_ptr_, 2, 4};
        if (index == 4 && d_pcount_ptr_) return {d_pcount_ptr_, 2, (uint32_t)sizeof(TInput)};
        if (index == 5 && d_pcount_ptr_) return {d_pcount_ptr_, 2, 4};
        return {};
    }
    void setInverse(bool inv) override { is_inverse_ = inv; }
    bool isInverse() const override { return is_inverse_; }

    uint16_t getStageTypeId() const override { return (uint16_t)StageType::LORENZO_QUANT; }
    uint8_t getOutputDataType(size_t i) const override {
        switch (i) {
        case 0: return (uint8_t)dataTypeOf<TCode>();
        case 1: case 4: return (uint8_t)dataTypeOf<TInput>();
        case 2: case 3: case 5: case 6: return (uint8_t)DataType::UINT32;
        default: return (uint8_t)DataType::UINT8;
        }
    }
    uint8_t getInputDataType(size_t) const override { return (uint8_t)dataTypeOf<TInput>(); }
    size_t serializeHeader(size_t output_index, uint8_t* buf, size_t max_size) const override;
    void deserializeHeader(const uint8_t* buf, size_t size) override;
    size_t getMaxHeaderSize(size_t) const override { return sizeof(LorenzoQuantConfig); }

  private:
    size_t maxOutliers(size_t n) const { return (size_t)std::ceil((double)n * config_.outlier_capacity); }
    size_t maxPatches(size_t n) const {
        return std::max<size_t>(256, (size_t)std::ceil((double)n * config_.patch_capacity));
    }
    size_t numPorts() const { return config_.strict_error_bound ? 7 : 4; }
    void resolveDims(size_t n, size_t& nx, size_t& ny, size_t& nz) const;
    void executeF64(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                    const std::vector<void*>& outputs, size_t n, size_t valid, size_t nx, size_t ny, size_t nz,
                    size_t max_out);

    Config config_;
    bool is_inverse_ = false;
    std::vector<size_t> actual_output_sizes_ = std::vector<size_t>(4, 0), saved_output_sizes_;
    size_t num_elements_ = 0;
    uint32_t actual_outlier_count_ = 0;
    TInput computed_abs_eb_ = 0;  // in the input type (lorenzo_quant.h:310)
    float computed_value_base_ = 0.0f;
    size_t valid_elems_ = 0;             // > 0: input elements past this read as zero (one call)
    const void* d_count_ptr_ = nullptr;  // forward: outlier count output
    const void* d_pcount_ptr_ = nullptr; // forward (strict): patch count output
    bool params_pending_ = false;        // REL/NOA resolved on the device
    uint32_t h_count_ = 0, h_pcount_ = 0;
    TInput h_params_[4] = {0, 0, 0, 0};
    DeviceScratch ws_, inv_ws_, params_;
};

// ---------------------------------------------------------------------------
// Bitshuffle
// ---------------------------------------------------------------------------
class BitshuffleStage : public Stage {
  public:
    void setInverse(bool inv) override { is_inverse_ = inv; }
    bool isInverse() const override { return is_inverse_; }
    void setBlockSize(size_t bytes) { block_size_ = (uint32_t)bytes; }
    void setElementWidth(size_t bytes) { element_width_ = (uint8_t)bytes; }
    size_t getBlockSize() const { return block_size_; }
    size_t getElementWidth() const { return element_width_; }
    size_t getRequiredInputAlignment() const override { return block_size_; }

    void execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                 const std::vector<void*>& outputs, const std::vector<size_t>& sizes) override;
    std::string getName() const override { return "Bitshuffle"; }
    size_t getNumInputs() const override { return 1; }
    size_t getNumOutputs() const override { return 1; }
    std::vector<size_t> estimateOutputSizes(const std::vector<size_t>& in) const override { return {in[0]}; }
    std::unordered_map<std::string, size_t> getActualOutputSizesByName() const override {
        return {{"output", actual_output_size_}};
    }
    size_t getActualOutputSize(int i) const override { return i == 0 ? actual_output_size_ : 0; }
    uint16_t getStageTypeId() const override { return (uint16_t)StageType::BITSHUFFLE; }
    uint8_t getOutputDataType(size_t) const override { return (uint8_t)DataType::UINT8; }
    // [0..3] block_size u32 LE, [4] element_width  (bitshuffle_stage.h:98-113)
    size_t serializeHeader(size_t, uint8_t* buf, size_t max_size) const override {
        if (max_size < 5) return 0;
        std::memcpy(buf, &block_size_, 4);
        buf[4] = element_width_;
        return 5;
    }
    void deserializeHeader(const uint8_t* buf, size_t size) override {
        if (size >= 4) std::memcpy(&block_size_, buf, 4);
        if (size >= 5) element_width_ = buf[4];
    }
    size_t getMaxHeaderSize(size_t) const override { return 5; }
    void saveState() override { saved_ = {block_size_, element_width_, actual_output_size_}; }
    void restoreState() override {
        block_size_ = saved_.bs;
        element_width_ = saved_.ew;
        actual_output_size_ = saved_.out;
    }

  private:
    bool is_inverse_ = false;
    uint32_t block_size_ = 16384;
    uint8_t element_width_ = 4;
    size_t actual_output_size_ = 0;
    struct Saved { uint32_t bs; uint8_t ew; size_t out; } saved_{16384, 4, 0};
};

// ---------------------------------------------------------------------------
// RZE
// ---------------------------------------------------------------------------
class RZEStage : public Stage {
  public:
    void setInverse(bool inv) override { is_inverse_ = inv; }
    bool isInverse() const override { return is_inverse_; }
    void setChunkSize(size_t bytes) { chunk_size_ = (uint32_t)bytes; }
    void setLevels(int n) { levels_ = (uint8_t)n; }
    size_t getChunkSize() const { return chunk_size_; }
    int getLevels() const { return levels_; }
    uint32_t getCachedOrigBytes() const { return cached_orig_bytes_; }
    size_t getRequiredInputAlignment() const override { return chunk_size_; }
    // Set by the pipeline when a BitshuffleStage (16 KiB blocks, element
    // width 1 or 2) feeds this stage: the input is then the un-shuffled data
    // and the shuffle runs inside the RZE kernels (same stream bytes).
    void setPreShuffle(int elem_width) { pre_shuffle_ = elem_width; }
    int getPreShuffle() const { return pre_shuffle_; }
    // Set by the pipeline when this stage's output is the concat's first
    // slot: execute() then leaves the pack step to the pipeline's final
    // concat + readback launch, and deferredPack() describes it (until the
    // next execute).
    void setDeferPack(bool on) { defer_pack_ = on; }
    const void* deferredPack() const { return pack_deferred_ ? pack_desc_ : nullptr; }

    void execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                 const std::vector<void*>& outputs, const std::vector<size_t>& sizes) override;
    void postStreamSync(hipStream_t stream) override;
    std::string getName() const override { return "RZE"; }
    size_t getNumInputs() const override { return 1; }
    size_t getNumOutputs() const override { return 1; }
    std::vector<size_t> estimateOutputSizes(const std::vector<size_t>& in) const override;
    std::unordered_map<std::string, size_t> getActualOutputSizesByName() const override {
        return {{"output", actual_output_size_}};
    }
    size_t getActualOutputSize(int i) const override { return i == 0 ? actual_output_size_ : 0; }
    uint16_t getStageTypeId() const override { return (uint16_t)StageType::RZE; }
    uint8_t getOutputDataType(size_t) const override { return (uint8_t)DataType::UINT8; }
    // [0..3] chunk_size, [4] levels, [5..8] cached_orig_bytes  (rze_stage.h:169-190)
    size_t serializeHeader(size_t, uint8_t* buf, size_t max_size) const override {
        if (max_size < 9) return 0;
        std::memcpy(buf, &chunk_size_, 4);
        buf[4] = levels_;
        std::memcpy(buf + 5, &cached_orig_bytes_, 4);
        return 9;
    }
    void deserializeHeader(const uint8_t* buf, size_t size) override {
        if (size >= 4) std::memcpy(&chunk_size_, buf, 4);
        if (size >= 5) levels_ = buf[4];
        if (size >= 9) std::memcpy(&cached_orig_bytes_, buf + 5, 4);
    }
    size_t getMaxHeaderSize(size_t) const override { return 9; }
    void saveState() override { saved_ = {chunk_size_, levels_, cached_orig_bytes_}; }
    void restoreState() override {
        chunk_size_ = saved_.cs;
        levels_ = saved_.lv;
        cached_orig_bytes_ = saved_.ob;
    }

  private:
    bool is_inverse_ = false;
    int pre_shuffle_ = 0;
    bool defer_pack_ = false, pack_deferred_ = false;
    alignas(8) unsigned char pack_desc_[96] = {};
    uint32_t chunk_size_ = 16384;
    uint8_t levels_ = 4;
    uint32_t cached_orig_bytes_ = 0;
    size_t actual_output_size_ = 0;
    bool size_pending_ = false, status_pending_ = false;
    uint64_t h_size_ = 0;
    int32_t h_status_ = 0;
    struct Saved { uint32_t cs; uint8_t lv; uint32_t ob; } saved_{16384, 4, 0};
    DeviceScratch ws_, inv_ws_, small_;  // small_: [0] u64 out size, [8] i32 status
  public:
    DeviceSize getDeviceOutputSize(int index) const override {
        if (is_inverse_ || index != 0 || !size_pending_) return {};
        return {small_.ptr, 1, 1};
    }
};

// ---------------------------------------------------------------------------
// RLE
// ---------------------------------------------------------------------------
template <typename T = uint16_t>
class RLEStage : public Stage {
    static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4, "RLEStage: 1/2/4-byte elements");

  public:
    void setInverse(bool inv) override { is_inverse_ = inv; }
    bool isInverse() const override { return is_inverse_; }
    void execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                 const std::vector<void*>& outputs, const std::vector<size_t>& sizes) override;
    void postStreamSync(hipStream_t stream) override;
    std::string getName() const override { return "RLE"; }
    size_t getNumInputs() const override { return 1; }
    size_t getNumOutputs() const override { return 1; }
    std::vector<size_t> estimateOutputSizes(const std::vector<size_t>& in) const override {
        if (is_inverse_) return {cached_num_elements_ ? (size_t)cached_num_elements_ * sizeof(T) : in[0] * 2};
        const size_t n = in[0] / sizeof(T);
        return {4 + ((n * sizeof(T) + 3) & ~size_t(3)) + 4 * n};
    }
    std::unordered_map<std::string, size_t> getActualOutputSizesByName() const override {
        return {{"output", actual_output_size_}};
    }
    size_t getActualOutputSize(int i) const override { return i == 0 ? actual_output_size_ : 0; }
    uint16_t getStageTypeId() const override { return (uint16_t)StageType::RLE; }
    uint8_t getOutputDataType(size_t) const override { return (uint8_t)dataTypeOf<T>(); }
    uint8_t getInputDataType(size_t) const override { return (uint8_t)dataTypeOf<T>(); }
    // [0] DataType, [1..4] cached_num_elements  (rle.h:123-141)
    size_t serializeHeader(size_t, uint8_t* buf, size_t max_size) const override {
        if (max_size < 5) return 0;
        buf[0] = (uint8_t)dataTypeOf<T>();
        std::memcpy(buf + 1, &cached_num_elements_, 4);
        return 5;
    }
    void deserializeHeader(const uint8_t* buf, size_t size) override {
        if (size >= 5) std::memcpy(&cached_num_elements_, buf + 1, 4);
    }
    size_t getMaxHeaderSize(size_t) const override { return 5; }
    void saveState() override { saved_n_ = cached_num_elements_; saved_out_ = actual_output_size_; }
    void restoreState() override { cached_num_elements_ = saved_n_; actual_output_size_ = saved_out_; }

  private:
    bool is_inverse_ = false;
    uint32_t cached_num_elements_ = 0, saved_n_ = 0;
    size_t actual_output_size_ = 0, saved_out_ = 0;
    bool size_pending_ = false, status_pending_ = false;
    uint64_t h_size_ = 0;
    int32_t h_status_ = 0;
    DeviceScratch ws_, inv_ws_, small_;
  public:
    DeviceSize getDeviceOutputSize(int index) const override {
        if (is_inverse_ || index != 0 || !size_pending_) return {};
        return {small_.ptr, 1, 1};
    }
};

// ---------------------------------------------------------------------------
// Huffman (coarse mode)
// ---------------------------------------------------------------------------
enum class HuffmanEncodeMode { Coarse, Fine };

template <typename T = uint16_t>
class HuffmanStage : public Stage {
    static_assert(std::is_same_v<T, uint8_t> || std::is_same_v<T, uint16_t> || std::is_same_v<T, uint32_t>,
                  "HuffmanStage: T must be uint8_t, uint16_t, or uint32_t.");

  public:
    void setBklen(uint32_t b) { bklen_ = b; }
    uint32_t getBklen() const { return bklen_; }
    // Coarse: 768-symbol partitions.  Fine: 1024-symbol partitions (the
    // ReVISIT-lite layout) when no codeword exceeds 8 bits, else coarse
    // (hf_hl.cc:69-80).  The encode is device-side either way.
    void setEncodeMode(HuffmanEncodeMode m) { encode_mode_ = m; }
    HuffmanEncodeMode getEncodeMode() const { return encode_mode_; }
    void setInverse(bool inv) override { is_inverse_ = inv; }
    bool isInverse() const override { return is_inverse_; }
    // Extension: build the codebook on the device (coarse encoding, bklen <=
    // 1024), so the forward path enqueues device work only and can be
    // captured in a graph; the stream is byte for byte the host-book one.
    // Out-of-range symbols are then reported after the stream synchronises.
    // The device build is a serial heap replay: slower than the host
    // builder (ms for a full 1024-symbol alphabet), chosen for capture.
    void setDeviceCodebook(bool on) { device_book_ = on; }
    bool getDeviceCodebook() const { return device_book_; }
    // With the host codebook the histogram has to reach the host builder.
    bool isGraphCompatible() const override { return is_inverse_ || device_book_; }

    void execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                 const std::vector<void*>& outputs, const std::vector<size_t>& sizes) override;
    void postStreamSync(hipStream_t stream) override;
    std::string getName() const override { return "Huffman"; }
    size_t getNumInputs() const override { return 1; }
    size_t getNumOutputs() const override { return 1; }
    std::vector<size_t> estimateOutputSizes(const std::vector<size_t>& in) const override;
    std::unordered_map<std::string, size_t> getActualOutputSizesByName() const override {
        return {{"output", actual_output_size_}};
    }
    size_t getActualOutputSize(int i) const override { return i == 0 ? actual_output_size_ : 0; }
    uint16_t getStageTypeId() const override { return (uint16_t)StageType::HUFFMAN; }
    uint8_t getOutputDataType(size_t) const override { return (uint8_t)DataType::UNKNOWN; }
    uint8_t getInputDataType(size_t) const override { return (uint8_t)DataType::UNKNOWN; }
    // [0] DataType, [1..2] bklen u16, [3..10] original_len u64  (huffman_stage.h:174-193)
    size_t serializeHeader(size_t, uint8_t* buf, size_t max_size) const override {
        if (max_size < 11) return 0;
        buf[0] = (uint8_t)dataTypeOf<T>();
        const uint16_t bk = (uint16_t)bklen_;
        std::memcpy(buf + 1, &bk, 2);
        std::memcpy(buf + 3, &original_len_, 8);
        return 11;
    }
    void deserializeHeader(const uint8_t* buf, size_t size) override {
        if (size >= 3) {
            uint16_t bk;
            std::memcpy(&bk, buf + 1, 2);
            bklen_ = bk;
        }
        if (size >= 11) std::memcpy(&original_len_, buf + 3, 8);
    }
    size_t getMaxHeaderSize(size_t) const override { return 11; }
    void saveState() override { saved_ = {bklen_, original_len_, actual_output_size_}; }
    void restoreState() override {
        bklen_ = saved_.bk;
        original_len_ = saved_.ol;
        actual_output_size_ = saved_.out;
    }

  private:
    bool is_inverse_ = false;
    uint32_t bklen_ = std::is_same_v<T, uint8_t> ? 256 : 1024;
    HuffmanEncodeMode encode_mode_ = HuffmanEncodeMode::Coarse;
    uint64_t original_len_ = 0;
    size_t actual_output_size_ = 0;
    bool device_book_ = false;
    bool size_pending_ = false, status_pending_ = false, book_pending_ = false;
    uint64_t h_size_ = 0;
    int32_t h_status_ = 0, h_book_status_ = 0;
    struct Saved { uint32_t bk; uint64_t ol; size_t out; } saved_{1024, 0, 0};
    DeviceScratch ws_, small_, book_, freq_;
  public:
    DeviceSize getDeviceOutputSize(int index) const override {
        if (is_inverse_ || index != 0 || !size_pending_) return {};
        return {small_.ptr, 1, 1};
    }
};

// ---------------------------------------------------------------------------
// Quantizer (ABS / NOA)
// ---------------------------------------------------------------------------
struct QuantizerConfig {
    float abs_error_bound;
    float user_error_bound;
    float value_base;
    uint32_t quant_radius;
    uint32_t num_elements;
    uint32_t outlier_count;
    DataType input_type;
    DataType code_type;
    uint8_t eb_mode;
    uint8_t zigzag_codes;
    float outlier_threshold;
    uint8_t inplace_outliers;
    uint8_t pad[3];
    QuantizerConfig() {
        std::memset(this, 0, sizeof(*this));
        input_type = DataType::FLOAT32;
        code_type = DataType::UINT16;
        outlier_threshold = INFINITY;
    }
};
static_assert(sizeof(QuantizerConfig) == 36, "QuantizerConfig is 36 bytes");

template <typename TInput = float, typename TCode = uint16_t>
class QuantizerStage : public Stage {
    static_assert(std::is_same_v<TInput, float>, "QuantizerStage: float input on gfx950");

  public:
    struct Config {
        float error_bound = 1e-4f;
        int quant_radius = 32768;
        float outlier_capacity = 0.05f;
        ErrorBoundMode eb_mode = ErrorBoundMode::ABS;
        float precomputed_value_base = 0.0f;
        bool zigzag_codes = false;
        float outlier_threshold = INFINITY;
        // ABS/NOA: outliers as raw float bits in the codes (u32 codes, zigzag on)
        bool inplace_outliers = false;
        Config() = default;
        Config(float eb, int radius = 32768, float cap = 0.05f) : error_bound(eb), quant_radius(radius), outlier_capacity(cap) {}
    };
    explicit QuantizerStage(const Config& c = Config()) : config_(c) {}

    void execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                 const std::vector<void*>& outputs, const std::vector<size_t>& sizes) override;
    void postStreamSync(hipStream_t stream) override;
    std::string getName() const override { return "Quantizer"; }
    // quantizer.h:134-146: one port each way in in-place mode
    size_t getNumInputs() const override { return !is_inverse_ ? 1 : (isInplaceMode() ? 1 : 4); }
    size_t getNumOutputs() const override { return is_inverse_ ? 1 : (isInplaceMode() ? 1 : 4); }
    std::vector<std::string> getOutputNames() const override {
        if (is_inverse_) return {"reconstructed"};
        if (isInplaceMode()) return {"codes"};
        return {"codes", "outlier_vals", "outlier_idxs", "outlier_count"};
    }
    std::vector<size_t> estimateOutputSizes(const std::vector<size_t>& in) const override;
    std::unordered_map<std::string, size_t> getActualOutputSizesByName() const override;
    size_t getActualOutputSize(int i) const override {
        return (i >= 0 && i < (int)actual_output_sizes_.size()) ? actual_output_sizes_[i] : 0;
    }
    void setInverse(bool inv) override { is_inverse_ = inv; }
    bool isInverse() const override { return is_inverse_; }
    void saveState() override { saved_sizes_ = actual_output_sizes_; }
    void restoreState() override { actual_output_sizes_ = saved_sizes_; }
    void setErrorBound(TInput eb) { config_.error_bound = (float)eb; }
    void setQuantRadius(int r) { config_.quant_radius = r; }
    void setOutlierCapacity(float c) { config_.outlier_capacity = c; }
    void setErrorBoundMode(ErrorBoundMode m) { config_.eb_mode = m; }
    void setValueBase(float vb) { config_.precomputed_value_base = vb; }
    void setZigzagCodes(bool on) { config_.zigzag_codes = on; }
    void setOutlierThreshold(float t) { config_.outlier_threshold = t; }
    void setInplaceOutliers(bool on) { config_.inplace_outliers = on; }
    TInput getErrorBound() const { return (TInput)config_.error_bound; }
    int getQuantRadius() const { return config_.quant_radius; }
    ErrorBoundMode getErrorBoundMode() const { return config_.eb_mode; }
    float getValueBase() const { return config_.precomputed_value_base; }
    float getOutlierCapacity() const { return config_.outlier_capacity; }
    bool getZigzagCodes() const { return config_.zigzag_codes; }
    float getOutlierThreshold() const { return config_.outlier_threshold; }
    bool getInplaceOutliers() const { return config_.inplace_outliers; }
    uint16_t getStageTypeId() const override { return (uint16_t)StageType::QUANTIZER; }
    uint8_t getOutputDataType(size_t i) const override {
        if (is_inverse_) return (uint8_t)DataType::FLOAT32;
        if (isInplaceMode()) return (uint8_t)dataTypeOf<TCode>();
        switch (i) {
        case 0: return (uint8_t)dataTypeOf<TCode>();
        case 1: return (uint8_t)DataType::FLOAT32;
        default: return (uint8_t)DataType::UINT32;
        }
    }
    uint8_t getInputDataType(size_t) const override { return (uint8_t)DataType::FLOAT32; }
    size_t serializeHeader(size_t, uint8_t* buf, size_t max_size) const override;
    void deserializeHeader(const uint8_t* buf, size_t size) override;
    size_t getMaxHeaderSize(size_t) const override { return sizeof(QuantizerConfig); }

  private:
    bool isInplaceMode() const { return config_.inplace_outliers && config_.eb_mode != ErrorBoundMode::REL; }
    size_t maxOutliers(size_t n) const { return (size_t)std::ceil((double)n * config_.outlier_capacity); }
    Config config_;
    bool is_inverse_ = false;
    std::vector<size_t> actual_output_sizes_ = std::vector<size_t>(4, 0), saved_sizes_;
    size_t num_elements_ = 0;
    uint32_t outlier_count_ = 0;
    float computed_abs_eb_ = 0.0f, computed_value_base_ = 0.0f;
    const void* d_count_ptr_ = nullptr;
    uint32_t h_count_ = 0;
    DeviceScratch ws_, vb_;
  public:
    DeviceSize getDeviceOutputSize(int index) const override {
        if (is_inverse_ || (index != 1 && index != 2) || !d_count_ptr_) return {};
        return {d_count_ptr_, 2, 4};
    }
};

// ---------------------------------------------------------------------------
// Bitpack
// ---------------------------------------------------------------------------
template <typename T = uint16_t>
class BitpackStage : public Stage {
  public:
    void setInverse(bool inv) override { is_inverse_ = inv; }
    bool isInverse() const override { return is_inverse_; }
    void setNBits(uint8_t nbits) {
        if (nbits == 0 || nbits > 8 * sizeof(T) || (nbits & (nbits - 1)) != 0)
            throw std::invalid_argument("BitpackStage::setNBits: nbits must be a power of two in [1, " +
                                        std::to_string(8 * sizeof(T)) + "], got " + std::to_string(nbits));
        nbits_ = nbits;
    }
    uint8_t getNBits() const { return nbits_; }
    void setAutoDetect(bool on) { auto_detect_ = on; }
    bool isAutoDetect() const { return auto_detect_; }
    bool isGraphCompatible() const override { return !auto_detect_ || is_inverse_; }

    void execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                 const std::vector<void*>& outputs, const std::vector<size_t>& sizes) override;
    std::string getName() const override { return "Bitpack"; }
    size_t getNumInputs() const override { return 1; }
    size_t getNumOutputs() const override { return 1; }
    std::vector<size_t> estimateOutputSizes(const std::vector<size_t>& in) const override {
        if (is_inverse_) return {num_elements_ * sizeof(T)};
        return {in[0]};
    }
    std::unordered_map<std::string, size_t> getActualOutputSizesByName() const override {
        return {{"output", actual_output_size_}};
    }
    size_t getActualOutputSize(int i) const override { return i == 0 ? actual_output_size_ : 0; }
    uint16_t getStageTypeId() const override { return (uint16_t)StageType::BITPACK; }
    uint8_t getOutputDataType(size_t) const override { return (uint8_t)DataType::UINT8; }
    uint8_t getInputDataType(size_t) const override { return (uint8_t)dataTypeOf<T>(); }
    // [0] DataType, [1] nbits, [2..9] num_elements u64
    size_t serializeHeader(size_t, uint8_t* buf, size_t max_size) const override {
        if (max_size < 10) return 0;
        buf[0] = (uint8_t)dataTypeOf<T>();
        buf[1] = nbits_;
        const uint64_t n = num_elements_;
        std::memcpy(buf + 2, &n, 8);
        return 10;
    }
    void deserializeHeader(const uint8_t* buf, size_t size) override {
        if (size >= 2) nbits_ = buf[1];
        if (size >= 10) {
            uint64_t n;
            std::memcpy(&n, buf + 2, 8);
            num_elements_ = n;
        }
    }
    size_t getMaxHeaderSize(size_t) const override { return 10; }
    void saveState() override { saved_ = {nbits_, num_elements_, actual_output_size_}; }
    void restoreState() override {
        nbits_ = saved_.nb;
        num_elements_ = saved_.n;
        actual_output_size_ = saved_.out;
    }

  private:
    bool is_inverse_ = false, auto_detect_ = false;
    uint8_t nbits_ = 8 * sizeof(T);
    size_t num_elements_ = 0, actual_output_size_ = 0;
    struct Saved { uint8_t nb; size_t n; size_t out; } saved_{8 * sizeof(T), 0, 0};
    DeviceScratch ws_;
};

// ---------------------------------------------------------------------------
// Difference (modules/predictors/diff/diff.h:16-200): first-order
// differences, optionally reset at every chunk_size-byte boundary, optionally
// negabinary-coded (T signed, TOut its unsigned counterpart).  Inverse: the
// cumulative sum; floating-point sums in the reference's association.
// ---------------------------------------------------------------------------
template <typename T = float, typename TOut = T>
class DifferenceStage : public Stage {
    static_assert(std::is_same_v<T, TOut> ||
                      (std::is_integral_v<T> && std::is_signed_v<T> && std::is_integral_v<TOut> &&
                       std::is_unsigned_v<TOut> && sizeof(T) == sizeof(TOut)),
                  "DifferenceStage: TOut must equal T, or be the unsigned counterpart of a signed T");

  public:
    void setInverse(bool inv) override { is_inverse_ = inv; }
    bool isInverse() const override { return is_inverse_; }
    void setChunkSize(size_t bytes) { chunk_size_ = bytes; }
    size_t getChunkSize() const { return chunk_size_; }
    size_t getRequiredInputAlignment() const override { return chunk_size_ > 0 ? chunk_size_ : 1; }
    void execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                 const std::vector<void*>& outputs, const std::vector<size_t>& sizes) override;
    std::string getName() const override { return "Difference"; }
    size_t getNumInputs() const override { return 1; }
    size_t getNumOutputs() const override { return 1; }
    std::vector<size_t> estimateOutputSizes(const std::vector<size_t>& in) const override { return {in[0]}; }
    std::unordered_map<std::string, size_t> getActualOutputSizesByName() const override {
        return {{"output", actual_output_size_}};
    }
    size_t getActualOutputSize(int i) const override { return i == 0 ? actual_output_size_ : 0; }
    uint16_t getStageTypeId() const override { return (uint16_t)StageType::DIFFERENCE; }
    uint8_t getOutputDataType(size_t) const override { return (uint8_t)dataTypeOf<TOut>(); }
    uint8_t getInputDataType(size_t) const override { return (uint8_t)dataTypeOf<T>(); }
    // [0] DataType T, [1] DataType TOut, [2..5] chunk_size u32 (0 = none)
    size_t serializeHeader(size_t, uint8_t* buf, size_t max_size) const override {
        if (max_size < 6) return 0;
        buf[0] = (uint8_t)dataTypeOf<T>();
        buf[1] = (uint8_t)dataTypeOf<TOut>();
        const uint32_t cs = (uint32_t)chunk_size_;
        std::memcpy(buf + 2, &cs, 4);
        return 6;
    }
    void deserializeHeader(const uint8_t* buf, size_t size) override {
        if (size >= 6) {
            uint32_t cs;
            std::memcpy(&cs, buf + 2, 4);
            chunk_size_ = cs;
        }
    }
    size_t getMaxHeaderSize(size_t) const override { return 6; }
    void saveState() override { saved_ = {chunk_size_, actual_output_size_}; }
    void restoreState() override {
        chunk_size_ = saved_.cs;
        actual_output_size_ = saved_.out;
    }

  private:
    bool is_inverse_ = false;
    size_t chunk_size_ = 0, actual_output_size_ = 0;
    struct Saved { size_t cs, out; } saved_{0, 0};
    DeviceScratch ws_;
};

// ---------------------------------------------------------------------------
// Zigzag (modules/transforms/zigzag/zigzag_stage.h) and Negabinary
// (modules/transforms/negabinary/negabinary_stage.h): signed TIn <-> its
// unsigned TOut, element-wise; 2-byte header [TIn DataType, TOut DataType].
// ---------------------------------------------------------------------------
template <bool NEGABINARY, typename TIn, typename TOut>
class IntCodeStage : public Stage {
    static_assert(std::is_integral_v<TIn> && std::is_signed_v<TIn> && std::is_unsigned_v<TOut> &&
                      sizeof(TIn) == sizeof(TOut),
                  "signed TIn and its unsigned counterpart TOut");

  public:
    void setInverse(bool inv) override { is_inverse_ = inv; }
    bool isInverse() const override { return is_inverse_; }
    void execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                 const std::vector<void*>& outputs, const std::vector<size_t>& sizes) override;
    std::string getName() const override { return NEGABINARY ? "Negabinary" : "Zigzag"; }
    size_t getNumInputs() const override { return 1; }
    size_t getNumOutputs() const override { return 1; }
    std::vector<size_t> estimateOutputSizes(const std::vector<size_t>& in) const override { return {in[0]}; }
    std::unordered_map<std::string, size_t> getActualOutputSizesByName() const override {
        return {{"output", actual_output_size_}};
    }
    size_t getActualOutputSize(int i) const override { return i == 0 ? actual_output_size_ : 0; }
    uint16_t getStageTypeId() const override {
        return (uint16_t)(NEGABINARY ? StageType::NEGABINARY : StageType::ZIGZAG);
    }
    uint8_t getOutputDataType(size_t) const override {
        return (uint8_t)(is_inverse_ ? dataTypeOf<TIn>() : dataTypeOf<TOut>());
    }
    uint8_t getInputDataType(size_t) const override {
        return (uint8_t)(is_inverse_ ? dataTypeOf<TOut>() : dataTypeOf<TIn>());
    }
    size_t serializeHeader(size_t, uint8_t* buf, size_t max_size) const override {
        if (max_size < 2) return 0;
        buf[0] = (uint8_t)dataTypeOf<TIn>();
        buf[1] = (uint8_t)dataTypeOf<TOut>();
        return 2;
    }
    size_t getMaxHeaderSize(size_t) const override { return 2; }
    void saveState() override { saved_out_ = actual_output_size_; }
    void restoreState() override { actual_output_size_ = saved_out_; }

  private:
    bool is_inverse_ = false;
    size_t actual_output_size_ = 0, saved_out_ = 0;
};
template <typename TIn = int32_t, typename TOut = uint32_t> using ZigzagStage = IntCodeStage<false, TIn, TOut>;
template <typename TIn = int32_t, typename TOut = uint32_t> using NegabinaryStage = IntCodeStage<true, TIn, TOut>;

// ---------------------------------------------------------------------------
// Integer Lorenzo (modules/predictors/lorenzo/lorenzo_stage.h): 1D deltas
// inside 256-element blocks; 2D/3D inclusion-exclusion deltas over the whole
// field; inverse prefix sums along x, y, z.  16-byte header (LorenzoConfig).
// ---------------------------------------------------------------------------
struct LorenzoConfig {
    DataType data_type;
    uint8_t ndim;
    uint8_t reserved[2];
    uint32_t dim_x, dim_y, dim_z;
    LorenzoConfig() : data_type(DataType::INT32), ndim(1), reserved{0, 0}, dim_x(0), dim_y(1), dim_z(1) {}
};
static_assert(sizeof(LorenzoConfig) == 16, "LorenzoConfig is 16 bytes");

template <typename T = int32_t>
class LorenzoStage : public Stage {
    static_assert(std::is_integral_v<T> && std::is_signed_v<T>, "LorenzoStage requires a signed integer type");

  public:
    void setInverse(bool inv) override { is_inverse_ = inv; }
    bool isInverse() const override { return is_inverse_; }
    void setDims(const std::array<size_t, 3>& dims) override { dims_ = dims; }
    void setDims(size_t x, size_t y = 1, size_t z = 1) { dims_ = {x, y, z}; }
    std::array<size_t, 3> getDims() const { return dims_; }
    int ndim() const { return dims_[2] > 1 ? 3 : (dims_[1] > 1 ? 2 : 1); }
    void execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                 const std::vector<void*>& outputs, const std::vector<size_t>& sizes) override;
    std::string getName() const override { return "Lorenzo"; }
    size_t getNumInputs() const override { return 1; }
    size_t getNumOutputs() const override { return 1; }
    std::vector<size_t> estimateOutputSizes(const std::vector<size_t>& in) const override {
        return {in.empty() ? 0 : in[0]};
    }
    std::unordered_map<std::string, size_t> getActualOutputSizesByName() const override {
        return {{"output", actual_output_size_}};
    }
    size_t getActualOutputSize(int i) const override { return i == 0 ? actual_output_size_ : 0; }
    uint16_t getStageTypeId() const override { return (uint16_t)StageType::LORENZO; }
    uint8_t getOutputDataType(size_t) const override { return (uint8_t)dataTypeOf<T>(); }
    uint8_t getInputDataType(size_t) const override { return (uint8_t)dataTypeOf<T>(); }
    size_t serializeHeader(size_t, uint8_t* buf, size_t max_size) const override {
        if (max_size < sizeof(LorenzoConfig)) return 0;
        LorenzoConfig c;
        c.data_type = dataTypeOf<T>();
        c.ndim = (uint8_t)ndim();
        c.dim_x = (uint32_t)dims_[0];
        c.dim_y = (uint32_t)dims_[1];
        c.dim_z = (uint32_t)dims_[2];
        std::memcpy(buf, &c, sizeof(c));
        return sizeof(c);
    }
    void deserializeHeader(const uint8_t* buf, size_t size) override {
        if (size < sizeof(LorenzoConfig)) throw std::runtime_error("LorenzoStage: header too small");
        LorenzoConfig c;
        std::memcpy(&c, buf, sizeof(c));
        const int nd = c.ndim == 0 ? 1 : c.ndim;
        dims_ = {c.dim_x, nd >= 2 ? c.dim_y : 1u, nd >= 3 ? c.dim_z : 1u};
    }
    size_t getMaxHeaderSize(size_t) const override { return sizeof(LorenzoConfig); }
    void saveState() override { saved_out_ = actual_output_size_; }
    void restoreState() override { actual_output_size_ = saved_out_; }

  private:
    bool is_inverse_ = false;
    std::array<size_t, 3> dims_ = {0, 1, 1};
    size_t actual_output_size_ = 0, saved_out_ = 0;
    DeviceScratch ws_;
};

// Rebuilds a stage from its FZM record (include/stage/stage_factory.h:36-248).
Stage* createStage(StageType type, const uint8_t* config, size_t config_size);

}  // namespace fz

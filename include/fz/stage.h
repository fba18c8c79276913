// fz/stage.h — the stage plugin interface.
//
// Same virtual interface as the reference's fz::Stage
// (include/stage/stage.h:30-242), with hipStream_t in place of the CUDA stream:
// the pipeline talks to stages only through these methods, so a stage
// written against the reference interface drops in.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "fz/fzm_format.h"

namespace fz {

class MemoryPool;

// A small device value (size, count, status word) that a stage reads on the
// host in postStreamSync().
struct Readback {
    const void* d_src;
    void* h_dst;
    uint32_t bytes;  // multiple of 4
};

// Where an output's final byte count lives on the device, for a concat that
// runs before the host has seen it.  kind 0: the count is host-known
// (getActualOutputSize); 1: u64 byte count at ptr; 2: u32 element count at
// ptr, times elem bytes.
struct DeviceSize {
    const void* ptr = nullptr;
    uint32_t kind = 0;
    uint32_t elem = 1;
};

class Stage {
  public:
    virtual ~Stage() = default;

    // Enqueue the stage on `stream`.  inputs/outputs are device pointers,
    // sizes are the input byte counts.  Stages that must read device data on
    // the host return false from isGraphCompatible().
    virtual void execute(hipStream_t stream, MemoryPool* pool, const std::vector<void*>& inputs,
                         const std::vector<void*>& outputs, const std::vector<size_t>& sizes) = 0;

    virtual std::string getName() const = 0;
    virtual size_t getNumInputs() const = 0;
    virtual size_t getNumOutputs() const = 0;

    // LCM of these becomes the pipeline's input zero-padding granularity.
    virtual size_t getRequiredInputAlignment() const { return 1; }

    virtual std::vector<std::string> getOutputNames() const { return {"output"}; }

    int getOutputIndex(const std::string& name) const {
        const auto names = getOutputNames();
        for (size_t i = 0; i < names.size(); i++)
            if (names[i] == name) return (int)i;
        return -1;
    }

    // Safe upper bound of each output's size for the given input sizes.
    virtual std::vector<size_t> estimateOutputSizes(const std::vector<size_t>& input_sizes) const = 0;

    // Sizes after execute() (and after postStreamSync()), by output name.
    virtual std::unordered_map<std::string, size_t> getActualOutputSizesByName() const = 0;

    virtual size_t getActualOutputSize(int index) const {
        const auto names = getOutputNames();
        if (index < 0 || index >= (int)names.size()) return 0;
        const auto m = getActualOutputSizesByName();
        const auto it = m.find(names[index]);
        return it != m.end() ? it->second : 0;
    }

    virtual void setInverse(bool inverse) { (void)inverse; }
    virtual bool isInverse() const { return false; }

    virtual uint16_t getStageTypeId() const = 0;
    virtual uint8_t getOutputDataType(size_t output_index) const = 0;
    virtual uint8_t getInputDataType(size_t /*input_index*/) const { return (uint8_t)DataType::UNKNOWN; }

    // Stage config for the FZM header (at most 128 bytes).
    virtual size_t serializeHeader(size_t output_index, uint8_t* buf, size_t max_size) const {
        (void)output_index; (void)buf; (void)max_size;
        return 0;
    }
    virtual void deserializeHeader(const uint8_t* buf, size_t size) { (void)buf; (void)size; }

    virtual void saveState() {}
    virtual void restoreState() {}

    virtual void setDims(const std::array<size_t, 3>& dims) { (void)dims; }

    // Called once by Pipeline::finalize() with the estimated input size.
    virtual void onFinalize(size_t /*estimated_inlen*/, MemoryPool* /*pool*/) {}

    virtual size_t estimateDeviceFootprintBytes(size_t /*inlen*/) const { return 0; }
    virtual size_t estimatePinnedFootprintBytes(size_t /*inlen*/) const { return 0; }

    // After the pipeline synchronised the stream: small device->host reads.
    virtual void postStreamSync(hipStream_t stream) { (void)stream; }

    virtual size_t getMaxHeaderSize(size_t /*output_index*/) const { return 0; }

    // True if execute() only enqueues device work (no host reads of device data).
    virtual bool isGraphCompatible() const { return true; }

    virtual size_t estimateScratchBytes(const std::vector<size_t>& /*input_sizes*/) const { return 0; }

    // ---- Extensions of this build (no counterpart in the reference) ----

    // The pipeline input is zero-padded to a multiple of the stages'
    // alignment.  A source stage that returns true here reads only the first
    // valid_bytes of its input and treats the rest as zeros, so the pipeline
    // passes the caller's buffer instead of a padded copy.
    virtual bool acceptZeroPaddedInput(size_t valid_bytes) { (void)valid_bytes; return false; }

    virtual DeviceSize getDeviceOutputSize(int index) const { (void)index; return {}; }

    // Device values postStreamSync() needs, registered by execute().  A
    // pipeline copies all stages' values in one transfer and then calls
    // markReadbacksDone(); a stage driven on its own copies them itself.
    const std::vector<Readback>& pendingReadbacks() const { return readbacks_; }
    void markReadbacksDone() { readbacks_done_ = true; }

  protected:
    void clearReadbacks() { readbacks_.clear(); readbacks_done_ = false; }
    void addReadback(const void* d_src, void* h_dst, uint32_t bytes) { readbacks_.push_back({d_src, h_dst, bytes}); }
    // Copies the values unless a pipeline already did; re-arms for the next call.
    void completeReadbacks() {
        if (!readbacks_done_)
            for (const auto& r : readbacks_) {
                const hipError_t e = hipMemcpy(r.h_dst, r.d_src, r.bytes, hipMemcpyDeviceToHost);
                if (e != hipSuccess) throw std::runtime_error(std::string("readback: ") + hipGetErrorString(e));
            }
        readbacks_done_ = false;
    }

  private:
    std::vector<Readback> readbacks_;
    bool readbacks_done_ = false;
};

// Turns an fzh_* status code into the reference's error behaviour
// (std::runtime_error carrying the library's message).
void throwIfFzh(int rc, const char* what);

}  // namespace fz

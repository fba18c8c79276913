// fz/batch.h — batched compression of many fields with one persistent
// launch per stage (no counterpart in the reference, which compresses one
// field per Pipeline::compress, include/pipeline/compressor.h:175-181).
//
// The chain is LorenzoQuant -> Bitshuffle(16384, 2) -> RZE(16384, 4) with
// uint16 codes.  Field i's result is byte-identical to fz::Pipeline's
// compress() of that field alone with the same stage settings (and stage
// fusion, the default), so fz::Pipeline::decompress reads it back.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "fz/stages.h"

namespace fz {

struct BatchField {
    const float* d_in = nullptr;  // nx*ny*nz float32 values, 16-byte aligned
    size_t nx = 0, ny = 1, nz = 1;
};

struct BatchConfig {
    float error_bound = 1e-3f;
    ErrorBoundMode eb_mode = ErrorBoundMode::REL;
    uint32_t quant_radius = 32768;
    float outlier_capacity = 0.1f;
    bool zigzag = true;
    bool strict = false;
    float patch_capacity = 0.01f;
};

// Persistent-grid geometry (workgroups per CU per kernel; Lorenzo tiles per
// wave step; the RZE encoder takes one workgroup per chunk).  Changes timing
// only.  The Lorenzo default, a grid covering every tile with one tile per
// wave (lz_tiles_per_unit 1 and a grid at least the tile count), runs the
// loop-free kernels.
struct BatchLaunch {
    uint32_t vb_wg_per_cu = 4;
    uint32_t lz_wg_per_cu = 1u << 20;
    uint32_t lz_tiles_per_unit = 1;
    uint32_t pack_wg_per_cu = 4;
    uint32_t scatter_wg_per_cu = 2;
};

class BatchCompressor {
  public:
    explicit BatchCompressor(const BatchConfig& cfg);
    ~BatchCompressor();
    BatchCompressor(const BatchCompressor&) = delete;
    BatchCompressor& operator=(const BatchCompressor&) = delete;

    // Plans the batch: device tables, codes/scratch/output buffers.
    void setFields(const std::vector<BatchField>& fields);
    size_t numFields() const { return fields_.size(); }

    // Compresses every field on `stream` and waits for it; throws
    // std::runtime_error on a strict-mode overflow (as LorenzoQuantStage).
    void compress(hipStream_t stream);

    const uint8_t* output() const { return d_out_; }
    uint64_t outputCapacity() const { return out_cap_; }
    uint64_t fieldOffset(size_t i) const;
    uint64_t fieldBytes(size_t i) const;
    uint64_t totalInputBytes() const { return in_bytes_; }
    uint64_t totalOutputBytes() const;

    void setLaunch(const BatchLaunch& l);
    BatchLaunch launch() const { return launch_; }
    // Times each persistent kernel over its candidate geometries on the
    // planned batch and keeps the fastest (bytes are the same for all).
    BatchLaunch autotune(hipStream_t stream, int reps = 3);

    void enableGraphMode(bool on);
    void enableProfiling(bool on) { profiling_ = on; }
    // Device ms per stage of the last profiled (eager) compress.
    const std::vector<std::pair<std::string, float>>& stageTimes() const { return stage_ms_; }

  private:
    struct Impl;
    void release();
    void enqueue(hipStream_t s);
    void destroyGraph();

    BatchConfig cfg_;
    BatchLaunch launch_;
    std::vector<BatchField> fields_;
    std::vector<uint64_t> out_off_;
    Impl* impl_ = nullptr;
    uint8_t* d_out_ = nullptr;
    uint64_t out_cap_ = 0, in_bytes_ = 0;
    int cus_ = 0;
    bool graph_mode_ = true, profiling_ = false;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    hipStream_t graph_stream_ = nullptr;
    int calls_since_plan_ = 0;
    std::vector<std::pair<std::string, float>> stage_ms_;
};

}  // namespace fz

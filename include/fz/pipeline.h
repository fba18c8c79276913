// fz/pipeline.h — the stage pipeline: build, finalize, compress, decompress,
// FZM file I/O.
//
// Mirrors the reference's fz::Pipeline (include/pipeline/compressor.h:34-758)
// for the stage chains of this build: stages form a tree rooted at one source
// stage fed by the pipeline input; each stage output is either consumed by
// exactly one downstream stage or becomes a pipeline output.  Execution is on
// the caller's stream.  Every size the host needs comes back in ONE batched
// transfer after ONE stream synchronisation, and multi-output results are
// concatenated on the device from device-resident sizes, so a compress of a
// Lorenzo -> Bitshuffle -> RZE chain is capture-safe end to end; with graph
// mode on it replays as a single hipGraph launch.
#pragma once

#include <array>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "fz/mempool.h"
#include "fz/stage.h"

namespace fz {

struct StageTimingResult {
    std::string name;
    int level = 0;
    float elapsed_ms = 0.0f;
    size_t input_bytes = 0;
    size_t output_bytes = 0;
    float throughput_gbs() const noexcept {
        return elapsed_ms > 0 ? (float)((double)input_bytes / (elapsed_ms * 1e-3) / 1e9) : 0.0f;
    }
};

struct PipelinePerfResult {
    bool is_compress = true;
    float host_elapsed_ms = 0.0f;  // whole call, wall clock
    float dag_elapsed_ms = 0.0f;   // stage execution on the stream (events)
    size_t input_bytes = 0;
    size_t output_bytes = 0;
    std::vector<StageTimingResult> stages;
    // uncompressed bytes / dag_elapsed_ms  (src/pipeline/perf.cpp:114-124)
    float throughput_gbs() const noexcept;
    float pipeline_throughput_gbs() const noexcept;
};

class Pipeline {
  public:
    explicit Pipeline(size_t input_data_size = 0, MemoryStrategy strategy = MemoryStrategy::MINIMAL,
                      float pool_multiplier = 3.0f);
    ~Pipeline();
    Pipeline(const Pipeline&) = delete;
    Pipeline& operator=(const Pipeline&) = delete;

    void setMemoryStrategy(MemoryStrategy s) { strategy_ = s; }
    // One stream executes the chain; accepted for interface compatibility.
    void setNumStreams(int n) { num_streams_ = n > 0 ? n : 1; }
    void setDims(size_t x, size_t y = 1, size_t z = 1) { dims_ = {x, y, z}; }
    void setDims(std::array<size_t, 3> d) { dims_ = d; }
    std::array<size_t, 3> getDims() const { return dims_; }

    template <typename StageT, typename... Args>
    StageT* addStage(Args&&... args) {
        if (finalized_) throw std::runtime_error("Cannot add stages after finalize()");
        auto* s = new StageT(std::forward<Args>(args)...);
        addRawStage(s);
        return s;
    }
    Stage* addRawStage(Stage* stage);

    // Returns 0 on success (compressor.h:92-95).
    int connect(Stage* dependent, Stage* producer, const std::string& output_name = "output");
    int connect(Stage* dependent, const std::vector<Stage*>& producers);

    void finalize();
    bool isFinalized() const { return finalized_; }
    void warmup(hipStream_t stream = nullptr);
    void setWarmupOnFinalize(bool on) { warmup_on_finalize_ = on; }
    bool isWarmupOnFinalizeEnabled() const { return warmup_on_finalize_; }
    void setPoolManagedDecompOutput(bool on) { pool_managed_decomp_ = on; }
    bool isPoolManagedDecompOutput() const { return pool_managed_decomp_; }

    size_t getMaxCompressedSize(size_t input_bytes) const;
    size_t getLastUncompressedSize() const { return input_size_; }
    size_t getInputAlignment() const { return input_alignment_; }

    // Pool-owned output, valid until the next compress().
    void compress(const void* d_input, size_t input_size, void** d_output, size_t* output_size,
                  hipStream_t stream = nullptr);
    // Caller-owned output buffer.
    void compress(const void* d_input, size_t input_size, void* d_output_buf, size_t output_buf_capacity,
                  size_t* actual_output_size, hipStream_t stream = nullptr);
    // d_input is what compress() returned.  The result is pool-owned (valid
    // until the next decompress) if setPoolManagedDecompOutput(true), else a
    // hipMalloc'd buffer the caller frees with hipFree.
    void decompress(const void* d_input, size_t input_size, void** d_output, size_t* output_size,
                    hipStream_t stream = nullptr);
    void decompress(const void* d_input, size_t input_size, void* d_output_buf, size_t output_buf_capacity,
                    size_t* actual_output_size, hipStream_t stream = nullptr);
    void reset(hipStream_t stream = nullptr);
    // Extension (no counterpart in the reference): lets decompress() decode
    // a result that another Pipeline with the same stages produced from
    // `uncompressed_bytes` of input.  The stages' state must first be
    // restored from their recorded headers (Stage::deserializeHeader), as
    // decompressFromFile does from FZM stage records.
    void prepareDecompress(size_t uncompressed_bytes);

    void enableProfiling(bool on) { profiling_ = on; }
    bool isProfilingEnabled() const { return profiling_; }
    const PipelinePerfResult& getLastPerfResult() const { return last_perf_; }

    bool isMemPoolFallbackMode() const { return pool_ && pool_->isFallbackMode(); }
    size_t getPeakMemoryUsage() const;
    size_t getCurrentMemoryUsage() const;
    void printPipeline() const;

    // Graph mode: compress() is captured into a hipGraph on its first call
    // for a given (input pointer, size) and replayed afterwards.  Needs every
    // stage isGraphCompatible().
    // Stage fusion: a BitshuffleStage (16 KiB blocks, element width 1/2)
    // feeding an RZEStage runs as one kernel pass (identical bytes, no
    // shuffled intermediate).  On by default.
    void setStageFusion(bool on) { fusion_ = on; }
    bool isStageFusionEnabled() const { return fusion_; }

    void enableGraphMode(bool on);
    bool isGraphModeEnabled() const { return graph_mode_; }
    void captureGraph(hipStream_t stream = nullptr);
    bool isGraphCaptured() const { return graph_exec_ != nullptr; }

    struct OutputBuffer {
        void* d_ptr;
        size_t actual_size;
        size_t allocated_size;
        std::string name;
        int buffer_id;
    };
    std::vector<OutputBuffer> getOutputBuffers() const;

    struct FZMFileHeader {
        FZMHeaderCore core;
        std::vector<FZMStageInfo> stages;
        std::vector<FZMBufferEntry> buffers;
    };
    FZMFileHeader buildHeader() const;
    void writeToFile(const std::string& filename, hipStream_t stream = nullptr);
    static FZMFileHeader readHeader(const std::string& filename);
    // Rebuilds the inverse chain from the file's stage records; the output
    // is hipMalloc'd (caller frees with hipFree).  Its size is the source
    // stage's reconstructed input, i.e. the padded input size, as in the
    // reference (src/pipeline/compressor_io.cpp:628-640).
    static void decompressFromFile(const std::string& filename, void** d_output, size_t* output_size,
                                   hipStream_t stream = nullptr, PipelinePerfResult* perf_out = nullptr);

  private:
    struct Node {
        Stage* stage = nullptr;
        std::vector<int> out_buf;  // buffer id per output
        int in_buf = -1;           // buffer id of the single input
        int level = 0;
        int fused_from = -1;       // RZE node: the Bitshuffle node folded into it
        bool elided = false;       // Bitshuffle node run inside the next RZE node
    };
    struct Buffer {
        int producer = -1;  // node index, -1 = pipeline input
        int out_index = 0;
        int consumer = -1;  // node index, -1 = pipeline output
        void* ptr = nullptr;
        size_t cap = 0;
        size_t size = 0;
        std::string name;
        bool aliased = false;  // points into the concat buffer (not owned)
    };

    int nodeOf(const Stage* s) const;
    void ensureBuffers(size_t padded_input);
    void launchForward(const void* d_in, size_t padded, bool zero_view, hipStream_t stream);
    void finishCompress(hipStream_t stream, void** d_output, size_t* output_size);
    void runInverse(const std::vector<std::pair<const void*, size_t>>& outputs, hipStream_t stream,
                    void* d_final, size_t final_cap, void** d_result, size_t* result_size);
    std::vector<std::pair<const void*, size_t>> parseConcat(const void* d_input, size_t input_size,
                                                            hipStream_t stream);
    void collectReadbacks();
    void destroyGraph();
    bool updateFusion();

    std::unique_ptr<MemoryPool> pool_;
    MemoryStrategy strategy_ = MemoryStrategy::MINIMAL;
    int num_streams_ = 1;
    std::array<size_t, 3> dims_ = {0, 1, 1};
    std::vector<std::unique_ptr<Stage>> stages_;
    struct Edge { Stage* dep; Stage* prod; std::string name; };
    std::vector<Edge> edges_;
    std::vector<Node> nodes_;       // topological order
    std::vector<Buffer> buffers_;   // 0 = pipeline input
    std::vector<int> outputs_;      // buffer ids of pipeline outputs, in order
    bool finalized_ = false, warmup_on_finalize_ = false, pool_managed_decomp_ = false;
    bool profiling_ = false, graph_mode_ = false, compressed_ = false, fusion_ = true;
    size_t input_hint_ = 0, input_alignment_ = 1;
    size_t input_size_ = 0;       // padded input bytes of the last compress
    size_t orig_input_size_ = 0;  // caller's input bytes of the last compress
    size_t buffers_for_ = 0;
    std::vector<size_t> est_;     // per-buffer capacity estimate for buffers_for_
    bool capturing_ = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> stage_events_;
    struct RbMap { uint32_t off; void* h_dst; uint32_t bytes; };
    std::vector<RbMap> rb_map_;
    std::vector<uint8_t> concat_hdr_;
    uint64_t concat_total_ = 0;
    const void* warm_in_ = nullptr;
    size_t warm_size_ = 0;
    void* pad_buf_ = nullptr;
    size_t pad_cap_ = 0;
    void* concat_buf_ = nullptr;
    size_t concat_cap_ = 0;
    void* concat_ws_ = nullptr;   // concat offsets + u64 total
    uint8_t* rb_host_ = nullptr;  // pinned, device-mapped
    void* rb_dev_ = nullptr;
    size_t last_output_size_ = 0;
    std::vector<uint64_t> last_sizes_;
    // inverse-side buffers (one per forward buffer), pool-owned
    std::vector<void*> inv_buf_;
    std::vector<size_t> inv_cap_;
    void* decomp_out_ = nullptr;
    size_t decomp_cap_ = 0;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    const void* graph_in_ = nullptr;
    size_t graph_size_ = 0;
    hipStream_t graph_stream_ = nullptr;
    PipelinePerfResult last_perf_;
};

}  // namespace fz

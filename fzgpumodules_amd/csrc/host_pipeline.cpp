// host_pipeline.cpp — fz::Pipeline: topology, buffers, compress/decompress,
// graph capture and FZM files.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <stdexcept>
#include <string>

#include "fz/pipeline.h"
#include "fz/stages.h"
#include "fzh.h"
#include "fzh_internal.h"

namespace fz {

namespace {

void hipck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

constexpr size_t kReadbackBytes = 4096;
constexpr size_t kConcatWsBytes = 1024;  // [0, 512): offsets/sizes table; [512]: u64 total

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    float ms() const {
        return std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

uint32_t headerChecksum(const Pipeline::FZMFileHeader& fh) {
    FZMHeaderCore core = fh.core;
    core.header_checksum = 0;
    uint32_t crc = crc32Feed(0xFFFFFFFFu, &core, sizeof(core));
    if (!fh.stages.empty()) crc = crc32Feed(crc, fh.stages.data(), fh.stages.size() * sizeof(FZMStageInfo));
    if (!fh.buffers.empty()) crc = crc32Feed(crc, fh.buffers.data(), fh.buffers.size() * sizeof(FZMBufferEntry));
    return crc ^ 0xFFFFFFFFu;
}

std::string hex32(uint32_t v) {
    char b[16];
    std::snprintf(b, sizeof(b), "0x%08X", v);
    return b;
}

}  // namespace

float PipelinePerfResult::throughput_gbs() const noexcept {
    const size_t u = is_compress ? input_bytes : output_bytes;
    return dag_elapsed_ms > 0 ? (float)((double)u / (dag_elapsed_ms * 1e-3) / 1e9) : 0.0f;
}
float PipelinePerfResult::pipeline_throughput_gbs() const noexcept {
    const size_t u = is_compress ? input_bytes : output_bytes;
    return host_elapsed_ms > 0 ? (float)((double)u / (host_elapsed_ms * 1e-3) / 1e9) : 0.0f;
}

// ---------------------------------------------------------------------------
// Construction and topology
// ---------------------------------------------------------------------------
Pipeline::Pipeline(size_t input_data_size, MemoryStrategy strategy, float pool_multiplier)
    : strategy_(strategy), input_hint_(input_data_size) {
    pool_ = std::make_unique<MemoryPool>(MemoryPoolConfig(input_data_size, pool_multiplier));
}

Pipeline::~Pipeline() {
    (void)hipDeviceSynchronize();
    destroyGraph();
    if (graph_stream_) (void)hipStreamDestroy(graph_stream_);
    for (auto& b : buffers_)
        if (b.producer >= 0 && b.ptr && !b.aliased) pool_->freePersistentDevice(b.ptr);
    for (void* p : inv_buf_)
        if (p) pool_->freePersistentDevice(p);
    if (pad_buf_) pool_->freePersistentDevice(pad_buf_);
    if (concat_buf_) pool_->freePersistentDevice(concat_buf_);
    if (concat_ws_) pool_->freePersistentDevice(concat_ws_);
    if (decomp_out_) pool_->freePersistentDevice(decomp_out_);
    if (rb_host_) (void)hipHostFree(rb_host_);
    stages_.clear();  // stage scratch goes back to the pool before it is destroyed
}

Stage* Pipeline::addRawStage(Stage* stage) {
    if (finalized_) throw std::runtime_error("Cannot add stages after finalize()");
    if (!stage) throw std::invalid_argument("addStage: null stage");
    stages_.emplace_back(stage);
    return stage;
}

int Pipeline::connect(Stage* dependent, Stage* producer, const std::string& output_name) {
    if (finalized_) throw std::runtime_error("Cannot connect stages after finalize()");
    auto owned = [&](Stage* s) {
        for (auto& p : stages_)
            if (p.get() == s) return true;
        return false;
    };
    if (!owned(dependent) || !owned(producer))
        throw std::runtime_error("connect: both stages must be added to this pipeline first");
    if (dependent == producer) throw std::runtime_error("connect: a stage cannot consume its own output");
    if (producer->getOutputIndex(output_name) < 0)
        throw std::runtime_error("connect: stage '" + producer->getName() + "' has no output named '" + output_name + "'");
    edges_.push_back({dependent, producer, output_name});
    return 0;
}

int Pipeline::connect(Stage* dependent, const std::vector<Stage*>& producers) {
    if (producers.size() != 1)
        throw std::runtime_error("connect: stages of this build take one input; got " +
                                 std::to_string(producers.size()) + " producers");
    return connect(dependent, producers[0], producers[0]->getOutputNames().at(0));
}

int Pipeline::nodeOf(const Stage* s) const {
    for (size_t i = 0; i < nodes_.size(); i++)
        if (nodes_[i].stage == s) return (int)i;
    return -1;
}

void Pipeline::finalize() {
    if (finalized_) return;
    if (stages_.empty()) throw std::runtime_error("Pipeline has no stages");
    // one producer edge per dependent
    for (size_t i = 0; i < edges_.size(); i++)
        for (size_t j = i + 1; j < edges_.size(); j++)
            if (edges_[i].dep == edges_[j].dep)
                throw std::runtime_error("Stage '" + edges_[i].dep->getName() +
                                         "' has more than one producer; stages of this build take one input");
    std::vector<Stage*> sources;
    for (auto& sp : stages_) {
        bool dep = false;
        for (auto& e : edges_) dep |= e.dep == sp.get();
        if (!dep) sources.push_back(sp.get());
    }
    if (sources.size() != 1)
        throw std::runtime_error("Pipeline needs exactly one source stage; found " + std::to_string(sources.size()));
    // breadth-first topological order from the source (edges in connect order)
    nodes_.clear();
    nodes_.push_back(Node{sources[0], {}, 0, 0});
    for (size_t k = 0; k < nodes_.size(); k++)
        for (auto& e : edges_)
            if (e.prod == nodes_[k].stage) {
                if (nodeOf(e.dep) >= 0) throw std::runtime_error("Pipeline has a cycle");
                nodes_.push_back(Node{e.dep, {}, -1, nodes_[k].level + 1});
            }
    if (nodes_.size() != stages_.size())
        throw std::runtime_error("Pipeline has stages that are not connected to the source");

    if (dims_[0] > 0 || dims_[1] > 1 || dims_[2] > 1)
        for (auto& n : nodes_) n.stage->setDims(dims_);
    for (auto& n : nodes_) n.stage->setInverse(false);

    buffers_.clear();
    buffers_.push_back(Buffer{-1, 0, 0, nullptr, 0, 0, "input"});
    for (size_t k = 0; k < nodes_.size(); k++) {
        const auto names = nodes_[k].stage->getOutputNames();
        const size_t nout = nodes_[k].stage->getNumOutputs();
        for (size_t i = 0; i < nout; i++) {
            nodes_[k].out_buf.push_back((int)buffers_.size());
            buffers_.push_back(Buffer{(int)k, (int)i, -1, nullptr, 0, 0, i < names.size() ? names[i] : "output"});
        }
    }
    for (auto& e : edges_) {
        const int pk = nodeOf(e.prod), dk = nodeOf(e.dep);
        const int oi = e.prod->getOutputIndex(e.name);
        if (oi < 0 || oi >= (int)nodes_[pk].out_buf.size())
            throw std::runtime_error("connect: bad output '" + e.name + "'");
        Buffer& b = buffers_[nodes_[pk].out_buf[oi]];
        if (b.consumer >= 0)
            throw std::runtime_error("Output '" + e.name + "' of stage '" + e.prod->getName() + "' is consumed twice");
        b.consumer = dk;
        nodes_[dk].in_buf = nodes_[pk].out_buf[oi];
        const uint8_t pt = e.prod->getOutputDataType(oi), ct = e.dep->getInputDataType(0);
        if (pt != (uint8_t)DataType::UNKNOWN && ct != (uint8_t)DataType::UNKNOWN && pt != ct)
            throw std::runtime_error("Type mismatch: '" + e.prod->getName() + "." + e.name + "' produces " +
                                     dataTypeToString((DataType)pt) + " but '" + e.dep->getName() + "' expects " +
                                     dataTypeToString((DataType)ct));
    }
    // Pipeline outputs, deepest stage first (then stage and port order): the
    // first one -- the coder stream of a chain -- is written straight into
    // the concat buffer, the rest are gathered behind it.
    outputs_.clear();
    for (size_t b = 1; b < buffers_.size(); b++)
        if (buffers_[b].consumer < 0) outputs_.push_back((int)b);
    std::stable_sort(outputs_.begin(), outputs_.end(), [&](int a, int b) {
        return nodes_[buffers_[a].producer].level > nodes_[buffers_[b].producer].level;
    });
    if (outputs_.empty()) throw std::runtime_error("Pipeline has no outputs");
    if (outputs_.size() > FZH_CONCAT_MAX_SLOTS || outputs_.size() > FZM_MAX_BUFFERS)
        throw std::runtime_error("Pipeline has more than 32 outputs");

    // input zero-padding granularity: LCM of every stage's alignment
    // (src/pipeline/compressor.cpp:245-261)
    input_alignment_ = 1;
    for (auto& n : nodes_) {
        const size_t a = n.stage->getRequiredInputAlignment();
        if (a > 1) input_alignment_ = std::lcm(input_alignment_, a);
    }

    void* mapped = nullptr;
    hipck(hipHostMalloc(&mapped, kReadbackBytes, hipHostMallocMapped | hipHostMallocCoherent), "hipHostMalloc");
    rb_host_ = (uint8_t*)mapped;
    hipck(hipHostGetDevicePointer(&rb_dev_, mapped, 0), "hipHostGetDevicePointer");
    concat_ws_ = pool_->allocatePersistentDevice(kConcatWsBytes, "concat ws");
    if (!concat_ws_) throw std::runtime_error("concat workspace allocation failed");

    for (auto& n : nodes_) n.stage->onFinalize(input_hint_, pool_.get());
    finalized_ = true;
    if (input_hint_ > 0) ensureBuffers(round_up(input_hint_, input_alignment_));
    if (warmup_on_finalize_ && input_hint_ > 0) warmup(nullptr);
}

// Fold each BitshuffleStage that feeds an RZEStage with matching 16 KiB
// blocks/chunks into the RZE kernels.  Returns true if anything changed.
bool Pipeline::updateFusion() {
    bool changed = false;
    for (auto& n : nodes_) {
        auto* rze = dynamic_cast<RZEStage*>(n.stage);
        if (!rze) continue;
        int from = -1, ew = 0;
        if (fusion_ && n.in_buf > 0) {
            const int k = buffers_[n.in_buf].producer;
            auto* bs = k >= 0 ? dynamic_cast<BitshuffleStage*>(nodes_[k].stage) : nullptr;
            if (bs && bs->getBlockSize() == 16384 && rze->getChunkSize() == 16384 &&
                (bs->getElementWidth() == 1 || bs->getElementWidth() == 2)) {
                from = k;
                ew = (int)bs->getElementWidth();
            }
        }
        if (n.fused_from != from || rze->getPreShuffle() != ew) changed = true;
        if (n.fused_from >= 0) nodes_[n.fused_from].elided = false;
        n.fused_from = from;
        rze->setPreShuffle(ew);
        if (from >= 0) nodes_[from].elided = true;
    }
    if (changed) destroyGraph();
    return changed;
}

void Pipeline::ensureBuffers(size_t padded) {
    if (buffers_for_ == padded) return;
    destroyGraph();
    buffers_[0].cap = padded;
    std::vector<size_t> est(buffers_.size(), 0);
    est[0] = padded;
    for (auto& n : nodes_) {
        const auto e = n.stage->estimateOutputSizes({est[n.in_buf]});
        for (size_t i = 0; i < n.out_buf.size(); i++) est[n.out_buf[i]] = i < e.size() ? e[i] : 0;
    }
    const int primary = outputs_.size() > 1 ? outputs_[0] : -1;
    if (outputs_.size() > 1) {
        std::vector<fzh_concat_slot> slots(outputs_.size());
        for (size_t i = 0; i < outputs_.size(); i++) slots[i].cap = est[outputs_[i]];
        const size_t need = round_up(fzh_concat_bound(slots.data(), (int)slots.size()), 256);
        if (concat_cap_ < need) {
            if (buffers_[primary].ptr && !buffers_[primary].aliased) pool_->freePersistentDevice(buffers_[primary].ptr);
            buffers_[primary].ptr = nullptr;
            if (concat_buf_) pool_->freePersistentDevice(concat_buf_);
            concat_buf_ = pool_->allocatePersistentDevice(need, "concat");
            if (!concat_buf_) throw std::runtime_error("Pipeline: concat allocation failed");
            concat_cap_ = need;
        }
        // the primary output lives in its concat slot (the slot right after
        // the header, 16-byte aligned)
        Buffer& P = buffers_[primary];
        P.ptr = (uint8_t*)concat_buf_ + ((4 + 8 * outputs_.size() + 15) & ~size_t(15));
        P.cap = round_up(std::max<size_t>(est[primary], 16), 16);
        P.aliased = true;
    }
    for (size_t b = 1; b < buffers_.size(); b++) {
        if ((int)b == primary) continue;
        Buffer& B = buffers_[b];
        const size_t need = round_up(std::max<size_t>(est[b], 16), 256);
        if (B.ptr && B.cap >= need && !B.aliased) continue;
        if (B.ptr && !B.aliased) pool_->freePersistentDevice(B.ptr);
        B.aliased = false;
        B.ptr = pool_->allocatePersistentDevice(need, B.name);
        if (!B.ptr) throw std::runtime_error("Pipeline: buffer allocation failed (" + std::to_string(need) + " bytes)");
        B.cap = need;
    }
    // est[] is the per-call capacity of each output (what concat may copy)
    last_sizes_.assign(buffers_.size(), 0);
    for (size_t b = 0; b < buffers_.size(); b++) buffers_[b].size = est[b];
    est_ = est;
    buffers_for_ = padded;
}

size_t Pipeline::getMaxCompressedSize(size_t input_bytes) const {
    if (!finalized_) throw std::runtime_error("getMaxCompressedSize: finalize() first");
    const size_t padded = round_up(input_bytes, input_alignment_);
    std::vector<size_t> est(buffers_.size(), 0);
    est[0] = padded;
    for (auto& n : nodes_) {
        const auto e = n.stage->estimateOutputSizes({est[n.in_buf]});
        for (size_t i = 0; i < n.out_buf.size(); i++) est[n.out_buf[i]] = i < e.size() ? e[i] : 0;
    }
    if (outputs_.size() == 1) return est[outputs_[0]];
    std::vector<fzh_concat_slot> slots(outputs_.size());
    for (size_t i = 0; i < outputs_.size(); i++) slots[i].cap = est[outputs_[i]];
    return fzh_concat_bound(slots.data(), (int)slots.size());
}

// ---------------------------------------------------------------------------
// Compress
// ---------------------------------------------------------------------------
void Pipeline::launchForward(const void* d_in, size_t padded, bool zero_view, hipStream_t stream) {
    (void)zero_view;
    buffers_[0].ptr = const_cast<void*>(d_in);
    buffers_[0].size = padded;
    stage_events_.clear();
    // an RZE stage writing the concat's first slot leaves its pack step to
    // the final concat + readback launch
    RZEStage* defer_rze = nullptr;
    if (outputs_.size() > 1) {
        const Buffer& P = buffers_[outputs_[0]];
        if (P.aliased && P.producer >= 0 && !nodes_[P.producer].elided)
            defer_rze = dynamic_cast<RZEStage*>(nodes_[P.producer].stage);
    }
    for (auto& n : nodes_)
        if (auto* r = dynamic_cast<RZEStage*>(n.stage)) r->setDeferPack(r == defer_rze);
    for (auto& n : nodes_) {
        if (n.elided) {  // runs inside the consuming RZE node
            buffers_[n.out_buf[0]].size = buffers_[n.in_buf].size;
            stage_events_.push_back({nullptr, nullptr});
            continue;
        }
        const Buffer& in = buffers_[n.fused_from >= 0 ? nodes_[n.fused_from].in_buf : n.in_buf];
        std::vector<void*> outs;
        for (int b : n.out_buf) outs.push_back(buffers_[b].ptr);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (profiling_ && !capturing_) {
            hipck(hipEventCreate(&e0), "hipEventCreate");
            hipck(hipEventCreate(&e1), "hipEventCreate");
            hipck(hipEventRecord(e0, stream), "hipEventRecord");
        }
        n.stage->execute(stream, pool_.get(), {in.ptr}, outs, {in.size});
        if (e1) hipck(hipEventRecord(e1, stream), "hipEventRecord");
        stage_events_.push_back({e0, e1});
        for (size_t i = 0; i < n.out_buf.size(); i++) buffers_[n.out_buf[i]].size = n.stage->getActualOutputSize((int)i);
    }
    // device-side concat from device-resident sizes
    rb_map_.clear();
    std::vector<fzh_readback_entry> rb;
    uint32_t off = 0;
    auto add_rb = [&](const void* src, void* h_dst, uint32_t bytes) {
        if (rb.size() >= FZH_READBACK_MAX || off + bytes > kReadbackBytes)
            throw std::runtime_error("Pipeline: too many readback values");
        rb.push_back({src, bytes, off});
        rb_map_.push_back({off, h_dst, bytes});
        off = (uint32_t)round_up(off + bytes, 8);
    };
    if (outputs_.size() > 1) {
        std::vector<fzh_concat_slot> slots(outputs_.size());
        for (size_t i = 0; i < outputs_.size(); i++) {
            const Buffer& B = buffers_[outputs_[i]];
            const Node& P = nodes_[B.producer];
            const DeviceSize ds = P.stage->getDeviceOutputSize(B.out_index);
            fzh_concat_slot& s = slots[i];
            s.src = B.ptr;
            s.cap = est_[outputs_[i]];
            s.kind = ds.kind;
            s.size_ptr = ds.ptr;
            s.elem = ds.elem;
            s.size = ds.kind == 0 ? std::min<size_t>(B.size, s.cap) : 0;
        }
        concat_hdr_.assign(4 + 8 * outputs_.size(), 0);
        add_rb(concat_buf_, concat_hdr_.data(), (uint32_t)round_up(concat_hdr_.size(), 4));
        add_rb((uint8_t*)concat_ws_ + 512, &concat_total_, 8);
        for (auto& n : nodes_)
            for (const auto& r : n.stage->pendingReadbacks()) add_rb(r.d_src, r.h_dst, r.bytes);
        // concat and every readback in one launch (the header and total
        // entries see the values that launch writes), with the RZE pack step
        // when that stage deferred it
        const void* pd = defer_rze ? defer_rze->deferredPack() : nullptr;
        if (defer_rze) defer_rze->setDeferPack(false);
        if (pd)
            throwIfFzh(fzh::rze_pack_concat_readback(stream, *static_cast<const fzh::RzePackDesc*>(pd), slots.data(),
                                                     (int)slots.size(), concat_buf_, concat_cap_,
                                                     (uint64_t*)((uint8_t*)concat_ws_ + 512), rb.data(),
                                                     (int)rb.size(), rb_dev_),
                       "Pipeline concat");
        else
            throwIfFzh(fzh_concat_readback(stream, slots.data(), (int)slots.size(), concat_buf_, concat_cap_,
                                           (uint64_t*)((uint8_t*)concat_ws_ + 512), rb.data(), (int)rb.size(),
                                           rb_dev_),
                       "Pipeline concat");
        return;
    }
    for (auto& n : nodes_)
        for (const auto& r : n.stage->pendingReadbacks()) add_rb(r.d_src, r.h_dst, r.bytes);
    throwIfFzh(fzh_readback(stream, rb.data(), (int)rb.size(), rb_dev_), "Pipeline readback");
}

void Pipeline::collectReadbacks() {
    for (const auto& m : rb_map_) std::memcpy(m.h_dst, rb_host_ + m.off, m.bytes);
    for (auto& n : nodes_) n.stage->markReadbacksDone();
}

void Pipeline::finishCompress(hipStream_t stream, void** d_output, size_t* output_size) {
    hipck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    collectReadbacks();
    for (auto& n : nodes_) n.stage->postStreamSync(stream);
    for (auto& n : nodes_) {
        if (n.elided) continue;
        for (size_t i = 0; i < n.out_buf.size(); i++) buffers_[n.out_buf[i]].size = n.stage->getActualOutputSize((int)i);
    }
    if (outputs_.size() == 1) {
        const Buffer& B = buffers_[outputs_[0]];
        *d_output = B.ptr;
        *output_size = B.size;
    } else {
        // the device concat saw the same sizes the stages report after the sync
        for (size_t i = 0; i < outputs_.size(); i++) {
            uint64_t s;
            std::memcpy(&s, concat_hdr_.data() + 4 + 8 * i, 8);
            if (s != buffers_[outputs_[i]].size)
                throw std::runtime_error("Pipeline: device concat size of '" + buffers_[outputs_[i]].name +
                                         "' (" + std::to_string(s) + ") differs from the stage's (" +
                                         std::to_string(buffers_[outputs_[i]].size) + ")");
        }
        *d_output = concat_buf_;
        *output_size = concat_total_;
    }
    last_output_size_ = *output_size;
    compressed_ = true;
}

void Pipeline::compress(const void* d_input, size_t input_size, void** d_output, size_t* output_size,
                        hipStream_t stream) {
    if (!finalized_) throw std::runtime_error("Pipeline must be finalized before execution");
    if (!d_input) throw std::runtime_error("compress(): null device pointer passed as input");
    if (!d_output || !output_size) throw std::invalid_argument("compress(): null output argument");
    Timer wall;
    const size_t padded = round_up(input_size, input_alignment_);
    ensureBuffers(padded);
    updateFusion();
    input_size_ = padded;
    orig_input_size_ = input_size;
    Stage* src = nodes_[0].stage;

    const void* feed = d_input;
    bool zero_view = false;
    if (padded != input_size) {
        if (src->acceptZeroPaddedInput(input_size)) {
            zero_view = true;
        } else {
            if (pad_cap_ < padded) {
                if (pad_buf_) pool_->freePersistentDevice(pad_buf_);
                pad_buf_ = pool_->allocatePersistentDevice(padded, "input pad");
                if (!pad_buf_) throw std::runtime_error("Failed to allocate pipeline input pad buffer");
                pad_cap_ = padded;
            }
            // the reference's padded copy (src/pipeline/compressor_exec.cpp:35-54)
            hipck(hipMemcpyAsync(pad_buf_, d_input, input_size, hipMemcpyDeviceToDevice, stream), "pad copy");
            const size_t a16 = round_up(input_size, 16);
            if (a16 != input_size)
                hipck(hipMemsetAsync((uint8_t*)pad_buf_ + input_size, 0, a16 - input_size, stream), "pad zero");
            throwIfFzh(fzh_workspace_init(stream, (uint8_t*)pad_buf_ + a16, padded - a16), "pad zero");
            feed = pad_buf_;
        }
    }

    hipEvent_t g0 = nullptr, g1 = nullptr;
    if (profiling_) {
        hipck(hipEventCreate(&g0), "hipEventCreate");
        hipck(hipEventCreate(&g1), "hipEventCreate");
        hipck(hipEventRecord(g0, stream), "hipEventRecord");
    }
    bool all_graph_ok = true;
    for (auto& n : nodes_) all_graph_ok &= n.stage->isGraphCompatible();
    const bool use_graph = graph_mode_ && all_graph_ok && feed == d_input;
    if (use_graph && graph_exec_ && graph_in_ == d_input && graph_size_ == input_size) {
        hipck(hipGraphLaunch(graph_exec_, stream), "hipGraphLaunch");
    } else if (use_graph && warm_in_ == d_input && warm_size_ == input_size) {
        destroyGraph();
        if (!graph_stream_) hipck(hipStreamCreateWithFlags(&graph_stream_, hipStreamNonBlocking), "hipStreamCreate");
        if (zero_view) src->acceptZeroPaddedInput(input_size);
        capturing_ = true;
        hipck(hipStreamBeginCapture(graph_stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            launchForward(feed, padded, zero_view, graph_stream_);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(graph_stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            capturing_ = false;
            throw;
        }
        capturing_ = false;
        hipck(hipStreamEndCapture(graph_stream_, &graph_), "hipStreamEndCapture");
        hipck(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
        graph_in_ = d_input;
        graph_size_ = input_size;
        hipck(hipGraphLaunch(graph_exec_, stream), "hipGraphLaunch");
    } else {
        launchForward(feed, padded, zero_view, stream);
        warm_in_ = d_input;
        warm_size_ = input_size;
    }
    if (profiling_) hipck(hipEventRecord(g1, stream), "hipEventRecord");
    finishCompress(stream, d_output, output_size);

    if (profiling_) {
        PipelinePerfResult r;
        r.is_compress = true;
        hipck(hipEventElapsedTime(&r.dag_elapsed_ms, g0, g1), "hipEventElapsedTime");
        (void)hipEventDestroy(g0);
        (void)hipEventDestroy(g1);
        for (size_t k = 0; k < stage_events_.size() && k < nodes_.size(); k++) {
            StageTimingResult st;
            st.name = nodes_[k].stage->getName() + (nodes_[k].elided ? " (fused into RZE)" : "");
            st.level = nodes_[k].level;
            if (stage_events_[k].first)
                (void)hipEventElapsedTime(&st.elapsed_ms, stage_events_[k].first, stage_events_[k].second);
            st.input_bytes = buffers_[nodes_[k].in_buf].size;
            for (int b : nodes_[k].out_buf) st.output_bytes += buffers_[b].size;
            r.stages.push_back(st);
        }
        r.input_bytes = input_size;
        r.output_bytes = *output_size;
        r.host_elapsed_ms = wall.ms();
        last_perf_ = r;
    }
    for (auto& e : stage_events_) {
        if (e.first) (void)hipEventDestroy(e.first);
        if (e.second) (void)hipEventDestroy(e.second);
    }
    stage_events_.clear();
}

void Pipeline::compress(const void* d_input, size_t input_size, void* d_output_buf, size_t output_buf_capacity,
                        size_t* actual_output_size, hipStream_t stream) {
    if (!d_output_buf || !actual_output_size) throw std::invalid_argument("compress(): null output argument");
    void* out = nullptr;
    size_t sz = 0;
    compress(d_input, input_size, &out, &sz, stream);
    if (sz > output_buf_capacity)
        throw std::runtime_error("compress(): output buffer too small (" + std::to_string(output_buf_capacity) +
                                 " bytes, need " + std::to_string(sz) + ")");
    hipck(hipMemcpyAsync(d_output_buf, out, sz, hipMemcpyDeviceToDevice, stream), "hipMemcpyAsync");
    hipck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    *actual_output_size = sz;
}

void Pipeline::warmup(hipStream_t stream) {
    if (!finalized_) throw std::runtime_error("warmup: finalize() first");
    if (input_hint_ == 0) return;
    void* d = pool_->allocatePersistentDevice(input_hint_, "warmup input");
    if (!d) throw std::runtime_error("warmup: allocation failed");
    throwIfFzh(fzh_workspace_init(stream, d, input_hint_), "warmup");
    void* out = nullptr;
    size_t sz = 0;
    try {
        compress(d, input_hint_, &out, &sz, stream);
    } catch (...) {
        pool_->freePersistentDevice(d);
        throw;
    }
    pool_->freePersistentDevice(d);
    warm_in_ = nullptr;
    compressed_ = false;
}

void Pipeline::reset(hipStream_t stream) {
    if (stream) (void)hipStreamSynchronize(stream);
    compressed_ = false;
}

// ---------------------------------------------------------------------------
// Decompress
// ---------------------------------------------------------------------------
std::vector<std::pair<const void*, size_t>> Pipeline::parseConcat(const void* d_input, size_t input_size,
                                                                  hipStream_t stream) {
    std::vector<std::pair<const void*, size_t>> slots;
    if (outputs_.size() == 1) {
        slots.push_back({d_input, input_size});
        return slots;
    }
    const size_t n = outputs_.size();
    const size_t hdr = round_up(4 + 8 * n, 16);
    if (input_size < hdr) throw std::runtime_error("decompress(): input smaller than the concat header");
    std::vector<uint8_t> h(4 + 8 * n);
    if (reinterpret_cast<uintptr_t>(d_input) % 4 == 0 && round_up(h.size(), 4) <= kReadbackBytes) {
        // one small kernel into the host-mapped readback buffer: a shorter
        // round trip than a device-to-host copy
        const fzh_readback_entry e{d_input, (uint32_t)round_up(h.size(), 4), 0};
        throwIfFzh(fzh_readback(stream, &e, 1, rb_dev_), "decompress(): header readback");
        hipck(hipStreamSynchronize(stream), "hipStreamSynchronize");
        std::memcpy(h.data(), rb_host_, h.size());
    } else {
        hipck(hipMemcpyAsync(h.data(), d_input, h.size(), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
        hipck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    uint32_t cnt;
    std::memcpy(&cnt, h.data(), 4);
    if (cnt != n)
        throw std::runtime_error("decompress(): concat header holds " + std::to_string(cnt) +
                                 " buffers; this pipeline has " + std::to_string(n) + " outputs");
    size_t off = hdr;
    for (size_t i = 0; i < n; i++) {
        uint64_t s;
        std::memcpy(&s, h.data() + 4 + 8 * i, 8);
        if (off + s > input_size) throw std::runtime_error("decompress(): concat slot exceeds the input");
        slots.push_back({(const uint8_t*)d_input + off, (size_t)s});
        off += round_up(s, 16);
    }
    return slots;
}

void Pipeline::runInverse(const std::vector<std::pair<const void*, size_t>>& slots, hipStream_t stream,
                          void* d_final, size_t final_cap, void** d_result, size_t* result_size) {
    if (inv_buf_.size() != buffers_.size()) {
        inv_buf_.resize(buffers_.size(), nullptr);
        inv_cap_.resize(buffers_.size(), 0);
    }
    updateFusion();
    std::vector<const void*> content(buffers_.size(), nullptr);
    std::vector<size_t> csize(buffers_.size(), 0);
    for (size_t i = 0; i < outputs_.size(); i++) {
        content[outputs_[i]] = slots[i].first;
        csize[outputs_[i]] = slots[i].second;
    }
    for (auto& n : nodes_) {
        n.stage->saveState();
        n.stage->setInverse(true);
    }
    void* result = nullptr;
    size_t rsize = 0;
    try {
        for (size_t k = nodes_.size(); k-- > 0;) {
            Node& n = nodes_[k];
            if (n.elided) continue;  // its inverse ran inside the RZE decoder
            // a fused RZE node reconstructs the Bitshuffle node's input
            const int tgt_node = n.fused_from >= 0 ? n.fused_from : (int)k;
            const int tgt_buf = nodes_[tgt_node].in_buf;
            std::vector<void*> ins;
            std::vector<size_t> sizes;
            for (int b : n.out_buf) {
                ins.push_back(const_cast<void*>(content[b]));
                sizes.push_back(csize[b]);
            }
            const size_t need = n.stage->estimateOutputSizes(sizes).at(0);
            void* out;
            if (tgt_node == 0 && d_final && final_cap >= need) {
                out = d_final;
            } else {
                const int b = tgt_buf;
                const size_t cap = round_up(std::max<size_t>(need, 16), 256);
                if (inv_cap_[b] < cap) {
                    if (inv_buf_[b]) pool_->freePersistentDevice(inv_buf_[b]);
                    inv_buf_[b] = pool_->allocatePersistentDevice(cap, "inverse buffer");
                    if (!inv_buf_[b]) throw std::runtime_error("decompress(): allocation failed");
                    inv_cap_[b] = cap;
                }
                out = inv_buf_[b];
            }
            n.stage->execute(stream, pool_.get(), ins, {out}, sizes);
            content[tgt_buf] = out;
            csize[tgt_buf] = n.stage->getActualOutputSize(0);
            if (tgt_node == 0) {
                result = out;
                rsize = csize[tgt_buf];
            }
        }
        std::vector<fzh_readback_entry> rb;
        rb_map_.clear();
        uint32_t off = 0;
        for (auto& n : nodes_)
            for (const auto& r : n.stage->pendingReadbacks()) {
                if (rb.size() >= FZH_READBACK_MAX) throw std::runtime_error("too many readback values");
                rb.push_back({r.d_src, r.bytes, off});
                rb_map_.push_back({off, r.h_dst, r.bytes});
                off = (uint32_t)round_up(off + r.bytes, 8);
            }
        throwIfFzh(fzh_readback(stream, rb.data(), (int)rb.size(), rb_dev_), "decompress readback");
        hipck(hipStreamSynchronize(stream), "hipStreamSynchronize");
        collectReadbacks();
        for (auto& n : nodes_) n.stage->postStreamSync(stream);
    } catch (...) {
        (void)hipStreamSynchronize(stream);
        for (auto& n : nodes_) {
            n.stage->setInverse(false);
            n.stage->restoreState();
        }
        throw;
    }
    for (auto& n : nodes_) {
        n.stage->setInverse(false);
        n.stage->restoreState();
    }
    *d_result = result;
    *result_size = rsize;
}

void Pipeline::prepareDecompress(size_t uncompressed_bytes) {
    if (!finalized_) throw std::runtime_error("Pipeline must be finalized before execution");
    if (uncompressed_bytes == 0) throw std::invalid_argument("prepareDecompress: zero input size");
    const size_t padded = round_up(uncompressed_bytes, input_alignment_);
    ensureBuffers(padded);
    updateFusion();
    input_size_ = padded;
    orig_input_size_ = uncompressed_bytes;
    compressed_ = true;
}

void Pipeline::decompress(const void* d_input, size_t input_size, void** d_output, size_t* output_size,
                          hipStream_t stream) {
    if (!finalized_) throw std::runtime_error("Pipeline must be finalized before execution");
    if (!compressed_) throw std::runtime_error("decompress(): call compress() on this pipeline first");
    if (!d_input || !d_output || !output_size) throw std::invalid_argument("decompress(): null argument");
    Timer wall;
    hipEvent_t g0 = nullptr, g1 = nullptr;
    if (profiling_) {
        hipck(hipEventCreate(&g0), "hipEventCreate");
        hipck(hipEventCreate(&g1), "hipEventCreate");
        hipck(hipEventRecord(g0, stream), "hipEventRecord");
    }
    const auto slots = parseConcat(d_input, input_size, stream);
    void* res = nullptr;
    size_t rs = 0;
    if (pool_managed_decomp_) {
        if (decomp_cap_ < input_size_) {
            if (decomp_out_) pool_->freePersistentDevice(decomp_out_);
            decomp_out_ = pool_->allocatePersistentDevice(round_up(input_size_, 256), "decompressed");
            if (!decomp_out_) throw std::runtime_error("decompress(): allocation failed");
            decomp_cap_ = round_up(input_size_, 256);
        }
        runInverse(slots, stream, decomp_out_, decomp_cap_, &res, &rs);
    } else {
        void* d = nullptr;
        hipck(hipMalloc(&d, std::max<size_t>(input_size_, 16)), "hipMalloc");
        try {
            runInverse(slots, stream, d, input_size_, &res, &rs);
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        if (res != d) {
            hipck(hipMemcpyAsync(d, res, std::min(rs, input_size_), hipMemcpyDeviceToDevice, stream), "hipMemcpyAsync");
            hipck(hipStreamSynchronize(stream), "hipStreamSynchronize");
        }
        res = d;
    }
    *d_output = res;
    *output_size = std::min(rs, orig_input_size_ ? orig_input_size_ : rs);
    if (profiling_) {
        hipck(hipEventRecord(g1, stream), "hipEventRecord");
        hipck(hipEventSynchronize(g1), "hipEventSynchronize");
        PipelinePerfResult r;
        r.is_compress = false;
        hipck(hipEventElapsedTime(&r.dag_elapsed_ms, g0, g1), "hipEventElapsedTime");
        r.input_bytes = input_size;
        r.output_bytes = *output_size;
        r.host_elapsed_ms = wall.ms();
        last_perf_ = r;
    }
    if (g0) (void)hipEventDestroy(g0);
    if (g1) (void)hipEventDestroy(g1);
}

void Pipeline::decompress(const void* d_input, size_t input_size, void* d_output_buf, size_t output_buf_capacity,
                          size_t* actual_output_size, hipStream_t stream) {
    if (!finalized_) throw std::runtime_error("Pipeline must be finalized before execution");
    if (!compressed_) throw std::runtime_error("decompress(): call compress() on this pipeline first");
    if (!d_input || !d_output_buf || !actual_output_size) throw std::invalid_argument("decompress(): null argument");
    const size_t want = orig_input_size_ ? orig_input_size_ : input_size_;
    if (output_buf_capacity < want)
        throw std::runtime_error("decompress(): output buffer too small (" + std::to_string(output_buf_capacity) +
                                 " bytes, need " + std::to_string(want) + ")");
    const auto slots = parseConcat(d_input, input_size, stream);
    void* res = nullptr;
    size_t rs = 0;
    const bool direct = reinterpret_cast<uintptr_t>(d_output_buf) % 16 == 0;
    runInverse(slots, stream, direct ? d_output_buf : nullptr, direct ? output_buf_capacity : 0, &res, &rs);
    const size_t n = std::min(rs, want);
    if (res != d_output_buf) {
        hipck(hipMemcpyAsync(d_output_buf, res, n, hipMemcpyDeviceToDevice, stream), "hipMemcpyAsync");
        hipck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    *actual_output_size = n;
}

// ---------------------------------------------------------------------------
// Graph mode, introspection
// ---------------------------------------------------------------------------
void Pipeline::enableGraphMode(bool on) {
    graph_mode_ = on;
    if (!on) destroyGraph();
}

void Pipeline::captureGraph(hipStream_t) {
    if (!graph_mode_) throw std::runtime_error("captureGraph: enableGraphMode(true) first");
    // Capture happens on the second compress() of a given input (the first
    // one sizes every scratch buffer eagerly).
}

void Pipeline::destroyGraph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    graph_in_ = nullptr;
    graph_size_ = 0;
}

size_t Pipeline::getPeakMemoryUsage() const {
    return pool_->getPeakUsage() + pool_->getPersistentDeviceBytes();
}
size_t Pipeline::getCurrentMemoryUsage() const {
    return pool_->getCurrentUsage() + pool_->getPersistentDeviceBytes();
}

void Pipeline::printPipeline() const {
    std::printf("Pipeline: %zu stage(s), %zu output(s), input alignment %zu\n", nodes_.size(), outputs_.size(),
                input_alignment_);
    for (auto& n : nodes_) {
        std::printf("  [L%d] %s <- buffer %d ->", n.level, n.stage->getName().c_str(), n.in_buf);
        for (int b : n.out_buf) std::printf(" %d(%s%s)", b, buffers_[b].name.c_str(), buffers_[b].consumer < 0 ? ", out" : "");
        std::printf("\n");
    }
}

std::vector<Pipeline::OutputBuffer> Pipeline::getOutputBuffers() const {
    std::vector<OutputBuffer> r;
    for (int b : outputs_) r.push_back({buffers_[b].ptr, buffers_[b].size, buffers_[b].cap, buffers_[b].name, b});
    return r;
}

// ---------------------------------------------------------------------------
// FZM files (src/pipeline/compressor_io.cpp:90-475)
// ---------------------------------------------------------------------------
Pipeline::FZMFileHeader Pipeline::buildHeader() const {
    if (!finalized_) throw std::runtime_error("Cannot build header before finalization");
    if (!compressed_) throw std::runtime_error("Cannot build header before compress() is called");
    FZMFileHeader fh;
    fh.core.uncompressed_size = input_size_;
    fh.core.num_sources = 1;
    fh.core.source_uncompressed_sizes[0] = input_size_;
    for (auto& n : nodes_) {
        FZMStageInfo si;
        si.stage_type = (StageType)n.stage->getStageTypeId();
        si.stage_version = 1;
        si.num_inputs = 1;
        si.num_outputs = (uint8_t)n.out_buf.size();
        si.input_buffer_ids[0] = (uint16_t)n.in_buf;
        for (size_t i = 0; i < n.out_buf.size() && i < FZM_MAX_STAGE_OUTPUTS; i++)
            si.output_buffer_ids[i] = (uint16_t)n.out_buf[i];
        if (si.num_outputs > 0)
            si.config_size = (uint32_t)n.stage->serializeHeader(0, si.stage_config, FZM_STAGE_CONFIG_SIZE);
        fh.stages.push_back(si);
    }
    fh.core.num_stages = (uint32_t)fh.stages.size();
    uint64_t off = 0;
    for (int b : outputs_) {
        const Buffer& B = buffers_[b];
        Stage* p = nodes_[B.producer].stage;
        FZMBufferEntry e;
        e.stage_type = (StageType)p->getStageTypeId();
        e.stage_version = 1;
        e.data_type = (DataType)p->getOutputDataType(B.out_index);
        e.producer_output_idx = (uint8_t)B.out_index;
        e.dag_buffer_id = (uint16_t)b;
        std::strncpy(e.name, B.name.c_str(), FZM_MAX_NAME_LEN - 1);
        e.data_size = B.size;
        e.allocated_size = est_.empty() ? B.cap : est_[b];
        e.uncompressed_size = B.size;
        e.byte_offset = off;
        e.config_size = (uint32_t)p->serializeHeader(B.out_index, e.stage_config, FZM_STAGE_CONFIG_SIZE);
        off += B.size;
        fh.buffers.push_back(e);
    }
    fh.core.num_buffers = (uint16_t)fh.buffers.size();
    fh.core.compressed_size = off;
    fh.core.header_size = fh.core.computeHeaderSize();
    return fh;
}

void Pipeline::writeToFile(const std::string& filename, hipStream_t stream) {
    if (!finalized_) throw std::runtime_error("Cannot write to file before finalization");
    if (!compressed_) throw std::runtime_error("compress() must be called before writeToFile()");
    FZMFileHeader fh = buildHeader();
    std::vector<uint8_t> payload(fh.core.compressed_size);
    size_t off = 0;
    for (int b : outputs_) {
        if (buffers_[b].size)
            hipck(hipMemcpyAsync(payload.data() + off, buffers_[b].ptr, buffers_[b].size, hipMemcpyDeviceToHost, stream),
                  "hipMemcpyAsync");
        off += buffers_[b].size;
    }
    hipck(hipStreamSynchronize(stream), "hipStreamSynchronize");
    fh.core.data_checksum = crc32(payload.data(), payload.size());
    fh.core.flags |= FZM_FLAG_HAS_DATA_CHECKSUM | FZM_FLAG_HAS_HEADER_CHECKSUM;
    fh.core.header_checksum = headerChecksum(fh);
    std::ofstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("Failed to open file for writing: " + filename);
    f.write((const char*)&fh.core, sizeof(fh.core));
    if (!fh.stages.empty()) f.write((const char*)fh.stages.data(), fh.stages.size() * sizeof(FZMStageInfo));
    if (!fh.buffers.empty()) f.write((const char*)fh.buffers.data(), fh.buffers.size() * sizeof(FZMBufferEntry));
    f.write((const char*)payload.data(), payload.size());
    if (!f) throw std::runtime_error("Failed to write file: " + filename);
}

Pipeline::FZMFileHeader Pipeline::readHeader(const std::string& filename) {
    std::ifstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("Failed to open file for reading: " + filename);
    uint32_t magic = 0;
    uint16_t version = 0;
    f.read((char*)&magic, 4);
    f.read((char*)&version, 2);
    if (!f) throw std::runtime_error("Failed to read header from file: " + filename);
    if (magic != FZM_MAGIC) throw std::runtime_error("Invalid FZM file format (bad magic number)");
    const uint8_t major = fzmVersionMajor(version), minor = fzmVersionMinor(version);
    if (major != FZM_VERSION_MAJOR)
        throw std::runtime_error("Incompatible FZM major version: file has major " + std::to_string(major) +
                                 ", reader expects major " + std::to_string(FZM_VERSION_MAJOR));
    const bool exact = minor == FZM_VERSION_MINOR, legacy = minor == 0;
    f.seekg(0);
    FZMFileHeader fh;
    f.read((char*)&fh.core, legacy ? FZM_LEGACY_HEADER_CORE_SIZE : sizeof(FZMHeaderCore));
    if (!f) throw std::runtime_error("Failed to read header from file: " + filename);
    if (fh.core.num_stages > FZM_MAX_BUFFERS || fh.core.num_buffers > FZM_MAX_BUFFERS)
        throw std::runtime_error("FZM header has too many stages/buffers");
    fh.stages.resize(fh.core.num_stages);
    if (fh.core.num_stages) f.read((char*)fh.stages.data(), fh.stages.size() * sizeof(FZMStageInfo));
    fh.buffers.resize(fh.core.num_buffers);
    if (fh.core.num_buffers) f.read((char*)fh.buffers.data(), fh.buffers.size() * sizeof(FZMBufferEntry));
    if (!f) throw std::runtime_error("Failed to read stage/buffer records from file: " + filename);
    if (!legacy && exact && (fh.core.flags & FZM_FLAG_HAS_HEADER_CHECKSUM)) {
        const uint32_t c = headerChecksum(fh);
        if (c != fh.core.header_checksum)
            throw std::runtime_error("FZM header checksum mismatch in '" + filename + "': expected " +
                                     hex32(fh.core.header_checksum) + ", computed " + hex32(c) +
                                     " — file may be corrupted");
    }
    return fh;
}

void Pipeline::decompressFromFile(const std::string& filename, void** d_output, size_t* output_size,
                                  hipStream_t stream, PipelinePerfResult* perf_out) {
    if (!d_output || !output_size) throw std::invalid_argument("decompressFromFile: null output argument");
    Timer wall;
    FZMFileHeader fh = readHeader(filename);
    std::vector<uint8_t> payload(fh.core.compressed_size);
    {
        std::ifstream f(filename, std::ios::binary);
        f.seekg((std::streamoff)fh.core.header_size);
        f.read((char*)payload.data(), (std::streamsize)payload.size());
        if (!f) throw std::runtime_error("Failed to read compressed data from file: " + filename);
    }
    if (fh.core.flags & FZM_FLAG_HAS_DATA_CHECKSUM) {
        const uint32_t c = crc32(payload.data(), payload.size());
        if (c != fh.core.data_checksum)
            throw std::runtime_error("FZM data checksum mismatch in '" + filename + "': expected " +
                                     hex32(fh.core.data_checksum) + ", computed " + hex32(c) +
                                     " — compressed data may be corrupted");
    }
    for (const auto& e : fh.buffers)
        if (e.byte_offset + e.data_size > payload.size())
            throw std::runtime_error("FZM buffer '" + std::string(e.name) + "' exceeds the payload");

    // Stage records -> inverse stages; buffer ids route the data.
    std::vector<std::unique_ptr<Stage>> st;
    for (const auto& si : fh.stages) {
        st.emplace_back(createStage(si.stage_type, si.stage_config, si.config_size));
        st.back()->setInverse(true);
    }
    const size_t ns = st.size();
    if (ns == 0) throw std::runtime_error("decompressFromFile: file has no stages");
    std::vector<int> producer_of(65536, -1);
    for (size_t k = 0; k < ns; k++)
        for (int j = 0; j < fh.stages[k].num_outputs && j < (int)FZM_MAX_STAGE_OUTPUTS; j++)
            if (fh.stages[k].output_buffer_ids[j] != 0xFFFF) producer_of[fh.stages[k].output_buffer_ids[j]] = (int)k;
    int source = -1;
    for (size_t k = 0; k < ns && source < 0; k++) {
        const uint16_t in = fh.stages[k].input_buffer_ids[0];
        if (in == 0xFFFF || producer_of[in] < 0) source = (int)k;
    }
    if (source < 0) throw std::runtime_error("decompressFromFile: could not identify any source stage in topology");

    // Each 16-byte-aligned slot of one device copy of the payload.
    std::vector<size_t> slot_off(fh.buffers.size());
    size_t total = 0;
    for (size_t i = 0; i < fh.buffers.size(); i++) {
        slot_off[i] = total;
        total += round_up(fh.buffers[i].data_size, 16);
    }
    std::vector<uint8_t> staged(std::max<size_t>(total, 16), 0);
    for (size_t i = 0; i < fh.buffers.size(); i++)
        std::memcpy(staged.data() + slot_off[i], payload.data() + fh.buffers[i].byte_offset, fh.buffers[i].data_size);
    void* d_comp = nullptr;
    hipck(hipMalloc(&d_comp, staged.size()), "hipMalloc");
    std::vector<void*> owned{d_comp};
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(stream);
        for (void* p : owned) (void)hipFree(p);
    };
    try {
        hipck(hipMemcpyAsync(d_comp, staged.data(), staged.size(), hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
        std::vector<const void*> content(65536, nullptr);
        std::vector<size_t> csize(65536, 0);
        for (size_t i = 0; i < fh.buffers.size(); i++) {
            content[fh.buffers[i].dag_buffer_id] = (uint8_t*)d_comp + slot_off[i];
            csize[fh.buffers[i].dag_buffer_id] = fh.buffers[i].data_size;
        }
        // Inverse in reverse record order (records are in forward level order).
        std::vector<bool> done(ns, false);
        void* result = nullptr;
        size_t rsize = 0;
        MemoryPool* no_pool = nullptr;
        for (size_t pass = 0; pass < ns; pass++) {
            int pick = -1;
            for (size_t k = ns; k-- > 0;) {
                if (done[k]) continue;
                bool ready = true;
                for (int j = 0; j < fh.stages[k].num_outputs && j < (int)FZM_MAX_STAGE_OUTPUTS; j++) {
                    const uint16_t b = fh.stages[k].output_buffer_ids[j];
                    if (b != 0xFFFF && !content[b]) ready = false;
                }
                if (ready) { pick = (int)k; break; }
            }
            if (pick < 0) throw std::runtime_error("decompressFromFile: a stage's inputs are missing from the file");
            const FZMStageInfo& si = fh.stages[pick];
            std::vector<void*> ins;
            std::vector<size_t> sizes;
            for (int j = 0; j < si.num_outputs && j < (int)FZM_MAX_STAGE_OUTPUTS; j++) {
                const uint16_t b = si.output_buffer_ids[j];
                ins.push_back(const_cast<void*>(content[b]));
                sizes.push_back(csize[b]);
            }
            const size_t need = std::max<size_t>(st[pick]->estimateOutputSizes(sizes).at(0), 16);
            void* out = nullptr;
            hipck(hipMalloc(&out, round_up(need, 256)), "hipMalloc");
            owned.push_back(out);
            st[pick]->execute(stream, no_pool, ins, {out}, sizes);
            const uint16_t ib = si.input_buffer_ids[0];
            if (ib != 0xFFFF) {
                content[ib] = out;
                csize[ib] = st[pick]->getActualOutputSize(0);
            }
            if (pick == source) {
                result = out;
                rsize = st[pick]->getActualOutputSize(0);
            }
            done[pick] = true;
        }
        hipck(hipStreamSynchronize(stream), "hipStreamSynchronize");
        for (auto& s : st) s->postStreamSync(stream);
        // hand the result buffer to the caller; free the rest
        for (auto& p : owned)
            if (p == result) p = nullptr;
        cleanup();
        *d_output = result;
        *output_size = rsize;
        if (perf_out) {
            perf_out->is_compress = false;
            perf_out->host_elapsed_ms = wall.ms();
            perf_out->dag_elapsed_ms = perf_out->host_elapsed_ms;
            perf_out->input_bytes = fh.core.compressed_size;
            perf_out->output_bytes = rsize;
            perf_out->stages.clear();
        }
    } catch (...) {
        cleanup();
        throw;
    }
}

}  // namespace fz

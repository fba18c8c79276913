// host_batch.cpp — fz::BatchCompressor (include/fz/batch.h) and its C ABI
// (include/fzb.h): plans the device tables of a batch of fields, runs the
// persistent kernels of batch.hip (captured as one hipGraph from the second
// call on) and tunes their launch geometry.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fz/batch.h"
#include "fzb.h"
#include "fzh_batch.h"
#include "fzh_internal.h"

namespace fz {

namespace {

void hipck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void fzck(int rc, const char* what) {
    if (rc != FZH_OK) throw std::runtime_error(std::string(what) + ": " + fzh_last_error());
}

uint64_t align_to(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

}  // namespace

struct BatchCompressor::Impl {
    fzh::batch::Plan plan{};
    fzh::batch::Plan* d_plan = nullptr;  // device copy read by the kernels
    std::vector<void*> allocs;
    fzh::batch::Result* h_results = nullptr;  // host-mapped
    void* d_results_map = nullptr;
    hipEvent_t ev[9] = {};

    void* dev(size_t bytes) {
        void* p = nullptr;
        hipck(hipMalloc(&p, std::max<size_t>(bytes, 256)), "hipMalloc (batch)");
        allocs.push_back(p);
        return p;
    }
    ~Impl() {
        for (void* p : allocs) (void)hipFree(p);
        if (h_results) (void)hipHostFree(h_results);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

BatchCompressor::BatchCompressor(const BatchConfig& cfg) : cfg_(cfg) {
    if (!(cfg.error_bound > 0.0f) || !std::isfinite(cfg.error_bound))
        throw std::invalid_argument("BatchCompressor: error_bound must be positive and finite");
    if (cfg.quant_radius < 1 || cfg.quant_radius > 32768)
        throw std::invalid_argument("BatchCompressor: quant_radius must be in [1, 32768]");
    if (!(cfg.outlier_capacity >= 0.0f) || cfg.outlier_capacity > 1.0f)
        throw std::invalid_argument("BatchCompressor: outlier_capacity must be in [0, 1]");
    if (!(cfg.patch_capacity >= 0.0f) || cfg.patch_capacity > 1.0f)
        throw std::invalid_argument("BatchCompressor: patch_capacity must be in [0, 1]");
    int dev = 0;
    hipck(hipGetDevice(&dev), "hipGetDevice");
    hipck(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    cus_ = std::max(cus_, 1);
}

BatchCompressor::~BatchCompressor() {
    destroyGraph();
    if (graph_stream_) (void)hipStreamDestroy(graph_stream_);
    release();
}

void BatchCompressor::release() {
    delete impl_;
    impl_ = nullptr;
    d_out_ = nullptr;
    out_cap_ = 0;
}

void BatchCompressor::destroyGraph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
    calls_since_plan_ = 0;
}

void BatchCompressor::setFields(const std::vector<BatchField>& fields) {
    using namespace fzh::batch;
    if (fields.empty()) throw std::invalid_argument("BatchCompressor: no fields");
    if (fields.size() > (1u << 24)) throw std::invalid_argument("BatchCompressor: more than 2^24 fields");
    destroyGraph();
    release();
    fields_ = fields;
    const size_t F = fields.size();
    std::vector<Field> tab(F);
    std::vector<uint32_t> tile0(F + 1), chunk0(F + 1), unit0(F + 1);
    uint64_t T = 0, NC = 0, U = 0, codes = 0, out = 0;
    in_bytes_ = 0;
    out_off_.assign(F, 0);
    const bool strict = cfg_.strict;
    const uint32_t nout = strict ? 7 : 4;
    const uint32_t hdr = (uint32_t)align_to(4 + 8ull * nout, 16);
    for (size_t i = 0; i < F; i++) {
        const BatchField& b = fields[i];
        if (!b.d_in) throw std::invalid_argument("BatchCompressor: field " + std::to_string(i) + " has no data");
        if (reinterpret_cast<uintptr_t>(b.d_in) % 16)
            throw std::invalid_argument("BatchCompressor: field " + std::to_string(i) + " is not 16-byte aligned");
        if (b.nx == 0 || b.ny == 0 || b.nz == 0)
            throw std::invalid_argument("BatchCompressor: field " + std::to_string(i) + " is empty");
        const uint64_t nv = (uint64_t)b.nx * b.ny * b.nz;
        const uint64_t bytes = 4 * nv;
        const uint64_t n = align_to(bytes, kChunk) / 4;  // Pipeline pads the input to 16 KiB (Bitshuffle/RZE)
        if (n > 0xFFFFFFFFull) throw std::invalid_argument("BatchCompressor: field larger than 2^32 elements");
        Field& f = tab[i];
        std::memset(&f, 0, sizeof(f));
        f.in = b.d_in;
        f.n = n;
        f.nv = nv;
        f.nx = (uint32_t)b.nx;
        f.ny = (uint32_t)b.ny;
        f.nz = (uint32_t)b.nz;
        f.ndim = b.nz > 1 ? 3 : (b.ny > 1 ? 2 : 1);
        uint64_t ntiles;
        if (f.ndim == 1) {
            ntiles = n / 1024;
        } else if (f.ndim == 2) {
            f.gx = (uint32_t)fzh::cdiv(b.nx, 32);
            ntiles = (uint64_t)f.gx * fzh::cdiv(b.ny, 32);
        } else {
            f.gx = (uint32_t)fzh::cdiv(b.nx, 8);
            f.gy = (uint32_t)fzh::cdiv(b.ny, 8);
            ntiles = (uint64_t)f.gx * f.gy * fzh::cdiv(b.nz, 8);
        }
        const uint64_t nch = fzh::cdiv(2 * n, kChunk);
        const uint64_t nun = fzh::cdiv(nv, kVbUnit);
        // LorenzoQuantStage's capacities (maxOutliers / maxPatches)
        const uint64_t max_out = (uint64_t)std::ceil((double)n * cfg_.outlier_capacity);
        const uint64_t max_patch = std::max<uint64_t>(256, (uint64_t)std::ceil((double)n * cfg_.patch_capacity));
        f.tile0 = (uint32_t)T;
        f.ntiles = (uint32_t)ntiles;
        f.chunk0 = (uint32_t)NC;
        f.nchunks = (uint32_t)nch;
        f.unit0 = (uint32_t)U;
        f.nunits = (uint32_t)nun;
        f.max_out = (uint32_t)std::min<uint64_t>(max_out, 0xFFFFFFFFu);
        f.max_patch = strict ? (uint32_t)std::min<uint64_t>(max_patch, 0xFFFFFFFFu) : 0u;
        f.codes_off = codes;
        uint64_t cap = hdr + align_to(fzh_rze_enc_bound(2 * n), 16) + 2 * align_to(4 * max_out, 16) + 16;
        if (strict) cap += 2 * align_to(4 * max_patch, 16) + 16;
        f.out_off = out;
        f.out_cap = cap;
        out_off_[i] = out;
        tile0[i] = (uint32_t)T;
        chunk0[i] = (uint32_t)NC;
        unit0[i] = (uint32_t)U;
        T += ntiles;
        NC += nch;
        U += nun;
        codes += n;
        out += align_to(cap, 256);
        in_bytes_ += bytes;
        if (T >= 0xFFFFFFF0ull || NC >= 0x0FFFFFFFull || U >= 0xFFFFFFF0ull)
            throw std::invalid_argument("BatchCompressor: batch too large (tile/chunk index space)");
    }
    tile0[F] = (uint32_t)T;
    chunk0[F] = (uint32_t)NC;
    unit0[F] = (uint32_t)U;
    const uint64_t n_counts = T * (strict ? 2 : 1) + NC;
    if (n_counts >= 0xFFFFFFF0ull) throw std::invalid_argument("BatchCompressor: batch too large (count table)");

    impl_ = new Impl();
    Plan& p = impl_->plan;
    p.F = (uint32_t)F;
    p.T = (uint32_t)T;
    p.NC = (uint32_t)NC;
    p.U = (uint32_t)U;
    p.user_eb = cfg_.error_bound;
    p.mode = (int)cfg_.eb_mode;
    p.rad = (float)(uint16_t)cfg_.quant_radius;
    p.zigzag = cfg_.zigzag ? 1 : 0;
    p.strict = strict ? 1 : 0;
    p.hdr = hdr;
    p.nout = nout;
    Field* d_tab = (Field*)impl_->dev(sizeof(Field) * F);
    uint32_t* d_t0 = (uint32_t*)impl_->dev(4 * (F + 1));
    uint32_t* d_c0 = (uint32_t*)impl_->dev(4 * (F + 1));
    uint32_t* d_u0 = (uint32_t*)impl_->dev(4 * (F + 1));
    hipck(hipMemcpy(d_tab, tab.data(), sizeof(Field) * F, hipMemcpyHostToDevice), "hipMemcpy (fields)");
    hipck(hipMemcpy(d_t0, tile0.data(), 4 * (F + 1), hipMemcpyHostToDevice), "hipMemcpy (tiles)");
    hipck(hipMemcpy(d_c0, chunk0.data(), 4 * (F + 1), hipMemcpyHostToDevice), "hipMemcpy (chunks)");
    hipck(hipMemcpy(d_u0, unit0.data(), 4 * (F + 1), hipMemcpyHostToDevice), "hipMemcpy (units)");
    // tiles by dimensionality: per class, its fields, their tile prefix and
    // the class field of every 64th class tile (the Lorenzo kernels start
    // their walk there instead of a binary search)
    for (int d = 0; d < 3; d++) {
        std::vector<uint32_t> cf, ct0{0};
        for (size_t i = 0; i < F; i++)
            if (tab[i].ndim == (uint32_t)d + 1) {
                cf.push_back((uint32_t)i);
                ct0.push_back(ct0.back() + tab[i].ntiles);
            }
        const uint32_t Td = ct0.back();
        const size_t nb = std::max<size_t>((Td + 63) / 64, 1);
        std::vector<uint32_t> tf(nb, 0);
        for (size_t k = 0, j = 0; k < nb && !cf.empty(); k++) {
            while (j + 1 < cf.size() && ct0[j + 1] <= 64 * k) j++;
            tf[k] = (uint32_t)j;
        }
        if (cf.empty()) cf.push_back(0);
        uint32_t* d_cf = (uint32_t*)impl_->dev(4 * cf.size());
        uint32_t* d_ct0 = (uint32_t*)impl_->dev(4 * ct0.size());
        uint32_t* d_ctf = (uint32_t*)impl_->dev(4 * nb);
        hipck(hipMemcpy(d_cf, cf.data(), 4 * cf.size(), hipMemcpyHostToDevice), "hipMemcpy (class fields)");
        hipck(hipMemcpy(d_ct0, ct0.data(), 4 * ct0.size(), hipMemcpyHostToDevice), "hipMemcpy (class tiles)");
        hipck(hipMemcpy(d_ctf, tf.data(), 4 * nb, hipMemcpyHostToDevice), "hipMemcpy (class tile fields)");
        p.cls_field[d] = d_cf;
        p.cls_tile0[d] = d_ct0;
        p.cls_tfield[d] = d_ctf;
        p.cls_F[d] = (uint32_t)(ct0.size() - 1);
        p.cls_T[d] = Td;
        if (d == 2) {  // 3D tiles in x-adjacent pairs of one field: ceil(gx / 2) gy gz per field
            std::vector<uint32_t> pr0{0};
            for (size_t k = 0; k + 1 < ct0.size(); k++) {
                const Field& fk = tab[cf[k]];
                const uint64_t gz = fzh::cdiv(fk.nz, 8);
                pr0.push_back(pr0.back() + (uint32_t)((uint64_t)((fk.gx + 1) / 2) * fk.gy * gz));
            }
            const uint32_t P = pr0.back();
            const size_t pb = std::max<size_t>((P + 63) / 64, 1);
            std::vector<uint32_t> pf(pb, 0);
            for (size_t k = 0, j = 0; k < pb && P; k++) {
                while (j + 2 < pr0.size() && pr0[j + 1] <= 64 * k) j++;
                pf[k] = (uint32_t)j;
            }
            uint32_t* d_pr0 = (uint32_t*)impl_->dev(4 * pr0.size());
            uint32_t* d_pf = (uint32_t*)impl_->dev(4 * pb);
            hipck(hipMemcpy(d_pr0, pr0.data(), 4 * pr0.size(), hipMemcpyHostToDevice), "hipMemcpy (tile pairs)");
            hipck(hipMemcpy(d_pf, pf.data(), 4 * pb, hipMemcpyHostToDevice), "hipMemcpy (pair fields)");
            p.pair0 = d_pr0;
            p.pfield = d_pf;
            p.P = P;
        }
    }
    std::vector<uint32_t> cfield(std::max<uint64_t>(NC, 1), 0);
    for (uint32_t f = 0; f < F; f++)
        for (uint32_t c = chunk0[f]; c < chunk0[f + 1]; c++) cfield[c] = f;
    uint32_t* d_cf = (uint32_t*)impl_->dev(4 * cfield.size());
    hipck(hipMemcpy(d_cf, cfield.data(), 4 * cfield.size(), hipMemcpyHostToDevice), "hipMemcpy (chunk fields)");
    p.cfield = d_cf;
    std::vector<uint32_t> ufield(std::max<uint64_t>(U, 1), 0);
    for (uint32_t f = 0; f < F; f++)
        for (uint32_t u = unit0[f]; u < unit0[f + 1]; u++) ufield[u] = f;
    uint32_t* d_uf = (uint32_t*)impl_->dev(4 * ufield.size());
    hipck(hipMemcpy(d_uf, ufield.data(), 4 * ufield.size(), hipMemcpyHostToDevice), "hipMemcpy (unit fields)");
    p.ufield = d_uf;
    p.fields = d_tab;
    p.tile0 = d_t0;
    p.chunk0 = d_c0;
    p.unit0 = d_u0;
    p.params = (float*)impl_->dev(16 * F);
    p.vb_part = (float2*)impl_->dev(8 * std::max<uint64_t>(U, 1));
    p.codes = (uint16_t*)impl_->dev(2 * codes);
    p.lerr = (float*)impl_->dev(4ull * kListSlots * T);
    p.lidx = (uint32_t*)impl_->dev(4ull * kListSlots * T);
    if (strict) {
        p.lpval = (float*)impl_->dev(4ull * kListSlots * T);
        p.lpidx = (uint32_t*)impl_->dev(4ull * kListSlots * T);
    }
    p.n_counts = (uint32_t)n_counts;
    p.counts = (uint32_t*)impl_->dev(4 * n_counts);
    p.excl = (uint32_t*)impl_->dev(4 * (n_counts + 1));
    p.sizes = (uint32_t*)impl_->dev(4 * NC);
    p.chunk_scratch = (uint8_t*)impl_->dev((size_t)kChunk * NC);
    p.derived = (Derived*)impl_->dev(sizeof(Derived) * F);
    p.results = (Result*)impl_->dev(sizeof(Result) * F);
    p.scan_ws = impl_->dev(scan_ws_bytes_for(n_counts));
    d_out_ = (uint8_t*)impl_->dev(out);
    p.out = d_out_;
    out_cap_ = out;
    void* mapped = nullptr;
    hipck(hipHostMalloc(&mapped, sizeof(Result) * F, hipHostMallocMapped | hipHostMallocCoherent), "hipHostMalloc");
    impl_->h_results = (Result*)mapped;
    std::memset(mapped, 0, sizeof(Result) * F);
    hipck(hipHostGetDevicePointer(&impl_->d_results_map, mapped, 0), "hipHostGetDevicePointer");
    impl_->d_plan = (Plan*)impl_->dev(sizeof(Plan));
    hipck(hipMemcpy(impl_->d_plan, &p, sizeof(Plan), hipMemcpyHostToDevice), "hipMemcpy (plan)");
    for (auto& e : impl_->ev) hipck(hipEventCreate(&e), "hipEventCreate");
}

void BatchCompressor::setLaunch(const BatchLaunch& l) {
    auto fix = [](uint32_t v, uint32_t d) { return v ? v : d; };
    BatchLaunch d;
    launch_.vb_wg_per_cu = fix(l.vb_wg_per_cu, d.vb_wg_per_cu);
    launch_.lz_wg_per_cu = fix(l.lz_wg_per_cu, d.lz_wg_per_cu);
    launch_.lz_tiles_per_unit = fix(l.lz_tiles_per_unit, d.lz_tiles_per_unit);
    launch_.pack_wg_per_cu = fix(l.pack_wg_per_cu, d.pack_wg_per_cu);
    launch_.scatter_wg_per_cu = fix(l.scatter_wg_per_cu, d.scatter_wg_per_cu);
    destroyGraph();
}

void BatchCompressor::enableGraphMode(bool on) {
    graph_mode_ = on;
    if (!on) destroyGraph();
}

void BatchCompressor::enqueue(hipStream_t s) {
    using namespace fzh::batch;
    const Plan& p = impl_->plan;
    const Plan* dp = impl_->d_plan;
    const uint32_t cu = (uint32_t)cus_;
    const bool prof = profiling_;
    hipEvent_t* ev = impl_->ev;
    if (prof) hipck(hipEventRecord(ev[0], s), "hipEventRecord");
    fzck(launch_value_base(s, p, dp, cu * launch_.vb_wg_per_cu), "batch value base");
    if (prof) hipck(hipEventRecord(ev[1], s), "hipEventRecord");
    fzck(launch_lorenzo(s, p, dp, cu * launch_.lz_wg_per_cu, launch_.lz_tiles_per_unit), "batch lorenzo");
    if (prof) hipck(hipEventRecord(ev[2], s), "hipEventRecord");
    fzck(launch_rze(s, p, dp), "batch rze");
    if (prof) hipck(hipEventRecord(ev[3], s), "hipEventRecord");
    fzck(launch_scan(s, p), "batch scan");
    if (prof) hipck(hipEventRecord(ev[4], s), "hipEventRecord");
    fzck(launch_finalize(s, p, dp), "batch finalize");
    if (prof) hipck(hipEventRecord(ev[5], s), "hipEventRecord");
    fzck(launch_pack(s, p, dp, cu * launch_.pack_wg_per_cu), "batch pack");
    if (prof) hipck(hipEventRecord(ev[6], s), "hipEventRecord");
    fzck(launch_scatter(s, p, dp, cu * launch_.scatter_wg_per_cu), "batch scatter");
    if (prof) hipck(hipEventRecord(ev[7], s), "hipEventRecord");
    fzck(launch_readback(s, p, impl_->d_results_map), "batch readback");
}

void BatchCompressor::compress(hipStream_t stream) {
    if (!impl_) throw std::runtime_error("BatchCompressor: setFields() first");
    const bool use_graph = graph_mode_ && !profiling_;
    if (use_graph && graph_exec_) {
        hipck(hipGraphLaunch(graph_exec_, stream), "hipGraphLaunch");
    } else if (use_graph && calls_since_plan_ >= 1) {
        if (!graph_stream_) hipck(hipStreamCreateWithFlags(&graph_stream_, hipStreamNonBlocking), "hipStreamCreate");
        hipck(hipStreamBeginCapture(graph_stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            enqueue(graph_stream_);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(graph_stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hipck(hipStreamEndCapture(graph_stream_, &graph_), "hipStreamEndCapture");
        hipck(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
        hipck(hipGraphLaunch(graph_exec_, stream), "hipGraphLaunch");
    } else {
        enqueue(stream);
    }
    calls_since_plan_++;
    hipck(hipStreamSynchronize(stream), "hipStreamSynchronize (batch)");
    if (profiling_ && !use_graph) {
        static const char* names[7] = {"value_base", "lorenzo", "bitshuffle_rze", "scan", "finalize", "pack", "scatter"};
        stage_ms_.clear();
        for (int i = 0; i < 7; i++) {
            float ms = 0.0f;
            hipck(hipEventElapsedTime(&ms, impl_->ev[i], impl_->ev[i + 1]), "hipEventElapsedTime");
            stage_ms_.emplace_back(names[i], ms);
        }
    }
    using namespace fzh::batch;
    const Result* r = impl_->h_results;
    size_t over = 0;
    for (size_t i = 0; i < fields_.size(); i++) {
        const uint32_t fl = r[i].flags;
        if (fl & kOverflowCapacity)
            throw std::runtime_error("BatchCompressor: field " + std::to_string(i) + " exceeds its output capacity");
        if (cfg_.strict && (fl & kOverflowPatches))
            throw std::runtime_error("BatchCompressor: field " + std::to_string(i) + ": strict error bound needs " +
                                     std::to_string(r[i].kp) + " patched values; raise patch_capacity");
        if (cfg_.strict && (fl & kOverflowOutliers))
            throw std::runtime_error("BatchCompressor: field " + std::to_string(i) +
                                     ": strict error bound cannot hold with " + std::to_string(r[i].k) +
                                     " outliers; raise outlier_capacity");
        if (fl & kOverflowOutliers) over++;
    }
    if (over && cfg_.outlier_capacity != 0.0f)
        std::fprintf(stderr, "[fz] batch: outlier overflow in %zu field(s); the excess was dropped\n", over);
}

uint64_t BatchCompressor::fieldOffset(size_t i) const {
    if (i >= out_off_.size()) throw std::out_of_range("BatchCompressor: field index");
    return out_off_[i];
}

uint64_t BatchCompressor::fieldBytes(size_t i) const {
    if (!impl_ || i >= fields_.size()) throw std::out_of_range("BatchCompressor: field index");
    return impl_->h_results[i].bytes;
}

uint64_t BatchCompressor::totalOutputBytes() const {
    uint64_t t = 0;
    if (impl_)
        for (size_t i = 0; i < fields_.size(); i++) t += impl_->h_results[i].bytes;
    return t;
}

BatchLaunch BatchCompressor::autotune(hipStream_t s, int reps) {
    using namespace fzh::batch;
    if (!impl_) throw std::runtime_error("BatchCompressor: setFields() first");
    reps = std::max(reps, 1);
    const bool gm = graph_mode_, pr = profiling_;
    graph_mode_ = false;
    profiling_ = false;
    destroyGraph();
    compress(s);  // every buffer the kernels read now holds this batch's data
    const Plan& p = impl_->plan;
    const Plan* dp = impl_->d_plan;
    const uint32_t cu = (uint32_t)cus_;
    hipEvent_t e0 = impl_->ev[0], e1 = impl_->ev[1];
    auto time = [&](auto&& launch) {
        launch();  // warm
        hipck(hipEventRecord(e0, s), "hipEventRecord");
        for (int r = 0; r < reps; r++) launch();
        hipck(hipEventRecord(e1, s), "hipEventRecord");
        hipck(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0f;
        hipck(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return ms / reps;
    };
    BatchLaunch best = launch_;
    float bt = std::numeric_limits<float>::max();
    for (uint32_t w : {1u, 2u, 4u, 8u}) {
        const float t = time([&] { fzck(launch_value_base(s, p, dp, cu * w), "tune value base"); });
        if (t < bt) { bt = t; best.vb_wg_per_cu = w; }
    }
    bt = std::numeric_limits<float>::max();
    // 1 << 20 workgroups per CU exceeds every grid: one unit per wave (the
    // launch clamps the grid to the units), i.e. not persistent at all
    for (uint32_t w : {2u, 4u, 8u, 16u, 1u << 20})
        for (uint32_t g : {1u, 4u, 16u}) {
            const float t = time([&] { fzck(launch_lorenzo(s, p, dp, cu * w, g), "tune lorenzo"); });
            if (t < bt) { bt = t; best.lz_wg_per_cu = w; best.lz_tiles_per_unit = g; }
        }
    bt = std::numeric_limits<float>::max();
    for (uint32_t w : {1u, 2u, 4u, 8u}) {
        const float t = time([&] { fzck(launch_pack(s, p, dp, cu * w), "tune pack"); });
        if (t < bt) { bt = t; best.pack_wg_per_cu = w; }
    }
    bt = std::numeric_limits<float>::max();
    for (uint32_t w : {1u, 2u, 4u, 8u}) {
        const float t = time([&] { fzck(launch_scatter(s, p, dp, cu * w), "tune scatter"); });
        if (t < bt) { bt = t; best.scatter_wg_per_cu = w; }
    }
    hipck(hipStreamSynchronize(s), "hipStreamSynchronize (tune)");
    graph_mode_ = gm;
    profiling_ = pr;
    setLaunch(best);
    return best;
}

}  // namespace fz

// ===========================================================================
// C ABI
// ===========================================================================
struct fzb_batch {
    std::unique_ptr<fz::BatchCompressor> b;
};

namespace {

template <typename F>
int guard(F&& f) {
    try {
        f();
        return FZH_OK;
    } catch (const std::invalid_argument& e) {
        return fzh::fail(FZH_E_INVALID, e.what());
    } catch (const std::out_of_range& e) {
        return fzh::fail(FZH_E_INVALID, e.what());
    } catch (const std::exception& e) {
        return fzh::fail(FZH_E_RUNTIME, e.what());
    } catch (...) {
        return fzh::fail(FZH_E_RUNTIME, "unknown C++ exception");
    }
}

void need(const void* p) {
    if (!p) throw std::invalid_argument("null argument");
}

}  // namespace

extern "C" {

int fzb_create(const fzb_config* cfg, fzb_batch** out) {
    return guard([&] {
        need(cfg);
        need(out);
        if (cfg->eb_mode < 0 || cfg->eb_mode > 2) throw std::invalid_argument("eb_mode must be 0, 1 or 2");
        fz::BatchConfig c;
        c.error_bound = cfg->error_bound;
        c.eb_mode = (fz::ErrorBoundMode)cfg->eb_mode;
        c.quant_radius = cfg->quant_radius;
        c.outlier_capacity = cfg->outlier_capacity;
        c.zigzag = cfg->zigzag != 0;
        c.strict = cfg->strict != 0;
        c.patch_capacity = cfg->patch_capacity;
        auto h = std::make_unique<fzb_batch>();
        h->b = std::make_unique<fz::BatchCompressor>(c);
        *out = h.release();
    });
}

int fzb_destroy(fzb_batch* b) {
    return guard([&] { delete b; });
}

int fzb_set_fields(fzb_batch* b, const fzb_field* fields, size_t n) {
    return guard([&] {
        need(b);
        need(fields);
        std::vector<fz::BatchField> v(n);
        for (size_t i = 0; i < n; i++) {
            v[i].d_in = fields[i].d_in;
            v[i].nx = (size_t)fields[i].nx;
            v[i].ny = (size_t)fields[i].ny;
            v[i].nz = (size_t)fields[i].nz;
        }
        b->b->setFields(v);
    });
}

int fzb_compress(fzb_batch* b, void* stream) {
    return guard([&] {
        need(b);
        b->b->compress(fzh::S(stream));
    });
}

int fzb_output(fzb_batch* b, void** d_base, uint64_t* capacity) {
    return guard([&] {
        need(b);
        if (d_base) *d_base = (void*)b->b->output();
        if (capacity) *capacity = b->b->outputCapacity();
    });
}

int fzb_field_result(fzb_batch* b, size_t i, uint64_t* offset, uint64_t* bytes) {
    return guard([&] {
        need(b);
        if (offset) *offset = b->b->fieldOffset(i);
        if (bytes) *bytes = b->b->fieldBytes(i);
    });
}

int fzb_totals(fzb_batch* b, uint64_t* in_bytes, uint64_t* out_bytes) {
    return guard([&] {
        need(b);
        if (in_bytes) *in_bytes = b->b->totalInputBytes();
        if (out_bytes) *out_bytes = b->b->totalOutputBytes();
    });
}

static fz::BatchLaunch from_c(const fzb_launch& l) {
    fz::BatchLaunch r;
    r.vb_wg_per_cu = l.vb_wg_per_cu;
    r.lz_wg_per_cu = l.lz_wg_per_cu;
    r.lz_tiles_per_unit = l.lz_tiles_per_unit;
    r.pack_wg_per_cu = l.pack_wg_per_cu;
    r.scatter_wg_per_cu = l.scatter_wg_per_cu;
    return r;
}

static void to_c(const fz::BatchLaunch& l, fzb_launch* o) {
    o->vb_wg_per_cu = l.vb_wg_per_cu;
    o->lz_wg_per_cu = l.lz_wg_per_cu;
    o->lz_tiles_per_unit = l.lz_tiles_per_unit;
    o->pack_wg_per_cu = l.pack_wg_per_cu;
    o->scatter_wg_per_cu = l.scatter_wg_per_cu;
}

int fzb_set_launch(fzb_batch* b, const fzb_launch* l) {
    return guard([&] {
        need(b);
        need(l);
        b->b->setLaunch(from_c(*l));
    });
}

int fzb_get_launch(fzb_batch* b, fzb_launch* l) {
    return guard([&] {
        need(b);
        need(l);
        to_c(b->b->launch(), l);
    });
}

int fzb_autotune(fzb_batch* b, void* stream, int reps, fzb_launch* best) {
    return guard([&] {
        need(b);
        const fz::BatchLaunch r = b->b->autotune(fzh::S(stream), reps);
        if (best) to_c(r, best);
    });
}

int fzb_enable_graph_mode(fzb_batch* b, int on) {
    return guard([&] {
        need(b);
        b->b->enableGraphMode(on != 0);
    });
}

int fzb_enable_profiling(fzb_batch* b, int on) {
    return guard([&] {
        need(b);
        b->b->enableProfiling(on != 0);
    });
}

int fzb_stage_ms(fzb_batch* b, int i, char* name, size_t name_cap, float* ms) {
    return guard([&] {
        need(b);
        const auto& v = b->b->stageTimes();
        if (i < 0 || (size_t)i >= v.size()) throw std::invalid_argument("no such stage time");
        if (name && name_cap) {
            std::strncpy(name, v[i].first.c_str(), name_cap - 1);
            name[name_cap - 1] = 0;
        }
        if (ms) *ms = v[i].second;
    });
}

}  // extern "C"

// host_select.cpp — selecting compressor (include/fzs.h): a field is cut
// into equal chunks, each chunk gets the chain the selector picks, and the
// result is one container that records the choices.  No counterpart in the
// reference.
//
// Container FZS1 (little endian):
//   0   u32 magic 0x31535A46 ("FZS1")   4  u16 version (1)   6 u8 mode  7 u8 ndim
//   8   u64 nx, ny, nz (field)          32 u32 cx, cy, cz (chunk)
//   44  u32 n_chunks                    48 f32 rel_eb, f32 value_base, f32 abs_eb
//   60  u32 CRC32 of bytes [64, end)
//   64  u32 chain_mask (bit k: chain k is used), then for each used chain k
//       in order: u32 n_stages, FZMStageInfo x n_stages (the chain's stage
//       records as fz::Pipeline::buildHeader writes them to FZM)
//   ..  u8 chain[n_chunks], zero-padded to 8
//   ..  u64 bytes[n_chunks]
//   ..  chunk results back to back, each zero-padded to 16: the output of
//       fz::Pipeline::compress of that chunk with its chain
// Decode rebuilds each used chain (same settings as the labeller, fzs.h),
// checks the recorded stage types, restores every stage from its recorded
// header and runs fz::Pipeline::decompress per chunk (Bitpack's width, which
// differs per chunk, follows from the packed stream's size).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "fz/batch.h"
#include "fz/pipeline.h"
#include "fz/stages.h"
#include "fzh_internal.h"
#include "fzs.h"

namespace fz {
namespace {

constexpr uint32_t kMagic = 0x31535A46u;
constexpr uint16_t kVersion = 1;

void hipck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
void fzck(int rc, const char* what) {
    if (rc != FZH_OK) throw std::runtime_error(std::string(what) + ": " + fzh_last_error());
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t b) {
        if (bytes >= b && p) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        hipck(hipMalloc(&p, std::max<size_t>(b, 256)), "hipMalloc (selector)");
        bytes = std::max<size_t>(b, 256);
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// The candidate chains with the labeller's settings (fzs.h); `stages` lists
// the chain's stages in the order fz::Pipeline records them.
struct Chain {
    std::unique_ptr<Pipeline> p;
    std::vector<Stage*> stages;
    BitpackStage<uint16_t>* bitpack = nullptr;
};

Chain make_chain(int k, float abs_eb, uint32_t cx, uint32_t cy, uint32_t cz) {
    const size_t bytes = 4ull * cx * cy * cz;
    Chain c;
    c.p = std::make_unique<Pipeline>(bytes);
    Pipeline* p = c.p.get();
    if (k == 3) {
        auto* q = p->addStage<QuantizerStage<float, uint16_t>>();
        q->setErrorBound(abs_eb);
        q->setErrorBoundMode(ErrorBoundMode::ABS);
        q->setQuantRadius(FZS_Q_RADIUS);
        q->setOutlierCapacity(FZS_OUTLIER_CAP);
        q->setZigzagCodes(true);
        auto* b = p->addStage<BitpackStage<uint16_t>>();
        b->setAutoDetect(true);
        p->connect(b, q, "codes");
        c.stages = {q, b};
        c.bitpack = b;
    } else {
        auto* l = p->addStage<LorenzoQuantStage<float, uint16_t>>();
        l->setErrorBound(abs_eb);
        l->setErrorBoundMode(ErrorBoundMode::ABS);
        l->setQuantRadius((uint16_t)FZS_LZ_RADIUS);
        l->setOutlierCapacity(FZS_OUTLIER_CAP);
        l->setZigzagCodes(true);
        c.stages = {l};
        if (k == 0) {
            auto* r = p->addStage<RLEStage<uint16_t>>();
            p->connect(r, l, "codes");
            c.stages.push_back(r);
        } else if (k == 1) {
            auto* bs = p->addStage<BitshuffleStage>();
            bs->setBlockSize(16384);
            bs->setElementWidth(2);
            auto* rz = p->addStage<RZEStage>();
            rz->setChunkSize(16384);
            rz->setLevels(4);
            p->connect(bs, l, "codes");
            p->connect(rz, bs, "output");
            c.stages.push_back(bs);
            c.stages.push_back(rz);
        } else if (k == 2) {
            auto* h = p->addStage<HuffmanStage<uint16_t>>();
            h->setBklen(FZS_HUF_BKLEN);
            p->connect(h, l, "codes");
            c.stages.push_back(h);
        } else {
            throw std::invalid_argument("chain must be 0..3");
        }
    }
    p->setDims(cx, cy, cz);
    p->finalize();
    return c;
}

template <typename T>
void put(std::vector<uint8_t>& v, size_t off, const T& x) {
    std::memcpy(v.data() + off, &x, sizeof(T));
}
template <typename T>
T get(const uint8_t* p, size_t off) {
    T x;
    std::memcpy(&x, p + off, sizeof(T));
    return x;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

class ChainSelectCompressor {
  public:
    ChainSelectCompressor(uint32_t cx, uint32_t cy, uint32_t cz) : cx_(cx), cy_(cy), cz_(cz) {
        const uint64_t E = (uint64_t)cx * cy * cz;
        if (E == 0 || (4 * E) % 16384 != 0 || E > (1ull << 24))
            throw std::invalid_argument("selector: chunk bytes must be a nonzero multiple of 16384, at most 2^24 values");
    }

    void setModel(const float* params, const float* norm) {
        const size_t P = fzs_mlp_param_count();
        params_.ensure(4 * P);
        norm_.ensure(4 * 2 * FZS_NFEAT);
        hipck(hipMemcpy(params_.p, params, 4 * P, hipMemcpyHostToDevice), "hipMemcpy (model)");
        hipck(hipMemcpy(norm_.p, norm, 4 * 2 * FZS_NFEAT, hipMemcpyHostToDevice), "hipMemcpy (norm)");
        has_model_ = true;
    }

    const std::vector<uint8_t>& compress(hipStream_t s, const float* d_field, uint64_t nx, uint64_t ny, uint64_t nz,
                                         float rel_eb, int mode, int fixed) {
        const double t0 = now_ms();
        if (!d_field) throw std::invalid_argument("selector: null field");
        if (!(rel_eb > 0.0f)) throw std::invalid_argument("selector: rel_eb must be positive");
        if (nx % cx_ || ny % cy_ || nz % cz_ || nx * ny * nz == 0)
            throw std::invalid_argument("selector: the field must be a whole number of chunks");
        if (mode == FZS_MODE_LEARNED && !has_model_) throw std::runtime_error("selector: no model set");
        if (mode == FZS_MODE_FIXED && (fixed < 0 || fixed >= FZS_NCHAINS))
            throw std::invalid_argument("selector: fixed chain must be 0..3");
        if (mode < 0 || mode > 2) throw std::invalid_argument("selector: bad mode");
        const uint64_t E = (uint64_t)cx_ * cy_ * cz_;
        const uint64_t n = nx * ny * nz / E;
        // REL bound of the whole field (LorenzoQuantStage REL: max |x|)
        vb_ws_.ensure(256 + fzh_value_base_ws_bytes(nx * ny * nz));
        fzck(fzh_workspace_init(s, vb_ws_.p, vb_ws_.bytes), "value base workspace");
        fzck(fzh_value_base_f32(s, d_field, nx * ny * nz, 1, rel_eb, (float*)vb_ws_.p, (uint8_t*)vb_ws_.p + 256,
                                vb_ws_.bytes - 256),
             "value base");
        float prm[4];
        fzck(fzh::read_small(s, vb_ws_.p, 16, prm), "value base readback");
        const float abs_eb = prm[1];
        chunks_.ensure(4 * n * E);
        fzck(fzs_gather_chunks(s, d_field, nx, ny, nz, cx_, cy_, cz_, (float*)chunks_.p), "gather chunks");
        choice_.assign(n, (uint8_t)std::max(fixed, 0));
        if (mode != FZS_MODE_FIXED) {
            eb_.ensure(4 * n);
            std::vector<float> ebv(n, abs_eb);
            hipck(hipMemcpyAsync(eb_.p, ebv.data(), 4 * n, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
            if (mode == FZS_MODE_LEARNED) {
                feat_.ensure(4 * FZS_NFEAT * n);
                logits_.ensure(4 * FZS_NCHAINS * n);
                dchoice_.ensure(n);
                fzck(fzs_features_f32(s, (const float*)chunks_.p, n, cx_, cy_, cz_, (const float*)eb_.p,
                                      (float*)feat_.p),
                     "features");
                fzck(fzs_mlp_forward(s, (const float*)params_.p, (const float*)norm_.p, (const float*)feat_.p, n,
                                     (float*)logits_.p),
                     "mlp forward");
                fzck(fzs_select(s, (const float*)logits_.p, n, (uint8_t*)dchoice_.p), "select");
                fzck(fzh::read_small(s, dchoice_.p, n, choice_.data()), "choice readback");
            } else {
                label_ws_.ensure(fzs_label_ws_bytes(n, cx_, cy_, cz_));
                sizes_.assign(FZS_NCHAINS * n, 0);
                fzck(fzs_label_sizes(s, (const float*)chunks_.p, n, cx_, cy_, cz_, (const float*)eb_.p,
                                     sizes_.data(), label_ws_.p, label_ws_.bytes, 0),
                     "labeller");
                for (uint64_t c = 0; c < n; c++) {
                    const uint64_t* r = sizes_.data() + FZS_NCHAINS * c;
                    choice_[c] = (uint8_t)(std::min_element(r, r + FZS_NCHAINS) - r);
                }
            }
        }
        const double t1 = now_ms();
        // compress every chunk with its chain: kWorkers host threads, each
        // with its own stream and its own pipeline per chain, take the chunks
        // in turn (each fz::Pipeline::compress ends with a host wait, so one
        // thread would leave the GPU idle between chunks)
        uint32_t mask = 0;
        for (uint8_t k : choice_) mask |= 1u << k;
        hipck(hipStreamSynchronize(s), "hipStreamSynchronize");  // chunks_ and the choices are ready
        const int W = (int)std::max<uint64_t>(1, std::min<uint64_t>(kWorkers, n));
        while ((int)workers_.size() < W) {
            workers_.push_back(std::make_unique<Worker>());
            hipck(hipStreamCreate(&workers_.back()->s), "hipStreamCreate");
        }
        for (int w = 0; w < W; w++)
            for (int k = 0; k < FZS_NCHAINS; k++)
                if ((mask >> k) & 1u) {
                    Worker& wk = *workers_[w];
                    if (!wk.chain[k].p || wk.chain_eb[k] != abs_eb) {
                        wk.chain[k] = make_chain(k, abs_eb, cx_, cy_, cz_);
                        wk.chain_eb[k] = abs_eb;
                    }
                }
        // the chunks on chain 1 (LorenzoQuant -> Bitshuffle -> RZE) go through
        // fz::BatchCompressor, one launch per stage for all of them, whose
        // per-field bytes equal fz::Pipeline::compress of each chunk; the
        // last such chunk still runs through its pipeline on a worker, so the
        // container's stage records come from a real run and its bytes check
        // the batch's
        std::vector<uint64_t> c1;
        for (uint64_t c = 0; c < n; c++)
            if (choice_[c] == 1) c1.push_back(c);
        const bool batch1 = c1.size() >= kBatchMin;
        const uint64_t c1_last = batch1 ? c1.back() : 0;
        std::vector<std::vector<uint8_t>> blobs(n);
        std::vector<uint8_t> batch_last;
        std::vector<std::exception_ptr> errs(W);
        std::vector<std::thread> th;
        for (int w = 0; w < W; w++)
            th.emplace_back([&, w] {
                try {
                    Worker& wk = *workers_[w];
                    for (uint64_t c = (uint64_t)w; c < n; c += (uint64_t)W) {
                        if (batch1 && choice_[c] == 1 && c != c1_last) continue;
                        void* d_out = nullptr;
                        size_t bytes = 0;
                        wk.chain[choice_[c]].p->compress((const float*)chunks_.p + c * E, 4 * E, &d_out, &bytes, wk.s);
                        blobs[c].resize(bytes);
                        hipck(hipMemcpyAsync(blobs[c].data(), d_out, bytes, hipMemcpyDeviceToHost, wk.s),
                              "hipMemcpyAsync");
                        hipck(hipStreamSynchronize(wk.s), "hipStreamSynchronize");
                    }
                } catch (...) {
                    errs[w] = std::current_exception();
                }
            });
        std::exception_ptr batch_err;
        if (batch1) {
            try {
                run_batch1(s, abs_eb, c1, blobs, batch_last);
            } catch (...) {
                batch_err = std::current_exception();
            }
        }
        for (auto& t : th) t.join();
        for (auto& e : errs)
            if (e) std::rethrow_exception(e);
        if (batch_err) std::rethrow_exception(batch_err);
        if (batch1 && batch_last != blobs[c1_last])
            throw std::runtime_error("selector: batched chain-1 bytes differ from the pipeline's");
        const double t2 = now_ms();
        // container: chain k's stages are described by the pipeline that
        // compressed the last chunk choosing k (chunk c went to worker c % W)
        std::vector<uint64_t> last(FZS_NCHAINS, 0);
        for (uint64_t c = 0; c < n; c++) last[choice_[c]] = c;
        std::vector<std::vector<FZMStageInfo>> recs(FZS_NCHAINS);
        size_t hdr = 64 + 4;
        for (int k = 0; k < FZS_NCHAINS; k++)
            if ((mask >> k) & 1u) {
                recs[k] = workers_[last[k] % (uint64_t)W]->chain[k].p->buildHeader().stages;
                hdr += 4 + recs[k].size() * sizeof(FZMStageInfo);
            }
        const size_t o_choice = hdr;
        const size_t o_bytes = o_choice + ((n + 7) & ~7ull);
        size_t total = o_bytes + 8 * n;
        std::vector<size_t> o_blob(n);
        for (uint64_t c = 0; c < n; c++) {
            o_blob[c] = total;
            total += (blobs[c].size() + 15) & ~15ull;
        }
        out_.assign(total, 0);
        put<uint32_t>(out_, 0, kMagic);
        put<uint16_t>(out_, 4, kVersion);
        put<uint8_t>(out_, 6, (uint8_t)mode);
        put<uint8_t>(out_, 7, (uint8_t)(nz > 1 ? 3 : (ny > 1 ? 2 : 1)));
        put<uint64_t>(out_, 8, nx);
        put<uint64_t>(out_, 16, ny);
        put<uint64_t>(out_, 24, nz);
        put<uint32_t>(out_, 32, cx_);
        put<uint32_t>(out_, 36, cy_);
        put<uint32_t>(out_, 40, cz_);
        put<uint32_t>(out_, 44, (uint32_t)n);
        put<float>(out_, 48, rel_eb);
        put<float>(out_, 52, prm[0]);
        put<float>(out_, 56, abs_eb);
        size_t o = 64;
        put<uint32_t>(out_, o, mask);
        o += 4;
        for (int k = 0; k < FZS_NCHAINS; k++)
            if ((mask >> k) & 1u) {
                put<uint32_t>(out_, o, (uint32_t)recs[k].size());
                o += 4;
                for (const auto& r : recs[k]) {
                    put<FZMStageInfo>(out_, o, r);
                    o += sizeof(FZMStageInfo);
                }
            }
        std::memcpy(out_.data() + o_choice, choice_.data(), n);
        for (uint64_t c = 0; c < n; c++) {
            put<uint64_t>(out_, o_bytes + 8 * c, (uint64_t)blobs[c].size());
            std::memcpy(out_.data() + o_blob[c], blobs[c].data(), blobs[c].size());
        }
        put<uint32_t>(out_, 60, crc32(out_.data() + 64, out_.size() - 64));
        const double t3 = now_ms();
        select_ms_ = (float)(t1 - t0);
        compress_ms_ = (float)(t2 - t1);
        total_ms_ = (float)(t3 - t0);
        return out_;
    }

    void decompress(hipStream_t s, const uint8_t* in, size_t bytes, float* d_out, size_t out_elems) {
        if (!in || bytes < 68) throw std::invalid_argument("selector: container too short");
        if (get<uint32_t>(in, 0) != kMagic) throw std::runtime_error("selector: not an FZS1 container");
        if (get<uint16_t>(in, 4) != kVersion) throw std::runtime_error("selector: unsupported container version");
        if (get<uint32_t>(in, 60) != crc32(in + 64, bytes - 64)) throw std::runtime_error("selector: checksum mismatch");
        const uint64_t nx = get<uint64_t>(in, 8), ny = get<uint64_t>(in, 16), nz = get<uint64_t>(in, 24);
        const uint32_t cx = get<uint32_t>(in, 32), cy = get<uint32_t>(in, 36), cz = get<uint32_t>(in, 40);
        const uint32_t n = get<uint32_t>(in, 44);
        const float abs_eb = get<float>(in, 56);
        if (cx != cx_ || cy != cy_ || cz != cz_) throw std::runtime_error("selector: chunk shape differs from the handle's");
        const uint64_t E = (uint64_t)cx * cy * cz;
        if (nx % cx || ny % cy || nz % cz || nx * ny * nz / E != n) throw std::runtime_error("selector: bad geometry");
        if (out_elems < nx * ny * nz) throw std::invalid_argument("selector: output too small");
        size_t o = 64;
        const uint32_t mask = get<uint32_t>(in, o);
        o += 4;
        // each recorded chain: its stages get the recorded stage headers, as
        // Pipeline::decompressFromFile restores stages from FZM records
        std::vector<Chain> ch(FZS_NCHAINS);
        for (int k = 0; k < FZS_NCHAINS; k++)
            if ((mask >> k) & 1u) {
                if (o + 4 > bytes) throw std::runtime_error("selector: truncated stage records");
                const uint32_t ns = get<uint32_t>(in, o);
                o += 4;
                if (o + (size_t)ns * sizeof(FZMStageInfo) > bytes) throw std::runtime_error("selector: truncated stage records");
                ch[k] = make_chain(k, abs_eb, cx, cy, cz);
                if (ch[k].stages.size() != ns)
                    throw std::runtime_error("selector: recorded chain differs from chain " + std::to_string(k));
                for (uint32_t j = 0; j < ns; j++) {
                    const FZMStageInfo r = get<FZMStageInfo>(in, o + j * sizeof(FZMStageInfo));
                    if ((uint16_t)r.stage_type != ch[k].stages[j]->getStageTypeId() || r.config_size > FZM_STAGE_CONFIG_SIZE)
                        throw std::runtime_error("selector: recorded stage differs from chain " + std::to_string(k));
                    ch[k].stages[j]->deserializeHeader(r.stage_config, r.config_size);
                }
                o += (size_t)ns * sizeof(FZMStageInfo);
                ch[k].p->prepareDecompress(4 * E);
            }
        const size_t o_choice = o, o_bytes = o_choice + ((n + 7) & ~7ull);
        if (o_bytes + 8ull * n > bytes) throw std::runtime_error("selector: truncated chunk table");
        chunks_.ensure(4 * E * n);
        blob_.ensure(16);
        size_t ob = o_bytes + 8ull * n;
        for (uint32_t c = 0; c < n; c++) {
            const uint8_t k = in[o_choice + c];
            const uint64_t nb = get<uint64_t>(in, o_bytes + 8ull * c);
            if (k >= FZS_NCHAINS || !ch[k].p) throw std::runtime_error("selector: chunk uses an unrecorded chain");
            if (ob + nb > bytes || nb < 16) throw std::runtime_error("selector: truncated chunk data");
            if (ch[k].bitpack) {
                // Bitpack's width is per chunk: the packed stream holds E codes
                const uint64_t packed = get<uint64_t>(in, ob + 4);
                const uint64_t nbits = packed * 8 / E;
                if (nbits == 0 || nbits > 16 || (nbits & (nbits - 1)) || (E * nbits + 7) / 8 != packed)
                    throw std::runtime_error("selector: bad Bitpack stream size");
                ch[k].bitpack->setNBits((uint8_t)nbits);
            }
            blob_.ensure(nb + 16);
            hipck(hipMemcpyAsync(blob_.p, in + ob, nb, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
            size_t got = 0;
            ch[k].p->decompress(blob_.p, nb, (float*)chunks_.p + (uint64_t)c * E, 4 * E, &got, s);
            if (got != 4 * E) throw std::runtime_error("selector: chunk decoded to a wrong size");
            ob += (nb + 15) & ~15ull;
        }
        fzck(fzs_scatter_chunks(s, (const float*)chunks_.p, nx, ny, nz, cx, cy, cz, d_out), "scatter chunks");
        hipck(hipStreamSynchronize(s), "hipStreamSynchronize");
    }

    const std::vector<uint8_t>& choices() const { return choice_; }
    void timing(float* a, float* b, float* c) const {
        if (a) *a = select_ms_;
        if (b) *b = compress_ms_;
        if (c) *c = total_ms_;
    }

  private:
    // Compresses the chunks `c1` (all on chain 1) in one batch on stream s:
    // every result but the last into blobs, the last into `last`.  The batch
    // plan is kept while the chunk list and the bound stay the same.
    void run_batch1(hipStream_t s, float abs_eb, const std::vector<uint64_t>& c1,
                    std::vector<std::vector<uint8_t>>& blobs, std::vector<uint8_t>& last) {
        const uint64_t E = (uint64_t)cx_ * cy_ * cz_;
        if (!batch_ || batch_eb_ != abs_eb || batch_idx_ != c1 || batch_base_ != chunks_.p) {
            batch_.reset();
            BatchConfig cfg;
            cfg.error_bound = abs_eb;
            cfg.eb_mode = ErrorBoundMode::ABS;
            cfg.quant_radius = FZS_LZ_RADIUS;
            cfg.outlier_capacity = FZS_OUTLIER_CAP;
            cfg.zigzag = true;
            cfg.strict = false;
            auto b = std::make_unique<BatchCompressor>(cfg);
            std::vector<BatchField> f(c1.size());
            for (size_t i = 0; i < c1.size(); i++) {
                f[i].d_in = (const float*)chunks_.p + c1[i] * E;
                f[i].nx = cx_;
                f[i].ny = cy_;
                f[i].nz = cz_;
            }
            b->setFields(f);
            b->enableGraphMode(false);  // the plan changes with the choices; no capture to amortise
            batch_ = std::move(b);
            batch_eb_ = abs_eb;
            batch_idx_ = c1;
            batch_base_ = chunks_.p;
        }
        batch_->compress(s);
        // the results packed back to back (16-byte steps) by one gather
        // launch, then one copy to the host
        const uint8_t* d = batch_->output();
        const size_t F = c1.size();
        std::vector<fzh_copy_desc> desc(F);
        std::vector<uint64_t> pos(F + 1, 0);
        for (size_t i = 0; i < F; i++) {
            const uint64_t b = batch_->fieldBytes(i);
            pos[i + 1] = pos[i] + ((b + 15) & ~15ull);
        }
        pack_.ensure(pos[F]);
        pack_desc_.ensure(sizeof(fzh_copy_desc) * F);
        for (size_t i = 0; i < F; i++)
            desc[i] = {d + batch_->fieldOffset(i), (uint8_t*)pack_.p + pos[i], batch_->fieldBytes(i)};
        hipck(hipMemcpyAsync(pack_desc_.p, desc.data(), sizeof(fzh_copy_desc) * F, hipMemcpyHostToDevice, s),
              "hipMemcpyAsync (batch descriptors)");
        fzck(fzh_gather(s, (const fzh_copy_desc*)pack_desc_.p, (int)F), "gather (batch)");
        pack_host_.resize(pos[F]);
        hipck(hipMemcpyAsync(pack_host_.data(), pack_.p, pos[F], hipMemcpyDeviceToHost, s), "hipMemcpyAsync (batch)");
        hipck(hipStreamSynchronize(s), "hipStreamSynchronize (batch)");
        for (size_t i = 0; i < F; i++) {
            std::vector<uint8_t>& dst = i + 1 < F ? blobs[c1[i]] : last;
            dst.assign(pack_host_.data() + pos[i], pack_host_.data() + pos[i] + batch_->fieldBytes(i));
        }
    }

    static constexpr int kWorkers = 8;
    static constexpr size_t kBatchMin = 2;  // chain-1 chunks below this run per chunk
    struct Worker {
        hipStream_t s = nullptr;
        Chain chain[FZS_NCHAINS];
        float chain_eb[FZS_NCHAINS] = {0, 0, 0, 0};
        ~Worker() {
            for (auto& c : chain) c.p.reset();
            if (s) (void)hipStreamDestroy(s);
        }
    };
    std::vector<std::unique_ptr<Worker>> workers_;
    uint32_t cx_, cy_, cz_;
    bool has_model_ = false;
    DevBuf params_, norm_, chunks_, eb_, feat_, logits_, dchoice_, label_ws_, vb_ws_, blob_;
    std::vector<uint8_t> choice_, out_;
    std::vector<uint64_t> sizes_;
    float select_ms_ = 0, compress_ms_ = 0, total_ms_ = 0;
    std::unique_ptr<BatchCompressor> batch_;
    DevBuf pack_, pack_desc_;
    std::vector<uint8_t> pack_host_;
    float batch_eb_ = 0;
    std::vector<uint64_t> batch_idx_;
    const void* batch_base_ = nullptr;
};

}  // namespace fz

struct fzs_compressor {
    std::unique_ptr<fz::ChainSelectCompressor> c;
};

namespace {
template <typename F>
int guard(F&& f) {
    try {
        f();
        return FZH_OK;
    } catch (const std::invalid_argument& e) {
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

int fzs_create(uint32_t cx, uint32_t cy, uint32_t cz, fzs_compressor** out) {
    return guard([&] {
        need(out);
        auto h = std::make_unique<fzs_compressor>();
        h->c = std::make_unique<fz::ChainSelectCompressor>(cx, cy, cz);
        *out = h.release();
    });
}

int fzs_destroy(fzs_compressor* c) {
    return guard([&] { delete c; });
}

int fzs_set_model(fzs_compressor* c, const float* h_params, const float* h_norm) {
    return guard([&] {
        need(c);
        need(h_params);
        need(h_norm);
        c->c->setModel(h_params, h_norm);
    });
}

int fzs_compress_field(fzs_compressor* c, void* stream, const float* d_field, uint64_t nx, uint64_t ny, uint64_t nz,
                       float rel_eb, int mode, int fixed_chain, const uint8_t** h_out, size_t* out_bytes) {
    return guard([&] {
        need(c);
        need(h_out);
        need(out_bytes);
        const auto& v = c->c->compress(fzh::S(stream), d_field, nx, ny, nz, rel_eb, mode, fixed_chain);
        *h_out = v.data();
        *out_bytes = v.size();
    });
}

int fzs_decompress_field(fzs_compressor* c, void* stream, const uint8_t* h_in, size_t in_bytes, float* d_out,
                         size_t out_elems) {
    return guard([&] {
        need(c);
        need(d_out);
        c->c->decompress(fzh::S(stream), h_in, in_bytes, d_out, out_elems);
    });
}

int fzs_last_choices(fzs_compressor* c, const uint8_t** h_choices, size_t* n) {
    return guard([&] {
        need(c);
        if (h_choices) *h_choices = c->c->choices().data();
        if (n) *n = c->c->choices().size();
    });
}

int fzs_last_timing(fzs_compressor* c, float* select_ms, float* compress_ms, float* total_ms) {
    return guard([&] {
        need(c);
        c->c->timing(select_ms, compress_ms, total_ms);
    });
}

}  // extern "C"

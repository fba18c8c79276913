// fzp_capi.cpp — extern "C" wrappers around fz::Pipeline (include/fzp.h).
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "fz/pipeline.h"
#include "fz/stages.h"
#include "fzh_internal.h"
#include "fzp.h"

struct fzp_pipeline {
    std::unique_ptr<fz::Pipeline> p;
    std::vector<fz::Stage*> stages;
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

fz::ErrorBoundMode mode_of(int m) {
    if (m < 0 || m > 2) throw std::invalid_argument("eb_mode must be 0 (ABS), 1 (REL) or 2 (NOA)");
    return (fz::ErrorBoundMode)m;
}

fz::Stage* stage_at(fzp_pipeline* p, int id) {
    if (!p || id < 0 || id >= (int)p->stages.size()) throw std::invalid_argument("bad stage id");
    return p->stages[id];
}

int add(fzp_pipeline* p, fz::Stage* s, int* id) {
    p->p->addRawStage(s);
    p->stages.push_back(s);
    if (id) *id = (int)p->stages.size() - 1;
    return FZH_OK;
}

void need(const void* p) {
    if (!p) throw std::invalid_argument("null argument");
}

// The resolved absolute bound of a LorenzoQuant stage of any instantiation.
bool lorenzo_abs_eb(fz::Stage* s, double* v) {
    if (auto* a = dynamic_cast<fz::LorenzoQuantStage<float, uint16_t>*>(s)) *v = a->getComputedErrorBound();
    else if (auto* b = dynamic_cast<fz::LorenzoQuantStage<float, uint8_t>*>(s)) *v = b->getComputedErrorBound();
    else if (auto* c = dynamic_cast<fz::LorenzoQuantStage<double, uint16_t>*>(s)) *v = c->getComputedErrorBound();
    else if (auto* d = dynamic_cast<fz::LorenzoQuantStage<double, uint32_t>*>(s)) *v = d->getComputedErrorBound();
    else return false;
    return true;
}

}  // namespace

static int add_code(fzp_pipeline* p, int eb, bool nb, int* id);

extern "C" {

int fzp_create(size_t hint, float mult, fzp_pipeline** out) {
    return guard([&] {
        need(out);
        auto* h = new fzp_pipeline;
        try {
            h->p = std::make_unique<fz::Pipeline>(hint, fz::MemoryStrategy::MINIMAL, mult > 0 ? mult : 3.0f);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int fzp_destroy(fzp_pipeline* p) {
    return guard([&] { delete p; });
}

int fzp_add_lorenzo(fzp_pipeline* p, float eb, int eb_mode, uint32_t radius, float cap, int zigzag, int code_bytes,
                    float value_base, int* id) {
    return guard([&] {
        need(p);
        fz::Stage* s;
        if (code_bytes == 2) {
            auto* l = new fz::LorenzoQuantStage<float, uint16_t>();
            l->setErrorBound(eb);
            l->setErrorBoundMode(mode_of(eb_mode));
            l->setQuantRadius((uint16_t)radius);
            l->setOutlierCapacity(cap);
            l->setZigzagCodes(zigzag != 0);
            l->setValueBase(value_base);
            s = l;
        } else if (code_bytes == 1) {
            auto* l = new fz::LorenzoQuantStage<float, uint8_t>();
            l->setErrorBound(eb);
            l->setErrorBoundMode(mode_of(eb_mode));
            l->setQuantRadius((uint8_t)radius);
            l->setOutlierCapacity(cap);
            l->setZigzagCodes(zigzag != 0);
            l->setValueBase(value_base);
            s = l;
        } else {
            throw std::invalid_argument("code_bytes must be 1 or 2");
        }
        add(p, s, id);
    });
}

int fzp_add_lorenzo_f64(fzp_pipeline* p, float eb, int eb_mode, uint32_t radius, float cap, int zigzag,
                        int code_bytes, float value_base, int* id) {
    return guard([&] {
        need(p);
        auto setup = [&](auto* l, auto r) {
            l->setErrorBound(eb);
            l->setErrorBoundMode(mode_of(eb_mode));
            l->setQuantRadius(r);
            l->setOutlierCapacity(cap);
            l->setZigzagCodes(zigzag != 0);
            l->setValueBase(value_base);
        };
        fz::Stage* s;
        if (code_bytes == 2) {
            auto* l = new fz::LorenzoQuantStage<double, uint16_t>();
            setup(l, (uint16_t)radius);
            s = l;
        } else if (code_bytes == 4) {
            auto* l = new fz::LorenzoQuantStage<double, uint32_t>();
            setup(l, (uint32_t)radius);
            s = l;
        } else {
            throw std::invalid_argument("code_bytes must be 2 or 4 for FP64 input");
        }
        add(p, s, id);
    });
}

int fzp_set_lorenzo_strict(fzp_pipeline* p, int id, int on, float cap) {
    return guard([&] {
        fz::Stage* s = stage_at(p, id);
        if (auto* a = dynamic_cast<fz::LorenzoQuantStage<float, uint16_t>*>(s)) {
            a->setStrictErrorBound(on != 0);
            if (cap > 0) a->setPatchCapacity(cap);
        } else if (auto* b = dynamic_cast<fz::LorenzoQuantStage<float, uint8_t>*>(s)) {
            b->setStrictErrorBound(on != 0);
            if (cap > 0) b->setPatchCapacity(cap);
        } else if (dynamic_cast<fz::LorenzoQuantStage<double, uint16_t>*>(s) ||
                   dynamic_cast<fz::LorenzoQuantStage<double, uint32_t>*>(s)) {
            throw std::invalid_argument("the strict error bound is available for float input only");
        } else {
            throw std::invalid_argument("stage is not a LorenzoQuant stage");
        }
    });
}

int fzp_add_bitshuffle(fzp_pipeline* p, uint32_t block_size, int ew, int* id) {
    return guard([&] {
        need(p);
        auto* s = new fz::BitshuffleStage();
        s->setBlockSize(block_size);
        s->setElementWidth((size_t)ew);
        add(p, s, id);
    });
}

int fzp_add_rze(fzp_pipeline* p, uint32_t chunk, int levels, int* id) {
    return guard([&] {
        need(p);
        auto* s = new fz::RZEStage();
        s->setChunkSize(chunk);
        s->setLevels(levels);
        add(p, s, id);
    });
}

int fzp_add_rle(fzp_pipeline* p, int eb, int* id) {
    return guard([&] {
        need(p);
        fz::Stage* s;
        switch (eb) {
        case 1: s = new fz::RLEStage<uint8_t>(); break;
        case 2: s = new fz::RLEStage<uint16_t>(); break;
        case 4: s = new fz::RLEStage<uint32_t>(); break;
        default: throw std::invalid_argument("RLE elem_bytes must be 1, 2 or 4");
        }
        add(p, s, id);
    });
}

int fzp_add_huffman(fzp_pipeline* p, int eb, uint32_t bklen, int* id) {
    return guard([&] {
        need(p);
        fz::Stage* s;
        switch (eb) {
        case 1: { auto* h = new fz::HuffmanStage<uint8_t>(); if (bklen) h->setBklen(bklen); s = h; break; }
        case 2: { auto* h = new fz::HuffmanStage<uint16_t>(); if (bklen) h->setBklen(bklen); s = h; break; }
        case 4: { auto* h = new fz::HuffmanStage<uint32_t>(); if (bklen) h->setBklen(bklen); s = h; break; }
        default: throw std::invalid_argument("Huffman elem_bytes must be 1, 2 or 4");
        }
        add(p, s, id);
    });
}

int fzp_add_quantizer(fzp_pipeline* p, float eb, int eb_mode, uint32_t radius, float cap, int zigzag, int code_bytes,
                      float value_base, int* id) {
    return guard([&] {
        need(p);
        auto cfg = [&](auto* q) {
            q->setErrorBound(eb);
            q->setErrorBoundMode(mode_of(eb_mode));
            q->setQuantRadius((int)radius);
            q->setOutlierCapacity(cap);
            q->setZigzagCodes(zigzag != 0);
            q->setValueBase(value_base);
        };
        fz::Stage* s;
        switch (code_bytes) {
        case 1: { auto* q = new fz::QuantizerStage<float, uint8_t>(); cfg(q); s = q; break; }
        case 2: { auto* q = new fz::QuantizerStage<float, uint16_t>(); cfg(q); s = q; break; }
        case 4: { auto* q = new fz::QuantizerStage<float, uint32_t>(); cfg(q); s = q; break; }
        default: throw std::invalid_argument("Quantizer code_bytes must be 1, 2 or 4");
        }
        add(p, s, id);
    });
}

int fzp_set_quantizer_outliers(fzp_pipeline* p, int id, float threshold, int inplace) {
    return guard([&] {
        fz::Stage* s = stage_at(p, id);
        auto set = [&](auto* q) {
            q->setOutlierThreshold(threshold);
            q->setInplaceOutliers(inplace != 0);
        };
        if (auto* a = dynamic_cast<fz::QuantizerStage<float, uint8_t>*>(s)) set(a);
        else if (auto* b = dynamic_cast<fz::QuantizerStage<float, uint16_t>*>(s)) set(b);
        else if (auto* c = dynamic_cast<fz::QuantizerStage<float, uint32_t>*>(s)) set(c);
        else throw std::invalid_argument("stage is not a Quantizer stage");
    });
}

int fzp_add_bitpack(fzp_pipeline* p, int eb, int nbits, int* id) {
    return guard([&] {
        need(p);
        auto cfg = [&](auto* b) {
            if (nbits == 0) b->setAutoDetect(true);
            else b->setNBits((uint8_t)nbits);
        };
        fz::Stage* s;
        switch (eb) {
        case 1: { auto* b = new fz::BitpackStage<uint8_t>(); cfg(b); s = b; break; }
        case 2: { auto* b = new fz::BitpackStage<uint16_t>(); cfg(b); s = b; break; }
        case 4: { auto* b = new fz::BitpackStage<uint32_t>(); cfg(b); s = b; break; }
        default: throw std::invalid_argument("Bitpack elem_bytes must be 1, 2 or 4");
        }
        add(p, s, id);
    });
}

int fzp_add_difference(fzp_pipeline* p, int dtype, int negabinary, size_t chunk_bytes, int* id) {
    return guard([&] {
        need(p);
        fz::Stage* s = nullptr;
        auto mk = [&](auto* d) {
            d->setChunkSize(chunk_bytes);
            s = d;
        };
        if (negabinary) {
            switch (dtype) {
            case 4: mk(new fz::DifferenceStage<int8_t, uint8_t>()); break;
            case 5: mk(new fz::DifferenceStage<int16_t, uint16_t>()); break;
            case 6: mk(new fz::DifferenceStage<int32_t, uint32_t>()); break;
            case 7: mk(new fz::DifferenceStage<int64_t, uint64_t>()); break;
            default: throw std::invalid_argument("Difference with negabinary needs a signed integer type");
            }
        } else {
            switch (dtype) {
            case 0: mk(new fz::DifferenceStage<uint8_t>()); break;
            case 1: mk(new fz::DifferenceStage<uint16_t>()); break;
            case 2: mk(new fz::DifferenceStage<uint32_t>()); break;
            case 6: mk(new fz::DifferenceStage<int32_t>()); break;
            case 7: mk(new fz::DifferenceStage<int64_t>()); break;
            case 8: mk(new fz::DifferenceStage<float>()); break;
            case 9: mk(new fz::DifferenceStage<double>()); break;
            default: throw std::invalid_argument("Difference: unsupported data type");
            }
        }
        add(p, s, id);
    });
}

int fzp_add_zigzag(fzp_pipeline* p, int eb, int* id) { return add_code(p, eb, false, id); }
int fzp_add_negabinary(fzp_pipeline* p, int eb, int* id) { return add_code(p, eb, true, id); }

int fzp_add_lorenzo_int(fzp_pipeline* p, int eb, int* id) {
    return guard([&] {
        need(p);
        fz::Stage* s;
        switch (eb) {
        case 1: s = new fz::LorenzoStage<int8_t>(); break;
        case 2: s = new fz::LorenzoStage<int16_t>(); break;
        case 4: s = new fz::LorenzoStage<int32_t>(); break;
        case 8: s = new fz::LorenzoStage<int64_t>(); break;
        default: throw std::invalid_argument("elem_bytes must be 1, 2, 4 or 8");
        }
        add(p, s, id);
    });
}

int fzp_set_huffman_fine(fzp_pipeline* p, int sid, int fine) {
    return guard([&] {
        need(p);
        fz::Stage* s = stage_at(p, sid);
        const auto m = fine ? fz::HuffmanEncodeMode::Fine : fz::HuffmanEncodeMode::Coarse;
        if (auto* h = dynamic_cast<fz::HuffmanStage<uint8_t>*>(s)) h->setEncodeMode(m);
        else if (auto* h2 = dynamic_cast<fz::HuffmanStage<uint16_t>*>(s)) h2->setEncodeMode(m);
        else if (auto* h4 = dynamic_cast<fz::HuffmanStage<uint32_t>*>(s)) h4->setEncodeMode(m);
        else throw std::invalid_argument("stage is not a Huffman stage");
    });
}

int fzp_set_huffman_device_book(fzp_pipeline* p, int sid, int on) {
    return guard([&] {
        need(p);
        fz::Stage* s = stage_at(p, sid);
        if (auto* h = dynamic_cast<fz::HuffmanStage<uint8_t>*>(s)) h->setDeviceCodebook(on != 0);
        else if (auto* h2 = dynamic_cast<fz::HuffmanStage<uint16_t>*>(s)) h2->setDeviceCodebook(on != 0);
        else if (auto* h4 = dynamic_cast<fz::HuffmanStage<uint32_t>*>(s)) h4->setDeviceCodebook(on != 0);
        else throw std::invalid_argument("stage is not a Huffman stage");
    });
}

int fzp_connect(fzp_pipeline* p, int dep, int prod, const char* name) {
    return guard([&] {
        need(p);
        fz::Stage* d = stage_at(p, dep);
        fz::Stage* s = stage_at(p, prod);
        p->p->connect(d, s, name ? std::string(name) : s->getOutputNames().at(0));
    });
}

int fzp_set_dims(fzp_pipeline* p, size_t x, size_t y, size_t z) {
    return guard([&] {
        need(p);
        p->p->setDims(x, y, z);
    });
}

int fzp_finalize(fzp_pipeline* p) {
    return guard([&] {
        need(p);
        p->p->finalize();
    });
}

int fzp_enable_graph_mode(fzp_pipeline* p, int on) {
    return guard([&] {
        need(p);
        p->p->enableGraphMode(on != 0);
    });
}

int fzp_set_stage_fusion(fzp_pipeline* p, int on) {
    return guard([&] {
        need(p);
        p->p->setStageFusion(on != 0);
    });
}

int fzp_enable_profiling(fzp_pipeline* p, int on) {
    return guard([&] {
        need(p);
        p->p->enableProfiling(on != 0);
    });
}

int fzp_set_pool_managed_decomp_output(fzp_pipeline* p, int on) {
    return guard([&] {
        need(p);
        p->p->setPoolManagedDecompOutput(on != 0);
    });
}

int fzp_max_compressed_size(fzp_pipeline* p, size_t in_bytes, size_t* out) {
    return guard([&] {
        need(p);
        need(out);
        *out = p->p->getMaxCompressedSize(in_bytes);
    });
}

int fzp_compress(fzp_pipeline* p, void* stream, const void* d_in, size_t in_bytes, void** d_out, size_t* out_bytes) {
    return guard([&] {
        need(p);
        p->p->compress(d_in, in_bytes, d_out, out_bytes, (hipStream_t)stream);
    });
}

int fzp_decompress(fzp_pipeline* p, void* stream, const void* d_in, size_t in_bytes, void** d_out,
                   size_t* out_bytes) {
    return guard([&] {
        need(p);
        p->p->decompress(d_in, in_bytes, d_out, out_bytes, (hipStream_t)stream);
    });
}

int fzp_compress_into(fzp_pipeline* p, void* stream, const void* d_in, size_t in_bytes, void* d_out, size_t cap,
                      size_t* out_bytes) {
    return guard([&] {
        need(p);
        p->p->compress(d_in, in_bytes, d_out, cap, out_bytes, (hipStream_t)stream);
    });
}

int fzp_decompress_into(fzp_pipeline* p, void* stream, const void* d_in, size_t in_bytes, void* d_out, size_t cap,
                        size_t* out_bytes) {
    return guard([&] {
        need(p);
        p->p->decompress(d_in, in_bytes, d_out, cap, out_bytes, (hipStream_t)stream);
    });
}

int fzp_write_file(fzp_pipeline* p, const char* path) {
    return guard([&] {
        need(p);
        need(path);
        p->p->writeToFile(path);
    });
}

int fzp_decompress_file(const char* path, void* stream, void** d_out, size_t* out_bytes) {
    return guard([&] {
        need(path);
        fz::Pipeline::decompressFromFile(path, d_out, out_bytes, (hipStream_t)stream);
    });
}

int fzp_last_perf(fzp_pipeline* p, float* dag_ms, float* host_ms, size_t* in_bytes, size_t* out_bytes) {
    return guard([&] {
        need(p);
        const auto& r = p->p->getLastPerfResult();
        if (dag_ms) *dag_ms = r.dag_elapsed_ms;
        if (host_ms) *host_ms = r.host_elapsed_ms;
        if (in_bytes) *in_bytes = r.input_bytes;
        if (out_bytes) *out_bytes = r.output_bytes;
    });
}

int fzp_last_stage_perf(fzp_pipeline* p, int i, char* name, size_t cap, float* ms) {
    return guard([&] {
        need(p);
        const auto& r = p->p->getLastPerfResult();
        if (i < 0 || i >= (int)r.stages.size()) throw std::invalid_argument("stage index out of range");
        if (name && cap) {
            std::strncpy(name, r.stages[i].name.c_str(), cap - 1);
            name[cap - 1] = 0;
        }
        if (ms) *ms = r.stages[i].elapsed_ms;
    });
}

int fzp_num_outputs(fzp_pipeline* p, int* n) {
    return guard([&] {
        need(p);
        need(n);
        *n = (int)p->p->getOutputBuffers().size();
    });
}

int fzp_output_info(fzp_pipeline* p, int i, char* name, size_t cap, size_t* bytes) {
    return guard([&] {
        need(p);
        const auto b = p->p->getOutputBuffers();
        if (i < 0 || i >= (int)b.size()) throw std::invalid_argument("output index out of range");
        if (name && cap) {
            std::strncpy(name, b[i].name.c_str(), cap - 1);
            name[cap - 1] = 0;
        }
        if (bytes) *bytes = b[i].actual_size;
    });
}

int fzp_stage_abs_error_bound(fzp_pipeline* p, int id, float* abs_eb) {
    return guard([&] {
        need(abs_eb);
        double v;
        if (!lorenzo_abs_eb(stage_at(p, id), &v)) throw std::invalid_argument("stage is not a LorenzoQuant stage");
        *abs_eb = (float)v;
    });
}

int fzp_stage_abs_error_bound_f64(fzp_pipeline* p, int id, double* abs_eb) {
    return guard([&] {
        need(abs_eb);
        if (!lorenzo_abs_eb(stage_at(p, id), abs_eb)) throw std::invalid_argument("stage is not a LorenzoQuant stage");
    });
}

}  // extern "C"

static int add_code(fzp_pipeline* p, int eb, bool nb, int* id) {
    return guard([&] {
        need(p);
        fz::Stage* s;
        switch (eb) {
        case 1: s = nb ? (fz::Stage*)new fz::NegabinaryStage<int8_t, uint8_t>() : new fz::ZigzagStage<int8_t, uint8_t>(); break;
        case 2: s = nb ? (fz::Stage*)new fz::NegabinaryStage<int16_t, uint16_t>() : new fz::ZigzagStage<int16_t, uint16_t>(); break;
        case 4: s = nb ? (fz::Stage*)new fz::NegabinaryStage<int32_t, uint32_t>() : new fz::ZigzagStage<int32_t, uint32_t>(); break;
        case 8: s = nb ? (fz::Stage*)new fz::NegabinaryStage<int64_t, uint64_t>() : new fz::ZigzagStage<int64_t, uint64_t>(); break;
        default: throw std::invalid_argument("elem_bytes must be 1, 2, 4 or 8");
        }
        add(p, s, id);
    });
}

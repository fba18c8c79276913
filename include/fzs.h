/* fzs.h — C ABI of the learned chain selector (no counterpart in the
 * reference, which runs the one chain the caller configured).
 *
 * Chunks are equal-shaped float32 blocks (nx x ny x nz, x fastest) stored
 * back to back.  The selector reads FZS_NFEAT cheap statistics per chunk
 * (fzs_features_f32), feeds them to a 32 -> 64 -> 64 -> 4 MLP (ReLU, BF16
 * inputs to every layer, FP32 accumulation and FP32 master weights) and
 * picks one of FZS_NCHAINS candidate chains per chunk:
 *   0  LorenzoQuant -> RLE
 *   1  LorenzoQuant -> Bitshuffle -> RZE
 *   2  LorenzoQuant -> Huffman
 *   3  Quantizer (ABS) -> Bitpack
 * Training runs on the device too: fzs_mlp_train_step does forward +
 * backward over a batch and fzs_mlp_adam applies the update.  The labels
 * come from the exhaustive labeller (fzs_label_*): the exact compressed size
 * of every chunk under every chain.
 *
 * Every call returns 0 or a negative FZH_E_* code (fzh_last_error()).
 */
#ifndef FZS_H
#define FZS_H

#include <stddef.h>
#include <stdint.h>

#include "fzh.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FZS_NFEAT 32
#define FZS_H1 64
#define FZS_H2 64
#define FZS_NCHAINS 4

/* Settings of the candidate chains (the same for every chunk; the Lorenzo
 * dimensionality is the chunk's own):
 *   LorenzoQuant<float,uint16_t>  ABS eb = the chunk's bound, radius
 *                                 FZS_LZ_RADIUS, zigzag, outlier capacity
 *                                 FZS_OUTLIER_CAP
 *   RLEStage<uint16_t>; BitshuffleStage(16384, 2) -> RZEStage(16384, 4);
 *   HuffmanStage<uint16_t> with bklen FZS_HUF_BKLEN (2 x radius, as the
 *   reference CLI sets it);
 *   QuantizerStage ABS eb, radius FZS_Q_RADIUS, zigzag, capacity
 *   FZS_OUTLIER_CAP -> BitpackStage<uint16_t> (nbits from the maximum).   */
#define FZS_LZ_RADIUS 512
#define FZS_HUF_BKLEN 1024
#define FZS_Q_RADIUS 32768
#define FZS_OUTLIER_CAP 0.2f

/* ---- features -------------------------------------------------------------
 * d_feat[c * FZS_NFEAT + k] = feature k of chunk c, from the chunk's values
 * and its absolute error bound d_abs_eb[c] (see the feature list in
 * select.hip).  Deterministic (fixed reduction order).                     */
int fzs_features_f32(void* stream, const float* d_chunks, size_t n_chunks, uint32_t nx, uint32_t ny,
                     uint32_t nz, const float* d_abs_eb, float* d_feat);

/* ---- MLP ------------------------------------------------------------------
 * Parameters (FP32 master copy, rounded to BF16 where a layer reads them):
 *   W1[64][32] b1[64] W2[64][64] b2[64] W3[4][64] b3[4]   (fzs_mlp_param_count)
 * d_norm: per-feature shift[32] then scale[32]; a layer-1 input is
 * bf16((x - shift) * scale).                                               */
size_t fzs_mlp_param_count(void);
/* logits[n][4] (FP32). */
int fzs_mlp_forward(void* stream, const float* d_params, const float* d_norm, const float* d_feat, size_t n,
                    float* d_logits);
/* Forward + backward of a batch.  d_cost[n][4] = relative regret of each
 * chain (size / best size - 1); the loss is the cross entropy against the
 * soft target softmax(-cost / tau).  Writes the mean gradient d_grad[P] and
 * the mean loss *d_loss.  Workspace: fzs_mlp_train_ws_bytes(n).            */
size_t fzs_mlp_train_ws_bytes(size_t n);
int fzs_mlp_train_step(void* stream, const float* d_params, const float* d_norm, const float* d_feat,
                       const float* d_cost, size_t n, float tau, float* d_grad, float* d_loss, void* d_ws,
                       size_t ws_bytes);
/* Adam (decoupled weight decay) on the FP32 master parameters; step >= 1. */
int fzs_mlp_adam(void* stream, float* d_params, const float* d_grad, float* d_m, float* d_v, uint32_t step,
                 float lr, float beta1, float beta2, float eps, float weight_decay);
/* ---- exhaustive labeller --------------------------------------------------
 * h_sizes[c * FZS_NCHAINS + k] = bytes of fz::Pipeline::compress's result
 * for chunk c under chain k (the chains above; the whole multi-output
 * concat).  Chunks hold nx*ny*nz values with 4*nx*ny*nz a multiple of 16384
 * (no chain pads them) and at most 2^24 values.  Synchronous: the host
 * builds every chunk's Huffman codebook between two device passes, on
 * `host_threads` threads (0 = hardware concurrency, at most 16).          */
size_t fzs_label_ws_bytes(size_t n_chunks, uint32_t nx, uint32_t ny, uint32_t nz);
int fzs_label_sizes(void* stream, const float* d_chunks, size_t n_chunks, uint32_t nx, uint32_t ny,
                    uint32_t nz, const float* d_abs_eb, uint64_t* h_sizes, void* d_ws, size_t ws_bytes,
                    int host_threads);

/* d_choice[i] = argmax of logits row i (lowest index on ties). */
int fzs_select(void* stream, const float* d_logits, size_t n, uint8_t* d_choice);

/* Chunk-major copy of an nx*ny*nz field (x fastest) that is a whole number
 * of cx*cy*cz chunks: chunk (bz, by, bx) in row-major order of the chunk
 * grid, and back.                                                          */
int fzs_gather_chunks(void* stream, const float* d_field, uint64_t nx, uint64_t ny, uint64_t nz, uint32_t cx,
                      uint32_t cy, uint32_t cz, float* d_chunks);
int fzs_scatter_chunks(void* stream, const float* d_chunks, uint64_t nx, uint64_t ny, uint64_t nz, uint32_t cx,
                       uint32_t cy, uint32_t cz, float* d_field);

/* Seeded training/evaluation chunks generated on the device (no
 * counterpart in the reference): n_chunks 256 KiB chunks of nd = 2 (256 x
 * 256, 4 per 512 x 512 parent field) or nd = 3 (16 x 64 x 64, 8 per
 * 32 x 128 x 128 parent), parent f of regime f % 5 (0 GRF k^-4: a window at
 * a seeded offset of the base field d_base, bx x by x bz, e.g. from
 * fzh_synth_power_law_f32 with expo -4, wrapping periodically; 1 white
 * noise; 2 80% zeros; 3 piecewise-constant boxes; 4 Poisson(20)),
 * d_eb[c] = 1e-3 * max |parent| (1e-3 for an all-zero parent),
 * d_regime[c] (optional) = the regime.  n_chunks must be a multiple of the
 * chunks per parent.  Deterministic in (nd, seed, base, chunk index).      */
int fzs_synth_chunks(void* stream, int nd, uint64_t n_chunks, uint64_t seed, const float* d_base, uint32_t bx,
                     uint32_t by, uint32_t bz, float* d_chunks, float* d_eb, uint8_t* d_regime);

/* ---- selecting compressor -------------------------------------------------
 * Compresses a field chunk by chunk, each chunk with the chain the selector
 * picks (FZS_MODE_LEARNED: the MLP; FZS_MODE_EXHAUSTIVE: the smallest size
 * from the labeller; FZS_MODE_FIXED: `fixed_chain` everywhere).  The bound
 * is REL: rel_eb * max|field| for every chunk.  The result is one host
 * container (FZS1, layout in host_select.cpp) that records every chunk's
 * chain and the chains' stage records (FZMStageInfo), followed by each
 * chunk's fz::Pipeline result; fzs_decompress_field rebuilds the recorded
 * chains and decodes every chunk exactly as fz::Pipeline::decompress.     */
#define FZS_MODE_LEARNED 0
#define FZS_MODE_EXHAUSTIVE 1
#define FZS_MODE_FIXED 2
typedef struct fzs_compressor fzs_compressor;
int fzs_create(uint32_t cx, uint32_t cy, uint32_t cz, fzs_compressor** out);
int fzs_destroy(fzs_compressor* c);
/* Host arrays: fzs_mlp_param_count() parameters and 2 * FZS_NFEAT norm. */
int fzs_set_model(fzs_compressor* c, const float* h_params, const float* h_norm);
/* *h_out is owned by the handle, valid until its next compress.           */
int fzs_compress_field(fzs_compressor* c, void* stream, const float* d_field, uint64_t nx, uint64_t ny,
                       uint64_t nz, float rel_eb, int mode, int fixed_chain, const uint8_t** h_out,
                       size_t* out_bytes);
int fzs_decompress_field(fzs_compressor* c, void* stream, const uint8_t* h_in, size_t in_bytes, float* d_out,
                         size_t out_elems);
/* Chains of the last compress (one per chunk) and its timing split (ms):
 * choosing the chains, compressing the chunks, the whole call.            */
int fzs_last_choices(fzs_compressor* c, const uint8_t** h_choices, size_t* n);
int fzs_last_timing(fzs_compressor* c, float* select_ms, float* compress_ms, float* total_ms);

#ifdef __cplusplus
}
#endif

#endif /* FZS_H */

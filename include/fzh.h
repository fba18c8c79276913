/*
 * fzh.h — C ABI of the MI355X (gfx950) stage kernels.
 *
 * This is the drop-in boundary underneath the fz::Stage classes.  Every entry
 * point takes plain device pointers and byte/element counts, enqueues work on
 * the given HIP stream (passed as void*; NULL = the null stream), allocates no
 * device memory, throws nothing, and returns 0 on success or a negative
 * FZH_E* code (message in fzh_last_error(), thread-local).
 *
 * Workspaces: entry points that need scratch take (d_ws, ws_bytes).  Query the
 * size with the matching *_ws_bytes() function.  A workspace must be zeroed
 * with fzh_workspace_init() before its first use and again whenever the
 * problem size changes; between those, calls of the same size on one stream
 * reuse it with no clearing (the small scan state it keeps is tagged per
 * call), but two calls that may run concurrently never share one.  Stages
 * that emit variable-length lists (outliers, runs, compressed chunks,
 * Huffman cells) stage them in the workspace per tile/chunk and place them
 * with a separate kernel, so the workspace holds about as many bytes as the
 * stage's input.
 *
 * Reference interfaces replaced (file:line in szcompressor/FZGPUModules):
 *   fzh_lorenzo_quant_fwd_f32   launchLorenzoKernel / launchLorenzoKernel2D /
 *                               launchLorenzoKernel3D
 *                               (modules/fused/lorenzo_quant/lorenzo_quant.h:343-401)
 *   fzh_lorenzo_quant_inv_f32   launchLorenzoInverseKernel{,2D,3D}
 *                               (lorenzo_quant.h:354-414)
 *   fzh_value_base_f32          computeValueBase
 *                               (modules/predictors/predictor_utils.cuh:74-117)
 *   fzh_bitshuffle_{enc,dec}    BitshuffleStage::execute kernels
 *                               (modules/shufflers/bitshuffle/bitshuffle_stage.cu:355-475)
 *   fzh_rze_{enc,dec}           RZEStage::execute
 *                               (modules/coders/rze/rze_stage.cu:774-994)
 *   fzh_rle_{enc,dec}           RLEStage<T>::execute (modules/coders/rle/rle.cu:185-385)
 *   fzh_histogram               fz::module::GPU_histogram_generic
 *                               (modules/util/histogram/histogram.cu:100-110)
 *   fzh_huffman_build_book      phf_CPU_build_canonized_codebook_v2
 *                               (modules/coders/huffman/phf/hf_bk.cc:84-140)
 *   fzh_huffman_{encode,decode} phf::high_level<E>::encode / decode
 *                               (modules/coders/huffman/phf/hf_hl.cc:43-137)
 *   fzh_quant_abs_{fwd,inv}_f32 quantizer_abs_fwd_kernel / quantizer_abs_inv_kernel
 *                               (modules/quantizers/quantizer/quantizer.cu:30-137)
 *   fzh_bitpack_{enc,dec}       BitpackStage launchEncode / launchDecode
 *                               (modules/coders/bitpack/bitpack_stage.cu:130-165)
 *   fzh_max_u32                 cub::DeviceReduce::Max call in BitpackStage::execute
 *                               (bitpack_stage.cu:190-205)
 *   fzh_gather                  launch_gather_kernel (src/pipeline/concat_kernel.cu:63-71)
 */
#ifndef FZH_H
#define FZH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FZH_OK 0
#define FZH_E_INVALID (-1)   /* bad argument / unsupported configuration   */
#define FZH_E_RUNTIME (-2)   /* HIP runtime or launch failure              */
#define FZH_E_CAPACITY (-3)  /* output or workspace too small              */

/* ---- library ------------------------------------------------------------ */
const char* fzh_last_error(void);
const char* fzh_version(void);
/* Zero a workspace (fill kernel, safe under stream capture). */
int fzh_workspace_init(void* stream, void* d_ws, size_t ws_bytes);

/* ---- Lorenzo predict + quantize ------------------------------------------
 * dims (nx, ny, nz): ndim = nz>1 ? 3 : ny>1 ? 2 : 1 (lorenzo_quant.h:181-185).
 * n = element count of d_codes.  1D: elements [nx, n) of the input are read
 * as zeros when 0 < nx < n (the reference's zero padding, without a padded
 * copy; d_in needs only nx elements), else d_in holds n elements.  2D/3D
 * read nx*ny[*nz] elements and write zeros into codes[nx*ny*nz .. n).
 * ebx2_r = 1/(2*abs_eb) in float.  If d_ebx2_r is non-NULL the kernel reads
 * the value from device memory instead (set by fzh_value_base_f32).
 * code_bytes: 1 (u8) or 2 (u16) codes.  radius is cast to the code type.
 * Outliers: d_outlier_count receives the total count (may exceed
 * max_outliers; only the first max_outliers are stored), in tile-major,
 * row-major-within-tile order.                                              */
size_t fzh_lorenzo_quant_fwd_ws_bytes(size_t n, size_t nx, size_t ny, size_t nz);
int fzh_lorenzo_quant_fwd_f32(void* stream, const float* d_in, size_t n,
                              size_t nx, size_t ny, size_t nz,
                              float ebx2_r, const float* d_ebx2_r,
                              uint32_t radius, int zigzag, int code_bytes,
                              void* d_codes, float* d_outlier_err,
                              uint32_t* d_outlier_idx, uint32_t* d_outlier_count,
                              size_t max_outliers, void* d_ws, size_t ws_bytes);

/* Strict error bound (an extension; no counterpart in the reference): the
 * same codes and outliers as fzh_lorenzo_quant_fwd_f32, plus a patch list of
 * the elements the inverse would reconstruct with |x - x_hat| > eb (the
 * reference's FP32 arithmetic overshoots eb by a few ULPs on a few elements):
 * d_patch_val[i] = x[d_patch_idx[i]], ordered like the outliers, count in
 * *d_patch_count (may exceed max_patches; only max_patches stored).  eb and
 * ebx2 = 2*eb are read from d_eb_params[1] / [3] (fzh_value_base_f32's
 * layout) when it is non-NULL.  Guaranteed for finite inputs; a tile with
 * |x/(2 eb)| >= 2^21 anywhere is patched whole.                             */
size_t fzh_lorenzo_quant_fwd_strict_ws_bytes(size_t n, size_t nx, size_t ny, size_t nz);
int fzh_lorenzo_quant_fwd_strict_f32(void* stream, const float* d_in, size_t n,
                                     size_t nx, size_t ny, size_t nz,
                                     float ebx2_r, const float* d_ebx2_r,
                                     uint32_t radius, int zigzag, int code_bytes,
                                     void* d_codes, float* d_outlier_err,
                                     uint32_t* d_outlier_idx, uint32_t* d_outlier_count,
                                     size_t max_outliers, float eb, float ebx2,
                                     const float* d_eb_params, float* d_patch_val,
                                     uint32_t* d_patch_idx, uint32_t* d_patch_count,
                                     size_t max_patches, void* d_ws, size_t ws_bytes);
/* fzh_lorenzo_quant_inv_f32 followed by fzh_lorenzo_apply_patches (n_out =
 * n), for strict-mode streams (an extension; no counterpart in the
 * reference): with at most 256 outliers and 256 patches a single launch,
 * every tile's wave applying both short lists itself.  Same workspace as
 * fzh_lorenzo_quant_inv_f32.                                               */
int fzh_lorenzo_quant_inv_strict_f32(void* stream, const void* d_codes, size_t n,
                                     size_t nx, size_t ny, size_t nz,
                                     const float* d_outlier_err,
                                     const uint32_t* d_outlier_idx,
                                     const uint32_t* d_outlier_count, size_t max_outliers,
                                     float ebx2, int zigzag, int code_bytes,
                                     const float* d_patch_val, const uint32_t* d_patch_idx,
                                     const uint32_t* d_patch_count, size_t max_patches,
                                     float* d_out, void* d_ws, size_t ws_bytes);
/* After the inverse: d_out[idx[i]] = val[i] for i < min(*count, max_patches)
 * and idx[i] < n_out.                                                      */
int fzh_lorenzo_apply_patches(void* stream, float* d_out, const float* d_patch_val,
                              const uint32_t* d_patch_idx, const uint32_t* d_patch_count,
                              size_t max_patches, size_t n_out);

/* Inverse: uses min(*d_outlier_count, max_outliers) outliers in any order.
 * ebx2 = 2*abs_eb in float.                                                 */
size_t fzh_lorenzo_quant_inv_ws_bytes(size_t n, size_t nx, size_t ny, size_t nz,
                                      size_t max_outliers);
int fzh_lorenzo_quant_inv_f32(void* stream, const void* d_codes, size_t n,
                              size_t nx, size_t ny, size_t nz,
                              const float* d_outlier_err,
                              const uint32_t* d_outlier_idx,
                              const uint32_t* d_outlier_count,
                              size_t max_outliers, float ebx2, int zigzag,
                              int code_bytes, float* d_out, void* d_ws,
                              size_t ws_bytes);

/* The same inverse in two parts (together exactly fzh_lorenzo_quant_inv_f32
 * on one workspace): _bin sorts the outliers by tile into the workspace and
 * needs only the outlier arrays, so a caller can queue it before the codes
 * exist; _apply then reconstructs from the codes and the binned outliers
 * (same n, dims, outlier arrays, max_outliers and workspace).  With
 * max_outliers <= 256, _bin does nothing and every tile reads the outlier
 * list directly in _apply (no binning launches).                          */
int fzh_lorenzo_quant_inv_bin(void* stream, size_t n, size_t nx, size_t ny, size_t nz,
                              const float* d_outlier_err, const uint32_t* d_outlier_idx,
                              const uint32_t* d_outlier_count, size_t max_outliers,
                              void* d_ws, size_t ws_bytes);
int fzh_lorenzo_quant_inv_apply_f32(void* stream, const void* d_codes, size_t n,
                                    size_t nx, size_t ny, size_t nz,
                                    const float* d_outlier_err,
                                    const uint32_t* d_outlier_idx,
                                    const uint32_t* d_outlier_count, size_t max_outliers,
                                    float ebx2, int zigzag, int code_bytes, float* d_out,
                                    void* d_ws, size_t ws_bytes);

/* REL (mode 1: max|x|) / NOA (mode 2: max-min) value base, resolved on the
 * device.  d_params (4 floats) receives
 *   [0] value_base, [1] abs_eb, [2] ebx2_r = 1/(2*abs_eb), [3] ebx2 = 2*abs_eb
 * with abs_eb = user_eb*value_base, or user_eb when value_base <= 0
 * (lorenzo_quant.cu:612-644).  mode | FZH_VB_ZERO_PADDED folds a 0.0 into
 * the min/max, as the reference sees when it reduces a zero-padded copy.    */
#define FZH_VB_ZERO_PADDED 4
size_t fzh_value_base_ws_bytes(size_t n);
int fzh_value_base_f32(void* stream, const float* d_in, size_t n, int mode,
                       float user_eb, float* d_params, void* d_ws, size_t ws_bytes);
/* fzh_value_base_f32 without a workspace argument and without any
 * allocation (no counterpart in the reference; for callers that cannot
 * allocate: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same kernels, same d_params.  Its scratch is a
 * fixed 4 KiB device array per GPU (FZH_VB_NOALLOC_PARTIALS min/max pairs,
 * one per workgroup of the first pass; the grid is capped at that many
 * workgroups, so every n fits), kept until the process ends; every pair the
 * second pass reads is written by the same call, so nothing is cleared.
 * All calls on one GPU share that array: calls that may overlap on one GPU
 * (different streams, or host threads) race and must be serialised by the
 * caller.  The stream's device (the current device for NULL) is used.      */
#define FZH_VB_NOALLOC_PARTIALS 512
int fzh_value_base_f32_noalloc(void* stream, const float* d_in, size_t n, int mode,
                               float user_eb, float* d_params);

/* ---- Lorenzo predict + quantize, FP64 input --------------------------------
 * LorenzoQuantStage<double, uint16_t | uint32_t> (lorenzo_quant.h:335-339;
 * kernels lorenzo_quant.cu:262-364, lorenzo_quant_nd.cu:29-105,199-274 with
 * TInput = double).  The same tiles, dims rules, zero padding and outlier
 * order as the FP32 entry points; every step runs in double
 * (s = round(x * ebx2_r), the differences, |delta| < radius) and outlier
 * errors are doubles.  code_bytes: 2 (u16) or 4 (u32).  d_in and d_codes
 * 16-byte aligned.  If d_ebx2_r is non-NULL the kernel reads ebx2_r (a
 * double) from device memory (fzh_value_base_f64's d_params + 2).           */
size_t fzh_lorenzo_quant_fwd_f64_ws_bytes(size_t n, size_t nx, size_t ny, size_t nz);
int fzh_lorenzo_quant_fwd_f64(void* stream, const double* d_in, size_t n,
                              size_t nx, size_t ny, size_t nz,
                              double ebx2_r, const double* d_ebx2_r,
                              uint32_t radius, int zigzag, int code_bytes,
                              void* d_codes, double* d_outlier_err,
                              uint32_t* d_outlier_idx, uint32_t* d_outlier_count,
                              size_t max_outliers, void* d_ws, size_t ws_bytes);
/* Inverse (lorenzo_quant.cu:17-87,155-244; lorenzo_quant_nd.cu:112-190,
 * 281-364 in double): min(*d_outlier_count, max_outliers) outliers with
 * distinct indices, in any order; ebx2 = 2*abs_eb in double.  d_codes and
 * d_out 16-byte aligned.                                                    */
size_t fzh_lorenzo_quant_inv_f64_ws_bytes(size_t n, size_t nx, size_t ny, size_t nz,
                                          size_t max_outliers);
int fzh_lorenzo_quant_inv_f64(void* stream, const void* d_codes, size_t n,
                              size_t nx, size_t ny, size_t nz,
                              const double* d_outlier_err,
                              const uint32_t* d_outlier_idx,
                              const uint32_t* d_outlier_count,
                              size_t max_outliers, double ebx2, int zigzag,
                              int code_bytes, double* d_out, void* d_ws,
                              size_t ws_bytes);
/* REL / NOA value base of FP64 input (predictor_utils.cuh:75-116 with
 * TInput = double: min/max in double, the base cast to float) and the bound
 * in double (lorenzo_quant.cu:612-644).  d_params (4 doubles):
 *   [0] (double)value_base, [1] abs_eb = (double)user_eb * [0] (user_eb when
 *   value_base <= 0), [2] ebx2_r = 1/(2*abs_eb), [3] ebx2 = 2*abs_eb.
 * mode as fzh_value_base_f32 (| FZH_VB_ZERO_PADDED).  d_in 16-byte aligned. */
size_t fzh_value_base_f64_ws_bytes(size_t n);
int fzh_value_base_f64(void* stream, const double* d_in, size_t n, int mode,
                       float user_eb, double* d_params, void* d_ws, size_t ws_bytes);

/* ---- Bitshuffle (block_size multiple of 1024*elem_width; tail copied) ---- */
int fzh_bitshuffle_enc(void* stream, const void* d_in, void* d_out,
                       size_t nbytes, uint32_t block_size, int elem_width);
int fzh_bitshuffle_dec(void* stream, const void* d_in, void* d_out,
                       size_t nbytes, uint32_t block_size, int elem_width);

/* ---- RZE (chunk 16384, 4 levels) ------------------------------------------
 * Stream: [orig u32][n_chunks u32][size|raw-flag u32 x n][chunks], padded
 * with zeros to 4 bytes.  d_out_size (device u64) receives the stream size. */
size_t fzh_rze_enc_bound(size_t nbytes);
size_t fzh_rze_enc_ws_bytes(size_t nbytes);
int fzh_rze_enc(void* stream, const void* d_in, size_t nbytes, void* d_out,
                size_t out_cap, uint64_t* d_out_size, void* d_ws, size_t ws_bytes);
/* orig_bytes: expected decoded size (from the stage header).  *d_status is
 * set to 0 on success, nonzero if the stream is malformed or disagrees.    */
size_t fzh_rze_dec_ws_bytes(size_t orig_bytes);
int fzh_rze_dec(void* stream, const void* d_in, size_t in_bytes, void* d_out,
                size_t orig_bytes, int* d_status, void* d_ws, size_t ws_bytes);

/* Fused Bitshuffle(block 16384, element width 1 or 2) + RZE: the same stream
 * fzh_bitshuffle_enc followed by fzh_rze_enc produce, in one pass over the
 * input (no shuffled intermediate in HBM); the decoder likewise returns the
 * unshuffled bytes.  Workspaces as for fzh_rze_enc / fzh_rze_dec.  This pair
 * is what the pipeline runs for a BitshuffleStage feeding an RZEStage
 * (bitshuffle_stage.cu:355-475 + rze_stage.cu:774-994).                   */
int fzh_rze_enc_shuffled(void* stream, const void* d_in, size_t nbytes, int elem_width,
                         void* d_out, size_t out_cap, uint64_t* d_out_size, void* d_ws,
                         size_t ws_bytes);
int fzh_rze_dec_shuffled(void* stream, const void* d_in, size_t in_bytes, void* d_out,
                         size_t orig_bytes, int elem_width, int* d_status, void* d_ws,
                         size_t ws_bytes);

/* ---- RLE: [R u32][values T x R, padded to 4][lengths u32 x R] ------------ */
size_t fzh_rle_enc_bound(size_t n, int elem_bytes);
size_t fzh_rle_enc_ws_bytes(size_t n);
int fzh_rle_enc(void* stream, const void* d_in, size_t n, int elem_bytes,
                void* d_out, size_t out_cap, uint64_t* d_out_size,
                void* d_ws, size_t ws_bytes);
size_t fzh_rle_dec_ws_bytes(size_t n);
int fzh_rle_dec(void* stream, const void* d_in, size_t in_bytes,
                int elem_bytes, void* d_out, size_t n, int* d_status,
                void* d_ws, size_t ws_bytes);

/* ---- Histogram / Huffman (coarse 768- or fine 1024-symbol partitions) ------ */
int fzh_histogram(void* stream, const void* d_in, size_t n, int elem_bytes,
                  uint32_t bklen, uint32_t* d_freq);
/* Host routine: canonical codebook from a host histogram. */
size_t fzh_huffman_revbook_bytes(uint32_t bklen, int elem_bytes);
int fzh_huffman_build_book(const uint32_t* h_freq, uint32_t bklen,
                           int elem_bytes, uint32_t* h_bk4, uint8_t* h_revbook);
/* The same codebook built on the device (an extension; the reference builds
 * it on the host, huffman_stage.h:108): the reference builder's binary-heap
 * merges replayed in order by one lane with the heap in LDS, leaf depths
 * from parent links, canonisation in parallel; d_bk4 and d_revbook are byte
 * for byte those of fzh_huffman_build_book on the same histogram.  No host
 * round trip, so a Huffman encode can be captured in a graph.  bklen <= 1024
 * (256 for 1-byte symbols).  *d_status: 0, 1 when the histogram does not sum
 * to n_expected (symbols outside [0, bklen); the encoder then gives them no
 * code), 2 when a code would exceed 58 bits.  One workgroup whose time is
 * the replay's chain of dependent LDS accesses (3 heap operations of about
 * log2(m) steps per merge, m symbols present): on MI355X about 130 us for
 * 83 symbols and 2.3 ms for 1024, where the host builder takes 7-22 us --
 * it serves callers that cannot take a host round trip, not speed.        */
int fzh_huffman_build_book_device(void* stream, const uint32_t* d_freq, uint32_t bklen,
                                  int elem_bytes, uint64_t n_expected, uint32_t* d_bk4,
                                  uint8_t* d_revbook, int* d_status);
size_t fzh_huffman_enc_bound(size_t n, uint32_t bklen, int elem_bytes);
size_t fzh_huffman_enc_ws_bytes(size_t n);
int fzh_huffman_encode(void* stream, const void* d_in, size_t n, int elem_bytes,
                       uint32_t bklen, const uint32_t* d_bk4,
                       const uint8_t* d_revbook, void* d_out, size_t out_cap,
                       uint64_t* d_out_size, void* d_ws, size_t ws_bytes);
/* Fine mode (HuffmanStage Fine, huffman_stage.cu:41; hf_buf.cc:48-49): the
 * same stream with 1024-symbol partitions, as the reference's ReVISIT-lite
 * kernel lays it out (hf_kernels.cu:145-250).  The reference takes it only
 * when no codeword of the book is longer than 8 bits and falls back to the
 * coarse stream otherwise (hf_hl.cc:69-80); callers check that on the host
 * book (HuffmanStage does).  Same bound and workspace as fzh_huffman_encode. */
int fzh_huffman_encode_fine(void* stream, const void* d_in, size_t n, int elem_bytes,
                            uint32_t bklen, const uint32_t* d_bk4,
                            const uint8_t* d_revbook, void* d_out, size_t out_cap,
                            uint64_t* d_out_size, void* d_ws, size_t ws_bytes);
int fzh_huffman_decode(void* stream, const void* d_in, size_t in_bytes,
                       int elem_bytes, void* d_out, size_t n, int* d_status);

/* ---- Quantizer (ABS) ------------------------------------------------------ */
int fzh_quant_abs_fwd_f32(void* stream, const float* d_in, size_t n,
                          float ebx2_r, float threshold, uint32_t radius,
                          int zigzag, int code_bytes, void* d_codes,
                          float* d_outlier_val, uint32_t* d_outlier_idx,
                          uint32_t* d_outlier_count, size_t max_outliers,
                          void* d_ws, size_t ws_bytes);
size_t fzh_quant_abs_fwd_ws_bytes(size_t n);
int fzh_quant_abs_inv_f32(void* stream, const void* d_codes, size_t n,
                          float ebx2, int zigzag, int code_bytes,
                          const float* d_outlier_val,
                          const uint32_t* d_outlier_idx,
                          const uint32_t* d_outlier_count,
                          size_t max_outliers, float* d_out);
/* In-place outliers (QuantizerStage::setInplaceOutliers,
 * quantizer_abs_fwd_inplace_kernel / quantizer_abs_inv_inplace_kernel,
 * quantizer.cu:83-104, 147-168): u32 codes; a quantisable x stores
 * zigzag(q) = (q << 1) ^ (q >> 31), any other x its raw IEEE-754 bits; the
 * inverse reads (code >> 1) >= radius as raw bits.  No outlier lists.       */
int fzh_quant_abs_fwd_inplace_f32(void* stream, const float* d_in, size_t n,
                                  float ebx2_r, float threshold, uint32_t radius,
                                  uint32_t* d_codes);
int fzh_quant_abs_inv_inplace_f32(void* stream, const uint32_t* d_codes, size_t n,
                                  float ebx2, uint32_t radius, float* d_out);

/* ---- Quantizer (REL, log2-space bins) --------------------------------------
 * quantizer_rel_fwd_kernel / quantizer_rel_inv_kernel (quantizer.cu:171-352)
 * with the stage's constants (quantizer.cu:563-568): log2eb =
 * 2*log2f(1+eps), 1/log2eb, 1+eps, 1/(1+eps) (fzh_quant_rel_params).  A
 * normal finite x with |x| < threshold gets log_bin = rint(log2|x| *
 * (1/log2eb)) and the code ((|log_bin| << 2) | (log_bin < 0) << 1 |
 * (x < 0)) + 1 when |log_bin| < radius, the code fits the code type and
 * 2^(log_bin*log2eb) is within [|x|/(1+eps), |x|(1+eps)]; otherwise code 0
 * and an outlier (x, i), listed in index order.  log2 and 2^t are the
 * correctly rounded float results (the reference's __log2f / exp2f are
 * within 2 ulp of them).  Same workspace as fzh_quant_abs_fwd_f32.  The
 * inverse writes +-2^(log_bin*log2eb), 0 for code 0, then assigns the
 * outliers.                                                                */
void fzh_quant_rel_params(float eps, float* params4);
int fzh_quant_rel_fwd_f32(void* stream, const float* d_in, size_t n, float eps,
                          float threshold, uint32_t radius, int code_bytes,
                          void* d_codes, float* d_outlier_val,
                          uint32_t* d_outlier_idx, uint32_t* d_outlier_count,
                          size_t max_outliers, void* d_ws, size_t ws_bytes);
int fzh_quant_rel_inv_f32(void* stream, const void* d_codes, size_t n, float eps,
                          int code_bytes, const float* d_outlier_val,
                          const uint32_t* d_outlier_idx,
                          const uint32_t* d_outlier_count, size_t max_outliers,
                          float* d_out);

/* ---- Bitpack -------------------------------------------------------------- */
size_t fzh_max_u32_ws_bytes(void);
int fzh_max_u32(void* stream, const void* d_in, size_t n, int elem_bytes,
                uint32_t* d_max, void* d_ws, size_t ws_bytes);
int fzh_bitpack_enc(void* stream, const void* d_in, size_t n, int elem_bytes,
                    int nbits, void* d_out);
int fzh_bitpack_dec(void* stream, const void* d_in, size_t n, int elem_bytes,
                    int nbits, void* d_out);

/* ---- Gather (multi-output concat) ----------------------------------------- */
typedef struct fzh_copy_desc {
    const void* src;
    void* dst;
    uint64_t bytes;
} fzh_copy_desc;
int fzh_gather(void* stream, const fzh_copy_desc* d_descs, int n_segs);

/* Synthetic-data helper: an nx*ny*nz window of a periodic base field
 * (bx*by*bz, x fastest) starting at (x0, y0, z0), wrapping at the edges.   */
int fzh_periodic_window_f32(void* stream, const float* d_base, uint64_t bx, uint64_t by, uint64_t bz,
                            uint64_t x0, uint64_t y0, uint64_t z0, float* d_out, uint64_t nx, uint64_t ny,
                            uint64_t nz);

/* Synthetic-data helper (no counterpart in the reference): a seeded
 * Gaussian random field with power spectrum ~ k^expo in nx*ny*nz (x fastest;
 * d = 3 if nz > 1, else 2 if ny > 1, else 1): octaves of smoothly blended
 * standard-normal lattice noise at spacings 2^o, amplitude 2^(-o(expo+d)/2),
 * the largest amplitude scaled to 1.  Deterministic for a given seed.      */
int fzh_synth_power_law_f32(void* stream, float* d_out, uint64_t nx, uint64_t ny, uint64_t nz,
                            float expo, uint64_t seed);

/* Per-kernel device time (no counterpart in the reference, whose
 * benchmarks time whole stages).  While enabled, each launch of the chain's
 * main kernels outside stream capture is bracketed by a HIP event pair on
 * its own stream; fzh_kernel_timing_get waits for the recorded pairs and
 * returns the summed milliseconds and launch count under `name`:
 * "minmax_partial_kernel", "lz1d_fwd_kernel" / "lz2d_fwd_kernel" /
 * "lz3d_fwd_kernel", "place_kernel", "scan_small_kernel", "rze_enc_kernel",
 * "rze_pack_kernel", "rze_offsets_kernel", "rze_dec_kernel",
 * "bin_count_kernel", "bin_place_kernel", "lz1d_inv_kernel" /
 * "lz2d_inv_kernel" / "lz3d_inv_kernel", "apply_patches_kernel",
 * "bitshuffle_kernel", "copy_kernel".  Enabling or disabling clears the
 * totals.  The event records add a small gap between launches.            */
int fzh_kernel_timing_enable(int on);
int fzh_kernel_timing_get(const char* name, double* total_ms, uint64_t* count);

/* ---- Transforms (modules/transforms, modules/predictors/diff,
 *      modules/predictors/lorenzo) ------------------------------------------
 * Element types are given as the FZM DataType code: 0 u8, 1 u16, 2 u32,
 * 3 u64, 4 i8, 5 i16, 6 i32, 7 i64, 8 f32, 9 f64.  Buffers 16-byte aligned.
 *
 * Zigzag (zigzag.h:70-80: (x << 1) ^ (x >> (W-1))) and negabinary
 * (negabinary.h: (x + 0xAA..) ^ 0xAA..) codes of signed integers of
 * elem_bytes 1/2/4/8, element-wise, in place allowed; inverse = 1 decodes.  */
int fzh_zigzag(void* stream, const void* d_in, void* d_out, size_t n, int elem_bytes, int inverse);
int fzh_negabinary(void* stream, const void* d_in, void* d_out, size_t n, int elem_bytes, int inverse);

/* Difference coding (DifferenceStage, diff.cu:22-240): d_out[i] = in[i] -
 * in[i-1], the first element of every chunk of chunk_elems (a power of two;
 * 0 = one chunk) stored as is, in the element type's arithmetic (integers
 * wrap); negabinary = 1 (signed integer types) stores the negabinary codes.
 * Inverse: the cumulative sum per chunk (negabinary decoded first), d_in !=
 * d_out; floating-point sums follow the reference's association: 512-element
 * tiles of a chunk, each as cub::BlockScan<T, 512> (raking) sums it -- 32
 * segments of 16 elements added left to right, a Kogge-Stone scan of the
 * segment sums over 32 lanes, each segment's elements added left to right
 * onto the previous segments' sum -- plus the left-to-right sum of the
 * chunk's earlier tile totals.  Workspace: fzh_diff_inv_ws_bytes.           */
int fzh_diff_fwd(void* stream, const void* d_in, void* d_out, size_t n, int dtype, size_t chunk_elems,
                 int negabinary);
size_t fzh_diff_inv_ws_bytes(size_t n, int dtype, size_t chunk_elems);
int fzh_diff_inv(void* stream, const void* d_in, void* d_out, size_t n, int dtype, size_t chunk_elems,
                 int negabinary, void* d_ws, size_t ws_bytes);

/* Integer Lorenzo (LorenzoStage, lorenzo_stage.cu): signed elements of
 * elem_bytes 1/2/4/8, wrapping arithmetic, d_in != d_out.  1D (ny = nz = 1):
 * d_out[i] = in[i] - in[i-1] inside blocks of 256 elements (the first of a
 * block as is); 2D/3D: the inclusion-exclusion delta of every element over
 * the whole field, neighbours outside it 0.  Inverse: inclusive prefix sums
 * per 256-element block (1D), or along x, then y, then z.                   */
int fzh_lorenzo_int_fwd(void* stream, const void* d_in, void* d_out, size_t nx, size_t ny, size_t nz,
                        int elem_bytes);
size_t fzh_lorenzo_int_inv_ws_bytes(size_t nx, size_t ny, size_t nz);
int fzh_lorenzo_int_inv(void* stream, const void* d_in, void* d_out, size_t nx, size_t ny, size_t nz,
                        int elem_bytes, void* d_ws, size_t ws_bytes);

/* Exclusive scan of n u32 values: the offsets primitive under every stage
 * whose unit sizes are only known on the device (RZE chunk offsets, outlier
 * compaction, outlier binning, Huffman cell offsets); the reference takes
 * these offsets from CUB scans or host loops (rze_stage.cu:616-640).
 * d_out may alias d_in; *d_total = the sum.  Sums saturate at 0xFFFFFFFF
 * unless flags has FZH_SCAN_WRAP (modulo 2^32); FZH_SCAN_CLEAR zeroes d_in
 * as it is read (d_in != d_out).  Workspace: fzh_scan_ws_bytes(n).        */
#define FZH_SCAN_WRAP 1
#define FZH_SCAN_CLEAR 2
size_t fzh_scan_ws_bytes(uint64_t n);
int fzh_scan_u32(void* stream, uint32_t* d_in, uint32_t* d_out, uint64_t n, uint32_t* d_total, int flags,
                 void* d_ws, size_t ws_bytes);

/* Plain device copy (grid-stride 16-byte accesses; 16-byte-aligned
 * pointers), capture-safe.  Also the bandwidth reference for benchmarks.   */
int fzh_copy(void* stream, void* d_dst, const void* d_src, size_t bytes);

/* Device-side concat of up to FZH_CONCAT_MAX_SLOTS outputs whose sizes may
 * only be known on the device (replaces Pipeline::writeConcatBuffer,
 * src/pipeline/compressor.cpp:805-851, which needs the sizes on the host).
 * Layout: [u32 n][u64 size_0..n-1] zero-padded to 16 B, then each slot's
 * bytes zero-padded to 16 B (the reference leaves the pads unwritten).
 * Slot size: kind 0 = `size`; kind 1 = *(const uint64_t*)size_ptr;
 * kind 2 = *(const uint32_t*)size_ptr * elem; every size is clamped to `cap`
 * (the slot's source capacity).  slots is a host array (passed by value to
 * the kernels, so the call is capture-safe); src must be 16-byte aligned.
 * A slot whose src already is its place in d_out (a stage that wrote its
 * output in place) is not copied; only its pad bytes are zeroed.
 * d_total (u64) receives the concat length.  d_ws: fzh_concat_ws_bytes().  */
#define FZH_CONCAT_MAX_SLOTS 32
typedef struct fzh_concat_slot {
    const void* src;
    const void* size_ptr;
    uint64_t size;
    uint64_t cap;
    uint32_t kind;
    uint32_t elem;
} fzh_concat_slot;
size_t fzh_concat_ws_bytes(void);
uint64_t fzh_concat_bound(const fzh_concat_slot* slots, int n);
int fzh_concat(void* stream, const fzh_concat_slot* slots, int n, void* d_out,
               uint64_t out_cap, uint64_t* d_total, void* d_ws, size_t ws_bytes);

/* Copies up to FZH_READBACK_MAX small device values (4-byte multiples,
 * 4-byte aligned) into one host-mapped pinned buffer with a single kernel;
 * visible on the host after the stream is synchronised.  Entries are passed
 * by value (capture-safe).                                                  */
#define FZH_READBACK_MAX 32
typedef struct fzh_readback_entry {
    const void* src;
    uint32_t bytes;
    uint32_t dst_offset;
} fzh_readback_entry;
int fzh_readback(void* stream, const fzh_readback_entry* entries, int n, void* d_mapped_dst);

/* fzh_concat followed by fzh_readback, in one launch (the pipeline's last
 * step of compress: Pipeline::writeConcatBuffer, compressor.cpp:805-851,
 * plus the size readbacks the reference does with cudaMemcpy).  Entries may
 * read the concat header in d_out and the total; they see the values this
 * call writes.  No workspace.                                              */
int fzh_concat_readback(void* stream, const fzh_concat_slot* slots, int n, void* d_out,
                        uint64_t out_cap, uint64_t* d_total,
                        const fzh_readback_entry* entries, int n_entries,
                        void* d_mapped_dst);

/* ---- Device helpers for bindings (not on the hot path) ------------------- */
int fzh_device_alloc(void** d_ptr, size_t bytes);
int fzh_device_free(void* d_ptr);
int fzh_memcpy_htod(void* d_dst, const void* h_src, size_t bytes);
int fzh_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes);
int fzh_memset(void* d_dst, int value, size_t bytes);
int fzh_synchronize(void);
int fzh_stream_create(void** stream);
int fzh_stream_destroy(void* stream);
int fzh_stream_sync(void* stream);
int fzh_event_create(void** ev);
int fzh_event_destroy(void* ev);
int fzh_event_record(void* ev, void* stream);
int fzh_event_elapsed_ms(void* ev_start, void* ev_end, float* ms);
int fzh_device_count(int* count);

#ifdef __cplusplus
}
#endif
#endif /* FZH_H */

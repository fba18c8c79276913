/*
 * fz_oracle.h — CPU restatement of the FZGPUModules stage-chain semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libfzh.so, the fz::
 * C++ classes, bench.py's timed GPU legs) may link or call this library.
 * It is loaded by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, as the checker.
 *
 * Every function restates one reference kernel/host routine with the same
 * FP32 operation order (citations are to the reference repository):
 *
 *   orc_lorenzo_fwd        modules/fused/lorenzo_quant/lorenzo_quant.cu:262-364 (1D)
 *                          modules/fused/lorenzo_quant/lorenzo_quant_nd.cu:29-105 (2D)
 *                          modules/fused/lorenzo_quant/lorenzo_quant_nd.cu:199-274 (3D)
 *   orc_lorenzo_inv        lorenzo_quant.cu:17-87,155-244 (1D scan tree)
 *                          lorenzo_quant_nd.cu:112-190 (2D), 281-364 (3D)
 *   orc_value_base         modules/predictors/predictor_utils.cuh:27-117
 *   orc_bitshuffle_*       modules/shufflers/bitshuffle/bitshuffle_stage.cu:66-349
 *   orc_rze_*              modules/coders/rze/rze_stage.cu:135-606, 774-994
 *   orc_rle_*              modules/coders/rle/rle.cu:19-148, 196-384
 *   orc_histogram          modules/util/histogram/histogram.cu:26-60
 *   orc_huffman_book       modules/coders/huffman/phf/hf_bk.cc:84-140,
 *                          hf_bk_impl1.cc:55-163, hf_bk_internal.cc:58-101,
 *                          hf_canon.cc:86-131
 *   orc_huffman_encode     phf/hf_kernels.cu:59-132, 366-439; phf/hf_hl.cc:43-120;
 *                          phf/hf_buf.cc:151-176
 *   orc_huffman_decode     phf/hf_kernels.cu:258-312
 *   orc_quant_abs_*        modules/quantizers/quantizer/quantizer.cu:30-67, 116-137
 *   orc_quant_rel_*        quantizer.cu:171-352, 426-436, 563-580 (log2/exp2 correctly
 *                          rounded to float where the reference uses __log2f/exp2f)
 *   orc_bitpack_*          modules/coders/bitpack/bitpack_stage.cu:30-124
 *   orc_crc32              src/pipeline/compressor_io.cpp:22-68
 *
 * Outlier lists are emitted in one canonical order (tile-major, row-major
 * inside a tile); the reference's order is set by atomics and is compared as
 * a set sorted by index.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Lorenzo predict + quantize (float input, u16/u8/u32 codes) -------- */
/* code_bytes: 1, 2 or 4.  dims {nx,ny,nz}; ndim = nz>1 ? 3 : ny>1 ? 2 : 1.
 * n is the element count (1D uses n; 2D/3D touch only nx*ny*nz and leave the
 * rest of codes[] as written by the caller).  Returns total outlier count
 * (may exceed max_outliers; only the first max_outliers are stored). */
uint32_t orc_lorenzo_fwd_f32(const float* in, size_t n, size_t nx, size_t ny,
                             size_t nz, float ebx2_r, uint32_t radius,
                             int zigzag, int code_bytes, void* codes,
                             float* out_err, uint32_t* out_idx,
                             size_t max_outliers);

/* Inverse: scatters min(count,max_outliers) outliers, then runs the exact
 * reference scan tree and scales by ebx2. */
void orc_lorenzo_inv_f32(const void* codes, size_t n, size_t nx, size_t ny,
                         size_t nz, const float* err, const uint32_t* idx,
                         uint32_t count, size_t max_outliers, float ebx2,
                         int zigzag, int code_bytes, float* out);

/* FP64 input (LorenzoQuantStage<double, TCode>, lorenzo_quant.h:335-339):
 * the same tiles, order and scan trees with every step in double; outlier
 * errors are doubles. */
uint32_t orc_lorenzo_fwd_f64(const double* in, size_t n, size_t nx, size_t ny,
                             size_t nz, double ebx2_r, uint32_t radius,
                             int zigzag, int code_bytes, void* codes,
                             double* out_err, uint32_t* out_idx,
                             size_t max_outliers);
void orc_lorenzo_inv_f64(const void* codes, size_t n, size_t nx, size_t ny,
                         size_t nz, const double* err, const uint32_t* idx,
                         uint32_t count, size_t max_outliers, double ebx2,
                         int zigzag, int code_bytes, double* out);

/* mode 1 = REL (max |x|), 2 = NOA (max - min); computed in the input type
 * and returned as float (predictor_utils.cuh:110-115). */
float orc_value_base_f32(const float* in, size_t n, int mode);
float orc_value_base_f64(const double* in, size_t n, int mode);

/* ---- Bitshuffle ---------------------------------------------------------- */
void orc_bitshuffle_enc(const uint8_t* in, uint8_t* out, size_t nbytes,
                        uint32_t block_size, int elem_width);
void orc_bitshuffle_dec(const uint8_t* in, uint8_t* out, size_t nbytes,
                        uint32_t block_size, int elem_width);

/* ---- RZE (chunk 16384, 4 levels) ---------------------------------------- */
/* Returns the stream size (multiple of 4).  out must hold
 * nbytes + 8 + 4*n_chunks + 4 bytes. */
size_t orc_rze_enc(const uint8_t* in, size_t nbytes, uint8_t* out);
/* Returns the decoded size, or (size_t)-1 on a malformed stream. */
size_t orc_rze_dec(const uint8_t* in, size_t in_bytes, uint8_t* out,
                   size_t out_cap);

/* ---- RLE ----------------------------------------------------------------- */
/* elem_bytes 1, 2 or 4.  Returns the stream size. */
size_t orc_rle_enc(const void* in, size_t n, int elem_bytes, uint8_t* out);
/* Returns number of decoded elements, or (size_t)-1 on error. */
size_t orc_rle_dec(const uint8_t* in, size_t in_bytes, int elem_bytes,
                   void* out, size_t out_cap_elems);

/* ---- Histogram + Huffman (coarse mode, sublen 768, 32-bit cells) -------- */
void orc_histogram(const void* in, size_t n, int elem_bytes, uint32_t bklen,
                   uint32_t* freq);
/* bk4[bklen] canonical codebook words (bitcount in bits 31..27),
 * revbook = first[32] i32, entry[32] i32, keys[bklen] (elem_bytes each). */
void orc_huffman_book(const uint32_t* freq, uint32_t bklen, int elem_bytes,
                      uint32_t* bk4, uint8_t* revbook);
size_t orc_huffman_revbook_bytes(uint32_t bklen, int elem_bytes);
/* Full HuffmanStage forward (histogram -> book -> coarse encode -> merged
 * PHF stream).  Returns the stream size, 0 if a symbol is >= bklen. */
size_t orc_huffman_encode(const void* in, size_t n, int elem_bytes,
                          uint32_t bklen, uint8_t* out, size_t out_cap);
size_t orc_huffman_encode_mode(const void* in, size_t n, int elem_bytes, uint32_t bklen, int fine,
                               uint8_t* out, size_t cap);
size_t orc_huffman_encode_bound(size_t n, uint32_t bklen, int elem_bytes);
/* Decode a PHF stream into n symbols.  Returns 0 on success. */
int orc_huffman_decode(const uint8_t* in, size_t in_bytes, int elem_bytes,
                       void* out, size_t n);

/* ---- Quantizer (ABS) + Bitpack ------------------------------------------ */
uint32_t orc_quant_abs_fwd_f32(const float* in, size_t n, float ebx2_r,
                               float threshold, uint32_t radius, int zigzag,
                               int code_bytes, void* codes, float* out_val,
                               uint32_t* out_idx, size_t max_outliers);
void orc_quant_abs_inv_f32(const void* codes, size_t n, float ebx2, int zigzag,
                           int code_bytes, const float* vals,
                           const uint32_t* idx, uint32_t count,
                           size_t max_outliers, float* out);
/* Quantizer REL (log2-space bins): [log2eb, 1/log2eb, 1+eps, 1/(1+eps)]. */
void orc_quant_rel_params(float eps, float* params4);
uint32_t orc_quant_rel_fwd_f32(const float* in, size_t n, float eps, float threshold,
                               uint32_t radius, int code_bytes, void* codes,
                               float* out_val, uint32_t* out_idx, size_t max_outliers);
void orc_quant_rel_inv_f32(const void* codes, size_t n, float eps, int code_bytes,
                           const float* vals, const uint32_t* idx, uint32_t count,
                           size_t max_outliers, float* out);
int orc_bitpack_nbits_for_max(uint32_t max_val, int elem_bytes);
size_t orc_bitpack_enc(const void* in, size_t n, int elem_bytes, int nbits,
                       uint8_t* out);
void orc_bitpack_dec(const uint8_t* in, size_t n, int elem_bytes, int nbits,
                     void* out);

/* ---- CRC32 (IEEE 802.3, reflected) -------------------------------------- */
uint32_t orc_crc32(const uint8_t* data, size_t n);

/* ---- CPU baseline: whole Lorenzo(2D)->Bitshuffle->RZE chain, threaded --- */
/* Returns compressed RZE bytes; *n_outliers receives the outlier count. */
size_t orc_chain_l2d_bs_rze(const float* in, size_t n_padded, size_t nx,
                            size_t ny, float ebx2_r, uint32_t radius,
                            int zigzag, uint16_t* codes_tmp,
                            uint8_t* shuf_tmp, uint8_t* out, float* out_err,
                            uint32_t* out_idx, size_t max_outliers,
                            uint32_t* n_outliers);
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif

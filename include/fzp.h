/* fzp.h — C ABI over fz::Pipeline (include/fz/pipeline.h).
 *
 * For bindings that cannot use the C++ classes (ctypes, cgo, JNI).  Every
 * call returns 0 on success or a negative FZH_E_* code with the message in
 * fzh_last_error(); C++ exceptions never cross this boundary.  Handles are
 * not thread-safe (one pipeline per host thread, as in the reference,
 * include/mem/mempool.h:80).
 *
 * Entry point                     replaces (reference)
 *   fzp_create / fzp_destroy      Pipeline(size_t, MemoryStrategy, float)
 *                                 (include/pipeline/compressor.h:41-54)
 *   fzp_add_lorenzo ...           Pipeline::addStage<T>(...) (compressor.h:82-83,
 *                                 733-758) with the stage's setters
 *   fzp_connect                   Pipeline::connect (compressor.h:92)
 *   fzp_set_dims                  Pipeline::setDims (compressor.h:72)
 *   fzp_finalize                  Pipeline::finalize (compressor.h:102)
 *   fzp_compress                  Pipeline::compress (compressor.h:175-181)
 *   fzp_decompress                Pipeline::decompress (compressor.h:232-238)
 *   fzp_write_file                Pipeline::writeToFile (compressor.h:357)
 *   fzp_decompress_file           Pipeline::decompressFromFile (compressor.h:379-386)
 *   fzp_enable_graph_mode         Pipeline::enableGraphMode (compressor.h:327)
 *   fzp_enable_profiling /        Pipeline::enableProfiling /
 *   fzp_last_perf                 getLastPerfResult (compressor.h:278-282)
 */
#ifndef FZP_H
#define FZP_H

#include <stddef.h>
#include <stdint.h>

#include "fzh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fzp_pipeline fzp_pipeline;

/* error-bound modes (lorenzo_quant.h ErrorBoundMode) */
#define FZP_EB_ABS 0
#define FZP_EB_REL 1
#define FZP_EB_NOA 2

int fzp_create(size_t input_size_hint, float pool_multiplier, fzp_pipeline** out);
int fzp_destroy(fzp_pipeline* p);

/* Stage constructors; *stage_id receives the stage's index in this pipeline. */
int fzp_add_lorenzo(fzp_pipeline* p, float error_bound, int eb_mode, uint32_t quant_radius,
                    float outlier_capacity, int zigzag, int code_bytes, float value_base,
                    int* stage_id);
/* FP64 input: LorenzoQuantStage<double, uint16_t> (code_bytes 2) or
 * <double, uint32_t> (code_bytes 4) (lorenzo_quant.h:335-339). */
int fzp_add_lorenzo_f64(fzp_pipeline* p, float error_bound, int eb_mode, uint32_t quant_radius,
                        float outlier_capacity, int zigzag, int code_bytes, float value_base,
                        int* stage_id);
/* Extension (no counterpart in the reference): strict error bound on a
 * LorenzoQuant stage -- codes and outliers unchanged, plus a patch list so
 * every decompressed value is within eb (LorenzoQuantStage::setStrictErrorBound). */
int fzp_set_lorenzo_strict(fzp_pipeline* p, int stage_id, int on, float patch_capacity);
int fzp_add_bitshuffle(fzp_pipeline* p, uint32_t block_size, int elem_width, int* stage_id);
int fzp_add_rze(fzp_pipeline* p, uint32_t chunk_size, int levels, int* stage_id);
int fzp_add_rle(fzp_pipeline* p, int elem_bytes, int* stage_id);
int fzp_add_huffman(fzp_pipeline* p, int elem_bytes, uint32_t bklen, int* stage_id);
/* HuffmanStage::setEncodeMode: fine = 1 for HuffmanEncodeMode::Fine. */
int fzp_set_huffman_fine(fzp_pipeline* p, int stage_id, int fine);
/* Extension (no counterpart in the reference, which builds the codebook on
 * the host, huffman_stage.h:108): HuffmanStage::setDeviceCodebook -- the
 * codebook is built on the device, so a compress enqueues device work only
 * and graph mode captures it; coarse encoding, bklen <= 1024, the stream is
 * byte for byte that of the host codebook. */
int fzp_set_huffman_device_book(fzp_pipeline* p, int stage_id, int on);
int fzp_add_quantizer(fzp_pipeline* p, float error_bound, int eb_mode, uint32_t quant_radius,
                      float outlier_capacity, int zigzag, int code_bytes, float value_base,
                      int* stage_id);
/* QuantizerStage::setOutlierThreshold and ::setInplaceOutliers
 * (quantizer.h:216-219): |x| >= threshold is always an outlier (INFINITY =
 * off); inplace = 1 stores outliers as raw float bits in u32 zigzag codes
 * (code_bytes 4, zigzag 1) and leaves one output port.  Call before
 * finalize.                                                                */
int fzp_set_quantizer_outliers(fzp_pipeline* p, int stage_id, float threshold, int inplace);
/* nbits 0 = auto-detect from the maximum */
int fzp_add_bitpack(fzp_pipeline* p, int elem_bytes, int nbits, int* stage_id);
/* DifferenceStage<T, TOut> (diff.h:48-200): dtype = FZM DataType code of T
 * (0 u8 1 u16 2 u32 4 i8 5 i16 6 i32 7 i64 8 f32 9 f64); negabinary = 1
 * picks the fused signed->unsigned variant; chunk_bytes 0 = no chunking.  */
int fzp_add_difference(fzp_pipeline* p, int dtype, int negabinary, size_t chunk_bytes, int* stage_id);
/* ZigzagStage / NegabinaryStage<intN, uintN> (zigzag_stage.h, negabinary_stage.h). */
int fzp_add_zigzag(fzp_pipeline* p, int elem_bytes, int* stage_id);
int fzp_add_negabinary(fzp_pipeline* p, int elem_bytes, int* stage_id);
/* LorenzoStage<intN> (lorenzo_stage.h): the integer Lorenzo predictor; its
 * dims come from fzp_set_dims. */
int fzp_add_lorenzo_int(fzp_pipeline* p, int elem_bytes, int* stage_id);

int fzp_connect(fzp_pipeline* p, int dependent_id, int producer_id, const char* output_name);
int fzp_set_dims(fzp_pipeline* p, size_t x, size_t y, size_t z);
int fzp_finalize(fzp_pipeline* p);
int fzp_enable_graph_mode(fzp_pipeline* p, int on);
/* Stage fusion (Bitshuffle folded into RZE); on by default. */
int fzp_set_stage_fusion(fzp_pipeline* p, int on);
int fzp_enable_profiling(fzp_pipeline* p, int on);
int fzp_set_pool_managed_decomp_output(fzp_pipeline* p, int on);
int fzp_max_compressed_size(fzp_pipeline* p, size_t input_bytes, size_t* out);

/* Pool-owned result, valid until the next fzp_compress on this pipeline. */
int fzp_compress(fzp_pipeline* p, void* stream, const void* d_in, size_t in_bytes,
                 void** d_out, size_t* out_bytes);
/* Result is pool-owned if fzp_set_pool_managed_decomp_output(p, 1), else
 * caller-owned (release with fzh_device_free). */
int fzp_decompress(fzp_pipeline* p, void* stream, const void* d_in, size_t in_bytes,
                   void** d_out, size_t* out_bytes);
/* Caller-owned output buffers. */
int fzp_compress_into(fzp_pipeline* p, void* stream, const void* d_in, size_t in_bytes,
                      void* d_out, size_t out_cap, size_t* out_bytes);
int fzp_decompress_into(fzp_pipeline* p, void* stream, const void* d_in, size_t in_bytes,
                        void* d_out, size_t out_cap, size_t* out_bytes);

int fzp_write_file(fzp_pipeline* p, const char* path);
/* Result is caller-owned (fzh_device_free). */
int fzp_decompress_file(const char* path, void* stream, void** d_out, size_t* out_bytes);

/* After a profiled call: dag_ms = device time between the first and last
 * stage (events), host_ms = whole call. */
int fzp_last_perf(fzp_pipeline* p, float* dag_ms, float* host_ms, size_t* in_bytes,
                  size_t* out_bytes);
/* Name and device time of stage i of the last profiled eager compress. */
int fzp_last_stage_perf(fzp_pipeline* p, int i, char* name, size_t name_cap, float* ms);

/* Results of the last compress: pipeline output count and each output's
 * name/size, and the resolved absolute error bound of a Lorenzo/Quantizer
 * stage (REL/NOA resolved on the device). */
int fzp_num_outputs(fzp_pipeline* p, int* n);
int fzp_output_info(fzp_pipeline* p, int i, char* name, size_t name_cap, size_t* bytes);
int fzp_stage_abs_error_bound(fzp_pipeline* p, int stage_id, float* abs_eb);
/* The same bound in double: exact for an FP64 Lorenzo stage, whose bound is
 * a double (lorenzo_quant.h:310). */
int fzp_stage_abs_error_bound_f64(fzp_pipeline* p, int stage_id, double* abs_eb);

#ifdef __cplusplus
}
#endif

#endif /* FZP_H */

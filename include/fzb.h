/* fzb.h — C ABI of the batched compressor (fz::BatchCompressor,
 * include/fz/batch.h).
 *
 * No counterpart in the reference: the reference compresses one field per
 * Pipeline::compress (include/pipeline/compressor.h:175-181).  A batch runs
 * the chain LorenzoQuant -> Bitshuffle -> RZE over thousands of fields of
 * any size and dimensionality with ONE persistent launch per stage, driven
 * by device tables of tiles / chunks, and replays the sequence as a captured
 * hipGraph.  Field i's result is byte-identical to what fz::Pipeline::compress
 * returns for the same field with the same stage settings (stage fusion on,
 * which is the default): the multi-output concat
 *   [u32 n][u64 size x n] -> 16 B, then each output padded to 16 B,
 * outputs in the order rze stream, outlier_errors, outlier_indices,
 * outlier_count (+ patch_values, patch_indices, patch_count in strict mode),
 * which fz::Pipeline::decompress (fzp_decompress) reads back.
 *
 * Every call returns 0 or a negative FZH_E_* code, message in fzh_last_error().
 * A handle belongs to one host thread.
 */
#ifndef FZB_H
#define FZB_H

#include <stddef.h>
#include <stdint.h>

#include "fzh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fzb_batch fzb_batch;

/* One field: a device float32 array (16-byte aligned) of nx*ny*nz values,
 * x fastest.  ny = nz = 1 for 1D, nz = 1 for 2D. */
typedef struct fzb_field {
    const float* d_in;
    uint64_t nx, ny, nz;
} fzb_field;

/* Chain settings, shared by every field of the batch; the meaning and
 * defaults are LorenzoQuantStage::Config's (lorenzo_quant.h:89-109), with
 * BitshuffleStage(16384, elem_width) and RZEStage(16384, 4). */
typedef struct fzb_config {
    float error_bound;      /* user eb */
    int eb_mode;            /* FZP_EB_ABS 0 / FZP_EB_REL 1 / FZP_EB_NOA 2: REL/NOA per field */
    uint32_t quant_radius;  /* 1 .. 32768 */
    float outlier_capacity; /* fraction of the padded element count */
    int zigzag;
    int strict;             /* strict error bound (patch list), as fzp_set_lorenzo_strict */
    float patch_capacity;   /* strict mode: fraction, at least 256 slots */
} fzb_config;

/* Launch geometry of the persistent kernels (changes timing only, never
 * bytes).  0 = library default.  wg_per_cu: workgroups per compute unit of
 * each persistent grid (a grid larger than the work is clamped to one unit
 * per wave); tiles_per_unit: Lorenzo tiles taken per wave step.  Lorenzo
 * runs one launch per tile shape (1D, 2D, 3D fields); tiles_per_unit 1 with
 * a grid covering the tiles (the default) takes its loop-free kernels.
 * (The RZE encoder takes one workgroup per chunk and has no parameter.)    */
typedef struct fzb_launch {
    uint32_t vb_wg_per_cu;
    uint32_t lz_wg_per_cu;
    uint32_t lz_tiles_per_unit;
    uint32_t pack_wg_per_cu;
    uint32_t scatter_wg_per_cu;
} fzb_launch;

int fzb_create(const fzb_config* cfg, fzb_batch** out);
int fzb_destroy(fzb_batch* b);

/* Plans the batch: builds the device tile/chunk tables and allocates the
 * codes, scratch and output buffers (all owned by the handle).  Field data
 * is read at every fzb_compress. */
int fzb_set_fields(fzb_batch* b, const fzb_field* fields, size_t n_fields);

/* Compresses every field (graph replay from the second call on when graph
 * mode is on, the default), synchronises the stream and reads back the
 * per-field sizes.  Overflow of a strict-mode patch list or of a field's
 * outlier list in strict mode is an error, as in LorenzoQuantStage. */
int fzb_compress(fzb_batch* b, void* stream);

/* Device base of the output buffer (owned by the handle, valid until the
 * next fzb_set_fields / fzb_destroy) and its capacity. */
int fzb_output(fzb_batch* b, void** d_base, uint64_t* capacity);
/* Field i of the last compress: byte offset from the base and size. */
int fzb_field_result(fzb_batch* b, size_t i, uint64_t* offset, uint64_t* bytes);
/* Sum of the fields' compressed sizes and of their input bytes. */
int fzb_totals(fzb_batch* b, uint64_t* in_bytes, uint64_t* out_bytes);

int fzb_set_launch(fzb_batch* b, const fzb_launch* launch);
int fzb_get_launch(fzb_batch* b, fzb_launch* launch);
/* Autotuner: times each persistent kernel over its candidate geometries on
 * the planned batch (HIP events, `reps` launches each), keeps the fastest
 * and returns it.  The batch's bytes are the same for every candidate. */
int fzb_autotune(fzb_batch* b, void* stream, int reps, fzb_launch* best);

int fzb_enable_graph_mode(fzb_batch* b, int on);
/* Device time (ms) of each persistent stage in the last profiled compress:
 * names "value_base", "lorenzo", "bitshuffle_rze", "scan", "finalize",
 * "pack", "scatter". */
int fzb_enable_profiling(fzb_batch* b, int on);
int fzb_stage_ms(fzb_batch* b, int i, char* name, size_t name_cap, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* FZB_H */

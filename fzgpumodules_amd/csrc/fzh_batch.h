// fzh_batch.h — device tables and launchers of the batched compressor
// (batch.hip), driven by fz::BatchCompressor (host_batch.cpp).
//
// A batch is flattened into three global index spaces, each with a per-field
// start table (F + 1 entries, strictly increasing):
//   tiles   Lorenzo tiles of every field (1024 / 32x32 / 8x8x8 elements)
//   chunks  16 KiB chunks of every field's codes (Bitshuffle block = RZE chunk)
//   units   VB_UNIT-element slices of every field (value-base reduction)
// Each kernel walks one space -- with a fixed persistent grid, except the
// RZE encoder, which takes one workgroup per chunk -- mapping a global index
// to its field by binary search, and calls the same per-tile / per-chunk
// device functions as the single-field kernels.
#pragma once

#include "fzh_internal.h"

namespace fzh {
namespace batch {

constexpr uint32_t kListSlots = 16;      // per-tile list entries kept by the first pass
constexpr uint32_t kVbUnit = 65536;      // elements per value-base unit
constexpr uint32_t kChunk = 16384;       // Bitshuffle block = RZE chunk (bytes)

// Device table entry of one field (codes are uint16).
struct Field {
    const float* in;     // nv values, 16-byte aligned
    uint64_t n;          // padded element count (input bytes rounded up to 16 KiB)
    uint64_t nv;         // nx * ny * nz
    uint64_t codes_off;  // first code of the field in the codes buffer (elements)
    uint64_t out_off;    // byte offset of the field's result (256-byte aligned)
    uint64_t out_cap;    // bytes reserved for it
    uint32_t nx, ny, nz, ndim;
    uint32_t gx, gy;     // tile grid (2D: gx; 3D: gx, gy)
    uint32_t tile0, ntiles;
    uint32_t chunk0, nchunks;
    uint32_t unit0, nunits;
    uint32_t max_out, max_patch;
    uint32_t pad[2];
};

// Written by the finalize kernel for the pack / scatter kernels.
struct Derived {
    uint64_t o_err, o_idx, o_pval, o_pidx;  // byte offsets inside the field's result
    uint32_t ko, kpo;                       // stored outliers / patches
    uint32_t ex_t, ex_p, ex_c;              // global exclusive prefixes at the field's start
    uint32_t pad;
};

// Read back by the host after every compress.
struct Result {
    uint64_t bytes;  // size of the field's result
    uint32_t k;      // outliers found (may exceed max_out)
    uint32_t kp;     // strict-mode patches found
    uint32_t flags;  // kOverflowOutliers | kOverflowPatches | kOverflowCapacity
    uint32_t pad[3];
};
constexpr uint32_t kOverflowOutliers = 1u, kOverflowPatches = 2u, kOverflowCapacity = 4u;

struct Plan {
    const Field* fields;
    const uint32_t* tile0;   // F + 1
    const uint32_t* chunk0;  // F + 1
    const uint32_t* unit0;   // F + 1
    // Tiles by dimensionality (class d = ndim - 1), so each Lorenzo launch
    // holds one tile shape's registers: the class's fields in field order,
    // the prefix of their tile counts, and the class field holding class
    // tile 64 k (a start for the tile -> field walk).
    const uint32_t* cls_field[3];  // cls_F[d]
    const uint32_t* cls_tile0[3];  // cls_F[d] + 1
    const uint32_t* cls_tfield[3]; // ceil(cls_T[d] / 64)
    uint32_t cls_F[3], cls_T[3];
    // 3D fields' tiles in x-adjacent pairs (tiles (2 bx2, by, bz) and the
    // next, ceil(gx / 2) gy gz pairs per field): prefix of the pair counts
    // over the 3D class fields, and the class field holding pair 64 k.
    const uint32_t* pair0;   // cls_F[2] + 1
    const uint32_t* pfield;  // ceil(P / 64)
    uint32_t P;
    const uint32_t* cfield;  // NC: the field of each chunk
    const uint32_t* ufield;  // U: the field of each value-base unit
    uint32_t F, T, NC, U;
    // chain settings
    float user_eb;
    int mode;                // 0 ABS, 1 REL, 2 NOA
    float rad;               // (float)(uint16_t)radius
    int zigzag, strict;
    uint32_t hdr;            // concat header bytes (48, or 64 in strict mode)
    uint32_t nout;           // 4, or 7 in strict mode
    // buffers
    float* params;           // 4 per field
    float2* vb_part;         // U
    uint16_t* codes;
    float* lerr;             // T * kListSlots
    uint32_t* lidx;
    float* lpval;            // strict only
    uint32_t* lpidx;
    uint32_t* counts;        // [outliers per tile T | patches per tile T (strict) | chunk bytes NC]
    uint32_t* excl;          // exclusive prefixes of counts (+1: the total)
    uint32_t n_counts;
    uint32_t* sizes;         // NC, RZE size words (bit 31 = raw)
    uint8_t* chunk_scratch;  // NC * kChunk
    Derived* derived;        // F
    Result* results;         // F
    uint8_t* out;
    void* scan_ws;
};

size_t scan_ws_bytes_for(size_t n_counts);

// Launchers (grid = number of workgroups of the persistent kernel).  `p` is
// the host copy (grid sizes), `dp` the same Plan in device memory, which the
// kernels read (a kernel-argument copy of the whole Plan would sit in
// scalar registers across the persistent loops).
int launch_value_base(hipStream_t s, const Plan& p, const Plan* dp, uint32_t grid);
int launch_lorenzo(hipStream_t s, const Plan& p, const Plan* dp, uint32_t grid, uint32_t tiles_per_unit);
int launch_rze(hipStream_t s, const Plan& p, const Plan* dp);
int launch_scan(hipStream_t s, const Plan& p);
int launch_finalize(hipStream_t s, const Plan& p, const Plan* dp);
int launch_pack(hipStream_t s, const Plan& p, const Plan* dp, uint32_t grid);
int launch_scatter(hipStream_t s, const Plan& p, const Plan* dp, uint32_t grid);
// Copies results[F] to `dst` (host-mapped memory).
int launch_readback(hipStream_t s, const Plan& p, void* dst);

}  // namespace batch
}  // namespace fzh

// Selections over every row for a few queries (robust.hip): the exact
// all-rows answer of queries a regular step could not bound, and the
// reference GPU contestant's own ranking chain.
#pragma once
#include "common.h"

// Mode ROBUST_EXACT.  A regular step flags a query (FinalArgs::overflow)
// when its candidate list overflowed (bit 0: rows in an adversarial order),
// when more than 2048 rows fell within the error bound of its rank-R key
// (bit 1: e.g. thousands of duplicate vectors), or when an fp32 near-tie run
// at the rank-k boundary was longer than the fp64 window of finalize (bit 2).
// Those queries are answered again with no bounded candidate list:
//   1. d32 of every row in the reference's fp32 order (as finalize computes it)
//   2. the rank-k value t of d32 by a device-wide radix select (11+11+10 bits)
//   3. every row with d32 <= t * (1 + eps) collected: this window holds every
//      row whose exact distance can be among the k smallest
//   4. top-k of the collected rows by (fp64 distance, id), merged in chunks
//      of 2048 rows in LDS, so any number of tied rows up to `cap` is ranked.
// Output: ids in (fp64, id) order -- the exact oracle's -- and the rows' d32.
//
// Mode ROBUST_GPU_FILTER (competitors/gpu_filter/src/gpu_filter.cu:142-176):
//   1. key of every row = the reference's fp32 SQ8 distance, dims in order,
//      v = fma(code, scale, -(q - offset)), acc = fma(v, v, acc) (:52-63)
//   2. the r rows first in the stable ascending sort of (key, row) -- CUB's
//      SortPairs over idx = 0..n-1 (:154-155) -- by a radix select of the
//      64-bit (key, row) value: 6 digits, then every row at or below it
//   3. each re-ranked by the scalar fp32 sum acc += v * v, v = x - q, no
//      fused multiply-add (:161-170), top-k by (that value, row) (:173).
// Output: ids and the re-ranked values.
enum { ROBUST_EXACT = 0, ROBUST_GPU_FILTER = 1 };

struct RobustArgs {
    int mode;
    const float* X32;    // fp32 rows
    int64_t x32_stride;  // floats per row
    int64_t n;
    int d;
    const float* Q32;    // [* x d] queries of the step
    const int32_t* rows; // [F] query (and output) row of each query
    int F;
    int k;               // output columns; min(k, n) (gpu_filter: min(k, r)) written, the rest padded
    int r;               // gpu_filter: rows taken by the key (k <= r <= n)
    // gpu_filter mode: the SQ8 codes (raw bytes, kdim per row, the tiled
    // layout when tiled) and the per-dim scale / offset
    const uint8_t* codes;
    int kdim;
    int tiled;
    const float* scale;
    const float* offset;
    float* keys;         // [F x n] workspace: the key of every row
    uint32_t* hist;      // [F x 2048] workspace (zeroed by launch_robust)
    uint64_t* sel;       // [F x 4] workspace: prefix, mask, rank, threshold
    uint32_t* cnt;       // [F] workspace: collected rows
    uint32_t* cid;       // [F x cap] workspace: collected row ids
    int64_t cap;
    double eps_rel;      // exact mode: relative fp32 error window (finalize's)
    int64_t* out_ids;    // [* x k], row rows[f]
    float* out_d;        // [* x k]
    uint32_t* flags;     // [F] out: bit 0 set when more than cap rows were collected
    uint32_t* n_rescored;// [F] out: rows collected (re-ranked)
};
hipError_t launch_robust(const RobustArgs& a, hipStream_t st);

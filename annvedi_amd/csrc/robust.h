// Exact top-k over every row for the few queries a regular step could not
// answer exactly (robust.hip).
#pragma once
#include "common.h"

// A regular step flags a query (FinalArgs::overflow) when its candidate list
// overflowed (bit 0: rows in an adversarial order), when more than 2048 rows
// fell within the error bound of its rank-R key (bit 1: e.g. thousands of
// duplicate vectors), or when an fp32 near-tie run at the rank-k boundary
// was longer than the fp64 window of finalize (bit 2).  Those queries are
// answered again by this path, which needs no candidate list of bounded
// length:
//   1. d32 of every row in the reference's fp32 order (as finalize computes it)
//   2. the rank-k value t of d32 by a device-wide radix select (11+11+10 bits)
//   3. every row with d32 <= t * (1 + eps) collected: this window holds every
//      row whose exact distance can be among the k smallest
//   4. top-k of the collected rows by (fp64 distance, id), merged in chunks
//      of 2048 rows in LDS, so any number of tied rows up to `cap` is ranked.
// Output: ids in (fp64, id) order -- the exact oracle's -- and the rows' d32.
struct RobustArgs {
    const float* X32;    // fp32 rows
    int64_t x32_stride;  // floats per row
    int64_t n;
    int d;
    const float* Q32;    // [* x d] queries of the step
    const int32_t* rows; // [F] query (and output) row of each robust query
    int F;
    int k;               // output columns; min(k, n) are written, the rest padded
    float* keys;         // [F x n] workspace: d32 of every row
    uint32_t* hist;      // [F x 2048] workspace (zeroed by launch_robust)
    uint32_t* sel;       // [F x 4] workspace: prefix, mask, rank, collection threshold
    uint32_t* cnt;       // [F] workspace: collected rows
    uint32_t* cid;       // [F x cap] workspace: collected row ids
    int cap;
    double eps_rel;      // relative fp32 error window (finalize's)
    int64_t* out_ids;    // [* x k], row rows[f]
    float* out_d;        // [* x k]
    uint32_t* flags;     // [F] out: bit 0 set when more than cap rows fall in the window
    uint32_t* n_rescored;// [F] out: rows collected (re-scored in fp64)
};
hipError_t launch_robust(const RobustArgs& a, hipStream_t st);

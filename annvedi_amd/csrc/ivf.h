// Internal declarations of the IVF kernels (ivf.hip).
#pragma once
#include "common.h"

// Segment of one cluster's points in sorted order (k-means update).
struct KmSeg {
    int64_t pos_begin, pos_end;
};

// Work item of the list scan: rows [row_begin, row_end) of the list-ordered
// codes against queries gq[qoff .. qoff + nq), nq <= 16.
struct IvfItem {
    uint32_t row_begin, row_end, qoff, nq;
};

// Device-side plan of an IVF step, built from the probe ids with no host
// round trip: per list the queries that probe it (gq[qoff[l] ..
// qoff[l] + qcount[l])), and per pass the prefix over lists of the pass's
// work items, list l holding ceil(qcount[l] / 16) query groups x
// ceil(rows of the list in the pass / chunk) row chunks.  Pass 0 takes rows
// [0, S) of every probed list, pass p > 0 rows [S * 8^(p-1), S * 8^p).  A
// query whose probed lists hold fewer than kk rows probes every list.
struct IvfPlanArgs {
    const int64_t* probe;     // [B x nprobe] probed list ids (-1: none)
    int B, nprobe, nlist;
    int kk;
    const int64_t* list_off;  // [nlist + 1]
    uint32_t* qcount;         // [nlist] (zeroed by launch_ivf_plan)
    uint32_t* qoff;           // [nlist]
    uint32_t* qfill;          // [nlist]
    uint32_t* gq;             // [sum of qcount]
    uint8_t* every;           // [B] the query probes every list
    uint32_t* plan;           // [npass x (nlist + 1)]
    uint32_t* nitems;         // [npass]
    uint32_t* coarse;         // [npass x coarse_cap]: list of item 64 j of the pass
    int coarse_cap;
    int npass;
    int64_t S;
    int chunk;
};

// Row bounds [lo, hi) of pass p inside a list (see IvfPlanArgs).
__host__ __device__ inline void ivf_pass_bounds(int64_t S, int p, int64_t& lo, int64_t& hi) {
    lo = 0;
    hi = S;
    for (int i = 0; i < p; ++i) {
        lo = hi;
        hi *= 8;
    }
}

struct IvfScanArgs {
    const void* codes;        // list-ordered SQ8 codes
    int64_t xstride;          // bytes per row
    const float* xn;          // [n] norm term, list order
    const uint32_t* perm;     // list position -> original row id
    int kdim;                 // padded dims (multiple of 64)
    const void* Q;            // [B x kdim] int8 query weights
    const float* qn;
    const float* qscale;
    const float* thr;
    // this pass's items: plan[l] <= item < plan[l + 1] are list l's, in
    // (query group, row chunk) order; *nitems in all
    const uint32_t* plan;
    const uint32_t* nitems;
    const uint32_t* coarse;   // list of item 64 j, j < coarse_cap
    int coarse_cap;
    const int64_t* list_off;
    const uint32_t* qcount;
    const uint32_t* qoff;
    int nlist;
    int64_t lo, hi;           // rows of each list in this pass
    int chunk;                // rows per item
    const uint32_t* gq;       // query ids of the item groups
    uint32_t* next;           // work-queue counter (zeroed before launch)
    CandBuf cand;
};

hipError_t launch_gather_rows_f32(const float* src, int64_t stride, const uint32_t* idx, int64_t n, int kdim,
                                  float* dst, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long* p, unsigned long long v, int64_t n, hipStream_t st);
hipError_t launch_iota_u32(uint32_t* p, int64_t n, hipStream_t st);
hipError_t launch_argmin_to_assign(const unsigned long long* am, int64_t n, int nlist, uint32_t* assign,
                                   uint32_t* counts, hipStream_t st);
hipError_t sort_pairs_u32(void* tmp, size_t& tmp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                          const uint32_t* vals_in, uint32_t* vals_out, int64_t n, int key_bits, hipStream_t st);
hipError_t exclusive_sum_u32(void* tmp, size_t& tmp_bytes, const uint32_t* in, uint32_t* out, int64_t n,
                             hipStream_t st);
hipError_t launch_segment_sums(const float* X, int64_t stride, const uint32_t* perm, const KmSeg* segs, int nseg,
                               int d, double* partial, hipStream_t st);
hipError_t launch_centroid_update(const double* partial, const int* seg_first, const uint32_t* counts, int nlist,
                                  int d, int kdim, const float* X, int64_t stride, const uint32_t* rand_idx,
                                  int64_t n, float* cent, hipStream_t st);
hipError_t launch_ivf_plan(const IvfPlanArgs& a, hipStream_t st);
hipError_t launch_ivf_scan(const IvfScanArgs& a, hipStream_t st);

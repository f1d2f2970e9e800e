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
    const IvfItem* items;
    int nitems;
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
hipError_t launch_ivf_scan(const IvfScanArgs& a, hipStream_t st);

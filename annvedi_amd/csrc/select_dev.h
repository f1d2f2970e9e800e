// Block-wide selection helpers shared by the selection kernels (256-thread
// workgroups; LDS scratch passed in by the caller).
#pragma once
#include "common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kBins = 2048;

// Exclusive scan helper: returns, for the bin histogram h[kBins] in LDS, the
// bin b where the running count first reaches `target` (1-based rank) and the
// count of keys in bins below b.  All threads get the result.
__device__ inline void find_bin(const uint32_t* h, uint32_t target, uint32_t* sh, int* out_bin, uint32_t* out_below) {
    // each thread owns 8 consecutive bins
    const int t = threadIdx.x;
    uint32_t local = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) local += h[t * 8 + j];
    // inclusive scan across 256 threads (wave scan + cross-wave)
    uint32_t v = local;
    const int lane = t & 63;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t u = __shfl_up(v, o);
        if (lane >= o) v += u;
    }
    if (lane == 63) sh[t >> 6] = v;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < (t >> 6); ++w) before += sh[w];
    uint32_t incl = v + before;
    uint32_t excl = incl - local;
    if (excl < target && target <= incl) {
        uint32_t c = excl;
        for (int j = 0; j < 8; ++j) {
            uint32_t hj = h[t * 8 + j];
            if (c + hj >= target) {
                sh[4] = (uint32_t)(t * 8 + j);
                sh[5] = c;
                break;
            }
            c += hj;
        }
    }
    __syncthreads();
    *out_bin = (int)sh[4];
    *out_below = sh[5];
    __syncthreads();
}

// Key of rank `rank` (1-based) among keys[0..n), n >= rank >= 1.
__device__ inline uint32_t radix_select(const float* keys, int n, uint32_t rank, uint32_t* hist, uint32_t* sh) {
    uint32_t prefix = 0, prefix_mask = 0;
    const int shifts[3] = {21, 10, 0};
    const int widths[3] = {11, 11, 10};
    for (int p = 0; p < 3; ++p) {
        for (int b = threadIdx.x; b < kBins; b += kThreads) hist[b] = 0;
        __syncthreads();
        const int sft = shifts[p];
        const uint32_t wmask = (1u << widths[p]) - 1u;
        // four key loads per lane in flight
        constexpr int U = 4;
        for (int i0 = 0; i0 < n; i0 += kThreads * U) {
            uint32_t k[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + u * kThreads + (int)threadIdx.x;
                k[u] = i < n ? f2key(keys[i]) : 0u;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + u * kThreads + (int)threadIdx.x;
                if (i < n && (k[u] & prefix_mask) == prefix) atomicAdd(&hist[(k[u] >> sft) & wmask], 1u);
            }
        }
        __syncthreads();
        int bin;
        uint32_t below;
        find_bin(hist, rank, sh, &bin, &below);
        rank -= below;
        prefix |= (uint32_t)bin << sft;
        prefix_mask |= wmask << sft;
    }
    return prefix;
}

// Key of rank `rank` (1-based) among n ordered keys already in LDS.
__device__ inline uint32_t radix_select_lds(const uint32_t* keys, int n, uint32_t rank, uint32_t* hist, uint32_t* sh) {
    uint32_t prefix = 0, prefix_mask = 0;
    const int shifts[3] = {21, 10, 0};
    const int widths[3] = {11, 11, 10};
    for (int p = 0; p < 3; ++p) {
        for (int b = threadIdx.x; b < kBins; b += kThreads) hist[b] = 0;
        __syncthreads();
        const int sft = shifts[p];
        const uint32_t wmask = (1u << widths[p]) - 1u;
        for (int i = threadIdx.x; i < n; i += kThreads) {
            const uint32_t k = keys[i];
            if ((k & prefix_mask) == prefix) atomicAdd(&hist[(k >> sft) & wmask], 1u);
        }
        __syncthreads();
        int bin;
        uint32_t below;
        find_bin(hist, rank, sh, &bin, &below);
        rank -= below;
        prefix |= (uint32_t)bin << sft;
        prefix_mask |= wmask << sft;
    }
    return prefix;
}

// Bitonic sort of n (power of two) 64-bit keys in LDS, ascending.
__device__ inline void bitonic_sort(uint64_t* v, int n) {
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < n / 2; i += kThreads) {
                int lo = 2 * i - (i & (stride - 1));
                int hi = lo + stride;
                bool asc = (lo & size) == 0;
                uint64_t a = v[lo], b = v[hi];
                if ((a > b) == asc) {
                    v[lo] = b;
                    v[hi] = a;
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace

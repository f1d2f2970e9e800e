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

// fp32 squared distances of NR rows in the reference order, each computed by
// a 16-lane group; every lane of the group returns the values.  Loads are
// issued in batches of 8 per row ahead of their FMAs (the FMA order per lane
// is unchanged: dims l, l+16, l+32, ...).
template <int NR>
__device__ void l2_ref_order(const float* const (&x)[NR], const float* q, int d, int l16, float (&out)[NR]) {
    const int d16 = d & ~15;
    float s[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) s[r] = 0.f;
    int j = l16;
    for (; j + 16 * 7 < d16; j += 16 * 8) {
        float xv[NR][8];
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int u = 0; u < 8; ++u) xv[r][u] = x[r][j + 16 * u];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const float qv = q[j + 16 * u];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const float diff = xv[r][u] - qv;
                s[r] = fmaf(diff, diff, s[r]);
            }
        }
    }
    for (; j < d16; j += 16) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const float diff = x[r][j] - q[j];
            s[r] = fmaf(diff, diff, s[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        // tree: h_j = s_j + s_{j+8}; (h0+h1), (h2+h3), ...; then pairs of pairs
        float v = s[r];
        v += __shfl_xor(v, 8, 16);
        v += __shfl_xor(v, 1, 16);
        v += __shfl_xor(v, 2, 16);
        v += __shfl_xor(v, 4, 16);
        // tail (d % 16 dims): 8-wide block when >= 8 remain, then scalar FMAs
        int i = d16;
        if (d - i >= 8) {
            // separately rounded products and sums (no contraction into FMA)
            float p[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                float diff = x[r][i + t] - q[i + t];
                p[t] = __fmul_rn(diff, diff);
            }
            float g0 = __fadd_rn(p[0], p[4]), g1 = __fadd_rn(p[1], p[5]);
            float g2 = __fadd_rn(p[2], p[6]), g3 = __fadd_rn(p[3], p[7]);
            v = __fadd_rn(v, __fadd_rn(__fadd_rn(g0, g2), __fadd_rn(g1, g3)));
            i += 8;
        }
        for (; i < d; ++i) {
            float diff = x[r][i] - q[i];
            v = fmaf(diff, diff, v);
        }
        out[r] = v;
    }
}

// Exact squared L2 in fp64, sequential over dims, no FMA (every term
// (x - q)^2 of two floats is exact in fp64; only the sums round).  Loads are
// issued 16 at a time ahead of the sums.
__device__ inline double l2_f64(const float* x, const float* q, int d) {
    double acc = 0.0;
    int j = 0;
    for (; j + 16 <= d; j += 16) {
        float xv[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) xv[u] = x[j + u];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            double diff = (double)xv[u] - (double)q[j + u];
            acc = __dadd_rn(acc, __dmul_rn(diff, diff));
        }
    }
    for (; j < d; ++j) {
        double diff = (double)x[j] - (double)q[j];
        acc = __dadd_rn(acc, __dmul_rn(diff, diff));
    }
    return acc;
}

}  // namespace

// Bit-exact GPU reproduction of the reference's exact scan, including its
// early-abandon ranks: src/linear_index.cpp:27-121 with src/simd.h:32-64.
//
// The reference walks rows in order keeping the `limit = min(k, 64)` smallest
// values seen; each row's distance is abandoned at the first checkpoint
// (every 96 dims) whose partial sum reaches the current threshold, and the
// partial, not the full sum, is what it stores and ranks.  Rows enter the
// kept set exactly when their full value is below the threshold, so the
// threshold in effect at row i is the limit-th smallest full value among
// rows 0..i-1 -- a prefix statistic, computed here in parallel:
//
//   li_rows      every row's fp32 partials at the checkpoints, its pre-tail
//                and its full value, in the reference's summation order;
//   li_block_top 64 smallest full values of each block of 4096 rows;
//   li_prefix    one wave: running 64 smallest before each block (bitonic
//                merge of sorted lists);
//   li_walk      one wave per block: replays the insertions of the few rows
//                that can still enter the kept set, in row order, giving the
//                threshold each row saw, hence the value the reference
//                stored; block-local top-k of (value, id);
//   li_final     top-k of the block candidates by (value, id)
//                (linear_index.cpp:111 partial_sort of pair<float,int>).
#include <cfloat>

#include "common.h"
#include "select_dev.h"

namespace {

constexpr int kBlockRows = 4096;

__device__ inline float red16(float s) {
    s += __shfl_xor(s, 8, 16);
    s += __shfl_xor(s, 1, 16);
    s += __shfl_xor(s, 2, 16);
    s += __shfl_xor(s, 4, 16);
    return s;
}

struct LiArgs {
    const float* X;
    int64_t xstride;  // floats per row
    int64_t n;
    int d, nck, W;    // checkpoints, floats per row of P (nck + 2)
    const float* q;   // device query
    float* P;         // [n x W]
    int limit;        // min(k, 64)
    int G;            // blocks of kBlockRows rows
    float* BT;        // [G x 64]
    float* R;         // [G x 64]
    float* bkey;      // [G x k]
    uint32_t* bid;    // [G x k]
    int k;
};

__global__ __launch_bounds__(256) void li_rows(LiArgs a) {
    const int64_t row = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int l = threadIdx.x & 15;
    if (row >= a.n) return;  // uniform per 16-lane group
    const float* x = a.X + row * a.xstride;
    const float* q = a.q;
    float* P = a.P + row * a.W;
    float s = 0.f;
    int c = 0;
    int i = 0;
    for (; i + 15 < a.d; i += 16) {
        float diff = x[i + l] - q[i + l];
        s = fmaf(diff, diff, s);
        if (i % 96 == 80) {
            float t = red16(s);
            if (l == 0) P[c] = t;
            ++c;
        }
    }
    const float pt = red16(s);
    float f = pt;
    if (a.d - i >= 8) {
        float p[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            float diff = x[i + j] - q[i + j];
            p[j] = diff * diff;
        }
        float g0 = p[0] + p[4], g1 = p[1] + p[5], g2 = p[2] + p[6], g3 = p[3] + p[7];
        f = f + ((g0 + g2) + (g1 + g3));
        i += 8;
    }
    for (; i < a.d; ++i) {
        float diff = x[i] - q[i];
        f = fmaf(diff, diff, f);
    }
    if (l == 0) {
        P[a.nck] = pt;
        P[a.nck + 1] = f;
    }
}

// The value compute_l2_distance returns under `thr` (simd.h:44-56): the
// first checkpoint partial >= thr, else the pre-tail sum if >= thr, else
// the full value.
__device__ inline float stored_value(const float* P, int nck, float thr) {
    for (int c = 0; c < nck; ++c)
        if (P[c] >= thr) return P[c];
    if (P[nck] >= thr) return P[nck];
    return P[nck + 1];
}

template <class T>
__device__ void bitonic_sort_t(T* v, int n, int nthreads) {
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < n / 2; i += nthreads) {
                int lo = 2 * i - (i & (stride - 1));
                int hi = lo + stride;
                bool asc = (lo & size) == 0;
                T x = v[lo], y = v[hi];
                if ((x > y) == asc) {
                    v[lo] = y;
                    v[hi] = x;
                }
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(256) void li_block_top(LiArgs a) {
    __shared__ uint32_t v[kBlockRows];
    const int64_t r0 = (int64_t)blockIdx.x * kBlockRows;
    for (int j = threadIdx.x; j < kBlockRows; j += 256) {
        const int64_t row = r0 + j;
        v[j] = row < a.n ? f2key(a.P[row * a.W + a.nck + 1]) : f2key(__builtin_inff());
    }
    __syncthreads();
    bitonic_sort_t(v, kBlockRows, 256);
    if (threadIdx.x < 64) a.BT[(int64_t)blockIdx.x * 64 + threadIdx.x] = key2f(v[threadIdx.x]);
}

// One wave: R[b] = 64 smallest full values of rows before block b (FLT_MAX
// where fewer, as the reference's slots start at FLT_MAX, linear_index.cpp:35).
__global__ __launch_bounds__(64) void li_prefix(LiArgs a) {
    const int lane = threadIdx.x;
    float r = FLT_MAX;
    for (int b = 0; b < a.G; ++b) {
        a.R[(int64_t)b * 64 + lane] = r;
        float x = a.BT[(int64_t)b * 64 + 63 - lane];
        r = fminf(r, x);  // ascending vs descending: a bitonic sequence holding the 64 smallest
        for (int s = 32; s > 0; s >>= 1) {
            float o = __shfl_xor(r, s);
            r = (lane & s) ? fmaxf(r, o) : fminf(r, o);
        }
    }
}

__global__ __launch_bounds__(64) void li_walk(LiArgs a) {
    __shared__ unsigned long long sk[kBlockRows];
    const int lane = threadIdx.x;
    const int b = blockIdx.x;
    const int64_t r0 = (int64_t)b * kBlockRows;
    const int64_t r1 = r0 + kBlockRows < a.n ? r0 + kBlockRows : a.n;
    float r = a.R[(int64_t)b * 64 + lane];  // kept values, ascending
    float thr = __shfl(r, a.limit - 1);
    for (int64_t base = r0; base < r0 + kBlockRows; base += 64) {
        const int64_t row = base + lane;
        const bool ok = row < r1;
        const float f = ok ? a.P[row * a.W + a.nck + 1] : __builtin_inff();
        uint64_t mask = __ballot(ok && f < thr);
        float my_thr = thr;
        while (mask) {
            const int c = __builtin_ctzll(mask);
            mask &= mask - 1;
            const float fc = __shfl(f, c);
            if (fc < thr) {  // linear_index.cpp:50-58 sorted insertion
                const int p = __popcll(__ballot(r <= fc));
                const float up = __shfl_up(r, 1);
                r = lane < p ? r : (lane == p ? fc : up);
                thr = __shfl(r, a.limit - 1);
                if (lane > c) my_thr = thr;
            }
        }
        unsigned long long key = ~0ull;
        if (ok) {
            const float s = stored_value(a.P + row * a.W, a.nck, my_thr);
            key = ((unsigned long long)f2key(s) << 32) | (uint32_t)row;
        }
        sk[base - r0 + lane] = key;
    }
    __syncthreads();
    bitonic_sort_t(sk, kBlockRows, 64);
    for (int j = lane; j < a.k; j += 64) {
        const unsigned long long kv = j < kBlockRows ? sk[j] : ~0ull;
        const int64_t o = (int64_t)b * a.k + j;
        a.bkey[o] = kv == ~0ull ? __builtin_inff() : key2f((uint32_t)(kv >> 32));
        a.bid[o] = (uint32_t)kv;
    }
}

constexpr int kFinalMax = 2048;

__global__ __launch_bounds__(256) void li_final(LiArgs a, int kk, int64_t* out_ids, float* out_vals, uint32_t* flag) {
    __shared__ uint32_t hist[kBins];
    __shared__ uint32_t sh[8];
    __shared__ unsigned long long col[kFinalMax];
    __shared__ uint32_t ncol;
    const int m = a.G * a.k;
    const uint32_t pk = radix_select(a.bkey, m, (uint32_t)kk, hist, sh);
    const float pv = key2f(pk);
    if (threadIdx.x == 0) ncol = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += 256) {
        const float v = a.bkey[i];
        if (v <= pv) {
            uint32_t pos = atomicAdd(&ncol, 1u);
            if (pos < kFinalMax) col[pos] = ((unsigned long long)f2key(v) << 32) | a.bid[i];
        }
    }
    __syncthreads();
    uint32_t nc = ncol < kFinalMax ? ncol : kFinalMax;
    if (threadIdx.x == 0 && ncol > kFinalMax) *flag = 1u;
    int np = 1;
    while (np < (int)nc) np <<= 1;
    for (int i = (int)nc + threadIdx.x; i < np; i += 256) col[i] = ~0ull;
    __syncthreads();
    bitonic_sort_t(col, np, 256);
    for (int j = threadIdx.x; j < kk; j += 256) {
        out_ids[j] = (int64_t)(uint32_t)col[j];
        out_vals[j] = key2f((uint32_t)(col[j] >> 32));
    }
}

}  // namespace

// Workspace: P [n x (nck+2)], BT/R [G x 64], bkey/bid [G x k]; all device.
hipError_t launch_linear_index_query(const float* X, int64_t xstride, int64_t n, int d, const float* dq, int k,
                                     float* P, float* BT, float* R, float* bkey, uint32_t* bid, int64_t* d_ids,
                                     float* d_vals, uint32_t* d_flag, hipStream_t st) {
    LiArgs a{};
    a.X = X;
    a.xstride = xstride;
    a.n = n;
    a.d = d;
    a.nck = d / 96;
    a.W = a.nck + 2;
    a.q = dq;
    a.P = P;
    a.limit = k < 64 ? k : 64;
    a.G = (int)((n + kBlockRows - 1) / kBlockRows);
    a.BT = BT;
    a.R = R;
    a.bkey = bkey;
    a.bid = bid;
    a.k = k;
    const int kk = (int64_t)k < n ? k : (int)n;
    hipLaunchKernelGGL(li_rows, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, st, a);
    hipLaunchKernelGGL(li_block_top, dim3((unsigned)a.G), dim3(256), 0, st, a);
    hipLaunchKernelGGL(li_prefix, dim3(1), dim3(64), 0, st, a);
    hipLaunchKernelGGL(li_walk, dim3((unsigned)a.G), dim3(64), 0, st, a);
    hipLaunchKernelGGL(li_final, dim3(1), dim3(256), 0, st, a, kk, d_ids, d_vals, d_flag);
    return hipGetLastError();
}

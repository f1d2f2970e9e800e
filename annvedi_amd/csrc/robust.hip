// Exact top-k over every row for queries a regular step flagged (robust.h).
//
// Rare path: the regular pipeline bounds its candidate lists (65536 keys per
// query) and its re-score set (2048 rows); adversarial row orders, thousands
// of duplicate vectors or long fp32 near-tie runs exceed those bounds.  This
// path ranks a query against all n rows with no such bound: the fp32 value
// of every row (reference summation order, the value finalize reports), a
// device-wide radix select of the rank-k value, a collection of the rows
// within the fp32 error window of it, and a chunked fp64 ranking by
// (distance, id) -- the order src/linear_index.cpp:110-118 defines with the
// value taken as the exact distance.
#include "robust.h"
#include "select_dev.h"

namespace {

constexpr int kChunk = 2048;      // collected rows ranked per LDS merge
constexpr int kMerge = 2 * kChunk;

__global__ __launch_bounds__(kThreads) void rb_init(RobustArgs a) {
    const int f = blockIdx.x;
    for (int b = threadIdx.x; b < kBins; b += kThreads) a.hist[(int64_t)f * kBins + b] = 0u;
    if (threadIdx.x == 0) {
        a.sel[f * 4 + 0] = 0u;  // prefix
        a.sel[f * 4 + 1] = 0u;  // mask
        a.sel[f * 4 + 2] = (uint32_t)((int64_t)a.k < a.n ? (int64_t)a.k : a.n);  // rank (1-based)
        a.sel[f * 4 + 3] = 0u;
        a.cnt[f] = 0u;
    }
}

// d32 of every row: one 16-lane group per row, two rows at a time.
__global__ __launch_bounds__(kThreads) void rb_keys(RobustArgs a) {
    extern __shared__ float qs[];
    const int f = blockIdx.y;
    const float* q = a.Q32 + (int64_t)a.rows[f] * a.d;
    for (int j = threadIdx.x; j < a.d; j += kThreads) qs[j] = q[j];
    __syncthreads();
    const int grp = threadIdx.x >> 4, l16 = threadIdx.x & 15;
    constexpr int G = kThreads / 16;
    float* keys = a.keys + (int64_t)f * a.n;
    for (int64_t r0 = (int64_t)blockIdx.x * 2 * G + grp; r0 < a.n; r0 += (int64_t)gridDim.x * 2 * G) {
        const int64_t r1 = r0 + G < a.n ? r0 + G : r0;
        const float* x[2] = {a.X32 + r0 * a.x32_stride, a.X32 + r1 * a.x32_stride};
        float v[2];
        l2_ref_order<2>(x, qs, a.d, l16, v);
        if (l16 == 0) {
            keys[r0] = v[0];
            if (r1 != r0) keys[r1] = v[1];
        }
    }
}

// One radix pass: histogram of the next digit of the keys that match the
// prefix so far, LDS first, then added into the query's global histogram.
__global__ __launch_bounds__(kThreads) void rb_hist(RobustArgs a, int sft, int width) {
    __shared__ uint32_t h[kBins];
    const int f = blockIdx.y;
    for (int b = threadIdx.x; b < kBins; b += kThreads) h[b] = 0u;
    __syncthreads();
    const uint32_t prefix = a.sel[f * 4 + 0], mask = a.sel[f * 4 + 1];
    const uint32_t wmask = (1u << width) - 1u;
    const float* keys = a.keys + (int64_t)f * a.n;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kThreads) {
        const uint32_t k = f2key(keys[i]);
        if ((k & mask) == prefix) atomicAdd(&h[(k >> sft) & wmask], 1u);
    }
    __syncthreads();
    uint32_t* g = a.hist + (int64_t)f * kBins;
    for (int b = threadIdx.x; b < kBins; b += kThreads)
        if (h[b]) atomicAdd(&g[b], h[b]);
}

// The digit holding the rank; clears the histogram for the next pass.  After
// the last pass the prefix is the rank-k key and the collection threshold
// t * (1 + eps) (rounded up) is stored.
__global__ __launch_bounds__(kThreads) void rb_pick(RobustArgs a, int sft, int width, int last) {
    __shared__ uint32_t h[kBins];
    __shared__ uint32_t sh[8];
    const int f = blockIdx.x;
    uint32_t* g = a.hist + (int64_t)f * kBins;
    for (int b = threadIdx.x; b < kBins; b += kThreads) {
        h[b] = g[b];
        g[b] = 0u;
    }
    __syncthreads();
    int bin;
    uint32_t below;
    find_bin(h, a.sel[f * 4 + 2], sh, &bin, &below);
    if (threadIdx.x == 0) {
        const uint32_t wmask = (1u << width) - 1u;
        const uint32_t prefix = a.sel[f * 4 + 0] | ((uint32_t)bin << sft);
        a.sel[f * 4 + 0] = prefix;
        a.sel[f * 4 + 1] |= wmask << sft;
        a.sel[f * 4 + 2] -= below;
        if (last) {
            const float t = key2f(prefix);
            float tc = t;
            if (t == t && t < __builtin_inff()) {
                tc = (float)((double)t * (1.0 + a.eps_rel));
                tc = next_up(tc < t ? t : tc);
            }
            a.sel[f * 4 + 3] = __builtin_bit_cast(uint32_t, tc);
        }
    }
}

__global__ __launch_bounds__(kThreads) void rb_collect(RobustArgs a) {
    const int f = blockIdx.y;
    const float tc = __builtin_bit_cast(float, a.sel[f * 4 + 3]);
    const float* keys = a.keys + (int64_t)f * a.n;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kThreads) {
        if (keys[i] <= tc) {
            const uint32_t pos = atomicAdd(&a.cnt[f], 1u);
            if (pos < (uint32_t)a.cap) a.cid[(int64_t)f * a.cap + pos] = (uint32_t)i;
        }
    }
}

// Ascending bitonic sort of n (power of two) entries by (k64, id), moving v.
__device__ void bitonic_pairs(uint64_t* k64, uint32_t* id, float* v, int n) {
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < n / 2; i += kThreads) {
                const int lo = 2 * i - (i & (stride - 1));
                const int hi = lo + stride;
                const bool asc = (lo & size) == 0;
                const uint64_t ka = k64[lo], kb = k64[hi];
                const uint32_t ia = id[lo], ib = id[hi];
                const bool gt = ka > kb || (ka == kb && ia > ib);
                if (gt == asc) {
                    k64[lo] = kb;
                    k64[hi] = ka;
                    id[lo] = ib;
                    id[hi] = ia;
                    const float t = v[lo];
                    v[lo] = v[hi];
                    v[hi] = t;
                }
            }
            __syncthreads();
        }
    }
}

// Top-k of the collected rows by (fp64 distance, id): chunks of up to 2048
// rows are merged into the running top-k in LDS.
__global__ __launch_bounds__(kThreads) void rb_final(RobustArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    uint64_t* k64 = reinterpret_cast<uint64_t*>(smem);
    uint32_t* id = reinterpret_cast<uint32_t*>(smem + kMerge * sizeof(uint64_t));
    float* v = reinterpret_cast<float*>(smem + kMerge * (sizeof(uint64_t) + sizeof(uint32_t)));
    float* qs = v + kMerge;
    const int f = blockIdx.x;
    const int64_t row = a.rows[f];
    for (int j = threadIdx.x; j < a.d; j += kThreads) qs[j] = a.Q32[row * a.d + j];
    const int kk = (int)((int64_t)a.k < a.n ? (int64_t)a.k : a.n);
    uint32_t total = a.cnt[f];
    uint32_t flag = 0u;
    if (total > (uint32_t)a.cap) {
        flag = 1u;
        total = (uint32_t)a.cap;
    }
    const float* keys = a.keys + (int64_t)f * a.n;
    const uint32_t* cid = a.cid + (int64_t)f * a.cap;
    __syncthreads();
    int r = 0;  // entries [0, r) hold the running top-k, sorted
    for (uint32_t c0 = 0; c0 < total; c0 += kChunk) {
        const int m = (int)(total - c0 < (uint32_t)kChunk ? total - c0 : (uint32_t)kChunk);
        for (int j = threadIdx.x; j < m; j += kThreads) {
            const uint32_t rid = cid[c0 + j];
            const double d64 = l2_f64(a.X32 + (int64_t)rid * a.x32_stride, qs, a.d);
            k64[r + j] = __builtin_bit_cast(uint64_t, d64);  // non-negative: bits order as values
            id[r + j] = rid;
            v[r + j] = keys[rid];
        }
        int np = 1;
        while (np < r + m) np <<= 1;
        for (int j = r + m + (int)threadIdx.x; j < np; j += kThreads) {
            k64[j] = ~0ull;
            id[j] = ~0u;
            v[j] = __builtin_inff();
        }
        __syncthreads();
        bitonic_pairs(k64, id, v, np);
        r = r + m < kk ? r + m : kk;
    }
    int64_t* oid = a.out_ids + row * a.k;
    float* od = a.out_d + row * a.k;
    for (int j = threadIdx.x; j < a.k; j += kThreads) {
        if (j < r) {
            oid[j] = (int64_t)id[j];
            od[j] = v[j];
        } else {
            oid[j] = -1;
            od[j] = __builtin_inff();
        }
    }
    if (threadIdx.x == 0) {
        a.flags[f] = flag;
        a.n_rescored[f] = total;
    }
}

}  // namespace

size_t robust_final_smem_bytes(int d) {
    return (size_t)kMerge * (sizeof(uint64_t) + sizeof(uint32_t) + sizeof(float)) + (size_t)d * sizeof(float);
}

hipError_t launch_robust(const RobustArgs& a, hipStream_t st) {
    if (a.F <= 0 || a.n <= 0) return hipSuccess;
    const int64_t gb = (a.n + 4095) / 4096;
    const unsigned gx = (unsigned)(gb < 1024 ? gb : 1024);
    static const bool attr = [] {  // the merge buffers take more than the default 64 KB of LDS
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&rb_final), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        return true;
    }();
    (void)attr;
    ANNMI_LAUNCH(rb_init, dim3((unsigned)a.F), dim3(kThreads), 0, st, a);
    ANNMI_LAUNCH(rb_keys, dim3(gx, (unsigned)a.F), dim3(kThreads), (size_t)a.d * sizeof(float), st, a);
    const int sft[3] = {21, 10, 0}, wid[3] = {11, 11, 10};
    for (int p = 0; p < 3; ++p) {
        ANNMI_LAUNCH(rb_hist, dim3(gx, (unsigned)a.F), dim3(kThreads), 0, st, a, sft[p], wid[p]);
        ANNMI_LAUNCH(rb_pick, dim3((unsigned)a.F), dim3(kThreads), 0, st, a, sft[p], wid[p], p == 2 ? 1 : 0);
    }
    ANNMI_LAUNCH(rb_collect, dim3(gx, (unsigned)a.F), dim3(kThreads), 0, st, a);
    ANNMI_LAUNCH(rb_final, dim3((unsigned)a.F), dim3(kThreads), robust_final_smem_bytes(a.d), st, a);
    return hipGetLastError();
}

// Selections over every row for a few queries (robust.h).
//
// Rare paths with no bounded candidate list: the exact answer of queries the
// regular pipeline flagged (adversarial row orders, thousands of duplicate
// vectors, long fp32 near-tie runs), and the reference GPU contestant's
// ranking chain for its drop-in entry point (gf_query).  Both rank a query
// against all n rows: a key per row, a device-wide radix select of the
// rank-k (or rank-r) value over the 64-bit (key, row) image, a collection of
// the rows at or below it, and a chunked ranking in LDS by (value, row).
#include "robust.h"
#include "select_dev.h"

namespace {

constexpr int kChunk = 2048;      // collected rows ranked per LDS merge
constexpr int kMerge = 2 * kChunk;

// 64-bit ordered image of (key, row): ascending order of (key, row).
__device__ inline uint64_t comp_key(float key, int64_t row) {
    return ((uint64_t)f2key(key) << 32) | (uint64_t)(uint32_t)row;
}

__device__ inline int64_t rank_of(const RobustArgs& a) {
    const int64_t want = a.mode == ROBUST_GPU_FILTER ? (int64_t)a.r : (int64_t)a.k;
    return want < a.n ? want : a.n;
}

__global__ __launch_bounds__(kThreads) void rb_init(RobustArgs a) {
    const int f = blockIdx.x;
    for (int b = threadIdx.x; b < kBins; b += kThreads) a.hist[(int64_t)f * kBins + b] = 0u;
    if (threadIdx.x == 0) {
        a.sel[f * 4 + 0] = 0ull;  // prefix
        a.sel[f * 4 + 1] = 0ull;  // mask
        a.sel[f * 4 + 2] = (uint64_t)rank_of(a);  // rank (1-based)
        a.sel[f * 4 + 3] = 0ull;
        a.cnt[f] = 0u;
    }
}

// Exact mode: d32 of every row, one 16-lane group per row, two at a time.
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

// gpu_filter mode: the reference's SQ8 key of every row (gpu_filter.cu:52-63),
// one row per thread, dims in order; the 64 code bytes of a chunk are read
// as four 16-byte loads (contiguous across the 16 rows of a tile).
__global__ __launch_bounds__(kThreads) void rb_gf_keys(RobustArgs a) {
    extern __shared__ float sm[];
    float* qsh = sm;          // q - offset (gpu_filter.cu:148-149)
    float* sc = sm + a.d;     // scale
    const int f = blockIdx.y;
    const float* q = a.Q32 + (int64_t)a.rows[f] * a.d;
    for (int j = threadIdx.x; j < a.d; j += kThreads) {
        qsh[j] = __fsub_rn(q[j], a.offset[j]);
        sc[j] = a.scale[j];
    }
    __syncthreads();
    float* keys = a.keys + (int64_t)f * a.n;
    const int64_t stride = a.kdim;  // bytes per row
    for (int64_t row = (int64_t)blockIdx.x * kThreads + threadIdx.x; row < a.n; row += (int64_t)gridDim.x * kThreads) {
        float acc = 0.f;
        for (int c0 = 0; c0 < a.d; c0 += 64) {
            const int64_t off = a.tiled ? tiled_off(row, c0, stride, 64) : row * stride + c0;
            const uint4* p = reinterpret_cast<const uint4*>(a.codes + off);
            uint4 w[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) w[u] = p[u];
            const uint8_t* b = reinterpret_cast<const uint8_t*>(w);
            const int m = a.d - c0 < 64 ? a.d - c0 : 64;
            const uint8_t flip = a.tiled ? 0x80 : 0;  // the tiled copy stores c ^ 0x80
            for (int j = 0; j < m; ++j) {
                const float v = fmaf((float)(uint8_t)(b[j] ^ flip), sc[c0 + j], -qsh[c0 + j]);
                acc = fmaf(v, v, acc);
            }
        }
        keys[row] = acc;
    }
}

// One radix pass over the (key, row) image: histogram of the digit at
// [sft, sft + width) of the values that match the prefix so far, LDS first,
// then added into the query's global histogram.
__global__ __launch_bounds__(kThreads) void rb_hist(RobustArgs a, int sft, int width) {
    __shared__ uint32_t h[kBins];
    const int f = blockIdx.y;
    for (int b = threadIdx.x; b < kBins; b += kThreads) h[b] = 0u;
    __syncthreads();
    const uint64_t prefix = a.sel[f * 4 + 0], mask = a.sel[f * 4 + 1];
    const uint64_t wmask = (1ull << width) - 1ull;
    const float* keys = a.keys + (int64_t)f * a.n;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kThreads) {
        const uint64_t c = comp_key(keys[i], i);
        if ((c & mask) == prefix) atomicAdd(&h[(uint32_t)((c >> sft) & wmask)], 1u);
    }
    __syncthreads();
    uint32_t* g = a.hist + (int64_t)f * kBins;
    for (int b = threadIdx.x; b < kBins; b += kThreads)
        if (h[b]) atomicAdd(&g[b], h[b]);
}

// The digit holding the rank; clears the histogram for the next pass.  After
// the last pass: exact mode stores the collection threshold t * (1 + eps)
// (rounded up) of the rank-k key t; gpu_filter mode the rank-r (key, row)
// value itself.
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
    find_bin(h, (uint32_t)a.sel[f * 4 + 2], sh, &bin, &below);
    if (threadIdx.x == 0) {
        const uint64_t wmask = (1ull << width) - 1ull;
        const uint64_t prefix = a.sel[f * 4 + 0] | ((uint64_t)bin << sft);
        a.sel[f * 4 + 0] = prefix;
        a.sel[f * 4 + 1] |= wmask << sft;
        a.sel[f * 4 + 2] -= below;
        if (last) {
            if (a.mode == ROBUST_GPU_FILTER) {
                a.sel[f * 4 + 3] = prefix;
            } else {
                const float t = key2f((uint32_t)(prefix >> 32));
                float tc = t;
                if (t == t && t < __builtin_inff()) {
                    tc = (float)((double)t * (1.0 + a.eps_rel));
                    tc = next_up(tc < t ? t : tc);
                }
                a.sel[f * 4 + 3] = (uint64_t)__builtin_bit_cast(uint32_t, tc);
            }
        }
    }
}

__global__ __launch_bounds__(kThreads) void rb_collect(RobustArgs a) {
    const int f = blockIdx.y;
    const uint64_t lim = a.sel[f * 4 + 3];
    const float tc = __builtin_bit_cast(float, (uint32_t)lim);
    const bool gf = a.mode == ROBUST_GPU_FILTER;
    const float* keys = a.keys + (int64_t)f * a.n;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kThreads) {
        const bool take = gf ? comp_key(keys[i], i) <= lim : keys[i] <= tc;
        if (take) {
            const uint32_t pos = atomicAdd(&a.cnt[f], 1u);
            if ((int64_t)pos < a.cap) a.cid[(int64_t)f * a.cap + pos] = (uint32_t)i;
        }
    }
}

// gpu_filter.cu:161-170: acc += v * v in float, dims in order, the product
// rounded on its own (no fused multiply-add).
__device__ inline float l2_scalar_f32(const float* x, const float* q, int d) {
    float acc = 0.f;
    for (int j = 0; j < d; ++j) {
        const float v = __fsub_rn(x[j], q[j]);
        acc = __fadd_rn(acc, __fmul_rn(v, v));
    }
    return acc;
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

// Top-k of the collected rows, chunks of up to 2048 rows merged into the
// running top-k in LDS: exact mode by (fp64 distance, row) reporting d32;
// gpu_filter mode by (scalar fp32 value, row) reporting that value.
__global__ __launch_bounds__(kThreads) void rb_final(RobustArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    uint64_t* k64 = reinterpret_cast<uint64_t*>(smem);
    uint32_t* id = reinterpret_cast<uint32_t*>(smem + kMerge * sizeof(uint64_t));
    float* v = reinterpret_cast<float*>(smem + kMerge * (sizeof(uint64_t) + sizeof(uint32_t)));
    float* qs = v + kMerge;
    const bool gf = a.mode == ROBUST_GPU_FILTER;
    const int f = blockIdx.x;
    const int64_t row = a.rows[f];
    for (int j = threadIdx.x; j < a.d; j += kThreads) qs[j] = a.Q32[row * a.d + j];
    const int64_t lim = gf ? ((int64_t)a.r < a.n ? (int64_t)a.r : a.n) : a.n;
    const int kk = (int)((int64_t)a.k < lim ? (int64_t)a.k : lim);
    uint32_t total = a.cnt[f];
    uint32_t flag = 0u;
    if ((int64_t)total > a.cap) {
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
            const float* x = a.X32 + (int64_t)rid * a.x32_stride;
            if (gf) {
                const float dv = l2_scalar_f32(x, qs, a.d);
                k64[r + j] = (uint64_t)f2key(dv);
                v[r + j] = dv;
            } else {
                const double d64 = l2_f64(x, qs, a.d);
                k64[r + j] = __builtin_bit_cast(uint64_t, d64);  // non-negative: bits order as values
                v[r + j] = keys[rid];
            }
            id[r + j] = rid;
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
    if (a.mode == ROBUST_GPU_FILTER && (!a.codes || (a.kdim & 63) || a.d > a.kdim || a.r < 1))
        return hipErrorInvalidValue;
    const int64_t gb = (a.n + 4095) / 4096;
    const unsigned gx = (unsigned)(gb < 1024 ? gb : 1024);
    static const bool attr = [] {  // the merge buffers take more than the default 64 KB of LDS
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&rb_final), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        return true;
    }();
    (void)attr;
    ANNMI_LAUNCH(rb_init, dim3((unsigned)a.F), dim3(kThreads), 0, st, a);
    if (a.mode == ROBUST_GPU_FILTER) {
        const int64_t gr = (a.n + kThreads - 1) / kThreads;
        const unsigned gk = (unsigned)(gr < 4096 ? gr : 4096);
        ANNMI_LAUNCH(rb_gf_keys, dim3(gk, (unsigned)a.F), dim3(kThreads), (size_t)a.d * 2 * sizeof(float), st, a);
    } else {
        ANNMI_LAUNCH(rb_keys, dim3(gx, (unsigned)a.F), dim3(kThreads), (size_t)a.d * sizeof(float), st, a);
    }
    // digits of the 64-bit (key, row) image, most significant first: the key
    // alone (three digits) for the exact mode's window, key and row (six)
    // for gpu_filter's exact rank-r cut
    const int sft[6] = {53, 42, 32, 21, 10, 0}, wid[6] = {11, 11, 10, 11, 11, 10};
    const int passes = a.mode == ROBUST_GPU_FILTER ? 6 : 3;
    for (int p = 0; p < passes; ++p) {
        ANNMI_LAUNCH(rb_hist, dim3(gx, (unsigned)a.F), dim3(kThreads), 0, st, a, sft[p], wid[p]);
        ANNMI_LAUNCH(rb_pick, dim3((unsigned)a.F), dim3(kThreads), 0, st, a, sft[p], wid[p], p == passes - 1 ? 1 : 0);
    }
    ANNMI_LAUNCH(rb_collect, dim3(gx, (unsigned)a.F), dim3(kThreads), 0, st, a);
    ANNMI_LAUNCH(rb_final, dim3((unsigned)a.F), dim3(kThreads), robust_final_smem_bytes(a.d), st, a);
    return hipGetLastError();
}

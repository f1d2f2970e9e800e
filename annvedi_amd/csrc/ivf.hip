// IVF index: k-means build kernels and the list-scan kernel.
//
// Build (competitors/ivf_lsh/src/lsh_index_optimized.cpp:551-616, reorder
// :155-173): centroids start at rows of a std::shuffle(iota, mt19937(1337))
// permutation (done on the host), 15 rounds of assign + mean update, an
// empty cluster restarts at row rand_indices[(c*37) % n].  Assignment is the
// argmin scan (scan.hip, argmin mode: smallest fp32 distance, then lowest
// centroid id).  The update is deterministic: points are sorted by (cluster,
// id) with a stable radix sort, each cluster's rows are summed in that order
// in fp64 over fixed segments, segments are summed in order.  The list layout
// is the sorted order (cluster-contiguous, ascending original id inside a
// cluster, as the reference's cluster_to_points push_back order).
//
// Query: the list-scan kernel walks work items (row chunk of one list x
// group of up to 16 queries that probe that list), pulled from an atomic
// queue by every wave, so Zipf-skewed list lengths only change the number of
// items.  Each item is one 16-query x 16-row int8 MFMA tile stream over the
// list's SQ8 codes, with the flat scan's threshold filter and append.
#include <hipcub/hipcub.hpp>

#include "common.h"
#include "ivf.h"

namespace {

__global__ void gather_rows_f32(const float* src, int64_t stride, const uint32_t* idx, int64_t n, int kdim,
                                float* dst) {
    const int64_t i = blockIdx.x;
    if (i >= n) return;
    const float* s = src + (int64_t)idx[i] * stride;
    for (int j = threadIdx.x; j < kdim; j += blockDim.x) dst[i * kdim + j] = s[j];
}

__global__ void fill_u32_k(uint32_t* p, uint32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_u64(unsigned long long* p, unsigned long long v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void iota_u32(uint32_t* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}

__global__ void argmin_to_assign(const unsigned long long* am, int64_t n, int nlist, uint32_t* assign,
                                 uint32_t* counts) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c = (uint32_t)(am[i] & 0xFFFFFFFFull);
    if (c >= (uint32_t)nlist) c = 0;  // a row whose distances were all NaN
    assign[i] = c;
    atomicAdd(&counts[c], 1u);
}

// One workgroup per segment [pos_begin, pos_end) of one cluster's sorted
// points: fp64 column sums in sorted order.
__global__ __launch_bounds__(256) void segment_sums(const float* X, int64_t stride, const uint32_t* perm,
                                                   const KmSeg* segs, int d, double* partial) {
    const KmSeg s = segs[blockIdx.x];
    double* out = partial + (int64_t)blockIdx.x * d;
    for (int j0 = 0; j0 < d; j0 += 256) {
        const int j = j0 + threadIdx.x;
        double acc = 0.0;
        if (j < d)
            for (int64_t p = s.pos_begin; p < s.pos_end; ++p) acc += (double)X[(int64_t)perm[p] * stride + j];
        if (j < d) out[j] = acc;
    }
}

// centroid[c] = (sum of its segments, in order) / count; empty cluster ->
// row rand_idx[(c * 37) % n] (lsh_index_optimized.cpp:609-613).
__global__ __launch_bounds__(256) void centroid_update(const double* partial, const int* seg_first,
                                                       const uint32_t* counts, int d, int kdim, const float* X,
                                                       int64_t stride, const uint32_t* rand_idx, int64_t n,
                                                       float* cent) {
    const int c = blockIdx.x;
    const uint32_t cnt = counts[c];
    float* o = cent + (int64_t)c * kdim;
    if (cnt == 0) {
        const float* src = X + (int64_t)rand_idx[((int64_t)c * 37) % n] * stride;
        for (int j = threadIdx.x; j < kdim; j += 256) o[j] = j < d ? src[j] : 0.f;
        return;
    }
    const int s0 = seg_first[c], s1 = seg_first[c + 1];
    for (int j = threadIdx.x; j < kdim; j += 256) {
        if (j >= d) {
            o[j] = 0.f;
            continue;
        }
        double acc = 0.0;
        for (int s = s0; s < s1; ++s) acc += partial[(int64_t)s * d + j];
        o[j] = (float)(acc / (double)cnt);
    }
}

// ------------------------------------------------------------ step plan --
// Exclusive prefix sum over a 1024-thread block (wave scans, then the wave
// totals); returns this thread's prefix and the block total in *total.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        if (w < wave) before += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// Queries per list: each query's probes, or every list when its probes hold
// fewer than kk rows (lsh_index_optimized.cpp:452-457).
__global__ __launch_bounds__(256) void ivf_plan_count(IvfPlanArgs a) {
    __shared__ int64_t part[4];
    const int q = blockIdx.x;
    const int64_t* pr = a.probe + (int64_t)q * a.nprobe;
    int64_t rows = 0;
    for (int j = threadIdx.x; j < a.nprobe; j += blockDim.x) {
        const int64_t l = pr[j];
        if (l >= 0) rows += a.list_off[l + 1] - a.list_off[l];
    }
#pragma unroll
    for (int o = 32; o; o >>= 1) rows += __shfl_xor(rows, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = rows;
    __syncthreads();
    const bool every = part[0] + part[1] + part[2] + part[3] < (int64_t)a.kk;
    if (every) {
        for (int l = threadIdx.x; l < a.nlist; l += blockDim.x) atomicAdd(&a.qcount[l], 1u);
    } else {
        for (int j = threadIdx.x; j < a.nprobe; j += blockDim.x)
            if (pr[j] >= 0) atomicAdd(&a.qcount[pr[j]], 1u);
    }
    if (threadIdx.x == 0) a.every[q] = every ? 1 : 0;
}

// One workgroup: query offsets per list, then per pass the item prefix.
__global__ __launch_bounds__(1024) void ivf_plan_scan(IvfPlanArgs a) {
    __shared__ uint32_t wsum[16];
    uint32_t carry = 0, total = 0;
    for (int b = 0; b < a.nlist; b += 1024) {
        const int i = b + (int)threadIdx.x;
        const uint32_t v = i < a.nlist ? a.qcount[i] : 0u;
        const uint32_t ex = block_exclusive_scan(v, wsum, &total);
        if (i < a.nlist) {
            a.qoff[i] = carry + ex;
            a.qfill[i] = 0u;
        }
        carry += total;
    }
    for (int p = 0; p < a.npass; ++p) {
        int64_t lo, hi;
        ivf_pass_bounds(a.S, p, lo, hi);
        uint32_t* plan = a.plan + (int64_t)p * (a.nlist + 1);
        carry = 0;
        for (int b = 0; b < a.nlist; b += 1024) {
            const int i = b + (int)threadIdx.x;
            uint32_t v = 0;
            if (i < a.nlist && a.qcount[i] > 0) {
                const int64_t len = a.list_off[i + 1] - a.list_off[i];
                const int64_t rows = (len < hi ? len : hi) - lo;
                if (rows > 0) v = ((a.qcount[i] + 15u) / 16u) * (uint32_t)((rows + a.chunk - 1) / a.chunk);
            }
            const uint32_t ex = block_exclusive_scan(v, wsum, &total);
            if (i < a.nlist) plan[i] = carry + ex;
            carry += total;
        }
        if (threadIdx.x == 0) {
            plan[a.nlist] = carry;
            a.nitems[p] = carry;
        }
    }
}

// For every 64th item of each pass, the list holding it: the scan kernel's
// item search then runs between two neighbouring entries only.
__global__ __launch_bounds__(256) void ivf_plan_coarse(IvfPlanArgs a) {
    const int p = blockIdx.y;
    const uint32_t* plan = a.plan + (int64_t)p * (a.nlist + 1);
    uint32_t* co = a.coarse + (int64_t)p * a.coarse_cap;
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < a.nlist; l += gridDim.x * blockDim.x) {
        const uint32_t b = (plan[l] + 63u) >> 6, e = (plan[l + 1] + 63u) >> 6;
        for (uint32_t j = b; j < e && j < (uint32_t)a.coarse_cap; ++j) co[j] = (uint32_t)l;
    }
}

// Query ids into their lists' slots (the order inside a list does not
// change the answers: a pass's filter is fixed while it runs).
__global__ __launch_bounds__(256) void ivf_plan_scatter(IvfPlanArgs a) {
    const int q = blockIdx.x;
    if (a.every[q]) {
        for (int l = threadIdx.x; l < a.nlist; l += blockDim.x) a.gq[a.qoff[l] + atomicAdd(&a.qfill[l], 1u)] = (uint32_t)q;
    } else {
        const int64_t* pr = a.probe + (int64_t)q * a.nprobe;
        for (int j = threadIdx.x; j < a.nprobe; j += blockDim.x) {
            const int64_t l = pr[j];
            if (l >= 0) a.gq[a.qoff[l] + atomicAdd(&a.qfill[l], 1u)] = (uint32_t)q;
        }
    }
}

// ------------------------------------------------------------ list scan --
constexpr int kIvfThreads = 256;

// Work item `it` of a pass: its list l (plan[l] <= it < plan[l + 1], binary
// search), then the query group and row chunk inside the list.
__device__ inline IvfItem ivf_item(const IvfScanArgs& a, uint32_t it) {
    int lo = 0, hi = a.nlist;
    const uint32_t j = it >> 6;
    if (j < (uint32_t)a.coarse_cap) {
        // plan[coarse[j]] <= 64 j <= it; item 64 (j + 1), when it exists,
        // lies in list coarse[j + 1], so it < plan[coarse[j + 1] + 1]
        lo = (int)a.coarse[j];
        if (j + 1 < (uint32_t)a.coarse_cap && (j + 1) * 64u < *a.nitems) hi = (int)a.coarse[j + 1] + 1;
    }
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.plan[mid] <= it) lo = mid;
        else hi = mid;
    }
    const uint32_t local = it - a.plan[lo];
    const int64_t b0 = a.list_off[lo], len = a.list_off[lo + 1] - b0;
    const int64_t e = len < a.hi ? len : a.hi;
    const uint32_t nch = (uint32_t)((e - a.lo + a.chunk - 1) / a.chunk);
    const uint32_t grp = local / nch, c = local - grp * nch;
    IvfItem t;
    t.row_begin = (uint32_t)(b0 + a.lo + (int64_t)c * a.chunk);
    t.row_end = (uint32_t)(b0 + e < (int64_t)t.row_begin + a.chunk ? b0 + e : (int64_t)t.row_begin + a.chunk);
    t.qoff = a.qoff[lo] + 16u * grp;
    const uint32_t left = a.qcount[lo] - 16u * grp;
    t.nq = left < 16u ? left : 16u;
    return t;
}

__global__ __launch_bounds__(kIvfThreads) void ivf_scan_kernel(IvfScanArgs a) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4;
    const int n = lane & 15;
    const uint8_t* X = static_cast<const uint8_t*>(a.codes);
    const int nchunk = a.kdim / 64;
    for (;;) {
        uint32_t it = 0;
        if (lane == 0) it = atomicAdd(a.next, 1u);
        it = __shfl(it, 0);
        if (it >= *a.nitems) break;
        const IvfItem item = ivf_item(a, it);
        // A operand: query n of the group (16 queries per tile)
        const int qa = n < (int)item.nq ? (int)a.gq[item.qoff + n] : -1;
        // epilogue slots: queries 4g + i of the group
        int qe[4];
        float thr[4], qn[4], qs[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = 4 * g + i;
            qe[i] = m < (int)item.nq ? (int)a.gq[item.qoff + m] : -1;
            thr[i] = qe[i] >= 0 ? a.thr[qe[i]] : -__builtin_inff();
            qn[i] = qe[i] >= 0 ? a.qn[qe[i]] : 0.f;
            qs[i] = qe[i] >= 0 ? a.qscale[qe[i]] : 0.f;
        }
        const uint8_t* qrow = qa >= 0 ? static_cast<const uint8_t*>(a.Q) + (int64_t)qa * a.kdim : nullptr;
        // rows of up to kRegChunks chunks (d <= 1024): the item's query
        // fragments stay in registers across its tiles and every row chunk
        // of a tile is requested before the first MFMA; longer rows take the
        // chunk loop below
        constexpr int kRegChunks = 16;
        const bool reg = nchunk <= kRegChunks;
        i32x4 qf[kRegChunks];
        if (reg) {
#pragma unroll
            for (int c = 0; c < kRegChunks; ++c)
                qf[c] = (qrow && c < nchunk) ? *reinterpret_cast<const i32x4*>(qrow + 16 * g + c * 64)
                                             : (i32x4){0, 0, 0, 0};
        }
        for (uint32_t r0 = item.row_begin; r0 < item.row_end; r0 += 16) {
            const uint32_t row = r0 + n;
            const bool rok = row < item.row_end;
            const int64_t lrow = rok ? row : item.row_end - 1;
            i32x4 acc = (i32x4){0, 0, 0, 0};
            const uint8_t* xr = X + lrow * a.xstride + 16 * g;
            const float xn = a.xn[lrow];
            if (reg) {
                i32x4 b[kRegChunks];
#pragma unroll
                for (int c = 0; c < kRegChunks; ++c)
                    if (c < nchunk) b[c] = *reinterpret_cast<const i32x4*>(xr + c * 64);
#pragma unroll
                for (int c = 0; c < kRegChunks; ++c)
                    if (c < nchunk) {
                        const i32x4 bs =
                            b[c] ^ (i32x4){(int)0x80808080, (int)0x80808080, (int)0x80808080, (int)0x80808080};
                        acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(qf[c], bs, acc, 0, 0, 0);
                    }
            } else {
                const uint8_t* qr = qrow ? qrow + 16 * g : nullptr;
                // row and query fragments requested U chunks at a time before
                // their MFMAs (the chunk count is a runtime value, so the
                // compiler would otherwise wait on every load in turn)
                constexpr int U = 4;
                int c = 0;
                for (; c + U <= nchunk; c += U) {
                    i32x4 b[U], q[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        b[u] = *reinterpret_cast<const i32x4*>(xr + (c + u) * 64);
                        q[u] = qr ? *reinterpret_cast<const i32x4*>(qr + (c + u) * 64) : (i32x4){0, 0, 0, 0};
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const i32x4 bs =
                            b[u] ^ (i32x4){(int)0x80808080, (int)0x80808080, (int)0x80808080, (int)0x80808080};
                        acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(q[u], bs, acc, 0, 0, 0);
                    }
                }
                for (; c < nchunk; ++c) {
                    i32x4 b = *reinterpret_cast<const i32x4*>(xr + c * 64);
                    b ^= (i32x4){(int)0x80808080, (int)0x80808080, (int)0x80808080, (int)0x80808080};
                    i32x4 q = qr ? *reinterpret_cast<const i32x4*>(qr + c * 64) : (i32x4){0, 0, 0, 0};
                    acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(q, b, acc, 0, 0, 0);
                }
            }
            // Two phases: the tile's list reservations (one returning atomic
            // per query and 16-row group with passing rows) are all issued
            // before any result is used, so their round trips overlap (the
            // first pass, with an open filter, appends every row)
            float dist[4];
            uint32_t gm[4], base[4];
            uint32_t live = 0;  // bit i: some lane of the wave passes (wave-uniform)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                dist[i] = xn + qn[i] - 2.0f * qs[i] * (float)acc[i];
                const bool pass = rok && dist[i] <= thr[i];
                const uint64_t mb = __ballot(pass);
                gm[i] = (uint32_t)(mb >> (16 * g)) & 0xFFFFu;
                base[i] = 0;
                if (mb != 0) {
                    live |= 1u << i;
                    if (n == 0 && gm[i]) base[i] = atomicAdd(&a.cand.cnt[qe[i]], (uint32_t)__builtin_popcount(gm[i]));
                }
            }
            if (live == 0) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (!((live >> i) & 1u)) continue;
                const uint32_t bb = __shfl(base[i], lane & ~15);
                if ((gm[i] >> n) & 1u) {
                    const uint32_t off = bb + __builtin_popcount(gm[i] & ((1u << n) - 1u));
                    if (off < (uint32_t)a.cand.cap) {
                        const int64_t sl = (int64_t)qe[i] * a.cand.cap + off;
                        a.cand.key[sl] = dist[i];
                        a.cand.rid[sl] = a.perm[row];
                    }
                }
            }
        }
    }
}

}  // namespace

hipError_t launch_gather_rows_f32(const float* src, int64_t stride, const uint32_t* idx, int64_t n, int kdim,
                                  float* dst, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_f32, dim3((unsigned)n), dim3(256), 0, st, src, stride, idx, n, kdim, dst);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, unsigned long long v, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_iota_u32(uint32_t* p, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(iota_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_argmin_to_assign(const unsigned long long* am, int64_t n, int nlist, uint32_t* assign,
                                   uint32_t* counts, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(argmin_to_assign, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, am, n, nlist, assign,
                       counts);
    return hipGetLastError();
}

hipError_t sort_pairs_u32(void* tmp, size_t& tmp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                          const uint32_t* vals_in, uint32_t* vals_out, int64_t n, int key_bits, hipStream_t st) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out, (int)n, 0,
                                              key_bits, st);
}

hipError_t exclusive_sum_u32(void* tmp, size_t& tmp_bytes, const uint32_t* in, uint32_t* out, int64_t n,
                             hipStream_t st) {
    return hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, (int)n, st);
}

hipError_t launch_segment_sums(const float* X, int64_t stride, const uint32_t* perm, const KmSeg* segs, int nseg,
                               int d, double* partial, hipStream_t st) {
    if (nseg <= 0) return hipSuccess;
    hipLaunchKernelGGL(segment_sums, dim3((unsigned)nseg), dim3(256), 0, st, X, stride, perm, segs, d, partial);
    return hipGetLastError();
}

hipError_t launch_centroid_update(const double* partial, const int* seg_first, const uint32_t* counts, int nlist,
                                  int d, int kdim, const float* X, int64_t stride, const uint32_t* rand_idx,
                                  int64_t n, float* cent, hipStream_t st) {
    if (nlist <= 0) return hipSuccess;
    hipLaunchKernelGGL(centroid_update, dim3((unsigned)nlist), dim3(256), 0, st, partial, seg_first, counts, d, kdim,
                       X, stride, rand_idx, n, cent);
    return hipGetLastError();
}

hipError_t launch_ivf_plan(const IvfPlanArgs& a, hipStream_t st) {
    if (a.B <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_k, dim3((unsigned)((a.nlist + 255) / 256)), dim3(256), 0, st, a.qcount, 0u,
                       (int64_t)a.nlist);
    ANNMI_LAUNCH(ivf_plan_count, dim3((unsigned)a.B), dim3(256), 0, st, a);
    ANNMI_LAUNCH(ivf_plan_scan, dim3(1), dim3(1024), 0, st, a);
    ANNMI_LAUNCH(ivf_plan_coarse, dim3((unsigned)((a.nlist + 255) / 256), (unsigned)a.npass), dim3(256), 0, st, a);
    ANNMI_LAUNCH(ivf_plan_scatter, dim3((unsigned)a.B), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_ivf_scan(const IvfScanArgs& a, hipStream_t st) {
    // persistent: a few workgroups per CU, every wave pulls items until the
    // pass's queue (its count is on the device) is empty
    const int ncu = device_cu_count();
    ANNMI_LAUNCH(ivf_scan_kernel, dim3((unsigned)(ncu * 8)), dim3(kIvfThreads), 0, st, a);
    return hipGetLastError();
}

// Per-query selection over the candidate lists written by the scan:
//
//  * select_threshold -- k-th (or R-th) smallest approximate distance among
//    a query's candidates so far, the filter for the next row range;
//  * finalize -- pick the candidates to re-score, re-score them in fp32 in
//    the reference's own summation order (src/simd.h:32-64: 16 lanes, lane l
//    takes dims l, l+16, ... with FMA, tree reduction, 8-wide + scalar tail),
//    sort by (distance, id), resolve fp32 near-ties in fp64, write top-k.
//
// The exact top-k rule being reproduced is src/linear_index.cpp:110-118
// (partial_sort by (value, id): ties to the lower id) with the value taken
// as the exact distance; gpu_filter.cu:161-174 is the over-fetch + rerank
// shape.  The selection itself is a block-wide radix select on orderable
// float keys: 11 + 11 + 10 bits, histograms in LDS.
#include "common.h"
#include "select_dev.h"

namespace {

constexpr int kRMax = 2048;   // most candidates one query re-scores

// Lists of up to kStageKeys candidates are staged in LDS as ordered keys
// (eight loads a lane in flight) and selected there: a block reading the
// list from global memory once per radix pass waited on load latency for
// most of its time.  After the threshold the list is compacted in place to
// the candidates at or below it: the threshold only falls from range to
// range and finalize collects keys at or below the final rank-rsel key plus
// the same slack, so no dropped candidate could be collected later, and the
// next selections and finalize see ~rsel * rho keys instead of everything
// appended so far.
constexpr int kStageKeys = 8192;

__global__ __launch_bounds__(kThreads) void select_threshold(SelectArgs a) {
    __shared__ uint32_t hist[kBins];
    __shared__ uint32_t sh[8];
    __shared__ uint32_t wsum[kThreads / 64];
    __shared__ float tsh;
    extern __shared__ uint32_t skeys[];  // [kStageKeys] keys, then [kStageKeys] row ids
    uint32_t* srids = skeys + kStageKeys;
    const int q = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t cnt = a.cand.cnt[q];
    if (cnt > (uint32_t)a.cand.cap) {
        if (threadIdx.x == 0) a.overflow[q] |= 1u;
        cnt = (uint32_t)a.cand.cap;
    }
    if (cnt < (uint32_t)a.rsel) return;  // not enough rows seen yet: keep +inf
    float* keys = a.cand.key + (int64_t)q * a.cand.cap;
    uint32_t* rids = a.cand.rid + (int64_t)q * a.cand.cap;
    const bool staged = cnt <= (uint32_t)kStageKeys;
    uint32_t k;
    if (staged) {
        // keys and row ids staged together (the compaction below then reads
        // only LDS; with ids read from global memory per 256-entry round it
        // waited on a load latency per round)
        constexpr int U = 8;
        for (int i0 = 0; i0 < (int)cnt; i0 += kThreads * U) {
            float v[U];
            uint32_t r[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + u * kThreads + (int)threadIdx.x;
                v[u] = i < (int)cnt ? keys[i] : 0.f;
                r[u] = i < (int)cnt ? rids[i] : 0u;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + u * kThreads + (int)threadIdx.x;
                if (i < (int)cnt) {
                    skeys[i] = f2key(v[u]);
                    srids[i] = r[u];
                }
            }
        }
        __syncthreads();
        k = radix_select_lds(skeys, (int)cnt, (uint32_t)a.rsel, hist, sh);
    } else {
        k = radix_select(keys, (int)cnt, (uint32_t)a.rsel, hist, sh);
    }
    if (threadIdx.x == 0) {
        float t = key2f(k) + a.slack[q];
        // round the threshold up by one ulp so that the fp32 add never cuts a
        // candidate whose exact sum would be admitted
        t = next_up(t);
        if (t < a.thr[q]) a.thr[q] = t;
        tsh = a.thr[q];
    }
    __syncthreads();
    // compaction in place, order kept: a chunk's reads all happen before its
    // writes, and a kept entry moves to an index no larger than its own
    const float t = tsh;
    uint32_t base = 0;
    if (staged) {
        // from LDS, four consecutive entries a thread per round
        constexpr int V = 4;
        for (int i0 = 0; i0 < (int)cnt; i0 += kThreads * V) {
            const int i1 = i0 + (int)threadIdx.x * V;
            bool keep[V];
            uint32_t nk = 0;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                keep[v] = i1 + v < (int)cnt && key2f(skeys[i1 + v]) <= t;
                nk += keep[v] ? 1u : 0u;
            }
            // exclusive prefix of nk over the block: in the wave, then the waves
            uint32_t incl = nk;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t u = __shfl_up(incl, o);
                if (lane >= o) incl += u;
            }
            if (lane == 63) wsum[wave] = incl;
            __syncthreads();
            uint32_t off = base + incl - nk, total = 0;
#pragma unroll
            for (int w = 0; w < kThreads / 64; ++w) {
                if (w < wave) off += wsum[w];
                total += wsum[w];
            }
#pragma unroll
            for (int v = 0; v < V; ++v)
                if (keep[v]) {
                    keys[off] = key2f(skeys[i1 + v]);
                    rids[off] = srids[i1 + v];
                    ++off;
                }
            base += total;
            __syncthreads();
        }
        if (threadIdx.x == 0) a.cand.cnt[q] = base;
        return;
    }
    for (int i0 = 0; i0 < (int)cnt; i0 += kThreads) {
        const int i = i0 + (int)threadIdx.x;
        float kf = 0.f;
        uint32_t rid = 0;
        bool keep = false;
        if (i < (int)cnt) {
            kf = staged ? key2f(skeys[i]) : keys[i];
            keep = kf <= t;
            if (keep) rid = rids[i];
        }
        const uint64_t m = __ballot(keep);
        if (lane == 0) wsum[wave] = (uint32_t)__builtin_popcountll(m);
        __syncthreads();
        uint32_t off = base, total = 0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w) {
            if (w < wave) off += wsum[w];
            total += wsum[w];
        }
        off += (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull));
        if (keep) {
            keys[off] = kf;
            rids[off] = rid;
        }
        base += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) a.cand.cnt[q] = base;
}

// ---------------------------------------------------------------- finalize
struct FinalSmem {
    uint32_t hist[kBins];
    uint32_t sh[8];
    uint32_t ncol;
    uint32_t id[kRMax];
    float d32[kRMax];
    uint64_t sk[kRMax];
    uint8_t link[kRMax];
};

__global__ __launch_bounds__(kThreads) void finalize(FinalArgs a) {
    extern __shared__ __align__(16) uint8_t smem_raw[];
    FinalSmem& S = *reinterpret_cast<FinalSmem*>(smem_raw);
    float* qs = reinterpret_cast<float*>(smem_raw + sizeof(FinalSmem));
    const int q = blockIdx.x;
    const int lane = threadIdx.x & 63;
    for (int j = threadIdx.x; j < a.d; j += kThreads) qs[j] = a.Q32[(int64_t)q * a.d + j];

    uint32_t cnt = a.cand.cnt[q];
    uint32_t flags = 0;
    if (cnt > (uint32_t)a.cand.cap) {
        flags |= 1u;
        cnt = (uint32_t)a.cand.cap;
    }
    const float* keys = a.cand.key + (int64_t)q * a.cand.cap;
    const uint32_t* rids = a.cand.rid + (int64_t)q * a.cand.cap;

    // Lists of up to 2 * kRMax keys (the usual case: the range selections
    // compact them) are staged once in the sort buffer, free until step 4,
    // as ordered keys for the selection and the collection.
    uint32_t* skeys = reinterpret_cast<uint32_t*>(S.sk);
    const bool staged = cnt <= (uint32_t)(2 * kRMax);
    if (staged) {
        constexpr int U = 8;
        for (int i0 = 0; i0 < (int)cnt; i0 += kThreads * U) {
            float v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + u * kThreads + (int)threadIdx.x;
                v[u] = i < (int)cnt ? keys[i] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + u * kThreads + (int)threadIdx.x;
                if (i < (int)cnt) skeys[i] = f2key(v[u]);
            }
        }
        __syncthreads();
    }
    // 1. collection threshold
    float thr = __builtin_inff();
    if (cnt > (uint32_t)a.rsel) {
        uint32_t k = staged ? radix_select_lds(skeys, (int)cnt, (uint32_t)a.rsel, S.hist, S.sh)
                            : radix_select(keys, (int)cnt, (uint32_t)a.rsel, S.hist, S.sh);
        thr = key2f(k);
        if (a.slack[q] > 0.f) thr = next_up(thr + a.slack[q]);
    }
    // 2. collect candidates with key <= thr (ties at the pivot included)
    if (threadIdx.x == 0) S.ncol = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < (int)cnt; i += kThreads) {
        if ((staged ? key2f(skeys[i]) : keys[i]) <= thr) {
            uint32_t pos = atomicAdd(&S.ncol, 1u);
            if (pos < (uint32_t)kRMax) S.id[pos] = rids[i];
        }
    }
    __syncthreads();
    uint32_t ncol = S.ncol;
    if (ncol > (uint32_t)kRMax) {
        flags |= 2u;
        ncol = kRMax;
    }
    // 3. fp32 re-score, one row per 16-lane group
    // (four rows per group at a time, so four times the loads are in flight:
    // one query's ~130 rows take two rounds of row loads instead of five)
    const int grp = threadIdx.x >> 4, l16 = threadIdx.x & 15;
    constexpr int G = kThreads / 16;
    for (uint32_t c = grp; c < ncol; c += 4 * G) {
        uint32_t cr[4];
        const float* x[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cr[r] = c + r * G < ncol ? c + r * G : c;
            x[r] = a.X32 + (int64_t)S.id[cr[r]] * a.x32_stride;
        }
        float v[4];
        l2_ref_order<4>(x, qs, a.d, l16, v);
        if (l16 == 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) S.d32[cr[r]] = v[r];
        }
    }
    __syncthreads();
    // 4. sort by (d32, id)
    int np = 1;
    while (np < (int)ncol) np <<= 1;
    for (int i = threadIdx.x; i < np; i += kThreads)
        S.sk[i] = i < (int)ncol ? (((uint64_t)f2key(S.d32[i]) << 32) | S.id[i]) : ~0ull;
    __syncthreads();
    bitonic_sort(S.sk, np);
    // 5. fp64 boundary step: runs of neighbours closer than eps_rel (fp32
    // error bound of both values) that start before rank k are re-scored in
    // fp64 and re-ordered by (d64, id).
    for (int i = threadIdx.x; i < (int)ncol; i += kThreads) {
        uint8_t l = 0;
        if (i + 1 < (int)ncol) {
            float d0 = key2f((uint32_t)(S.sk[i] >> 32)), d1 = key2f((uint32_t)(S.sk[i + 1] >> 32));
            l = (double)d1 - (double)d0 <= a.eps_rel * fabs((double)d1) ? 1 : 0;
        }
        S.link[i] = l;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < (int)ncol && i < a.k; i += kThreads) {
        if (!S.link[i] || (i > 0 && S.link[i - 1])) continue;  // not a run start
        int e = i;
        while (e + 1 < (int)ncol && S.link[e]) ++e;            // run [i, e]
        // insertion sort of the run by (d64, id); at most a handful of rows
        double dv[32];
        uint64_t kv[32];
        int m = e - i + 1;
        if (m > 32) {
            m = 32;
            atomicOr(&a.overflow[q], 4u);  // run longer than the fp64 window: left in (fp32, id) order
        }
        for (int j = 0; j < m; ++j) {
            kv[j] = S.sk[i + j];
            dv[j] = l2_f64(a.X32 + (int64_t)(uint32_t)kv[j] * a.x32_stride, qs, a.d);
        }
        for (int j = 1; j < m; ++j) {
            double dj = dv[j];
            uint64_t kj = kv[j];
            int t = j - 1;
            while (t >= 0 && (dv[t] > dj || (dv[t] == dj && (uint32_t)kv[t] > (uint32_t)kj))) {
                dv[t + 1] = dv[t];
                kv[t + 1] = kv[t];
                --t;
            }
            dv[t + 1] = dj;
            kv[t + 1] = kj;
        }
        for (int j = 0; j < m; ++j) S.sk[i + j] = kv[j];
    }
    __syncthreads();
    // 6. output
    for (int j = threadIdx.x; j < a.k; j += kThreads) {
        int64_t o = (int64_t)q * a.k + j;
        if (j < (int)ncol) {
            a.out_ids[o] = (int64_t)(uint32_t)S.sk[j];
            a.out_d[o] = key2f((uint32_t)(S.sk[j] >> 32));
        } else {
            a.out_ids[o] = -1;
            a.out_d[o] = __builtin_inff();
        }
    }
    if (threadIdx.x == 0) {
        a.overflow[q] |= flags;
        a.n_reranked[q] = ncol;
    }
    (void)lane;
}

}  // namespace

hipError_t launch_select_threshold(const SelectArgs& a, hipStream_t st) {
    if (a.B <= 0) return hipSuccess;
    ANNMI_LAUNCH(select_threshold, dim3((unsigned)a.B), dim3(kThreads), 2 * kStageKeys * sizeof(uint32_t), st, a);
    return hipGetLastError();
}

size_t finalize_smem_bytes(int d) { return sizeof(FinalSmem) + (size_t)d * sizeof(float); }

const void* finalize_kernel_fn() { return reinterpret_cast<const void*>(&finalize); }

hipError_t launch_finalize(const FinalArgs& a, hipStream_t st) {
    if (a.B <= 0) return hipSuccess;
    size_t lds = finalize_smem_bytes(a.d);
    ANNMI_LAUNCH(finalize, dim3((unsigned)a.B), dim3(kThreads), lds, st, a);
    return hipGetLastError();
}

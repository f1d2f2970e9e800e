// CPU restatement of the reference k-NN search path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (annvedi_amd/) links,
// loads or calls this file; only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it, and only as the checker.
//
// Compiled with -ffp-contract=off: every fused multiply-add below is an
// explicit fmaf()/fma(), every other product/sum is rounded on its own, so the
// bits do not depend on the compiler.
//
// Functions and the reference code each one restates:
//   orc_l2_linear_index      src/simd.h:32-64   (AVX-512 branch of compute_l2_distance,
//                            as compiled by src/setup.py:10 flags: lane FMAs, tree reduce)
//   orc_linear_index_query   src/linear_index.cpp:27-60, 90-121 (64-slot early-abandon
//                            scan storing partial sums, partial_sort by (value, id))
//   orc_exact_topk_f64       closed form: sum_j (double(x_j) - double(q_j))^2 summed in
//                            index order without FMA; ranked by (d64, id).  This is the
//                            "exact" oracle; there is no single reference function for it
//                            (prepare_data.py:39-51 uses an unstable float32 argsort).
//   orc_sq8_train/encode     competitors/ivf_lsh/src/lsh_index_optimized.cpp:186-214
//                            (std::round, constant dim -> scale 1 offset 0) and
//                            competitors/gpu_filter/src/gpu_filter.cu:99-127 (lrintf,
//                            constant dim -> scale 1 offset min)
//   orc_sq8_l2               competitors/ivf_lsh/src/simd.h:118-147 (AVX-512 branch,
//                            no threshold): fmsub then fmadd, 16 lanes, tree reduce
//   orc_sq4_train/encode     submission/src/hnsw.cpp:147-183 (nibble layout)
//   orc_sq4_l2               submission/src/simd.h:292-318 (AVX-512 branch, no threshold)
//   orc_f32_direct_l2        gpu_filter.cu:161-170 rerank loop (scalar acc += v*v)
//   orc_gf_keys              gpu_filter.cu:52-63 (serial fmaf SQ8 key per row)
//   orc_kmeans_init          lsh_index_optimized.cpp:553-561 (initial centroid rows)
//   orc_ivf_kmeans           lsh_index_optimized.cpp:551-616, :32-73 (k-means, one thread)
//
// Parity pins for each function are listed in the header of oracle/oracle.py.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <queue>
#include <random>
#include <thread>
#include <utility>
#include <vector>

namespace {

// h_j = s_j + s_{j+8}; ((h0+h1)+(h2+h3)) + ((h4+h5)+(h6+h7)).
// src/simd.h:21-28 reduce_add_ps_512; the tree is the order g++ -O3 -ffast-math
// emits for the 8-term scalar sum (vshufps/vunpckhps pairs, then pair sums).
inline float reduce16(const float s[16]) {
    float h[8];
    for (int j = 0; j < 8; ++j) h[j] = s[j] + s[j + 8];
    float a = h[0] + h[1];
    float b = h[2] + h[3];
    float c = h[4] + h[5];
    float e = h[6] + h[7];
    return (a + b) + (c + e);
}

// Remainder of 8..15 dims after the 16-wide loop: the compiler's 8-wide
// vectorised tail: products p_j, g_j = p_j + p_{j+4}, (g0+g2) + (g1+g3).
inline float tail8(const float p[8]) {
    float g[4];
    for (int j = 0; j < 4; ++j) g[j] = p[j] + p[j + 4];
    return (g[0] + g[2]) + (g[1] + g[3]);
}

}  // namespace

extern "C" {

// src/simd.h:32-64: squared L2 with early abandon after 96, 192, ... dims.
float orc_l2_linear_index(const float* a, const float* b, int64_t dim, float threshold) {
    float s[16];
    for (int l = 0; l < 16; ++l) s[l] = 0.0f;
    int64_t i = 0;
    for (; i + 15 < dim; i += 16) {
        for (int l = 0; l < 16; ++l) {
            float diff = a[i + l] - b[i + l];
            s[l] = fmaf(diff, diff, s[l]);
        }
        if (i % 96 == 80) {
            float sum = reduce16(s);
            if (sum >= threshold) return sum;
        }
    }
    float sum = reduce16(s);
    if (sum >= threshold) return sum;
    int64_t r = dim - i;
    if (r >= 8) {
        float p[8];
        for (int j = 0; j < 8; ++j) {
            float diff = a[i + j] - b[i + j];
            p[j] = diff * diff;
        }
        sum = sum + tail8(p);
        i += 8;
    }
    for (; i < dim; ++i) {
        float diff = a[i] - b[i];
        sum = fmaf(diff, diff, sum);
    }
    return sum;
}

// Partial sums of orc_l2_linear_index at every checkpoint, no early exit.
// out[c] = reduced sum after 96*(c+1) dims for c < dim/96 (16-dim groups
// only), and out[n_ck] = the full value.  Returns n_ck.
int orc_l2_linear_index_checkpoints(const float* a, const float* b, int64_t dim, float* out) {
    float s[16];
    for (int l = 0; l < 16; ++l) s[l] = 0.0f;
    int64_t i = 0;
    int n = 0;
    for (; i + 15 < dim; i += 16) {
        for (int l = 0; l < 16; ++l) {
            float diff = a[i + l] - b[i + l];
            s[l] = fmaf(diff, diff, s[l]);
        }
        if (i % 96 == 80) out[n++] = reduce16(s);
    }
    float full = orc_l2_linear_index(a, b, dim, std::numeric_limits<float>::infinity());
    out[n] = full;
    return n;
}

// src/linear_index.cpp:90-121 with linear_worker :27-60 (single thread).
// Writes min(k, n) ids and their stored values (which can be early-abandon
// partial sums beyond rank 64).  Returns the count written.
int orc_linear_index_query(const float* base, int64_t n, int64_t dim, const float* q, int k,
                           int64_t* out_ids, float* out_vals) {
    std::vector<std::pair<float, int>> dist_id((size_t)n);
    std::pair<float, int> top[64];
    int limit_k = std::min(k, 64);
    for (int j = 0; j < limit_k; ++j) top[j] = {std::numeric_limits<float>::max(), -1};
    for (int64_t i = 0; i < n; ++i) {
        float threshold = top[limit_k - 1].first;
        float sum = orc_l2_linear_index(base + i * dim, q, dim, threshold);
        dist_id[(size_t)i] = {sum, (int)i};
        if (sum < threshold) {
            int pos = limit_k - 1;
            while (pos > 0 && sum < top[pos - 1].first) {
                top[pos] = top[pos - 1];
                pos--;
            }
            top[pos] = {sum, (int)i};
        }
    }
    int kk = k > (int)n ? (int)n : k;
    std::partial_sort(dist_id.begin(), dist_id.begin() + kk, dist_id.end());
    for (int j = 0; j < kk; ++j) {
        out_ids[j] = dist_id[(size_t)j].second;
        out_vals[j] = dist_id[(size_t)j].first;
    }
    return kk;
}

// Batched form of orc_linear_index_query over nq queries with nthreads host
// threads (one query per thread at a time); row-major outputs nq x k.
int orc_linear_index_batch(const float* base, int64_t n, int64_t dim, const float* Q, int nq, int k,
                           int64_t* out_ids, float* out_vals, int nthreads) {
    int kk = k > (int)n ? (int)n : k;
    if (nthreads < 1) nthreads = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) {
        th.emplace_back([=]() {
            for (int qi = t; qi < nq; qi += nthreads)
                orc_linear_index_query(base, n, dim, Q + (int64_t)qi * dim, k, out_ids + (int64_t)qi * k,
                                       out_vals + (int64_t)qi * k);
        });
    }
    for (auto& x : th) x.join();
    return kk;
}

// Exact squared L2 in double: sequential over dims, no FMA.
double orc_l2_f64(const float* a, const float* b, int64_t dim) {
    double acc = 0.0;
    for (int64_t j = 0; j < dim; ++j) {
        double diff = (double)a[j] - (double)b[j];
        double sq = diff * diff;
        acc = acc + sq;
    }
    return acc;
}

// Exact top-k by (d64, id) over rows [0, n) of base, for nq queries.
int orc_exact_topk_f64(const float* base, int64_t n, int64_t dim, const float* Q, int nq, int k,
                       int64_t* out_ids, double* out_d, int nthreads) {
    int kk = k > (int)n ? (int)n : k;
    if (nthreads < 1) nthreads = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) {
        th.emplace_back([=]() {
            typedef std::pair<double, int64_t> E;
            for (int qi = t; qi < nq; qi += nthreads) {
                const float* q = Q + (int64_t)qi * dim;
                std::priority_queue<E> heap;  // max-heap on (d, id)
                for (int64_t i = 0; i < n; ++i) {
                    E e(orc_l2_f64(base + i * dim, q, dim), i);
                    if ((int)heap.size() < kk) {
                        heap.push(e);
                    } else if (e < heap.top()) {
                        heap.pop();
                        heap.push(e);
                    }
                }
                for (int j = kk - 1; j >= 0; --j) {
                    out_ids[(int64_t)qi * k + j] = heap.top().second;
                    out_d[(int64_t)qi * k + j] = heap.top().first;
                    heap.pop();
                }
            }
        });
    }
    for (auto& x : th) x.join();
    return kk;
}

// Exact top-k restricted to a candidate subset (row ids cand[0..nc)).
int orc_exact_topk_subset_f64(const float* base, int64_t dim, const int64_t* cand, int64_t nc, const float* q,
                              int k, int64_t* out_ids, double* out_d) {
    std::vector<std::pair<double, int64_t>> v((size_t)nc);
    for (int64_t j = 0; j < nc; ++j) v[(size_t)j] = {orc_l2_f64(base + cand[j] * dim, q, dim), cand[j]};
    int kk = k > (int)nc ? (int)nc : k;
    std::partial_sort(v.begin(), v.begin() + kk, v.end());
    for (int j = 0; j < kk; ++j) {
        out_ids[j] = v[(size_t)j].second;
        out_d[j] = v[(size_t)j].first;
    }
    return kk;
}

// gpu_filter.cu:161-170: scalar float rerank, acc += v*v (separately rounded
// product; the reference is compiled by nvcc/host g++ without -ffast-math).
float orc_f32_direct_l2(const float* a, const float* b, int64_t dim) {
    float acc = 0.0f;
    for (int64_t j = 0; j < dim; ++j) {
        float v = a[j] - b[j];
        float sq = v * v;
        acc = acc + sq;
    }
    return acc;
}

// gpu_filter.cu:52-63 (sq8pd_scan): the key of each of n dim-ordered rows,
// v = fma(code, scale, -q_shift) and acc = fma(v, v, acc) over dims in order
// (one CUDA thread per row, serial fmaf).  codes row-major [n x dim].
void orc_gf_keys(const uint8_t* codes, int64_t n, int64_t dim, const float* q_shift, const float* scale,
                 float* keys) {
    for (int64_t i = 0; i < n; ++i) {
        float acc = 0.0f;
        for (int64_t d = 0; d < dim; ++d) {
            float v = fmaf((float)codes[i * dim + d], scale[d], -q_shift[d]);
            acc = fmaf(v, v, acc);
        }
        keys[i] = acc;
    }
}

// ---------------------------------------------------------------- SQ8 ----
// Per-dim min/max training.  variant 0 = ivf_lsh/hnsw (lsh_index_optimized.cpp:189-204):
// constant dim -> scale 1, offset 0.  variant 1 = gpu_filter (gpu_filter.cu:102-118):
// constant dim -> scale 1, offset min.  Range threshold 1e-8 in both.
void orc_sq8_train(const float* x, int64_t n, int64_t dim, int variant, float* scale, float* offset) {
    for (int64_t d = 0; d < dim; ++d) {
        float mn = std::numeric_limits<float>::max();
        float mx = std::numeric_limits<float>::lowest();
        for (int64_t i = 0; i < n; ++i) {
            float v = x[i * dim + d];
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
        if (mx - mn > 1e-8f) {
            scale[d] = (mx - mn) / 255.0f;
            offset[d] = mn;
        } else {
            scale[d] = 1.0f;
            offset[d] = variant == 0 ? 0.0f : mn;
        }
    }
}

// Row-major codes.  variant 0: int(std::round(...)) (half away from zero),
// lsh_index_optimized.cpp:208-213.  variant 1: lrintf (half to even),
// gpu_filter.cu:121-126.  Both clamp to [0, 255].
void orc_sq8_encode(const float* x, int64_t n, int64_t dim, int variant, const float* scale,
                    const float* offset, uint8_t* codes) {
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t d = 0; d < dim; ++d) {
            float t = (x[i * dim + d] - offset[d]) / scale[d];
            int qv = variant == 0 ? (int)std::round(t) : (int)lrintf(t);
            codes[i * dim + d] = (uint8_t)std::max(0, std::min(255, qv));
        }
    }
}

// competitors/ivf_lsh/src/simd.h:118-147 (AVX-512 branch) without a threshold:
// diff = fma(code, scale, -q_shifted); lane acc = fma(diff, diff, acc).
float orc_sq8_l2(const uint8_t* a, const float* q_shifted, const float* scale, int64_t dim) {
    float s[16];
    for (int l = 0; l < 16; ++l) s[l] = 0.0f;
    int64_t i = 0;
    for (; i + 15 < dim; i += 16) {
        for (int l = 0; l < 16; ++l) {
            float diff = fmaf((float)a[i + l], scale[i + l], -q_shifted[i + l]);
            s[l] = fmaf(diff, diff, s[l]);
        }
    }
    float sum = reduce16(s);
    // dim % 16 tail: contraction of the reference's scalar tail is compiler
    // dependent and not pinned; every configuration here has dim % 16 == 0.
    for (; i < dim; ++i) {
        float diff = fmaf((float)a[i], scale[i], -q_shifted[i]);
        sum = fmaf(diff, diff, sum);
    }
    return sum;
}

// Exact (double) SQ8 distance of the dequantised row: sum (c*s - q')^2.
double orc_sq8_l2_f64(const uint8_t* a, const float* q_shifted, const float* scale, int64_t dim) {
    double acc = 0.0;
    for (int64_t j = 0; j < dim; ++j) {
        double diff = (double)a[j] * (double)scale[j] - (double)q_shifted[j];
        acc = acc + diff * diff;
    }
    return acc;
}

// ---------------------------------------------------------------- SQ4 ----
// submission/src/hnsw.cpp:147-169: dim_pad = ceil(dim/32)*32, per-dim /15
// scale; constant dim -> scale 1, offset min; padded dims scale 0 offset 0.
int64_t orc_sq4_train(const float* x, int64_t n, int64_t dim, float* scale_pad, float* offset_pad) {
    int64_t dim_pad = (dim + 31) / 32 * 32;
    for (int64_t d = 0; d < dim_pad; ++d) {
        scale_pad[d] = 0.0f;
        offset_pad[d] = 0.0f;
    }
    for (int64_t d = 0; d < dim; ++d) {
        float mn = std::numeric_limits<float>::max();
        float mx = std::numeric_limits<float>::lowest();
        for (int64_t i = 0; i < n; ++i) {
            float v = x[i * dim + d];
            if (v < mn) mn = v;
            if (v > mx) mx = v;
        }
        if (mx - mn > 1e-8f) {
            scale_pad[d] = (mx - mn) / 15.0f;
            offset_pad[d] = mn;
        } else {
            scale_pad[d] = 1.0f;
            offset_pad[d] = mn;
        }
    }
    return dim_pad;
}

// submission/src/hnsw.cpp:170-182: byte j of 16-byte block blk holds dim
// 32*blk+j in the low nibble and dim 32*blk+16+j in the high nibble.
void orc_sq4_encode(const float* x, int64_t n, int64_t dim, const float* scale_pad, const float* offset_pad,
                    uint8_t* codes) {
    int64_t dim_pad = (dim + 31) / 32 * 32;
    int64_t stride = dim_pad / 2;
    std::memset(codes, 0, (size_t)(n * stride));
    for (int64_t i = 0; i < n; ++i) {
        uint8_t* c = codes + i * stride;
        for (int64_t d = 0; d < dim; ++d) {
            int nib = (int)std::round((x[i * dim + d] - offset_pad[d]) / scale_pad[d]);
            nib = std::max(0, std::min(15, nib));
            int64_t blk = d / 32, r = d % 32;
            uint8_t& byte = c[blk * 16 + (r < 16 ? r : r - 16)];
            byte = (r < 16) ? (uint8_t)(byte | nib) : (uint8_t)(byte | (nib << 4));
        }
    }
}

// submission/src/simd.h:297-318 (AVX-512 branch, no threshold): per 32-dim
// block, low-nibble run then high-nibble run, each fmsub + fmadd into the
// same 16 lanes; tree reduce.
float orc_sq4_l2(const uint8_t* a, const float* q_shifted_pad, const float* scale_pad, int64_t dim_pad) {
    float s[16];
    for (int l = 0; l < 16; ++l) s[l] = 0.0f;
    for (int64_t d = 0; d < dim_pad; d += 32) {
        const uint8_t* blk = a + d / 2;
        for (int l = 0; l < 16; ++l) {
            float lo = (float)(blk[l] & 0x0F);
            float dl = fmaf(lo, scale_pad[d + l], -q_shifted_pad[d + l]);
            s[l] = fmaf(dl, dl, s[l]);
        }
        for (int l = 0; l < 16; ++l) {
            float hi = (float)(blk[l] >> 4);
            float dh = fmaf(hi, scale_pad[d + 16 + l], -q_shifted_pad[d + 16 + l]);
            s[l] = fmaf(dh, dh, s[l]);
        }
    }
    return reduce16(s);
}

// Nibble of dim d in an SQ4 row (layout above).
int orc_sq4_nibble(const uint8_t* a, int64_t d) {
    int64_t blk = d / 32, r = d % 32;
    uint8_t byte = a[blk * 16 + (r < 16 ? r : r - 16)];
    return r < 16 ? (byte & 0x0F) : (byte >> 4);
}

// competitors/ivf_lsh/src/lsh_index_optimized.cpp:553-561: initial centroid
// rows = shuffle(iota(n), mt19937(seed))[c % n] (libstdc++ std::shuffle, as the
// reference build uses).  Also returns rand_indices for the empty-cluster rule.
void orc_kmeans_init(int64_t n, int nlist, uint32_t seed, int64_t* init_rows, int64_t* rand_indices) {
    std::vector<int> r((size_t)n);
    std::iota(r.begin(), r.end(), 0);
    std::shuffle(r.begin(), r.end(), std::mt19937(seed));
    for (int c = 0; c < nlist; ++c) init_rows[c] = r[(size_t)(c % n)];
    if (rand_indices)
        for (int64_t i = 0; i < n; ++i) rand_indices[i] = r[(size_t)i];
}

// competitors/ivf_lsh/src/lsh_index_optimized.cpp:551-616 (run_kmeans) with
// the workers :32-73, one thread: initial centroids from orc_kmeans_init;
// each round assigns every row to the argmin of compute_l2_distance
// (src/simd.h:32-64 order, strict <: the lowest centroid wins a tie), sums
// the rows of each cluster in float in row order, and sets the centroid to
// sum / count, or to row rand_indices[(c * 37) % n] when the cluster is
// empty.  The reference sums per thread and then over threads, so its float
// sums match these only where they are exact (e.g. integer coordinates with
// sums below 2^24); its assignment matches wherever the centroids do.
// n_clusters is clamped to n (:140).  centroids: [nlist x dim], assign: [n].
void orc_ivf_kmeans(const float* data, int64_t n, int64_t dim, int nlist, int iters, uint32_t seed,
                    float* centroids, int32_t* assign) {
    if (nlist > n) nlist = (int)n;
    std::vector<int64_t> init((size_t)nlist), ri((size_t)n);
    orc_kmeans_init(n, nlist, seed, init.data(), ri.data());
    for (int c = 0; c < nlist; ++c)
        std::memcpy(centroids + (int64_t)c * dim, data + init[(size_t)c] * dim, (size_t)dim * sizeof(float));
    std::vector<float> sums((size_t)nlist * dim);
    std::vector<int64_t> counts((size_t)nlist);
    for (int it = 0; it < iters; ++it) {
        for (int64_t i = 0; i < n; ++i) {
            float best = std::numeric_limits<float>::max();
            int bc = 0;
            for (int c = 0; c < nlist; ++c) {
                const float dd = orc_l2_linear_index(data + i * dim, centroids + (int64_t)c * dim, dim,
                                                     std::numeric_limits<float>::max());
                if (dd < best) {
                    best = dd;
                    bc = c;
                }
            }
            assign[i] = bc;
        }
        std::fill(sums.begin(), sums.end(), 0.0f);
        std::fill(counts.begin(), counts.end(), 0);
        for (int64_t i = 0; i < n; ++i) {
            const int c = assign[i];
            counts[(size_t)c]++;
            for (int64_t d = 0; d < dim; ++d) sums[(size_t)(c * dim + d)] += data[i * dim + d];
        }
        for (int c = 0; c < nlist; ++c) {
            if (counts[(size_t)c] > 0) {
                for (int64_t d = 0; d < dim; ++d)
                    centroids[(int64_t)c * dim + d] = sums[(size_t)(c * dim + d)] / (float)counts[(size_t)c];
            } else {
                const int64_t r = ri[(size_t)((c * 37) % n)];
                std::memcpy(centroids + (int64_t)c * dim, data + r * dim, (size_t)dim * sizeof(float));
            }
        }
    }
}

}  // extern "C"

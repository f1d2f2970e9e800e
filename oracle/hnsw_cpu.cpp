// CPU HNSW baseline: the reference's HNSW (submission/src/hnsw.cpp) restated
// with 64-bit offsets, so it builds and answers on 10M x 384 rows, where the
// reference's int arithmetic (node * dim_, npts_ * dim_ at
// submission/src/hnsw.cpp:86, 389, 410) overflows past 2^31 floats.
//
// TEST INFRASTRUCTURE ONLY: a CPU baseline timed beside the GPU path by
// bench.py's cpu_baseline leg and checked by tests/; the product
// (annvedi_amd/) never links or loads it.
//
// The algorithm follows the reference step for step:
//   fit       submission/src/hnsw.cpp:60-212   SQ8 codes (global scale, "sq8";
//                                              per-dim, "sq8pd"), first node
//                                              alone, the rest in parallel
//   levels    :389-393    floor(-ln(U + 1e-10) / ln(M)), at most 32
//   search    :399-487    beam search: candidate min-heap, result max-heap of
//                         ef, stop when the nearest candidate is farther than
//                         the worst result; distances in the build metric
//                         (exact fp32, or SQ8 with early abandon at the ef-th
//                         distance)
//   select    :526-553    heuristic (Malkov & Yashunin Alg. 4): keep a
//                         candidate closer to the base than to every kept
//                         neighbour, by exact fp32 distances; back-fill with
//                         the nearest discarded (keepPrunedConnections)
//   prune     :570-582    a full list and a new link re-selected by exact
//                         fp32 distances to the node
//   insert    :587-645    descent with ef 1, then per layer a beam of
//                         ef_construction, M (2M on layer 0) neighbours,
//                         links both ways under per-node locks
//   reorder   :335-386    node ids renumbered in BFS order over layer 0 from
//                         the entry point; answers map back to input ids
//   query     :197-243    descent with ef 1, beam of max(ef, k) on layer 0,
//                         the window re-ranked by exact fp32 distance in the
//                         quantized modes, sorted, first k
//
// Own layout: layer 0 adjacency is one flat array of n x (1 + 2M) int32
// (count, then neighbours) instead of a vector per node and layer; upper
// layers (about n / M nodes) keep one small array per node.  Node levels
// are drawn up front from one mt19937 seeded by the caller (the reference
// seeds each build thread from std::random_device), so a build is
// reproducible up to the order in which threads link nodes.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <queue>
#include <random>
#include <thread>
#include <utility>
#include <vector>

namespace {

enum Metric { kFloat = 0, kSq8 = 1, kSq8pd = 2 };

typedef std::pair<float, int32_t> DistId;  // (distance, node): ordered by distance, then node
typedef std::priority_queue<DistId> FarFirst;
typedef std::priority_queue<DistId, std::vector<DistId>, std::greater<DistId>> NearFirst;

constexpr float kInf = std::numeric_limits<float>::max();
constexpr int kLocks = 1 << 16;

// Squared L2 between two fp32 rows, stopping once a partial sum over a
// multiple of 64 dims reaches `limit` (the value returned is then that
// partial sum, >= limit, as the reference's early abandon returns it).
inline float l2_f32(const float* a, const float* b, int64_t d, float limit) {
    float s = 0.f;
    int64_t j = 0;
    for (; j + 64 <= d; j += 64) {
        float t = 0.f;
        for (int u = 0; u < 64; ++u) {
            const float v = a[j + u] - b[j + u];
            t += v * v;
        }
        s += t;
        if (s >= limit) return s;
    }
    for (; j < d; ++j) {
        const float v = a[j] - b[j];
        s += v * v;
    }
    return s;
}

// Global-scale SQ8: sum (c * scale + offset - q)^2 (simd.h:118-165).
inline float l2_sq8(const uint8_t* c, const float* q, int64_t d, float scale, float offset, float limit) {
    float s = 0.f;
    int64_t j = 0;
    for (; j + 64 <= d; j += 64) {
        float t = 0.f;
        for (int u = 0; u < 64; ++u) {
            const float v = (float)c[j + u] * scale + offset - q[j + u];
            t += v * v;
        }
        s += t;
        if (s >= limit) return s;
    }
    for (; j < d; ++j) {
        const float v = (float)c[j] * scale + offset - q[j];
        s += v * v;
    }
    return s;
}

// Per-dim SQ8: sum (c_d * scale_d - (q_d - offset_d))^2 with the query
// shifted once per search (lsh_index_optimized.cpp:320-323).
inline float l2_sq8pd(const uint8_t* c, const float* qs, int64_t d, const float* scale, float limit) {
    float s = 0.f;
    int64_t j = 0;
    for (; j + 64 <= d; j += 64) {
        float t = 0.f;
        for (int u = 0; u < 64; ++u) {
            const float v = (float)c[j + u] * scale[j + u] - qs[j + u];
            t += v * v;
        }
        s += t;
        if (s >= limit) return s;
    }
    for (; j < d; ++j) {
        const float v = (float)c[j] * scale[j] - qs[j];
        s += v * v;
    }
    return s;
}

// Visited marks by generation (the reference's SearchTracker), and a
// neighbour-list buffer of the searching thread.
struct Visited {
    std::vector<uint32_t> gen;
    uint32_t cur = 0;
    std::vector<int32_t> nb;
    Visited(int64_t n, int m0) : gen((size_t)n, 0u), nb((size_t)m0) {}
    void next() {
        if (++cur == 0) {
            std::fill(gen.begin(), gen.end(), 0u);
            cur = 1;
        }
    }
};

struct Graph {
    int64_t n = 0;
    int64_t d = 0;
    int M = 16, M0 = 32, efc = 100;
    bool heuristic = false;
    int metric = kFloat;
    std::vector<float> x;        // n x d
    std::vector<uint8_t> codes;  // n x d (quantized metrics)
    float g_scale = 1.f, g_offset = 0.f;
    std::vector<float> pd_scale, pd_offset;
    std::vector<int32_t> l0;                  // n x (1 + M0): count, neighbours
    std::vector<int8_t> level;                // top layer of each node
    std::vector<std::vector<int32_t>> upper;  // per node: layers 1..level, (1 + M) each
    std::vector<int32_t> new2orig;            // after reorder: node -> input id
    int32_t entry = 0;
    int top = 0;
    bool building = false;
    std::mutex glock;
    std::vector<std::mutex> locks = std::vector<std::mutex>(kLocks);
    std::atomic<int64_t> ndist{0};

    const float* row(int64_t i) const { return x.data() + i * d; }
    const uint8_t* code(int64_t i) const { return codes.data() + i * d; }
    int32_t* list(int64_t i, int layer) {
        return layer == 0 ? l0.data() + i * (1 + M0) : upper[(size_t)i].data() + (int64_t)(layer - 1) * (1 + M);
    }
    int cap(int layer) const { return layer == 0 ? M0 : M; }

    // distance in the build metric; qs is the shifted query (sq8pd)
    float dist(int64_t i, const float* q, const float* qs, float limit) const {
        switch (metric) {
            case kSq8: return l2_sq8(code(i), q, d, g_scale, g_offset, limit);
            case kSq8pd: return l2_sq8pd(code(i), qs, d, pd_scale.data(), limit);
            default: return l2_f32(row(i), q, d, limit);
        }
    }

    // a node's neighbour list at a layer, copied under its lock while building
    int neighbours(int64_t c, int layer, int32_t* out) {
        if (building) {
            std::lock_guard<std::mutex> lk(locks[(size_t)(c % kLocks)]);
            const int32_t* l = list(c, layer);
            std::memcpy(out, l + 1, sizeof(int32_t) * (size_t)l[0]);
            return l[0];
        }
        const int32_t* l = list(c, layer);
        std::memcpy(out, l + 1, sizeof(int32_t) * (size_t)l[0]);
        return l[0];
    }

    std::vector<DistId> search_layer(const float* q, const float* qs, int32_t ep, int ef, int layer, bool count,
                                     Visited& vis) {
        vis.next();
        vis.gen[(size_t)ep] = vis.cur;
        const float d_ep = dist(ep, q, qs, kInf);
        int64_t nd = 1;
        NearFirst cand;
        FarFirst win;
        cand.push({d_ep, ep});
        win.push({d_ep, ep});
        int32_t* nb = vis.nb.data();
        while (!cand.empty()) {
            const DistId c = cand.top();
            cand.pop();
            if (c.first > win.top().first) break;
            const int cnt = neighbours(c.second, layer, nb);
            for (int t = 0; t < cnt; ++t) {
                const int32_t e = nb[t];
                if (vis.gen[(size_t)e] == vis.cur) continue;
                vis.gen[(size_t)e] = vis.cur;
                const float lim = (int)win.size() >= ef ? win.top().first : kInf;
                const float de = dist(e, q, qs, lim);
                ++nd;
                if ((int)win.size() < ef || de < win.top().first) {
                    cand.push({de, e});
                    win.push({de, e});
                    if ((int)win.size() > ef) win.pop();
                }
            }
        }
        if (count) ndist += nd;
        std::vector<DistId> out;
        out.reserve(win.size());
        while (!win.empty()) {
            out.push_back(win.top());
            win.pop();
        }
        return out;
    }

    // neighbour selection over candidates sorted nearest first (build-metric
    // distances to the base); the heuristic compares exact fp32 distances
    void select(const std::vector<DistId>& cand, int mmax, std::vector<int32_t>& out) const {
        out.clear();
        if (!heuristic || (int)cand.size() <= mmax) {
            for (const auto& p : cand) {
                if ((int)out.size() >= mmax) break;
                out.push_back(p.second);
            }
            return;
        }
        std::vector<int32_t> dropped;
        for (const auto& p : cand) {
            if ((int)out.size() >= mmax) break;
            const float* pe = row(p.second);
            bool keep = true;
            for (int32_t s : out)
                if (l2_f32(row(s), pe, d, p.first) < p.first) {
                    keep = false;
                    break;
                }
            (keep ? out : dropped).push_back(p.second);
        }
        for (size_t i = 0; i < dropped.size() && (int)out.size() < mmax; ++i) out.push_back(dropped[i]);
    }

    void shift(const float* q, std::vector<float>& qs) const {
        if (metric != kSq8pd) return;
        qs.resize((size_t)d);
        for (int64_t j = 0; j < d; ++j) qs[(size_t)j] = q[j] - pd_offset[(size_t)j];
    }

    void insert(int64_t idx, Visited& vis, std::vector<float>& qs) {
        const int l = level[(size_t)idx];
        int32_t ep;
        int maxl;
        {
            std::lock_guard<std::mutex> lk(glock);
            ep = entry;
            maxl = top;
        }
        const float* q = row(idx);
        shift(q, qs);
        for (int lc = maxl; lc > l; --lc) ep = search_layer(q, qs.data(), ep, 1, lc, false, vis)[0].second;
        std::vector<int32_t> chosen;
        for (int lc = std::min(l, maxl); lc >= 0; --lc) {
            std::vector<DistId> w = search_layer(q, qs.data(), ep, efc, lc, false, vis);
            std::sort(w.begin(), w.end());
            select(w, cap(lc), chosen);
            ep = chosen[0];
            {
                std::lock_guard<std::mutex> lk(locks[(size_t)(idx % kLocks)]);
                int32_t* li = list(idx, lc);
                li[0] = (int32_t)chosen.size();
                std::memcpy(li + 1, chosen.data(), sizeof(int32_t) * chosen.size());
            }
            for (int32_t nb : chosen) {
                std::lock_guard<std::mutex> lk(locks[(size_t)(nb % kLocks)]);
                int32_t* ln = list(nb, lc);
                if (ln[0] < cap(lc)) {
                    ln[1 + ln[0]] = (int32_t)idx;
                    ++ln[0];
                } else {
                    // full: the list and the new link are re-selected (the
                    // reference appends, then prunes back to the cap,
                    // hnsw.cpp:637-640, 570-582)
                    std::vector<DistId> cand;
                    cand.reserve((size_t)ln[0] + 1);
                    const float* pn = row(nb);
                    for (int t = 0; t < ln[0]; ++t) cand.push_back({l2_f32(row(ln[1 + t]), pn, d, kInf), ln[1 + t]});
                    cand.push_back({l2_f32(q, pn, d, kInf), (int32_t)idx});
                    std::sort(cand.begin(), cand.end());
                    std::vector<int32_t> keep;
                    select(cand, cap(lc), keep);
                    ln[0] = (int32_t)keep.size();
                    std::memcpy(ln + 1, keep.data(), sizeof(int32_t) * keep.size());
                }
            }
        }
        if (l > maxl) {
            std::lock_guard<std::mutex> lk(glock);
            if (l > top) {
                entry = (int32_t)idx;
                top = l;
            }
        }
    }

    void encode() {
        const int64_t nd = n * d;
        if (metric == kSq8) {
            float lo = kInf, hi = -kInf;
            for (int64_t i = 0; i < nd; ++i) {
                lo = std::min(lo, x[(size_t)i]);
                hi = std::max(hi, x[(size_t)i]);
            }
            if (hi - lo > 1e-8f) {
                g_scale = (hi - lo) / 255.0f;
                g_offset = lo;
            } else {
                g_scale = 1.f;
                g_offset = 0.f;
            }
            codes.resize((size_t)nd);
            for (int64_t i = 0; i < nd; ++i) {
                const int v = (int)std::round((x[(size_t)i] - g_offset) / g_scale);
                codes[(size_t)i] = (uint8_t)std::max(0, std::min(255, v));
            }
        } else if (metric == kSq8pd) {
            pd_scale.assign((size_t)d, 1.f);
            pd_offset.assign((size_t)d, 0.f);
            std::vector<float> lo((size_t)d, kInf), hi((size_t)d, -kInf);
            for (int64_t i = 0; i < n; ++i)
                for (int64_t j = 0; j < d; ++j) {
                    const float v = x[(size_t)(i * d + j)];
                    lo[(size_t)j] = std::min(lo[(size_t)j], v);
                    hi[(size_t)j] = std::max(hi[(size_t)j], v);
                }
            for (int64_t j = 0; j < d; ++j)
                if (hi[(size_t)j] - lo[(size_t)j] > 1e-8f) {
                    pd_scale[(size_t)j] = (hi[(size_t)j] - lo[(size_t)j]) / 255.0f;
                    pd_offset[(size_t)j] = lo[(size_t)j];
                }
            codes.resize((size_t)nd);
            for (int64_t i = 0; i < n; ++i)
                for (int64_t j = 0; j < d; ++j) {
                    const int v = (int)std::round((x[(size_t)(i * d + j)] - pd_offset[(size_t)j]) / pd_scale[(size_t)j]);
                    codes[(size_t)(i * d + j)] = (uint8_t)std::max(0, std::min(255, v));
                }
        }
    }

    void reorder() {
        std::vector<int32_t> order;
        order.reserve((size_t)n);
        std::vector<uint8_t> seen((size_t)n, 0);
        order.push_back(entry);
        seen[(size_t)entry] = 1;
        for (size_t h = 0; h < order.size(); ++h) {
            const int32_t* l = list(order[h], 0);
            for (int t = 0; t < l[0]; ++t)
                if (!seen[(size_t)l[1 + t]]) {
                    seen[(size_t)l[1 + t]] = 1;
                    order.push_back(l[1 + t]);
                }
        }
        for (int64_t i = 0; i < n; ++i)
            if (!seen[(size_t)i]) order.push_back((int32_t)i);
        std::vector<int32_t> nid((size_t)n);
        for (int64_t i = 0; i < n; ++i) nid[(size_t)order[(size_t)i]] = (int32_t)i;
        {
            std::vector<float> t((size_t)(n * d));
            for (int64_t i = 0; i < n; ++i) std::memcpy(&t[(size_t)(i * d)], row(order[(size_t)i]), sizeof(float) * (size_t)d);
            x.swap(t);
        }
        if (!codes.empty()) {
            std::vector<uint8_t> t((size_t)(n * d));
            for (int64_t i = 0; i < n; ++i) std::memcpy(&t[(size_t)(i * d)], code(order[(size_t)i]), (size_t)d);
            codes.swap(t);
        }
        {
            std::vector<int32_t> t((size_t)(n * (1 + M0)));
            for (int64_t i = 0; i < n; ++i) {
                const int32_t* s = l0.data() + (int64_t)order[(size_t)i] * (1 + M0);
                int32_t* o = t.data() + i * (1 + M0);
                o[0] = s[0];
                for (int k = 0; k < s[0]; ++k) o[1 + k] = nid[(size_t)s[1 + k]];
            }
            l0.swap(t);
        }
        {
            std::vector<std::vector<int32_t>> t((size_t)n);
            std::vector<int8_t> lv((size_t)n);
            for (int64_t i = 0; i < n; ++i) {
                t[(size_t)i] = std::move(upper[(size_t)order[(size_t)i]]);
                lv[(size_t)i] = level[(size_t)order[(size_t)i]];
                for (int layer = 1; layer <= lv[(size_t)i]; ++layer) {
                    int32_t* l = t[(size_t)i].data() + (int64_t)(layer - 1) * (1 + M);
                    for (int k = 0; k < l[0]; ++k) l[1 + k] = nid[(size_t)l[1 + k]];
                }
            }
            upper.swap(t);
            level.swap(lv);
        }
        entry = nid[(size_t)entry];
        new2orig = std::move(order);
    }

    int query(const float* q, int k, int ef, int64_t* out, Visited& vis, std::vector<float>& qs) {
        ef = std::max(ef, k);
        shift(q, qs);
        int32_t ep = entry;
        for (int lc = top; lc > 0; --lc) ep = search_layer(q, qs.data(), ep, 1, lc, true, vis)[0].second;
        std::vector<DistId> w = search_layer(q, qs.data(), ep, ef, 0, true, vis);
        if (metric != kFloat)
            for (auto& p : w) p.first = l2_f32(row(p.second), q, d, kInf);
        std::sort(w.begin(), w.end());
        const int m = std::min(k, (int)w.size());
        for (int i = 0; i < m; ++i) out[i] = new2orig.empty() ? w[(size_t)i].second : new2orig[(size_t)w[(size_t)i].second];
        return m;
    }
};

}  // namespace

extern "C" {

// Build an HNSW over n x d fp32 rows (copied).  metric: 0 fp32, 1 SQ8 with
// one global scale ("sq8"), 2 per-dim SQ8 ("sq8pd").  threads <= 0: all
// hardware threads.  progress > 0: a line on stderr each time another
// `progress` nodes are linked (long builds).  Returns NULL on bad arguments.
void* hnsw_cpu_build(const float* x, int64_t n, int32_t d, int32_t M, int32_t efc, int32_t metric,
                     int32_t heuristic, int32_t reorder, int32_t threads, uint64_t seed, int64_t progress) {
    if (!x || n <= 0 || n > INT32_MAX || d <= 0 || M < 2 || efc < 1 || metric < 0 || metric > 2) return nullptr;
    Graph* g = new Graph();
    g->n = n;
    g->d = d;
    g->M = M;
    g->M0 = 2 * M;
    g->efc = efc;
    g->heuristic = heuristic != 0;
    g->metric = metric;
    g->x.assign(x, x + n * (int64_t)d);
    g->encode();
    g->l0.assign((size_t)(n * (1 + g->M0)), 0);
    g->level.resize((size_t)n);
    g->upper.resize((size_t)n);
    std::mt19937 rng((uint32_t)seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    const double ml = 1.0 / std::log((double)M);
    for (int64_t i = 0; i < n; ++i) {
        const int l = std::min((int)(-std::log(uni(rng) + 1e-10) * ml), 32);
        g->level[(size_t)i] = (int8_t)l;
        if (l > 0) g->upper[(size_t)i].assign((size_t)l * (1 + M), 0);
    }
    g->entry = 0;
    g->top = g->level[0];
    g->building = true;
    int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    if (nt <= 0) nt = 4;
    std::atomic<int64_t> next{1};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([g, &next, progress]() {
            Visited vis(g->n, g->M0);
            std::vector<float> qs;
            for (int64_t i = next++; i < g->n; i = next++) {
                g->insert(i, vis, qs);
                if (progress > 0 && i % progress == 0) {
                    fprintf(stderr, "[hnsw_cpu] linking node %lld of %lld\n", (long long)i, (long long)g->n);
                    fflush(stderr);
                }
            }
        });
    for (auto& th : pool) th.join();
    g->building = false;
    if (reorder) g->reorder();
    return g;
}

// One query: up to k input ids (nearest first) into out; returns how many.
int32_t hnsw_cpu_query(void* h, const float* q, int32_t k, int32_t ef, int64_t* out) {
    if (!h || !q || !out || k <= 0) return -1;
    Graph* g = static_cast<Graph*>(h);
    // the calling thread's visited marks, kept across calls (n words)
    thread_local std::unique_ptr<Visited> vis;
    thread_local std::vector<float> qs;
    if (!vis || vis->gen.size() != (size_t)g->n || vis->nb.size() != (size_t)g->M0) vis.reset(new Visited(g->n, g->M0));
    return g->query(q, k, ef, out, *vis, qs);
}

// nq queries (rows of Q) on `threads` threads, one query at a time each;
// out is nq x k (rows past a short answer hold -1).  Returns 0.
int32_t hnsw_cpu_query_batch(void* h, const float* Q, int64_t nq, int32_t k, int32_t ef, int64_t* out,
                             int32_t threads) {
    if (!h || !Q || !out || k <= 0 || nq < 0) return -1;
    Graph* g = static_cast<Graph*>(h);
    const int nt = threads > 0 ? threads : 1;
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&]() {
            Visited vis(g->n, g->M0);
            std::vector<float> qs;
            for (int64_t i = next++; i < nq; i = next++) {
                int64_t* o = out + i * k;
                const int m = g->query(Q + i * g->d, k, ef, o, vis, qs);
                for (int j = m; j < k; ++j) o[j] = -1;
            }
        });
    for (auto& th : pool) th.join();
    return 0;
}

int64_t hnsw_cpu_ndist(void* h) { return h ? static_cast<Graph*>(h)->ndist.load() : -1; }

void hnsw_cpu_free(void* h) { delete static_cast<Graph*>(h); }

}  // extern "C"

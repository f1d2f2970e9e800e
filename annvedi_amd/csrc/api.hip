// C ABI of the k-NN search library (declared in include/annmi.h).
//
// Query pipeline per step of up to max_batch queries, all on the handle's
// stream, one host synchronisation at the end:
//   prep_queries                 scan operand, norms, exact-mode slack
//   scan(rows [b0,b1)) with thr = +inf      every row of the first range
//   select_threshold + scan(rows [b_j, b_j+1))  geometric ranges: the filter is
//                                the rsel-th smallest candidate so far
//   finalize                     collect, fp32 re-score, sort, fp64 near-ties
//
// Correctness of the filter: a row is dropped from a range only when its
// approximate distance exceeds the rsel-th smallest approximate distance
// among rows already scanned (plus, in exact mode, twice the rigorous error
// bound of the approximation), so no row of the true top-k is ever dropped.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/annmi.h"
#include "common.h"
#include "ivf.h"
#include "robust.h"

#include <numeric>
#include <random>

hipError_t launch_encode(int kind, const float* src, int64_t sstride, int64_t n, int d, int kdim, void* dst, float* xn,
                         uint32_t* maxkey, float alpha, const float* scale, const float* offset, int variant,
                         hipStream_t st, const uint32_t* gather = nullptr, int tile_cb = 0);
hipError_t launch_dim_minmax(const float* src, int64_t sstride, int64_t n, int d, uint32_t* mn, uint32_t* mx,
                             hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st);
hipError_t launch_fill_f32(float* p, float v, int64_t n, hipStream_t st);
hipError_t launch_prep_queries(int kind, const float* Q, int B, int d, int kdim, void* Qscan, float* qn, float* qscale,
                               float* margin, float maxnorm, float alpha_x, const float* scale, const float* offset,
                               hipStream_t st, uint32_t* cnt = nullptr, uint32_t* flags = nullptr,
                               float* thr = nullptr);
hipError_t launch_linear_index_query(const float* X, int64_t xstride, int64_t n, int d, const float* dq, int k,
                                     float* P, float* BT, float* R, float* bkey, uint32_t* bid, int64_t* d_ids,
                                     float* d_vals, uint32_t* d_flag, hipStream_t st);

// armed timing events of the next ANNMI_LAUNCH on this thread (common.h)
thread_local hipEvent_t t_launch_ev[2] = {nullptr, nullptr};

namespace {

thread_local char g_err[512] = "";

void set_err(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

#define HIP_OR(expr, ret)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            set_err("HIP error: %s at %s:%d (%s)", hipGetErrorString(e_), __FILE__, __LINE__, #expr); \
            return ret;                                                                           \
        }                                                                                         \
    } while (0)

constexpr int kMaxDim = 4096;
constexpr int kCapDefault = 65536;

int round_up(int x, int m) { return (x + m - 1) / m * m; }

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

int scan_kind_of(int mode) {
    switch (mode) {
        case ANNMI_EXACT_F32: return SCAN_F32;
        case ANNMI_BF16: return SCAN_BF16;
        case ANNMI_FP8: return SCAN_FP8;
        case ANNMI_SQ8: return SCAN_SQ8;
        case ANNMI_SQ4: return SCAN_SQ4;
        case ANNMI_IVF_SQ8: return SCAN_SQ8;
    }
    return -1;
}

int chunk_of(int kind) {
    switch (kind) {
        case SCAN_F32: return 16;
        case SCAN_BF16: return 32;
        default: return 64;
    }
}

int64_t scan_row_bytes(int kind, int kdim) {
    switch (kind) {
        case SCAN_F32: return (int64_t)kdim * 4;
        case SCAN_BF16: return (int64_t)kdim * 2;
        case SCAN_SQ4: return kdim / 2;
        default: return kdim;
    }
}

int64_t scan_q_bytes(int kind, int kdim) {
    switch (kind) {
        case SCAN_F32: return (int64_t)kdim * 4;
        case SCAN_BF16: return (int64_t)kdim * 2;
        default: return kdim;
    }
}

// Chunk bytes of the tiled layout for a scan element type.
int tile_cb_of(int kind) { return (int)scan_row_bytes(kind, chunk_of(kind)); }

// Host conversions between the row-major and the tiled layout (n rows of
// `stride` bytes; the tiled buffer holds whole 16-row tiles).  The tiled SQ8
// copy holds each code as c ^ 0x80 (the int8 c - 128 its MFMAs read); flip
// is 0x80 for it and 0 for the other types.
int tile_flip_of(int kind) { return kind == SCAN_SQ8 ? 0x80 : 0; }
void untile_rows(const uint8_t* tiled, uint8_t* rows, int64_t n, int64_t stride, int cb, int flip) {
    for (int64_t r = 0; r < n; ++r)
        for (int64_t b = 0; b < stride; b += cb) memcpy(rows + r * stride + b, tiled + tiled_off(r, b, stride, cb), (size_t)cb);
    if (flip)
        for (int64_t i = 0; i < n * stride; ++i) rows[i] ^= (uint8_t)flip;
}
void tile_rows(const uint8_t* rows, uint8_t* tiled, int64_t n, int64_t stride, int cb, int flip) {
    for (int64_t r = 0; r < n; ++r)
        for (int64_t b = 0; b < stride; b += cb) {
            uint8_t* dst = tiled + tiled_off(r, b, stride, cb);
            memcpy(dst, rows + r * stride + b, (size_t)cb);
            if (flip)
                for (int i = 0; i < cb; ++i) dst[i] ^= (uint8_t)flip;
        }
}

struct Handle {
    annmi_params p{};
    int dev = 0;  // device current at creation (its stream's device)
    int64_t n = 0;
    int d = 0;
    int kdim32 = 0;
    float* X32 = nullptr;  // rerank base [n x kdim32]
    bool own_x32 = false;
    float* xn32 = nullptr;
    float maxnorm = 0.f;
    // min of xn and max of |xn| over the rows, and the prefilter mode they
    // select (the wide scan's epilogue)
    float xn_lo = -__builtin_inff(), xn_amax = __builtin_inff();
    int prefilter = 1;
    int kind = 0;
    int kdim = 0;
    void* Xs = nullptr;  // scan base
    int tiled = 0;       // Xs in the tiled layout (quantized flat modes)
    bool own_xs = false;
    int64_t xs_stride = 0;
    float* xn = nullptr;
    float alpha_x = 1.f;
    float* sq_scale = nullptr;
    float* sq_offset = nullptr;
    hipStream_t st = nullptr;
    bool own_stream = true;  // false: the parent IVF handle's stream (centroid index)
    int64_t bytes = 0;
    // workspace
    int ws_B = 0;
    int cap = kCapDefault;
    float* dQ32 = nullptr;
    void* dQs = nullptr;
    float* qn = nullptr;
    float* qscale = nullptr;
    float* slack = nullptr;
    float* thr = nullptr;
    uint32_t* cnt = nullptr;
    float* ckey = nullptr;
    uint32_t* crid = nullptr;
    uint32_t* ovf = nullptr;
    uint32_t* nrr = nullptr;
    uint32_t* hpin = nullptr;  // pinned host words: a step's flags, re-score counts, probe flags (3 x ws_B)
    int64_t* d_ids = nullptr;
    float* d_dist = nullptr;
    uint32_t* scratch_keys = nullptr;
    std::vector<uint32_t> last_nrr, last_flags;
    // exact re-run of flagged queries (robust.hip): per-query workspace for
    // up to rb_F queries at a time
    int rb_F = 0;
    float* rb_keys = nullptr;
    uint32_t* rb_hist = nullptr;
    uint64_t* rb_sel = nullptr;
    uint32_t* gf_cid = nullptr;  // gf_query's collected rows when r exceeds the shared list
    int64_t gf_cap = 0;
    uint32_t* rb_cnt = nullptr;
    uint32_t* rb_cid = nullptr;
    int32_t* rb_rows = nullptr;
    uint32_t* rb_flags = nullptr;
    uint32_t* rb_nres = nullptr;
    // optional per-kernel timing with HIP events on the handle's stream
    bool prof = false;
    double prof_acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // scan launches, scan ms, scan bytes, scan flops, select ms, finalize ms, prep ms, steps
    std::vector<hipEvent_t> ev_pool;
    struct KStat {
        double launches = 0, ms = 0, bytes = 0, flops = 0;
        std::vector<float> l_ms, l_bytes, l_flops;  // per launch (the first 4096)
    };
    std::map<std::string, KStat> kstats;
    std::atomic<int64_t> ndist{0};
    std::mutex mu;
    // IVF (mode ANNMI_IVF_SQ8): lists are contiguous runs of the list-ordered
    // scan rows Xs; perm maps list position -> original row id.
    bool ivf = false;
    int nlist = 0;
    uint32_t* perm = nullptr;
    std::vector<int64_t> list_off;  // host, nlist + 1
    float* centroids = nullptr;     // [nlist x kdim32]
    struct Handle* cent = nullptr;  // exact index over the centroids (probe step)
    int64_t* d_probe = nullptr;     // [ws_B x probe_cap] probed list ids
    float* d_probe_d = nullptr;
    int probe_cap = 0;
    // device plan of an IVF step (ivf.h IvfPlanArgs); sized for ws_B queries
    int64_t* d_list_off = nullptr;  // [nlist + 1] (index array, shared by clones)
    int64_t max_list_len = 0;
    uint32_t* d_qcount = nullptr;   // [nlist]
    uint32_t* d_qoff = nullptr;     // [nlist]
    uint32_t* d_qfill = nullptr;    // [nlist]
    uint8_t* d_every = nullptr;     // [plan_B]
    uint32_t* d_gq = nullptr;       // [plan_B x nlist]
    uint32_t* d_plan = nullptr;     // [kIvfMaxPass x (nlist + 1)]
    uint32_t* d_nitems = nullptr;   // [kIvfMaxPass]
    uint32_t* d_next = nullptr;     // [kIvfMaxPass]
    uint32_t* d_coarse = nullptr;   // [kIvfMaxPass x kIvfCoarse]
    int plan_B = 0;
    // a clone shares every index array with its source and owns only its
    // stream and workspace
    bool is_clone = false;
    // linear_index-compatible query workspace
    float* li_P = nullptr;
    float* li_BT = nullptr;
    float* li_R = nullptr;
    float* li_bkey = nullptr;
    uint32_t* li_bid = nullptr;
    float* li_q = nullptr;
    int64_t* li_ids = nullptr;
    float* li_vals = nullptr;
    uint32_t* li_flag = nullptr;
    int li_k = 0;
    // Captured steps (flat modes): one executable graph per (B, k, rsel,
    // output buffers), captured the second time that shape is seen and
    // replayed from then on.  Workspace pointers are baked in, so the cache
    // is dropped whenever the workspace is reallocated.
    struct StepGraph {
        int B, k, rsel, nprobe;
        const void* ids;      // output buffers the graph writes (its finalize node's arguments)
        const void* dist;
        hipGraphExec_t exec;  // null: seen once, not captured yet
        bool failed;          // capture or instantiation failed: run eagerly
        uint64_t last_use;
        hipGraph_t graph;     // the captured graph (kept: its finalize node is retargeted)
        hipGraphNode_t fin;   // the finalize kernel node (null: none found, not retargetable)
        FinalArgs fa;         // its arguments as captured
    };
    std::vector<StepGraph> graphs;
    uint64_t graph_clock = 0;
    bool tuning = false;  // the autotuner times eager steps (a graph would fix one shape)
    // Set on a handle that has been cloned and on its clones: handles used
    // from several threads never capture steps (a capture on one thread and
    // an allocation, free or legacy-stream operation on another conflict).
    std::atomic<bool> shared{false};
    // Scan kernel shape per batch size: `fixed` (annmi_set_scan_shape) for
    // every batch, else the autotuned entry of the nearest power of two.
    struct ScanShape {
        int kernel = -1;  // -1 built-in rule, 0 streaming kernel, 1 wide kernel
        int cfg = 0;      // wide configuration (kWideCfg*)
        int64_t chunk = 0;
        int min_batch = 0;  // fixed shapes: wide from this batch size up (0: built-in)
    };
    ScanShape fixed;
    std::vector<ScanShape> tuned;  // index = log2 of the batch size
    std::string tune_report;
};

Handle::ScanShape shape_for(const Handle* h, int B) {
    if (!h->tuned.empty()) {
        int lg = 0;
        while (lg + 1 < (int)h->tuned.size() && (1 << lg) * 3 < 2 * B) ++lg;  // nearest power of two (geometric)
        return h->tuned[(size_t)lg];
    }
    return h->fixed;
}

void apply_shape(const Handle::ScanShape& sh, ScanArgs& a) {
    a.wide_cfg = sh.cfg;
    a.chunk_max = sh.chunk;
    if (sh.kernel == 0) a.wide_min_batch = 1 << 30;
    else if (sh.kernel == 1) a.wide_min_batch = 1;
    else a.wide_min_batch = sh.min_batch;
}

void drop_graphs(Handle* h) {
    for (auto& g : h->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    h->graphs.clear();
}

// The finalize kernel node of a captured step that writes the caller's
// output buffer `ids` (an IVF step also holds the probe's finalize, which
// writes the handle's own probe buffers), and its arguments.
bool find_final_node(hipGraph_t g, const int64_t* ids, hipGraphNode_t* node, FinalArgs* fa) {
    size_t nn = 0;
    if (hipGraphGetNodes(g, nullptr, &nn) != hipSuccess || nn == 0) return false;
    std::vector<hipGraphNode_t> nodes(nn);
    if (hipGraphGetNodes(g, nodes.data(), &nn) != hipSuccess) return false;
    for (size_t i = 0; i < nn; ++i) {
        hipGraphNodeType t;
        if (hipGraphNodeGetType(nodes[i], &t) != hipSuccess || t != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams kp{};
        if (hipGraphKernelNodeGetParams(nodes[i], &kp) != hipSuccess) continue;
        if (kp.func != finalize_kernel_fn() || !kp.kernelParams || !kp.kernelParams[0]) continue;
        FinalArgs f;
        std::memcpy(&f, kp.kernelParams[0], sizeof(FinalArgs));
        if (f.out_ids != ids) continue;
        *node = nodes[i];
        *fa = f;
        return true;
    }
    return false;
}

// Points a captured step at other output buffers: the finalize node's
// arguments in the executable graph are replaced (no other node writes
// them), so steps of one shape share one graph whatever buffers the caller
// passes.
bool graph_debug() {  // ANNMI_GRAPH_DEBUG=1: report why a step shape runs eagerly (stderr)
    static const bool on = getenv("ANNMI_GRAPH_DEBUG") != nullptr;
    return on;
}

bool retarget_graph(Handle::StepGraph& g, int64_t* ids, float* dist) {
    if (!g.fin) {
        if (graph_debug()) fprintf(stderr, "annmi graphs: B=%d no finalize node to retarget\n", g.B);
        return false;
    }
    hipKernelNodeParams kp{};
    hipError_t e = hipGraphKernelNodeGetParams(g.fin, &kp);
    if (e == hipSuccess) {
        FinalArgs fa = g.fa;
        fa.out_ids = ids;
        fa.out_d = dist;
        void* args[1] = {&fa};
        kp.kernelParams = args;
        kp.extra = nullptr;
        e = hipGraphExecKernelNodeSetParams(g.exec, g.fin, &kp);
    }
    if (e != hipSuccess) {
        if (graph_debug()) fprintf(stderr, "annmi graphs: B=%d retarget failed: %s\n", g.B, hipGetErrorString(e));
        return false;
    }
    g.ids = ids;
    g.dist = dist;
    return true;
}

void free_ws(Handle* h) {
    drop_graphs(h);
    dfree(h->dQ32);
    dfree(h->dQs);
    dfree(h->qn);
    dfree(h->qscale);
    dfree(h->slack);
    dfree(h->thr);
    dfree(h->cnt);
    dfree(h->ckey);
    dfree(h->crid);
    dfree(h->ovf);
    dfree(h->nrr);
    dfree(h->d_ids);
    dfree(h->d_dist);
    if (h->hpin) (void)hipHostFree(h->hpin);
    h->hpin = nullptr;
    h->dQ32 = nullptr;
    h->dQs = nullptr;
    h->qn = h->qscale = h->slack = h->thr = nullptr;
    h->cnt = h->crid = h->ovf = h->nrr = nullptr;
    h->ckey = nullptr;
    h->d_ids = nullptr;
    h->d_dist = nullptr;
    h->ws_B = 0;
}

void free_handle(Handle* h) {
    if (!h) return;
    if (h->st) (void)hipStreamSynchronize(h->st);
    free_ws(h);
    if (!h->is_clone) {
        if (h->own_x32) dfree(h->X32);
        dfree(h->xn32);
        if (h->own_xs) dfree(h->Xs);
        if (h->xn != h->xn32) dfree(h->xn);
        dfree(h->sq_scale);
        dfree(h->sq_offset);
        dfree(h->perm);
        dfree(h->centroids);
        dfree(h->d_list_off);
    }
    dfree(h->scratch_keys);
    dfree(h->rb_keys);
    dfree(h->rb_hist);
    dfree(h->rb_sel);
    dfree(h->rb_cnt);
    dfree(h->rb_cid);
    dfree(h->rb_rows);
    dfree(h->rb_flags);
    dfree(h->rb_nres);
    dfree(h->gf_cid);
    dfree(h->li_P);
    dfree(h->li_BT);
    dfree(h->li_R);
    dfree(h->li_bkey);
    dfree(h->li_bid);
    dfree(h->li_q);
    dfree(h->li_ids);
    dfree(h->li_vals);
    dfree(h->li_flag);
    dfree(h->d_probe);
    dfree(h->d_probe_d);
    dfree(h->d_qcount);
    dfree(h->d_qoff);
    dfree(h->d_qfill);
    dfree(h->d_every);
    dfree(h->d_gq);
    dfree(h->d_plan);
    dfree(h->d_nitems);
    dfree(h->d_next);
    dfree(h->d_coarse);
    if (h->cent) free_handle(h->cent);
    for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
    if (h->st && h->own_stream) (void)hipStreamDestroy(h->st);
    delete h;
}

bool ensure_ws(Handle* h, int B, int k) {
    if (h->ws_B >= B) return true;
    free_ws(h);
    const int64_t Bq = B;
    HIP_OR(hipMalloc(&h->dQ32, Bq * h->d * sizeof(float)), false);
    HIP_OR(hipMalloc(&h->dQs, Bq * scan_q_bytes(h->kind, h->kdim)), false);
    HIP_OR(hipMalloc(&h->qn, Bq * sizeof(float)), false);
    HIP_OR(hipMalloc(&h->qscale, Bq * sizeof(float)), false);
    HIP_OR(hipMalloc(&h->slack, Bq * sizeof(float)), false);
    HIP_OR(hipMalloc(&h->thr, Bq * sizeof(float)), false);
    HIP_OR(hipMalloc(&h->cnt, Bq * sizeof(uint32_t)), false);
    HIP_OR(hipMalloc(&h->ckey, Bq * h->cap * sizeof(float)), false);
    HIP_OR(hipMalloc(&h->crid, Bq * h->cap * sizeof(uint32_t)), false);
    HIP_OR(hipMalloc(&h->ovf, Bq * sizeof(uint32_t)), false);
    HIP_OR(hipMalloc(&h->nrr, Bq * sizeof(uint32_t)), false);
    HIP_OR(hipMalloc(&h->d_ids, Bq * 2048 * sizeof(int64_t)), false);
    HIP_OR(hipMalloc(&h->d_dist, Bq * 2048 * sizeof(float)), false);
    // the per-step read-back lands in pinned memory: an asynchronous copy,
    // not a staged one that waits on the stream by itself
    HIP_OR(hipHostMalloc(reinterpret_cast<void**>(&h->hpin), 3 * Bq * sizeof(uint32_t), hipHostMallocDefault), false);
    h->ws_B = B;
    (void)k;
    return true;
}

// Geometric row ranges [b_j, b_{j+1}): the first holds every row of the
// first S0 rows, each later one grows by rho, so each adds about
// rsel * (rho - 1) candidates on exchangeable data.  (A longer first range
// saves launches at small batches but costs more in the selections over its
// candidates than it saves: measured at one query on 1M and 10M rows.)
// Writes the bounds (at most `maxb`: the factor is >= 4, so 2^32 rows take
// fewer than 20) into b and returns how many; no allocation.
int range_bounds(int64_t n, int rsel, int cap, int B, int64_t* b, int maxb) {
    // Large batches append rsel * (rho - 1) candidates per query per range
    // (each an LDS and a global atomic), small ones pay a launch and a
    // selection per range: large batches take a short first range and a
    // small growth factor, small batches few ranges.  ANNMI_RANGE_S0 /
    // ANNMI_RANGE_RHO / ANNMI_RANGE_BATCH override the first range, the
    // factor and the batch size from which the large-batch rule applies
    // (tuning).
    static const int64_t env_s0 = getenv("ANNMI_RANGE_S0") ? atoll(getenv("ANNMI_RANGE_S0")) : 0;
    static const int64_t env_rho = getenv("ANNMI_RANGE_RHO") ? atoll(getenv("ANNMI_RANGE_RHO")) : 0;
    static const int64_t env_b = getenv("ANNMI_RANGE_BATCH") ? atoll(getenv("ANNMI_RANGE_BATCH")) : 0;
    // (10M x 384 SQ8 a step: 16 queries 0.948 -> 0.844 ms and 256 queries
    // 1.497 -> 1.382 ms with the small-batch rule; 512 queries 2.483 ->
    // 2.331 ms with the large-batch one)
    const bool large = B >= (env_b > 0 ? env_b : 384);
    int nb = 0;
    b[nb++] = 0;
    int64_t s0 = env_s0 > 0 ? env_s0 : (large ? std::max<int64_t>(1024, 8 * (int64_t)rsel)
                                              : std::max<int64_t>(4096, 8 * (int64_t)rsel));
    s0 = (s0 + 15) / 16 * 16;  // ranges start on 16-row tiles of the tiled layout
    s0 = std::min<int64_t>(s0, cap / 2 / 16 * 16);
    s0 = std::max<int64_t>(s0, (rsel + 15) / 16 * 16);
    int64_t rho = std::max<int64_t>(4, std::min<int64_t>(32, cap / (8 * (int64_t)rsel)));
    if (large) rho = std::min<int64_t>(rho, 4);
    if (env_rho >= 2) rho = env_rho;
    int64_t cur = std::min<int64_t>(n, s0);
    b[nb++] = cur;
    while (cur < n && nb < maxb) {
        cur = nb + 1 < maxb ? std::min<int64_t>(n, cur * rho) : n;
        b[nb++] = cur;
    }
    return nb;
}

// Event bookkeeping shared by the flat and IVF steps.
struct Profiler {
    Handle* h;
    int& ev_used;
    std::vector<std::pair<int, int>>* scan_ev;
    std::vector<std::pair<int, int>>* sel_ev;
    std::vector<double>* scan_rows;
    std::vector<std::string>* scan_names;
    int ev_prep0, ev_prep1;
    std::vector<double> bytes_override, flops_override;  // per scan launch, when set
    int take() {
        if (ev_used == (int)h->ev_pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return -1;
            h->ev_pool.push_back(e);
        }
        return ev_used++;
    }
    // an event recorded on the stream now (spans of several operations)
    int ev() {
        if (!h->prof) return -1;
        const int i = take();
        if (i >= 0) (void)hipEventRecord(h->ev_pool[i], h->st);
        return i;
    }
    // times the one kernel the next ANNMI_LAUNCH issues: the dispatch
    // records the pair itself (no event packets between the kernels)
    void arm(int& e0, int& e1) {
        e0 = e1 = -1;
        if (!h->prof) return;
        e0 = take();
        e1 = take();
        if (e0 < 0 || e1 < 0) {
            e0 = e1 = -1;
            return;
        }
        t_launch_ev[0] = h->ev_pool[e0];
        t_launch_ev[1] = h->ev_pool[e1];
    }
    // after the launch call: a pair no launch consumed (nothing launched)
    // times nothing
    void disarm(int& e0, int& e1) {
        if (t_launch_ev[0]) e0 = e1 = -1;
        t_launch_ev[0] = t_launch_ev[1] = nullptr;
    }
};

// finalize launch (enqueue only), common to both step kinds.
bool launch_final(Handle* h, int B, int k, int rsel, int64_t* ids_out, float* d_out, const CandBuf& cb, Profiler& pr,
                  int& ev_f0, int& ev_f1) {
    hipStream_t st = h->st;
    FinalArgs fa{};
    fa.cand = cb;
    fa.B = B;
    fa.k = k;
    fa.rsel = rsel;
    fa.slack = h->slack;
    fa.X32 = h->X32;
    fa.x32_stride = h->kdim32;
    fa.d = h->d;
    fa.Q32 = h->dQ32;
    fa.out_ids = ids_out;
    fa.out_d = d_out;
    fa.overflow = h->ovf;
    fa.n_reranked = h->nrr;
    // fp32 re-score error bound (relative, both neighbours), doubled
    const double u = 5.9604644775390625e-8;
    fa.eps_rel = 4.0 * ((h->d + 15) / 16 + 28) * u;
    pr.arm(ev_f0, ev_f1);
    const hipError_t le = launch_finalize(fa, st);
    pr.disarm(ev_f0, ev_f1);
    HIP_OR(le, false);
    return true;
}

constexpr int kRobustCap = 65536;  // rows a robust query ranks within the fp32 error window of rank k

// Workspace of the all-rows selections (robust.hip) for up to rb_F queries
// at a time: the key of every row is kept per query, at most 512 MB of it.
bool ensure_robust(Handle* h) {
    if (h->rb_F) return true;
    const int64_t n = h->n;
    const int64_t fit = std::max<int64_t>(1, ((int64_t)512 << 20) / (n * 4));
    const int F = (int)std::min<int64_t>(64, fit);
    HIP_OR(hipMalloc(&h->rb_keys, (size_t)F * n * 4), false);
    HIP_OR(hipMalloc(&h->rb_hist, (size_t)F * 2048 * 4), false);
    HIP_OR(hipMalloc(&h->rb_sel, (size_t)F * 4 * 8), false);
    HIP_OR(hipMalloc(&h->rb_cnt, (size_t)F * 4), false);
    HIP_OR(hipMalloc(&h->rb_cid, (size_t)F * kRobustCap * 4), false);
    HIP_OR(hipMalloc(&h->rb_rows, (size_t)F * 4), false);
    HIP_OR(hipMalloc(&h->rb_flags, (size_t)F * 4), false);
    HIP_OR(hipMalloc(&h->rb_nres, (size_t)F * 4), false);
    h->rb_F = F;
    return true;
}

RobustArgs robust_args(Handle* h, int mode, int F, int k, int64_t* ids_out, float* d_out) {
    RobustArgs ra{};
    ra.mode = mode;
    ra.X32 = h->X32;
    ra.x32_stride = h->kdim32;
    ra.n = h->n;
    ra.d = h->d;
    ra.Q32 = h->dQ32;
    ra.rows = h->rb_rows;
    ra.F = F;
    ra.k = k;
    ra.r = k;
    ra.keys = h->rb_keys;
    ra.hist = h->rb_hist;
    ra.sel = h->rb_sel;
    ra.cnt = h->rb_cnt;
    ra.cid = h->rb_cid;
    ra.cap = kRobustCap;
    const double u = 5.9604644775390625e-8;
    ra.eps_rel = 4.0 * ((h->d + 15) / 16 + 28) * u;  // finalize's near-tie window
    ra.out_ids = ids_out;
    ra.out_d = d_out;
    ra.flags = h->rb_flags;
    ra.n_rescored = h->rb_nres;
    return ra;
}

// Re-runs queries `qs` of the step just completed (queries in h->dQ32,
// outputs rows of ids_out / d_out) with the exact all-rows path of
// robust.hip, a bounded number at a time.  rflags / rnres receive, per
// query, bit 0 when more than kRobustCap rows fell in the window and the
// rows collected.
bool run_robust(Handle* h, const std::vector<int32_t>& qs, int k, int64_t* ids_out, float* d_out,
                std::vector<uint32_t>& rflags, std::vector<uint32_t>& rnres) {
    hipStream_t st = h->st;
    if (!ensure_robust(h)) return false;
    rflags.assign(qs.size(), 0u);
    rnres.assign(qs.size(), 0u);
    for (size_t f0 = 0; f0 < qs.size(); f0 += (size_t)h->rb_F) {
        const int F = (int)std::min<size_t>((size_t)h->rb_F, qs.size() - f0);
        HIP_OR(hipMemcpyAsync(h->rb_rows, qs.data() + f0, (size_t)F * 4, hipMemcpyHostToDevice, st), false);
        const RobustArgs ra = robust_args(h, ROBUST_EXACT, F, k, ids_out, d_out);
        HIP_OR(launch_robust(ra, st), false);
        HIP_OR(hipMemcpyAsync(rflags.data() + f0, h->rb_flags, (size_t)F * 4, hipMemcpyDeviceToHost, st), false);
        HIP_OR(hipMemcpyAsync(rnres.data() + f0, h->rb_nres, (size_t)F * 4, hipMemcpyDeviceToHost, st), false);
        HIP_OR(hipStreamSynchronize(st), false);
    }
    return true;
}

// flag read-back, host synchronisation, event timing and accounting.
// Queries the step flagged (candidate overflow, more than 2048 rows in the
// near-tie window, an fp64 run longer than 32) are answered again by the
// exact all-rows path (flag bit 3 marks them in annmi_last_batch_stats).
bool complete_step(Handle* h, int B, int k, int64_t* ids_out, float* d_out, bool exact, Profiler& pr, int ev_f0,
                   int ev_f1, const uint32_t* probe_flags) {
    hipStream_t st = h->st;
    uint32_t* hf = h->hpin;
    HIP_OR(hipMemcpyAsync(hf, h->ovf, B * sizeof(uint32_t), hipMemcpyDeviceToHost, st), false);
    HIP_OR(hipMemcpyAsync(hf + B, h->nrr, B * sizeof(uint32_t), hipMemcpyDeviceToHost, st), false);
    if (probe_flags)  // the IVF probe's own flags: a query whose probe could not be bounded
        HIP_OR(hipMemcpyAsync(hf + 2 * B, probe_flags, B * sizeof(uint32_t), hipMemcpyDeviceToHost, st), false);
    HIP_OR(hipStreamSynchronize(st), false);
    std::vector<uint32_t> flags(hf, hf + B), nrr(hf + B, hf + 2 * B);
    if (probe_flags)
        for (int i = 0; i < B; ++i) flags[i] |= hf[2 * B + i] & 7u;
    if (h->prof && ev_f1 >= 0) {
        auto& scan_ev = *pr.scan_ev;
        auto& sel_ev = *pr.sel_ev;
        auto& scan_rows = *pr.scan_rows;
        auto& scan_names = *pr.scan_names;
        const int ev_prep0 = pr.ev_prep0, ev_prep1 = pr.ev_prep1;
        auto ms = [&](int a0, int a1) -> double {
            float t = 0.f;
            if (a0 < 0 || a1 < 0 || hipEventElapsedTime(&t, h->ev_pool[a0], h->ev_pool[a1]) != hipSuccess) return 0.0;
            return (double)t;
        };
        const double row_bytes = (double)h->xs_stride;
        for (size_t j = 0; j < scan_ev.size(); ++j) {
            const double t = ms(scan_ev[j].first, scan_ev[j].second);
            const double by = j < pr.bytes_override.size()
                                  ? pr.bytes_override[j]
                                  : scan_rows[j] * row_bytes + (double)B * scan_q_bytes(h->kind, h->kdim);
            const double fl = j < pr.flops_override.size() ? pr.flops_override[j]
                                                           : 2.0 * (double)B * scan_rows[j] * (double)h->d;
            h->prof_acc[0] += 1;
            h->prof_acc[1] += t;
            h->prof_acc[2] += by;
            h->prof_acc[3] += fl;
            Handle::KStat& ks = h->kstats[scan_names[j]];
            ks.launches += 1;
            ks.ms += t;
            ks.bytes += by;
            ks.flops += fl;
            if (ks.l_ms.size() < 4096) {
                ks.l_ms.push_back((float)t);
                ks.l_bytes.push_back((float)by);
                ks.l_flops.push_back((float)fl);
            }
        }
        for (auto& e : sel_ev) h->prof_acc[4] += ms(e.first, e.second);
        h->prof_acc[5] += ms(ev_f0, ev_f1);
        h->prof_acc[6] += ms(ev_prep0, ev_prep1);
        h->prof_acc[7] += 1;
    }
    std::vector<int32_t> redo;
    for (int i = 0; i < B; ++i)
        if (flags[i] & 7u) redo.push_back(i);
    std::vector<bool> full(B, false);
    if (!redo.empty()) {
        std::vector<uint32_t> rflags, rnres;
        if (!run_robust(h, redo, k, ids_out, d_out, rflags, rnres)) return false;
        for (size_t j = 0; j < redo.size(); ++j) {
            const int i = redo[j];
            if (rflags[j] & 1u) {
                set_err("query %d: more than %d rows lie within the fp32 error bound of the rank-%d distance", i,
                        kRobustCap, k);
                return false;
            }
            flags[i] |= 8u;
            nrr[i] = rnres[j];
            full[i] = true;
        }
    }
    int64_t nd = 0;
    for (int i = 0; i < B; ++i) nd += (exact || full[i]) ? h->n : (int64_t)nrr[i];
    h->ndist += nd;
    h->last_nrr.insert(h->last_nrr.end(), nrr.begin(), nrr.end());
    h->last_flags.insert(h->last_flags.end(), flags.begin(), flags.end());
    return true;
}



bool ensure_ivf_ws(Handle* h, int nprobe);
bool enqueue_ivf(Handle* h, int B, int k, int rsel, int nprobe, int64_t* ids_out, float* d_out, Profiler& pr,
                 std::vector<std::pair<int, int>>& scan_ev, std::vector<std::pair<int, int>>& sel_ev,
                 std::vector<double>& scan_rows, std::vector<std::string>& scan_names, int& ev_f0, int& ev_f1);
int32_t batch_impl(Handle* h, const float* Q, int32_t B, int32_t k, int32_t R, int64_t* ids, float* dists,
                   bool device, int32_t nprobe);
Handle* create_impl(const float* base, int64_t n, int32_t d, const annmi_params* params, int64_t src_stride,
                    const uint8_t* saved = nullptr);

// Enqueues one flat step (prep, geometric range scans with their filter
// selections, finalize) on the handle's stream; no host synchronisation, so
// it can be captured into a graph.
bool enqueue_flat(Handle* h, int B, int k, int rsel, int64_t* ids_out, float* d_out, Profiler& pr,
                  std::vector<std::pair<int, int>>& scan_ev, std::vector<std::pair<int, int>>& sel_ev,
                  std::vector<double>& scan_rows, std::vector<std::string>& scan_names, int& ev_f0, int& ev_f1) {
    const bool exact = h->kind == SCAN_F32;
    hipStream_t st = h->st;
    pr.arm(pr.ev_prep0, pr.ev_prep1);
    const hipError_t pe = launch_prep_queries(h->kind, h->dQ32, B, h->d, h->kdim, h->dQs, h->qn, h->qscale, h->slack,
                                              h->maxnorm, h->alpha_x, h->sq_scale, h->sq_offset, st, h->cnt, h->ovf,
                                              h->thr);
    pr.disarm(pr.ev_prep0, pr.ev_prep1);
    HIP_OR(pe, false);
    CandBuf cb{h->cnt, h->ckey, h->crid, h->cap};
    // no host allocation here unless the step is timed (the allocation-free
    // entry point runs this with timing off)
    int64_t bounds[64];
    const int nb = range_bounds(h->n, rsel, h->cap, B, bounds, 64);
    for (int j = 0; j + 1 < nb; ++j) {
        if (j > 0) {
            SelectArgs sa{cb, B, rsel, h->slack, h->thr, h->ovf};
            int e0, e1;
            pr.arm(e0, e1);
            const hipError_t le = launch_select_threshold(sa, st);
            pr.disarm(e0, e1);
            HIP_OR(le, false);
            if (h->prof) sel_ev.push_back({e0, e1});
        }
        ScanArgs a{};
        a.kind = h->kind;
        a.X = h->Xs;
        a.xstride = h->xs_stride;
        a.xn = h->xn;
        a.row_begin = bounds[j];
        a.row_end = bounds[j + 1];
        a.d = h->d;
        a.kdim = h->kdim;
        a.Q = h->dQs;
        a.qn = h->qn;
        a.qscale = exact ? nullptr : h->qscale;
        a.thr = h->thr;
        a.B = B;
        a.cand = cb;
        a.xn_lo = h->xn_lo;
        a.xn_amax = h->xn_amax;
        a.prefilter = h->prefilter;
        a.tiled = h->tiled;
        apply_shape(shape_for(h, B), a);
        int e0, e1;
        const char* vname = nullptr;
        pr.arm(e0, e1);
        const hipError_t le = launch_scan(a, st, &vname);
        pr.disarm(e0, e1);
        HIP_OR(le, false);
        if (h->prof) {
            scan_ev.push_back({e0, e1});
            scan_rows.push_back((double)(a.row_end - a.row_begin));
            scan_names.push_back(vname ? vname : "");
        }
    }
    return launch_final(h, B, k, rsel, ids_out, d_out, cb, pr, ev_f0, ev_f1);
}

bool graphs_enabled() {
    static const bool on = [] {  // ANNMI_GRAPHS=0: always launch eagerly (A/B timing)
        const char* e = getenv("ANNMI_GRAPHS");
        return !(e && e[0] == '0');
    }();
    return on;
}

constexpr size_t kMaxGraphs = 32;

// One step over <= ws_B queries already in h->dQ32; results in d_ids/d_dist.
// Steps of a shape seen before replay a captured graph (no per-kernel launch
// cost); the first step of a shape runs eagerly, the second captures.  IVF
// steps are planned on the device and captured the same way.
bool run_step(Handle* h, int B, int k, int R, int nprobe, int64_t* ids_out, float* d_out) {
    const bool exact = h->kind == SCAN_F32 && !h->ivf;
    const int kk = (int)std::min<int64_t>(k, h->n);
    int rsel = exact ? kk : std::max(kk, R);
    rsel = (int)std::min<int64_t>(rsel, h->n);
    if (!exact && rsel > 2048) {
        set_err("rerank R=%d exceeds the 2048-row re-score limit", rsel);
        return false;
    }
    if (h->ivf) {
        if (nprobe <= 0) nprobe = h->p.nprobe > 0 ? h->p.nprobe : 8;
        nprobe = std::min(nprobe, h->nlist);
        if (nprobe > 2048) {
            set_err("nprobe=%d exceeds 2048", nprobe);
            return false;
        }
        if (!ensure_ivf_ws(h, nprobe)) return false;
    } else {
        nprobe = 0;
    }
    hipStream_t st = h->st;
    std::vector<std::pair<int, int>> scan_ev, sel_ev;  // (start, end) event indices
    std::vector<double> scan_rows;
    std::vector<std::string> scan_names;
    int ev_used = 0;
    Profiler pr{h, ev_used, &scan_ev, &sel_ev, &scan_rows, &scan_names, -1, -1};
    int ev_f0 = -1, ev_f1 = -1;
    auto enqueue = [&]() {
        return h->ivf ? enqueue_ivf(h, B, k, rsel, nprobe, ids_out, d_out, pr, scan_ev, sel_ev, scan_rows, scan_names,
                                    ev_f0, ev_f1)
                      : enqueue_flat(h, B, k, rsel, ids_out, d_out, pr, scan_ev, sel_ev, scan_rows, scan_names, ev_f0,
                                     ev_f1);
    };
    const uint32_t* probe_flags = h->ivf ? h->cent->ovf : nullptr;
    Handle::StepGraph* sg = nullptr;
    if (graphs_enabled() && !h->prof && !h->tuning && !h->shared.load()) {
        // one graph per step shape: a graph captured for other output
        // buffers is retargeted before it is replayed
        for (auto& g : h->graphs)
            if (g.B == B && g.k == k && g.rsel == rsel && g.nprobe == nprobe) sg = &g;
        if (!sg) {
            if (h->graphs.size() >= kMaxGraphs) {  // evict the least recently used
                size_t lru = 0;
                for (size_t i = 1; i < h->graphs.size(); ++i)
                    if (h->graphs[i].last_use < h->graphs[lru].last_use) lru = i;
                if (h->graphs[lru].exec) (void)hipGraphExecDestroy(h->graphs[lru].exec);
                if (h->graphs[lru].graph) (void)hipGraphDestroy(h->graphs[lru].graph);
                h->graphs.erase(h->graphs.begin() + (long)lru);
            }
            h->graphs.push_back(Handle::StepGraph{B, k, rsel, nprobe, ids_out, d_out, nullptr, false,
                                                  ++h->graph_clock, nullptr, nullptr, FinalArgs{}});
        } else {
            sg->last_use = ++h->graph_clock;
            if (!sg->exec && !sg->failed) {
                // second sighting of this shape: capture the step, then replay it
                HIP_OR(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), false);
                const bool ok = enqueue();
                hipGraph_t graph = nullptr;
                const hipError_t ec = hipStreamEndCapture(st, &graph);
                if (!ok) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return false;
                }
                hipGraphExec_t exec = nullptr;
                const hipError_t ei = (ec == hipSuccess && graph)
                                          ? hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0)
                                          : hipErrorUnknown;
                if (graph_debug() && ei != hipSuccess)
                    fprintf(stderr, "annmi graphs: B=%d capture %s, instantiate %s\n", B, hipGetErrorString(ec),
                            hipGetErrorString(ei));
                if (ei == hipSuccess) {
                    sg->exec = exec;
                    sg->graph = graph;
                    sg->ids = ids_out;
                    sg->dist = d_out;
                    if (!find_final_node(graph, ids_out, &sg->fin, &sg->fa)) {
                        sg->fin = nullptr;
                        if (graph_debug()) fprintf(stderr, "annmi graphs: B=%d finalize node not found\n", B);
                    }
                } else {
                    sg->failed = true;
                    if (graph) (void)hipGraphDestroy(graph);
                }
                (void)hipGetLastError();
                scan_ev.clear();
                sel_ev.clear();
                scan_rows.clear();
                scan_names.clear();
            }
            if (sg->exec && (sg->ids != ids_out || sg->dist != d_out) && !retarget_graph(*sg, ids_out, d_out)) {
                // not retargetable: this shape runs eagerly from now on
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(sg->exec);
                if (sg->graph) (void)hipGraphDestroy(sg->graph);
                sg->exec = nullptr;
                sg->graph = nullptr;
                sg->failed = true;
            }
            if (sg->exec) {
                HIP_OR(hipGraphLaunch(sg->exec, st), false);
                return complete_step(h, B, k, ids_out, d_out, exact, pr, -1, -1, probe_flags);
            }
        }
    }
    if (!enqueue()) return false;
    return complete_step(h, B, k, ids_out, d_out, exact, pr, ev_f0, ev_f1, probe_flags);
}

bool ivf_build(Handle* h);

// Bounds of the scan's row norm term over all rows.
bool xn_bounds(Handle* h) {
    uint32_t* mm = nullptr;
    HIP_OR(hipMalloc(&mm, 2 * sizeof(uint32_t)), false);
    bool ok = launch_fill_u32(mm, 0xFFFFFFFFu, 1, h->st) == hipSuccess &&
              launch_fill_u32(mm + 1, 0u, 1, h->st) == hipSuccess &&
              launch_dim_minmax(h->xn, 1, h->n, 1, mm, mm + 1, h->st) == hipSuccess;
    uint32_t hm[2] = {0, 0};
    ok = ok && hipMemcpyAsync(hm, mm, sizeof(hm), hipMemcpyDeviceToHost, h->st) == hipSuccess &&
         hipStreamSynchronize(h->st) == hipSuccess;
    dfree(mm);
    if (!ok) {
        set_err("row norm bounds failed: %s", hipGetErrorString(hipGetLastError()));
        return false;
    }
    const float lo = key2f(hm[0]), hi = key2f(hm[1]);
    if (lo == lo && hi == hi) {  // NaN norms leave the prefilter off
        h->xn_lo = lo;
        h->xn_amax = std::max(std::fabs(lo), std::fabs(hi));
        // equal row norms to 0.5%: one bound per query on the raw dot is
        // about as tight as the per-row bound and cheaper.  Wider spreads let
        // rows just past a query's threshold through the global bound
        // (measured: fp8 on normalised data, spread ~5%, 11% slower).
        // Integer dots (SQ8, SQ4) fold each row's norm excess over xn_lo into
        // the accumulator's start value (mode 3): the cost of the global
        // bound with the tightness of the per-row one.
        h->prefilter = (hi - lo) <= 0.005f * h->xn_amax ? 2 : 1;
        if (h->kind == SCAN_SQ8 || h->kind == SCAN_SQ4) h->prefilter = 3;
        if (const char* e = getenv("ANNMI_PREFILTER")) {  // 1, 2 or 3: force a mode (tuning, tests)
            const int v = atoi(e);
            if (v >= 1 && v <= 3) h->prefilter = v;
        }
    }
    return true;
}

// A handle's centroid index runs on its parent's stream (its steps are part
// of the parent's IVF step and are captured with it).
void share_stream(Handle* child, Handle* parent) {
    if (child->st && child->own_stream) (void)hipStreamDestroy(child->st);
    child->st = parent->st;
    child->own_stream = false;
}

// IVF index arrays beside the lists: the exact index over the centroids (the
// probe step, lsh_index_optimized.cpp:325-333) and the list offsets on the
// device for the step plan.
bool ivf_attach(Handle* h) {
    annmi_params cp;
    annmi_default_params(&cp, ANNMI_EXACT_F32);
    cp.base_on_device = 1;
    cp.max_batch = h->p.max_batch;
    h->cent = create_impl(h->centroids, h->nlist, h->d, &cp, h->kdim32);
    if (!h->cent) return false;
    share_stream(h->cent, h);
    h->max_list_len = 0;
    for (int l = 0; l < h->nlist; ++l)
        h->max_list_len = std::max<int64_t>(h->max_list_len, h->list_off[(size_t)l + 1] - h->list_off[(size_t)l]);
    HIP_OR(hipMalloc(&h->d_list_off, h->list_off.size() * 8), false);
    HIP_OR(hipMemcpyAsync(h->d_list_off, h->list_off.data(), h->list_off.size() * 8, hipMemcpyHostToDevice, h->st),
           false);
    HIP_OR(hipStreamSynchronize(h->st), false);
    return true;
}

bool setup_mode(Handle* h) {
    hipStream_t st = h->st;
    const int64_t n = h->n;
    const int d = h->d;
    if (h->ivf && !ivf_build(h)) return false;
    if (h->kind == SCAN_F32) {
        h->kdim = h->kdim32;
        h->Xs = h->X32;
        h->xs_stride = (int64_t)h->kdim * 4;
        h->xn = h->xn32;
        return true;
    }
    h->kdim = round_up(d, chunk_of(h->kind));
    h->xs_stride = scan_row_bytes(h->kind, h->kdim);
    h->tiled = h->ivf ? 0 : 1;  // IVF lists start at any row: row-major
    const int64_t n_alloc = h->tiled ? (n + 15) / 16 * 16 : n;
    HIP_OR(hipMalloc(&h->Xs, n_alloc * h->xs_stride), false);
    h->own_xs = true;
    HIP_OR(hipMalloc(&h->xn, n * sizeof(float)), false);
    h->bytes += n_alloc * h->xs_stride + n * 4;
    float* sc = nullptr;
    float* of = nullptr;
    if (h->kind == SCAN_FP8 || h->kind == SCAN_SQ8 || h->kind == SCAN_SQ4) {
        uint32_t *mn, *mx;
        HIP_OR(hipMalloc(&mn, d * sizeof(uint32_t)), false);
        HIP_OR(hipMalloc(&mx, d * sizeof(uint32_t)), false);
        HIP_OR(launch_fill_u32(mn, 0xFFFFFFFFu, d, st), false);
        HIP_OR(launch_fill_u32(mx, 0u, d, st), false);
        HIP_OR(launch_dim_minmax(h->X32, h->kdim32, n, d, mn, mx, st), false);
        std::vector<uint32_t> hmn(d), hmx(d);
        HIP_OR(hipMemcpyAsync(hmn.data(), mn, d * 4, hipMemcpyDeviceToHost, st), false);
        HIP_OR(hipMemcpyAsync(hmx.data(), mx, d * 4, hipMemcpyDeviceToHost, st), false);
        HIP_OR(hipStreamSynchronize(st), false);
        dfree(mn);
        dfree(mx);
        std::vector<float> fmn(d), fmx(d);
        for (int j = 0; j < d; ++j) {
            fmn[j] = key2f(hmn[j]);
            fmx[j] = key2f(hmx[j]);
        }
        if (h->kind == SCAN_FP8) {
            float m = 0.f;
            for (int j = 0; j < d; ++j) m = std::max(m, std::max(std::fabs(fmn[j]), std::fabs(fmx[j])));
            // power of two putting the largest |x| in [128, 256): e4m3 max is 448
            float a = m > 0.f ? std::exp2(std::floor(std::log2(255.0f / m))) : 1.f;
            while (m * a >= 256.f) a *= 0.5f;
            h->alpha_x = a;
        } else {
            // scale/offset per the reference encoders (float arithmetic as there)
            const int kd = h->kdim;
            std::vector<float> hs(kd, 0.f), ho(kd, 0.f);
            const float levels = h->kind == SCAN_SQ8 ? 255.0f : 15.0f;
            for (int j = 0; j < d; ++j) {
                if (fmx[j] - fmn[j] > 1e-8f) {
                    hs[j] = (fmx[j] - fmn[j]) / levels;
                    ho[j] = fmn[j];
                } else {
                    hs[j] = 1.0f;
                    // SQ4 (submission/src/hnsw.cpp:166-168) and the gpu_filter SQ8
                    // encoder keep offset = min; the ivf/hnsw SQ8 encoder uses 0.
                    ho[j] = (h->kind == SCAN_SQ8 && h->p.sq_variant == 0) ? 0.0f : fmn[j];
                }
            }
            HIP_OR(hipMalloc(&h->sq_scale, kd * sizeof(float)), false);
            HIP_OR(hipMalloc(&h->sq_offset, kd * sizeof(float)), false);
            HIP_OR(hipMemcpyAsync(h->sq_scale, hs.data(), kd * 4, hipMemcpyHostToDevice, st), false);
            HIP_OR(hipMemcpyAsync(h->sq_offset, ho.data(), kd * 4, hipMemcpyHostToDevice, st), false);
            HIP_OR(hipStreamSynchronize(st), false);
            sc = h->sq_scale;
            of = h->sq_offset;
        }
    }
    HIP_OR(launch_encode(h->kind, h->X32, h->kdim32, n, d, h->kdim, h->Xs, h->xn, nullptr, h->alpha_x, sc, of,
                         h->p.sq_variant, st, h->ivf ? h->perm : nullptr, h->tiled ? tile_cb_of(h->kind) : 0),
           false);
    HIP_OR(hipStreamSynchronize(st), false);
    if (h->ivf && !ivf_attach(h)) return false;
    return true;
}

// Saved index (annmi_save / annmi_load): a header, then 256-byte aligned
// sections of the derived device arrays, so that loading skips the IVF
// k-means build and the encoders.  The fp32 rows are not saved: they are
// the base passed to annmi_load (the re-score and exact-scan rows).
struct SavedHeader {
    char magic[8];  // "ANNMIIDX"
    int32_t version;
    int32_t mode;
    int64_t n;
    int32_t d, kdim;
    int64_t row_stride;
    int32_t sq_variant, nlist;
    float alpha_x;
    int32_t reserved;
    // byte offsets: 0 scale, 1 offset, 2 scan rows, 3 row norms, 4 perm,
    // 5 list offsets, 6 centroids, 7 end (a section is empty when its offset
    // equals the next one)
    int64_t off[8];
};
constexpr int32_t kSavedVersion = 1;

bool load_mode(Handle* h, const uint8_t* blob);

Handle* create_impl(const float* base, int64_t n, int32_t d, const annmi_params* params, int64_t src_stride,
                    const uint8_t* saved) {
    g_err[0] = 0;
    if (!base || n <= 0 || d <= 0) {
        set_err("empty index: n=%lld d=%d", (long long)n, d);
        return nullptr;
    }
    if (d > kMaxDim) {
        set_err("dim %d > %d unsupported", d, kMaxDim);
        return nullptr;
    }
    if (n > 0xFFFFFFF0ll) {
        set_err("n=%lld exceeds the 32-bit row id range", (long long)n);
        return nullptr;
    }
    int ndev = 0;
    hipError_t derr = hipGetDeviceCount(&ndev);
    if (derr != hipSuccess || ndev == 0) {
        set_err("no HIP device visible (%s)", derr == hipSuccess ? "0 devices" : hipGetErrorString(derr));
        return nullptr;
    }
    annmi_params p;
    if (params)
        p = *params;
    else
        annmi_default_params(&p, ANNMI_EXACT_F32);
    int kind = scan_kind_of(p.mode);
    if (kind < 0) {
        set_err("unknown mode %d", p.mode);
        return nullptr;
    }
    {
        // the scan stages at least one 16-query tile in LDS (at most the 160
        // KB of a CU); IVF also scans fp32 centroids (k-means, probe)
        const int kd = round_up(d, chunk_of(kind));
        int64_t lds = 16 * (scan_q_bytes(kind, kd) + 16);
        if (p.mode == ANNMI_IVF_SQ8)
            lds = std::max<int64_t>(lds, 16 * (scan_q_bytes(SCAN_F32, round_up(d, chunk_of(SCAN_F32))) + 16));
        if (lds > 160 * 1024) {
            set_err("dim %d too large for mode %d (query tile needs %lld bytes of LDS)", d, p.mode, (long long)lds);
            return nullptr;
        }
    }
    Handle* h = new Handle();
    h->p = p;
    (void)hipGetDevice(&h->dev);
    h->n = n;
    h->d = d;
    h->kind = kind;
    h->ivf = p.mode == ANNMI_IVF_SQ8;
    h->kdim32 = round_up(d, 16);
    if (p.max_batch <= 0) h->p.max_batch = 4096;
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
        set_err("hipStreamCreate failed");
        delete h;
        return nullptr;
    }
    auto fail = [&]() -> Handle* {
        std::string keep = g_err;
        free_handle(h);
        snprintf(g_err, sizeof(g_err), "%s", keep.c_str());
        return nullptr;
    };
    // fp32 rows (re-score base; the exact scan base)
    if (p.base_on_device && h->kdim32 == d && (src_stride == 0 || src_stride == d) &&
        (reinterpret_cast<uintptr_t>(base) & 15) == 0) {
        h->X32 = const_cast<float*>(base);
        h->own_x32 = false;
    } else {
        if (hipMalloc(&h->X32, n * (int64_t)h->kdim32 * sizeof(float)) != hipSuccess) {
            set_err("hipMalloc of %lld bytes for the fp32 rows failed", (long long)(n * (int64_t)h->kdim32 * 4));
            return fail();
        }
        h->own_x32 = true;
        h->bytes += n * (int64_t)h->kdim32 * 4;
        hipMemcpyKind kindc = p.base_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        const size_t spitch = (size_t)(src_stride > 0 ? src_stride : d) * 4;
        if (hipMemcpy2DAsync(h->X32, (size_t)h->kdim32 * 4, base, spitch, (size_t)d * 4, (size_t)n, kindc,
                             h->st) != hipSuccess) {
            set_err("copy of the base rows failed");
            return fail();
        }
        if (h->kdim32 != d) {
            // zero the padding columns
            if (hipMemset2DAsync(h->X32 + d, (size_t)h->kdim32 * 4, 0, (size_t)(h->kdim32 - d) * 4, (size_t)n, h->st) !=
                hipSuccess) {
                set_err("padding clear failed");
                return fail();
            }
        }
    }
    if (hipMalloc(&h->xn32, n * sizeof(float)) != hipSuccess) {
        set_err("hipMalloc norms failed");
        return fail();
    }
    h->bytes += n * 4;
    uint32_t* mk = nullptr;
    if (hipMalloc(&mk, sizeof(uint32_t)) != hipSuccess) return fail();
    uint32_t zero = 0;
    (void)hipMemcpyAsync(mk, &zero, 4, hipMemcpyHostToDevice, h->st);
    if (launch_encode(SCAN_F32, h->X32, h->kdim32, n, d, h->kdim32, h->X32, h->xn32, mk, 1.f, nullptr, nullptr, 0,
                      h->st) != hipSuccess) {
        set_err("norm kernel launch failed");
        dfree(mk);
        return fail();
    }
    uint32_t hk = 0;
    (void)hipMemcpyAsync(&hk, mk, 4, hipMemcpyDeviceToHost, h->st);
    if (hipStreamSynchronize(h->st) != hipSuccess) {
        set_err("index build failed: %s", hipGetErrorString(hipGetLastError()));
        dfree(mk);
        return fail();
    }
    dfree(mk);
    h->maxnorm = key2f(hk);
    if (!(saved ? load_mode(h, saved) : setup_mode(h))) return fail();
    if (!xn_bounds(h)) return fail();
    return h;
}

// Sections of a saved index for this handle (header offsets), total size.
int64_t saved_layout(const Handle* h, SavedHeader& hd) {
    memset(&hd, 0, sizeof(hd));
    memcpy(hd.magic, "ANNMIIDX", 8);
    hd.version = kSavedVersion;
    hd.mode = h->p.mode;
    hd.n = h->n;
    hd.d = h->d;
    hd.kdim = h->kdim;
    hd.row_stride = h->xs_stride;
    hd.sq_variant = h->p.sq_variant;
    hd.nlist = h->ivf ? h->nlist : 0;
    hd.alpha_x = h->alpha_x;
    const bool flat32 = h->kind == SCAN_F32 && !h->ivf;
    const int64_t sz[7] = {
        h->sq_scale ? (int64_t)h->kdim * 4 : 0,
        h->sq_offset ? (int64_t)h->kdim * 4 : 0,
        flat32 ? 0 : h->n * h->xs_stride,
        flat32 ? 0 : h->n * 4,
        h->ivf ? h->n * 4 : 0,
        h->ivf ? ((int64_t)h->nlist + 1) * 8 : 0,
        h->ivf ? (int64_t)h->nlist * h->kdim32 * 4 : 0,
    };
    int64_t at = 256;
    for (int i = 0; i < 7; ++i) {
        hd.off[i] = at;
        at += (sz[i] + 255) / 256 * 256;
    }
    hd.off[7] = at;
    return at;
}

// The derived arrays of a saved index (header already checked by the caller).
bool load_mode(Handle* h, const uint8_t* blob) {
    SavedHeader hd;
    memcpy(&hd, blob, sizeof(hd));
    hipStream_t st = h->st;
    const int64_t n = h->n;
    auto sec = [&](int i) { return hd.off[i + 1] - hd.off[i]; };
    auto up = [&](void** dst, int i, int64_t bytes) -> bool {
        if (bytes <= 0) return true;
        if (sec(i) < bytes) {
            set_err("saved index: section %d holds %lld bytes, %lld needed", i, (long long)sec(i), (long long)bytes);
            return false;
        }
        HIP_OR(hipMalloc(dst, (size_t)bytes), false);
        HIP_OR(hipMemcpyAsync(*dst, blob + hd.off[i], (size_t)bytes, hipMemcpyHostToDevice, st), false);
        return true;
    };
    if (h->ivf) {
        h->nlist = hd.nlist;
        h->list_off.resize((size_t)hd.nlist + 1);
        if (sec(5) < (int64_t)h->list_off.size() * 8) {
            set_err("saved index: list offsets truncated");
            return false;
        }
        memcpy(h->list_off.data(), blob + hd.off[5], h->list_off.size() * 8);
        int64_t prev = 0;
        for (size_t i = 0; i < h->list_off.size(); ++i) {
            if (h->list_off[i] < prev || h->list_off[i] > n) {
                set_err("saved index: list offsets are not a partition of the %lld rows", (long long)n);
                return false;
            }
            prev = h->list_off[i];
        }
        if (h->list_off.front() != 0 || h->list_off.back() != n) {
            set_err("saved index: list offsets are not a partition of the %lld rows", (long long)n);
            return false;
        }
        if (!up((void**)&h->perm, 4, n * 4) || !up((void**)&h->centroids, 6, (int64_t)hd.nlist * h->kdim32 * 4))
            return false;
        h->bytes += n * 4 + (int64_t)hd.nlist * h->kdim32 * 4;
    }
    if (h->kind == SCAN_F32 && !h->ivf) {
        h->kdim = h->kdim32;
        h->Xs = h->X32;
        h->xs_stride = (int64_t)h->kdim * 4;
        h->xn = h->xn32;
        return hipStreamSynchronize(st) == hipSuccess;
    }
    h->kdim = hd.kdim;
    h->xs_stride = hd.row_stride;
    h->alpha_x = hd.alpha_x;
    h->tiled = h->ivf ? 0 : 1;
    if (h->tiled) {
        // the image holds row-major rows: tile them on the host
        if (sec(2) < n * h->xs_stride) {
            set_err("saved index: section 2 holds %lld bytes, %lld needed", (long long)sec(2),
                    (long long)(n * h->xs_stride));
            return false;
        }
        const int64_t n_alloc = (n + 15) / 16 * 16;
        std::vector<uint8_t> t((size_t)(n_alloc * h->xs_stride), 0);
        tile_rows(blob + hd.off[2], t.data(), n, h->xs_stride, tile_cb_of(h->kind), tile_flip_of(h->kind));
        HIP_OR(hipMalloc(&h->Xs, t.size()), false);
        HIP_OR(hipMemcpyAsync(h->Xs, t.data(), t.size(), hipMemcpyHostToDevice, st), false);
        HIP_OR(hipStreamSynchronize(st), false);
        if (!up((void**)&h->xn, 3, n * 4)) return false;
        h->bytes += n_alloc * h->xs_stride + n * 4;
    } else {
        if (!up(&h->Xs, 2, n * h->xs_stride) || !up((void**)&h->xn, 3, n * 4)) return false;
        h->bytes += n * h->xs_stride + n * 4;
    }
    h->own_xs = true;
    if (h->kind == SCAN_SQ8 || h->kind == SCAN_SQ4) {
        if (!up((void**)&h->sq_scale, 0, (int64_t)h->kdim * 4) || !up((void**)&h->sq_offset, 1, (int64_t)h->kdim * 4))
            return false;
    }
    HIP_OR(hipStreamSynchronize(st), false);
    if (h->ivf && !ivf_attach(h)) return false;
    return true;
}

int32_t batch_impl(Handle* h, const float* Q, int32_t B, int32_t k, int32_t R, int64_t* ids, float* dists,
                   bool device, int32_t nprobe) {
    g_err[0] = 0;
    if (!h) {
        set_err("null handle");
        return -1;
    }
    if (B < 0 || k <= 0 || (B > 0 && (!Q || !ids || !dists))) {
        set_err("bad arguments: B=%d k=%d", B, k);
        return -1;
    }
    if (k > 2048) {
        set_err("k=%d exceeds 2048", k);
        return -1;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    const int kk = (int)std::min<int64_t>(k, h->n);
    if (R <= 0) R = h->p.rerank > 0 ? h->p.rerank : 4 * k;
    h->last_nrr.clear();
    h->last_flags.clear();
    const int step = (int)std::min<int64_t>(h->p.max_batch, std::max(B, 1));
    if (!ensure_ws(h, step, k)) return -1;
    for (int s = 0; s < B; s += step) {
        const int b = std::min(step, B - s);
        hipMemcpyKind kin = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        HIP_OR(hipMemcpyAsync(h->dQ32, Q + (int64_t)s * h->d, (size_t)b * h->d * 4, kin, h->st), -1);
        int64_t* oid = device ? ids + (int64_t)s * k : h->d_ids;
        float* od = device ? dists + (int64_t)s * k : h->d_dist;
        if (!run_step(h, b, k, R, nprobe, oid, od)) return -1;
        if (!device) {
            HIP_OR(hipMemcpyAsync(ids + (int64_t)s * k, h->d_ids, (size_t)b * k * 8, hipMemcpyDeviceToHost, h->st), -1);
            HIP_OR(hipMemcpyAsync(dists + (int64_t)s * k, h->d_dist, (size_t)b * k * 4, hipMemcpyDeviceToHost, h->st),
                   -1);
            HIP_OR(hipStreamSynchronize(h->st), -1);
        }
    }
    return kk;
}

// src/linear_index.cpp:90-121 bit for bit (compat.hip), one query at a time.
int32_t linear_index_impl(Handle* h, const float* Q, int32_t B, int32_t k, int64_t* ids, float* vals) {
    g_err[0] = 0;
    if (!h || B < 0 || k <= 0 || (B > 0 && (!Q || !ids || !vals))) {
        set_err("bad arguments");
        return -1;
    }
    if (h->kind != SCAN_F32 || h->ivf) {
        set_err("the linear_index query needs an exact (fp32) index");
        return -1;
    }
    if (k > 2048) {
        set_err("k=%d exceeds 2048", k);
        return -1;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    const int64_t n = h->n;
    const int d = h->d;
    const int W = d / 96 + 2;
    const int64_t G = (n + 4095) / 4096;
    if (!h->li_P) {
        HIP_OR(hipMalloc(&h->li_P, (size_t)n * W * 4), -1);
        HIP_OR(hipMalloc(&h->li_BT, (size_t)G * 64 * 4), -1);
        HIP_OR(hipMalloc(&h->li_R, (size_t)G * 64 * 4), -1);
        HIP_OR(hipMalloc(&h->li_q, (size_t)d * 4), -1);
        HIP_OR(hipMalloc(&h->li_flag, 4), -1);
    }
    if (h->li_k < k) {
        dfree(h->li_bkey);
        dfree(h->li_bid);
        dfree(h->li_ids);
        dfree(h->li_vals);
        h->li_bkey = nullptr;
        h->li_bid = nullptr;
        h->li_ids = nullptr;
        h->li_vals = nullptr;
        h->li_k = k;
        HIP_OR(hipMalloc(&h->li_bkey, (size_t)G * k * 4), -1);
        HIP_OR(hipMalloc(&h->li_bid, (size_t)G * k * 4), -1);
        HIP_OR(hipMalloc(&h->li_ids, (size_t)k * 8), -1);
        HIP_OR(hipMalloc(&h->li_vals, (size_t)k * 4), -1);
    }
    const int kk = (int)std::min<int64_t>(k, n);
    hipStream_t st = h->st;
    for (int b = 0; b < B; ++b) {
        uint32_t zero = 0, flag = 0;
        HIP_OR(hipMemcpyAsync(h->li_q, Q + (int64_t)b * d, (size_t)d * 4, hipMemcpyHostToDevice, st), -1);
        HIP_OR(hipMemcpyAsync(h->li_flag, &zero, 4, hipMemcpyHostToDevice, st), -1);
        HIP_OR(launch_linear_index_query(h->X32, h->kdim32, n, d, h->li_q, k, h->li_P, h->li_BT, h->li_R, h->li_bkey,
                                         h->li_bid, h->li_ids, h->li_vals, h->li_flag, st),
               -1);
        HIP_OR(hipMemcpyAsync(ids + (int64_t)b * k, h->li_ids, (size_t)kk * 8, hipMemcpyDeviceToHost, st), -1);
        HIP_OR(hipMemcpyAsync(vals + (int64_t)b * k, h->li_vals, (size_t)kk * 4, hipMemcpyDeviceToHost, st), -1);
        HIP_OR(hipMemcpyAsync(&flag, h->li_flag, 4, hipMemcpyDeviceToHost, st), -1);
        HIP_OR(hipStreamSynchronize(st), -1);
        if (flag) {
            set_err("query %d: more than 2048 rows tie at the rank-%d value", b, kk);
            return -1;
        }
        for (int j = kk; j < k; ++j) {
            ids[(int64_t)b * k + j] = -1;
            vals[(int64_t)b * k + j] = __builtin_inff();
        }
        h->ndist += n;  // linear_index.cpp:108
    }
    return kk;
}

// Share an index: same device arrays, new stream and workspace.
Handle* clone_impl(Handle* src) {
    Handle* h = new Handle();
    h->p = src->p;
    h->dev = src->dev;
    h->n = src->n;
    h->d = src->d;
    h->kdim32 = src->kdim32;
    h->X32 = src->X32;
    h->xn32 = src->xn32;
    h->maxnorm = src->maxnorm;
    h->xn_lo = src->xn_lo;
    h->xn_amax = src->xn_amax;
    h->prefilter = src->prefilter;
    h->fixed = src->fixed;
    h->tuned = src->tuned;
    h->kind = src->kind;
    h->kdim = src->kdim;
    h->Xs = src->Xs;
    h->xs_stride = src->xs_stride;
    h->tiled = src->tiled;
    h->xn = src->xn;
    h->alpha_x = src->alpha_x;
    h->sq_scale = src->sq_scale;
    h->sq_offset = src->sq_offset;
    h->cap = src->cap;
    h->ivf = src->ivf;
    h->nlist = src->nlist;
    h->perm = src->perm;
    h->list_off = src->list_off;
    h->d_list_off = src->d_list_off;
    h->max_list_len = src->max_list_len;
    h->centroids = src->centroids;
    h->is_clone = true;
    h->shared = true;
    src->shared = true;
    drop_graphs(src);  // the source now runs eagerly too (caller holds its lock)
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
        set_err("hipStreamCreate failed");
        delete h;
        return nullptr;
    }
    if (src->cent) {
        h->cent = clone_impl(src->cent);
        if (!h->cent) {
            free_handle(h);
            return nullptr;
        }
        share_stream(h->cent, h);
    }
    return h;
}

// ------------------------------------------------------------------ IVF --
// k-means (lsh_index_optimized.cpp:551-616) on the GPU; see ivf.hip.
bool ivf_build(Handle* h) {
    hipStream_t st = h->st;
    const int64_t n = h->n;
    const int d = h->d, kd = h->kdim32;
    int nlist = h->p.nlist > 0 ? h->p.nlist : 256;
    if (nlist > n) nlist = (int)n;
    const int iters = h->p.kmeans_iters > 0 ? h->p.kmeans_iters : 15;
    const uint32_t seed = h->p.kmeans_seed > 0 ? (uint32_t)h->p.kmeans_seed : 1337u;
    h->nlist = nlist;
    // initial centroids: rows rand_indices[c % n] of shuffle(iota(n), mt19937(seed)) (:553-561)
    std::vector<uint32_t> ri((size_t)n);
    std::iota(ri.begin(), ri.end(), 0u);
    {
        std::vector<int> tmp((size_t)n);
        std::iota(tmp.begin(), tmp.end(), 0);
        std::shuffle(tmp.begin(), tmp.end(), std::mt19937(seed));
        for (int64_t i = 0; i < n; ++i) ri[(size_t)i] = (uint32_t)tmp[(size_t)i];
    }
    std::vector<uint32_t> init((size_t)nlist);
    for (int c = 0; c < nlist; ++c) init[(size_t)c] = ri[(size_t)(c % n)];
    uint32_t *d_rand = nullptr, *d_init = nullptr, *assign = nullptr, *counts = nullptr, *offs = nullptr,
             *keys_sorted = nullptr, *perm = nullptr, *iota = nullptr;
    unsigned long long* am = nullptr;
    float *cn = nullptr, *thr_inf = nullptr;
    void *sort_tmp = nullptr, *scan_tmp = nullptr;
    KmSeg* d_segs = nullptr;
    int* d_segfirst = nullptr;
    double* partial = nullptr;
    struct Guard {
        std::vector<void*> ptrs;
        ~Guard() {
            for (void* p : ptrs) dfree(p);
        }
    } guard;
    auto alloc = [&](void** p, size_t bytes) -> bool {
        if (hipMalloc(p, bytes) != hipSuccess) {
            set_err("hipMalloc of %zu bytes failed in the IVF build", bytes);
            return false;
        }
        guard.ptrs.push_back(*p);
        return true;
    };
    if (!alloc((void**)&d_rand, n * 4) || !alloc((void**)&d_init, (size_t)nlist * 4) || !alloc((void**)&assign, n * 4) ||
        !alloc((void**)&counts, (size_t)nlist * 4) || !alloc((void**)&offs, (size_t)nlist * 4) ||
        !alloc((void**)&keys_sorted, n * 4) || !alloc((void**)&iota, n * 4) || !alloc((void**)&am, n * 8) ||
        !alloc((void**)&cn, (size_t)nlist * 4) || !alloc((void**)&thr_inf, 4096 * 4))
        return false;
    HIP_OR(hipMalloc(&h->perm, n * 4), false);
    perm = h->perm;
    HIP_OR(hipMalloc(&h->centroids, (size_t)nlist * kd * 4), false);
    h->bytes += n * 4 + (int64_t)nlist * kd * 4;
    HIP_OR(hipMemcpyAsync(d_rand, ri.data(), n * 4, hipMemcpyHostToDevice, st), false);
    HIP_OR(hipMemcpyAsync(d_init, init.data(), (size_t)nlist * 4, hipMemcpyHostToDevice, st), false);
    HIP_OR(launch_gather_rows_f32(h->X32, kd, d_init, nlist, kd, h->centroids, st), false);
    HIP_OR(launch_iota_u32(iota, n, st), false);
    HIP_OR(launch_fill_f32(thr_inf, __builtin_inff(), 4096, st), false);
    int key_bits = 1;
    while ((1 << key_bits) < nlist) ++key_bits;
    size_t sort_bytes = 0, scan_bytes = 0;
    HIP_OR(sort_pairs_u32(nullptr, sort_bytes, assign, keys_sorted, iota, perm, n, key_bits, st), false);
    HIP_OR(exclusive_sum_u32(nullptr, scan_bytes, counts, offs, nlist, st), false);
    if (!alloc(&sort_tmp, sort_bytes + 16) || !alloc(&scan_tmp, scan_bytes + 16)) return false;
    const int64_t kSeg = 8192;
    const int max_segs = nlist + (int)(n / kSeg) + 1;
    if (!alloc((void**)&d_segs, (size_t)max_segs * sizeof(KmSeg)) || !alloc((void**)&d_segfirst, (size_t)(nlist + 1) * 4) ||
        !alloc((void**)&partial, (size_t)max_segs * d * 8))
        return false;
    std::vector<uint32_t> hoff((size_t)nlist), hcnt((size_t)nlist);
    for (int it = 0; it < iters; ++it) {
        // assign: argmin over centroids (scan.hip argmin mode), 4096 points per launch
        HIP_OR(launch_encode(SCAN_F32, h->centroids, kd, nlist, d, kd, h->centroids, cn, nullptr, 1.f, nullptr, nullptr,
                             0, st),
               false);
        HIP_OR(launch_fill_u64(am, ~0ull, n, st), false);
        for (int64_t s0 = 0; s0 < n; s0 += 4096) {
            const int b = (int)std::min<int64_t>(4096, n - s0);
            ScanArgs a{};
            a.kind = SCAN_F32;
            a.X = h->centroids;
            a.xstride = (int64_t)kd * 4;
            a.xn = cn;
            a.row_begin = 0;
            a.row_end = nlist;
            a.d = d;
            a.kdim = kd;
            a.Q = h->X32 + s0 * kd;
            a.qn = h->xn32 + s0;
            a.qscale = nullptr;
            a.thr = thr_inf;
            a.B = b;
            a.cand = CandBuf{nullptr, nullptr, nullptr, 0};
            a.argmin = am + s0;
            HIP_OR(launch_scan(a, st), false);
        }
        HIP_OR(launch_fill_u32(counts, 0u, nlist, st), false);
        HIP_OR(launch_argmin_to_assign(am, n, nlist, assign, counts, st), false);
        // stable sort by cluster: the list layout, ascending row id inside a list
        HIP_OR(sort_pairs_u32(sort_tmp, sort_bytes, assign, keys_sorted, iota, perm, n, key_bits, st), false);
        HIP_OR(exclusive_sum_u32(scan_tmp, scan_bytes, counts, offs, nlist, st), false);
        HIP_OR(hipMemcpyAsync(hoff.data(), offs, (size_t)nlist * 4, hipMemcpyDeviceToHost, st), false);
        HIP_OR(hipMemcpyAsync(hcnt.data(), counts, (size_t)nlist * 4, hipMemcpyDeviceToHost, st), false);
        HIP_OR(hipStreamSynchronize(st), false);
        // update: fixed segments summed in order, then per-cluster means
        std::vector<KmSeg> segs;
        std::vector<int> first((size_t)nlist + 1);
        for (int c = 0; c < nlist; ++c) {
            first[(size_t)c] = (int)segs.size();
            const int64_t b0 = hoff[(size_t)c], b1 = b0 + hcnt[(size_t)c];
            for (int64_t p = b0; p < b1; p += kSeg) segs.push_back(KmSeg{p, std::min(b1, p + kSeg)});
        }
        first[(size_t)nlist] = (int)segs.size();
        HIP_OR(hipMemcpyAsync(d_segs, segs.data(), segs.size() * sizeof(KmSeg), hipMemcpyHostToDevice, st), false);
        HIP_OR(hipMemcpyAsync(d_segfirst, first.data(), first.size() * 4, hipMemcpyHostToDevice, st), false);
        HIP_OR(launch_segment_sums(h->X32, kd, perm, d_segs, (int)segs.size(), d, partial, st), false);
        HIP_OR(launch_centroid_update(partial, d_segfirst, counts, nlist, d, kd, h->X32, kd, d_rand, n, h->centroids, st),
               false);
        HIP_OR(hipStreamSynchronize(st), false);
    }
    h->list_off.assign((size_t)nlist + 1, 0);
    for (int c = 0; c < nlist; ++c) h->list_off[(size_t)c] = hoff[(size_t)c];
    h->list_off[(size_t)nlist] = n;
    return true;
}

// IVF step, enqueued with no host round trip (so it can be captured):
//   prep_queries             SQ8 operands; resets counts, flags, thresholds
//   probe                    exact top-nprobe over the centroids (the
//                            centroid index's flat step on this stream)
//   plan                     queries per list (every list for a query whose
//                            probes hold fewer than k rows), item prefixes
//   list scan per pass       pass 0: the first S rows of every probed list
//                            with an open filter; pass p: rows [S 8^(p-1),
//                            S 8^p) against the rsel-th smallest key so far
//   finalize                 fp32 re-score of rsel candidates, original ids
constexpr int kIvfMaxPass = 16;
constexpr int kIvfChunk = 256;         // rows of a list per work item
constexpr int kIvfCoarse = 1 << 18;    // coarse item table entries a pass (16M items)

// First-pass rows of each probed list: enough for ~4 rsel candidates over
// the probes, within half the candidate list.
int64_t ivf_first_pass(const Handle* h, int rsel, int nprobe) {
    int64_t S = std::max<int64_t>(256, (4 * (int64_t)rsel + nprobe - 1) / nprobe);
    return std::max<int64_t>(16, std::min<int64_t>(S, h->cap / (2 * (int64_t)nprobe)));
}

int ivf_passes(const Handle* h, int64_t S) {
    int np = 1;
    int64_t lo, hi;
    ivf_pass_bounds(S, 0, lo, hi);
    while (hi < h->max_list_len && np < kIvfMaxPass) ivf_pass_bounds(S, np++, lo, hi);
    return np;
}

// Buffers of the IVF step for ws_B queries probing up to nprobe lists, and
// the centroid index's workspace (all before any capture).
bool ensure_ivf_ws(Handle* h, int nprobe) {
    const int B = h->ws_B;
    // captured steps hold these buffers' addresses: any reallocation drops them
    if (h->probe_cap < B * nprobe || h->plan_B < B || !h->d_qcount || h->cent->ws_B < B) drop_graphs(h);
    if (h->probe_cap < B * nprobe) {
        dfree(h->d_probe);
        dfree(h->d_probe_d);
        h->d_probe = nullptr;
        h->d_probe_d = nullptr;
        h->probe_cap = B * nprobe;
        HIP_OR(hipMalloc(&h->d_probe, (size_t)h->probe_cap * 8), false);
        HIP_OR(hipMalloc(&h->d_probe_d, (size_t)h->probe_cap * 4), false);
    }
    if (h->plan_B < B) {
        dfree(h->d_every);
        dfree(h->d_gq);
        h->d_every = nullptr;
        h->d_gq = nullptr;
        HIP_OR(hipMalloc(&h->d_every, (size_t)B), false);
        HIP_OR(hipMalloc(&h->d_gq, (size_t)B * h->nlist * 4), false);
        h->plan_B = B;
    }
    if (!h->d_qcount) {
        HIP_OR(hipMalloc(&h->d_qcount, (size_t)h->nlist * 4), false);
        HIP_OR(hipMalloc(&h->d_qoff, (size_t)h->nlist * 4), false);
        HIP_OR(hipMalloc(&h->d_qfill, (size_t)h->nlist * 4), false);
        HIP_OR(hipMalloc(&h->d_plan, (size_t)kIvfMaxPass * (h->nlist + 1) * 4), false);
        HIP_OR(hipMalloc(&h->d_nitems, (size_t)kIvfMaxPass * 4), false);
        HIP_OR(hipMalloc(&h->d_next, (size_t)kIvfMaxPass * 4), false);
        HIP_OR(hipMalloc(&h->d_coarse, (size_t)kIvfMaxPass * kIvfCoarse * 4), false);
    }
    return ensure_ws(h->cent, B, nprobe);
}

// Algorithmic bytes and flops of each pass's list-scan launch (timing only):
// a list's rows in the pass once per group of 16 queries probing it, and
// 2 * rows * d per query.  Reads the step's per-list query counts back.
bool ivf_pass_work(Handle* h, int npass, int64_t S, std::vector<double>& pbytes, std::vector<double>& pflops) {
    std::vector<uint32_t> qc((size_t)h->nlist);
    HIP_OR(hipMemcpyAsync(qc.data(), h->d_qcount, qc.size() * 4, hipMemcpyDeviceToHost, h->st), false);
    HIP_OR(hipStreamSynchronize(h->st), false);
    pbytes.assign((size_t)npass, 0.0);
    pflops.assign((size_t)npass, 0.0);
    for (int l = 0; l < h->nlist; ++l) {
        if (!qc[(size_t)l]) continue;
        const int64_t len = h->list_off[(size_t)l + 1] - h->list_off[(size_t)l];
        for (int p = 0; p < npass; ++p) {
            int64_t lo, hi;
            ivf_pass_bounds(S, p, lo, hi);
            const int64_t rows = std::min(len, hi) - lo;
            if (rows <= 0) break;
            pbytes[(size_t)p] += (double)rows * h->xs_stride * ((qc[(size_t)l] + 15) / 16);
            pflops[(size_t)p] += 2.0 * rows * (double)qc[(size_t)l] * h->d;
        }
    }
    return true;
}

bool enqueue_ivf(Handle* h, int B, int k, int rsel, int nprobe, int64_t* ids_out, float* d_out, Profiler& pr,
                 std::vector<std::pair<int, int>>& scan_ev, std::vector<std::pair<int, int>>& sel_ev,
                 std::vector<double>& scan_rows, std::vector<std::string>& scan_names, int& ev_f0, int& ev_f1) {
    hipStream_t st = h->st;
    const int kk = (int)std::min<int64_t>(k, h->n);
    pr.ev_prep0 = pr.ev();
    HIP_OR(launch_prep_queries(SCAN_SQ8, h->dQ32, B, h->d, h->kdim, h->dQs, h->qn, h->qscale, h->slack, h->maxnorm,
                               1.f, h->sq_scale, h->sq_offset, st, h->cnt, h->ovf, h->thr),
           false);
    // probe: the nprobe nearest centroids in exact order, lower id on ties
    Handle* c = h->cent;
    HIP_OR(hipMemcpyAsync(c->dQ32, h->dQ32, (size_t)B * h->d * 4, hipMemcpyDeviceToDevice, st), false);
    {
        std::vector<std::pair<int, int>> cs, cl;
        std::vector<double> cr;
        std::vector<std::string> cn;
        int cu = 0, c0 = -1, c1 = -1;
        Profiler cpr{c, cu, &cs, &cl, &cr, &cn, -1, -1};
        if (!enqueue_flat(c, B, nprobe, nprobe, h->d_probe, h->d_probe_d, cpr, cs, cl, cr, cn, c0, c1)) return false;
    }
    const int64_t S = ivf_first_pass(h, rsel, nprobe);
    const int npass = ivf_passes(h, S);
    IvfPlanArgs pa{};
    pa.probe = h->d_probe;
    pa.B = B;
    pa.nprobe = nprobe;
    pa.nlist = h->nlist;
    pa.kk = kk;
    pa.list_off = h->d_list_off;
    pa.qcount = h->d_qcount;
    pa.qoff = h->d_qoff;
    pa.qfill = h->d_qfill;
    pa.gq = h->d_gq;
    pa.every = h->d_every;
    pa.plan = h->d_plan;
    pa.nitems = h->d_nitems;
    pa.coarse = h->d_coarse;
    pa.coarse_cap = kIvfCoarse;
    pa.npass = npass;
    pa.S = S;
    pa.chunk = kIvfChunk;
    HIP_OR(launch_ivf_plan(pa, st), false);
    HIP_OR(launch_fill_u32(h->d_next, 0u, kIvfMaxPass, st), false);
    if (h->prof && !ivf_pass_work(h, npass, S, pr.bytes_override, pr.flops_override)) return false;
    pr.ev_prep1 = pr.ev();
    CandBuf cb{h->cnt, h->ckey, h->crid, h->cap};
    IvfScanArgs ia{};
    ia.codes = h->Xs;
    ia.xstride = h->xs_stride;
    ia.xn = h->xn;
    ia.perm = h->perm;
    ia.kdim = h->kdim;
    ia.Q = h->dQs;
    ia.qn = h->qn;
    ia.qscale = h->qscale;
    ia.thr = h->thr;
    ia.list_off = h->d_list_off;
    ia.qcount = h->d_qcount;
    ia.qoff = h->d_qoff;
    ia.nlist = h->nlist;
    ia.chunk = kIvfChunk;
    ia.gq = h->d_gq;
    ia.cand = cb;
    for (int pass = 0; pass < npass; ++pass) {
        if (pass > 0) {
            SelectArgs sa{cb, B, rsel, h->slack, h->thr, h->ovf};
            int e0, e1;
            pr.arm(e0, e1);
            const hipError_t le = launch_select_threshold(sa, st);
            pr.disarm(e0, e1);
            HIP_OR(le, false);
            if (h->prof) sel_ev.push_back({e0, e1});
        }
        ia.plan = h->d_plan + (int64_t)pass * (h->nlist + 1);
        ia.nitems = h->d_nitems + pass;
        ia.coarse = h->d_coarse + (int64_t)pass * kIvfCoarse;
        ia.coarse_cap = kIvfCoarse;
        ia.next = h->d_next + pass;
        ivf_pass_bounds(S, pass, ia.lo, ia.hi);
        int e0, e1;
        pr.arm(e0, e1);
        const hipError_t le = launch_ivf_scan(ia, st);
        pr.disarm(e0, e1);
        HIP_OR(le, false);
        if (h->prof) {
            scan_ev.push_back({e0, e1});
            scan_rows.push_back(0.0);
            scan_names.push_back("ivf_scan_kernel");
        }
    }
    return launch_final(h, B, k, rsel, ids_out, d_out, cb, pr, ev_f0, ev_f1);
}

// Autotuner: for batch sizes 1, 2, 4, ... up to max_batch, time whole query
// steps with every candidate kernel shape (streaming kernel up to 128
// queries; each wide configuration that fits the LDS at the default row
// chunk, then the best one at other chunk sizes) on queries drawn from the
// index's own rows, and keep the fastest per batch size.  Answers do not
// depend on the shape (every shape computes the same keys), only the time.
bool autotune_impl(Handle* h, int max_batch, int reps) {
    if (h->ivf) {
        set_err("autotune: IVF handles take one list-scan kernel; nothing to tune");
        return false;
    }
    max_batch = std::max(1, std::min<int>(max_batch, (int)std::min<int64_t>(h->p.max_batch, 4096)));
    reps = std::max(1, std::min(reps, 10));
    if (!ensure_ws(h, max_batch, 100)) return false;
    // queries: rows spread over the base (fp32 copy, padded row stride)
    const int64_t stride_rows = std::max<int64_t>(1, h->n / max_batch);
    HIP_OR(hipMemcpy2DAsync(h->dQ32, (size_t)h->d * 4, h->X32, (size_t)stride_rows * h->kdim32 * 4, (size_t)h->d * 4,
                            (size_t)std::min<int64_t>(max_batch, h->n), hipMemcpyDeviceToDevice, h->st),
           false);
    HIP_OR(hipStreamSynchronize(h->st), false);
    const bool prof = h->prof;
    h->prof = false;
    h->tuning = true;
    drop_graphs(h);
    h->tuned.clear();
    const Handle::ScanShape saved = h->fixed;
    const int k = (int)std::min<int64_t>(100, h->n);
    const int R = h->p.rerank > 0 ? h->p.rerank : 4 * k;
    auto time_step = [&](int B, const Handle::ScanShape& shp, double& best) -> bool {
        h->fixed = shp;
        best = 1e30;
        for (int r = 0; r <= reps; ++r) {  // the first run warms the variant up
            const auto t0 = std::chrono::steady_clock::now();
            if (!run_step(h, B, k, R, 0, h->d_ids, h->d_dist)) return false;
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (r > 0) best = std::min(best, ms);
        }
        return true;
    };
    std::string rep = "[";
    std::vector<Handle::ScanShape> table;
    bool ok = true;
    for (int B = 1; B <= max_batch && ok; B *= 2) {
        Handle::ScanShape def{}, best{};
        double t_def = 0, t_best = 1e30, t = 0;
        if (!(ok = time_step(B, def, t_def))) break;
        best = def;
        t_best = t_def;
        std::vector<Handle::ScanShape> cands;
        if (B <= 128) cands.push_back(Handle::ScanShape{0, 0, 0, 0});
        for (int c = 1; c < kWideCfgCount; ++c)
            if (wide_cfg_fits(h->kind, h->kdim, c)) cands.push_back(Handle::ScanShape{1, c, 0, 0});
        // the built-in wide configuration at other row-chunk sizes
        for (int64_t ch : {(int64_t)8192, (int64_t)16384, (int64_t)65536}) cands.push_back(Handle::ScanShape{1, 0, ch, 0});
        // a candidate replaces the built-in shape only when clearly faster
        // (2%: run-to-run noise of a step is about 1%)
        for (const auto& c : cands) {
            if (!(ok = time_step(B, c, t))) break;
            if (t < t_best && t < 0.98 * t_def) {
                t_best = t;
                best = c;
            }
        }
        if (ok && best.kernel == 1) {
            const Handle::ScanShape base = best;
            for (int64_t ch : {(int64_t)8192, (int64_t)16384, (int64_t)65536}) {
                Handle::ScanShape c = base;
                c.chunk = ch;
                if (!(ok = time_step(B, c, t))) break;
                if (t < t_best) {
                    t_best = t;
                    best = c;
                }
            }
        }
        // confirm a switch: time the built-in shape and the winner once more
        // each and keep the winner only if its best time still beats the
        // built-in one's by 2% (single timings of a step vary by a few
        // percent, which otherwise decides some picks)
        if (ok && best.kernel != -1) {
            double t_def2 = 0, t_best2 = 0;
            if (!(ok = time_step(B, def, t_def2)) || !(ok = time_step(B, best, t_best2))) break;
            t_def = std::min(t_def, t_def2);
            t_best = std::min(t_best, t_best2);
            if (!(t_best < 0.98 * t_def)) {
                best = def;
                t_best = t_def;
            }
        }
        if (!ok) break;
        table.push_back(best);
        char line[200];
        snprintf(line, sizeof(line), "%s{\"B\": %d, \"kernel\": \"%s\", \"cfg\": %d, \"chunk\": %lld, \"ms\": %.4f, \"default_ms\": %.4f}",
                 B == 1 ? "" : ", ", B, best.kernel == 0 ? "streaming" : (best.kernel == 1 ? "wide" : "default"),
                 best.cfg, (long long)best.chunk, t_best, t_def);
        rep += line;
    }
    h->fixed = saved;
    h->prof = prof;
    h->tuning = false;
    drop_graphs(h);
    if (!ok) return false;
    h->tuned = table;
    h->tune_report = rep + "]";
    return true;
}

// ------------------------------------------------ allocation-free queries --
// annmi_batch_query_noalloc runs a flat step on a static device array instead
// of the handle's workspace: up to kNoAllocMaxBatch queries, candidate lists
// of kNoAllocCap keys.  Sections (256-byte aligned): fp32 queries, scan
// operands, seven per-query words, candidate keys, candidate ids.
constexpr int kNoAllocMaxBatch = 256;
constexpr int kNoAllocCap = 16384;
constexpr size_t na_sec(size_t bytes) { return (bytes + 255) / 256 * 256; }
constexpr size_t kNaQ32 = na_sec((size_t)kNoAllocMaxBatch * kMaxDim * 4);
constexpr size_t kNaQs = na_sec((size_t)kNoAllocMaxBatch * kMaxDim * 4);
constexpr size_t kNaWord = na_sec((size_t)kNoAllocMaxBatch * 4);
constexpr size_t kNaList = na_sec((size_t)kNoAllocMaxBatch * kNoAllocCap * 4);
constexpr size_t kScratchBytes = kNaQ32 + kNaQs + 7 * kNaWord + 2 * kNaList;  // 40 MiB + 7 KiB

std::mutex g_noalloc_mu;         // one call at a time per process uses the array
void* g_scratch_dev[64] = {};    // the array's address on each device (hipGetSymbolAddress)

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

// Address of `scratch` on device `dev`, looked up once per device with that
// device current (the caller holds g_noalloc_mu and has made it current).
void* scratch_on(int dev) {
    if (dev < 0 || dev >= 64) {
        set_err("device %d out of range for the allocation-free entry point", dev);
        return nullptr;
    }
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_err("hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_dev[dev] = p;
    }
    return g_scratch_dev[dev];
}

// The handle's workspace pointers, swapped for sections of the scratch array
// for one call and restored on scope exit.
struct WorkspaceSwap {
    Handle* h;
    float *dQ32, *qn, *qscale, *slack, *thr, *ckey;
    void* dQs;
    uint32_t *cnt, *crid, *ovf, *nrr;
    int cap;
    bool prof;
    WorkspaceSwap(Handle* hh, uint8_t* base) : h(hh) {
        dQ32 = h->dQ32, qn = h->qn, qscale = h->qscale, slack = h->slack, thr = h->thr, ckey = h->ckey;
        dQs = h->dQs, cnt = h->cnt, crid = h->crid, ovf = h->ovf, nrr = h->nrr, cap = h->cap, prof = h->prof;
        uint8_t* at = base;
        auto take = [&](size_t bytes) {
            uint8_t* p = at;
            at += bytes;
            return p;
        };
        h->dQ32 = reinterpret_cast<float*>(take(kNaQ32));
        h->dQs = take(kNaQs);
        h->qn = reinterpret_cast<float*>(take(kNaWord));
        h->qscale = reinterpret_cast<float*>(take(kNaWord));
        h->slack = reinterpret_cast<float*>(take(kNaWord));
        h->thr = reinterpret_cast<float*>(take(kNaWord));
        h->cnt = reinterpret_cast<uint32_t*>(take(kNaWord));
        h->ovf = reinterpret_cast<uint32_t*>(take(kNaWord));
        h->nrr = reinterpret_cast<uint32_t*>(take(kNaWord));
        h->ckey = reinterpret_cast<float*>(take(kNaList));
        h->crid = reinterpret_cast<uint32_t*>(take(kNaList));
        h->cap = kNoAllocCap;
        h->prof = false;  // event timing would create events
    }
    ~WorkspaceSwap() {
        h->dQ32 = dQ32, h->qn = qn, h->qscale = qscale, h->slack = slack, h->thr = thr, h->ckey = ckey;
        h->dQs = dQs, h->cnt = cnt, h->crid = crid, h->ovf = ovf, h->nrr = nrr, h->cap = cap, h->prof = prof;
    }
};

int32_t noalloc_impl(Handle* h, const float* dQ, int32_t B, int32_t k, int32_t R, int64_t* ids, float* dists) {
    g_err[0] = 0;
    if (!h) {
        set_err("null handle");
        return -1;
    }
    if (h->ivf) {
        set_err("the allocation-free entry point serves the flat modes (exact, bf16, fp8, sq8, sq4), not IVF");
        return -1;
    }
    if (B < 0 || k <= 0 || (B > 0 && (!dQ || !ids || !dists))) {
        set_err("bad arguments: B=%d k=%d", B, k);
        return -1;
    }
    if (B > kNoAllocMaxBatch) {
        set_err("B=%d exceeds the %d queries the allocation-free entry point's scratch holds", B, kNoAllocMaxBatch);
        return -1;
    }
    if (k > 2048) {
        set_err("k=%d exceeds 2048", k);
        return -1;
    }
    const bool exact = h->kind == SCAN_F32;
    const int kk = (int)std::min<int64_t>(k, h->n);
    if (R <= 0) R = h->p.rerank > 0 ? h->p.rerank : 4 * k;
    int rsel = exact ? kk : std::max(kk, R);
    rsel = (int)std::min<int64_t>(rsel, h->n);
    if (!exact && rsel > 2048) {
        set_err("rerank R=%d exceeds the 2048-row re-score limit", rsel);
        return -1;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    std::lock_guard<std::mutex> lk2(g_noalloc_mu);
    int cur = 0;
    HIP_OR(hipGetDevice(&cur), -1);
    if (cur != h->dev) HIP_OR(hipSetDevice(h->dev), -1);
    struct DeviceBack {
        int cur, dev;
        ~DeviceBack() {
            if (cur != dev) (void)hipSetDevice(cur);
        }
    } back{cur, h->dev};
    uint8_t* base = static_cast<uint8_t*>(scratch_on(h->dev));
    if (!base) return -1;
    h->last_nrr.clear();
    h->last_flags.clear();
    if (B == 0) return kk;
    uint32_t flags[kNoAllocMaxBatch], nrr[kNoAllocMaxBatch];
    {
        WorkspaceSwap ws(h, base);
        HIP_OR(hipMemcpyAsync(h->dQ32, dQ, (size_t)B * h->d * 4, hipMemcpyDeviceToDevice, h->st), -1);
        std::vector<std::pair<int, int>> scan_ev, sel_ev;  // stay empty: nothing is timed
        std::vector<double> scan_rows;
        std::vector<std::string> scan_names;
        int ev_used = 0, ev_f0 = -1, ev_f1 = -1;
        Profiler pr{h, ev_used, &scan_ev, &sel_ev, &scan_rows, &scan_names, -1, -1};
        if (!enqueue_flat(h, B, k, rsel, ids, dists, pr, scan_ev, sel_ev, scan_rows, scan_names, ev_f0, ev_f1))
            return -1;
        HIP_OR(hipMemcpyAsync(flags, h->ovf, (size_t)B * 4, hipMemcpyDeviceToHost, h->st), -1);
        HIP_OR(hipMemcpyAsync(nrr, h->nrr, (size_t)B * 4, hipMemcpyDeviceToHost, h->st), -1);
        HIP_OR(hipStreamSynchronize(h->st), -1);
    }
    int64_t nd = 0;
    for (int i = 0; i < B; ++i) {
        if (flags[i] & 7u) {
            set_err("query %d needs the exact all-rows path (flags %u), which allocates workspace: answer it with "
                    "annmi_batch_query_device",
                    i, flags[i]);
            return -1;
        }
        nd += exact ? h->n : (int64_t)nrr[i];
    }
    h->ndist += nd;
    return kk;
}

}  // namespace

extern "C" {

void annmi_default_params(annmi_params* p, int32_t mode) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->mode = mode;
    p->rerank = 0;
    p->sq_variant = 0;
    p->base_on_device = 0;
    p->max_batch = 4096;
    p->nlist = 256;
    p->nprobe = 8;
    p->kmeans_iters = 15;
    p->kmeans_seed = 1337;
}

void* annmi_create(const float* base, int64_t n, int32_t d, const annmi_params* params) {
    return create_impl(base, n, d, params, 0);
}

int32_t annmi_query(void* h, const float* q, int32_t k, int32_t rerank, int64_t* out_ids, float* out_dists) {
    return batch_impl(static_cast<Handle*>(h), q, 1, k, rerank, out_ids, out_dists, false, 0);
}

int32_t annmi_batch_query(void* h, const float* Q, int32_t B, int32_t k, int32_t rerank, int64_t* out_ids,
                          float* out_dists) {
    return batch_impl(static_cast<Handle*>(h), Q, B, k, rerank, out_ids, out_dists, false, 0);
}

int32_t annmi_batch_query_ex(void* h, const float* Q, int32_t B, int32_t k, int32_t rerank, int32_t nprobe,
                             int64_t* out_ids, float* out_dists, int32_t on_device) {
    return batch_impl(static_cast<Handle*>(h), Q, B, k, rerank, out_ids, out_dists, on_device != 0, nprobe);
}

int32_t annmi_batch_query_device(void* h, const float* dQ, int32_t B, int32_t k, int32_t rerank, int64_t* d_ids,
                                 float* d_dists) {
    return batch_impl(static_cast<Handle*>(h), dQ, B, k, rerank, d_ids, d_dists, true, 0);
}

int32_t annmi_batch_query_noalloc(void* h, const float* dQ, int32_t B, int32_t k, int32_t rerank, int64_t* d_ids,
                                  float* d_dists) {
    return noalloc_impl(static_cast<Handle*>(h), dQ, B, k, rerank, d_ids, d_dists);
}

int64_t annmi_ndist(void* h) { return h ? static_cast<Handle*>(h)->ndist.load() : 0; }

int32_t annmi_last_batch_stats(void* hv, uint32_t* n_rescored, uint32_t* flags, int32_t max_q) {
    Handle* h = static_cast<Handle*>(hv);
    if (!h) {
        set_err("null handle");
        return -1;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    int32_t m = (int32_t)std::min<size_t>(h->last_nrr.size(), (size_t)std::max(max_q, 0));
    for (int i = 0; i < m; ++i) {
        if (n_rescored) n_rescored[i] = h->last_nrr[i];
        if (flags) flags[i] = h->last_flags[i];
    }
    return m;
}

void annmi_free(void* h) { free_handle(static_cast<Handle*>(h)); }

int32_t annmi_query_linear_index(void* h, const float* Q, int32_t B, int32_t k, int64_t* out_ids,
                                 float* out_values) {
    return linear_index_impl(static_cast<Handle*>(h), Q, B, k, out_ids, out_values);
}

void* annmi_clone(void* h) {
    if (!h) {
        set_err("null handle");
        return nullptr;
    }
    Handle* src = static_cast<Handle*>(h);
    std::lock_guard<std::mutex> lk(src->mu);
    return clone_impl(src);
}

const char* annmi_last_error(void) { return g_err; }

int64_t annmi_index_bytes(void* h) { return h ? static_cast<Handle*>(h)->bytes : 0; }

int64_t annmi_save(void* hv, void* buf, int64_t bytes) {
    g_err[0] = 0;
    Handle* h = static_cast<Handle*>(hv);
    if (!h) {
        set_err("null handle");
        return -1;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    SavedHeader hd;
    const int64_t need = saved_layout(h, hd);
    if (!buf || bytes < need) return need;
    uint8_t* out = static_cast<uint8_t*>(buf);
    memset(out, 0, (size_t)need);
    memcpy(out, &hd, sizeof(hd));
    const void* src[7] = {h->sq_scale, h->sq_offset, h->Xs, h->xn, h->perm, nullptr, h->centroids};
    for (int i = 0; i < 7; ++i) {
        const int64_t len = hd.off[i + 1] - hd.off[i];
        if (len == 0) continue;
        if (i == 5) {
            memcpy(out + hd.off[5], h->list_off.data(), h->list_off.size() * 8);
            continue;
        }
        if (i == 2 && h->tiled) {  // rows leave the tiled layout: the image is row-major
            const int64_t n_alloc = (h->n + 15) / 16 * 16;
            std::vector<uint8_t> t((size_t)(n_alloc * h->xs_stride));
            HIP_OR(hipMemcpyAsync(t.data(), h->Xs, t.size(), hipMemcpyDeviceToHost, h->st), -1);
            HIP_OR(hipStreamSynchronize(h->st), -1);
            untile_rows(t.data(), out + hd.off[2], h->n, h->xs_stride, tile_cb_of(h->kind), tile_flip_of(h->kind));
            continue;
        }
        int64_t m = 0;  // bytes of the array (sections are padded)
        switch (i) {
            case 0: case 1: m = (int64_t)h->kdim * 4; break;
            case 2: m = h->n * h->xs_stride; break;
            case 3: case 4: m = h->n * 4; break;
            case 6: m = (int64_t)h->nlist * h->kdim32 * 4; break;
        }
        HIP_OR(hipMemcpyAsync(out + hd.off[i], src[i], (size_t)m, hipMemcpyDeviceToHost, h->st), -1);
        HIP_OR(hipStreamSynchronize(h->st), -1);
    }
    return need;
}

void* annmi_load(const float* base, int64_t n, int32_t d, const annmi_params* params, const void* saved,
                 int64_t bytes) {
    g_err[0] = 0;
    if (!saved || bytes < (int64_t)sizeof(SavedHeader)) {
        set_err("saved index: %lld bytes is too short", (long long)bytes);
        return nullptr;
    }
    SavedHeader hd;
    memcpy(&hd, saved, sizeof(hd));
    if (memcmp(hd.magic, "ANNMIIDX", 8) != 0 || hd.version != kSavedVersion) {
        set_err("saved index: not an annmi index of version %d", kSavedVersion);
        return nullptr;
    }
    annmi_params p;
    if (params)
        p = *params;
    else
        annmi_default_params(&p, hd.mode);
    const int kind = scan_kind_of(hd.mode);
    if (hd.mode != p.mode || hd.n != n || hd.d != d || kind < 0) {
        set_err("saved index is mode %d, %lld x %d; asked for mode %d, %lld x %d", hd.mode, (long long)hd.n, hd.d,
                p.mode, (long long)n, d);
        return nullptr;
    }
    const int kdim = kind == SCAN_F32 ? round_up(d, 16) : round_up(d, chunk_of(kind));
    if (hd.kdim != kdim || (kind != SCAN_F32 && hd.row_stride != scan_row_bytes(kind, kdim)) ||
        (hd.mode == ANNMI_IVF_SQ8 && (hd.nlist <= 0 || hd.nlist > n)) || hd.off[0] < (int64_t)sizeof(SavedHeader)) {
        set_err("saved index: layout (kdim %d, row stride %lld, nlist %d) does not match this library", hd.kdim,
                (long long)hd.row_stride, hd.nlist);
        return nullptr;
    }
    for (int i = 0; i < 7; ++i)
        if (hd.off[i + 1] < hd.off[i]) {
            set_err("saved index: corrupt section table");
            return nullptr;
        }
    if (hd.off[7] > bytes) {
        set_err("saved index: %lld bytes given, %lld in the index", (long long)bytes, (long long)hd.off[7]);
        return nullptr;
    }
    if (hd.mode == ANNMI_IVF_SQ8) {
        // checked on the host before anything reaches the device: the list
        // permutation becomes candidate row ids, so every entry must be a row
        // (< n) and appear exactly once, or a kernel would read past the base
        const uint8_t* blob = static_cast<const uint8_t*>(saved);
        if (hd.off[5] - hd.off[4] < n * 4) {
            set_err("saved index: list permutation truncated");
            return nullptr;
        }
        std::vector<uint8_t> seen((size_t)n, 0);
        for (int64_t i = 0; i < n; ++i) {
            uint32_t v;
            memcpy(&v, blob + hd.off[4] + i * 4, 4);
            if ((int64_t)v >= n || seen[v]) {
                set_err("saved index: list permutation entry %lld (%u) is not a permutation of the %lld rows",
                        (long long)i, v, (long long)n);
                return nullptr;
            }
            seen[v] = 1;
        }
    }
    p.sq_variant = hd.sq_variant;
    p.nlist = hd.nlist;
    return create_impl(base, n, d, &p, 0, static_cast<const uint8_t*>(saved));
}

int32_t annmi_autotune(void* hv, int32_t max_batch, int32_t reps, char* report, int32_t len) {
    g_err[0] = 0;
    Handle* h = static_cast<Handle*>(hv);
    if (!h) {
        set_err("null handle");
        return -1;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    if (max_batch > 0 && !autotune_impl(h, max_batch, reps)) return -1;
    const int32_t need = (int32_t)h->tune_report.size() + 1;
    if (report && len > 0) {
        const int32_t m = std::min(need, len);
        memcpy(report, h->tune_report.c_str(), (size_t)m - 1);
        report[m - 1] = 0;
    }
    return need;
}

int32_t annmi_set_scan_shape(void* hv, int32_t wide_min_batch, int32_t wide_cfg, int64_t chunk_rows) {
    g_err[0] = 0;
    Handle* h = static_cast<Handle*>(hv);
    if (!h) {
        set_err("null handle");
        return -1;
    }
    if (wide_min_batch < 0 || wide_cfg < 0 || wide_cfg >= kWideCfgCount || chunk_rows < 0 ||
        (chunk_rows > 0 && chunk_rows < 2048) || (wide_cfg > 0 && !wide_cfg_fits(h->kind, h->kdim, wide_cfg))) {
        set_err("bad scan shape: min_batch=%d cfg=%d chunk=%lld (cfg must fit the LDS, chunk 0 or >= 2048)",
                wide_min_batch, wide_cfg, (long long)chunk_rows);
        return -1;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    drop_graphs(h);
    h->tuned.clear();
    h->tune_report.clear();
    h->fixed = Handle::ScanShape{-1, wide_cfg, chunk_rows, wide_min_batch};
    return 0;
}

int32_t annmi_graph_count(void* hv) {
    if (!hv) return -1;
    Handle* h = static_cast<Handle*>(hv);
    std::lock_guard<std::mutex> lk(h->mu);
    int32_t c = 0;
    for (auto& g : h->graphs) c += g.exec ? 1 : 0;
    return c;
}

int64_t annmi_export(void* hv, int32_t what, void* out, int64_t bytes) {
    Handle* h = static_cast<Handle*>(hv);
    if (!h) {
        set_err("null handle");
        return -1;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    const void* src = nullptr;
    int64_t need = 0;
    switch (what) {
        case 0: src = h->Xs; need = h->n * h->xs_stride; break;
        case 1: src = h->sq_scale; need = h->sq_scale ? (int64_t)h->kdim * 4 : 0; break;
        case 2: src = h->sq_offset; need = h->sq_offset ? (int64_t)h->kdim * 4 : 0; break;
        case 3: src = h->xn; need = h->n * 4; break;
        case 5: src = h->perm; need = h->perm ? h->n * 4 : 0; break;
        case 6:
            need = (int64_t)h->list_off.size() * 8;
            if (out && bytes >= need && need) std::memcpy(out, h->list_off.data(), (size_t)need);
            return need;
        case 7: src = h->centroids; need = h->centroids ? (int64_t)h->nlist * h->kdim32 * 4 : 0; break;
        case 4: {
            // scalar parameters: kdim, row stride (bytes), fp8 alpha, max row norm
            if (out && bytes >= 32) {
                double v[4] = {(double)h->kdim, (double)h->xs_stride, (double)h->alpha_x, (double)h->maxnorm};
                std::memcpy(out, v, 32);
            }
            return 32;
        }
        default: set_err("unknown export %d", what); return -1;
    }
    if (!out || bytes < need) return need;
    if (what == 0 && h->tiled) {  // row-major rows, whatever the device layout
        const int64_t n_alloc = (h->n + 15) / 16 * 16;
        std::vector<uint8_t> t((size_t)(n_alloc * h->xs_stride));
        HIP_OR(hipMemcpyAsync(t.data(), h->Xs, t.size(), hipMemcpyDeviceToHost, h->st), -1);
        HIP_OR(hipStreamSynchronize(h->st), -1);
        untile_rows(t.data(), static_cast<uint8_t*>(out), h->n, h->xs_stride, tile_cb_of(h->kind), tile_flip_of(h->kind));
        return need;
    }
    HIP_OR(hipMemcpyAsync(out, src, (size_t)need, hipMemcpyDeviceToHost, h->st), -1);
    HIP_OR(hipStreamSynchronize(h->st), -1);
    return need;
}

int32_t annmi_profile(void* hv, int32_t enable, double* out8) {
    Handle* h = static_cast<Handle*>(hv);
    if (!h) {
        set_err("null handle");
        return -1;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    if (out8)
        for (int i = 0; i < 8; ++i) out8[i] = h->prof_acc[i];
    if (enable >= 0) {
        h->prof = enable != 0;
        for (int i = 0; i < 8; ++i) h->prof_acc[i] = 0;
        h->kstats.clear();
    }
    return 8;
}

int32_t annmi_profile_kernels(void* hv, char* buf, int32_t len) {
    Handle* h = static_cast<Handle*>(hv);
    if (!h) {
        set_err("null handle");
        return -1;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    std::string out = "[";
    char line[256];
    bool first = true;
    auto arr = [&](const char* name, const std::vector<float>& v) {
        out += ",\"";
        out += name;
        out += "\":[";
        for (size_t i = 0; i < v.size(); ++i) {
            snprintf(line, sizeof(line), "%s%.6g", i ? "," : "", (double)v[i]);
            out += line;
        }
        out += "]";
    };
    for (auto& kv : h->kstats) {
        snprintf(line, sizeof(line), "%s{\"kernel\":\"%s\",\"launches\":%.0f,\"ms\":%.6f,\"bytes\":%.6g,\"flops\":%.6g",
                 first ? "" : ",", kv.first.c_str(), kv.second.launches, kv.second.ms, kv.second.bytes, kv.second.flops);
        out += line;
        arr("launch_ms", kv.second.l_ms);
        arr("launch_bytes", kv.second.l_bytes);
        arr("launch_flops", kv.second.l_flops);
        out += "}";
        first = false;
    }
    out += "]";
    if (buf && len > 0) snprintf(buf, (size_t)len, "%s", out.c_str());
    return (int32_t)out.size() + 1;
}

// ------------------------------------------------ gpu_filter drop-in ABI --
void* gf_create(const float* train, long n, long dim) {
    annmi_params p;
    annmi_default_params(&p, ANNMI_SQ8);
    p.sq_variant = 1;  // gpu_filter.cu:110-126 encoder
    p.rerank = 500;    // competitors/gpu_filter/algorithm.py:43 default r
    if (dim > kMaxDim) {
        set_err("dim %ld > %d unsupported", dim, kMaxDim);
        return nullptr;
    }
    return create_impl(train, (int64_t)n, (int32_t)dim, &p, 0);
}

int gf_query(void* handle, const float* q, int k, int r, long long* out_ids) {
    g_err[0] = 0;
    Handle* h = static_cast<Handle*>(handle);
    if (!h) {
        set_err("null handle");
        return -1;
    }
    if (!q || !out_ids || k <= 0) {
        set_err("bad arguments: k=%d", k);
        return -1;
    }
    if (h->kind != SCAN_SQ8 || h->ivf) {
        set_err("gf_query needs a gf_create (per-dim SQ8) handle");
        return -1;
    }
    if (k > 2048) {
        set_err("k=%d exceeds 2048", k);
        return -1;
    }
    // gpu_filter.cu:145-146: r clamped to [k, n]
    if (r < k) r = k;
    if (r > h->n) r = (int)h->n;
    const int out_k = std::min(k, r);
    std::lock_guard<std::mutex> lk(h->mu);
    hipStream_t st = h->st;
    if (!ensure_ws(h, 1, k) || !ensure_robust(h)) return -1;
    RobustArgs ra = robust_args(h, ROBUST_GPU_FILTER, 1, k, h->d_ids, h->d_dist);
    ra.r = r;
    ra.codes = static_cast<const uint8_t*>(h->Xs);
    ra.kdim = h->kdim;
    ra.tiled = h->tiled;
    ra.scale = h->sq_scale;
    ra.offset = h->sq_offset;
    // the r rows of the cut are collected whole: the shared list holds
    // rb_F * kRobustCap ids, a larger r takes a list of its own
    ra.cap = (int64_t)h->rb_F * kRobustCap;
    if (r > ra.cap) {
        if (h->gf_cap < r) {
            dfree(h->gf_cid);
            h->gf_cid = nullptr;
            h->gf_cap = 0;
            HIP_OR(hipMalloc(&h->gf_cid, (size_t)r * 4), -1);
            h->gf_cap = r;
        }
        ra.cid = h->gf_cid;
        ra.cap = h->gf_cap;
    }
    const int32_t row0 = 0;
    uint32_t flag = 0;
    HIP_OR(hipMemcpyAsync(h->dQ32, q, (size_t)h->d * 4, hipMemcpyHostToDevice, st), -1);
    HIP_OR(hipMemcpyAsync(h->rb_rows, &row0, 4, hipMemcpyHostToDevice, st), -1);
    HIP_OR(launch_robust(ra, st), -1);
    static_assert(sizeof(long long) == sizeof(int64_t), "");
    HIP_OR(hipMemcpyAsync(out_ids, h->d_ids, (size_t)out_k * 8, hipMemcpyDeviceToHost, st), -1);
    HIP_OR(hipMemcpyAsync(&flag, h->rb_flags, 4, hipMemcpyDeviceToHost, st), -1);
    HIP_OR(hipStreamSynchronize(st), -1);
    if (flag) {
        set_err("gf_query: collected more than %lld rows", (long long)ra.cap);
        return -1;
    }
    h->ndist += r;  // gpu_filter.cu:171
    return out_k;
}

long long gf_ndist(void* handle) { return handle ? (long long)static_cast<Handle*>(handle)->ndist.load() : 0; }

void gf_free(void* handle) { free_handle(static_cast<Handle*>(handle)); }

const char* gf_last_error(void) { return g_err; }

}  // extern "C"

/*
 * annmi -- MI355X (gfx950) k-NN search library: exact FP32 scan, BF16/FP8/SQ8/SQ4
 * candidate scans with FP32 re-score, each followed by top-k selection.
 *
 * Plain C ABI: pointers and sizes only.  One handle = one device-resident
 * index.  Results are row-major [B x k], ascending (distance, id); distances
 * are squared L2 in fp32.  Calls on one handle are serialised internally.
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef ANNMI_H
#define ANNMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* index modes (candidate-generation element type) */
enum {
    ANNMI_EXACT_F32 = 0, /* plain fp32 exact scan (src/linear_index.cpp)                   */
    ANNMI_BF16 = 1,      /* bf16 MFMA candidates + fp32 re-score of R                       */
    ANNMI_FP8 = 2,       /* e4m3 MFMA candidates + fp32 re-score of R                       */
    ANNMI_SQ8 = 3,       /* per-dim SQ8 codes, int8 MFMA + fp32 re-score (gpu_filter.cu)    */
    ANNMI_SQ4 = 4,       /* per-dim SQ4 nibbles, int8 MFMA + fp32 re-score                  */
    ANNMI_IVF_SQ8 = 5    /* k-means IVF, per-dim SQ8 codes per list, list scan + fp32 re-score
                            (competitors/ivf_lsh/src/lsh_index_optimized.cpp)               */
};

typedef struct annmi_params {
    int32_t mode;           /* ANNMI_* */
    int32_t rerank;         /* default over-fetch R for the quantized modes (0 -> 4k)         */
    int32_t sq_variant;     /* SQ8 encoder: 0 = lsh_index_optimized.cpp:186-214 (std::round,
                               constant dim offset 0), 1 = gpu_filter.cu:99-127 (lrintf,
                               constant dim offset min)                                          */
    int32_t base_on_device; /* 1: `base` is a device pointer, borrowed for the handle's life    */
    int64_t max_batch;      /* queries per internal step (0 -> 4096)                           */
    int32_t nlist;          /* IVF lists (0 -> 256, lsh_index_optimized.cpp:128 default)       */
    int32_t nprobe;         /* IVF default lists probed per query (0 -> 8, :304 default)       */
    int32_t kmeans_iters;   /* IVF k-means rounds (0 -> 15, :564)                              */
    int32_t kmeans_seed;    /* IVF initial-centroid shuffle seed (0 -> 1337, :555)             */
    int64_t reserved[2];
} annmi_params;

/* Fills `p` with the defaults for `mode`. */
void annmi_default_params(annmi_params* p, int32_t mode);

/* Build an index over n x d row-major float32 `base`.
 * Replaces: LinearIndex::fit        src/linear_index.cpp:73-88
 *           gf_create               competitors/gpu_filter/src/gpu_filter.cu:79-140
 * Returns NULL on error (annmi_last_error()).  A host `base` is copied; a
 * device `base` (params->base_on_device) is borrowed when d % 16 == 0.
 * d is limited by the LDS a 16-query tile of the scan takes (160 KB): fp32
 * (exact mode, and IVF, whose k-means and probe scan fp32 centroids) up to
 * d = 2544; the other modes up to 4096 (the library's maximum).  Larger d
 * fails here with an error naming the limit. */
void* annmi_create(const float* base, int64_t n, int32_t d, const annmi_params* params);

/* One query (host pointers).  Writes min(k, n) ids and distances.
 * Replaces: LinearIndex::query      src/linear_index.cpp:90-121
 *           gf_query                gpu_filter.cu:142-176 (rerank = r)
 * rerank <= 0 uses the handle default.  Returns the count written or -1. */
int32_t annmi_query(void* h, const float* q, int32_t k, int32_t rerank, int64_t* out_ids, float* out_dists);

/* B queries, host pointers, outputs [B x k].  No reference counterpart: the
 * batched form of annmi_query (the reference answers one query per call,
 * orthogonal-competition/harness.py:159-174).  Returns min(k, n) or -1. */
int32_t annmi_batch_query(void* h, const float* Q, int32_t B, int32_t k, int32_t rerank, int64_t* out_ids,
                          float* out_dists);

/* B queries with an explicit IVF probe count (nprobe <= 0: handle default;
 * ignored by the flat modes).  Replaces IVFLSHIndex::query
 * lsh_index_optimized.cpp:304 (n_probe_clusters, refine_r).  on_device != 0:
 * Q, ids and dists are device pointers. */
int32_t annmi_batch_query_ex(void* h, const float* Q, int32_t B, int32_t k, int32_t rerank, int32_t nprobe,
                             int64_t* out_ids, float* out_dists, int32_t on_device);

/* The reference's exact scan reproduced bit for bit, B host queries one at a
 * time: ids AND the fp32 values it ranks, including its early-abandon
 * partial sums beyond rank 64 (64-slot threshold, checkpoints every 96
 * dims).  Replaces LinearIndex::query src/linear_index.cpp:90-121 with
 * linear_worker :27-60 and compute_l2_distance src/simd.h:32-64 (AVX-512
 * branch).  Needs an ANNMI_EXACT_F32 handle.  Returns min(k, n) or -1. */
int32_t annmi_query_linear_index(void* h, const float* Q, int32_t B, int32_t k, int64_t* out_ids, float* out_values);

/* Same with device pointers (inputs must be ready when called); synchronous. */
int32_t annmi_batch_query_device(void* h, const float* dQ, int32_t B, int32_t k, int32_t rerank, int64_t* d_ids,
                                 float* d_dists);

/* Same as annmi_batch_query_device (device pointers, synchronous) but
 * allocates nothing: no hipMalloc / hipFree and no allocating library
 * routine, on the device or the host.  No counterpart in the reference: it is
 * for callers that cannot allocate once serving has started (tight serving
 * loops, harnesses that forbid allocation after set-up).  The step's
 * workspace is a static device array of 40 MiB held by every device the
 * library runs on until the process ends; each call uses it whole, so calls
 * are serialised across the process (calls that overlapped on one device
 * would race on it).  Limits: flat modes only (exact, bf16, fp8, sq8, sq4;
 * not IVF), B <= 256 queries per call, k <= 2048, candidate lists of 16384
 * keys per query.  A query the regular step flags (annmi_last_batch_stats
 * bits 0-2) would need the exact all-rows path, which allocates: the call
 * then returns -1 with an error naming the query.  Returns min(k, n) or -1.
 * annmi_last_batch_stats reports nothing for these calls. */
int32_t annmi_batch_query_noalloc(void* h, const float* dQ, int32_t B, int32_t k, int32_t rerank, int64_t* d_ids,
                                  float* d_dists);

/* Full-precision distance evaluations so far.
 * Replaces: LinearIndex::total_distances_count  src/linear_index.cpp:123 (N per query)
 *           gf_ndist                             gpu_filter.cu:178 (R per query) */
int64_t annmi_ndist(void* h);

/* Per-query diagnostics of the last batch: rows re-scored and flags
 * (bit0 candidate-list overflow, bit1 more than 2048 rows within the error
 * bound of the rank-R key, bit2 an fp32 near-tie run at rank k longer than the
 * fp64 window; bit3 the query was answered again by the exact all-rows path,
 * which any of bits 0-2 triggers).  Returns count or -1. */
int32_t annmi_last_batch_stats(void* h, uint32_t* n_rescored, uint32_t* flags, int32_t max_q);

/* Release the index.  NULL-safe.  Replaces gf_free gpu_filter.cu:180-190. */
void annmi_free(void* h);

/* A second handle on the same device-resident index (no copy) with its own
 * HIP stream and workspace, so batches on different handles run
 * concurrently on the device.  The source must outlive its clones.  No
 * reference counterpart (the reference serves one query at a time). */
void* annmi_clone(void* h);

/* Thread-local message of the last failure on this thread.
 * Replaces gf_last_error gpu_filter.cu:77 (process-global there). */
const char* annmi_last_error(void);

/* Bytes of device memory held by the index (base copies, codes, norms). */
int64_t annmi_index_bytes(void* h);

/* Copy index state to host memory (bytes >= the returned size), e.g. to
 * check encoders or persist an index.  No reference counterpart (the
 * reference keeps its codes private and has no persistence, README.md:12-13).
 *   what 0: scan rows (n x row-stride bytes: fp32 / bf16 / e4m3 / SQ8 codes /
 *           SQ4 nibbles in the submission/src/hnsw.cpp:170-182 layout)
 *        1: SQ per-dim scale (kdim floats)   2: SQ per-dim offset (kdim floats)
 *        3: per-row norm term (n floats)     4: 4 doubles {kdim, row stride
 *           bytes, fp8 scale, max row norm}
 *        5: IVF list position -> original row id (n uint32)
 *        6: IVF list offsets (nlist + 1 int64)
 *        7: IVF centroids (nlist x ceil(d/16)*16 floats)
 * With out == NULL or bytes too small, returns the size without copying;
 * -1 on error. */
int64_t annmi_export(void* h, int32_t what, void* out, int64_t bytes);

/* Per-kernel timing with HIP events on the handle's own stream, recorded by
 * the timed kernel dispatches themselves (start/stop events of
 * hipExtLaunchKernelGGL; the IVF step's prep span with stream events).  No
 * reference counterpart (the reference's stage timers are
 * lsh_index_optimized.cpp:535-548 / experiments/cuda/gpu_filter.cu:144-162).
 * Copies the accumulated counters into out8 (when not NULL):
 *   [0] scan launches  [1] scan ms  [2] scan algorithmic bytes  [3] scan
 *   algorithmic flops (2*B*rows*d)  [4] threshold-select ms  [5] finalize ms
 *   [6] query-prep ms  [7] steps
 * then, when enable >= 0, switches timing on/off and zeroes the counters. */
int32_t annmi_profile(void* h, int32_t enable, double* out8);

/* The same counters split by scan kernel instantiation, as a JSON array of
 * {kernel, launches, ms, bytes, flops, launch_ms[], launch_bytes[],
 * launch_flops[]} (the per-launch lists hold the first 4096 launches;
 * bytes and flops are algorithmic) written to buf (len bytes).  Returns the
 * length needed including the terminator, -1 on error. */
int32_t annmi_profile_kernels(void* h, char* buf, int32_t len);

/* Save an index: writes a self-describing byte image (header, SQ scale and
 * offset, scan rows -- SQ4 nibbles in the submission/src/hnsw.cpp:170-182
 * layout --, row norms, and for IVF the list permutation, list offsets and
 * centroids) to buf when bytes is large enough.  The fp32 base is not part
 * of it.  Returns the size needed, -1 on error.  No reference counterpart:
 * the reference rebuilds every index in fit (README.md:12-13). */
int64_t annmi_save(void* h, void* buf, int64_t bytes);

/* Recreate a saved index over the same n x d float32 base (host or device,
 * as in annmi_create) without the IVF k-means build or the encoders.  params
 * (NULL: defaults for the saved mode) must name the saved mode; the saved
 * encoder variant and list count are used.  Returns NULL (annmi_last_error)
 * when the image is truncated, corrupt, of another mode or shape, or of a
 * layout this library does not use.  No reference counterpart. */
void* annmi_load(const float* base, int64_t n, int32_t d, const annmi_params* params, const void* saved,
                 int64_t bytes);

/* Autotuner.  For batch sizes 1, 2, 4, ... up to max_batch (<= 4096), times
 * whole query steps (k = 100) with every candidate scan-kernel shape on
 * queries drawn from the index's own rows: the streaming kernel (up to 128
 * queries), each wide-kernel tile configuration that fits the LDS, the
 * built-in one at other row-chunk sizes, then the best one at other chunk
 * sizes.  A shape replaces the built-in choice only when at least 2% faster,
 * confirmed by a second timing of both (eager steps, no graph capture while
 * tuning); the result per batch size is
 * kept (a batch uses the entry of the nearest power of two) until the next
 * annmi_autotune or annmi_set_scan_shape.  Shapes change timing only: every
 * shape computes the same candidate keys, so answers are unchanged.  reps:
 * timed runs per candidate (best kept).  Writes a JSON report (one object
 * per batch size: kernel, cfg, chunk, ms, default_ms) to report (len bytes);
 * max_batch = 0 only writes the last report.  Returns the length needed
 * including the terminator, -1 on error (IVF handles: nothing to tune).  No
 * reference counterpart. */
int32_t annmi_autotune(void* h, int32_t max_batch, int32_t reps, char* report, int32_t len);

/* Fix the scan-kernel shape for every batch size (drops an autotuned
 * table): batches of wide_min_batch queries and more take the wide kernel
 * (0: built-in crossover), with tile configuration wide_cfg (0: built-in
 * choice; 1: 128 queries x 64 rows a wave on 4 waves; 2: 128 x 32 on 8
 * waves; 3: 64 x 64 on 8 waves; SQ8 and BF16 only, 32 x 32 MFMA tiles:
 * 4: 128 x 32 on 8 waves, 5: 128 x 64 on 8 waves, 6 and 7: SQ8 128 x 32 and
 * 128 x 64 on 8 waves with a compact epilogue, BF16 128 x 128 and 128 x 64
 * on 4 waves; 8: SQ8 with d padded to 128 or 256 only, 512 queries in
 * registers over rows staged in LDS) and at most chunk_rows
 * rows per row chunk
 * (0: built-in, else >= 2048).  Returns 0, or -1 for a configuration that does
 * not fit the LDS at this dim.  No reference counterpart (tuning, tests). */
int32_t annmi_set_scan_shape(void* h, int32_t wide_min_batch, int32_t wide_cfg, int64_t chunk_rows);

/* Number of step shapes (batch size, k, re-score count, IVF probe count)
 * whose query step this handle replays as a captured HIP graph.  A step --
 * flat or IVF: the IVF probe, work plan and list scan are all enqueued on
 * the device with no host round trip -- runs eagerly the first time its
 * shape is seen, is captured the second time and replayed from then on,
 * into whatever output buffers the call passes (the replay rewrites the
 * captured finalize kernel's output arguments).  ANNMI_GRAPHS=0 disables
 * this; steps are never captured while annmi_profile timing is on, nor on
 * cloned handles; growing the workspace (a larger batch than any before)
 * drops the captured steps.  No reference counterpart (diagnostics).  -1 on
 * a NULL handle. */
int32_t annmi_graph_count(void* h);

/* ------------------------------------------------------------------
 * Drop-in entry points with the exact signatures of the reference's GPU
 * contestant, so competitors/gpu_filter/algorithm.py:6-15 binds this
 * library unchanged: per-dim SQ8 (gpu_filter encoder), quantized scan,
 * top-r, fp32 re-score of r on the GPU, top-k.
 * ------------------------------------------------------------------ */
void* gf_create(const float* train, long n, long dim);                      /* gpu_filter.cu:79  */
int gf_query(void* handle, const float* q, int k, int r, long long* out_ids); /* gpu_filter.cu:142 */
long long gf_ndist(void* handle);                                           /* gpu_filter.cu:178 */
void gf_free(void* handle);                                                 /* gpu_filter.cu:180 */
const char* gf_last_error(void);                                            /* gpu_filter.cu:77  */

#ifdef __cplusplus
}
#endif
#endif /* ANNMI_H */

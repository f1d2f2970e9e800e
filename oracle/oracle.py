"""CPU oracle for the k-NN search path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only to check or time; the product
(``annvedi_amd``) never imports it and has no CPU fallback.

Thin numpy/ctypes layer over ``oracle/liboracle.so`` (built from
``oracle/oracle.cpp`` by ``oracle/Makefile``) plus numpy restatements of the
evaluator's recall.  Reference file:line for each function is in the C++
header and repeated below with what pins it.

Parity pins (each a check a plausible mistake would fail):

* ``linear_index_query`` (src/linear_index.cpp:27-121, src/simd.h:32-64):
  golden ids/values written by the reference module compiled in this
  container (tests/golden/linear_ref_*.npz), bit-exact; ranks 1-64 also equal
  the FP64 exact oracle away from FP32 near-ties (tests/test_oracle.py).
* ``l2_linear_index``: bit-exact against the compiled reference on
  checkpoint-sensitive thresholds through the golden vectors; value equals
  the plain float sum within 1e-5 relative; early-abandon returns a partial
  that is >= threshold and <= the full value (invariants).
* ``exact_topk_f64``: brute force in numpy float64 on tiny inputs, duplicate
  rows (ties broken by lower id), k > n clamp, self-query at distance 0.
* ``sq8_*`` (lsh_index_optimized.cpp:186-214, gpu_filter.cu:99-127): closed
  forms -- rounding mode (half away vs half even) on planted .5 cases, the
  constant-dimension conventions, clamp to [0, 255]; distance equals the
  float64 dequantised distance within 1e-5 relative.
* ``sq4_*`` (submission/src/hnsw.cpp:147-183, submission/src/simd.h:292-318):
  nibble layout against a hand-built 32-dim block, padding inert, distance vs
  float64 dequantised distance.
* ``recall`` (orthogonal-competition/evaluator.py:149-157): hand-computed
  cases including ties at the threshold.
* ``kmeans_init``, ``ivf_kmeans``, ``ivf_reference_query``
  (lsh_index_optimized.cpp:304-516, 551-616): the compiled reference's IVF
  answers on integer-coordinate data, where every float sum and distance is
  exact and so the whole chain -- the mt19937 shuffle's initial rows, 15
  rounds, the probe order, the fall-back to all rows, the fp32 ranking -- must
  match bit for bit (tests/golden/ivf_ref.npz); and on separated clusters
  seeded one per cluster, where only the partition must match.
* ``exact_topk_subset_f64``: brute force over the subset in numpy float64.
* ``f32_direct_l2`` (gpu_filter.cu:161-170): integer-valued rows (every sum
  exact), a sum that only the sequential order absorbs, a product that FMA
  contraction would not round.
* ``HnswCpu`` (submission/src/hnsw.cpp, oracle/hnsw_cpu.cpp; the CPU HNSW
  baseline, not a checker of the GPU path): with a beam as wide as the base
  the answer is the brute-force top-k over the nodes reachable on layer 0;
  renumbering in BFS order changes no answer; its recall equals the compiled
  reference's within 0.02 on the same data and parameters
  (tests/test_hnsw_cpu.py).
* ``gf_keys``, ``gf_query_reference`` (gpu_filter.cu:52-63, 142-176): the
  whole chain brute-forced in integer arithmetic on rows whose codes equal
  their values (every key and re-rank exact), duplicates for the stable
  order of the cut and the (value, id) order of the answer; a fused
  multiply-add a separate product would round differently; a sequential
  chain a tree would not absorb.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64
_i32 = ctypes.c_int

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_l2_linear_index.restype = ctypes.c_float
        L.orc_l2_linear_index.argtypes = [_f32p, _f32p, _i64, ctypes.c_float]
        L.orc_l2_linear_index_checkpoints.restype = _i32
        L.orc_l2_linear_index_checkpoints.argtypes = [_f32p, _f32p, _i64, _f32p]
        L.orc_linear_index_query.restype = _i32
        L.orc_linear_index_query.argtypes = [_f32p, _i64, _i64, _f32p, _i32, _i64p, _f32p]
        L.orc_linear_index_batch.restype = _i32
        L.orc_linear_index_batch.argtypes = [_f32p, _i64, _i64, _f32p, _i32, _i32, _i64p, _f32p, _i32]
        L.orc_l2_f64.restype = ctypes.c_double
        L.orc_l2_f64.argtypes = [_f32p, _f32p, _i64]
        L.orc_exact_topk_f64.restype = _i32
        L.orc_exact_topk_f64.argtypes = [_f32p, _i64, _i64, _f32p, _i32, _i32, _i64p, _f64p, _i32]
        L.orc_exact_topk_subset_f64.restype = _i32
        L.orc_exact_topk_subset_f64.argtypes = [_f32p, _i64, _i64p, _i64, _f32p, _i32, _i64p, _f64p]
        L.orc_f32_direct_l2.restype = ctypes.c_float
        L.orc_f32_direct_l2.argtypes = [_f32p, _f32p, _i64]
        L.orc_sq8_train.restype = None
        L.orc_sq8_train.argtypes = [_f32p, _i64, _i64, _i32, _f32p, _f32p]
        L.orc_sq8_encode.restype = None
        L.orc_sq8_encode.argtypes = [_f32p, _i64, _i64, _i32, _f32p, _f32p, _u8p]
        L.orc_sq8_l2.restype = ctypes.c_float
        L.orc_sq8_l2.argtypes = [_u8p, _f32p, _f32p, _i64]
        L.orc_sq8_l2_f64.restype = ctypes.c_double
        L.orc_sq8_l2_f64.argtypes = [_u8p, _f32p, _f32p, _i64]
        L.orc_sq4_train.restype = _i64
        L.orc_sq4_train.argtypes = [_f32p, _i64, _i64, _f32p, _f32p]
        L.orc_sq4_encode.restype = None
        L.orc_sq4_encode.argtypes = [_f32p, _i64, _i64, _f32p, _f32p, _u8p]
        L.orc_sq4_l2.restype = ctypes.c_float
        L.orc_sq4_l2.argtypes = [_u8p, _f32p, _f32p, _i64]
        L.orc_kmeans_init.restype = None
        L.orc_kmeans_init.argtypes = [_i64, _i32, ctypes.c_uint32, _i64p, _i64p]
        L.orc_gf_keys.restype = None
        L.orc_gf_keys.argtypes = [_u8p, _i64, _i64, _f32p, _f32p, _f32p]
        L.orc_ivf_kmeans.restype = None
        L.orc_ivf_kmeans.argtypes = [_f32p, _i64, _i64, _i32, _i32, ctypes.c_uint32, _f32p,
                                     np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")]
        L.orc_sq4_nibble.restype = _i32
        L.orc_sq4_nibble.argtypes = [_u8p, _i64]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# --------------------------------------------------------------- linear ----
def l2_linear_index(a, b, threshold=np.finfo(np.float32).max) -> np.float32:
    a, b = _f32(a), _f32(b)
    return np.float32(lib().orc_l2_linear_index(a, b, a.shape[0], float(threshold)))


def l2_linear_index_checkpoints(a, b) -> np.ndarray:
    a, b = _f32(a), _f32(b)
    out = np.zeros(a.shape[0] // 96 + 2, dtype=np.float32)
    n = lib().orc_l2_linear_index_checkpoints(a, b, a.shape[0], out)
    return out[: n + 1]


def linear_index_query(base, q, k):
    base, q = _f32(base), _f32(q)
    n, d = base.shape
    ids = np.empty(k, dtype=np.int64)
    vals = np.empty(k, dtype=np.float32)
    got = lib().orc_linear_index_query(base, n, d, q, int(k), ids, vals)
    return ids[:got], vals[:got]


def linear_index_batch(base, Q, k, nthreads=1):
    base, Q = _f32(base), _f32(np.atleast_2d(Q))
    n, d = base.shape
    nq = Q.shape[0]
    ids = np.empty((nq, k), dtype=np.int64)
    vals = np.empty((nq, k), dtype=np.float32)
    got = lib().orc_linear_index_batch(base, n, d, Q, nq, int(k), ids, vals, int(nthreads))
    return ids[:, :got], vals[:, :got]


# ---------------------------------------------------------------- exact ----
def l2_f64(a, b) -> float:
    a, b = _f32(a), _f32(b)
    return float(lib().orc_l2_f64(a, b, a.shape[0]))


def exact_topk_f64(base, Q, k, nthreads=None):
    """(ids int64 [nq,k'], d64 float64 [nq,k']) ranked by (d64, id); k' = min(k, n)."""
    base, Q = _f32(base), _f32(np.atleast_2d(Q))
    n, d = base.shape
    nq = Q.shape[0]
    if nthreads is None:
        nthreads = max(1, min(8, len(os.sched_getaffinity(0))))
    ids = np.empty((nq, k), dtype=np.int64)
    dd = np.empty((nq, k), dtype=np.float64)
    got = lib().orc_exact_topk_f64(base, n, d, Q, nq, int(k), ids, dd, int(nthreads))
    return ids[:, :got], dd[:, :got]


def exact_topk_subset_f64(base, cand, q, k):
    base, q = _f32(base), _f32(q)
    cand = np.ascontiguousarray(cand, dtype=np.int64)
    ids = np.empty(k, dtype=np.int64)
    dd = np.empty(k, dtype=np.float64)
    got = lib().orc_exact_topk_subset_f64(base, base.shape[1], cand, cand.shape[0], q, int(k), ids, dd)
    return ids[:got], dd[:got]


def f32_direct_l2(a, b) -> np.float32:
    a, b = _f32(a), _f32(b)
    return np.float32(lib().orc_f32_direct_l2(a, b, a.shape[0]))


# ------------------------------------------------------------------ SQ8 ----
def sq8_train(x, variant=0):
    x = _f32(x)
    n, d = x.shape
    s = np.empty(d, dtype=np.float32)
    o = np.empty(d, dtype=np.float32)
    lib().orc_sq8_train(x, n, d, int(variant), s, o)
    return s, o


def sq8_encode(x, scale, offset, variant=0):
    x = _f32(x)
    n, d = x.shape
    codes = np.empty((n, d), dtype=np.uint8)
    lib().orc_sq8_encode(x, n, d, int(variant), _f32(scale), _f32(offset), codes)
    return codes


def sq8_l2(code_row, q_shifted, scale) -> np.float32:
    c = np.ascontiguousarray(code_row, dtype=np.uint8)
    return np.float32(lib().orc_sq8_l2(c, _f32(q_shifted), _f32(scale), c.shape[0]))


def sq8_l2_f64(code_row, q_shifted, scale) -> float:
    c = np.ascontiguousarray(code_row, dtype=np.uint8)
    return float(lib().orc_sq8_l2_f64(c, _f32(q_shifted), _f32(scale), c.shape[0]))


def gf_keys(codes, q_shift, scale):
    """The reference GPU contestant's SQ8 key of every row (gpu_filter.cu:52-63)."""
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    n, d = codes.shape
    out = np.empty(n, dtype=np.float32)
    lib().orc_gf_keys(codes, n, d, _f32(q_shift), _f32(scale), out)
    return out


def gf_query_reference(base, q, k, r):
    """competitors/gpu_filter/src/gpu_filter.cu:79-176 for one query: the
    per-dim SQ8 encoder of gf_create (:99-127, lrintf, constant dim offset
    min); r clamped to [k, n] (:145-146); keys of every row (:52-63); the
    first r rows of the stable ascending sort of (key, row) (CUB SortPairs
    over idx = 0..n-1, :154-155); each re-ranked by the scalar float sum
    acc += v * v, v = row - q (:161-170); partial_sort by (value, row) and
    the first min(k, r) rows (:172-175).  Returns (ids int64, values float32)."""
    base, q = _f32(base), _f32(q)
    n = base.shape[0]
    r = max(int(r), int(k))
    r = min(r, n)
    scale, offset = sq8_train(base, 1)
    codes = sq8_encode(base, scale, offset, 1)
    keys = gf_keys(codes, (q - offset).astype(np.float32), scale)
    cand = np.argsort(keys, kind="stable")[:r]
    vals = np.array([f32_direct_l2(base[c], q) for c in cand], dtype=np.float32)
    order = np.lexsort((cand, vals))[:min(int(k), r)]
    return cand[order].astype(np.int64), vals[order]


# ------------------------------------------------------------------ SQ4 ----
def sq4_train(x):
    x = _f32(x)
    n, d = x.shape
    dp = (d + 31) // 32 * 32
    s = np.empty(dp, dtype=np.float32)
    o = np.empty(dp, dtype=np.float32)
    lib().orc_sq4_train(x, n, d, s, o)
    return s, o


def sq4_encode(x, scale_pad, offset_pad):
    x = _f32(x)
    n, d = x.shape
    dp = scale_pad.shape[0]
    codes = np.empty((n, dp // 2), dtype=np.uint8)
    lib().orc_sq4_encode(x, n, d, _f32(scale_pad), _f32(offset_pad), codes)
    return codes


def sq4_l2(code_row, q_shifted_pad, scale_pad) -> np.float32:
    c = np.ascontiguousarray(code_row, dtype=np.uint8)
    return np.float32(lib().orc_sq4_l2(c, _f32(q_shifted_pad), _f32(scale_pad), c.shape[0] * 2))


def sq4_unpack(codes, dim):
    """Nibbles as an n x dim uint8 array (inverse of the packing layout)."""
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    n, stride = codes.shape
    out = np.empty((n, stride * 2), dtype=np.uint8)
    blocks = codes.reshape(n, -1, 16)
    out.reshape(n, -1, 32)[:, :, :16] = blocks & 0x0F
    out.reshape(n, -1, 32)[:, :, 16:] = blocks >> 4
    return out[:, :dim]


# ------------------------------------------------------------------ IVF ----
def kmeans_init(n, nlist, seed=1337):
    """(init rows [nlist], rand_indices [n]) per lsh_index_optimized.cpp:553-561."""
    init = np.empty(nlist, dtype=np.int64)
    ri = np.empty(n, dtype=np.int64)
    lib().orc_kmeans_init(int(n), int(nlist), int(seed), init, ri)
    return init, ri


def ivf_kmeans(x, nlist, iters=15, seed=1337):
    """(centroids float32 [nlist', d], assign int32 [n]) of the reference's
    k-means (lsh_index_optimized.cpp:551-616) run on one thread; nlist' =
    min(nlist, n)."""
    x = _f32(x)
    n, d = x.shape
    m = min(int(nlist), n)
    cent = np.empty((m, d), dtype=np.float32)
    assign = np.empty(n, dtype=np.int32)
    lib().orc_ivf_kmeans(x, n, d, m, int(iters), int(seed), cent, assign)
    return cent, assign


def ivf_lists(assign, nlist):
    """(perm, list offsets): rows grouped by cluster, ascending id inside a
    cluster (the reordered layout of lsh_index_optimized.cpp:155-173)."""
    assign = np.asarray(assign, dtype=np.int64)
    perm = np.argsort(assign, kind="stable").astype(np.int64)
    off = np.zeros(nlist + 1, dtype=np.int64)
    np.cumsum(np.bincount(assign, minlength=nlist), out=off[1:])
    return perm, off


def ivf_reference_query(base, centroids, perm, list_off, Q, k, nprobe):
    """lsh_index_optimized.cpp:304-516 with n_tables 1, n_bits 0 (one bucket
    per list) and refine_r at least the candidate count, so that every
    candidate is refined:
      probe  the nprobe smallest (fp32 compute_l2_distance(q, centroid), list)
             pairs (:325-333);
      gather the probed lists' rows; every row when they hold fewer than k
             (:452-457);
      rank   candidates by fp32 compute_l2_distance(row, q) (:485-496).
    The reference's last two sorts compare distances only, so the order of
    exactly equal fp32 distances is unpinned; here they go by id.  Returns
    int64 ids [nq, k], padded with -1 past the candidates (:506-511)."""
    base = _f32(base)
    centroids = _f32(centroids)
    Q = _f32(np.atleast_2d(Q))
    n = base.shape[0]
    out = np.full((Q.shape[0], k), -1, dtype=np.int64)
    for qi, q in enumerate(Q):
        cd = np.array([l2_linear_index(q, c) for c in centroids], dtype=np.float32)
        probes = np.lexsort((np.arange(len(cd)), cd))[:min(nprobe, len(cd))]
        rows = np.concatenate([perm[list_off[l]:list_off[l + 1]] for l in probes]).astype(np.int64)
        if rows.shape[0] < k:
            rows = np.arange(n, dtype=np.int64)
        dd = np.array([l2_linear_index(base[r], q) for r in rows], dtype=np.float32)
        order = np.lexsort((rows, dd))[:k]
        out[qi, :len(order)] = rows[order]
    return out


def ivf_probe_exact(centroids, Q, nprobe):
    """nprobe nearest centroids by (fp64 distance, id)."""
    ids, _ = exact_topk_f64(centroids, Q, nprobe)
    return ids


def ivf_search_exact(base, perm, list_off, centroids, Q, k, nprobe):
    """Exact top-k (fp64, lower id on ties) over the union of the nprobe
    nearest lists: the result an IVF list scan with a complete re-score must
    return (lsh_index_optimized.cpp:325-333 probe, full-list candidates).
    When the probed lists hold fewer than k rows every row is a candidate
    (lsh_index_optimized.cpp:452-457)."""
    probes = ivf_probe_exact(centroids, Q, nprobe)
    out_ids, out_d = [], []
    for qi, q in enumerate(np.atleast_2d(Q)):
        rows = np.concatenate([perm[list_off[l]:list_off[l + 1]] for l in probes[qi]]).astype(np.int64)
        if rows.shape[0] < k:
            rows = np.arange(base.shape[0], dtype=np.int64)
        i, dd = exact_topk_subset_f64(base, rows, q, k)
        out_ids.append(i)
        out_d.append(dd)
    return out_ids, out_d


# ---------------------------------------------------------------- recall ---
def recall(true_dist, pred_dist, k):
    """orthogonal-competition/evaluator.py:149-157: mean(pred[:k] <= true[k-1])."""
    true_dist = np.asarray(true_dist)
    pred_dist = np.asarray(pred_dist)
    return np.array([np.mean(pred_dist[i][:k] <= true_dist[i][k - 1]) for i in range(true_dist.shape[0])])


def euclid_f32(base, Q, ids):
    """evaluator.py:455-460: float32 np.linalg.norm of (train[ids] - q)."""
    base = np.asarray(base, dtype=np.float32)
    Q = np.atleast_2d(np.asarray(Q, dtype=np.float32))
    return np.stack([np.linalg.norm(base[ids[i]] - Q[i], axis=1) for i in range(Q.shape[0])])


def recall_at_k(base, Q, pred_ids, true_ids, k):
    """Evaluator-identical recall given predicted and ground-truth id lists."""
    return recall(euclid_f32(base, Q, true_ids), euclid_f32(base, Q, pred_ids), k)


# ------------------------------------------------------------ reference ----
def ref_module(name):
    """Import a reference module built by oracle/ref.mk into oracle/_ref, or None."""
    import importlib
    import sys

    if not os.path.isdir(REF_DIR):
        return None
    if REF_DIR not in sys.path:
        sys.path.insert(0, REF_DIR)
    try:
        return importlib.import_module(name)
    except ImportError:
        return None


class HnswCpu:
    """The CPU HNSW baseline (oracle/hnsw_cpu.cpp, built into
    oracle/libhnsw_cpu.so): submission/src/hnsw.cpp restated with 64-bit
    offsets.  metric: "float", "sq8" (one global scale) or "sq8pd"."""

    _L = None

    @classmethod
    def _lib(cls):
        if cls._L is None:
            path = os.path.join(HERE, "libhnsw_cpu.so")
            if not os.path.exists(path):
                build()
            L = ctypes.CDLL(path)
            L.hnsw_cpu_build.restype = ctypes.c_void_p
            L.hnsw_cpu_build.argtypes = [_f32p, _i64, _i32, _i32, _i32, _i32, _i32, _i32, _i32, ctypes.c_uint64,
                                         _i64]
            L.hnsw_cpu_query.restype = _i32
            L.hnsw_cpu_query.argtypes = [ctypes.c_void_p, _f32p, _i32, _i32, _i64p]
            L.hnsw_cpu_query_batch.restype = _i32
            L.hnsw_cpu_query_batch.argtypes = [ctypes.c_void_p, _f32p, _i64, _i32, _i32, _i64p, _i32]
            L.hnsw_cpu_ndist.restype = _i64
            L.hnsw_cpu_ndist.argtypes = [ctypes.c_void_p]
            L.hnsw_cpu_free.restype = None
            L.hnsw_cpu_free.argtypes = [ctypes.c_void_p]
            cls._L = L
        return cls._L

    def __init__(self, x, M=16, ef_construction=100, metric="float", heuristic=False, reorder=False, threads=0,
                 seed=42, progress=0):
        L = self._lib()
        x = _f32(x)
        if x.ndim != 2:
            raise RuntimeError("Input must be a non-empty 2D array")
        mcode = {"float": 0, "sq8": 1, "sq8pd": 2}[metric]
        self.d = x.shape[1]
        self._h = L.hnsw_cpu_build(x, x.shape[0], x.shape[1], int(M), int(ef_construction), mcode, int(heuristic),
                                   int(reorder), int(threads), int(seed), int(progress))
        if not self._h:
            raise RuntimeError("hnsw_cpu_build refused its arguments")

    def query(self, q, k, ef=100):
        out = np.empty(int(k), dtype=np.int64)
        m = self._lib().hnsw_cpu_query(self._h, _f32(q).ravel(), int(k), int(ef), out)
        return out[:m]

    def query_batch(self, Q, k, ef=100, threads=1):
        Q = _f32(Q)
        out = np.empty((Q.shape[0], int(k)), dtype=np.int64)
        self._lib().hnsw_cpu_query_batch(self._h, Q, Q.shape[0], int(k), int(ef), out, int(threads))
        return out

    @property
    def n_distances(self):
        return int(self._lib().hnsw_cpu_ndist(self._h))

    def close(self):
        if getattr(self, "_h", None):
            self._lib().hnsw_cpu_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

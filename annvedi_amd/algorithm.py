"""The competition plugin surface on the MI355X search library.

Mirrors the reference's ``Algorithm`` facade (algorithm.py:12-90) and its
single-backend bundles (competitors/linear/algorithm.py:10-26,
competitors/gpu_filter/algorithm.py:30-59): ``fit(train, **index_params)``,
``query(q, k, **query_params) -> int64[k]`` and ``get_n_distances()``, plus
``batch_query(Q, k, **query_params) -> (ids int64[B, k], dists float32[B, k])``.

Selected in scenarios.yaml with ``backend: mi355x`` and ``mode`` in
{exact, linear_index, bf16, fp8, sq8, sq4, ivf}; ``rerank`` (aliases ``r`` as in
competitors/gpu_filter/scenarios.yaml, ``refine_r`` as in the ivf_lsh fork)
is the over-fetch for the quantized modes; IVF takes ``nlist`` (alias
``n_clusters``) at fit and ``nprobe`` (alias ``n_probe_clusters``) at query,
as competitors/ivf_lsh/algorithm.py passes them.  ``ds_size`` truncates the
base as algorithm.py:19-21 does.  ``autotune: true`` times the scan-kernel
shapes on the fitted index (up to ``max_batch`` queries) before the first
query; ``index_file`` names a saved index (``Index.save``) to load instead of
building one, and is written after the build when it does not exist yet.

``backend: gpu`` is the reference hybrid's GPU backend
(competitors/hybrid/algorithm.py:54-95): the per-dim SQ8 index of
``gf_create`` and the reference GPU contestant's ranking chain per query
(``gf_query``: top ``r`` by the SQ8 key, default 150, fp32 re-rank), its
answers identical to the reference's; the HNSW parameters of that backend
block (M, ef_construction, mode, heuristic, reorder, ef) are ignored.

``get_n_distances`` counts full-precision distance evaluations: N per query
in exact mode (src/linear_index.cpp:108), the re-scored rows per query in the
quantized modes (gpu_filter.cu:171, "only full Euclidean distances count").
"""
from __future__ import annotations

import os

import numpy as np

from .index import Index

BACKENDS = ("mi355x", "gpu", "linear")


class Algorithm:
    def __init__(self):
        self._backend = "mi355x"
        self._index = None
        self._rerank = 0
        self._nprobe = 0
        self._compat = False
        self._gf = False

    def fit(self, train: np.ndarray, **index_params) -> None:
        ds_size = index_params.pop("ds_size", None)
        if ds_size is not None:
            train = train[: int(ds_size)]
        self._backend = str(index_params.pop("backend", "mi355x")).lower()
        if self._backend not in BACKENDS:
            raise ValueError(f"backend {self._backend!r} is not served by this plugin ({BACKENDS})")
        self._gf = self._backend == "gpu"
        if self._gf:
            # the hybrid's gpu backend: gf_create's index, gf_query's chain
            for unused in ("M", "ef_construction", "mode", "heuristic", "reorder", "ef"):
                index_params.pop(unused, None)
            index_params["mode"] = "sq8"
            index_params["sq_variant"] = 1
        # mode 'linear_index': the reference's LinearIndex.query semantics bit
        # for bit (early-abandon values beyond rank 64 included)
        mode = str(index_params.pop("mode", "exact")).lower()
        self._compat = mode == "linear_index"
        if self._compat:
            mode = "exact"
        self._rerank = int(index_params.pop("rerank", index_params.pop("r", index_params.pop("refine_r", 0))) or 0)
        sq_variant = int(index_params.pop("sq_variant", 0))
        max_batch = int(index_params.pop("max_batch", 4096))
        # IVF (competitors/ivf_lsh: fit(n_clusters) :128, query(n_probe_clusters, refine_r) :304)
        nlist = int(index_params.pop("nlist", index_params.pop("n_clusters", 0)) or 0)
        self._nprobe = int(index_params.pop("nprobe", index_params.pop("n_probe_clusters", 0)) or 0)
        kmeans_iters = int(index_params.pop("kmeans_iters", 0) or 0)
        autotune = bool(index_params.pop("autotune", False))
        index_file = index_params.pop("index_file", None)
        for unused in ("n_tables", "n_bits"):   # LSH buckets of the ivf_lsh fork: a full-list scan here
            index_params.pop(unused, None)
        if index_params:
            raise ValueError(f"unknown index_params: {sorted(index_params)}")
        kw = dict(mode=mode, rerank=self._rerank, sq_variant=sq_variant, max_batch=max_batch, nlist=nlist,
                  nprobe=self._nprobe, kmeans_iters=kmeans_iters)
        if index_file is not None and os.path.exists(str(index_file)):
            self._index = Index.load(str(index_file), train, **kw)
        else:
            self._index = Index(train, **kw)
            if index_file is not None:
                self._index.save(str(index_file))
        if autotune and mode != "ivf":
            self._index.autotune(max_batch=max_batch)

    def _need_index(self):
        if self._index is None:
            raise RuntimeError("index not trained: call fit() first")

    def query(self, query: np.ndarray, k: int, **query_params) -> np.ndarray:
        self._need_index()
        q = np.asarray(query, dtype=np.float32)
        if q.ndim != 1 or q.shape[0] != self._index.d:
            raise RuntimeError(f"query must be a 1-D vector of length {self._index.d}")
        if self._compat:
            return self._index.search_linear_index(q[None, :], int(k))[0][0]
        if self._gf:
            return self._index.search_gpu_filter(q, int(k), int(query_params.get("r", 150)))
        r, npb = self._qp(query_params)
        ids, _ = self._index.search(q[None, :], int(k), rerank=r, nprobe=npb)
        return ids[0]

    def _qp(self, qp):
        r = int(qp.get("rerank", qp.get("r", qp.get("refine_r", self._rerank))) or 0)
        npb = int(qp.get("nprobe", qp.get("n_probe_clusters", self._nprobe)) or 0)
        return max(r, 0), max(npb, 0)

    def batch_query(self, queries, k: int, **query_params):
        self._need_index()
        if self._compat:
            if _is_device(queries):
                queries = queries.cpu().numpy()
            return self._index.search_linear_index(np.asarray(queries, dtype=np.float32), int(k))
        if self._gf:  # the reference contestant answers one query per call
            if _is_device(queries):
                queries = queries.cpu().numpy()
            Q = np.asarray(queries, dtype=np.float32)
            ids = np.stack([self._index.search_gpu_filter(q, int(k), int(query_params.get("r", 150))) for q in Q]) \
                if len(Q) else np.empty((0, min(int(k), self._index.n)), dtype=np.int64)
            return ids, None
        r, npb = self._qp(query_params)
        if _is_device(queries):
            return self._index.search_device(queries, int(k), rerank=r, nprobe=npb)
        Q = np.asarray(queries, dtype=np.float32)
        if Q.ndim != 2 or Q.shape[1] != self._index.d:
            raise RuntimeError(f"queries must have shape (B, {self._index.d})")
        return self._index.search(Q, int(k), rerank=r, nprobe=npb)

    def get_n_distances(self) -> int:
        if self._index is None:
            return 0
        return self._index.n_distances


def _is_device(x) -> bool:
    try:
        import torch
    except ImportError:
        return False
    return isinstance(x, torch.Tensor) and x.is_cuda

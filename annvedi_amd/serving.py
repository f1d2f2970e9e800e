"""Continuous-batching request loop over a device-resident index.

The reference serves one query per call inside the organisers' harness loop
(orthogonal-competition/harness.py:159-174), timing each call.  Here requests
(1..max_batch queries each) arrive on their own schedule and are served by
``workers`` handles on the SAME device-resident index (``Index.clone``: one
copy of the data, one HIP stream and workspace per handle), so batches on
different handles overlap on the GPU.

Admission by HBM budget.  ``hbm_budget`` is the device memory the server may
use: the index (``Index.nbytes``) plus the workspace its requests hold while
they are on the GPU, ``SLOT_BYTES_PER_QUERY`` a query.  A budget below the
index, or below the index plus one full slot (``max_batch`` queries), is
refused at construction; the worker count is capped by the slots that fit.
Each request holds its bytes from the moment a worker takes it until its
answers are back.  A free worker takes the oldest waiting requests, in
arrival order, as long as their queries fit its slot and the bytes held by
every request in flight stay within the budget; otherwise they wait until
memory frees up.  ``set_budget`` changes the budget while serving (requests
in flight keep their memory; new ones are admitted under the new budget).
A request that no budget-sized slot can hold is refused at ``submit``.
There is no fixed batching window: a request that arrives while others run
joins the next batch a worker forms.

Requests are host arrays or torch device tensors; device requests are
batched and answered on the device (``Index.search_device``).

Every result depends only on its own query (per-query thresholds, the same
kernel arithmetic whatever the batch), so an answer is identical served
alone or in any mix; tests/test_gpu_serving.py checks this.
"""
from __future__ import annotations

import math
import threading
import time
from collections import deque
from concurrent.futures import Future

import numpy as np

# workspace bytes per query slot of a handle (candidate lists cap=65536 x 8 B,
# operands, outputs), the unit of the HBM admission budget
SLOT_BYTES_PER_QUERY = 65536 * 8 + 64 * 1024


def _is_device(x) -> bool:
    try:
        import torch
    except ImportError:
        return False
    return isinstance(x, torch.Tensor) and x.is_cuda


class Request:
    __slots__ = ("queries", "k", "t_arrival", "t_start", "t_done", "future", "worker")

    def __init__(self, queries, k, t_arrival):
        self.queries = queries
        self.k = k
        self.t_arrival = t_arrival
        self.t_start = None
        self.t_done = None
        self.future = Future()
        self.worker = None

    @property
    def nq(self) -> int:
        return int(self.queries.shape[0])


class Server:
    def __init__(self, index, workers: int = 2, max_batch: int = 512, hbm_budget: int | None = None,
                 rerank: int = 0, nprobe: int = 0, search=None):
        """``index``: an ``annvedi_amd.Index``.  ``search``: optional
        callable(handle, Q, k) -> (ids, dists), for host-only tests of the
        scheduling logic."""
        self.max_batch = int(max_batch)
        self.index_bytes = int(getattr(index, "nbytes", 0) or 0)
        self.slot_bytes = self.max_batch * SLOT_BYTES_PER_QUERY
        if hbm_budget is not None:
            self._check_budget(int(hbm_budget))
            workers = max(1, min(workers, (int(hbm_budget) - self.index_bytes) // self.slot_bytes))
        self._budget = None if hbm_budget is None else int(hbm_budget)
        self._reserved = 0           # workspace bytes of the requests on the GPU
        self.peak_reserved = 0
        self.rerank = int(rerank)
        self.nprobe = int(nprobe)
        self.handles = [index] + ([index.clone() for _ in range(workers - 1)] if search is None else
                                  [index] * (workers - 1))
        self._search = search
        self._q: deque[Request] = deque()
        self._cv = threading.Condition()
        self._stop = False
        self.batches: list[tuple[int, int, float, float, int]] = []  # (worker, queries, t_start, t_end, requests)
        self._threads = [threading.Thread(target=self._worker, args=(i,), daemon=True) for i in range(workers)]
        for t in self._threads:
            t.start()

    # ----------------------------------------------------------- interface
    def _check_budget(self, nbytes: int):
        if nbytes < self.index_bytes:
            raise ValueError(f"HBM budget of {nbytes} bytes is below the index's {self.index_bytes} bytes")
        if nbytes < self.index_bytes + self.slot_bytes:
            raise ValueError(f"HBM budget of {nbytes} bytes does not hold the index ({self.index_bytes} bytes) "
                             f"and one worker slot of {self.max_batch} queries ({self.slot_bytes} bytes)")

    @property
    def budget(self):
        return self._budget

    def set_budget(self, nbytes: int | None):
        """Change the HBM budget while serving (None: unlimited).  Requests in
        flight keep their memory; waiting ones are admitted under the new
        budget as memory frees up."""
        if nbytes is not None:
            self._check_budget(int(nbytes))
        with self._cv:
            self._budget = None if nbytes is None else int(nbytes)
            self._cv.notify_all()

    def submit(self, queries, k: int, t_arrival: float | None = None) -> Future:
        if _is_device(queries):
            Q = queries if queries.dim() == 2 else queries.reshape(1, -1)
            Q = Q.contiguous()
        else:
            Q = np.ascontiguousarray(np.atleast_2d(np.asarray(queries, dtype=np.float32)))
        if Q.shape[0] > self.max_batch:
            raise ValueError(f"request of {Q.shape[0]} queries exceeds the slot size {self.max_batch}")
        r = Request(Q, int(k), time.perf_counter() if t_arrival is None else t_arrival)
        with self._cv:
            if self._stop:
                raise RuntimeError("server closed")
            self._q.append(r)
            self._cv.notify()
        r.future.request = r
        return r.future

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------- workers
    def _take(self):
        """Oldest requests, in order, that fit one slot (same k, same kind of
        input) and the budget left by the requests in flight.  Caller holds
        the lock."""
        batch, nq = [], 0
        room = None if self._budget is None else self._budget - self.index_bytes - self._reserved
        while self._q:
            r = self._q[0]
            if batch and (nq + r.nq > self.max_batch or r.k != batch[0].k or
                          _is_device(r.queries) != _is_device(batch[0].queries)):
                break
            if room is not None and (nq + r.nq) * SLOT_BYTES_PER_QUERY > room:
                break
            batch.append(self._q.popleft())
            nq += r.nq
        return batch, nq * SLOT_BYTES_PER_QUERY

    def _worker(self, wi):
        h = self.handles[wi]
        while True:
            with self._cv:
                while True:
                    if self._stop and not self._q:
                        return
                    batch, held = self._take() if self._q else ([], 0)
                    if batch:
                        break
                    self._cv.wait()
                self._reserved += held
                self.peak_reserved = max(self.peak_reserved, self._reserved)
            t0 = time.perf_counter()
            for r in batch:
                r.t_start = t0
                r.worker = wi
            dev = _is_device(batch[0].queries)
            if len(batch) == 1:
                Q = batch[0].queries
            elif dev:
                import torch

                Q = torch.cat([r.queries for r in batch])
            else:
                Q = np.concatenate([r.queries for r in batch])
            try:
                if self._search is not None:
                    ids, dd = self._search(h, Q, batch[0].k)
                elif dev:
                    ids, dd = h.search_device(Q, batch[0].k, rerank=self.rerank, nprobe=self.nprobe)
                else:
                    ids, dd = h.search(Q, batch[0].k, rerank=self.rerank, nprobe=self.nprobe)
            except Exception as e:  # deliver the failure to every request of the batch
                with self._cv:
                    self._reserved -= held
                    self._cv.notify_all()
                for r in batch:
                    r.future.set_exception(e)
                continue
            t1 = time.perf_counter()
            with self._cv:
                self._reserved -= held
                self._cv.notify_all()
            self.batches.append((wi, int(Q.shape[0]), t0, t1, len(batch)))
            off = 0
            for r in batch:
                b = r.nq
                r.t_done = t1
                r.future.set_result((ids[off:off + b], dd[off:off + b]))
                off += b


def lognormal_sizes(n, median=8, sigma=1.0, lo=1, hi=512, seed=4000):
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.clip(np.round(rng.lognormal(math.log(median), sigma, size=n)), lo, hi).astype(int)


def loguniform_sizes(n, lo=1, hi=4096, seed=4002):
    """Request sizes drawn log-uniform in [lo, hi] (the batch mix of the
    precision + batching workload)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), size=n))).astype(int).clip(lo, hi)


def run_open_loop(server: Server, Q, sizes, qps: float, k: int = 100, seed: int = 4001):
    """Poisson arrivals at ``qps`` queries/s (requests of ``sizes`` queries,
    drawn from Q in order; Q a host array or a device tensor).  Returns
    per-request latencies from arrival to completion (s), the achieved
    throughput and the answers in submission order."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mean_size = float(np.mean(sizes))
    gaps = rng.exponential(mean_size / qps, size=len(sizes))
    futs = []
    t0 = time.perf_counter()
    t_next = t0
    off = 0
    for i, b in enumerate(sizes):
        t_next += gaps[i]
        delay = t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        futs.append(server.submit(Q[off:off + b], k, t_arrival=t_next))
        off += b
    out = [f.result() for f in futs]
    t_end = max(f.request.t_done for f in futs)
    lat = np.array([f.request.t_done - f.request.t_arrival for f in futs])
    nq = int(np.sum(sizes))
    return {"latency_s": lat, "qps": nq / (t_end - t0), "queries": nq, "answers": out,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3)}

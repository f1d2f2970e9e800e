"""Device-resident k-NN index over the C ABI (numpy and torch I/O).

``Index`` owns one ``annmi`` handle.  Host arrays are copied to HBM at
construction (the reference copies into a ``std::vector``,
src/linear_index.cpp:81-82); a torch device tensor is borrowed instead and
kept alive by the index.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class Index:
    def __init__(self, base, mode: str = "exact", rerank: int = 0, sq_variant: int = 0, max_batch: int = 4096,
                 nlist: int = 0, nprobe: int = 0, kmeans_iters: int = 0, kmeans_seed: int = 0, saved=None):
        """saved: bytes of Index.save_bytes() (or a path via Index.load) for
        the same base: recreates that index without re-encoding or k-means."""
        L = _lib.load()
        if mode not in _lib.MODES:
            raise ValueError(f"unknown mode {mode!r}; one of {sorted(_lib.MODES)}")
        p = _lib.Params()
        L.annmi_default_params(ctypes.byref(p), _lib.MODES[mode])
        p.rerank = int(rerank)
        p.sq_variant = int(sq_variant)
        p.max_batch = int(max_batch)
        if nlist:
            p.nlist = int(nlist)
        if nprobe:
            p.nprobe = int(nprobe)
        if kmeans_iters:
            p.kmeans_iters = int(kmeans_iters)
        if kmeans_seed:
            p.kmeans_seed = int(kmeans_seed)
        self.mode = mode
        self._keep = None
        if _is_torch_cuda(base):
            import torch

            if base.dim() != 2:
                raise RuntimeError("base must be a 2-D array")
            base = base.contiguous().to(torch.float32)
            torch.cuda.current_stream().synchronize()
            p.base_on_device = 1
            self._keep = base
            n, d = base.shape
            ptr = ctypes.c_void_p(base.data_ptr())
        else:
            arr = np.asarray(base)
            if arr.ndim != 2:
                raise RuntimeError("base must be a 2-D array")
            arr = np.ascontiguousarray(arr, dtype=np.float32)
            n, d = arr.shape
            ptr = arr.ctypes.data_as(ctypes.c_void_p)
        self.n, self.d = int(n), int(d)
        if saved is not None:
            blob = np.frombuffer(saved, dtype=np.uint8)
            h = L.annmi_load(ptr, self.n, self.d, ctypes.byref(p), blob.ctypes.data_as(ctypes.c_void_p), blob.size)
        else:
            h = L.annmi_create(ptr, self.n, self.d, ctypes.byref(p))
        if not h:
            raise RuntimeError(_lib.last_error())
        self._h = ctypes.c_void_p(h)
        self._L = L

    def save_bytes(self) -> bytes:
        """The index as a byte image (annmi_save): derived arrays only; the
        fp32 base is passed again to load it."""
        need = self._L.annmi_save(self._h, None, 0)
        if need < 0:
            raise RuntimeError(_lib.last_error())
        buf = np.empty(int(need), dtype=np.uint8)
        if self._L.annmi_save(self._h, buf.ctypes.data_as(ctypes.c_void_p), int(need)) < 0:
            raise RuntimeError(_lib.last_error())
        return buf.tobytes()

    def save(self, path: str) -> None:
        with open(path, "wb") as f:
            f.write(self.save_bytes())

    @classmethod
    def load(cls, path: str, base, mode: str, **kw) -> "Index":
        """An index written by save() over the same base."""
        with open(path, "rb") as f:
            return cls(base, mode=mode, saved=f.read(), **kw)

    def clone(self) -> "Index":
        """Another handle on the same device-resident index with its own HIP
        stream and workspace (concurrent batches).  Keeps this index alive."""
        h = self._L.annmi_clone(self._h)
        if not h:
            raise RuntimeError(_lib.last_error())
        c = Index.__new__(Index)
        c.mode, c.n, c.d = self.mode, self.n, self.d
        c._keep = self
        c._h = ctypes.c_void_p(h)
        c._L = self._L
        return c

    # ---------------------------------------------------------------- queries
    def _check_q(self, Q):
        if Q.ndim != 2 or Q.shape[1] != self.d:
            raise RuntimeError(f"queries must have shape (B, {self.d}), got {tuple(Q.shape)}")

    def search(self, Q, k: int, rerank: int = 0, nprobe: int = 0):
        """(ids int64 [B, k], dists float32 [B, k]) for a host batch."""
        Q = np.ascontiguousarray(np.asarray(Q, dtype=np.float32))
        if Q.ndim == 1:
            Q = Q[None, :]
        self._check_q(Q)
        B = Q.shape[0]
        ids = np.empty((B, k), dtype=np.int64)
        dd = np.empty((B, k), dtype=np.float32)
        got = self._L.annmi_batch_query_ex(self._h, Q.ctypes.data_as(ctypes.c_void_p), B, int(k), int(rerank),
                                           int(nprobe), ids.ctypes.data_as(ctypes.c_void_p),
                                           dd.ctypes.data_as(ctypes.c_void_p), 0)
        if got < 0:
            raise RuntimeError(_lib.last_error())
        return ids[:, :got], dd[:, :got]

    def search_gpu_filter(self, q, k: int, r: int = 150):
        """gf_query on this index (built as mode sq8, sq_variant 1, as
        gf_create builds it): the reference GPU contestant's ids for one
        query (competitors/gpu_filter/src/gpu_filter.cu:142-176)."""
        q = np.ascontiguousarray(np.asarray(q, dtype=np.float32)).ravel()
        if q.shape[0] != self.d:
            raise RuntimeError(f"query must be a 1-D vector of length {self.d}")
        out = np.empty(int(k), dtype=np.int64)
        got = self._L.gf_query(self._h, q.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), int(k), int(r),
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
        if got < 0:
            raise RuntimeError((self._L.gf_last_error() or b"").decode(errors="replace"))
        return out[:got]

    def search_linear_index(self, Q, k: int):
        """The reference's LinearIndex.query semantics bit for bit (ids and
        the fp32 values it ranks, early-abandon partial sums included)."""
        Q = np.ascontiguousarray(np.atleast_2d(np.asarray(Q, dtype=np.float32)))
        self._check_q(Q)
        B = Q.shape[0]
        ids = np.empty((B, k), dtype=np.int64)
        vals = np.empty((B, k), dtype=np.float32)
        got = self._L.annmi_query_linear_index(self._h, Q.ctypes.data_as(ctypes.c_void_p), B, int(k),
                                               ids.ctypes.data_as(ctypes.c_void_p),
                                               vals.ctypes.data_as(ctypes.c_void_p))
        if got < 0:
            raise RuntimeError(_lib.last_error())
        return ids[:, :got], vals[:, :got]

    def search_device(self, Q, k: int, rerank: int = 0, nprobe: int = 0, out=None):
        """Same for a torch device tensor; returns device tensors.  out: an
        optional (ids int64 [B, k], dists float32 [B, k]) pair of contiguous
        device tensors to write (a serving loop that reuses its buffers)."""
        import torch

        self._check_q(Q)
        Q = Q.contiguous().to(torch.float32)
        B = Q.shape[0]
        if out is None:
            ids = torch.empty((B, k), dtype=torch.int64, device=Q.device)
            dd = torch.empty((B, k), dtype=torch.float32, device=Q.device)
        else:
            ids, dd = out
            if (tuple(ids.shape) != (B, k) or tuple(dd.shape) != (B, k) or ids.dtype != torch.int64
                    or dd.dtype != torch.float32 or not ids.is_contiguous() or not dd.is_contiguous()
                    or ids.device != Q.device or dd.device != Q.device):
                raise RuntimeError(f"out must be contiguous ({B}, {k}) int64 and float32 tensors on {Q.device}")
        torch.cuda.current_stream().synchronize()
        got = self._L.annmi_batch_query_ex(self._h, ctypes.c_void_p(Q.data_ptr()), B, int(k), int(rerank),
                                           int(nprobe), ctypes.c_void_p(ids.data_ptr()),
                                           ctypes.c_void_p(dd.data_ptr()), 1)
        if got < 0:
            raise RuntimeError(_lib.last_error())
        return ids[:, :got], dd[:, :got]

    def last_batch_stats(self, B: int):
        n = np.zeros(B, dtype=np.uint32)
        f = np.zeros(B, dtype=np.uint32)
        got = self._L.annmi_last_batch_stats(self._h, n.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                             f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), B)
        return n[:got], f[:got]

    def export(self, what: str):
        """Host copy of index state: 'rows' (scan rows as uint8 [n, stride]),
        'scale', 'offset' (SQ per-dim, float32), 'norms' (float32 [n]),
        'meta' (dict)."""
        codes = {"rows": 0, "scale": 1, "offset": 2, "norms": 3, "meta": 4, "perm": 5, "list_offsets": 6,
                 "centroids": 7}
        w = codes[what]
        need = self._L.annmi_export(self._h, w, None, 0)
        if need < 0:
            raise RuntimeError(_lib.last_error())
        if w == 4:
            buf = np.zeros(4, dtype=np.float64)
            self._L.annmi_export(self._h, 4, buf.ctypes.data_as(ctypes.c_void_p), 32)
            return {"kdim": int(buf[0]), "row_stride": int(buf[1]), "fp8_alpha": float(buf[2]),
                    "max_norm": float(buf[3])}
        buf = np.empty(int(need), dtype=np.uint8)
        if need and self._L.annmi_export(self._h, w, buf.ctypes.data_as(ctypes.c_void_p), int(need)) < 0:
            raise RuntimeError(_lib.last_error())
        if w in (1, 2, 3):
            return buf.view(np.float32)
        if w == 5:
            return buf.view(np.uint32)
        if w == 6:
            return buf.view(np.int64)
        if w == 7:
            kd = (self.d + 15) // 16 * 16
            return buf.view(np.float32).reshape(-1, kd)[:, :self.d]
        stride = self.export("meta")["row_stride"]
        return buf.reshape(self.n, stride)

    def profile(self, enable: int = -1) -> dict:
        """Per-kernel event timing counters (see annmi_profile); enable=1 starts
        a fresh measurement, 0 stops, -1 only reads."""
        out = (ctypes.c_double * 8)()
        if self._L.annmi_profile(self._h, int(enable), out) < 0:
            raise RuntimeError(_lib.last_error())
        keys = ["scan_launches", "scan_ms", "scan_bytes", "scan_flops", "select_ms", "finalize_ms", "prep_ms", "steps"]
        return dict(zip(keys, list(out)))

    def profile_kernels(self) -> list:
        """Per scan-kernel instantiation: launches, ms, algorithmic bytes and flops."""
        import json

        need = self._L.annmi_profile_kernels(self._h, None, 0)
        if need < 0:
            raise RuntimeError(_lib.last_error())
        buf = ctypes.create_string_buffer(int(need))
        self._L.annmi_profile_kernels(self._h, buf, int(need))
        return json.loads(buf.value.decode())

    def autotune(self, max_batch: int = 4096, reps: int = 2) -> list:
        """Time every scan-kernel shape per batch size (1, 2, 4, ... max_batch)
        on this index and keep the fastest (see annmi_autotune).  Returns the
        per-batch-size report."""
        import json

        if int(max_batch) <= 0:
            raise ValueError("max_batch must be positive")
        need = self._L.annmi_autotune(self._h, int(max_batch), int(reps), None, 0)
        if need < 0:
            raise RuntimeError(_lib.last_error())
        buf = ctypes.create_string_buffer(int(need))
        if self._L.annmi_autotune(self._h, 0, 0, buf, int(need)) < 0:  # read the report back
            raise RuntimeError(_lib.last_error())
        return json.loads(buf.value.decode())

    def set_scan_shape(self, wide_min_batch: int = 0, wide_cfg: int = 0, chunk_rows: int = 0):
        """Fix the scan-kernel shape for every batch (0s: built-in rules)."""
        if self._L.annmi_set_scan_shape(self._h, int(wide_min_batch), int(wide_cfg), int(chunk_rows)) < 0:
            raise RuntimeError(_lib.last_error())

    @property
    def graph_count(self) -> int:
        """Step shapes this handle replays as captured HIP graphs."""
        return int(self._L.annmi_graph_count(self._h))

    @property
    def n_distances(self) -> int:
        return int(self._L.annmi_ndist(self._h))

    @property
    def nbytes(self) -> int:
        return int(self._L.annmi_index_bytes(self._h))

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.annmi_free(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _is_torch_cuda(x) -> bool:
    try:
        import torch
    except ImportError:
        return False
    return isinstance(x, torch.Tensor) and x.is_cuda

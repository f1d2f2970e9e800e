"""ctypes binding of ``libannmi.so`` (the C ABI in ``include/annmi.h``).

The library is the only compute path: there is no CPU fallback.  If the
shared object is missing the import of this module raises, and on a host
without a visible GPU ``annmi_create`` fails with ``RuntimeError``.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# ANNMI_LIB: another build of the same library (A/B timing of kernel variants)
LIB_PATH = os.environ.get("ANNMI_LIB") or os.path.join(HERE, "libannmi.so")

ANNMI_EXACT_F32 = 0
ANNMI_BF16 = 1
ANNMI_FP8 = 2
ANNMI_SQ8 = 3
ANNMI_SQ4 = 4
ANNMI_IVF_SQ8 = 5

MODES = {"exact": ANNMI_EXACT_F32, "fp32": ANNMI_EXACT_F32, "bf16": ANNMI_BF16, "fp8": ANNMI_FP8,
         "sq8": ANNMI_SQ8, "sq4": ANNMI_SQ4, "ivf_sq8": ANNMI_IVF_SQ8, "ivf": ANNMI_IVF_SQ8}

# every function declared in include/annmi.h
EXPORTS = [
    "annmi_default_params", "annmi_create", "annmi_query", "annmi_batch_query", "annmi_batch_query_device",
    "annmi_batch_query_ex", "annmi_query_linear_index", "annmi_batch_query_noalloc",
    "annmi_ndist", "annmi_last_batch_stats", "annmi_free", "annmi_clone", "annmi_last_error", "annmi_index_bytes",
    "annmi_profile", "annmi_profile_kernels", "annmi_export", "annmi_graph_count",
    "annmi_autotune", "annmi_set_scan_shape", "annmi_save", "annmi_load",
    "gf_create", "gf_query", "gf_ndist", "gf_free", "gf_last_error",
]


class Params(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32),
        ("rerank", ctypes.c_int32),
        ("sq_variant", ctypes.c_int32),
        ("base_on_device", ctypes.c_int32),
        ("max_batch", ctypes.c_int64),
        ("nlist", ctypes.c_int32),
        ("nprobe", ctypes.c_int32),
        ("kmeans_iters", ctypes.c_int32),
        ("kmeans_seed", ctypes.c_int32),
        ("reserved", ctypes.c_int64 * 2),
    ]


_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load and type the library (cached).  Raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `make -f annvedi_amd/csrc/Makefile` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    # One HIP runtime per process: torch ships its own libamdhip64 whose
    # dependents name it "libamdhip64.so", which does not match the soname this
    # library pulls in from /opt/rocm.  Loading torch first makes both share
    # torch's runtime; loading this library first would leave torch with a
    # second runtime that cannot open the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    fp = ctypes.POINTER(ctypes.c_float)
    i64p = ctypes.POINTER(ctypes.c_int64)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    L.annmi_default_params.restype = None
    L.annmi_default_params.argtypes = [ctypes.POINTER(Params), i32]
    L.annmi_create.restype = vp
    L.annmi_create.argtypes = [vp, i64, i32, ctypes.POINTER(Params)]
    L.annmi_query.restype = i32
    L.annmi_query.argtypes = [vp, fp, i32, i32, i64p, fp]
    L.annmi_batch_query.restype = i32
    L.annmi_batch_query.argtypes = [vp, fp, i32, i32, i32, i64p, fp]
    L.annmi_batch_query_ex.restype = i32
    L.annmi_batch_query_ex.argtypes = [vp, vp, i32, i32, i32, i32, vp, vp, i32]
    L.annmi_query_linear_index.restype = i32
    L.annmi_query_linear_index.argtypes = [vp, vp, i32, i32, vp, vp]
    L.annmi_batch_query_device.restype = i32
    L.annmi_batch_query_device.argtypes = [vp, vp, i32, i32, i32, vp, vp]
    L.annmi_batch_query_noalloc.restype = i32
    L.annmi_batch_query_noalloc.argtypes = [vp, vp, i32, i32, i32, vp, vp]
    L.annmi_ndist.restype = i64
    L.annmi_ndist.argtypes = [vp]
    L.annmi_last_batch_stats.restype = i32
    L.annmi_last_batch_stats.argtypes = [vp, u32p, u32p, i32]
    L.annmi_free.restype = None
    L.annmi_free.argtypes = [vp]
    L.annmi_clone.restype = vp
    L.annmi_clone.argtypes = [vp]
    L.annmi_last_error.restype = ctypes.c_char_p
    L.annmi_last_error.argtypes = []
    L.annmi_index_bytes.restype = i64
    L.annmi_index_bytes.argtypes = [vp]
    L.annmi_profile_kernels.restype = i32
    L.annmi_profile_kernels.argtypes = [vp, ctypes.c_char_p, i32]
    L.annmi_graph_count.restype = i32
    L.annmi_graph_count.argtypes = [vp]
    L.annmi_autotune.restype = i32
    L.annmi_autotune.argtypes = [vp, i32, i32, ctypes.c_char_p, i32]
    L.annmi_set_scan_shape.restype = i32
    L.annmi_set_scan_shape.argtypes = [vp, i32, i32, i64]
    L.annmi_save.restype = i64
    L.annmi_save.argtypes = [vp, vp, i64]
    L.annmi_load.restype = vp
    L.annmi_load.argtypes = [vp, i64, i32, ctypes.POINTER(Params), vp, i64]
    L.annmi_export.restype = i64
    L.annmi_export.argtypes = [vp, i32, vp, i64]
    L.annmi_profile.restype = i32
    L.annmi_profile.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_double)]
    # gpu_filter-compatible ABI, same argtypes as competitors/gpu_filter/algorithm.py:6-15
    L.gf_create.restype = vp
    L.gf_create.argtypes = [fp, ctypes.c_long, ctypes.c_long]
    L.gf_query.restype = ctypes.c_int
    L.gf_query.argtypes = [vp, fp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)]
    L.gf_ndist.restype = ctypes.c_longlong
    L.gf_ndist.argtypes = [vp]
    L.gf_free.restype = None
    L.gf_free.argtypes = [vp]
    L.gf_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def last_error() -> str:
    return (load().annmi_last_error() or b"").decode(errors="replace")

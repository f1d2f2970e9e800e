"""C-ABI boundary and host logic (CPU only: no compute calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from annvedi_amd import _lib
from annvedi_amd.algorithm import Algorithm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "annmi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b((?:annmi|gf)_[a-z_]+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    declared = header_functions()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(L, name), f"{name} declared in include/annmi.h but not exported"
    assert sorted(_lib.EXPORTS) == declared


def test_params_struct_layout_matches_header():
    # int32 x4, int64 max_batch, int64[4]
    assert ctypes.sizeof(_lib.Params) == 4 * 4 + 8 + 4 * 8
    L = _lib.load()
    p = _lib.Params()
    L.annmi_default_params(ctypes.byref(p), _lib.ANNMI_SQ8)
    assert p.mode == _lib.ANNMI_SQ8 and p.max_batch == 4096 and p.base_on_device == 0


def test_missing_library_fails_loudly(tmp_path):
    saved = _lib._lib
    _lib._lib = None
    try:
        with pytest.raises(ImportError):
            _lib.load(str(tmp_path / "nope.so"))
    finally:
        _lib._lib = saved


def test_create_rejects_bad_shapes_without_device_work():
    L = _lib.load()
    p = _lib.Params()
    L.annmi_default_params(ctypes.byref(p), 0)
    x = np.zeros((4, 8), dtype=np.float32)
    assert not L.annmi_create(x.ctypes.data_as(ctypes.c_void_p), 0, 8, ctypes.byref(p))
    assert "empty" in _lib.last_error()
    assert not L.annmi_create(x.ctypes.data_as(ctypes.c_void_p), 4, 5000, ctypes.byref(p))
    assert "unsupported" in _lib.last_error()
    # gf_free / annmi_free are NULL-safe (gpu_filter.cu:180-182)
    L.gf_free(None)
    L.annmi_free(None)
    assert L.annmi_ndist(None) == 0


def test_algorithm_errors_before_fit_and_on_bad_params():
    a = Algorithm()
    assert a.get_n_distances() == 0
    with pytest.raises(RuntimeError):
        a.query(np.zeros(8, np.float32), 5)
    with pytest.raises(ValueError):
        a.fit(np.zeros((10, 8), np.float32), backend="hnsw")


def test_no_device_create_raises_runtime_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    a = Algorithm()
    with pytest.raises(RuntimeError, match="device"):
        a.fit(np.random.rand(50, 16).astype(np.float32))


def test_saved_image_header_checked_before_any_device_call():
    """annmi_load refuses a truncated or foreign image with a message (no
    device needed to get there)."""
    L = _lib.load()
    p = _lib.Params()
    L.annmi_default_params(ctypes.byref(p), _lib.ANNMI_SQ8)
    base = np.zeros((10, 16), dtype=np.float32)
    junk = np.zeros(64, dtype=np.uint8)
    h = L.annmi_load(base.ctypes.data_as(ctypes.c_void_p), 10, 16, ctypes.byref(p),
                     junk.ctypes.data_as(ctypes.c_void_p), junk.size)
    assert not h and "too short" in _lib.last_error()
    junk = np.zeros(4096, dtype=np.uint8)
    h = L.annmi_load(base.ctypes.data_as(ctypes.c_void_p), 10, 16, ctypes.byref(p),
                     junk.ctypes.data_as(ctypes.c_void_p), junk.size)
    assert not h and "not an annmi index" in _lib.last_error()

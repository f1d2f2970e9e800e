"""annmi_batch_query_noalloc: the allocation-free entry point (a static
device scratch array instead of the handle's workspace) answers exactly as
annmi_batch_query_device does, on consecutive calls of different sizes and
across handles that share the array; out-of-range calls are refused."""
import ctypes

import numpy as np
import pytest

from annvedi_amd import Index, _lib, data

pytestmark = pytest.mark.gpu


def _call(L, fn, h, Q, k, rerank):
    import torch

    B = Q.shape[0]
    ids = torch.full((B, k), -7, dtype=torch.int64, device=Q.device)
    dd = torch.full((B, k), -7.0, dtype=torch.float32, device=Q.device)
    torch.cuda.synchronize()
    got = fn(h, ctypes.c_void_p(Q.data_ptr()), B, k, rerank, ctypes.c_void_p(ids.data_ptr()),
             ctypes.c_void_p(dd.data_ptr()))
    return got, ids.cpu().numpy(), dd.cpu().numpy()


@pytest.mark.parametrize("mode,rerank", [("exact", 0), ("sq8", 300), ("bf16", 200)])
def test_noalloc_equals_device_entry_point(mode, rerank):
    import torch

    print(f"[case] noalloc {mode}", flush=True)
    L = _lib.load()
    base, Q = data.workload("exact-1m-384", n=60_000, n_queries=200)
    idx = Index(torch.from_numpy(base).cuda(), mode=mode, rerank=rerank)
    other = Index(base[:20_000], mode="sq4", rerank=256)
    dQ = torch.from_numpy(Q).cuda()
    for B in (200, 37):  # the second call is smaller: scratch left from the first must not leak in
        q = dQ[:B].contiguous()
        g0, i0, d0 = _call(L, L.annmi_batch_query_device, idx._h, q, 100, rerank)
        g1, i1, d1 = _call(L, L.annmi_batch_query_noalloc, idx._h, q, 100, rerank)
        assert g0 == 100 and g1 == 100, _lib.last_error()
        np.testing.assert_array_equal(i1, i0)
        np.testing.assert_array_equal(d1, d0)
        # another handle on the same device uses the same array in between
        h0, j0, e0 = _call(L, L.annmi_batch_query_device, other._h, q[:5].contiguous(), 10, 256)
        h1, j1, e1 = _call(L, L.annmi_batch_query_noalloc, other._h, q[:5].contiguous(), 10, 256)
        assert h0 == 10 and h1 == 10, _lib.last_error()
        np.testing.assert_array_equal(j1, j0)
        np.testing.assert_array_equal(e1, e0)


def test_noalloc_refuses_what_its_scratch_cannot_hold():
    import torch

    L = _lib.load()
    base, Q = data.workload("exact-1m-384", n=5_000, n_queries=257)
    idx = Index(base, mode="exact")
    dQ = torch.from_numpy(Q).cuda()
    got, _, _ = _call(L, L.annmi_batch_query_noalloc, idx._h, dQ, 10, 0)
    assert got == -1 and "256" in _lib.last_error()
    got, _, _ = _call(L, L.annmi_batch_query_noalloc, idx._h, dQ[:4].contiguous(), 4096, 0)
    assert got == -1 and "2048" in _lib.last_error()
    ivf = Index(base, mode="ivf", nlist=16)
    got, _, _ = _call(L, L.annmi_batch_query_noalloc, ivf._h, dQ[:4].contiguous(), 10, 0)
    assert got == -1 and "IVF" in _lib.last_error()
    # the handle still answers afterwards, through both entry points
    g0, i0, _ = _call(L, L.annmi_batch_query_device, idx._h, dQ[:3].contiguous(), 10, 0)
    g1, i1, _ = _call(L, L.annmi_batch_query_noalloc, idx._h, dQ[:3].contiguous(), 10, 0)
    assert g0 == g1 == 10
    np.testing.assert_array_equal(i0, i1)

"""Edge shapes through the C ABI: tiny and odd n and d, k of 1 and beyond n,
empty batches, every mode.  With every row re-scored the quantized modes must
return the exact answer; rows past n are padded with id -1."""
import numpy as np
import pytest

from annvedi_amd import Index
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(1, 8), (17, 3), (300, 1), (300, 129), (2049, 40), (500, 1100)]


@pytest.mark.parametrize("mode", ["exact", "bf16", "fp8", "sq8", "sq4"])
@pytest.mark.parametrize("n,d", SHAPES)
def test_tiny_and_odd_shapes(mode, n, d):
    rng = np.random.Generator(np.random.PCG64(n * 1000 + d))
    base = rng.standard_normal((n, d), dtype=np.float32)
    Q = rng.standard_normal((5, d), dtype=np.float32)
    print(f"[case] edges {mode} n={n} d={d}", flush=True)
    idx = Index(base, mode=mode, rerank=min(n, 2048))
    for k in (1, 50, 1000):
        ids, dd = idx.search(Q, k)
        kk = min(k, n)
        assert ids.shape == (5, kk)
        ref_ids, ref_d = O.exact_topk_f64(base, Q, k)
        if mode == "exact" or n <= 2048:
            np.testing.assert_array_equal(ids, ref_ids)
            rel = np.abs(dd.astype(np.float64) - ref_d) / np.maximum(ref_d, 1e-30)
            assert rel.max() <= 1e-5
    # one query, and an empty batch
    i1, _ = idx.search(Q[:1], 10)
    np.testing.assert_array_equal(i1, O.exact_topk_f64(base, Q[:1], 10)[0])
    i0, d0 = idx.search(np.zeros((0, d), dtype=np.float32), 10)
    assert i0.shape[0] == 0 and d0.shape[0] == 0


def test_k_beyond_n_pads_through_the_raw_abi():
    import ctypes

    from annvedi_amd import _lib

    L = _lib.load()
    rng = np.random.Generator(np.random.PCG64(5))
    base = rng.standard_normal((7, 16), dtype=np.float32)
    q = rng.standard_normal(16, dtype=np.float32)
    p = _lib.Params()
    L.annmi_default_params(ctypes.byref(p), _lib.ANNMI_SQ8)
    h = L.annmi_create(base.ctypes.data_as(ctypes.c_void_p), 7, 16, ctypes.byref(p))
    assert h, _lib.last_error()
    ids = np.full(12, 99, dtype=np.int64)
    dd = np.zeros(12, dtype=np.float32)
    fp, i64p = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int64)
    got = L.annmi_query(h, q.ctypes.data_as(fp), 12, 7, ids.ctypes.data_as(i64p), dd.ctypes.data_as(fp))
    assert got == 7
    np.testing.assert_array_equal(ids[:7], O.exact_topk_f64(base, q[None, :], 7)[0][0])
    assert np.all(ids[7:] == -1) and np.all(np.isinf(dd[7:]))
    L.annmi_free(h)


def test_dimension_limits():
    """A 16-query scan tile must fit the 160 KB of LDS: fp32 rows up to
    d = 2544, the other modes up to the library's 4096; beyond that
    creation fails with an error instead of a failed launch."""
    rng = np.random.Generator(np.random.PCG64(11))
    for mode, d_ok, d_bad, msg in (("exact", 2544, 2545, "too large"), ("sq8", 4096, 4097, "unsupported")):
        print(f"[case] dimension limit {mode} d={d_ok}", flush=True)
        base = rng.standard_normal((40, d_ok), dtype=np.float32)
        Q = rng.standard_normal((3, d_ok), dtype=np.float32)
        ids, _ = Index(base, mode=mode, rerank=40).search(Q, 10)
        np.testing.assert_array_equal(ids, O.exact_topk_f64(base, Q, 10)[0])
        with pytest.raises(RuntimeError, match=msg):
            Index(rng.standard_normal((40, d_bad), dtype=np.float32), mode=mode)


@pytest.mark.parametrize("mode", ["exact", "sq8"])
def test_batches_beyond_max_batch_split_into_steps(mode):
    """150 queries on a handle built for 64 run as steps of 64, 64 and 22
    (the last shapes repeat, so steps are captured and replayed); host and
    device buffers give the single-step answers."""
    import torch

    rng = np.random.Generator(np.random.PCG64(21))
    base = rng.standard_normal((20000, 96), dtype=np.float32)
    Q = rng.standard_normal((150, 96), dtype=np.float32)
    print(f"[case] split batch {mode} B=150 max_batch=64", flush=True)
    whole = Index(base, mode=mode, rerank=200)
    ref_ids, ref_d = whole.search(Q, 100)
    small = Index(base, mode=mode, rerank=200, max_batch=64)
    for _ in range(3):
        ids, dd = small.search(Q, 100)
        np.testing.assert_array_equal(ids, ref_ids)
        np.testing.assert_array_equal(dd, ref_d)
    di, dd2 = small.search_device(torch.from_numpy(Q).to("cuda:0"), 100)
    np.testing.assert_array_equal(di.cpu().numpy(), ref_ids)
    np.testing.assert_array_equal(dd2.cpu().numpy(), ref_d)
    if mode == "exact":
        np.testing.assert_array_equal(ref_ids, O.exact_topk_f64(base, Q, 100)[0])

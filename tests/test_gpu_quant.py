"""Quantized candidate scans (BF16 / FP8 / SQ8 / SQ4) + FP32 re-score on the GPU."""
import ctypes

import numpy as np
import pytest

from annvedi_amd import Index, _lib, data
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ds():
    base, Q = data.workload("sq-1m-768", n=40000, n_queries=64)
    ref_ids, ref_d = O.exact_topk_f64(base, Q, 100)
    return base, Q, ref_ids, ref_d


@pytest.mark.parametrize("mode,R,min_recall", [("bf16", 400, 0.99), ("fp8", 1000, 0.99),
                                               ("sq8", 1000, 0.99), ("sq4", 1000, 0.99)])
def test_quantized_recall_and_rescore(ds, mode, R, min_recall):
    base, Q, ref_ids, ref_d = ds
    print(f"[case] {mode} R={R}", flush=True)
    idx = Index(base, mode=mode, rerank=R)
    ids, dd = idx.search(Q, 100)
    rec = O.recall_at_k(base, Q, ids, ref_ids, 100).mean()
    print(f"{mode} recall@100 = {rec:.4f}")
    assert rec >= min_recall
    # returned distances are fp32 re-scores of the returned rows, sorted
    for i in range(0, 64, 9):
        for j in range(0, 100, 11):
            assert dd[i, j] == O.l2_linear_index(base[ids[i, j]], Q[i])
    assert np.all(np.diff(dd, axis=1) >= 0)
    nres, flags = idx.last_batch_stats(64)
    assert np.all(flags == 0) and np.all(nres >= R)
    assert idx.n_distances == int(nres.sum())  # only fp32 re-scores count (gpu_filter.cu:171)


@pytest.mark.parametrize("mode", ["bf16", "fp8", "sq8", "sq4"])
def test_rescore_everything_equals_exact(mode):
    """With R = n every row is re-scored, so the answer must be the exact one."""
    base, Q = data.workload("sq-1m-768", n=1500, n_queries=12)
    idx = Index(base, mode=mode, rerank=1500)
    ids, dd = idx.search(Q, 100)
    ref_ids, ref_d = O.exact_topk_f64(base, Q, 100)
    np.testing.assert_array_equal(ids, ref_ids)


def test_gpu_filter_abi_drop_in():
    """Bind gf_* exactly as competitors/gpu_filter/algorithm.py:6-15 does."""
    L = _lib.load()
    base, Q = data.workload("exact-1m-384", n=20000, n_queries=8)
    h = L.gf_create(base.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), base.shape[0], base.shape[1])
    assert h, L.gf_last_error()
    ref_ids, _ = O.exact_topk_f64(base, Q, 100)
    out = np.empty(100, dtype=np.int64)
    recs = []
    for i, q in enumerate(Q):
        got = L.gf_query(h, q.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 100, 500,
                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
        assert got == 100
        recs.append(len(set(out.tolist()) & set(ref_ids[i].tolist())) / 100)
    assert np.mean(recs) >= 0.99
    assert L.gf_ndist(h) >= 8 * 500
    L.gf_free(h)


@pytest.mark.parametrize("mode", ["bf16", "fp8", "sq8", "sq4"])
def test_large_batch_quantized_consistent(mode):
    """Query-stationary kernel (B = 256) vs streaming kernel (B = 1): a
    query's answer does not depend on the batch it is in (same keys, same
    per-query thresholds), and with R = 600 the re-scored rows hold the true
    top-100 on this data."""
    print(f"[case] batch invariance {mode}", flush=True)
    base, Q = data.workload("exact-1m-384", n=120_000, n_queries=256)
    idx = Index(base, mode=mode, rerank=600)
    ids, dd = idx.search(Q, 100)
    ref_ids, _ = O.exact_topk_f64(base, Q[:8], 100)
    rec = O.recall_at_k(base, Q[:8], ids[:8], ref_ids, 100).mean()
    assert rec == 1.0, rec
    for i in (0, 200, 255):
        i1, d1 = idx.search(Q[i:i + 1], 100)
        np.testing.assert_array_equal(i1[0], ids[i])
        np.testing.assert_array_equal(d1[0], dd[i])


@pytest.mark.parametrize("mode", ["bf16", "fp8", "sq8", "sq4"])
@pytest.mark.parametrize("n", [5, 16, 33, 1039])
def test_partial_tiles_of_the_tiled_layout(mode, n):
    """The quantized scan copies are stored in 16-row tiles: row counts below,
    at and past a tile boundary give the exact answer with a full re-score,
    and the exported rows equal a row-major encoding (byte-exact)."""
    print(f"[case] partial tiles {mode} n={n}", flush=True)
    base, Q = data.workload("sq-1m-768", n=n, n_queries=7)
    k = min(100, n)
    idx = Index(base, mode=mode, rerank=n)
    ids, dd = idx.search(Q, k)
    ref_ids, _ = O.exact_topk_f64(base, Q, k)
    np.testing.assert_array_equal(ids, ref_ids)
    rows = idx.export("rows")
    assert rows.shape[0] == n
    if mode == "sq8":
        sc, of = O.sq8_train(base, 0)
        np.testing.assert_array_equal(rows[:, :base.shape[1]], O.sq8_encode(base, sc, of, 0))


@pytest.mark.parametrize("B", [1, 17, 300])
def test_tiled_scan_on_unaligned_ranges_and_batches(B):
    """Row counts that end mid-tile, with the wide kernel on small ranges
    (large batches) and the streaming kernel (one query): answers equal the
    full re-score."""
    n = 70_000 + 13
    print(f"[case] tiled ranges sq8 n={n} B={B}", flush=True)
    base, Q = data.workload("exact-1m-384", n=n, n_queries=B)
    idx = Index(base, mode="sq8", rerank=400)
    ids, _ = idx.search(Q, 100)
    full = Index(base, mode="sq8", rerank=2048)
    fids, _ = full.search(Q, 100)
    ref_ids, _ = O.exact_topk_f64(base, Q[:4], 100)
    for i in range(min(B, 4)):
        assert len(set(ids[i]) & set(ref_ids[i])) >= 97
        assert len(set(fids[i]) & set(ref_ids[i])) >= 99

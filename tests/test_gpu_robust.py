"""Queries the regular step cannot bound: adversarial row order (candidate
lists overflow), thousands of duplicate rows (more than 2048 rows within the
error bound of rank k) and long fp32-equal / fp64-distinct runs at the rank-k
boundary.  The library answers them again over every row; the answer must be
the exact oracle's, ids ranked by (fp64 distance, id), where the reference
returns k ids too (src/linear_index.cpp:110-118)."""
import numpy as np
import pytest

from annvedi_amd import Index, data
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _check_exact(base, Q, ids, dd, k=100, exact_rows=None):
    """Flagged queries (and every query of an exact index) must return the
    oracle's ids; the other queries of a quantized index its recall."""
    ref_ids, ref_d = O.exact_topk_f64(base, Q, k)
    rows = range(Q.shape[0]) if exact_rows is None else exact_rows
    for i in rows:
        np.testing.assert_array_equal(ids[i], ref_ids[i])
        rel = np.abs(dd[i].astype(np.float64) - ref_d[i]) / np.maximum(ref_d[i], 1e-30)
        assert rel.max() <= 1e-5, rel.max()
    assert O.recall_at_k(base, Q, ids, ref_ids, k).mean() >= 0.99


@pytest.mark.parametrize("mode", ["exact", "sq8"])
def test_rows_sorted_by_descending_distance(mode):
    """Every row is closer to query 0 than every row before it: each range's
    filter admits the whole range and the candidate list overflows."""
    print(f"[case] robust sorted base {mode}", flush=True)
    base, Q = data.workload("exact-1m-384", n=120_000, n_queries=3)
    dist = ((base.astype(np.float64) - Q[0].astype(np.float64)) ** 2).sum(1)
    base = np.ascontiguousarray(base[np.argsort(-dist, kind="stable")])
    idx = Index(base, mode=mode, rerank=400)
    ids, dd = idx.search(Q, 100)
    _check_exact(base, Q, ids, dd, exact_rows=None if mode == "exact" else [0])
    _, flags = idx.last_batch_stats(3)
    assert flags[0] & 8, flags  # query 0 was answered by the all-rows path
    assert flags[0] & 1, flags


@pytest.mark.parametrize("mode", ["exact", "sq8", "bf16"])
def test_five_thousand_duplicates(mode):
    """5000 copies of one vector next to the query: they tie at every rank,
    so the answer is the 100 lowest ids among them."""
    print(f"[case] robust duplicates {mode}", flush=True)
    rng = np.random.Generator(np.random.PCG64(77))
    base, Q = data.workload("exact-1m-384", n=40_000, n_queries=4)
    dup_rows = np.sort(rng.choice(40_000, size=5000, replace=False))
    v = Q[1] + np.float32(0.01) * rng.standard_normal(384, dtype=np.float32)
    base[dup_rows] = v
    idx = Index(base, mode=mode, rerank=1000)
    ids, dd = idx.search(Q, 100)
    _check_exact(base, Q, ids, dd, exact_rows=None if mode == "exact" else [1])
    np.testing.assert_array_equal(ids[1], dup_rows[:100])
    _, flags = idx.last_batch_stats(4)
    assert flags[1] & 8 and flags[1] & 2, flags


def _planted_run(run_len, n_near):
    """q = 0; n_near rows at squared distance < 0.6; run_len rows at fp32
    distance exactly 1.0 (1 + e^2 with e^2 below half an ulp of 1) but
    distinct fp64 distances, their ids in the reverse of their fp64 order;
    far rows beyond distance 4."""
    rng = np.random.Generator(np.random.PCG64(run_len))
    d = 32
    n = 3000
    base = np.empty((n, d), dtype=np.float32)
    far = rng.standard_normal((n, d)).astype(np.float32)
    far *= (np.sqrt(rng.uniform(4, 9, size=n)) / np.linalg.norm(far, axis=1)).astype(np.float32)[:, None]
    base[:] = far
    near_rows = rng.choice(np.arange(0, 1000), size=n_near, replace=False)
    near = rng.standard_normal((n_near, d)).astype(np.float32)
    near *= (np.sqrt(rng.uniform(0.1, 0.6, size=n_near)) / np.linalg.norm(near, axis=1)).astype(np.float32)[:, None]
    base[near_rows] = near
    run_rows = 2000 + np.arange(run_len)[::-1]  # higher fp64 distance -> lower id
    for j, r in enumerate(run_rows):
        base[r] = 0
        base[r, 0] = 1.0
        base[r, 1] = np.float32(5e-6 * (j + 1))
    Q = np.zeros((2, d), dtype=np.float32)
    Q[1] = rng.standard_normal(d).astype(np.float32) * np.float32(0.05)
    return base, Q


@pytest.mark.parametrize("run_len,n_near", [(20, 90), (40, 80)])
def test_fp32_equal_fp64_distinct_run_at_rank_k(run_len, n_near):
    """A run of rows with equal fp32 and distinct fp64 distances straddles
    rank 100: finalize orders runs of up to 32 rows in fp64; a longer run is
    answered by the all-rows path.  Either way the ids follow (fp64, id)."""
    print(f"[case] robust planted run {run_len}", flush=True)
    base, Q = _planted_run(run_len, n_near)
    q0 = Q[0]
    d32 = np.array([O.l2_linear_index(base[r], q0) for r in range(2000, 2000 + run_len)])
    assert np.all(d32 == np.float32(1.0))  # the plant: equal fp32 values
    idx = Index(base, mode="exact")
    ids, dd = idx.search(Q, 100)
    _check_exact(base, Q, ids, dd)
    _, flags = idx.last_batch_stats(2)
    if run_len > 32:
        assert flags[0] & 4 and flags[0] & 8, flags
    else:
        assert flags[0] == 0, flags

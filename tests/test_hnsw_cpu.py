"""The CPU HNSW baseline (oracle/hnsw_cpu.cpp, submission/src/hnsw.cpp
restated with 64-bit offsets) -- CPU tests.

Pins: with a beam as wide as the base the layer-0 search visits every
reachable node, so the answer is the brute-force top-k (fp32 metric) or the
brute-force top-k of the exact re-rank of every node (SQ8 metrics: the
window is re-ranked in fp32, hnsw.cpp:220-228); a single-threaded build is
deterministic, and renumbering the nodes in BFS order (hnsw.cpp:335-386)
changes no answer; on the same data and parameters its recall@100 is the
compiled reference's within a small margin (both builds are randomised)."""
import numpy as np
import pytest

from annvedi_amd import data
from oracle import oracle as O


def _base(n=3000, d=48, seed=91):
    return data.mixture(n, d, 24, 0.3, seed, center_seed=seed + 1), data.mixture(20, d, 24, 0.3, seed + 2,
                                                                                 center_seed=seed + 1)


@pytest.mark.parametrize("metric", ["float", "sq8", "sq8pd"])
def test_full_beam_is_brute_force_over_the_reachable_nodes(metric):
    """ef = n: the beam never stops early, so it takes every node reachable
    on layer 0 (the heuristic leaves a few nodes without incoming links, as
    the reference's does: hnsw.cpp:351-352), and the answer is the
    brute-force top-k over exactly those nodes."""
    n = 600
    base, Q = _base(n, 40)
    h = O.HnswCpu(base, M=8, ef_construction=64, metric=metric, heuristic=True, threads=2)
    for i in range(Q.shape[0]):
        reach = h.query(Q[i], n, ef=n)
        assert len(reach) >= 0.95 * n and len(set(reach.tolist())) == len(reach)
        sub = np.sort(reach)
        want, _ = O.exact_topk_f64(base[sub], Q[i:i + 1], 10)
        np.testing.assert_array_equal(h.query(Q[i], 10, ef=n), sub[want[0]], err_msg=str(i))
    assert h.n_distances >= Q.shape[0] * 0.95 * n


def test_reorder_changes_no_answer():
    base, Q = _base(2500, 32)
    a = O.HnswCpu(base, M=12, ef_construction=80, metric="sq8", heuristic=True, reorder=False, threads=1, seed=7)
    b = O.HnswCpu(base, M=12, ef_construction=80, metric="sq8", heuristic=True, reorder=True, threads=1, seed=7)
    for i in range(Q.shape[0]):
        np.testing.assert_array_equal(a.query(Q[i], 50, ef=60), b.query(Q[i], 50, ef=60))


def test_batch_queries_equal_single_queries():
    base, Q = _base(2000, 32)
    h = O.HnswCpu(base, M=12, ef_construction=80, metric="sq8", heuristic=True, reorder=True, threads=4)
    out = h.query_batch(Q, 30, ef=50, threads=3)
    for i in range(Q.shape[0]):
        np.testing.assert_array_equal(out[i], h.query(Q[i], 30, ef=50))


def test_recall_beside_the_compiled_reference():
    ref = O.ref_module("hnsw_cpp")
    if ref is None:
        pytest.skip("the reference's hnsw_cpp is not built here (oracle/ref.mk needs the reference tree)")
    base, Q = _base(6000, 48, seed=95)
    true_ids, _ = O.exact_topk_f64(base, Q, 100)
    td = np.linalg.norm(base[true_ids] - Q[:, None, :], axis=2).astype(np.float32)

    def recall(ids):
        pd = np.linalg.norm(base[ids] - Q[:, None, :], axis=2).astype(np.float32)
        return float(np.mean(pd <= td[:, 99:100]))

    # (five builds of each measured here: 0.9955-0.9965 vs 0.9925-0.996)
    r = ref.HNSWIndex()
    r.fit(base, 12, 64, "sq8", True, True)
    rr = recall(np.stack([np.asarray(r.query(Q[i], 100, 150)) for i in range(Q.shape[0])]))
    h = O.HnswCpu(base, M=12, ef_construction=64, metric="sq8", heuristic=True, reorder=True)
    hr = recall(h.query_batch(Q, 100, ef=150, threads=2))
    assert abs(hr - rr) <= 0.02, (hr, rr)
    assert hr >= 0.98


def test_refuses_bad_arguments():
    with pytest.raises(RuntimeError):
        O.HnswCpu(np.zeros((0, 4), np.float32))
    with pytest.raises(RuntimeError):
        O.HnswCpu(np.zeros(8, np.float32))

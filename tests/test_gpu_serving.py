"""Continuous batching on the GPU: every answer equals the same request
served alone, and batches on different handles overlap in time."""
import numpy as np
import pytest

from annvedi_amd import Index, data
from annvedi_amd.serving import Server, lognormal_sizes, run_open_loop

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["exact", "bf16", "ivf"])
def test_served_answers_equal_alone_answers(mode):
    base, Q = data.workload("exact-1m-384", n=200_000, n_queries=3000)
    kw = dict(nlist=128, kmeans_iters=5) if mode == "ivf" else {}
    idx = Index(base, mode=mode, rerank=300, **kw)
    sizes = lognormal_sizes(120, seed=11)
    sizes = sizes[np.cumsum(sizes) <= Q.shape[0]]
    with Server(idx, workers=3, max_batch=512, rerank=300, nprobe=16) as srv:
        res = run_open_loop(srv, Q, sizes, qps=30000, k=100)
        batches = list(srv.batches)
    off = 0
    for b, (ids, dd) in zip(sizes, res["answers"]):
        a_ids, a_dd = idx.search(Q[off:off + b], 100, rerank=300, nprobe=16)
        np.testing.assert_array_equal(ids, a_ids)
        np.testing.assert_array_equal(dd, a_dd)
        off += b
    # requests that arrived while others ran shared a batch, and batches on
    # different workers overlapped in time
    assert max(nr for *_, nr in batches) > 1
    overl = sum(1 for i, a in enumerate(batches) for b in batches[i + 1:]
                if a[0] != b[0] and a[2] < b[3] and b[2] < a[3])
    print(f"{mode}: {len(batches)} batches, {overl} overlapping pairs, QPS {res['qps']:.0f}, "
          f"p50 {res['p50_ms']:.2f} ms, p99 {res['p99_ms']:.2f} ms")
    assert overl > 0


@pytest.mark.parametrize("mode", ["sq8", "ivf"])
def test_repeated_shapes_on_cloned_handles_across_threads(mode):
    """Regression: several threads serving on clones of one index with the
    same request shapes again and again (the shapes a single handle would
    capture as graphs).  Clones and their source run steps eagerly: no step
    fails on a capture in one thread meeting copies or frees in another, and
    every answer equals the answer computed alone."""
    base, Q = data.workload("exact-1m-384", n=120_000, n_queries=4096)
    kw = dict(nlist=64, kmeans_iters=4) if mode == "ivf" else {}
    print(f"[case] repeated shapes on clones {mode}", flush=True)
    idx = Index(base, mode=mode, rerank=200, **kw)
    sizes = np.array([1, 4, 8] * 120)
    with Server(idx, workers=4, max_batch=64, rerank=200, nprobe=8) as srv:
        res = run_open_loop(srv, Q, sizes, qps=200000, k=100)
    assert idx.graph_count == 0
    off = 0
    for b, (ids, dd) in zip(sizes, res["answers"]):
        a_ids, a_dd = idx.search(Q[off:off + b], 100, rerank=200, nprobe=8)
        np.testing.assert_array_equal(ids, a_ids)
        off += b


def test_device_requests_under_an_hbm_budget():
    """Requests as device tensors, served under a budget of the index plus
    two slots: budgets below the index or below one slot are refused, the
    workers are capped at the slots that fit, the bytes held stay within the
    budget and every answer equals the one computed alone."""
    import torch

    from annvedi_amd.serving import SLOT_BYTES_PER_QUERY

    print("[case] served device requests under a budget", flush=True)
    base, Q = data.workload("exact-1m-384", n=100_000, n_queries=1000)
    idx = Index(base, mode="sq8", rerank=200)
    slot = 64 * SLOT_BYTES_PER_QUERY
    assert idx.nbytes > 0
    with pytest.raises(ValueError, match="below the index"):
        Server(idx, workers=2, max_batch=64, hbm_budget=idx.nbytes - 1)
    with pytest.raises(ValueError, match="one worker slot"):
        Server(idx, workers=2, max_batch=64, hbm_budget=idx.nbytes + slot - 1)
    dQ = torch.from_numpy(Q).cuda()
    sizes = lognormal_sizes(100, seed=5, hi=64)
    sizes = sizes[np.cumsum(sizes) <= Q.shape[0]]
    with Server(idx, workers=3, max_batch=64, hbm_budget=idx.nbytes + 2 * slot, rerank=200) as srv:
        assert len(srv.handles) == 2
        res = run_open_loop(srv, dQ, sizes, qps=20000, k=100)
        assert 0 < srv.peak_reserved <= 2 * slot
    off = 0
    for b, (ids, dd) in zip(sizes, res["answers"]):
        assert ids.is_cuda
        a_ids, a_dd = idx.search(Q[off:off + b], 100, rerank=200)
        np.testing.assert_array_equal(ids.cpu().numpy(), a_ids)
        np.testing.assert_array_equal(dd.cpu().numpy(), a_dd)
        off += b

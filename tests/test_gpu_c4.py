"""The precision + batching configuration: the 10M x 384 workload's
generator (4096 clusters, sigma 0.3, unit rows) at 200k rows, BF16 and FP8
MFMA candidates, fp32 re-score of the top 400 with the fp64 near-tie step
at rank 100, over a mix of batch sizes that runs both the streaming and the
query-stationary scan kernels, eager and replayed from captured graphs.
Recall@100 and distance error against the fp64 oracle."""
import numpy as np
import pytest

from annvedi_amd import Index, data
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BATCHES = [1, 2, 5, 35, 200]  # 243 queries; bf16 streams up to 7, fp8 up to 3


@pytest.fixture(scope="module")
def c4():
    base, Q = data.workload("mixed-10m-384", n=200_000, n_queries=sum(BATCHES), config_id=3)
    ref_ids, ref_d = O.exact_topk_f64(base, Q, 100)
    return base, Q, ref_ids, ref_d


@pytest.mark.parametrize("mode", ["bf16", "fp8"])
def test_c4_precision_and_batching(c4, mode):
    base, Q, ref_ids, ref_d = c4
    print(f"[case] C4 {mode} R=400 n=200000 batches={BATCHES}", flush=True)
    idx = Index(base, mode=mode, rerank=400)
    # both scan kernels run over the mix (timed pass: no capture while timing)
    idx.profile(1)
    off = 0
    for b in BATCHES:
        idx.search(Q[off:off + b], 100)
        off += b
    names = {k["kernel"].split("<")[0] for k in idx.profile_kernels()}
    idx.profile(0)
    assert {"scan_kernel", "scan_wide_kernel"} <= names, names
    ids = np.empty((Q.shape[0], 100), dtype=np.int64)
    dd = np.empty((Q.shape[0], 100), dtype=np.float32)
    off = 0
    for b in BATCHES:
        first = idx.search(Q[off:off + b], 100)  # eager, then captured, then replayed
        for _ in range(3):
            again = idx.search(Q[off:off + b], 100)
            np.testing.assert_array_equal(again[0], first[0])
            np.testing.assert_array_equal(again[1], first[1])
        ids[off:off + b], dd[off:off + b] = first
        off += b
    assert idx.graph_count >= len(BATCHES)
    rec = O.recall_at_k(base, Q, ids, ref_ids, 100).mean()
    print(f"C4 {mode} recall@100 = {rec:.4f}")
    assert rec >= 0.99
    # returned distances: the rows' fp32 values, within 1e-5 of fp64
    d64 = np.array([[O.l2_f64(base[r], Q[i]) for r in ids[i]] for i in range(Q.shape[0])])
    err = np.abs(dd.astype(np.float64) - d64) / np.maximum(d64, 1e-30)
    print(f"C4 {mode} max distance error vs fp64 = {err.max():.2e}")
    assert err.max() <= 1e-5

"""Captured query steps: a step shape seen for the second time is captured as
a HIP graph and replayed afterwards.  Every replay must give the answer the
eager path gives for the same queries (per-query answers do not depend on the
batch they arrive in), with other eager work and host synchronisations
between the replays."""
import numpy as np
import pytest

from annvedi_amd import Index, data
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["exact", "sq8", "bf16"])
def test_replayed_steps_equal_eager_steps(mode):
    n, d, B, k = 60000, 96, 5, 100
    print(f"[case] graphs mode={mode} n={n} d={d} B={B}", flush=True)
    base = data.mixture(n, d, 64, 0.3, 41, center_seed=42)
    Q = data.mixture(6 * B, d, 64, 0.3, 43, center_seed=42)
    idx = Index(base, mode=mode, rerank=300)
    # eager answers: one batch of a shape that is only ever seen once
    ref_ids, ref_d = idx.search(Q, k)
    other = data.mixture(7, d, 64, 0.3, 44, center_seed=42)
    for i in range(6):
        ids, dd = idx.search(Q[i * B:(i + 1) * B], k)  # 1st eager, 2nd captured, then replays
        np.testing.assert_array_equal(ids, ref_ids[i * B:(i + 1) * B], err_msg=f"step {i}")
        np.testing.assert_array_equal(dd, ref_d[i * B:(i + 1) * B], err_msg=f"step {i}")
        idx.search(other[: 1 + i % 3], k)  # eager work of other shapes in between
    assert idx.graph_count >= 1
    if mode == "exact":
        o_ids, _ = O.exact_topk_f64(base, Q, k)
        np.testing.assert_array_equal(ref_ids, o_ids)


def test_replays_with_device_tensors():
    import torch

    n, d, k = 40000, 128, 10
    base = data.mixture(n, d, 32, 0.3, 51, center_seed=52)
    Q = data.mixture(12, d, 32, 0.3, 53, center_seed=52)
    print(f"[case] graphs device n={n} d={d}", flush=True)
    idx = Index(torch.from_numpy(base).cuda(), mode="exact")
    o_ids, _ = O.exact_topk_f64(base, Q, k)
    Qd = torch.from_numpy(Q).cuda()
    for rep in range(4):
        for i in range(4):
            ids, _ = idx.search_device(Qd[3 * i:3 * i + 3], k)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(ids.cpu().numpy(), o_ids[3 * i:3 * i + 3], err_msg=f"rep {rep} batch {i}")

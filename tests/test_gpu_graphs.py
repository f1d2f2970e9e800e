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


@pytest.mark.parametrize("mode", ["exact", "sq8", "ivf"])
def test_replays_retargeted_to_new_output_buffers(mode):
    """One graph per step shape: a replay into output buffers other than the
    ones it was captured with rewrites its finalize node's arguments.  Every
    step writes fresh device tensors (all kept alive, so no two share an
    address) and each equals the eager answer; the buffers of earlier steps
    keep their answers."""
    import torch

    n, d, k, B = 50000, 64, 100, 3
    base = data.mixture(n, d, 48, 0.3, 81, center_seed=82)
    Q = data.mixture(8 * B, d, 48, 0.3, 83, center_seed=82)
    print(f"[case] graph retarget mode={mode} n={n} d={d} B={B}", flush=True)
    kw = dict(nlist=64, nprobe=8) if mode == "ivf" else {}
    idx = Index(base, mode=mode, rerank=300, **kw)
    other = Index(base, mode="exact")
    ref_ids, ref_d = idx.search(Q, k)
    Qd = torch.from_numpy(Q).cuda()
    outs = []
    for i in range(8):
        ids = torch.full((B, k), -7, dtype=torch.int64, device="cuda")
        dd = torch.full((B, k), -7.0, dtype=torch.float32, device="cuda")
        got = idx.search_device(Qd[i * B:(i + 1) * B], k, out=(ids, dd))
        assert got[0].data_ptr() == ids.data_ptr()
        outs.append((ids, dd))
        # eager work in between: another index's search and torch work
        other.search(Q[:1 + i % 2], k)
        float((Qd * 2.0).sum())
    # the one shape seen here was captured once and replayed, retargeted, six
    # times (a failed retarget would have dropped the graph)
    assert idx.graph_count == 1
    for i, (ids, dd) in enumerate(outs):
        np.testing.assert_array_equal(ids.cpu().numpy(), ref_ids[i * B:(i + 1) * B], err_msg=f"step {i}")
        np.testing.assert_array_equal(dd.cpu().numpy(), ref_d[i * B:(i + 1) * B], err_msg=f"step {i}")


@pytest.mark.parametrize("nprobe", [4, 16])
def test_ivf_steps_are_captured_and_replayed(nprobe):
    """IVF steps are planned on the device (probe, queries per list, work
    items), so they capture like flat steps: the second sighting of a shape
    is captured and later steps replay it.  Every replay gives the eager
    answer, with eager work of other shapes and host synchronisations in
    between, including a query whose probed lists hold fewer than k rows
    (it scans every list inside the captured step)."""
    n, d, B, k = 40000, 64, 6, 100
    print(f"[case] ivf graphs nprobe={nprobe}", flush=True)
    kw = dict(center_seed=8000, normalize=False, zipf_s=1.1)
    base = data.mixture(n, d, 96, 0.3, 8001, **kw)
    Q = data.mixture(6 * B, d, 96, 0.3, 8002, **kw)
    idx = Index(base, mode="ivf", nlist=128, kmeans_iters=5, rerank=400)
    ref_ids, ref_d = idx.search(Q, k, nprobe=nprobe)
    assert np.all(ref_ids >= 0)
    other = data.mixture(9, d, 96, 0.3, 8003, **kw)
    for i in range(6):
        ids, dd = idx.search(Q[i * B:(i + 1) * B], k, nprobe=nprobe)  # 1st eager, 2nd captured, then replays
        np.testing.assert_array_equal(ids, ref_ids[i * B:(i + 1) * B], err_msg=f"step {i}")
        np.testing.assert_array_equal(dd, ref_d[i * B:(i + 1) * B], err_msg=f"step {i}")
        idx.search(other[: 1 + i % 4], k, nprobe=nprobe)
    assert idx.graph_count >= 1
    # the probed-lists answer of the library equals brute force over the same lists
    perm, off, cent = idx.export("perm"), idx.export("list_offsets"), idx.export("centroids")
    full = Index(base, mode="ivf", nlist=128, kmeans_iters=5, rerank=2048)
    f_ids, _ = full.search(Q[:8], k, nprobe=nprobe)
    want_ids, _ = O.ivf_search_exact(base, perm, off, cent, Q[:8], k, nprobe)
    for i in range(8):
        probes = O.ivf_probe_exact(cent, Q[i:i + 1], nprobe)[0]
        if sum(int(off[l + 1] - off[l]) for l in probes) <= 2048:
            np.testing.assert_array_equal(f_ids[i], want_ids[i])

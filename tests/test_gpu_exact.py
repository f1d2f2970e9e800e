"""Exact FP32 scan on the GPU vs the CPU oracle (ids bit-exact, distances
within 1e-5 relative of FP64) and vs the reference's golden ids."""
import os

import numpy as np
import pytest

from annvedi_amd import Algorithm, Index, data
from oracle import oracle as O
from tests.golden.make_golden import LINEAR_CASES, case_inputs, dup_case

pytestmark = pytest.mark.gpu
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "linear_ref.npz"))
REL_TOL = 1e-5


def check_exact(base, Q, k, ids, dd):
    ref_ids, ref_d = O.exact_topk_f64(base, Q, k)
    kk = ref_ids.shape[1]
    np.testing.assert_array_equal(ids[:, :kk], ref_ids)
    rel = np.abs(dd[:, :kk].astype(np.float64) - ref_d) / np.maximum(ref_d, 1e-30)
    assert rel.max() <= REL_TOL, rel.max()
    # reported fp32 distances are the reference's own fp32 values (src/simd.h order)
    for i in range(0, Q.shape[0], max(1, Q.shape[0] // 8)):
        for j in range(0, kk, 7):
            assert dd[i, j] == O.l2_linear_index(base[ids[i, j]], Q[i]), (i, j)
    if kk < ids.shape[1]:
        assert np.all(ids[:, kk:] == -1)


@pytest.mark.parametrize("case", LINEAR_CASES, ids=[c[0] for c in LINEAR_CASES])
def test_exact_matches_oracle_and_reference(case):
    name, n, d, C, sigma, norm, nq, k, off = case
    base, Q = case_inputs(n, d, C, sigma, norm, nq, off)
    print(f"[case] exact {name} n={n} d={d} nq={nq} k={k}", flush=True)
    idx = Index(base, mode="exact")
    ids, dd = idx.search(Q, k)
    check_exact(base, Q, k, ids, dd)
    # the reference's ranks 1..min(k,64) are exact (64-slot threshold) -> equal
    gold = GOLD[f"{name}__ids"]
    m = min(64, gold.shape[1])
    np.testing.assert_array_equal(ids[:, :m], gold[:, :m])
    assert idx.n_distances == n * nq  # src/linear_index.cpp:108


def test_duplicates_ties_lower_id_first():
    base, Q = dup_case()
    print("[case] exact duplicates", flush=True)
    idx = Index(base, mode="exact")
    ids, dd = idx.search(Q, 30)
    check_exact(base, Q, 30, ids, dd)
    np.testing.assert_array_equal(ids[0, :3], [0, 79, 80])
    np.testing.assert_array_equal(ids, GOLD["dups__ids"])  # k=30 < 64: all ranks exact in the reference


def test_reference_invariants_through_algorithm():
    """tests/test_logic.py:12-42 of the reference, on the GPU backend."""
    rng = np.random.Generator(np.random.PCG64(3))
    data50 = rng.random((50, 16), dtype=np.float32)
    a = Algorithm()
    a.fit(data50)
    assert len(a.query(data50[0], 5)) == 5
    assert a.query(data50[10], 1)[0] == 10
    d20 = rng.random((20, 8), dtype=np.float32)
    b = Algorithm()
    b.fit(d20)
    b.query(d20[0], 3)
    assert b.get_n_distances() == 20


def test_batch_invariance_alone_vs_mixed():
    base, Q = data.workload("linear-cpu-100k", n=30000, n_queries=300)
    idx = Index(base, mode="exact")
    ids_all, d_all = idx.search(Q, 100)
    for i in (0, 17, 299):
        ids1, d1 = idx.search(Q[i:i + 1], 100)
        np.testing.assert_array_equal(ids1[0], ids_all[i])
        np.testing.assert_array_equal(d1[0], d_all[i])
    ids_s, d_s = idx.search(Q[5:70], 100)
    np.testing.assert_array_equal(ids_s, ids_all[5:70])


def test_device_tensors_and_borrowed_base():
    import torch

    base, Q = data.workload("linear-cpu-100k", n=25000, n_queries=40)
    idx = Index(torch.from_numpy(base).cuda(), mode="exact")
    ids, dd = idx.search_device(torch.from_numpy(Q).cuda(), 100)
    check_exact(base, Q, 100, ids.cpu().numpy(), dd.cpu().numpy())


@pytest.fixture(scope="module")
def c2_base():
    """The 1M x 384 plain-fp32 workload (1024 clusters, sigma 0.3, unit
    rows), built in HBM, and its host copy for the oracle."""
    import torch

    base = data.mixture_torch(1_000_000, 384, 1024, 0.3, 2001, center_seed=1001)
    yield base, base.cpu().numpy()
    del base
    torch.cuda.empty_cache()


@pytest.mark.parametrize("B", [1, 64, 1024, 10000])
def test_exact_one_million_rows(c2_base, B):
    """Full C2 base size (1M x 384) at the C2 batch sizes (10000 queries run
    as three steps of the 4096-query workspace); oracle on a sample of each
    batch, size-independent properties on every query."""
    import torch

    print(f"[case] exact 1M x 384 B={B}", flush=True)
    base, hb = c2_base
    Q = data.mixture_torch(B, 384, 1024, 0.3, 3001 + B, center_seed=1001)
    idx = Index(base, mode="exact")
    ids, dd = idx.search_device(Q, 100)
    ids, dd = ids.cpu().numpy(), dd.cpu().numpy()
    hq = Q.cpu().numpy()
    sel = sorted({0, B // 2, B - 1, min(B - 1, 4095), min(B - 1, 4096), min(B - 1, 8193)})
    check_exact(hb, hq[sel], 100, ids[sel], dd[sel])
    # size-independent properties on every query: ranked by (fp64, id) with
    # the reference's fp32 values reported, so a reported value may step
    # down by a rounding error only where the fp64 order says so
    steps = np.argwhere(np.diff(dd, axis=1) < 0)
    print(f"fp32 values stepping down inside a near-tie run: {len(steps)}")
    for i, j in steps:
        a, b = O.l2_f64(hb[ids[i, j]], hq[i]), O.l2_f64(hb[ids[i, j + 1]], hq[i])
        assert a < b or (a == b and ids[i, j] < ids[i, j + 1]), (i, j)
        assert dd[i, j] - dd[i, j + 1] <= 1e-6 * dd[i, j]
    assert np.all((ids >= 0) & (ids < 1_000_000))
    assert all(len(set(r)) == 100 for r in ids)
    if B == 64:
        # self query: a base row is its own nearest neighbour at distance 0
        ids2, dd2 = idx.search_device(base[[5, 999_999]], 10)
        assert ids2[:, 0].tolist() == [5, 999_999] and float(dd2[:, 0].abs().max()) == 0.0
    del idx
    torch.cuda.empty_cache()


def test_large_batch_kernel_matches_small_batch_and_oracle():
    """B >= 33 over >= 65536-row ranges runs the query-stationary kernel;
    B = 1 runs the streaming kernel.  Both must give the exact answer."""
    base, Q = data.workload("exact-1m-384", n=150_000, n_queries=300)
    idx = Index(base, mode="exact")
    ids, dd = idx.search(Q, 100)
    sel = list(range(0, 300, 19))
    check_exact(base, Q[sel], 100, ids[sel], dd[sel])
    for i in (0, 128, 299):
        i1, d1 = idx.search(Q[i:i + 1], 100)
        np.testing.assert_array_equal(i1[0], ids[i])
        np.testing.assert_array_equal(d1[0], dd[i])


@pytest.mark.parametrize("qcap", ["1024", "0"])
def test_large_batch_kernel_candidate_queue_paths(qcap, monkeypatch):
    """The wide kernel stages passing pairs in a per-wave LDS queue.  With the
    queue capped at 1024 entries it is flushed many times per launch; at 0
    the kernel appends straight to the lists.  Both give the answer of the
    default queue size, which is the exact one."""
    base, Q = data.workload("exact-1m-384", n=150_000, n_queries=256)
    print(f"[case] exact wide queue cap={qcap} n=150000 B=256", flush=True)
    idx = Index(base, mode="exact")
    ids0, dd0 = idx.search(Q, 100)
    monkeypatch.setenv("ANNMI_WIDE_QCAP", qcap)
    ids, dd = idx.search(Q, 100)
    np.testing.assert_array_equal(ids, ids0)
    np.testing.assert_array_equal(dd, dd0)
    check_exact(base, Q[::51], 100, ids[::51], dd[::51])
    bf = Index(base, mode="bf16", rerank=300)
    monkeypatch.delenv("ANNMI_WIDE_QCAP")
    ib0, db0 = bf.search(Q, 100)
    monkeypatch.setenv("ANNMI_WIDE_QCAP", qcap)
    ib, db = bf.search(Q, 100)
    np.testing.assert_array_equal(ib, ib0)
    np.testing.assert_array_equal(db, db0)


@pytest.mark.parametrize("mode", ["exact", "bf16", "sq8", "sq4"])
def test_large_batch_kernel_prefilter_modes(mode, monkeypatch):
    """The wide kernel's epilogue prefilter (per-row bound; one bound per
    query when row norms are nearly equal; for integer dots one bound per
    query on dot minus each row's scaled norm excess) only decides which
    pairs take the exact filter test: every mode gives identical answers."""
    base, Q = data.workload("exact-1m-384", n=150_000, n_queries=256)
    print(f"[case] wide prefilter modes {mode} n=150000 B=256", flush=True)
    out = []
    for pf in ("1", "2", "3"):
        monkeypatch.setenv("ANNMI_PREFILTER", pf)
        idx = Index(base, mode=mode, rerank=300)
        out.append(idx.search(Q, 100))
        del idx
    for o in out[1:]:
        np.testing.assert_array_equal(out[0][0], o[0])
        np.testing.assert_array_equal(out[0][1], o[1])
    if mode == "exact":
        check_exact(base, Q[::64], 100, out[1][0][::64], out[1][1][::64])

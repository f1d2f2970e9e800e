"""IVF build and list scan on the GPU vs the oracle."""
import numpy as np
import pytest

from annvedi_amd import Algorithm, Index, data
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def zipf_data(n, d, C, nq, seed=0):
    kw = dict(center_seed=7000 + seed, normalize=False, zipf_s=1.1)
    return data.mixture(n, d, C, 0.3, 7100 + seed, **kw), data.mixture(nq, d, C, 0.3, 7200 + seed, **kw)


def lists_of(idx):
    perm = idx.export("perm")
    off = idx.export("list_offsets")
    return perm, off


def test_kmeans_first_round_follows_reference_init():
    """One k-means round: lists = argmin over the initial centroids, which
    are the rows shuffle(iota(n), mt19937(1337))[c] (lsh_index_optimized.cpp:
    553-561); centroids afterwards = list means."""
    base, _ = zipf_data(20000, 64, 48, 1)
    nlist = 40
    idx = Index(base, mode="ivf", nlist=nlist, kmeans_iters=1)
    perm, off = lists_of(idx)
    assert off[0] == 0 and off[-1] == 20000 and np.all(np.diff(off) >= 0)
    assert sorted(perm.tolist()) == list(range(20000))
    for l in range(nlist):
        assert np.all(np.diff(perm[off[l]:off[l + 1]].astype(np.int64)) > 0)  # ascending ids in a list
    init, ri = O.kmeans_init(20000, nlist)
    C0 = base[init].astype(np.float64)
    X = base.astype(np.float64)
    d2 = ((X[:, None, :] - C0[None, :, :]) ** 2).sum(-1)
    want = d2.argmin(1)
    srt = np.sort(d2, axis=1)
    near_tie = (srt[:, 1] - srt[:, 0]) <= 1e-5 * srt[:, 1]
    got = np.empty(20000, dtype=np.int64)
    for l in range(nlist):
        got[perm[off[l]:off[l + 1]]] = l
    assert np.all((got == want) | near_tie)
    cent = idx.export("centroids")
    for l in range(nlist):
        rows = perm[off[l]:off[l + 1]]
        if len(rows):
            np.testing.assert_allclose(cent[l], X[rows].mean(0), rtol=1e-5, atol=1e-6)
        else:
            np.testing.assert_array_equal(cent[l], base[ri[(l * 37) % 20000]])


@pytest.mark.parametrize("d", [64, 768, 1100])
def test_ivf_full_rescore_equals_bruteforce_over_probed_lists(d):
    """d = 768: 12 code chunks a row (query fragments held in registers);
    d = 1100: 18 chunks (the list scan's chunk loop)."""
    print(f"[case] ivf brute force d={d}", flush=True)
    base, Q = zipf_data(16000, d, 32, 24, seed=1)
    idx = Index(base, mode="ivf", nlist=32, kmeans_iters=4)
    perm, off = lists_of(idx)
    cent = idx.export("centroids")
    for nprobe in (1, 3):
        ids, dd = idx.search(Q, 100, rerank=2048, nprobe=nprobe)
        want_ids, want_d = O.ivf_search_exact(base, perm, off, cent, Q, 100, nprobe)
        for i in range(Q.shape[0]):
            m = len(want_ids[i])
            if m <= 2048 and off is not None:
                probes = O.ivf_probe_exact(cent, Q[i:i + 1], nprobe)[0]
                tot = sum(int(off[l + 1] - off[l]) for l in probes)
                if tot <= 2048:
                    np.testing.assert_array_equal(ids[i, :m], want_ids[i])
                    np.testing.assert_allclose(dd[i, :m], want_d[i], rtol=1e-5)


def test_ivf_recall_rises_with_nprobe():
    base, Q = zipf_data(120000, 128, 256, 64, seed=2)
    ex_ids, _ = O.exact_topk_f64(base, Q, 100)
    idx = Index(base, mode="ivf", nlist=256, rerank=400)
    recs = []
    for nprobe in (4, 16, 64):
        ids, _ = idx.search(Q, 100, nprobe=nprobe)
        recs.append(O.recall_at_k(base, Q, ids, ex_ids, 100).mean())
    print("ivf recall@100 by nprobe 4/16/64:", recs)
    assert recs[0] <= recs[1] + 1e-9 <= recs[2] + 2e-9
    assert recs[2] >= 0.95
    nres, flags = idx.last_batch_stats(64)
    assert np.all(flags == 0)


def test_ivf_through_plugin_with_reference_param_names():
    base, Q = zipf_data(30000, 64, 64, 8, seed=3)
    a = Algorithm()
    a.fit(base, backend="mi355x", mode="ivf", n_clusters=64, n_tables=1, n_bits=0)
    res = a.query(Q[0], 100, n_probe_clusters=8, refine_r=400)
    assert res.shape == (100,) and len(set(res.tolist())) == 100
    assert a.query(base[123], 1, n_probe_clusters=8)[0] == 123
    ids, dd = a.batch_query(Q, 10, n_probe_clusters=8)
    assert ids.shape == (8, 10) and np.all(np.diff(dd, axis=1) >= 0)


def test_ivf_falls_back_to_every_row_when_probed_lists_are_short():
    """nprobe = 1 on small lists: a query whose probed list holds fewer than
    k rows takes every row as a candidate (lsh_index_optimized.cpp:452-457),
    so no answer is padded with -1 while n >= k, and with every row
    re-scored (n <= R) the answer is the exact one."""
    print("[case] ivf fallback", flush=True)
    base, Q = zipf_data(2000, 48, 40, 30, seed=5)
    idx = Index(base, mode="ivf", nlist=40, kmeans_iters=6)
    perm, off = lists_of(idx)
    cent = idx.export("centroids")
    probes = O.ivf_probe_exact(cent, Q, 1)[:, 0]
    short = [int(off[l + 1] - off[l]) < 100 for l in probes]
    assert sum(short) >= 1  # the fall-back case occurs
    ids, dd = idx.search(Q, 100, rerank=2000, nprobe=1)
    assert np.all(ids >= 0)
    want_ids, want_d = O.ivf_search_exact(base, perm, off, cent, Q, 100, 1)
    ex_ids, _ = O.exact_topk_f64(base, Q, 100)
    for i in range(Q.shape[0]):
        np.testing.assert_array_equal(ids[i], want_ids[i])
        if short[i]:
            np.testing.assert_array_equal(ids[i], ex_ids[i])


def test_ivf_answers_equal_the_compiled_reference():
    """tests/golden/ivf_ref.npz holds the reference's own IVF answers
    (competitors/ivf_lsh/src/lsh_index_optimized.cpp built from its sources,
    n_tables 1, n_bits 0, refine_r = n) on 24 separated clusters, each seeded
    by one of the reference's initial centroid rows: the GPU build (same
    initial rows, 15 rounds) must find the same lists, probe the same ones,
    fall back to every row where they hold fewer than k, and rank as the
    reference does.  The reference ranks by its fp32 value and this library
    by (fp64, id), so a position may differ only between rows whose
    distances agree to fp32 rounding."""
    import os

    from tests.golden.make_golden import checksum
    from tests.golden.make_ivf_golden import CASES, K, NPROBES

    print("[case] ivf vs compiled reference", flush=True)
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "ivf_ref.npz"))
    base, Q, nlist = CASES["clusters"]()
    assert checksum(base, Q) == str(gold["clusters__checksum"])
    want = gold["clusters__ids"]
    idx = Index(base, mode="ivf", nlist=nlist)  # reference defaults: 15 rounds, seed 1337
    n_equal = 0
    for j, p in enumerate(NPROBES):
        ids, dd = idx.search(Q, K, rerank=2048, nprobe=p)
        assert np.all(ids >= 0)
        for i in range(Q.shape[0]):
            if np.array_equal(ids[i], want[i, j]):
                n_equal += 1
                continue
            d_got = np.array([O.l2_f64(base[r], Q[i]) for r in ids[i]])
            d_want = np.array([O.l2_f64(base[r], Q[i]) for r in want[i, j]])
            np.testing.assert_allclose(d_got, d_want, rtol=2e-6, err_msg=f"query {i} nprobe {p}")
    print(f"identical answers: {n_equal} of {Q.shape[0] * len(NPROBES)}")
    assert n_equal >= 0.9 * Q.shape[0] * len(NPROBES)

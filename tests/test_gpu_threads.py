"""Searches from several host threads at once.  Each index owns its stream
and serialises calls on itself; the launch-timing events, the kernel-name
record and the error text are per thread, and graph capture is thread-local.
So concurrent searches on different indexes, and on one shared index, must
give exactly the answers the same searches give one at a time, including the
steps that are captured and replayed as graphs while other threads run."""
import threading

import numpy as np
import pytest

from annvedi_amd import Index, data

pytestmark = pytest.mark.gpu


def _run_threads(jobs, timeout=90.0):
    errs = []

    def wrap(fn):
        def body():
            try:
                fn()
            except BaseException as e:  # re-raised on the main thread
                errs.append(e)
        return body

    ts = [threading.Thread(target=wrap(fn), daemon=True) for fn in jobs]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout)
        assert not t.is_alive(), "search thread did not finish"
    if errs:
        raise errs[0]


def _batches(Q):
    # shapes repeat, so each thread's second and later steps of a shape run
    # as captured graphs
    cuts = [(0, 5), (5, 12), (12, 17), (17, 24), (24, 29), (29, 36), (36, 41)]
    return [Q[a:b] for a, b in cuts]


def test_concurrent_searches_on_separate_indexes():
    d, k = 128, 100
    specs = [("exact", 0), ("sq8", 0), ("bf16", 0), ("ivf", 16)]
    idxs, Qs = [], []
    for j, (mode, nprobe) in enumerate(specs):
        base = data.mixture(50000, d, 64, 0.3, 900 + j, center_seed=950 + j)
        Q = data.mixture(41, d, 64, 0.3, 910 + j, center_seed=950 + j)
        kw = dict(nlist=64) if mode == "ivf" else {}
        idxs.append(Index(base, mode=mode, rerank=200, **kw))
        Qs.append(Q)
    print("[case] threads: separate indexes exact/sq8/bf16/ivf n=50000 d=128", flush=True)
    # serial answers, on fresh indexes' first sight of each shape
    ref = [[idx.search(q, k, nprobe=npb) for q in _batches(Q)]
           for idx, Q, (_, npb) in zip(idxs, Qs, specs)]
    got = [[None] * 3 * len(_batches(Q)) for Q in Qs]

    def job(j):
        def fn():
            idx, Q, npb = idxs[j], Qs[j], specs[j][1]
            for rep in range(3):
                for i, q in enumerate(_batches(Q)):
                    got[j][rep * len(_batches(Q)) + i] = idx.search(q, k, nprobe=npb)
        return fn

    _run_threads([job(j) for j in range(len(specs))])
    for j, (mode, _) in enumerate(specs):
        nb = len(ref[j])
        for s, (ids, dd) in enumerate(got[j]):
            r_ids, r_d = ref[j][s % nb]
            np.testing.assert_array_equal(ids, r_ids, err_msg=f"{mode} step {s}")
            np.testing.assert_array_equal(dd, r_d, err_msg=f"{mode} step {s}")
    assert any(idx.graph_count >= 1 for idx in idxs)


def test_concurrent_searches_on_one_index():
    d, k = 96, 50
    base = data.mixture(40000, d, 32, 0.3, 960, center_seed=961)
    Q = data.mixture(41, d, 32, 0.3, 962, center_seed=961)
    idx = Index(base, mode="sq8", rerank=150)
    print("[case] threads: one sq8 index n=40000 d=96, 4 threads", flush=True)
    bs = _batches(Q)
    ref = [idx.search(q, k) for q in bs]
    got = {}

    def job(t):
        def fn():
            for rep in range(2):
                for i in range(len(bs)):
                    s = (i + t) % len(bs)  # threads walk the batches out of phase
                    got[(t, rep, s)] = idx.search(bs[s], k)
        return fn

    _run_threads([job(t) for t in range(4)])
    assert len(got) == 4 * 2 * len(bs)
    for (t, rep, s), (ids, dd) in got.items():
        np.testing.assert_array_equal(ids, ref[s][0], err_msg=f"thread {t} rep {rep} batch {s}")
        np.testing.assert_array_equal(dd, ref[s][1], err_msg=f"thread {t} rep {rep} batch {s}")

"""Scheduling logic of the continuous-batching server (host only: a stand-in
search function replaces the GPU call)."""
import threading
import time

import numpy as np
import pytest

from annvedi_amd.serving import SLOT_BYTES_PER_QUERY, Server, lognormal_sizes, run_open_loop


class FakeIndex:
    d = 4


def fake_search_factory(delay=0.0, log=None):
    def search(h, Q, k):
        if log is not None:
            log.append(Q.shape[0])
        time.sleep(delay)
        # ids encode the query content: row i -> [Q[i,0]*1000 + j]
        ids = (Q[:, :1] * 1000).astype(np.int64) + np.arange(k)[None, :]
        return ids, np.zeros((Q.shape[0], k), np.float32)
    return search


def test_results_return_to_their_requests():
    srv = Server(FakeIndex(), workers=3, max_batch=64, search=fake_search_factory())
    futs = []
    for i in range(50):
        q = np.full((1 + i % 5, 4), float(i), np.float32)
        futs.append(srv.submit(q, 7))
    for i, f in enumerate(futs):
        ids, dd = f.result(timeout=10)
        assert ids.shape == (1 + i % 5, 7)
        assert np.all(ids[:, 0] == i * 1000)
    srv.close()


def test_requests_arriving_during_a_batch_form_the_next_batch():
    log = []
    srv = Server(FakeIndex(), workers=1, max_batch=100, search=fake_search_factory(0.2, log))
    f0 = srv.submit(np.zeros((3, 4), np.float32), 5)
    time.sleep(0.05)  # the worker is busy with the first request
    fs = [srv.submit(np.ones((2, 4), np.float32), 5) for _ in range(6)]
    for f in [f0] + fs:
        f.result(timeout=10)
    srv.close()
    assert log == [3, 12]  # no fixed window: batch 1 alone, then everything that arrived


def test_slot_capacity_and_fifo():
    log = []
    srv = Server(FakeIndex(), workers=1, max_batch=10, search=fake_search_factory(0.1, log))
    srv.submit(np.zeros((1, 4), np.float32), 5)
    time.sleep(0.03)
    futs = [srv.submit(np.zeros((b, 4), np.float32), 5) for b in (4, 4, 4, 9, 1)]
    for f in futs:
        f.result(timeout=10)
    srv.close()
    assert log == [1, 8, 4, 9, 1] or log == [1, 8, 4, 10]
    with pytest.raises(ValueError):
        Server(FakeIndex(), workers=1, max_batch=2, search=fake_search_factory()).submit(np.zeros((3, 4)), 1)


def test_hbm_budget_caps_concurrent_slots():
    srv = Server(FakeIndex(), workers=8, max_batch=256, hbm_budget=3 * 256 * SLOT_BYTES_PER_QUERY + 1,
                 search=fake_search_factory())
    assert len(srv.handles) == 3
    srv.close()


def test_open_loop_poisson_accounting():
    srv = Server(FakeIndex(), workers=2, max_batch=512, search=fake_search_factory(0.002))
    sizes = lognormal_sizes(200)
    assert sizes.min() >= 1 and sizes.max() <= 512 and 4 <= np.median(sizes) <= 12
    Q = np.repeat(np.arange(int(sizes.sum()), dtype=np.float32)[:, None], 4, axis=1)
    res = run_open_loop(srv, Q, sizes, qps=20000, k=3)
    srv.close()
    assert res["queries"] == int(sizes.sum())
    off = 0
    for b, (ids, _) in zip(sizes, res["answers"]):
        np.testing.assert_array_equal(ids[:, 0], np.arange(off, off + b) * 1000)
        off += b
    assert 0 < res["p50_ms"] <= res["p99_ms"]


class SizedIndex:
    d = 4
    nbytes = 10_000


def test_budget_below_the_index_or_one_slot_is_refused():
    S = 8 * SLOT_BYTES_PER_QUERY
    with pytest.raises(ValueError, match="below the index"):
        Server(SizedIndex(), workers=2, max_batch=8, hbm_budget=9_999, search=fake_search_factory())
    with pytest.raises(ValueError, match="one worker slot"):
        Server(SizedIndex(), workers=2, max_batch=8, hbm_budget=10_000 + S - 1, search=fake_search_factory())
    srv = Server(SizedIndex(), workers=5, max_batch=8, hbm_budget=10_000 + 2 * S, search=fake_search_factory())
    assert len(srv.handles) == 2  # the slots that fit next to the index
    with pytest.raises(ValueError):
        srv.set_budget(10_000 + S - 1)
    srv.close()


def test_live_budget_limits_the_requests_in_flight():
    """Two workers, then a budget of one slot: requests of 6 queries (two do
    not share an 8-query slot) run one at a time, and the bytes held never
    exceed the budget left by the index."""
    S = 8 * SLOT_BYTES_PER_QUERY
    active, peak, lock = [0], [0], threading.Lock()

    def search(h, Q, k):
        with lock:
            active[0] += 1
            peak[0] = max(peak[0], active[0])
        time.sleep(0.02)
        with lock:
            active[0] -= 1
        return np.zeros((Q.shape[0], k), np.int64), np.zeros((Q.shape[0], k), np.float32)

    srv = Server(SizedIndex(), workers=2, max_batch=8, hbm_budget=10_000 + 2 * S, search=search)
    futs = [srv.submit(np.zeros((6, 4), np.float32), 3) for _ in range(4)]
    for f in futs:
        f.result(timeout=10)
    assert peak[0] == 2  # two slots: two requests on the GPU at once
    srv.set_budget(10_000 + S)
    peak[0] = 0
    srv.peak_reserved = 0
    futs = [srv.submit(np.zeros((6, 4), np.float32), 3) for _ in range(6)]
    for f in futs:
        f.result(timeout=10)
    srv.close()
    assert peak[0] == 1
    assert srv.peak_reserved <= S

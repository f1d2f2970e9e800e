"""The harness-faithful runner (orthogonal-competition/harness.py:117-196,
evaluator.py:149-157, 485-511) on the host: scenario resolution, the
per-query loop and its checks, the metrics."""
import numpy as np
import pytest

from annvedi_amd import runner


class BruteForce:
    """A stand-in plugin (numpy brute force) for the loop's host logic."""

    def __init__(self, bad_at=None):
        self.n = 0
        self.bad_at = bad_at
        self.calls = 0

    def fit(self, train, **p):
        self.x = np.asarray(train, np.float32)
        self.p = p

    def query(self, q, k, **qp):
        self.calls += 1
        if self.bad_at is not None and self.calls - 1 == self.bad_at:
            return np.arange(k - 1)
        self.n += self.x.shape[0]
        return np.argsort(((self.x - q) ** 2).sum(1), kind="stable")[:k]

    def get_n_distances(self):
        return self.n


def test_scenario_resolution(tmp_path):
    p = tmp_path / "s.yaml"
    p.write_text("scenarios:\n  a:\n    default: {index_params: {mode: sq8}, query_params: {r: 5}}\n"
                 "    ds1: {index_params: {mode: exact}}\n  b: {}\n")
    assert runner.load_scenario("a", "ds1", str(p)) == {"index_params": {"mode": "exact"}, "query_params": {}}
    assert runner.load_scenario("a", "other", str(p)) == {"index_params": {"mode": "sq8"}, "query_params": {"r": 5}}
    with pytest.raises(ValueError, match="not found"):
        runner.load_scenario("c", "ds1", str(p))
    with pytest.raises(ValueError, match="no 'default'"):
        runner.load_scenario("b", "ds1", str(p))
    # the shipped scenario file resolves
    assert runner.load_scenario("high_recall", "any")["index_params"]["mode"] == "exact"


def test_loop_fields_and_metrics():
    rng = np.random.Generator(np.random.PCG64(3))
    train = rng.standard_normal((500, 8)).astype(np.float32)
    queries = rng.standard_normal((20, 8)).astype(np.float32)
    res = runner.run(BruteForce(), train, queries, 10, {"mode": "x"}, {})
    assert res["neighbors"].dtype == np.int32 and res["neighbors"].shape == (20, 10)
    assert res["query_times"].shape == (20,) and np.all(res["query_times"] > 0)
    assert res["n_dist_queries"] == 500 * 20 and res["build_time"] >= 0
    true_d = np.sort(np.linalg.norm(train[None] - queries[:, None], axis=2), axis=1)[:, :10]
    m = runner.evaluate(train, queries, true_d, res, 10)
    assert m["avg_recall"] == 1.0 and m["qps"] > 0
    assert m["latency_ms"]["median"] <= m["latency_ms"]["p99"] <= m["latency_ms"]["max"]
    # a wrong-length answer stops the run (harness.py:169-173)
    with pytest.raises(ValueError, match="length k=10"):
        runner.run(BruteForce(bad_at=3), train, queries, 10, {}, {})


def test_recall_counts_ties_at_the_kth_distance():
    true_d = np.array([[1.0, 2.0, 3.0]])
    assert runner.recalls(true_d, np.array([[1.0, 3.0, 3.0]]), 3)[0] == 1.0
    assert runner.recalls(true_d, np.array([[1.0, 3.0, 3.5]]), 3)[0] == pytest.approx(2 / 3)

"""The organisers' evaluation loop, run against this plugin.

Restates orthogonal-competition/harness.py:117-196 (one scenario: untimed
load, timed ``fit`` with the resident-set growth as index memory, every
query timed on its own through ``Algorithm.query``, int32 neighbours of
shape (n_queries, k), ``get_n_distances`` after the loop) and the metrics of
orthogonal-competition/evaluator.py:149-157, 485-511 (recall against the
true k-th distance with float32 Euclidean distances, QPS over the summed
query times, latency mean / median / p95 / p99 / p99.9 / max).

h5py is not part of this environment, so datasets and results are ``.npz``
files with the reference's dataset names: a dataset holds ``train``,
``test`` and (for the metrics) ``distances``; a result holds ``neighbors``,
``build_time``, ``query_times``, ``n_dist_queries`` and ``index_mem_mb``.
``index_mem_mb`` is host memory, as the harness measures it (the device
index is reported apart as ``device_index_mb``).

    DATASET_PATH=d.npz RESULTS_PATH=r.npz SCENARIO_NAME=fast DATASET_NAME=d \\
        QUERY_K=100 python -m annvedi_amd.runner
"""
from __future__ import annotations

import os
import resource
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SCENARIOS_FILE = os.path.join(HERE, "scenarios.yaml")


def load_scenario(scenario_name: str, dataset_name: str, path: str = SCENARIOS_FILE) -> dict:
    """index_params / query_params of a scenario: the dataset's own block,
    else the ``default`` block (harness.py:72-114)."""
    import yaml

    with open(path) as f:
        raw = yaml.safe_load(f)
    if "scenarios" not in raw or not isinstance(raw["scenarios"], dict):
        raise ValueError("scenarios.yaml must contain a top-level 'scenarios' mapping.")
    if scenario_name not in raw["scenarios"]:
        raise ValueError(f"Scenario {scenario_name!r} not found in scenarios.yaml. "
                         f"Available: {', '.join(raw['scenarios'])}")
    cfg = raw["scenarios"][scenario_name] or {}
    if dataset_name in cfg:
        block = cfg[dataset_name]
    elif "default" in cfg:
        block = cfg["default"]
    else:
        raise ValueError(f"Scenario {scenario_name!r} has no entry for dataset {dataset_name!r} "
                         "and no 'default' fallback.")
    block = block or {}
    return {"index_params": dict(block.get("index_params") or {}),
            "query_params": dict(block.get("query_params") or {})}


def run(algo, train: np.ndarray, queries: np.ndarray, k: int, index_params: dict, query_params: dict) -> dict:
    """One scenario: the result fields of harness.py:189-194."""
    start_mem_kb = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss
    t0 = time.perf_counter()
    algo.fit(train, **index_params)
    build_time = time.perf_counter() - t0
    index_mem_mb = (resource.getrusage(resource.RUSAGE_SELF).ru_maxrss - start_mem_kb) / 1024
    n = queries.shape[0]
    neighbors = np.empty((n, k), dtype=np.int32)
    query_times = np.empty(n, dtype=np.float64)
    for i in range(n):
        t0 = time.perf_counter()
        result = algo.query(queries[i], k=k, **query_params)
        query_times[i] = time.perf_counter() - t0
        result = np.asarray(result, dtype=np.int32)
        if result.shape != (k,):
            raise ValueError(f"algo.query() must return a 1-D array of length k={k}, "
                             f"got shape {result.shape} on query index {i}.")
        neighbors[i] = result
    out = {"neighbors": neighbors, "build_time": float(build_time), "query_times": query_times,
           "n_dist_queries": int(algo.get_n_distances()), "index_mem_mb": float(index_mem_mb)}
    idx = getattr(algo, "_index", None)
    if idx is not None and hasattr(idx, "nbytes"):
        out["device_index_mb"] = idx.nbytes / 2 ** 20
    return out


def recalls(true_distances: np.ndarray, predicted_distances: np.ndarray, k: int) -> np.ndarray:
    """evaluator.py:149-157: per query, the share of the first k predicted
    distances at or below the true k-th distance."""
    return np.array([np.mean(predicted_distances[i][:k] <= true_distances[i][k - 1])
                     for i in range(true_distances.shape[0])])


def evaluate(train: np.ndarray, queries: np.ndarray, true_distances: np.ndarray, result: dict, k: int) -> dict:
    """evaluator.py:450-511: float32 Euclidean distances of the predicted
    rows, recall, QPS over the summed per-query times, latency statistics."""
    nb = result["neighbors"]
    n = queries.shape[0]
    if nb.shape[0] != n or result["query_times"].shape[0] != n:
        raise ValueError(f"result holds {nb.shape[0]} answers and {result['query_times'].shape[0]} times "
                         f"for {n} queries")
    pred = np.array([np.linalg.norm(train[nb[i]] - queries[i], axis=1) for i in range(n)])
    rec = recalls(true_distances, pred, k)
    total = float(result["query_times"].sum())
    lat = result["query_times"] * 1e3
    return {"build_time_s": result["build_time"], "total_query_time_s": total, "qps": n / total,
            "n_dist_queries": result["n_dist_queries"], "index_mem_mb": result["index_mem_mb"],
            "avg_recall": float(rec.mean()),
            "latency_ms": {"mean": float(np.mean(lat)), "median": float(np.median(lat)),
                           "p95": float(np.percentile(lat, 95)), "p99": float(np.percentile(lat, 99)),
                           "p999": float(np.percentile(lat, 99.9)), "max": float(np.max(lat))}}


def main() -> int:
    from .algorithm import Algorithm

    dataset_path = os.environ["DATASET_PATH"]
    results_path = os.environ["RESULTS_PATH"]
    scenario = os.environ["SCENARIO_NAME"]
    dataset = os.environ["DATASET_NAME"]
    k = int(os.environ.get("QUERY_K", 10))
    with np.load(dataset_path) as f:
        train = np.ascontiguousarray(f["train"], dtype=np.float32)
        queries = np.ascontiguousarray(f["test"], dtype=np.float32)
    cfg = load_scenario(scenario, dataset, os.environ.get("SCENARIOS_FILE", SCENARIOS_FILE))
    print(f"[runner] {train.shape[0]} train | {queries.shape[0]} queries | dim={train.shape[1]} | k={k} | "
          f"scenario {scenario} {cfg}", flush=True)
    res = run(Algorithm(), train, queries, k, cfg["index_params"], cfg["query_params"])
    t = res["query_times"]
    print(f"[runner] build {res['build_time']:.3f}s; queries total={t.sum():.4f}s QPS={len(t) / t.sum():.1f} "
          f"median={np.median(t) * 1e3:.3f}ms p99={np.percentile(t, 99) * 1e3:.3f}ms "
          f"n_dist_queries={res['n_dist_queries']}", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(results_path)), exist_ok=True)
    np.savez(results_path, **res)
    print(f"[runner] results written to {results_path}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

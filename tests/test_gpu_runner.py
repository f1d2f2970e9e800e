"""The harness-faithful runner end to end on the GPU: environment variables,
the shipped scenarios, a .npz dataset and result file."""
import numpy as np
import pytest

from annvedi_amd import data, runner
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scenario", ["high_recall", "fast", "memory"])
def test_runner_main_with_the_shipped_scenarios(tmp_path, monkeypatch, scenario):
    print(f"[case] runner {scenario}", flush=True)
    train, test = data.workload("exact-1m-384", n=30000, n_queries=40)
    ids, _ = O.exact_topk_f64(train, test, 100)
    # ground-truth distances as the reference's dataset preparation writes
    # them: float32 np.linalg.norm of the neighbours (prepare_data.py:39-51)
    gt_d = np.stack([np.linalg.norm(train[ids[i]] - test[i], axis=1) for i in range(test.shape[0])])
    ds = tmp_path / "mix30k.npz"
    np.savez(ds, train=train, test=test, neighbors=ids.astype(np.int32), distances=gt_d.astype(np.float32))
    out = tmp_path / "res" / "r.npz"
    for k, v in dict(DATASET_PATH=str(ds), RESULTS_PATH=str(out), SCENARIO_NAME=scenario,
                     DATASET_NAME="mix30k", QUERY_K="100").items():
        monkeypatch.setenv(k, v)
    assert runner.main() == 0
    with np.load(out) as r:
        res = {k: r[k] for k in r.files}
    assert res["neighbors"].dtype == np.int32 and res["neighbors"].shape == (40, 100)
    assert res["query_times"].shape == (40,)
    with np.load(ds) as f:
        m = runner.evaluate(f["train"], f["test"], f["distances"], res, 100)
    print(scenario, m)
    if scenario == "high_recall":
        assert m["avg_recall"] == 1.0
        assert int(res["n_dist_queries"]) == 30000 * 40  # N per query (src/linear_index.cpp:108)
    else:
        assert m["avg_recall"] >= 0.99

"""bench.py's contract pieces that need no GPU."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_batch_mix_is_fixed_and_log_uniform():
    a, b = bench.batch_sizes(23), bench.batch_sizes(23)
    np.testing.assert_array_equal(a, b)
    assert a.min() >= 1 and a.max() <= 4096
    big = bench.batch_sizes(20000, seed=7)
    # log-uniform: each octave holds about the same share
    counts = np.histogram(np.log2(big), bins=12, range=(0, 12))[0]
    assert counts.min() > 0.6 * counts.mean()


def test_bench_script_names_only_defined_names():
    """A name bench.py loads but never binds (a renamed variable) shows up
    only when the GPU run reaches that line; catch it here."""
    import ast
    import builtins

    tree = ast.parse(open(os.path.join(ROOT, "bench.py")).read())
    bound = {"__file__", "__name__"}
    for n in ast.walk(tree):
        if isinstance(n, ast.Name) and isinstance(n.ctx, (ast.Store, ast.Del)):
            bound.add(n.id)
        elif isinstance(n, (ast.FunctionDef, ast.ClassDef)):
            bound.add(n.name)
        elif isinstance(n, ast.arg):
            bound.add(n.arg)
        elif isinstance(n, (ast.Import, ast.ImportFrom)):
            bound.update((a.asname or a.name).split(".")[0] for a in n.names)
        elif isinstance(n, ast.ExceptHandler) and n.name:
            bound.add(n.name)
    used = {n.id for n in ast.walk(tree) if isinstance(n, ast.Name) and isinstance(n.ctx, ast.Load)}
    assert sorted(u for u in used - bound if not hasattr(builtins, u)) == []


def test_workload_shapes_default_to_the_named_workload():
    assert bench.workload_shape("mixed-10m-384") == (10_000_000, 384)
    assert bench.workload_shape("ivf-10m-768") == (10_000_000, 768)
    assert bench.workload_shape("sq-1m-768", 5000, 0) == (5000, 768)
    assert bench.workload_shape("exact-1m-384", 0, 96) == (1_000_000, 96)


def test_per_launch_roofline():
    # one HBM-bound launch at exactly the HBM peak, one MFMA-bound at half
    ms = [1.0, 2.0]
    by = [bench.HBM_PEAK_GBS * 1e9 * 1e-3, 1.0]
    ops = [1.0, 5000e12 * 1e-3]
    frac, nh = bench.per_launch_roofline(ms, by, ops, 5000.0)
    assert nh == 1 and abs(frac - 2.0 / 3.0) < 1e-4
    assert bench.per_launch_roofline([], [], [], 5000.0) is None


def test_traffic_profiles_are_consistent_and_scaled_by_algorithmic_bytes():
    for mode in ("sq8", "exact"):
        path = os.path.join(ROOT, "profiles", f"traffic_mixed-10m-384_{mode}.json")
        t = json.load(open(path))
        cal = t["calibration"]
        assert abs(t["fetch_bytes_per_launch"] - t["fetch_kib_per_launch"] * 1024 * cal["fetch_scale"]) < 1e-4 * \
            t["fetch_bytes_per_launch"]
        assert abs(t["write_bytes_per_launch"] - t["write_kib_per_launch"] * 1024 * cal["write_scale"]) < 1e-4 * \
            max(t["write_bytes_per_launch"], 1)
        assert abs(t["fetch_ratio"] - t["fetch_bytes_per_launch"] / t["fetch_run_algorithmic_bytes_per_launch"]) < 1e-4
        assert abs(t["write_ratio"] - t["write_bytes_per_launch"] / t["write_run_algorithmic_bytes_per_launch"]) < 1e-4
        assert 0.9 < t["fetch_ratio"] < 2.0  # each row read at least about once, and not wildly more
        assert os.path.exists(os.path.join(ROOT, t["summary"]))
        alg = t["fetch_run_algorithmic_bytes_per_launch"]
        hit = bench.measured_traffic("mixed-10m-384", mode, t["kernel"], alg)
        want = (t["fetch_ratio"] + t["write_ratio"]) * alg
        assert abs(hit["bytes_per_launch"] - want) <= 1
        # scales with this run's algorithmic bytes; another kernel gets none
        assert abs(bench.measured_traffic("mixed-10m-384", mode, t["kernel"], 2 * alg)["bytes_per_launch"] - 2 * want) <= 2
        assert bench.measured_traffic("mixed-10m-384", mode, "another_kernel", alg) is None

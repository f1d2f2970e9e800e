"""The library's per-launch timing (annmi_profile / annmi_profile_kernels),
which bench.py's roofline is computed from: counts, durations and
algorithmic bytes per launch are consistent, timing changes no answer and
captures no step graph, and the IVF step is timed too."""
import numpy as np
import pytest

from annvedi_amd import Index, data

pytestmark = pytest.mark.gpu


def test_flat_step_timing_is_consistent_and_changes_nothing():
    base, Q = data.workload("mixed-10m-384", n=200_000, n_queries=300)
    print("[case] profile flat sq8 n=200000", flush=True)
    idx = Index(base, mode="sq8", rerank=128)
    ref = [idx.search(Q[:b], 100) for b in (1, 37, 300)]
    idx.profile(1)
    got = [idx.search(Q[:b], 100) for b in (1, 37, 300)]
    assert idx.graph_count == 0  # timed steps run eagerly
    ks = idx.profile_kernels()
    pr = idx.profile(0)
    for (i0, d0), (i1, d1) in zip(ref, got):
        np.testing.assert_array_equal(i0, i1)
        np.testing.assert_array_equal(d0, d1)
    assert pr["steps"] == 3
    assert pr["scan_launches"] == sum(k["launches"] for k in ks) > 0
    assert pr["select_ms"] > 0 and pr["finalize_ms"] > 0 and pr["prep_ms"] > 0
    assert abs(pr["scan_ms"] - sum(k["ms"] for k in ks)) < 1e-6 * max(pr["scan_ms"], 1.0) + 1e-9
    total_rows = 0.0
    for k in ks:
        assert k["ms"] > 0 and len(k["launch_ms"]) == k["launches"]
        assert all(t > 0 for t in k["launch_ms"])
        assert abs(sum(k["launch_ms"]) - k["ms"]) < 1e-3 * k["ms"] + 1e-6
        assert abs(sum(k["launch_bytes"]) - k["bytes"]) < 1e-3 * k["bytes"]
        # int8 MFMA work: 2 * B * rows * d per launch, bytes rows * d + B * d
        for by, fl in zip(k["launch_bytes"], k["launch_flops"]):
            assert fl > 0 and by > 0
        total_rows += sum(fl / (2.0 * 384) for fl in k["launch_flops"])
    # every query scans every row once: sum over launches of B * rows
    assert abs(total_rows - 200_000 * (1 + 37 + 300)) < 1e-3 * total_rows


def test_ivf_step_is_timed():
    base = data.mixture(30000, 64, 64, 0.3, 7301, center_seed=7300, normalize=False, zipf_s=1.1)
    Q = data.mixture(40, 64, 64, 0.3, 7302, center_seed=7300, normalize=False, zipf_s=1.1)
    print("[case] profile ivf n=30000", flush=True)
    idx = Index(base, mode="ivf", nlist=64, rerank=200)
    ref = idx.search(Q, 100, nprobe=8)
    idx.profile(1)
    got = idx.search(Q, 100, nprobe=8)
    ks = idx.profile_kernels()
    pr = idx.profile(0)
    np.testing.assert_array_equal(ref[0], got[0])
    assert pr["steps"] >= 1 and pr["finalize_ms"] > 0
    ivf = [k for k in ks if "ivf" in k["kernel"]]
    assert ivf and all(k["ms"] > 0 and k["bytes"] > 0 for k in ivf)

"""Scan-kernel shapes: every wide-kernel tile configuration and chunk size,
the streaming kernel, and the autotuner's table give the same answers as the
built-in choice (a shape changes timing only)."""
import numpy as np
import pytest

from annvedi_amd import Index, data
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N, D = 150_000, 96


@pytest.fixture(scope="module")
def inputs():
    base = data.mixture(N, D, 128, 0.3, 61, center_seed=62)
    Q = data.mixture(40, D, 128, 0.3, 63, center_seed=62)
    return base, Q


@pytest.mark.parametrize("mode", ["exact", "bf16", "fp8", "sq8", "sq4"])
def test_every_scan_shape_gives_the_same_answers(inputs, mode):
    base, Q = inputs
    print(f"[case] shapes mode={mode} n={N} d={D}", flush=True)
    idx = Index(base, mode=mode, rerank=200)
    ref_ids, ref_d = idx.search(Q, 100)
    ref_small = [idx.search(Q[:b], 100) for b in (1, 3)]  # tiny batches: streaming or wide by type
    if mode == "exact":
        np.testing.assert_array_equal(ref_ids, O.exact_topk_f64(base, Q, 100)[0])
    shapes = [(1 << 30, 0, 0), (1, 0, 0), (1, 0, 4096)] + [(1, c, ch) for c in (1, 2, 3, 4, 5, 6, 7, 8) for ch in (0, 8192)]
    tried = 0
    for mb, cfg, ch in shapes:
        try:
            idx.set_scan_shape(mb, cfg, ch)
        except RuntimeError as e:  # configuration does not fit the LDS at this dim
            assert "does not fit" in str(e) or "bad scan shape" in str(e)
            continue
        tried += 1
        ids, dd = idx.search(Q, 100)
        np.testing.assert_array_equal(ids, ref_ids, err_msg=f"shape {(mb, cfg, ch)}")
        np.testing.assert_array_equal(dd, ref_d, err_msg=f"shape {(mb, cfg, ch)}")
        for b, (ri, rd) in zip((1, 3), ref_small):
            ids, dd = idx.search(Q[:b], 100)
            np.testing.assert_array_equal(ids, ri, err_msg=f"shape {(mb, cfg, ch)} B={b}")
            np.testing.assert_array_equal(dd, rd, err_msg=f"shape {(mb, cfg, ch)} B={b}")
    assert tried >= 5


def test_autotune_table_keeps_answers(inputs):
    base, Q = inputs
    print(f"[case] autotune sq8 n={N} d={D}", flush=True)
    idx = Index(base, mode="sq8", rerank=200)
    ref_ids, ref_d = idx.search(Q, 100)
    rep = idx.autotune(max_batch=64, reps=1)
    assert [r["B"] for r in rep] == [1, 2, 4, 8, 16, 32, 64]
    for r in rep:
        assert r["kernel"] in ("streaming", "wide", "default") and r["ms"] > 0 and r["ms"] <= r["default_ms"] * 1.0001
    for B in (1, 3, 40):
        ids, dd = idx.search(Q[:B], 100)
        np.testing.assert_array_equal(ids, ref_ids[:B])
        np.testing.assert_array_equal(dd, ref_d[:B])
    with pytest.raises(RuntimeError):
        idx.set_scan_shape(1, 9, 0)  # past the last configuration


@pytest.mark.parametrize("B", [1, 3, 255, 256, 300, 385])
def test_sq8_384_specialisation_equals_the_runtime_dim_kernel(B):
    """SQ8 at d = 384 runs a wide-scan instantiation with the dim fixed at
    compile time (12 waves a workgroup from 256 queries, 8 below); the
    8-wave tile shape with the dim read at run time (wide configuration 2)
    gives bit-identical answers, on a row count that ends mid-tile, with the
    built-in and a small row chunk, on both sides of the 256-query switch
    and of the 384-query range-schedule switch (385: one live query in the
    last 128-query tile)."""
    n = 150_003
    print(f"[case] sq8 d=384 specialisation B={B}", flush=True)
    base = data.mixture(n, 384, 512, 0.3, 64, center_seed=65)
    Q = data.mixture(B, 384, 512, 0.3, 66, center_seed=65)
    idx = Index(base, mode="sq8", rerank=200)
    for ch in (0, 8192):
        idx.set_scan_shape(1, 0, ch)
        ref_ids, ref_d = idx.search(Q, 100)
        idx.set_scan_shape(1, 2, ch)
        ids, dd = idx.search(Q, 100)
        np.testing.assert_array_equal(ids, ref_ids, err_msg=f"chunk {ch}")
        np.testing.assert_array_equal(dd, ref_d, err_msg=f"chunk {ch}")
    o_ids, _ = O.exact_topk_f64(base, Q[:2], 100)
    for i in range(min(B, 2)):
        assert len(set(ref_ids[i]) & set(o_ids[i])) >= 99


def test_sq8_768_every_shape_fits_or_is_refused():
    """At d = 768 the 32 x 32 lean-epilogue shapes need more LDS than a
    workgroup has: setting them is refused up front, every other shape gives
    the built-in answers, and the autotuner (which tries every shape that
    fits) completes."""
    print("[case] sq8 d=768 shapes and autotune", flush=True)
    base = data.mixture(30_000, 768, 64, 0.3, 67, center_seed=68)
    Q = data.mixture(200, 768, 64, 0.3, 69, center_seed=68)
    idx = Index(base, mode="sq8", rerank=300)
    ref_ids, ref_d = idx.search(Q, 100)
    refused = []
    for cfg in range(1, 9):
        try:
            idx.set_scan_shape(1, cfg, 0)
        except RuntimeError as e:
            assert "does not fit" in str(e) or "bad scan shape" in str(e)
            refused.append(cfg)
            continue
        ids, dd = idx.search(Q, 100)
        np.testing.assert_array_equal(ids, ref_ids, err_msg=f"cfg {cfg}")
        np.testing.assert_array_equal(dd, ref_d, err_msg=f"cfg {cfg}")
    assert 6 in refused and 7 in refused and 8 in refused, refused
    idx.set_scan_shape(0, 0, 0)
    table = idx.autotune(max_batch=256, reps=1)
    assert len(table) > 0
    ids, dd = idx.search(Q, 100)
    np.testing.assert_array_equal(ids, ref_ids)

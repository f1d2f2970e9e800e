"""SQ8 wide scan with rows staged in LDS and queries in registers (wide
configuration 8): the same candidates, hence bit-identical answers, as the
built-in kernels, for every padded dim it takes (128, 256), batches that
fill one, two and a partial query tile of 512, row counts that end mid-tile,
and the queue's flush and spill paths."""
import numpy as np
import pytest

from annvedi_amd import Index, data
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CFG_RL = 8


def _rl_launches(idx):
    return sum(r["launches"] for r in idx.profile_kernels() if r["kernel"].startswith("scan_rl_kernel"))


@pytest.mark.parametrize("d,B", [(256, 600), (256, 1024), (200, 513), (100, 200), (256, 40)])
def test_row_staged_scan_equals_the_built_in_kernels(d, B):
    n = 200_003
    print(f"[case] rowstage sq8 n={n} d={d} B={B}", flush=True)
    base = data.mixture(n, d, 1024, 0.3, 71, center_seed=72)
    Q = data.mixture(B, d, 1024, 0.3, 73, center_seed=72)
    idx = Index(base, mode="sq8", rerank=200)
    ref_ids, ref_d = idx.search(Q, 100)
    idx.set_scan_shape(1, CFG_RL, 0)
    idx.profile(1)
    ids, dd = idx.search(Q, 100)
    assert _rl_launches(idx) >= 1
    idx.profile(0)
    np.testing.assert_array_equal(ids, ref_ids)
    np.testing.assert_array_equal(dd, ref_d)
    # the answers are the oracle's top-100 wherever the re-scored rows hold it
    o_ids, _ = O.exact_topk_f64(base, Q[:3], 100)
    for i in range(3):
        assert len(set(ids[i]) & set(o_ids[i])) >= 99


@pytest.mark.parametrize("qcap", ["0", "256"])
def test_row_staged_scan_spill_and_flush_paths(monkeypatch, qcap):
    """A zero-length LDS queue sends every candidate straight to the lists; a
    short one flushes every stage.  Neither changes the answer."""
    n = 150_000
    print(f"[case] rowstage queue cap {qcap}", flush=True)
    base = data.mixture(n, 256, 1024, 0.3, 76, center_seed=72)
    Q = data.mixture(520, 256, 1024, 0.3, 77, center_seed=72)
    idx = Index(base, mode="sq8", rerank=128)
    ref_ids, ref_d = idx.search(Q, 100)
    monkeypatch.setenv("ANNMI_WIDE_QCAP", qcap)
    idx2 = Index(base, mode="sq8", rerank=128)
    idx2.set_scan_shape(1, CFG_RL, 0)
    ids, dd = idx2.search(Q, 100)
    np.testing.assert_array_equal(ids, ref_ids)
    np.testing.assert_array_equal(dd, ref_d)


@pytest.mark.parametrize("d", [384, 500])
def test_row_staged_scan_refuses_other_dims(d):
    base = data.mixture(5000, d, 64, 0.3, 74, center_seed=75)
    idx = Index(base, mode="sq8", rerank=128)
    with pytest.raises(RuntimeError):
        idx.set_scan_shape(1, CFG_RL, 0)

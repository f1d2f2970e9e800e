"""GPU encoders vs the oracle's restatement of the reference encoders:
byte-exact codes, bit-exact per-dim scale/offset."""
import numpy as np
import pytest

from annvedi_amd import Index, data
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def sq_inputs():
    base, _ = data.workload("sq-1m-768", n=6000, n_queries=1)
    base = base[:, :200].copy()               # d = 200: exercises padding to 256
    base[:, 7] = 0.5                          # constant dimension
    base[:, 9] = -1.25                        # constant negative dimension
    base[:13, 11] = np.float32(3.0)           # a dimension with ties at the max
    return base


@pytest.mark.parametrize("variant", [0, 1])
def test_sq8_codes_byte_exact(variant):
    base = sq_inputs()
    idx = Index(base, mode="sq8", sq_variant=variant)
    s, o = O.sq8_train(base, variant=variant)
    gs, go = idx.export("scale"), idx.export("offset")
    d = base.shape[1]
    np.testing.assert_array_equal(gs[:d], s)
    np.testing.assert_array_equal(go[:d], o)
    codes = O.sq8_encode(base, s, o, variant=variant)
    rows = idx.export("rows")
    np.testing.assert_array_equal(rows[:, :d], codes)
    assert np.all(rows[:, d:] == 0)
    # the constant dims follow each reference's convention
    assert go[7] == (0.0 if variant == 0 else 0.5) and gs[7] == 1.0


def test_sq4_codes_byte_exact_reference_layout():
    base = sq_inputs()
    idx = Index(base, mode="sq4")
    s, o = O.sq4_train(base)
    dp = s.shape[0]                           # 224 = ceil(200/32)*32
    np.testing.assert_array_equal(idx.export("scale")[:dp], s)
    np.testing.assert_array_equal(idx.export("offset")[:dp], o)
    codes = O.sq4_encode(base, s, o)          # n x 112 bytes
    rows = idx.export("rows")                 # n x 128 bytes (kdim 256)
    np.testing.assert_array_equal(rows[:, :dp // 2], codes)
    assert np.all(rows[:, dp // 2:] == 0)


def test_bf16_and_fp8_codes_match_torch_rounding():
    import torch

    base, _ = data.workload("exact-1m-384", n=3000, n_queries=1)
    ib = Index(base, mode="bf16")
    rows = ib.export("rows").view(np.uint16)[:, :384]
    want = torch.from_numpy(base).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(rows, want)
    i8 = Index(base, mode="fp8")
    alpha = i8.export("meta")["fp8_alpha"]
    assert alpha == 2.0 ** round(np.log2(alpha))
    assert 128 <= np.abs(base).max() * alpha < 256
    rows = i8.export("rows")[:, :384]
    want = (torch.from_numpy(base) * alpha).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    np.testing.assert_array_equal(rows, want)


def test_norm_terms():
    base = sq_inputs()
    idx = Index(base, mode="sq8")
    s, o = O.sq8_train(base)
    codes = O.sq8_encode(base, s, o)
    want = ((codes.astype(np.float64) * s + o) ** 2).sum(1)  # ||reconstructed row||^2
    np.testing.assert_allclose(idx.export("norms"), want, rtol=1e-5)
    ex = Index(base, mode="exact")
    np.testing.assert_allclose(ex.export("norms"), (base.astype(np.float64) ** 2).sum(1), rtol=1e-5)

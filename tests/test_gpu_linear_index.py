"""The reference's LinearIndex.query reproduced on the GPU bit for bit:
all 100 ranks of the golden ids (early-abandon partial sums beyond rank 64
included) and the fp32 values the reference ranks."""
import os

import numpy as np
import pytest

from annvedi_amd import Algorithm, Index, data
from oracle import oracle as O
from tests.golden.make_golden import LINEAR_CASES, case_inputs, dup_case

pytestmark = pytest.mark.gpu
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "linear_ref.npz"))


@pytest.mark.parametrize("case", LINEAR_CASES, ids=[c[0] for c in LINEAR_CASES])
def test_linear_index_mode_bit_exact_vs_reference(case):
    name, n, d, C, sigma, norm, nq, k, off = case
    base, Q = case_inputs(n, d, C, sigma, norm, nq, off)
    print(f"[case] linear_index {name}", flush=True)
    idx = Index(base, mode="exact")
    ids, vals = idx.search_linear_index(Q, k)
    np.testing.assert_array_equal(ids, GOLD[f"{name}__ids"])
    o_ids, o_vals = O.linear_index_batch(base, Q, k, nthreads=8)
    np.testing.assert_array_equal(ids, o_ids)
    np.testing.assert_array_equal(vals, o_vals)
    assert idx.n_distances == n * nq


def test_linear_index_mode_duplicates():
    base, Q = dup_case()
    idx = Index(base, mode="exact")
    ids, vals = idx.search_linear_index(Q, 30)
    np.testing.assert_array_equal(ids, GOLD["dups__ids"])


def test_linear_index_mode_many_blocks_and_plugin():
    """Several 4096-row blocks (the threshold carried across blocks)."""
    base, Q = data.workload("linear-cpu-100k", n=50_000, n_queries=6)
    a = Algorithm()
    a.fit(base, mode="linear_index")
    ids, vals = a.batch_query(Q, 100)
    o_ids, o_vals = O.linear_index_batch(base, Q, 100, nthreads=8)
    np.testing.assert_array_equal(ids, o_ids)
    np.testing.assert_array_equal(vals, o_vals)
    ref = O.ref_module("linear_ann_cpp")
    if ref is not None:  # the compiled reference itself, when present
        r = ref.LinearIndex()
        r.fit(base)
        for i in range(Q.shape[0]):
            np.testing.assert_array_equal(ids[i], r.query(Q[i], 100))
    assert a.query(Q[0], 100).tolist() == ids[0].tolist()

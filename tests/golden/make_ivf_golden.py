"""Write tests/golden/ivf_ref.npz from the reference's own IVF code.

Runs only in a container that holds the reference tree: it needs
oracle/_ref/ivf_lsh_cpp, which ``make -f oracle/ref.mk`` compiles from
competitors/ivf_lsh/src/lsh_index_optimized.cpp.  The reference is fitted
with n_tables=1, n_bits=0 (one LSH bucket per list, so a probed list is
scanned whole) and queried with refine_r = n (every candidate refined, and
the candidate cap max(2000, 4 * refine_r), :357-358, never reached).  Only
the reference's answers (ids) and an input checksum are stored; the tests
regenerate the inputs from the seeds below.

Two cases:

* ``line``: integer coordinates on one axis (the other dims 0).  Every
  float sum of the k-means update is an exact integer below 2^24 and every
  distance is one rounded square, so the reference's answers do not depend
  on its thread count or FMA contraction: the oracle's k-means and query
  restatement must reproduce them bit for bit (initial rows of the
  mt19937(1337) shuffle, 15 rounds, probe order, fall-back to all rows).
* ``clusters``: 24 separated Gaussian clusters, each seeded by one of the
  reference's initial centroid rows, Zipf-skewed sizes (some lists hold
  fewer than k rows).  k-means keeps the generating partition whatever the
  float rounding, so any IVF build (the GPU one included) must give the
  reference's answers.

    python tests/golden/make_ivf_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests.golden.make_golden import checksum  # noqa: E402

K = 100
NPROBES = (1, 2, 4)


def line_case():
    """(base, Q, nlist): 1200 distinct integer positions in [0, 10000) on
    axis 0 of 16 dims; queries at integers + 0.37 (no two rows at equal
    distance)."""
    rng = np.random.Generator(np.random.PCG64(110_000))
    n, d, nq = 1200, 16, 60
    base = np.zeros((n, d), dtype=np.float32)
    base[:, 0] = rng.choice(10_000, size=n, replace=False).astype(np.float32)
    Q = np.zeros((nq, d), dtype=np.float32)
    Q[:, 0] = (rng.integers(0, 10_000, size=nq) + 0.37).astype(np.float32)
    return base, Q, 30


def clusters_case():
    """(base, Q, nlist): 3000 rows in 24 clusters (centers 10 * N(0,1)^64,
    sigma 0.5), cluster sizes ~ Zipf(1.1), the reference's initial centroid
    row c placed in cluster c; 40 queries from the same clusters."""
    rng = np.random.Generator(np.random.PCG64(120_000))
    n, d, C, nq = 3000, 64, 24, 40
    centers = (10.0 * rng.standard_normal((C, d))).astype(np.float32)
    p = np.arange(1, C + 1, dtype=np.float64) ** -1.1
    z = rng.choice(C, size=n, p=p / p.sum())
    init, _ = O.kmeans_init(n, C)
    z[init] = np.arange(C)
    base = centers[z] + np.float32(0.5) * rng.standard_normal((n, d)).astype(np.float32)
    zq = rng.integers(0, C, size=nq)
    Q = centers[zq] + np.float32(0.5) * rng.standard_normal((nq, d)).astype(np.float32)
    return np.ascontiguousarray(base, dtype=np.float32), np.ascontiguousarray(Q, dtype=np.float32), C


CASES = {"line": line_case, "clusters": clusters_case}


def reference_answers(ref, base, Q, nlist):
    idx = ref.LSHIndex()
    idx.fit(base, n_tables=1, n_bits=0, n_clusters=nlist)
    n = base.shape[0]
    out = np.empty((Q.shape[0], len(NPROBES), K), dtype=np.int32)
    for qi, q in enumerate(Q):
        for j, p in enumerate(NPROBES):
            out[qi, j] = idx.query(q, K, n_probes=0, n_probe_clusters=p, refine_r=n)
    return out


def main():
    ref = O.ref_module("ivf_lsh_cpp")
    if ref is None:
        raise SystemExit("oracle/_ref/ivf_lsh_cpp not built: make -f oracle/ref.mk")
    out = {}
    for name, make in CASES.items():
        base, Q, nlist = make()
        ids = reference_answers(ref, base, Q, nlist)
        out[f"{name}__ids"] = ids
        out[f"{name}__checksum"] = np.asarray(checksum(base, Q))
        # the restatement agrees (the tests check it again from the fixture)
        cent, assign = O.ivf_kmeans(base, nlist)
        perm, off = O.ivf_lists(assign, nlist)
        for j, p in enumerate(NPROBES):
            got = O.ivf_reference_query(base, cent, perm, off, Q, K, p)
            print(name, "nprobe", p, "oracle == reference:", np.array_equal(got, ids[:, j]),
                  "queries falling back to all rows:",
                  int(sum(sum(off[l + 1] - off[l] for l in
                              np.lexsort((np.arange(nlist), [O.l2_linear_index(q, c) for c in cent]))[:p]) < K
                          for q in Q)))
    np.savez_compressed(os.path.join(HERE, "ivf_ref.npz"), **out)
    print("wrote", os.path.join(HERE, "ivf_ref.npz"))


if __name__ == "__main__":
    main()

"""Write the golden vectors under tests/golden/ from the reference's own code.

Runs only in a container that holds the reference tree: it needs the modules
that ``make -f oracle/ref.mk`` compiles from the reference sources into
``oracle/_ref`` (src/linear_index.cpp -> linear_ann_cpp).  The fixtures it
writes are data only -- inputs are regenerated from seeds by the tests, and
only the reference's outputs (ids) plus an input checksum are stored.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from annvedi_amd import data  # noqa: E402
from oracle import oracle as O  # noqa: E402

# (name, n, d, clusters, sigma, normalize, n_queries, k, seed offset)
LINEAR_CASES = [
    ("mix20k_d384", 20000, 384, 256, 0.30, True, 200, 100, 0),
    ("mix5k_d768", 5000, 768, 64, 0.25, False, 50, 100, 1),
    ("mix3k_d100", 3000, 100, 32, 0.30, True, 50, 100, 2),   # checkpoint at 96 + 4-dim tail
    ("mix3k_d29", 3000, 29, 32, 0.30, False, 50, 100, 3),    # 16 + 8-wide tail + 5 scalar
    ("mix2k_d20", 2000, 20, 16, 0.30, False, 50, 100, 4),    # 16 + 4 scalar, no checkpoint
    ("mix50_d16_k100", 50, 16, 4, 0.30, False, 10, 100, 5),  # k > n clamp, n < 64
    ("mix500_d64_k10", 500, 64, 8, 0.30, True, 20, 10, 6),   # k < 64 threshold slots
]


def case_inputs(n, d, C, sigma, normalize, nq, seed_off):
    kw = dict(center_seed=50_000 + seed_off, normalize=normalize)
    base = data.mixture(n, d, C, sigma, 60_000 + seed_off, **kw)
    Q = data.mixture(nq, d, C, sigma, 70_000 + seed_off, **kw)
    return base, Q


def checksum(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def dup_case():
    """Duplicate rows: ties in every distance, broken by lower id."""
    rng = np.random.Generator(np.random.PCG64(80_000))
    u = rng.standard_normal((40, 48), dtype=np.float32)
    base = np.concatenate([u, u[::-1], u[:10]], axis=0)  # 90 rows, many exact duplicates
    Q = np.concatenate([u[:5], rng.standard_normal((5, 48), dtype=np.float32)], axis=0)
    return base, Q


def order_pairs(trials=4000, d=384):
    """Rows that are permutations of one vector, query 0: the exact sums are
    equal, so which row the reference ranks first exposes how it rounds."""
    rng = np.random.Generator(np.random.PCG64(90_000))
    V = rng.standard_normal((trials, d), dtype=np.float32)
    P = np.stack([rng.permutation(d) for _ in range(trials)])
    return V, P


def main():
    ref = O.ref_module("linear_ann_cpp")
    if ref is None:
        raise SystemExit("oracle/_ref/linear_ann_cpp not built: make -f oracle/ref.mk")
    out = {}
    for name, n, d, C, sigma, norm, nq, k, off in LINEAR_CASES:
        base, Q = case_inputs(n, d, C, sigma, norm, nq, off)
        idx = ref.LinearIndex()
        idx.fit(base)
        ids = np.stack([idx.query(q, k) for q in Q]).astype(np.int32)
        assert idx.total_distances_count() == n * nq
        out[name] = dict(ids=ids, checksum=checksum(base, Q), k=k)
        print(name, ids.shape)
    base, Q = dup_case()
    idx = ref.LinearIndex()
    idx.fit(base)
    out["dups"] = dict(ids=np.stack([idx.query(q, 30) for q in Q]).astype(np.int32),
                       checksum=checksum(base, Q), k=30)

    V, P = order_pairs()
    first = np.empty(V.shape[0], dtype=np.int8)
    zero = np.zeros(V.shape[1], dtype=np.float32)
    for t in range(V.shape[0]):
        pair = np.stack([V[t], V[t][P[t]]])
        idx = ref.LinearIndex()
        idx.fit(pair)
        first[t] = idx.query(zero, 2)[0]
    out["order_pairs"] = dict(first=first, checksum=checksum(V, P))

    np.savez_compressed(
        os.path.join(HERE, "linear_ref.npz"),
        **{f"{c}__{key}": np.asarray(v) for c, dct in out.items() for key, v in dct.items()},
    )
    print("wrote", os.path.join(HERE, "linear_ref.npz"))


if __name__ == "__main__":
    main()

"""Saved indexes (annmi_save / annmi_load): a loaded index holds the same
derived arrays as the one saved and answers every query identically; the
IVF load skips k-means.  Corrupt or mismatched images are refused."""
import numpy as np
import pytest

from annvedi_amd import Index, data

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["exact", "bf16", "fp8", "sq8", "sq4", "ivf"])
def test_saved_index_answers_identically(mode, tmp_path):
    n, d = 30000, 100
    print(f"[case] persist mode={mode} n={n} d={d}", flush=True)
    base = data.mixture(n, d, 64, 0.3, 71, center_seed=72)
    Q = data.mixture(33, d, 64, 0.3, 73, center_seed=72)
    kw = dict(nlist=64, nprobe=8) if mode == "ivf" else {}
    a = Index(base, mode=mode, rerank=200, **kw)
    ids_a, d_a = a.search(Q, 100)
    path = str(tmp_path / "idx.annmi")
    a.save(path)
    b = Index.load(path, base, mode=mode, rerank=200, **kw)
    ids_b, d_b = b.search(Q, 100)
    np.testing.assert_array_equal(ids_a, ids_b)
    np.testing.assert_array_equal(d_a, d_b)
    for what in ("rows", "scale", "offset", "norms", "perm", "list_offsets", "centroids"):
        if mode == "exact" and what in ("rows",):
            continue
        np.testing.assert_array_equal(a.export(what), b.export(what), err_msg=what)
    assert a.save_bytes() == b.save_bytes()


def test_bad_images_are_refused():
    n, d = 5000, 64
    base = data.mixture(n, d, 16, 0.3, 81, center_seed=82)
    a = Index(base, mode="sq8")
    blob = a.save_bytes()
    with pytest.raises(RuntimeError, match="too short|bytes given"):
        Index(base, mode="sq8", saved=blob[:len(blob) // 2])
    with pytest.raises(RuntimeError, match="mode"):
        Index(base, mode="sq4", saved=blob)
    with pytest.raises(RuntimeError, match="5000|4999"):
        Index(base[:4999], mode="sq8", saved=blob)
    bad = bytearray(blob)
    bad[0:8] = b"NOTANIDX"
    with pytest.raises(RuntimeError, match="not an annmi index"):
        Index(base, mode="sq8", saved=bytes(bad))


def test_plugin_index_file_and_autotune(tmp_path):
    """Algorithm.fit(index_file=...) writes the index after the first build and
    loads it on the next fit; autotune: true tunes at fit; answers agree."""
    from annvedi_amd import Algorithm

    n, d = 40000, 64
    print(f"[case] plugin index_file n={n} d={d}", flush=True)
    base = data.mixture(n, d, 64, 0.3, 91, center_seed=92)
    Q = data.mixture(20, d, 64, 0.3, 93, center_seed=92)
    path = str(tmp_path / "plugin.annmi")
    a = Algorithm()
    a.fit(base, backend="mi355x", mode="sq8", rerank=300, index_file=path, autotune=True, max_batch=64)
    import os
    assert os.path.exists(path)
    b = Algorithm()
    b.fit(base, backend="mi355x", mode="sq8", rerank=300, index_file=path)
    ia, _ = a.batch_query(Q, 100)
    ib, _ = b.batch_query(Q, 100)
    np.testing.assert_array_equal(ia, ib)
    np.testing.assert_array_equal(a.query(Q[0], 100), ib[0])


def test_corrupt_ivf_permutation_is_refused_before_upload():
    """The list permutation of a saved IVF index becomes candidate row ids:
    an entry past the base, or a repeated one, is refused on the host."""
    import struct

    n, d = 6000, 64
    base = data.mixture(n, d, 16, 0.3, 85, center_seed=86)
    a = Index(base, mode="ivf", nlist=16)
    blob = a.save_bytes()
    # header: magic[8] version mode n d kdim row_stride sq_variant nlist alpha reserved, then off[8]
    off = struct.unpack_from("<8q", blob, 56)
    perm_at = off[4]
    assert struct.unpack_from("<I", blob, perm_at)[0] < n
    bad = bytearray(blob)
    struct.pack_into("<I", bad, perm_at + 4 * 17, n)  # past the last row
    with pytest.raises(RuntimeError, match="permutation"):
        Index(base, mode="ivf", nlist=16, saved=bytes(bad))
    bad = bytearray(blob)
    first = struct.unpack_from("<I", blob, perm_at)[0]
    struct.pack_into("<I", bad, perm_at + 4 * 5, first)  # a repeated row
    with pytest.raises(RuntimeError, match="permutation"):
        Index(base, mode="ivf", nlist=16, saved=bytes(bad))
    ok = Index(base, mode="ivf", nlist=16, saved=blob)  # the untouched image still loads
    assert ok.n == n

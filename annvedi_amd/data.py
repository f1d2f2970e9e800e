"""Seeded synthetic Gaussian-mixture vectors for the k-NN search path.

The reference ships no datasets (its public HDF5 files are downloaded at run
time, orthogonal-competition/prepare_data.py:17-36), so every workload here is
synthetic: centers ~ N(0, 1)^d, a point is center[z] + sigma * N(0, 1)^d with
z uniform over the clusters (or truncated Zipf(s) when ``zipf_s`` is set),
optionally multiplied per dimension by s_d ~ U[lo, hi] and optionally
L2-normalised.  Queries come from the same centers with an independent seed.

Two generators share these semantics:

* :func:`mixture` -- numpy ``Generator(PCG64(seed))``, streamed in chunks,
  used for anything a CPU oracle has to see;
* :func:`mixture_torch` -- torch Philox on a device, used by the benchmark
  to build 10M-row bases directly in HBM.  It draws different numbers from the
  numpy one (same distribution, different stream).
"""
from __future__ import annotations

import numpy as np

# Seeds per workload: centers 1000+c, base 2000+c, queries 3000+c.
SEED_CENTERS = 1000
SEED_BASE = 2000
SEED_QUERIES = 3000

WORKLOADS = {
    # name: (n, d, clusters, sigma, normalize, dim_scale, zipf_s)
    "linear-cpu-100k": (100_000, 384, 256, 0.30, True, None, None),
    "exact-1m-384": (1_000_000, 384, 1024, 0.30, True, None, None),
    "sq-1m-768": (1_000_000, 768, 1024, 0.25, False, (0.5, 2.0), None),
    "mixed-10m-384": (10_000_000, 384, 4096, 0.30, True, None, None),
    "ivf-10m-768": (10_000_000, 768, 4096, 0.30, False, None, 1.1),
}


def _cluster_probs(n_clusters: int, zipf_s: float | None) -> np.ndarray | None:
    if zipf_s is None:
        return None
    p = (np.arange(1, n_clusters + 1, dtype=np.float64)) ** (-float(zipf_s))
    return p / p.sum()


def centers(n_clusters: int, d: int, seed: int) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.standard_normal((n_clusters, d), dtype=np.float32)


def mixture(n: int, d: int, n_clusters: int, sigma: float, seed: int, *, center_seed: int,
            normalize: bool = True, dim_scale: tuple[float, float] | None = None,
            zipf_s: float | None = None, chunk: int = 1 << 18) -> np.ndarray:
    """n x d float32 points of the mixture (C-contiguous)."""
    c = centers(n_clusters, d, center_seed)
    scale = None
    if dim_scale is not None:
        srng = np.random.Generator(np.random.PCG64(center_seed + 7))
        scale = srng.uniform(dim_scale[0], dim_scale[1], size=d).astype(np.float32)
    p = _cluster_probs(n_clusters, zipf_s)
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        m = e - s
        z = rng.choice(n_clusters, size=m, p=p) if p is not None else rng.integers(0, n_clusters, size=m)
        x = c[z]
        x += np.float32(sigma) * rng.standard_normal((m, d), dtype=np.float32)
        if scale is not None:
            x *= scale
        if normalize:
            x /= np.linalg.norm(x, axis=1, keepdims=True).astype(np.float32)
        out[s:e] = x
    return out


def workload(name: str, n: int | None = None, n_queries: int = 1000, config_id: int = 0):
    """(base, queries) for a named workload, optionally with a smaller n."""
    wn, d, C, sigma, norm, dscale, zs = WORKLOADS[name]
    n = wn if n is None else n
    kw = dict(center_seed=SEED_CENTERS + config_id, normalize=norm, dim_scale=dscale, zipf_s=zs)
    base = mixture(n, d, C, sigma, SEED_BASE + config_id, **kw)
    queries = mixture(n_queries, d, C, sigma, SEED_QUERIES + config_id, **kw)
    return base, queries


def mixture_torch(n: int, d: int, n_clusters: int, sigma: float, seed: int, *, center_seed: int,
                  normalize: bool = True, dim_scale: tuple[float, float] | None = None,
                  zipf_s: float | None = None, device: str = "cuda", chunk: int = 1 << 20):
    """Same distribution as :func:`mixture`, generated on ``device``."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(center_seed)
    c = torch.randn((n_clusters, d), generator=g, device=device, dtype=torch.float32)
    scale = None
    if dim_scale is not None:
        scale = torch.empty(d, device=device).uniform_(dim_scale[0], dim_scale[1], generator=g)
    probs = None
    if zipf_s is not None:
        probs = torch.arange(1, n_clusters + 1, device=device, dtype=torch.float64) ** (-float(zipf_s))
        probs = (probs / probs.sum()).float()
    g.manual_seed(seed)
    out = torch.empty((n, d), device=device, dtype=torch.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        m = e - s
        if probs is not None:
            z = torch.multinomial(probs, m, replacement=True, generator=g)
        else:
            z = torch.randint(0, n_clusters, (m,), generator=g, device=device)
        x = c[z]
        x.add_(torch.randn((m, d), generator=g, device=device, dtype=torch.float32), alpha=sigma)
        if scale is not None:
            x.mul_(scale)
        if normalize:
            x.div_(torch.linalg.vector_norm(x, dim=1, keepdim=True))
        out[s:e] = x
    return out

"""Synthetic data sets for the EFANNA workloads (deterministic, numpy only).

Stream order (numpy ``default_rng(seed)``): mixture centres, then base labels,
then base noise in row chunks, then query labels, then query noise.
"""
from __future__ import annotations

import numpy as np

_CHUNK = 65536


def uniform_int(n: int, d: int, nq: int, seed: int = 0):
    """i.i.d. uniform integers 0..255 stored as float32 (the 10K x 128 CPU case)."""
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 256, size=(n, d)).astype(np.float32)
    q = rng.integers(0, 256, size=(nq, d)).astype(np.float32)
    return base, q


def _mixture_rows(rng, centres, n, sigma, lo, hi, rounded):
    d = centres.shape[1]
    labels = rng.integers(0, centres.shape[0], size=n)
    out = np.empty((n, d), np.float32)
    for s in range(0, n, _CHUNK):
        e = min(n, s + _CHUNK)
        x = centres[labels[s:e]] + rng.normal(0.0, sigma, size=(e - s, d)).astype(np.float32)
        np.clip(x, lo, hi, out=x)
        if rounded:
            np.rint(x, out=x)
        out[s:e] = x
    return out


def gaussian_mixture(n: int, d: int, nq: int, centres: int = 1000, sigma: float = 20.0, hi: float = 255.0,
                     rounded: bool = True, seed: int = 0, uniform_query_frac: float = 0.0):
    """Gaussian mixture: centres uniform in [0, hi]^d, isotropic sigma, clipped to
    [0, hi] (and rounded to integers when ``rounded``).  A ``uniform_query_frac``
    share of the queries is drawn uniformly from [0, hi]^d instead (rounded
    likewise); those queries come last."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, hi, size=(centres, d)).astype(np.float32)
    base = _mixture_rows(rng, c, n, sigma, 0.0, hi, rounded)
    nu = int(round(nq * uniform_query_frac))
    q = _mixture_rows(rng, c, nq - nu, sigma, 0.0, hi, rounded)
    if nu:
        u = rng.uniform(0.0, hi, size=(nu, d)).astype(np.float32)
        if rounded:
            np.rint(u, out=u)
        q = np.concatenate([q, u])
    return base, q


def mixture_queries(d: int, nq: int, centres: int = 1000, sigma: float = 20.0, hi: float = 255.0,
                    rounded: bool = True, seed: int = 0, query_seed: int = 1, uniform_query_frac: float = 0.0):
    """Extra queries from the mixture of ``gaussian_mixture(..., seed)``: the
    centres are drawn again from ``seed``, the queries from an independent
    generator seeded with ``query_seed`` (uniform ones last)."""
    c = np.random.default_rng(seed).uniform(0.0, hi, size=(centres, d)).astype(np.float32)
    rng = np.random.default_rng(query_seed)
    nu = int(round(nq * uniform_query_frac))
    q = _mixture_rows(rng, c, nq - nu, sigma, 0.0, hi, rounded)
    if nu:
        u = rng.uniform(0.0, hi, size=(nu, d)).astype(np.float32)
        if rounded:
            np.rint(u, out=u)
        q = np.concatenate([q, u])
    return q


def workload(name: str, n: int | None = None, nq: int | None = None):
    """Named workloads with their sizes; n/nq override for scaled-down tests."""
    if name == "uniform10k":
        return uniform_int(n or 10_000, 128, nq or 100, seed=0)
    if name == "mixture1m":
        return gaussian_mixture(n or 1_000_000, 128, nq or 10_000, centres=1000, sigma=20.0, hi=255.0, rounded=True)
    if name == "mixture1m_batched":
        return gaussian_mixture(n or 1_000_000, 128, nq or 100_000, centres=1000, sigma=20.0, hi=255.0,
                                rounded=True, uniform_query_frac=0.3)
    if name == "mixture1m_d960":
        return gaussian_mixture(n or 1_000_000, 960, nq or 10_000, centres=500, sigma=0.05, hi=1.0, rounded=False)
    raise KeyError(name)


def lsh_codes(base, queries, tables: int, bits: int, seed: int = 0):
    """Hash codes for a hashing index: per table `bits` random hyperplanes
    through the base mean, bit b set when the row lies on the positive side
    of hyperplane b.  Returns (tables x n, tables x nq) uint64 codes, the
    layout of the reference's code files (one file per table)."""
    base = np.asarray(base, np.float32)
    queries = np.asarray(queries, np.float32)
    mean = base.mean(axis=0, dtype=np.float64).astype(np.float32)
    rng = np.random.default_rng(seed)
    out = [np.zeros((tables, x.shape[0]), np.uint64) for x in (base, queries)]
    for t in range(tables):
        r = rng.standard_normal((base.shape[1], bits)).astype(np.float32)
        for codes, x in zip(out, (base, queries)):
            for c0 in range(0, x.shape[0], 65536):
                s = (x[c0:c0 + 65536] - mean) @ r > 0
                w = (s.astype(np.uint64) << np.arange(bits, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)
                codes[t, c0:c0 + 65536] = w
    return out[0], out[1]

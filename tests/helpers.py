"""Shared fixtures: small data sets and oracle-built indexes."""
from __future__ import annotations

import functools

import numpy as np

import oracle
from efanna_amd import datasets


@functools.lru_cache(maxsize=None)
def uniform(n=10_000, d=128, nq=100, seed=0):
    return datasets.uniform_int(n, d, nq, seed)


@functools.lru_cache(maxsize=None)
def mixture(n=20_000, d=128, nq=200, centres=50, sigma=20.0, hi=255.0, rounded=True, seed=3):
    return datasets.gaussian_mixture(n, d, nq, centres=centres, sigma=sigma, hi=hi, rounded=rounded, seed=seed)


@functools.lru_cache(maxsize=None)
def oracle_index(kind="uniform", trees=8, ml=8, epochs=8, L=30, check=25, K=10, S=10, **kw):
    base, q = uniform(**kw) if kind == "uniform" else mixture(**kw)
    ix = oracle.Index(base, trees=trees, ml=ml, epochs=epochs, check=check, L=L, K=K, S=S)
    ix.build_index()
    return base, q, ix


def recall(ids, gt, k):
    ids = np.asarray(ids)[:, :k]
    gt = np.asarray(gt)[:, :k]
    hit = sum(len(set(a.tolist()) & set(b.tolist())) for a, b in zip(ids, gt))
    return hit / (ids.shape[0] * k)

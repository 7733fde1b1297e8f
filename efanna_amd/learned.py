"""Learned stopping rule: network helpers over the C ABI (training, logits,
gradients) and the pipeline that trains a stopping network for an index on
traces of its own searches.  No counterpart in the reference: it stands in
for the fixed epoch/pool stopping rule of getNeighbors_nnexp
(kdtreeub_index.hpp:444-484)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib

N_FEAT = 16


def _p(a, t):
    return a.ctypes.data_as(t)


def param_count() -> int:
    return int(lib().efg_mlp_param_count())


def init_params(seed: int = 0) -> np.ndarray:
    w = np.empty(param_count(), np.float32)
    check(lib().efg_mlp_init(_p(w, _lib.c_f32p), ctypes.c_uint64(seed)))
    return w


def feature_norm(feat: np.ndarray) -> np.ndarray:
    """mean[16] | 1 / deviation[16] of the feature rows (deviation 0 -> 1)."""
    f = np.asarray(feat, np.float64).reshape(-1, N_FEAT)
    mu = f.mean(axis=0)
    sd = f.std(axis=0)
    sd[sd <= 0] = 1.0
    return np.concatenate([mu, 1.0 / sd]).astype(np.float32)


def logits(params, norm, feat) -> np.ndarray:
    params = np.ascontiguousarray(params, np.float32)
    norm = np.ascontiguousarray(norm, np.float32)
    feat = np.ascontiguousarray(feat, np.float32).reshape(-1, N_FEAT)
    z = np.empty(feat.shape[0], np.float32)
    check(lib().efg_mlp_logits(_p(params, _lib.c_f32p), _p(norm, _lib.c_f32p), _p(feat, _lib.c_f32p),
                               feat.shape[0], _p(z, _lib.c_f32p)))
    return z


def gradient(params, norm, feat, label):
    params = np.ascontiguousarray(params, np.float32)
    norm = np.ascontiguousarray(norm, np.float32)
    feat = np.ascontiguousarray(feat, np.float32).reshape(-1, N_FEAT)
    label = np.ascontiguousarray(label, np.uint8)
    g = np.empty(param_count(), np.float32)
    loss = ctypes.c_float()
    check(lib().efg_mlp_gradient(_p(params, _lib.c_f32p), _p(norm, _lib.c_f32p), _p(feat, _lib.c_f32p),
                                 _p(label, _lib.c_u8p), feat.shape[0], _p(g, _lib.c_f32p), ctypes.byref(loss)))
    return g, float(loss.value)


def train(feat, label, norm, params, epochs=20, batch=1024, lr=1e-3, seed=0):
    """Adam on the GPU; returns (trained parameters, mean loss per epoch)."""
    feat = np.ascontiguousarray(feat, np.float32).reshape(-1, N_FEAT)
    label = np.ascontiguousarray(label, np.uint8)
    norm = np.ascontiguousarray(norm, np.float32)
    w = np.array(params, np.float32, copy=True)
    hist = np.zeros(max(epochs, 1), np.float32)
    check(lib().efg_mlp_train(_p(feat, _lib.c_f32p), _p(label, _lib.c_u8p), feat.shape[0], _p(norm, _lib.c_f32p),
                              _p(w, _lib.c_f32p), int(epochs), int(batch), ctypes.c_float(lr), ctypes.c_uint64(seed),
                              _p(hist, _lib.c_f32p)))
    return w, hist[:epochs]


def min_match(k: int, agree: float) -> int:
    """Top-k ids that must agree with the plain search for a positive label."""
    return max(1, int(np.ceil(agree * k - 1e-9)))


def traces_to_examples(feat, label, k=None, agree=1.0):
    """Rows of the checkpoints the probe search reached, labelled 1 when at
    least min_match(k, agree) of the probe's top k are in the plain top k
    (with k None the labels are used as they are)."""
    f = np.asarray(feat).reshape(-1, N_FEAT)
    y = np.asarray(label).reshape(-1)
    keep = y != 255
    y = y[keep]
    if k is not None:
        y = (y >= min_match(k, agree)).astype(np.uint8)
    return np.ascontiguousarray(f[keep]), np.ascontiguousarray(y)


def train_for_index(index, queries, k, agree=1.0, epochs=20, batch=1024, lr=1e-3, seed=0):
    """Traces of the probe search on `queries` labelled against the plain
    search on the same index (positive: at least a fraction `agree` of the
    top k agrees), then a network trained on them.  Returns (params, norm,
    loss per epoch, positive fraction)."""
    index.setSearchMode(0)
    index.knnSearch(k, queries)
    plain, _ = index.results()
    feat, label = index.searchTraces(k, queries, plain)
    x, y = traces_to_examples(feat, label, k, agree)
    norm = feature_norm(x)
    w0 = init_params(seed)
    w, hist = train(x, y, norm, w0, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return w, norm, hist, float(y.mean()) if y.size else 0.0

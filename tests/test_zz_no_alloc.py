"""efg_knn_search_device_noalloc (the search without device allocation) against
efg_knn_search_device, called through the C ABI: identical ids and distance
bits on two calls in a row, the second with fewer queries, for FP32 rows
with queries that outgrow the LDS visited table (the global-table pass then
runs on tables cut from the fixed scratch array), BF16 rows and method 1;
then the refusals (batch too large for the scratch, index never searched)."""
import ctypes

import numpy as np
import pytest

from efanna_amd import _lib
from efanna_amd.index import FIndex, KDTreeUbIndexParams, L2DistanceAVX
from tests import helpers

pytestmark = pytest.mark.gpu


def _index(base):
    ix = FIndex(base, L2DistanceAVX(), KDTreeUbIndexParams(True, 16, 6, 6, 25, 30, 10, 8, 10))
    ix.buildIndex()
    return ix


def _queries(base, q):
    far = np.random.default_rng(5).integers(0, 256, (90, base.shape[1])).astype(np.float32)
    return np.ascontiguousarray(np.concatenate([q, far])[np.random.default_rng(6).permutation(q.shape[0] + 90)])


def _run(fn, ix, k, qd, nq, stream):
    import torch
    ids = torch.full((nq, k), -7, dtype=torch.int32, device=qd.device)
    dd = torch.full((nq, k), -7.0, dtype=torch.float32, device=qd.device)
    args = [ix.h, k, ctypes.c_void_p(qd.data_ptr()), nq, qd.shape[1], ctypes.c_void_p(ids.data_ptr()),
            ctypes.c_void_p(dd.data_ptr())]
    if fn == "efg_knn_search_device":
        args.append(None)  # per-query distance counts
    rc = getattr(_lib.lib(), fn)(*args, ctypes.c_void_p(stream.cuda_stream))
    stream.synchronize()
    return rc, ids.cpu().numpy(), dd.cpu().numpy()


@pytest.mark.parametrize("precision,method,pool,extend", [(0, 0, 1600, 400), (1, 0, 1200, 200), (0, 1, 200, 50)])
def test_noalloc_equals_device_search(cuda, precision, method, pool, extend):
    import torch
    base, q = helpers.mixture(n=20_000, nq=200, centres=50)
    ix = _index(base)
    qq = _queries(base, q)
    st = torch.cuda.Stream()
    qd = torch.from_numpy(qq).to(cuda)
    k = 10
    ix.setSearchParams(4, pool, extend, 16, -1, method)
    ix.setSearchPrecision(precision)
    for nq in (qq.shape[0], 97):  # two calls in a row, the second smaller
        rc, ref_ids, ref_d = _run("efg_knn_search_device", ix, k, qd[:nq], nq, st)
        assert rc == 0, _lib.lib().efg_last_error()
        if method == 0 and precision == 0 and nq == qq.shape[0]:
            assert ix.counters()["overflow"] > 0  # far queries take the global-table pass
        rc, ids, d = _run("efg_knn_search_device_noalloc", ix, k, qd[:nq], nq, st)
        assert rc == 0, _lib.lib().efg_last_error()
        np.testing.assert_array_equal(ids, ref_ids)
        np.testing.assert_array_equal(d.view(np.uint32), ref_d.view(np.uint32))


def test_noalloc_refusals(cuda):
    import torch
    base, q = helpers.mixture(n=20_000, nq=200, centres=50)
    ix = _index(base)
    st = torch.cuda.Stream()
    ix.setSearchParams(4, 1200, 200, 16, -1, 0)
    qd = torch.from_numpy(np.ascontiguousarray(q)).to(cuda)
    # never searched: the search trees are not laid out yet
    rc, _, _ = _run("efg_knn_search_device_noalloc", ix, 10, qd, q.shape[0], st)
    assert rc == -3 and "run one efg_knn_search" in _lib.lib().efg_last_error().decode()
    rc, _, _ = _run("efg_knn_search_device", ix, 10, qd, q.shape[0], st)
    assert rc == 0
    # 40,000 queries x 16 trees x 31 leaves do not fit the 64 MiB scratch
    big = qd.repeat(200, 1)[:40_000].contiguous()
    rc, _, _ = _run("efg_knn_search_device_noalloc", ix, 10, big, big.shape[0], st)
    msg = _lib.lib().efg_last_error().decode()
    assert rc == -2 and "fixed search scratch" in msg, msg
    limit = int(msg.split("at most ")[1].split()[0])
    assert 25_000 < limit < 40_000
    # a batch at the stated limit runs
    rc, ids, _ = _run("efg_knn_search_device_noalloc", ix, 10, big[:limit], limit, st)
    assert rc == 0, _lib.lib().efg_last_error()
    assert (ids >= 0).all()

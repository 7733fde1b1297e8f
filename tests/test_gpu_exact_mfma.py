"""GPU: exact kNN on BF16 / FP16 MFMA tiles with FP32 re-rank (precision form).

On integers 0..255 with d <= 128 the 16-bit values, the products and every
FP32 sum are exact, so the result must equal the FP32 oracle bit for bit.
On float data the re-ranked distances are the exact FP32 ones; recall against
the FP32 oracle and the error against FP64 are checked.
"""
import numpy as np
import pytest

import oracle
from efanna_amd.index import FIndex, KDTreeUbIndexParams, L2DistanceAVX, exact_knn_device
from tests import helpers

pytestmark = pytest.mark.gpu


def _run(cuda, base, q, k, exclude, precision):
    import torch
    gix = FIndex(base, L2DistanceAVX(), KDTreeUbIndexParams(True, 4))
    dptr, stride = gix.device_data()
    tq = torch.from_numpy(np.ascontiguousarray(q)).to(cuda)
    ids = torch.empty((q.shape[0], k), dtype=torch.int32, device=cuda)
    dd = torch.empty((q.shape[0], k), dtype=torch.float32, device=cuda)
    exact_knn_device(dptr, base.shape[0], base.shape[1], stride, tq.data_ptr(), q.shape[0], q.shape[1], k, exclude,
                     ids.data_ptr(), dd.data_ptr(), precision=precision)
    torch.cuda.synchronize()
    return ids.cpu().numpy(), dd.cpu().numpy()


@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("k,exclude", [(1, False), (10, False), (10, True), (100, False)])
def test_mfma_exact_integer_data_bitexact(cuda, precision, k, exclude):
    base, q = helpers.mixture(n=20_000, nq=200)
    if exclude:
        q = base[:150].copy()
    oids, od = oracle.exact_knn(base, q, k, exclude_self=exclude)
    ids, dd = _run(cuda, base, q, k, exclude, precision)
    np.testing.assert_array_equal(ids, oids)
    np.testing.assert_array_equal(dd.view(np.uint32), od.view(np.uint32))


@pytest.mark.parametrize("precision", [1, 2])
def test_mfma_exact_ragged_shapes(cuda, precision):
    # n, nq not multiples of the tiles; d not a multiple of the 64-element stage
    base, q = helpers.uniform(n=9_999, d=100, nq=77, seed=3)
    oids, od = oracle.exact_knn(base, q, 7)
    ids, dd = _run(cuda, base, q, 7, False, precision)
    np.testing.assert_array_equal(ids, oids)
    np.testing.assert_array_equal(dd.view(np.uint32), od.view(np.uint32))


@pytest.mark.parametrize("precision,d", [(1, 96), (2, 96), (1, 960)])
def test_mfma_exact_float_data(cuda, precision, d):
    n = 12_000 if d == 960 else 20_000
    base, q = helpers.mixture(n=n, nq=100, d=d, centres=20, sigma=0.05, hi=1.0, rounded=False)
    k = 10
    oids, od = oracle.exact_knn(base, q, k)
    ids, dd = _run(cuda, base, q, k, False, precision)
    # every returned distance is the exact FP32 distance of the returned id, ascending
    for i in range(0, q.shape[0], 9):
        for j in range(k):
            assert dd[i, j] == np.float32(oracle.l2(q[i], base[ids[i, j]]))
        assert (np.diff(dd[i]) >= 0).all()
    rec = helpers.recall(ids, oids, k)
    assert rec >= 0.995, rec
    d64 = ((q[:, None, :].astype(np.float64) - base[ids].astype(np.float64)) ** 2).sum(-1)
    rel = np.abs(dd - d64) / np.maximum(d64, 1e-30)
    assert rel.max() < 1e-5, rel.max()

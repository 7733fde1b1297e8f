"""GPU parity at the shapes of the high-dimensional and batched workloads,
scaled to sizes the oracle finishes in seconds, and on inputs that drive the
random-fill and capped branches.

  - 1M x 960 workload shape (K = 100 graph, L = check = 200, S = 10, trees
    for the graph a subset of the search trees) on 20K float rows: the FP32
    build equals the oracle stage by stage; the BF16 build keeps unique pools
    with exact FP32 distances and the FP32 build's graph recall; BF16-row
    search returns exact FP32 distances at the FP32 search's recall.
  - batched workload values: pools {200, 400, 800, 1600}, k {1, 10, 100},
    30 % uniform queries far from the data, so that queries outgrow the LDS
    visited table and the global-table pass runs inside the batch; every row
    equals the oracle's search with that query's parameters.
  - a hub row near every other row: reverse lists far above the cap of 100.
  - leaves of at most two rows: the search pool needs the random top-up.
"""
import numpy as np
import pytest

import oracle
from efanna_amd.index import FIndex, KDTreeUbIndexParams, L2DistanceAVX
from tests import helpers
from tests.test_gpu_build import _cmp_state, _cmp_trees
from tests.test_oracle_random import hub_data

pytestmark = pytest.mark.gpu


def _pair(base, trees, trees_build, ml, epochs, check, L, K, S, tns=10):
    params = KDTreeUbIndexParams(True, trees, ml, epochs, check, L, K, trees_build, S)
    params.TNS = tns
    g = FIndex(base, L2DistanceAVX(), params)
    o = oracle.Index(base, trees=trees, trees_build=trees_build, ml=ml, epochs=epochs, check=check, L=L, K=K, S=S,
                     tns=tns)
    return g, o


D960 = dict(trees=16, trees_build=8, ml=5, epochs=8, check=200, L=200, K=100, S=10)


def _d960():
    return helpers.mixture(n=20_000, nq=200, d=960, centres=30, sigma=0.05, hi=1.0, rounded=False)


def _graph_recall(base, graph, K, rows):
    """Recall@K of the graph rows against exact kNN (self excluded), oracle brute force."""
    gt, _ = oracle.exact_knn(base, base[rows], K + 1)
    hit = 0
    for r, g_row, t_row in zip(rows, graph[rows], gt):
        t = [x for x in t_row.tolist() if x != r][:K]
        hit += len(set(g_row[:K].tolist()) & set(t))
    return hit / (len(rows) * K)


def test_d960_shape_fp32_build_matches_oracle(cuda):
    base, _ = _d960()
    g, o = _pair(base, **D960)
    g.build_stage(0)
    o.build_trees()
    _cmp_trees(g, o)
    g.build_stage(1)
    o.merge_init()
    assert g.counters()["merge"] == o.counters()["merge"]
    _cmp_state(g, o)
    for it in range(D960["epochs"]):
        g.build_stage(2)
        o.refine_iter()
        _cmp_state(g, o)
    assert g.join_pairs() == o.join_pairs()
    g.build_stage(3)
    o.finalize()
    np.testing.assert_array_equal(g.graph_matrix(), o.graph_matrix())


def test_d960_shape_bf16_build(cuda):
    base, _ = _d960()
    g, _o = _pair(base, **D960)
    g.buildIndex()
    h, _o2 = _pair(base, **D960)
    h.setBuildPrecision(1)
    h.buildIndex()
    kh, _, nh, _ = h.pools()
    # approximate keys may differ between d(x, y) and d(y, x); no row may
    # still meet itself in a join pair and enter its own pool
    n = base.shape[0]
    live = np.arange(kh.shape[1])[None, :] < nh.astype(np.int64)[:, None]
    own = (kh & np.uint64(0xFFFFFFFF)).astype(np.int64) == np.arange(n)[:, None]
    assert not (own & live).any()
    assert not (h.graph_matrix() == np.arange(n)[:, None]).any()
    rng = np.random.default_rng(4)
    for p in rng.choice(base.shape[0], 300, replace=False):
        keys = kh[p, :nh[p]]
        assert (np.diff(keys.astype(np.uint64)) > 0).all()
        ids = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
        assert len(np.unique(ids)) == len(ids)
        for key in keys[:2 * D960["K"]:17]:
            i = int(key & 0xFFFFFFFF)
            assert np.uint32(int(key) >> 32).view(np.float32) == np.float32(oracle.l2(base[p], base[i]))
    rows = rng.choice(base.shape[0], 1000, replace=False)
    K = D960["K"]
    rg = _graph_recall(base, g.graph_matrix(), K, rows)
    rh = _graph_recall(base, h.graph_matrix(), K, rows)
    assert rh >= rg - 0.001, (rh, rg)


def test_d960_shape_bf16_rows_search(cuda):
    base, q = _d960()
    g, o = _pair(base, **D960)
    g.buildIndex()
    o.set_trees(g.trees())
    o.set_graph(*g.graph())
    k = 10
    g.setSearchParams(4, 1200, 200, 16, -1, 0)
    g.knnSearch(k, q)
    fi, fd = g.results()
    oi, od, _ = o.search(q, k, 4, 1200, 200, 16, 0)
    np.testing.assert_array_equal(fi, oi)
    np.testing.assert_array_equal(fd.view(np.uint32), od.view(np.uint32))
    g.setSearchPrecision(1)
    g.knnSearch(k, q)
    bi, bd = g.results()
    for qi in range(q.shape[0]):
        for j in range(k):
            assert bd[qi, j] == np.float32(oracle.l2(q[qi], base[bi[qi, j]]))
    gt, _ = oracle.exact_knn(base, q, k)
    rf, rb = helpers.recall(fi, gt, k), helpers.recall(bi, gt, k)
    assert rf >= 0.95 and rb >= rf - 0.005, (rf, rb)


@pytest.mark.parametrize("one_pass", [True, False])
def test_batch_workload_values_far_queries_match_oracle(cuda, one_pass):
    """The batched workload's values (pools to 1600, k to 100, 30 % uniform
    far queries), all classes in one pass or one pass per class: far queries
    outgrow the LDS visited table and continue in a spill table, and every
    row equals the oracle's search with that query's parameters."""
    base, q, oix = helpers.oracle_index("mixture", n=20_000, nq=200, centres=50, trees=16, ml=6)
    rng = np.random.default_rng(21)
    far = rng.integers(0, 256, (86, base.shape[1])).astype(np.float32)  # 30 % of the queries
    qq = np.concatenate([q, far])[rng.permutation(q.shape[0] + 86)]
    nq = qq.shape[0]
    gix = FIndex(base, L2DistanceAVX(), KDTreeUbIndexParams(True, 16, 6, 8))
    gix.set_trees(oix.trees())
    gix.set_graph(*oix.graph())
    epochs, trees = 4, 16
    gix.setSearchParams(epochs, 400, 66, trees, -1, 0)
    gix.setBatchForm(one_pass)
    pools = rng.choice([200, 400, 800, 1600], size=nq).astype(np.int32)
    ks = rng.choice([1, 10, 100], size=nq).astype(np.int32)
    ids, d = gix.knnSearchBatch(qq, pools, ks, out_stride=100)
    assert gix.counters()["overflow"] > 0  # queries outgrew the LDS table inside the batch
    assert gix.counters()["overflow2"] == 0
    for (p, k) in set(zip(pools.tolist(), ks.tolist())):
        sel = np.where((pools == p) & (ks == k))[0]
        oids, od, _ = oix.search(qq[sel], k, epochs, p, max(1, p // 6), trees, 0)
        np.testing.assert_array_equal(ids[sel, :k], oids, err_msg=f"pool {p} k {k}")
        np.testing.assert_array_equal(d[sel, :k].view(np.uint32), od.view(np.uint32))
        assert (ids[sel, k:] == -1).all()


def test_hub_reverse_lists_match_oracle(cuda):
    base = hub_data()
    g, o = _pair(base, trees=4, trees_build=4, ml=3, epochs=3, check=25, L=30, K=10, S=10)
    o.keep_reverse_lists()
    g.build_stage(0)
    o.build_trees()
    g.build_stage(1)
    o.merge_init()
    _cmp_state(g, o)
    capped = 0
    for it in range(3):
        g.build_stage(2)
        o.refine_iter()
        _cmp_state(g, o)
        for w in (0, 1):
            off, _ = o.reverse_lists(w)
            capped += int((np.diff(off.astype(np.int64)) > 100).sum())
    assert capped >= 2
    g.build_stage(3)
    o.finalize()
    np.testing.assert_array_equal(g.graph_matrix(), o.graph_matrix())


def test_search_topup_matches_oracle(cuda):
    base = np.random.default_rng(2).integers(0, 256, (2000, 16)).astype(np.float32)
    o = oracle.Index(base, trees=2, S=2, tns=2, ml=2, epochs=0)
    o.build_trees()
    rng = np.random.default_rng(3)
    o.set_graph(np.arange(2001, dtype=np.uint64) * 6, rng.integers(0, 2000, 2000 * 6).astype(np.uint32))
    g = FIndex(base, L2DistanceAVX(), KDTreeUbIndexParams(True, 2, 2, 0))
    g.set_trees(o.trees())
    g.set_graph(*o.graph())
    q = base[::50] + 0.5
    g.setSearchParams(3, 100, 20, 1, -1, 0)
    g.knnSearch(10, q)
    gi, gd = g.results()
    oi, od, ond = o.search(q, 10, 3, 100, 20, 1, 0)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))
    assert g.counters()["search_dist"] == int(ond.sum())

"""GPU parity: graph search (method 0) and exact kNN against the CPU oracle.

Search is deterministic given (graph, trees): the HIP kernels must return the
same ids, the same FP32 distances (same accumulation order) and the same
number of distance evaluations per query as the oracle.
"""
import numpy as np
import pytest

import oracle
from efanna_amd.index import FIndex, KDTreeUbIndexParams, L2DistanceAVX, Matrix
from tests import helpers

pytestmark = pytest.mark.gpu


def _gpu_index(base, trees, graph):
    ix = FIndex(Matrix(base.shape[0], base.shape[1], base), L2DistanceAVX(), KDTreeUbIndexParams(True, 8, 8, 8))
    ix.set_trees(trees)
    ix.set_graph(*graph)
    return ix


def _search_both(oix, gix, q, k, epochs, pool, extend, trees):
    gix.setSearchParams(epochs, pool, extend, trees, -1, 0)
    gix.knnSearch(k, q)
    gids, gd = gix.results()
    oids, od, ond = oix.search(q, k, epochs, pool, extend, trees, 0)
    return gids, gd, oids, od, ond


@pytest.mark.parametrize("epochs,pool,extend,trees,k", [
    (4, 120, 20, 4, 10),    # the 10K x 128 CPU configuration
    (0, 120, 20, 4, 10),    # seeds only
    (4, 120, 10, 8, 20),    # k > extend_to
    (6, 300, 300, 8, 10),   # extend_to == init_num
    (2, 60, 200, 2, 5),     # extend_to clamped to init_num
])
def test_search_matches_oracle_uniform(cuda, epochs, pool, extend, trees, k):
    base, q, oix = helpers.oracle_index("uniform")
    gix = _gpu_index(base, oix.trees(), oix.graph())
    gids, gd, oids, od, ond = _search_both(oix, gix, q, k, epochs, pool, extend, trees)
    np.testing.assert_array_equal(gids, oids)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))
    assert gix.counters()["search_dist"] == int(ond.sum())


def test_search_matches_oracle_float_mixture(cuda):
    # non-integer data: bit-identical distances need the same FMA order
    base, q, oix = helpers.oracle_index("mixture", n=20_000, nq=300, centres=40, sigma=0.05, hi=1.0,
                                        rounded=False, d=100, ml=4)
    gix = _gpu_index(base, oix.trees(), oix.graph())
    gids, gd, oids, od, ond = _search_both(oix, gix, q, 10, 4, 400, 60, 8)
    np.testing.assert_array_equal(gids, oids)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))
    assert gix.counters()["search_dist"] == int(ond.sum())


def test_search_large_pool_mixture(cuda):
    base, q, oix = helpers.oracle_index("mixture", n=40_000, nq=200, centres=40, trees=16, ml=8)
    gix = _gpu_index(base, oix.trees(), oix.graph())
    gids, gd, oids, od, ond = _search_both(oix, gix, q, 10, 4, 1200, 200, 16)
    np.testing.assert_array_equal(gids, oids)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))
    assert gix.counters()["search_dist"] == int(ond.sum())


def test_search_visited_table_fallback(cuda):
    # dense random graph: expansions overflow the LDS visited table and the
    # queries are redone with the global-memory table
    base, q = helpers.uniform(n=6000, d=32, nq=64)
    oix = oracle.Index(base, trees=4, ml=4, epochs=1)
    oix.build_trees()
    rng = np.random.default_rng(1)
    deg = 64
    ids = rng.integers(0, base.shape[0], size=base.shape[0] * deg).astype(np.uint32)
    off = np.arange(base.shape[0] + 1, dtype=np.uint64) * deg
    oix.set_graph(off, ids)
    gix = _gpu_index(base, oix.trees(), (off, ids))
    gids, gd, oids, od, ond = _search_both(oix, gix, q, 10, 4, 100, 100, 4)
    assert gix.counters()["overflow"] > 0
    np.testing.assert_array_equal(gids, oids)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))


def test_search_random_topup(cuda):
    # pool larger than what the leaves hold: the generator tops the pool up
    base, q = helpers.uniform(n=700, d=16, nq=40)
    oix = oracle.Index(base, trees=2, ml=2, epochs=2, S=2, K=2, L=6, check=4, tns=2)
    oix.build_index()
    gix = _gpu_index(base, oix.trees(), oix.graph())
    gids, gd, oids, od, ond = _search_both(oix, gix, q, 5, 3, 300, 20, 2)
    np.testing.assert_array_equal(gids, oids)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))


def test_search_errors(cuda):
    base, q, oix = helpers.oracle_index("uniform")
    gix = _gpu_index(base, oix.trees(), oix.graph())
    from efanna_amd._lib import EfgError
    gix.setSearchParams(4, 120, 20, 9, -1, 0)
    with pytest.raises(EfgError) as e:
        gix.knnSearch(10, q)
    assert e.value.code == -4 and "wrong tree number" in e.value.msg
    gix.setSearchParams(4, 120, 20, 0, -1, 0)
    with pytest.raises(EfgError):
        gix.knnSearch(10, q)
    gix.setSearchParams(4, 120, 20, 4, -1, 7)
    with pytest.raises(EfgError) as e:
        gix.knnSearch(10, q)
    assert e.value.code == -5


@pytest.mark.parametrize("k,exclude", [(10, False), (100, True), (1, False)])
def test_exact_knn_matches_oracle(cuda, k, exclude):
    import torch
    base, q = helpers.mixture(n=20_000, nq=200)
    if exclude:
        q = base[:150].copy()
    oids, od = oracle.exact_knn(base, q, k, exclude_self=exclude)
    gix = FIndex(base, L2DistanceAVX(), KDTreeUbIndexParams(True, 4))
    dptr, stride = gix.device_data()
    tq = torch.from_numpy(q).to(cuda)
    ids = torch.empty((q.shape[0], k), dtype=torch.int32, device=cuda)
    dd = torch.empty((q.shape[0], k), dtype=torch.float32, device=cuda)
    from efanna_amd.index import exact_knn_device
    exact_knn_device(dptr, base.shape[0], base.shape[1], stride, tq.data_ptr(), q.shape[0], q.shape[1], k, exclude,
                     ids.data_ptr(), dd.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ids.cpu().numpy(), oids)
    np.testing.assert_array_equal(dd.cpu().numpy().view(np.uint32), od.view(np.uint32))


def test_exact_knn_float_data_matches_oracle(cuda):
    import torch
    base, q = helpers.mixture(n=9_000, nq=64, d=96, centres=20, sigma=0.05, hi=1.0, rounded=False)
    oids, od = oracle.exact_knn(base, q, 16)
    gix = FIndex(base, L2DistanceAVX(), KDTreeUbIndexParams(True, 4))
    dptr, stride = gix.device_data()
    tq = torch.from_numpy(q).to(cuda)
    ids = torch.empty((q.shape[0], 16), dtype=torch.int32, device=cuda)
    dd = torch.empty((q.shape[0], 16), dtype=torch.float32, device=cuda)
    from efanna_amd.index import exact_knn_device
    exact_knn_device(dptr, base.shape[0], base.shape[1], stride, tq.data_ptr(), q.shape[0], q.shape[1], 16, False,
                     ids.data_ptr(), dd.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(ids.cpu().numpy(), oids)
    np.testing.assert_array_equal(dd.cpu().numpy().view(np.uint32), od.view(np.uint32))


@pytest.mark.parametrize("epochs,pool,extend,trees,k", [
    (4, 120, 20, 4, 10),    # the 10K x 128 CPU configuration, method 1
    (3, 60, 30, 4, 10),     # epochs * extend >= pool: unsorted seeds, pool raised to 90
    (2, 200, 10, 8, 20),    # k > extend_to
    (5, 300, 40, 8, 5),
    (4, 1200, 200, 8, 10),  # the 1M x 128 search parameters (lists in registers: k + extend_to + 1 <= 256)
    (2, 600, 300, 8, 10),   # k + extend_to + 1 > 256: the knn list in LDS
    (2, 400, 100, 8, 80),   # k + 1 > 64: the results list in LDS
])
def test_search_kgraph_matches_oracle(cuda, epochs, pool, extend, trees, k):
    base, q, oix = helpers.oracle_index("uniform")
    gix = _gpu_index(base, oix.trees(), oix.graph())
    gix.setSearchParams(epochs, pool, extend, trees, -1, 1)
    gix.knnSearch(k, q)
    gids, gd = gix.results()
    oids, od, ond = oix.search(q, k, epochs, pool, extend, trees, 1)
    np.testing.assert_array_equal(gids, oids)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))
    assert gix.counters()["search_dist"] == int(ond.sum())
    assert gix.getSearchParams()["init_num"] == oix.init_num_state


def test_search_kgraph_float_mixture(cuda):
    base, q, oix = helpers.oracle_index("mixture", n=20_000, nq=300, centres=40, sigma=0.05, hi=1.0,
                                        rounded=False, d=100, ml=4)
    gix = _gpu_index(base, oix.trees(), oix.graph())
    gix.setSearchParams(4, 400, 60, 8, -1, 1)
    gix.knnSearch(10, q)
    gids, gd = gix.results()
    oids, od, _ = oix.search(q, 10, 4, 400, 60, 8, 1)
    np.testing.assert_array_equal(gids, oids)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))


def test_search_bf16_rows_identical_on_integer_data(cuda):
    # BF16 holds 0..255 exactly: the precision form must return the plain results
    base, q, oix = helpers.oracle_index("mixture", n=40_000, nq=200, centres=40, trees=16, ml=8)
    gix = _gpu_index(base, oix.trees(), oix.graph())
    gix.setSearchParams(4, 1200, 200, 16, -1, 0)
    gix.knnSearch(10, q)
    plain_ids, plain_d = gix.results()
    gix.setSearchPrecision(1)
    gix.knnSearch(10, q)
    ids, d = gix.results()
    np.testing.assert_array_equal(ids, plain_ids)
    np.testing.assert_array_equal(d.view(np.uint32), plain_d.view(np.uint32))


def test_search_bf16_rows_float_data(cuda):
    base, q, oix = helpers.oracle_index("mixture", n=20_000, nq=300, centres=40, sigma=0.05, hi=1.0,
                                        rounded=False, d=100, ml=4)
    gix = _gpu_index(base, oix.trees(), oix.graph())
    gix.setSearchParams(4, 400, 60, 8, -1, 0)
    gix.setSearchPrecision(1)
    gix.knnSearch(10, q)
    ids, d = gix.results()
    # returned distances are the exact FP32 distances of the returned ids, ascending
    for i in range(0, q.shape[0], 7):
        for j in range(10):
            assert d[i, j] == np.float32(oracle.l2(q[i], base[ids[i, j]]))
        assert (np.diff(d[i]) >= 0).all()
    gt, _ = oracle.exact_knn(base, q, 10)
    oids, _, _ = oix.search(q, 10, 4, 400, 60, 8, 0)
    assert helpers.recall(ids, gt, 10) >= helpers.recall(oids, gt, 10) - 0.01


@pytest.mark.parametrize("threads", [64, 128, 256])
def test_search_threads_per_query_same_results(cuda, threads):
    # launch shape only: every width returns the oracle's results and counts
    base, q, oix = helpers.oracle_index("mixture", n=20_000, nq=200, centres=50)
    gix = _gpu_index(base, oix.trees(), oix.graph())
    gix.setSearchThreads(threads)
    for epochs, pool, extend, trees in [(4, 120, 20, 8), (4, 600, 100, 8), (2, 1000, 60, 4)]:
        if pool > 16 * threads:
            continue
        gids, gd, oids, od, ond = _search_both(oix, gix, q, 10, epochs, pool, extend, trees)
        np.testing.assert_array_equal(gids, oids)
        np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))


def test_autotune_keeps_results(cuda):
    base, q, oix = helpers.oracle_index("mixture", n=20_000, nq=200, centres=50)
    gix = _gpu_index(base, oix.trees(), oix.graph())
    gix.setSearchParams(4, 300, 50, 8, -1, 0)
    full, probe = gix.autotune(10, q)
    assert full in (64, 128, 256) and probe == 0
    gix.knnSearch(10, q)
    gids, _ = gix.results()
    oids, _, _ = oix.search(q, 10, 4, 300, 50, 8, 0)
    np.testing.assert_array_equal(gids, oids)


@pytest.mark.parametrize("one_pass", [True, False])
@pytest.mark.parametrize("tune", [False, True])
def test_batch_heterogeneous_matches_oracle(cuda, tune, one_pass):
    """Per-query pool, extend_to and k in one call, all classes in one pass
    or one pass per class: every row equals the oracle's search with that
    query's parameters."""
    base, q, oix = helpers.oracle_index("mixture", n=20_000, nq=200, centres=50)
    gix = _gpu_index(base, oix.trees(), oix.graph())
    gix.setBatchForm(one_pass)
    epochs, trees = 4, 8
    gix.setSearchParams(epochs, 300, 50, trees, -1, 0)
    rng = np.random.default_rng(11)
    pools = rng.choice([60, 120, 240, 480], size=q.shape[0]).astype(np.int32)
    ks = rng.choice([1, 10, 50], size=q.shape[0]).astype(np.int32)
    ks = np.minimum(ks, pools)
    ext = np.maximum(1, pools // 6).astype(np.int32)
    ext[::5] = pools[::5] // 3  # some explicit extend_to values
    ids, d = gix.knnSearchBatch(q, pools, ks, extend=ext, out_stride=64, tune=tune)
    assert ids.shape == (q.shape[0], 64)
    for (p, e, k) in set(zip(pools.tolist(), ext.tolist(), ks.tolist())):
        sel = np.where((pools == p) & (ext == e) & (ks == k))[0]
        oids, od, _ = oix.search(q[sel], k, epochs, p, e, trees, 0)
        np.testing.assert_array_equal(ids[sel, :k], oids)
        np.testing.assert_array_equal(d[sel, :k].view(np.uint32), od.view(np.uint32))
        assert (ids[sel, k:] == -1).all()


@pytest.mark.parametrize("one_pass", [True, False])
def test_batch_plan_graph_replays_match_eager(cuda, one_pass):
    """The batch captured into a HIP graph: every replay, with eager searches
    and synchronisations in between and new queries written into the same
    buffer, equals the eager batch on those queries."""
    import torch
    base, q, oix = helpers.oracle_index("mixture", n=20_000, nq=200, centres=50)
    gix = _gpu_index(base, oix.trees(), oix.graph())
    gix.setBatchForm(one_pass)
    gix.setSearchParams(4, 300, 50, 8, -1, 0)
    nq = q.shape[0]
    rng = np.random.default_rng(12)
    pools = rng.choice([60, 120, 240], size=nq).astype(np.int32)
    ks = rng.choice([1, 10, 40], size=nq).astype(np.int32)
    _, q2 = helpers.mixture(n=20_000, nq=2 * nq, centres=50)
    batches = [q, q2[nq:], q2[:nq]]
    stride = 40
    st = torch.cuda.Stream()
    qd = torch.from_numpy(np.ascontiguousarray(q)).to(cuda)
    ids = torch.empty((nq, stride), dtype=torch.int32, device=cuda)
    dd = torch.empty((nq, stride), dtype=torch.float32, device=cuda)
    plan = gix.batch_plan(qd.data_ptr(), nq, q.shape[1], pools, ks, stride, ids.data_ptr(), dd.data_ptr(),
                          stream=st.cuda_stream)
    try:
        for rep in range(6):
            qb = batches[rep % len(batches)]
            ei, ed = gix.knnSearchBatch(qb, pools, ks, out_stride=stride)  # eager, on the host path
            with torch.cuda.stream(st):
                qd.copy_(torch.from_numpy(np.ascontiguousarray(qb)).to(cuda))
            plan.run(st.cuda_stream)
            st.synchronize()
            np.testing.assert_array_equal(ids.cpu().numpy(), ei, err_msg=f"replay {rep}")
            np.testing.assert_array_equal(dd.cpu().numpy().view(np.uint32), ed.view(np.uint32))
            gix.knnSearch(10, qb)  # other eager work on the index between replays
            torch.cuda.synchronize()
    finally:
        plan.close()


def test_search_irregular_graph_matches_oracle(cuda):
    """Rows of different degrees (offsets not row * degree) take the CSR
    path of the neighbour gather; a K-regular graph takes the direct one."""
    base, q, oix = helpers.oracle_index("mixture", n=20_000, nq=200, centres=50)
    off, ids = oix.graph()
    off = off.astype(np.int64)
    deg = np.diff(off)
    keep = deg - (np.arange(deg.size) % 3 == 0)  # every third row loses its last neighbour
    noff = np.concatenate([[0], np.cumsum(keep)]).astype(np.uint64)
    nids = np.concatenate([ids[off[i]:off[i] + keep[i]] for i in range(deg.size)]).astype(np.uint32)
    o2 = oracle.Index(base, trees=8, ml=8, epochs=8, check=25, L=30, K=10, S=10)
    o2.set_trees(oix.trees())
    o2.set_graph(noff, nids)
    gix = _gpu_index(base, oix.trees(), (noff, nids))
    gids, gd, oids, od, ond = _search_both(o2, gix, q, 10, 4, 300, 50, 8)
    np.testing.assert_array_equal(gids, oids)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))
    assert gix.counters()["search_dist"] == int(ond.sum())


def _caterpillar_trees(base, leaf=2):
    """Two search trees over `base`.  Tree 0: each internal node splits off
    the `leaf` rows with the largest first coordinate as its right child (a
    leaf), so the tree is n / leaf levels deep and a query that goes left at
    the top has far more pending far children than a short stack holds.
    Tree 1: a single leaf holding every row."""
    n = base.shape[0]
    order = np.argsort(base[:, 0], kind="stable").astype(np.uint32)
    m = n // leaf - 1  # internal nodes of tree 0
    size = 2 * m + 1 + 1
    dim = np.zeros(size, np.int32)
    val = np.zeros(size, np.float32)
    left = np.full(size, -1, np.int32)
    right = np.full(size, -1, np.int32)
    start = np.zeros(size, np.uint32)
    end = np.zeros(size, np.uint32)
    for k in range(m):  # internal node k covers positions [0, n - k * leaf)
        hi = n - k * leaf
        val[k] = base[order[hi - leaf], 0]
        left[k] = k + 1 if k + 1 < m else 2 * m
        right[k] = m + k
        end[k] = hi
        start[m + k], end[m + k] = hi - leaf, hi
    end[2 * m] = n - m * leaf  # last left leaf: [0, n - m * leaf)
    end[2 * m + 1] = n  # tree 1
    leaves = np.stack([order, np.arange(n, dtype=np.uint32)])
    return dict(dim=dim, val=val, left=left, right=right, start=start, end=end,
                roots=np.array([0, 2 * m + 1], np.int32), leaf=leaves)


def test_search_deep_tree_parent_walk(cuda):
    """Seed leaves past the descent stack's depth: on the caterpillar tree the
    first leaf is 99 levels down, so after the 16 deepest far children the
    descent climbs parent links for every further leaf, and the one-leaf tree
    records its root and stops."""
    rng = np.random.default_rng(5)
    n, d = 200, 8
    base = rng.random((n, d), dtype=np.float32)
    q = rng.random((30, d), dtype=np.float32)
    q[:10, 0] *= 0.05  # left at the top of tree 0: the deepest walks
    trees = _caterpillar_trees(base)
    oix = oracle.Index(base, trees=2, ml=1, epochs=1, S=2, K=4, L=8, check=4)
    oix.set_trees(trees)
    gt, _ = oracle.exact_knn(base, base, 6, exclude_self=True)
    off = np.arange(n + 1, dtype=np.uint64) * 6
    ids = gt.astype(np.uint32).ravel()
    oix.set_graph(off, ids)
    gix = _gpu_index(base, trees, (off, ids))
    # pool 200 over 2 trees: 41 leaves per tree, 25 of them past the stack
    gids, gd, oids, od, ond = _search_both(oix, gix, q, 10, 2, 200, 40, 2)
    np.testing.assert_array_equal(gids, oids)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))
    assert gix.counters()["search_dist"] == int(ond.sum())

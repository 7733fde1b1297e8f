"""GPU parity of the graph build against the oracle, stage by stage.

With the counter-based generator both sides make the same random choices, so
trees, merge/init state, every NN-descent iteration and the final graph must
be identical.
"""
import numpy as np
import pytest

import oracle
from efanna_amd.index import FIndex, KDTreeUbIndexParams, L2DistanceAVX
from tests import helpers

pytestmark = pytest.mark.gpu


def _pair(base, trees=8, ml=8, epochs=8, check=25, L=30, K=10, S=10, tns=10):
    params = KDTreeUbIndexParams(True, trees, ml, epochs, check, L, K, trees, S)
    params.TNS = tns
    g = FIndex(base, L2DistanceAVX(), params)
    o = oracle.Index(base, trees=trees, ml=ml, epochs=epochs, check=check, L=L, K=K, S=S, tns=tns)
    return g, o


def _cmp_trees(g, o):
    tg, to = g.trees(), o.trees()
    for key in ("dim", "left", "right", "start", "end", "roots", "leaf"):
        np.testing.assert_array_equal(tg[key], to[key], err_msg=key)
    np.testing.assert_array_equal(tg["val"].view(np.uint32), to["val"].view(np.uint32))


def _cmp_state(g, o, lists=True):
    kg, fg, ng, rg = g.pools()
    ko, fo, no, ro = o.pools()
    np.testing.assert_array_equal(ng, no, err_msg="fill")
    np.testing.assert_array_equal(rg, ro, err_msg="range")
    np.testing.assert_array_equal(kg, ko, err_msg="pool keys")
    np.testing.assert_array_equal(fg, fo, err_msg="flags")
    if lists:
        for which in (0, 1):
            og, ig = g.lists(which)
            oo, io = o.lists(which)
            np.testing.assert_array_equal(og, oo, err_msg=f"list offsets {which}")
            for i in range(len(og) - 1):
                a = np.sort(ig[og[i]:og[i + 1]])
                b = np.sort(io[oo[i]:oo[i + 1]])
                if not np.array_equal(a, b):
                    raise AssertionError(f"list {which} of node {i} differs: {a} vs {b}")


@pytest.mark.parametrize("n,d,trees", [(10_000, 128, 8), (3000, 20, 3)])
def test_trees_match_oracle(cuda, n, d, trees):
    base, _ = helpers.uniform(n=n, d=d, nq=1)
    g, o = _pair(base, trees=trees)
    g.buildTrees()
    o.build_trees()
    _cmp_trees(g, o)


def test_trees_match_oracle_mixture_dfs_phase(cuda):
    # enough nodes per level to leave the BFS phase (|active| >= 1100)
    base, _ = helpers.mixture(n=40_000, nq=1)
    g, o = _pair(base, trees=8, S=6, tns=12)
    g.buildTrees()
    o.build_trees()
    _cmp_trees(g, o)


@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("rows", ["integer", "float"])
def test_precision_form_trees_match_oracle(cuda, precision, rows):
    """The precision form builds its trees from the 16-bit high parts of the
    split rows when every row is exact in them (integer rows 0..255, in BF16
    and in FP16), else from the FP32 rows (float rows); either way the trees
    are the oracle's, through both the BFS and the DFS phase."""
    if rows == "integer":
        base, _ = helpers.mixture(n=40_000, nq=1)
    else:
        base, _ = helpers.mixture(n=40_000, nq=1, d=96, centres=40, sigma=0.05, hi=1.0, rounded=False)
    g, o = _pair(base, trees=8, S=6, tns=12)
    g.setBuildPrecision(precision)
    g.buildTrees()
    o.build_trees()
    _cmp_trees(g, o)


@pytest.mark.parametrize("direct", [True, False])
def test_build_stages_match_oracle(cuda, direct):
    """Every stage equals the oracle, with the join's candidates routed
    straight into their targets' buckets (default) or staged and scattered."""
    base, _ = helpers.uniform()
    g, o = _pair(base)
    g.setJoinRouting(direct)
    g.build_stage(0)
    o.build_trees()
    _cmp_trees(g, o)
    g.build_stage(1)
    o.merge_init()
    assert g.counters()["merge"] == o.counters()["merge"]
    assert g.counters()["init"] == o.counters()["init"]
    _cmp_state(g, o)
    for it in range(8):
        g.build_stage(2)
        o.refine_iter()
        _cmp_state(g, o)
    assert g.join_pairs() == o.join_pairs()
    g.build_stage(3)
    o.finalize()
    np.testing.assert_array_equal(g.graph_matrix(), o.graph_matrix())


@pytest.mark.parametrize("L,K,check,S", [(12, 8, 12, 8), (80, 20, 40, 12), (300, 20, 40, 12)])
def test_build_pool_lengths_match_oracle(cuda, L, K, check, S):
    """Pool lengths on either side of the update kernels' lane groups (16
    lanes per node for L = 12; 64 lanes and several passes above 64) and of
    the register pool merge's list buffers (64, 256 and 512 entries)."""
    base, _ = helpers.uniform(n=4000)
    g, o = _pair(base, trees=4, ml=5, epochs=5, check=check, L=L, K=K, S=S)
    g.build_stage(0)
    o.build_trees()
    g.build_stage(1)
    o.merge_init()
    _cmp_state(g, o)
    for it in range(5):
        g.build_stage(2)
        o.refine_iter()
        _cmp_state(g, o)
    assert g.join_pairs() == o.join_pairs()
    g.build_stage(3)
    o.finalize()
    np.testing.assert_array_equal(g.graph_matrix(), o.graph_matrix())


@pytest.mark.parametrize("precision", [0, 1])
def test_build_join_staging_reruns_keep_results(cuda, precision):
    """A staging margin far below the kept candidate fraction makes the
    larger iterations join again with room for every candidate (the staging
    always has one 4096-slot run per wavefront besides, so this takes a few
    hundred thousand rows), here with every join kernel staged; the
    precision form's block tasks, staged beside direct routing, rerun through
    the same code.  Graph and join pairs equal the default build's (direct
    routing; the FP32 build equals the oracle in the tests above, the
    precision form equals it on integer rows)."""
    base, _ = helpers.mixture(n=300_000, nq=1)
    g, _o = _pair(base, trees=8, ml=8, epochs=6)
    g.buildIndex()
    h, _o2 = _pair(base, trees=8, ml=8, epochs=6)
    h.setBuildPrecision(precision)
    h.setJoinRouting(False)
    h.setJoinStagingMargin(0.01)
    h.buildIndex()
    assert g.counters()["join_reruns"] == 0
    assert h.counters()["join_reruns"] > 0
    np.testing.assert_array_equal(h.graph_matrix(), g.graph_matrix())
    assert h.join_pairs() == g.join_pairs()


@pytest.mark.parametrize("d,direct", [(100, True), (200, True), (200, False)])
def test_build_float_data_long_lists_matches_oracle(cuda, d, direct):
    # non-integer rows (every FMA rounds), d % 32 != 0 (tail chunk first),
    # more than 64 columns (several column chunks per join task) and lists
    # longer than one 32-row block (tasks with I != J); block tasks with the
    # candidates routed directly or staged
    base, _ = helpers.mixture(n=6000, nq=1, d=d, centres=30, sigma=0.05, hi=1.0, rounded=False)
    g, o = _pair(base, trees=4, ml=6, epochs=5, check=40, L=60, K=20, S=12)
    g.setJoinRouting(direct)
    g.build_stage(0)
    o.build_trees()
    g.build_stage(1)
    o.merge_init()
    _cmp_state(g, o)
    longest = 0
    for it in range(5):
        g.build_stage(2)
        o.refine_iter()
        _cmp_state(g, o)
        on, _ = g.lists(0)
        oo, _ = g.lists(1)
        longest = max(longest, int(np.max(np.diff(on) + np.diff(oo))))
    assert longest > 64
    assert g.join_pairs() == o.join_pairs()


def test_build_index_mixture_matches_oracle(cuda):
    base, _ = helpers.mixture(n=30_000, nq=1)
    g, o = _pair(base, trees=8, ml=8, epochs=6)
    g.buildIndex()
    o.build_index()
    np.testing.assert_array_equal(g.graph_matrix(), o.graph_matrix())


def test_merge_level_error(cuda):
    base, _ = helpers.uniform(n=300, d=8, nq=1)
    g, _o = _pair(base, trees=2, ml=12, epochs=1)
    from efanna_amd._lib import EfgError
    with pytest.raises(EfgError) as e:
        g.buildIndex()
    assert e.value.code == -3 and "merge level deeper than tree" in e.value.msg


@pytest.mark.parametrize("precision", [1, 2])
def test_build_precision_form_integer_data_equals_fp32_build(cuda, precision):
    """BF16 / FP16 MFMA merge and join tiles: on integers 0..255 with d = 128
    every distance is exact, so the build equals the FP32 build bit for bit
    (graph, pools and per-iteration join pairs)."""
    base, _ = helpers.mixture(n=30_000, nq=1)
    g, _o = _pair(base, trees=8, ml=8, epochs=6)
    g.buildIndex()
    h, _o2 = _pair(base, trees=8, ml=8, epochs=6)
    h.setBuildPrecision(precision)
    h.buildIndex()
    np.testing.assert_array_equal(h.graph_matrix(), g.graph_matrix())
    assert h.join_pairs() == g.join_pairs()
    kg, fg, ng, _ = g.pools()
    kh, fh, nh, _ = h.pools()
    np.testing.assert_array_equal(nh, ng)
    np.testing.assert_array_equal(kh, kg)


def test_build_precision_form_float_data(cuda):
    """On float rows the BF16 tiles change the build slightly: the graph's
    recall against exact kNN stays within 0.01 of the FP32 build's, and the
    final pools hold exact FP32 distances in ascending order."""
    base, _ = helpers.mixture(n=8000, nq=1, d=200, centres=30, sigma=0.05, hi=1.0, rounded=False)
    K = 20
    g, _o = _pair(base, trees=4, ml=4, epochs=6, check=40, L=40, K=K, S=12)
    g.buildIndex()
    h, _o2 = _pair(base, trees=4, ml=4, epochs=6, check=40, L=40, K=K, S=12)
    h.setBuildPrecision(1)
    h.buildIndex()
    gt, _ = oracle.exact_knn(base, base, K, exclude_self=True)
    rg = helpers.recall(g.graph_matrix(), gt, K)
    rh = helpers.recall(h.graph_matrix(), gt, K)
    assert rh >= rg - 0.01, (rh, rg)
    kh, _, nh, _ = h.pools()
    L = kh.shape[1]
    rows = np.random.default_rng(1).choice(base.shape[0], 200, replace=False)
    for p in rows:
        keys = kh[p, :nh[p]]
        assert (np.diff(keys.astype(np.uint64)) > 0).all()
        for key in keys[:5]:
            i = int(key & 0xFFFFFFFF)
            dist = np.uint32(int(key) >> 32).view(np.float32)
            assert dist == np.float32(oracle.l2(base[p], base[i]))


def test_build_precision_form_exact_rows_large_norms(cuda):
    """Rows exact in BF16 (multiples of 256 up to 65280) skip the low parts,
    but their squared norms are far above 2^22, so the FP32 sums of the tiles
    round and the build keeps its re-rank and id dedup: pools unique, no row
    in its own pool, exact FP32 distances in the final pools, recall close to
    the FP32 build's."""
    base, _ = helpers.mixture(n=6000, nq=1, d=64, centres=30, sigma=20.0, hi=255.0, rounded=True)
    base = base * np.float32(256.0)
    K = 16
    g, _o = _pair(base, trees=4, ml=4, epochs=6, check=32, L=32, K=K, S=10)
    g.buildIndex()
    h, _o2 = _pair(base, trees=4, ml=4, epochs=6, check=32, L=32, K=K, S=10)
    h.setBuildPrecision(1)
    h.buildIndex()
    kh, _, nh, _ = h.pools()
    n = base.shape[0]
    live = np.arange(kh.shape[1])[None, :] < nh.astype(np.int64)[:, None]
    ids = (kh & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert not ((ids == np.arange(n)[:, None]) & live).any()
    for p in range(0, n, 11):
        row = ids[p, :nh[p]]
        assert len(np.unique(row)) == len(row)
        keys = kh[p, :nh[p]]
        assert (np.diff(keys.astype(np.uint64)) > 0).all()
        for key in keys[:4]:
            i = int(key & 0xFFFFFFFF)
            assert np.uint32(int(key) >> 32).view(np.float32) == np.float32(oracle.l2(base[p], base[i]))
    gt, _ = oracle.exact_knn(base, base, K, exclude_self=True)
    rg = helpers.recall(g.graph_matrix(), gt, K)
    rh = helpers.recall(h.graph_matrix(), gt, K)
    assert rh >= rg - 0.01, (rh, rg)


@pytest.mark.parametrize("precision", [1, 2])
def test_build_precision_form_pools_unique(cuda, precision):
    """A pair's 16-bit-tile distance can depend on which block holds which
    row, so the precision form's merge drops duplicate ids: no pool keeps one
    id twice."""
    base, _ = helpers.mixture(n=8000, nq=1, d=200, centres=30, sigma=0.05, hi=1.0, rounded=False)
    h, _o = _pair(base, trees=4, ml=4, epochs=3, check=40, L=40, K=20, S=12)
    h.setBuildPrecision(precision)
    h.build_stage(0)
    h.build_stage(1)
    for it in range(3):
        h.build_stage(2)
        kh, _, nh, _ = h.pools()
        for p in range(0, base.shape[0], 7):
            ids = (kh[p, :nh[p]] & np.uint64(0xFFFFFFFF)).astype(np.int64)
            assert len(np.unique(ids)) == len(ids), (it, p)


@pytest.mark.parametrize("precision", [1, 2])
def test_build_precision_form_long_lists_equals_fp32_build(cuda, precision):
    """Lists of more than 64 rows: the MFMA block tasks (beside the whole-node
    MFMA kernel) give the FP32 build's pools and graph bit for bit on integer
    rows (the FP16 form with its candidates staged, the others routed
    directly)."""
    base, _ = helpers.mixture(n=20_000, nq=1)
    kw = dict(trees=4, ml=6, epochs=4, check=40, L=80, K=40, S=20)
    g, _o = _pair(base, **kw)
    g.build_stage(0)
    g.build_stage(1)
    h, _o2 = _pair(base, **kw)
    h.setBuildPrecision(precision)
    h.setJoinRouting(precision == 1)
    h.build_stage(0)
    h.build_stage(1)
    longest = 0
    for it in range(4):
        g.build_stage(2)
        h.build_stage(2)
        on, _ = g.lists(0)
        oo, _ = g.lists(1)
        longest = max(longest, int(np.max(np.diff(on) + np.diff(oo))))
        kg, fg, ng, _ = g.pools()
        kh, fh, nh, _ = h.pools()
        np.testing.assert_array_equal(nh, ng)
        np.testing.assert_array_equal(kh, kg)
        np.testing.assert_array_equal(fh, fg)
    assert longest > 64
    assert h.join_pairs() == g.join_pairs()

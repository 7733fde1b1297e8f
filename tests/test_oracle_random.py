"""Pins the oracle's random-fill branches by what the mathematics fixes.

The reference draws these from rand()/mt19937, so no fixture can pin the ids
themselves; what every correct implementation shares is checked instead:
  GenRandom (kdtreeub_index.hpp:1121-1135): `size` distinct ids in [0, N),
      a cyclic window when N = size + 1, no preferred region;
  init top-up (1399-1405): the pool holds the merge's candidates plus
      distinct extra ids, S of them, with their true distances, and the
      nn_new tail is GenRandom's (1389-1392);
  search top-up (411-417): the pool is the leaf members in sequence order,
      then distinct ids outside them, init_num in all;
  reverse lists (base_index.hpp:426-434): at most 100 entries, a subset of
      the uncapped list, equal to it when it is short, not simply its first
      100 members.
"""
import numpy as np
import pytest

import oracle
from tests.test_oracle import _py_merge

SEED = 1998


# --------------------------------------------------------------------------
# GenRandom
# --------------------------------------------------------------------------
@pytest.mark.parametrize("N,size", [(2, 1), (3, 2), (21, 20), (25, 20), (40, 20), (1000, 20), (10 ** 6, 20),
                                    (100, 99), (7, 3)])
def test_gen_random_distinct_in_range(N, size):
    for node in range(0, 400, 7):
        a = oracle.gen_random(SEED, node, size, N)
        assert len(a) == size
        assert (a < N).all()
        assert len(set(a.tolist())) == size, (node, a)


@pytest.mark.parametrize("size", [1, 5, 20])
def test_gen_random_window_when_one_spare(size):
    # N - size = 1: every draw is 0, the fix-up makes 0..size-1, the offset
    # rotates that window; so the ids are consecutive modulo N
    N = size + 1
    for node in range(50):
        a = oracle.gen_random(SEED, node, size, N)
        off = int(a[0])
        assert a.tolist() == [(off + i) % N for i in range(size)]


def test_gen_random_covers_the_range():
    # 4000 nodes x 20 ids over N = 1000: every decile within 20% of its share
    # (an offset or modulus mistake piles ids into part of the range)
    N, size = 1000, 20
    h = np.zeros(10)
    for node in range(4000):
        h += np.bincount(oracle.gen_random(7, node, size, N) * 10 // N, minlength=10)
    share = 4000 * size / 10
    assert (np.abs(h - share) < 0.2 * share).all(), h


# --------------------------------------------------------------------------
# init top-up and nn_new fill
# --------------------------------------------------------------------------
@pytest.mark.parametrize("n,ml", [(200, 4), (300, 5)])
def test_init_topup(n, ml):
    S = 10
    base = np.random.default_rng(0).integers(0, 256, (n, 16)).astype(np.float32)
    o = oracle.Index(base, trees=2, ml=ml, epochs=0, S=S, K=S, L=20, check=15)
    o.build_trees()
    o.merge_init()
    assert o.counters()["init"] > 0  # the top-up ran
    heaps, extra = _py_merge(base, o.trees(), 2, ml, S)
    keys, flags, fill, rng = o.pools()
    off, nn_new = o.lists(0)
    topped = 0
    for p in range(n):
        k = keys[p, :S]
        d = (k >> np.uint64(32)).astype(np.uint32).view(np.float32)
        ids = (k & np.uint64(0xFFFFFFFF)).astype(np.int64)
        assert fill[p] == S and rng[p] == S and (flags[p, :S] == 1).all()
        assert (np.diff(k.astype(np.uint64).view(np.int64)) > 0).all()  # sorted by (dist, id), distinct
        assert len(set(ids.tolist())) == S and (ids >= 0).all() and (ids < n).all()
        for j in range(S):
            assert d[j] == np.float32(oracle.l2(base[p], base[ids[j]]))
        nw = nn_new[off[p]:off[p + 1]]
        assert len(nw) == 2 * S and nw[:S].tolist() == ids.tolist()
        if len(heaps[p]) < S:
            topped += 1
            # every merge candidate stays; the rest are extra ids
            assert {c[1] for c in heaps[p]} <= set(ids.tolist())
            # no merge overflow, so nn_new was filled by GenRandom and its
            # first S entries overwritten with the pool
            assert len(extra[p]) == 0
            assert nw[S:].tolist() == oracle.gen_random(SEED, p, 2 * S, n)[S:].tolist()
    assert topped > 0


# --------------------------------------------------------------------------
# search top-up
# --------------------------------------------------------------------------
def _leaf_members(t, tree, q, lsize):
    """First lsize leaves of a depth-first descent, query side first (kdtreeub_index.hpp:332-346)."""
    out = []

    def dfs(x):
        if len(out) >= lsize:
            return
        if t["left"][x] >= 0:
            a, b = (t["left"][x], t["right"][x]) if q[t["dim"][x]] < t["val"][x] else (t["right"][x], t["left"][x])
            dfs(a)
            dfs(b)
        else:
            out.append(x)

    dfs(t["roots"][tree])
    return out


def test_search_topup():
    # leaves of at most 2 rows: 41 leaves reach at most 82 ids, so a pool of
    # 100 needs the random top-up
    base = np.random.default_rng(2).integers(0, 256, (2000, 16)).astype(np.float32)
    o = oracle.Index(base, trees=1, S=2, tns=2, ml=2, epochs=0)
    o.build_trees()
    t = o.trees()
    init = 100
    lsize = init * 2 // 5 + 1
    for qi in range(20):
        q = base[qi * 37] + 0.5
        seq = []
        for lf in _leaf_members(t, 0, q, lsize):
            for j in range(t["start"][lf], t["end"][lf]):
                m = int(t["leaf"][0, j])
                if m not in seq:
                    seq.append(m)
        assert len(seq) < init
        seeds = o.search_seeds(q, qi, init, 1)
        assert seeds[:len(seq)].tolist() == seq
        rest = seeds[len(seq):].tolist()
        assert len(set(rest)) == len(rest) and not set(rest) & set(seq)
        assert all(0 <= x < 2000 for x in rest)
    # the search returns its whole pool when k = init_num and no epoch runs
    g = np.tile(np.arange(4, dtype=np.uint32), 2000)
    o.set_graph(np.arange(2001, dtype=np.uint64) * 4, g)
    ids, d, _ = o.search(base[:3] + 0.5, init, 0, init, 10, 1)
    for qi in range(3):
        assert sorted(ids[qi].tolist()) == sorted(o.search_seeds(base[qi] + 0.5, qi, init, 1).tolist())


# --------------------------------------------------------------------------
# reverse lists around the cap of 100
# --------------------------------------------------------------------------
def hub_data(n=600, d=32, seed=1):
    """Rows on shells of radius 20-40 around row 0: row 0 is near everyone,
    so its reverse lists run far past 100."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, d))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = rng.uniform(20, 40, size=(n, 1))
    base = (128 + u * r).astype(np.float32)
    base[0] = 128
    return base


def _rows(off, ids):
    return [ids[off[i]:off[i + 1]].tolist() for i in range(len(off) - 1)]


def test_reverse_list_cap():
    base = hub_data()
    o = oracle.Index(base, trees=4, ml=3, epochs=3, S=10, K=10, L=30, check=25)
    o.keep_reverse_lists()
    o.build_trees()
    o.merge_init()
    capped = 0
    for _ in range(3):
        o.refine_iter()
        for w in range(2):
            raw = _rows(*o.reverse_lists(w))
            kept = _rows(*o.reverse_lists(w + 2))
            lists = _rows(*o.lists(w))
            for x in range(base.shape[0]):
                r, k = raw[x], kept[x]
                assert k == sorted(k) and len(set(k)) == len(k)
                if len(r) <= 100:
                    assert k == r
                else:
                    capped += 1
                    assert len(k) == 100 and set(k) <= set(r)
                    assert k != r[:100] and k != r[-100:]
                # the join list ends with the kept reverse list
                assert lists[x][len(lists[x]) - len(k):] == k
    assert capped >= 2

"""The oracle's hashing index (IEH seeds, algorithm/hashing_index.hpp) pinned
on the CPU: mask order against itertools' lexicographic combinations, the
seed sequence against a numpy restatement of the bucket walk, the hash-only
search against a brute-force top K of those candidates, and the nnexp
search over a full pool against exact kNN."""
import itertools

import numpy as np
import pytest

import oracle
from efanna_amd import datasets
from tests import helpers


def _hash_index(n=1500, d=16, nq=40, tables=3, bits=10, radius=2, seed=0):
    base, q = datasets.uniform_int(n, d, nq, seed)
    bc, qc = datasets.lsh_codes(base, q, tables, bits, seed)
    o = oracle.Index(base, trees=1)
    o.set_hash(bc, qc, bits, radius)
    return base, q, bc, qc, o


@pytest.mark.parametrize("bits,radius", [(6, 3), (10, 2), (12, 0), (40, 2)])
def test_masks_are_lexicographic_combinations_by_radius(bits, radius):
    base, q = datasets.uniform_int(200, 8, 4, 1)
    o = oracle.Index(base, trees=1)
    codes = np.zeros((1, 200), np.uint64)
    o.set_hash(codes, np.zeros((1, 4), np.uint64), bits, radius)
    want = [0]
    for r in range(1, radius + 1):
        for comb in itertools.combinations(range(bits), r):
            want.append(sum(1 << i for i in comb))
    assert o.hash_masks().tolist() == want


def _py_seeds(bc, qc, masks, qi, init_num):
    """Bucket walk: mask outer, table inner, members in id order, first
    occurrences, up to init_num."""
    seen, out = set(), []
    for m in masks:
        for t in range(bc.shape[0]):
            members = np.flatnonzero(bc[t] == (qc[t, qi] ^ m))
            for x in members.tolist():
                if len(out) >= init_num:
                    return out
                if x in seen:
                    continue
                seen.add(x)
                out.append(x)
    return out


@pytest.mark.parametrize("init_num", [30, 200])
def test_seed_sequence_matches_bucket_walk(init_num):
    base, q, bc, qc, o = _hash_index()
    masks = o.hash_masks()
    for qi in range(0, 40, 3):
        want = _py_seeds(bc, qc, masks, qi, init_num)
        got = o.hash_seeds(qi, init_num, topup_to=0)  # no top-up: the bucket sequence only
        assert got.tolist() == want, qi


def test_hash_only_search_is_top_k_of_candidates():
    base, q, bc, qc, o = _hash_index(radius=1)
    k, init_num = 10, 60
    ids, ds, nd = o.hash_search(q, k, 0, init_num, 20, method=1)  # no graph: hash-only
    masks = o.hash_masks()
    for qi in range(q.shape[0]):
        cand = _py_seeds(bc, qc, masks, qi, init_num)
        if len(cand) < k:  # the reference tops up to K only
            cand = o.hash_seeds(qi, init_num, topup_to=k).tolist()
        assert nd[qi] == len(cand)
        d = ((base[cand].astype(np.int64) - q[qi].astype(np.int64)) ** 2).sum(1)
        order = np.lexsort((np.array(cand), d))[:k]
        assert ids[qi].tolist() == [cand[i] for i in order]
        assert ds[qi].tolist() == d[order].astype(np.float32).tolist()


def test_topup_only_below_k():
    # one table, codes all distinct: radius 0 finds one candidate per query
    base, q = datasets.uniform_int(300, 8, 5, 2)
    o = oracle.Index(base, trees=1)
    bc = np.arange(300, dtype=np.uint64)[None] * 3
    qc = bc[:, :5].copy()
    o.set_hash(bc, qc, 12, 0)
    s = o.hash_seeds(0, 50, topup_to=4)
    assert len(s) == 4 and s[0] == 0
    s = o.hash_seeds(0, 50, topup_to=1)
    assert s.tolist() == [0]


def test_nnexp_full_pool_is_exact_knn():
    base, q, bc, qc, o = _hash_index(n=600)
    o.set_graph(np.arange(601, dtype=np.uint64) * 2, np.random.default_rng(0).integers(0, 600, 1200).astype(np.uint32))
    ids, ds, nd = o.hash_search(q, 10, 2, 600, 20, method=1)
    gt, gd = oracle.exact_knn(base, q, 10)
    np.testing.assert_array_equal(ids, gt)
    np.testing.assert_array_equal(ds, gd)


def test_hash_search_errors():
    base, q, bc, qc, o = _hash_index(nq=5)
    with pytest.raises(oracle.OracleError):
        o.hash_search(np.concatenate([q, q]), 10, 0, 50, 20)  # more queries than query codes
    o.set_graph(np.arange(base.shape[0] + 1, dtype=np.uint64), np.zeros(base.shape[0], np.uint32))
    with pytest.raises(oracle.OracleError):
        o.hash_search(q, 10, 2, 50, 20, method=2)

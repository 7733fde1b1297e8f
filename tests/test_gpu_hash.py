"""GPU parity of the hashing index (IEH seeds, algorithm/hashing_index.hpp)
against the oracle: the hash-only search (no graph), nnexp and kgraph
expansion from bucket seeds, 32- and 64-bit codes, the code-length shift,
the reference's code files, and the parameter checks."""
import struct

import numpy as np
import pytest

import oracle
from efanna_amd import datasets
from efanna_amd._lib import EfgError
from efanna_amd.index import FIndex, KDTreeUbIndexParams, L2DistanceAVX, Matrix

pytestmark = pytest.mark.gpu


def _pair(n=3000, d=16, nq=60, tables=3, bits=12, radius=2, lenshift=0, graph_k=0, seed=0):
    base, q = datasets.uniform_int(n, d, nq, seed)
    bc, qc = datasets.lsh_codes(base, q, tables, bits, seed)
    o = oracle.Index(base, trees=1)
    o.set_hash(bc >> np.uint64(lenshift), qc >> np.uint64(lenshift), bits - lenshift, radius)
    g = FIndex(Matrix(n, d, base), L2DistanceAVX(), KDTreeUbIndexParams(True, 8, 8, 8))
    g.setHashCodes(bc, qc, bits, upbits=min(4, bits - lenshift - 1), radius=radius, lenshift=lenshift)
    g.buildIndex()
    if graph_k:
        gt, _ = oracle.exact_knn(base, base, graph_k, exclude_self=True)
        off = np.arange(n + 1, dtype=np.uint64) * graph_k
        ids = gt.astype(np.uint32).ravel()
        o.set_graph(off, ids)
        g.set_graph(off, ids)
    return base, q, o, g


def _both(o, g, q, k, epochs, pool, extend, method):
    g.setSearchParams(epochs, pool, extend, 0, -1, method)
    g.knnSearch(k, q)
    gi, gd = g.results()
    oi, od, ond = o.hash_search(q, k, epochs, pool, extend, method)
    return gi, gd, oi, od, ond


# codes of up to 22 bits use the dense bucket table, longer ones the binary search
@pytest.mark.parametrize("bits,radius,pool,k", [(12, 2, 100, 10), (20, 1, 50, 10), (12, 0, 40, 5), (30, 1, 60, 10)])
def test_hash_only_search_matches_oracle(cuda, bits, radius, pool, k):
    base, q, o, g = _pair(bits=bits, radius=radius)
    gi, gd, oi, od, ond = _both(o, g, q, k, 4, pool, 20, 1)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))
    assert g.counters()["search_dist"] == int(ond.sum())


@pytest.mark.parametrize("method", [1, 0])
def test_hash_graph_search_matches_oracle(cuda, method):
    # method 1 = nnexp, 0 = kgraph on a hashing index
    base, q, o, g = _pair(graph_k=10)
    gi, gd, oi, od, ond = _both(o, g, q, 10, 4, 200, 40, method)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))
    assert g.counters()["search_dist"] == int(ond.sum())


@pytest.mark.parametrize("lenshift", [18, 4])
def test_hash_64bit_codes_and_shift(cuda, lenshift):
    base, q, o, g = _pair(bits=40, radius=1, lenshift=lenshift, graph_k=8)
    gi, gd, oi, od, ond = _both(o, g, q, 8, 3, 120, 30, 1)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd.view(np.uint32), od.view(np.uint32))


def _write_codes(prefix, codes, bits):
    for t in range(codes.shape[0]):
        with open(f"{prefix}_{t + 1}", "wb") as f:
            f.write(struct.pack("<iii", 1, bits, codes.shape[1]))
            f.write(codes[t].astype(np.uint32 if bits <= 32 else np.uint64).tobytes())


@pytest.mark.parametrize("bits", [12, 40])
def test_hash_code_files(cuda, tmp_path, bits):
    """The reference's code files load to the same index as the arrays."""
    n, d, nq, tables = 2000, 16, 30, 2
    base, q = datasets.uniform_int(n, d, nq, 4)
    bc, qc = datasets.lsh_codes(base, q, tables, bits, 4)
    _write_codes(tmp_path / "base", bc, bits)
    _write_codes(tmp_path / "query", qc, bits)
    g1 = FIndex(Matrix(n, d, base), L2DistanceAVX(), KDTreeUbIndexParams(True, 8, 8, 8))
    g1.loadHashCodes(tables, bits, 4, 1, tmp_path / "base", tmp_path / "query")
    g1.buildIndex()
    g2 = FIndex(Matrix(n, d, base), L2DistanceAVX(), KDTreeUbIndexParams(True, 8, 8, 8))
    g2.setHashCodes(bc, qc, bits, 4, 1)
    g2.buildIndex()
    for g in (g1, g2):
        g.setSearchParams(0, 60, 20, 0, -1, 1)
        g.knnSearch(10, q)
    np.testing.assert_array_equal(g1.results()[0], g2.results()[0])
    with pytest.raises(EfgError):
        g1.loadHashCodes(tables + 1, bits, 4, 1, tmp_path / "base", tmp_path / "query")  # no third file


def test_hash_errors(cuda):
    n, d = 500, 8
    base, q = datasets.uniform_int(n, d, 4, 5)
    g = FIndex(Matrix(n, d, base), L2DistanceAVX(), KDTreeUbIndexParams(True, 8, 8, 8))
    z = np.zeros((1, n), np.uint64)
    zq = np.zeros((1, 4), np.uint64)
    with pytest.raises(EfgError):
        g.setHashCodes(z, zq, 12, upbits=12, radius=1)  # upbits must stay below the code length
    big = z.copy()
    big[0, 3] = 1 << 12
    with pytest.raises(EfgError):
        g.setHashCodes(big, zq, 12, upbits=4, radius=1)  # code above 2^12 - 1
    with pytest.raises(EfgError):
        g.setHashCodes(np.zeros((2, n), np.uint64), np.zeros((2, 4), np.uint64), 32, upbits=8, radius=3)  # 5489 x 2 slots
    g.setHashCodes(z, zq, 12, upbits=4, radius=1)
    g.buildIndex()
    g.setSearchParams(0, 20, 10, 0, -1, 1)
    with pytest.raises(EfgError):
        g.knnSearch(5, np.concatenate([q, q]))  # 8 queries, 4 query codes

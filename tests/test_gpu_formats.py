"""Reference file formats and the command-line tools, on the GPU library."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from efanna_amd._lib import EfgError
from efanna_amd.index import FIndex, KDTreeUbIndexParams, L2DistanceAVX
from tests import formats, helpers

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "efanna_amd", "_lib", "bin")


@pytest.fixture(scope="module")
def built():
    base, q = helpers.uniform(n=6000, d=32, nq=50)
    ix = FIndex(base, L2DistanceAVX(), KDTreeUbIndexParams(True, 6, 4, 4, 20, 24, 8, 4, 8))
    ix.buildIndex()
    ix.setSearchParams(3, 100, 20, 6, -1, 0)
    ix.knnSearch(8, q)
    return base, q, ix


def _arrays_from_file(t):
    """Node arrays (all trees) from a parsed trees file."""
    dim, val, left, right, start, end, roots = [], [], [], [], [], [], []
    for tr in t["trees"]:
        off = len(dim)
        for r in tr["nodes"]:
            inner = r["left"] is not None
            dim.append(r["dim"] if inner else -1)
            val.append(r["val"])
            left.append(off + r["left"] if inner else -1)
            right.append(off + r["right"] if inner else -1)
            start.append(r["start"])
            end.append(r["end"])
        roots.append(off + tr["root"])
    return dict(dim=np.array(dim, np.int32), val=np.array(val, np.float32), left=np.array(left, np.int32),
                right=np.array(right, np.int32), start=np.array(start, np.uint32), end=np.array(end, np.uint32),
                roots=np.array(roots, np.int32), leaf=t["leaf"])


def test_save_graph_bytes(built, tmp_path):
    base, q, ix = built
    ix.saveGraph(tmp_path / "g.graph")
    g = ix.graph_matrix()
    assert (tmp_path / "g.graph").read_bytes() == formats.graph_bytes(g)
    assert ix.getGraphSize() == base.shape[0]
    assert ix.getGraphRow(17) == g[17].tolist()
    assert ix.getGraphRow(base.shape[0] + 5) == []


def test_save_trees_layout_and_reload(built, tmp_path):
    base, q, ix = built
    ix.saveTrees(tmp_path / "t.trees")
    t = formats.read_trees((tmp_path / "t.trees").read_bytes())
    assert (t["K"], t["T"], t["num"], t["dim"]) == (8, 6, base.shape[0], base.shape[1])
    assert all(tr["node_size"] == 48 for tr in t["trees"])
    mine = ix.trees()
    np.testing.assert_array_equal(t["leaf"], mine["leaf"])
    a = _arrays_from_file(t)
    # same trees: compare node by node along a joint traversal
    for tr in range(6):
        st = [(int(a["roots"][tr]), int(mine["roots"][tr]))]
        while st:
            x, y = st.pop()
            assert (a["start"][x], a["end"][x]) == (mine["start"][y], mine["end"][y])
            assert (a["left"][x] < 0) == (mine["left"][y] < 0)
            if a["left"][x] >= 0:
                assert a["dim"][x] == mine["dim"][y] and a["val"][x] == mine["val"][y]
                st += [(int(a["left"][x]), int(mine["left"][y])), (int(a["right"][x]), int(mine["right"][y]))]
    # reload into a fresh index: identical search
    ix2 = FIndex(base, L2DistanceAVX(), KDTreeUbIndexParams(True, 8, 8, 10, 25, 30, 10))
    ix2.loadTrees(tmp_path / "t.trees")
    ix.saveGraph(tmp_path / "g.graph")
    ix2.loadGraph(tmp_path / "g.graph")
    assert ix2.getSearchParams()["search_trees"] == 6  # SP.tree_num from the file
    ix2.setSearchParams(3, 100, 20, 6, -1, 0)
    ix2.knnSearch(8, q)
    np.testing.assert_array_equal(ix2.results()[0], ix.results()[0])


def test_save_load_index_and_results(built, tmp_path):
    base, q, ix = built
    ix.saveIndex(tmp_path / "all.index")
    ix2 = FIndex(base, L2DistanceAVX(), KDTreeUbIndexParams(True, 8))
    ix2.loadIndex(tmp_path / "all.index")
    ix2.setSearchParams(3, 100, 20, 6, -1, 0)
    ix2.knnSearch(8, q)
    np.testing.assert_array_equal(ix2.results()[0], ix.results()[0])
    ix.saveResults(tmp_path / "r.bin")
    assert (tmp_path / "r.bin").read_bytes() == formats.results_bytes(ix.results()[0])


def test_format_errors(built, tmp_path):
    base, q, ix = built
    ix2 = FIndex(base, L2DistanceAVX(), KDTreeUbIndexParams(True, 4))
    with pytest.raises(EfgError) as e:
        ix2.loadGraph(tmp_path / "missing.graph")
    assert e.value.code == -6
    ix.saveTrees(tmp_path / "t.trees")
    data = (tmp_path / "t.trees").read_bytes()
    (tmp_path / "cut.trees").write_bytes(data[: len(data) // 2])
    with pytest.raises(EfgError) as e:
        ix2.loadTrees(tmp_path / "cut.trees")
    assert e.value.code == -7
    small, _ = helpers.uniform(n=100, d=32, nq=1)
    other = FIndex(small, L2DistanceAVX(), KDTreeUbIndexParams(True, 4))
    with pytest.raises(EfgError) as e:
        other.loadTrees(tmp_path / "t.trees")  # built for another data set
    assert e.value.code == -7
    ix.saveGraph(tmp_path / "g.graph")
    with pytest.raises(EfgError):
        other.loadGraph(tmp_path / "g.graph")


def test_cli_pipeline_matches_oracle(tmp_path):
    base, q = helpers.uniform(n=5000, d=32, nq=40)
    formats.write_vecs(tmp_path / "base.fvecs", base)
    formats.write_vecs(tmp_path / "q.fvecs", q)
    run = lambda *a: subprocess.run([str(x) for x in a], capture_output=True, text=True, check=True).stdout
    run(os.path.join(BIN, "efanna_index_buildall"), tmp_path / "base.fvecs", tmp_path / "g.graph",
        tmp_path / "t.trees", 4, 4, 4, 30, 25, 10, 10)
    run(os.path.join(BIN, "efanna_search"), tmp_path / "base.fvecs", tmp_path / "t.trees", tmp_path / "g.graph",
        tmp_path / "q.fvecs", tmp_path / "r.bin", 4, 4, 120, 20, 10)
    res = np.fromfile(tmp_path / "r.bin", np.int32).reshape(40, 11)
    assert (res[:, 0] == 10).all()
    # the oracle on the files the tools wrote gives the same results
    o = oracle.Index(base, trees=4)
    o.set_trees(_arrays_from_file(formats.read_trees((tmp_path / "t.trees").read_bytes())))
    rows = formats.read_graph((tmp_path / "g.graph").read_bytes())
    off = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint64)
    o.set_graph(off, np.concatenate(rows))
    oids, _, _ = o.search(q, 10, 4, 120, 20, 4, 0)
    np.testing.assert_array_equal(res[:, 1:], oids)
    gt, _ = oracle.exact_knn(base, q, 10)
    formats.write_vecs(tmp_path / "gt.ivecs", gt, np.int32)
    formats.write_vecs(tmp_path / "res.ivecs", res[:, 1:], np.int32)
    out = run(os.path.join(BIN, "evaluate"), tmp_path / "res.ivecs", tmp_path / "gt.ivecs", 10)
    rec = float(out.strip().split()[-1])
    assert abs(rec - helpers.recall(oids, gt, 10)) < 1e-6

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


# --------------------------------------------------------------------------
# index image (efg_save_image / efg_load_image)
# --------------------------------------------------------------------------
def _image_pair(tmp_path):
    base, q = helpers.mixture(n=20_000, nq=100, centres=50)
    p = KDTreeUbIndexParams(True, 8, 6, 5, 25, 30, 10, 4, 10)
    a = FIndex(base, L2DistanceAVX(), p)
    a.buildIndex()
    a.setSearchParams(4, 300, 50, 8, -1, 0)
    a.knnSearch(10, q)
    ra = a.results()
    a.saveImage(tmp_path / "ix.img")
    return base, q, p, a, ra


def test_image_round_trip(tmp_path):
    base, q, p, a, (ia, da) = _image_pair(tmp_path)
    b = FIndex(base, L2DistanceAVX(), p)
    b.loadImage(tmp_path / "ix.img")
    b.setSearchParams(4, 300, 50, 8, -1, 0)
    b.knnSearch(10, q)
    ib, db = b.results()
    np.testing.assert_array_equal(ib, ia)
    np.testing.assert_array_equal(db.view(np.uint32), da.view(np.uint32))
    ta, tb = a.trees(), b.trees()
    for key in ta:
        np.testing.assert_array_equal(ta[key], tb[key], err_msg=key)
    np.testing.assert_array_equal(a.graph_matrix(), b.graph_matrix())
    # the reference formats written from the reloaded index are the same bytes
    a.saveTrees(tmp_path / "a.trees")
    b.saveTrees(tmp_path / "b.trees")
    ta_ = formats.read_trees((tmp_path / "a.trees").read_bytes())
    tb_ = formats.read_trees((tmp_path / "b.trees").read_bytes())
    assert [[{k: v for k, v in r.items() if k not in ("lp", "rp")} for r in t["nodes"]] for t in ta_["trees"]] == \
        [[{k: v for k, v in r.items() if k not in ("lp", "rp")} for r in t["nodes"]] for t in tb_["trees"]]
    np.testing.assert_array_equal(ta_["leaf"], tb_["leaf"])
    a.saveGraph(tmp_path / "a.graph")
    b.saveGraph(tmp_path / "b.graph")
    assert (tmp_path / "a.graph").read_bytes() == (tmp_path / "b.graph").read_bytes()
    # layout: the independent parser finds the documented sections, checksums included
    data = (tmp_path / "ix.img").read_bytes()
    hdr, secs = formats.read_image(data)
    assert hdr[0] == b"EFGIMG01" and hdr[3] == base.shape[0] and len(secs) == 12
    assert [s[0] for s in secs] == list(range(1, 13))
    for sid, elem, count, off, chk in secs:
        assert off % 4096 == 0 and formats.image_hash(data[off:off + elem * count]) == chk


def test_image_reloads_index_from_reference_files(tmp_path):
    """Trees and graph loaded from the reference formats, saved as an image
    and reloaded, search the same."""
    base, q, p, a, (ia, da) = _image_pair(tmp_path)
    a.saveTrees(tmp_path / "t.trees")
    a.saveGraph(tmp_path / "t.graph")
    b = FIndex(base, L2DistanceAVX(), p)
    b.loadTrees(tmp_path / "t.trees")
    b.loadGraph(tmp_path / "t.graph")
    b.saveImage(tmp_path / "b.img")
    c = FIndex(base, L2DistanceAVX(), p)
    c.loadImage(tmp_path / "b.img")
    c.setSearchParams(4, 300, 50, 8, -1, 0)
    c.knnSearch(10, q)
    ic, dc = c.results()
    np.testing.assert_array_equal(ic, ia)
    np.testing.assert_array_equal(dc.view(np.uint32), da.view(np.uint32))


def test_image_refusals(tmp_path):
    base, q, p, a, (ia, _) = _image_pair(tmp_path)
    good = (tmp_path / "ix.img").read_bytes()
    hdr, secs = formats.read_image(good)
    gids = secs[9]  # graph ids
    b = FIndex(base, L2DistanceAVX(), p)
    b.loadImage(tmp_path / "ix.img")
    b.setSearchParams(4, 300, 50, 8, -1, 0)

    def refused(data, what):
        (tmp_path / "bad.img").write_bytes(bytes(data))
        with pytest.raises(EfgError) as e:
            b.loadImage(tmp_path / "bad.img")
        assert e.value.code == -7 and what in e.value.msg, e.value.msg

    flipped = bytearray(good)
    flipped[gids[3] + 40] ^= 1
    refused(flipped, "checksum")
    refused(good[:gids[3] + 100], "inconsistent")
    refused(b"EFGIMG02" + good[8:], "not an index image")
    # an id out of range with valid checksums: the device range check refuses it
    tampered = bytearray(good)
    struct_at = gids[3] + 4 * 123
    tampered[struct_at:struct_at + 4] = np.uint32(base.shape[0] + 7).tobytes()
    formats.resign_image(tampered)
    refused(tampered, "out of range")
    # a tree link out of range, re-signed
    tampered = bytearray(good)
    sn = secs[5]  # search nodes: 32 bytes each, right child at byte 12
    tampered[sn[3] + 12:sn[3] + 16] = np.int32(10 ** 8).tobytes()
    formats.resign_image(tampered)
    refused(tampered, "out of range")
    # other data: rows differ
    other = FIndex(base[:15_000], L2DistanceAVX(), p)
    with pytest.raises(EfgError) as e:
        other.loadImage(tmp_path / "ix.img")
    assert e.value.code == -7 and "other data" in e.value.msg
    # after the refusals the index still holds the good image
    b.knnSearch(10, q)
    np.testing.assert_array_equal(b.results()[0], ia)

"""File formats and command-line tools that need no GPU."""
import os
import subprocess

import numpy as np

from tests import formats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "efanna_amd", "_lib", "bin")


def test_trees_reader_on_hand_built_file():
    # one tree: root [0,4) split dim 1 at 2.5 -> leaf [0,2), leaf [2,4); post-order
    recs = [dict(dim=0, val=0.0, start=0, end=2), dict(dim=0, val=0.0, start=2, end=4),
            dict(dim=1, val=2.5, start=0, end=4, lp=11, rp=22)]
    data = formats.trees_bytes(10, 4, 3, [recs], np.array([[3, 1, 0, 2]]))
    assert len(data) == 24 + 8 + 3 * 40 + 16
    t = formats.read_trees(data)
    tr = t["trees"][0]
    root = tr["nodes"][tr["root"]]
    assert tr["root"] == 2 and root["left"] == 0 and root["right"] == 1 and root["dim"] == 1
    assert t["leaf"].tolist() == [[3, 1, 0, 2]] and t["end"] == len(data)


def test_graph_and_results_layouts():
    rows = [[1, 2], [0], [0, 1, 3], []]
    b = formats.graph_bytes(rows)
    assert len(b) == 4 * (4 + 6)
    assert [r.tolist() for r in formats.read_graph(b)] == rows
    assert formats.results_bytes([[5, 6]]) == b"\x02\x00\x00\x00\x05\x00\x00\x00\x06\x00\x00\x00"


def test_tools_are_built():
    for tool in ("efanna_index_buildgraph", "efanna_index_buildtrees", "efanna_index_buildall", "efanna_search",
                 "evaluate"):
        assert os.access(os.path.join(BIN, tool), os.X_OK), tool


def test_evaluate_tool(tmp_path):
    rng = np.random.default_rng(0)
    gt = np.stack([rng.permutation(50)[:10] for _ in range(20)]).astype(np.int32)
    res = gt.copy()
    res[:5, :3] = 100 + np.arange(3)  # 15 misses of 200
    formats.write_vecs(tmp_path / "r.ivecs", res, np.int32)
    formats.write_vecs(tmp_path / "g.ivecs", gt, np.int32)
    out = subprocess.run([os.path.join(BIN, "evaluate"), str(tmp_path / "r.ivecs"), str(tmp_path / "g.ivecs"), "10"],
                         capture_output=True, text=True, check=True).stdout
    assert "10NN accuracy: 0.925" in out


def test_facade_compiles(tmp_path):
    src = tmp_path / "user.cc"
    src.write_text("""
#include <efanna.hpp>
using namespace efanna;
int main() {
  std::vector<float> d(16 * 8, 1.0f);
  Matrix<float> m(16, 5, d.data());
  L2DistanceAVX<float> l2;
  float x = l2.compare(m.get_row(0), m.get_row(1), 5);
  KDTreeUbIndexParams p(true, 4, 2, 2, 10, 12, 5, 0, 4);
  char* bf = (char*)"base_codes";
  char* qf = (char*)"query_codes";
  HASHINGIndexParams hp(32, 4, 16, 2, bf, qf, 0);
  (void)p; (void)x; (void)hp;
  if (0) { FIndex<float> ix(m, &l2, p); ix.buildIndex(); ix.setSearchParams(2, 10, 5, 4); ix.knnSearch(5, m); }
  if (0) { FIndex<float> ih(m, &l2, hp); ih.buildIndex(); ih.setSearchParams(2, 10, 5, 0, -1, 1); ih.knnSearch(5, m); }
  return 0;
}
""")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include", "efanna"), "-I",
                    os.path.join(ROOT, "include"), str(src), "-o", str(tmp_path / "user"), "-L",
                    os.path.join(ROOT, "efanna_amd", "_lib"), "-lefg"], check=True)

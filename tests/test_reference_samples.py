"""The reference's own sample programs compile and link unchanged against the
efanna facade (include/efanna) and libefg.so.

The sources are read where they lie in the reference tree; the test skips
when that tree is absent.  Boost is not installed, so tests/compat holds a
wall-clock stand-in for boost/timer/timer.hpp, the only Boost header the
samples include.  Nothing here runs the programs (that needs a GPU)."""
import glob
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SAMPLES = "/root/reference/samples"
LIBDIR = os.path.join(ROOT, "efanna_amd", "_lib")

SAMPLES = sorted(glob.glob(os.path.join(REF_SAMPLES, "*.cc")))


@pytest.mark.skipif(not SAMPLES, reason="reference samples not present")
@pytest.mark.parametrize("src", SAMPLES, ids=[os.path.basename(s) for s in SAMPLES])
def test_reference_sample_compiles_against_facade(src, tmp_path):
    out = tmp_path / os.path.basename(src).replace(".cc", "")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wno-unused-variable", "-Wno-sign-compare",
           "-I", os.path.join(ROOT, "tests", "compat"), "-I", os.path.join(ROOT, "include", "efanna"),
           "-I", os.path.join(ROOT, "include"), src, "-o", str(out), "-L", LIBDIR, "-lefg",
           "-Wl,-rpath," + LIBDIR]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    assert os.access(out, os.X_OK)


def test_facade_rejects_non_l2_distance(tmp_path):
    src = tmp_path / "other.cc"
    src.write_text(r"""
#include <efanna.hpp>
#include <cstdio>
using namespace efanna;
struct L1 : Distance<float> {
  float compare(const float* a, const float* b, size_t n) const { float s = 0; for (size_t i = 0; i < n; ++i) s += a[i] > b[i] ? a[i] - b[i] : b[i] - a[i]; return s; }
  float norm(const float*, size_t) const { return 0; }
  float dot(const float*, const float*, size_t) const { return 0; }
};
int main() {
  std::vector<float> d(64 * 8, 1.0f);
  Matrix<float> m(64, 8, d.data());
  L1 l1;
  try { FIndex<float> ix(m, &l1, KDTreeUbIndexParams(true, 2)); } catch (const std::invalid_argument& e) { std::printf("rejected: %s\n", e.what()); }
  try { FIndex<float> ix(m, nullptr, KDTreeUbIndexParams(true, 2)); } catch (const std::invalid_argument& e) { std::printf("rejected null\n"); }
  static_assert(KGRAPH_MATRIX_ALIGN % 32 == 0, "row alignment");
  return 0;
}
""")
    exe = tmp_path / "other"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include", "efanna"), "-I",
                    os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L", LIBDIR, "-lefg",
                    "-Wl,-rpath," + LIBDIR], check=True)
    # the rejection happens before any device call, so it runs without a GPU
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    assert "rejected: unsupported distance" in out and "rejected null" in out

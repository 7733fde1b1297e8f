"""Host sanitizer builds (CPU only): the oracle under AddressSanitizer and
UndefinedBehaviorSanitizer through a driver that runs every build stage,
both search methods, the hashing-index search and exact kNN; the C++ facade's
host-side checks under the same sanitizers.  OpenMP runtime races are not
covered (libgomp is not instrumented, so ThreadSanitizer would report false
positives)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "efanna_amd", "_lib")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]


def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitize", "-s"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", OMP_NUM_THREADS="4")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_build", "oracle_sanitize")], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0 and "sanitize run ok" in r.stdout, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]


def test_facade_host_checks_under_asan_ubsan(tmp_path):
    src = tmp_path / "facade.cc"
    src.write_text(r"""
#include <efanna.hpp>
#include <cstdio>
using namespace efanna;
struct Cosine : Distance<float> {
  float compare(const float*, const float*, size_t) const { return 0; }
  float norm(const float*, size_t) const { return 0; }
  float dot(const float*, const float*, size_t) const { return 0; }
};
int main() {
  std::vector<float> d(40 * 16, 2.0f);
  Matrix<float> m(40, 13, d.data());
  L2DistanceSSE<float> l2;
  volatile float x = l2.compare(m.get_row(1), m.get_row(2), 13);
  try { m.get_row(40); return 1; } catch (const std::runtime_error&) {}
  Cosine c;
  try { FIndex<float> ix(m, &c, KDTreeUbIndexParams(true, 2)); return 2; } catch (const std::invalid_argument&) {}
  std::printf("facade ok %g\n", (double)x);
  return 0;
}
""")
    exe = tmp_path / "facade"
    subprocess.run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "include", "efanna"), "-I",
                    os.path.join(ROOT, "include"), str(src), "-o", str(exe), "-L", LIBDIR, "-lefg",
                    "-Wl,-rpath," + LIBDIR], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    assert r.returncode == 0 and "facade ok 0" in r.stdout, r.stdout + r.stderr[-3000:]

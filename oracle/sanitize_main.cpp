// Host sanitizer run of the oracle (test infrastructure): builds a small
// index stage by stage, searches with both methods, runs exact kNN and the
// hashing-index search, under AddressSanitizer and UndefinedBehaviorSanitizer
// (make -C oracle sanitize).  Exit status 0 and no sanitizer report = pass.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

extern "C" {
const char* orc_last_error();
uint64_t orc_hash4(uint64_t, uint64_t, uint64_t, uint64_t);
void orc_set_threads(int);
void orc_exact_knn(const float*, size_t, size_t, size_t, const float*, size_t, size_t, int, int, int32_t*, float*);
void* orc_new(const float*, size_t, size_t, size_t);
void orc_free(void*);
void orc_set_params(void*, int, int, int, int, int, int, int, int, int, uint64_t, int);
void orc_keep_rev(void*, int);
int orc_build_trees(void*);
int orc_merge_init(void*);
int orc_refine_iter(void*);
int orc_finalize(void*);
int orc_search(void*, const float*, size_t, size_t, int, int, int*, int, int, int, int32_t*, float*, uint64_t*);
int orc_hash_set(void*, int, int, int, const uint64_t*, const uint64_t*, size_t);
int orc_hash_search(void*, const float*, size_t, size_t, int, int, int*, int, int, int32_t*, float*, uint64_t*);
}

static int fail(const char* what) {
  std::fprintf(stderr, "sanitize run: %s failed: %s\n", what, orc_last_error());
  return 1;
}

int main() {
  orc_set_threads(4);
  const size_t n = 3000, d = 20, nq = 40;  // d % 8 != 0: padded rows
  std::vector<float> x(n * d), q(nq * d);
  for (size_t i = 0; i < x.size(); ++i) x[i] = float(orc_hash4(1, 0, i, 0) % 256);
  for (size_t i = 0; i < q.size(); ++i) q[i] = float(orc_hash4(1, 1, i, 0) % 256);
  for (int mode = 0; mode < 2; ++mode) {
    void* h = orc_new(x.data(), n, d, d);
    orc_set_params(h, 4, 4, 4, 3, 20, 24, 10, 8, 10, 1998, mode);
    orc_keep_rev(h, 1);
    if (orc_build_trees(h)) return fail("trees");
    if (orc_merge_init(h)) return fail("merge");
    for (int e = 0; e < 3; ++e)
      if (orc_refine_iter(h)) return fail("refine");
    if (orc_finalize(h)) return fail("finalize");
    std::vector<int32_t> ids(nq * 10);
    std::vector<float> dd(nq * 10);
    std::vector<uint64_t> nd(nq);
    for (int method = 0; method < 2; ++method) {
      int init = 60;
      if (orc_search(h, q.data(), nq, d, 10, 3, &init, 12, 4, method, ids.data(), dd.data(), nd.data()))
        return fail("search");
    }
    if (mode == 0) {
      std::vector<uint64_t> bc(2 * n), qc(2 * nq);
      for (size_t i = 0; i < bc.size(); ++i) bc[i] = orc_hash4(2, 0, i, 0) & 0xfff;
      for (size_t i = 0; i < qc.size(); ++i) qc[i] = orc_hash4(2, 1, i, 0) & 0xfff;
      if (orc_hash_set(h, 2, 12, 2, bc.data(), qc.data(), nq)) return fail("hash set");
      int init = 60;
      if (orc_hash_search(h, q.data(), nq, d, 10, 2, &init, 12, 1, ids.data(), dd.data(), nd.data()))
        return fail("hash search");
    }
    orc_free(h);
  }
  std::vector<float> xp(n * 24, 0.f), qp(nq * 24, 0.f);
  for (size_t i = 0; i < n; ++i) std::memcpy(&xp[i * 24], &x[i * d], d * 4);
  for (size_t i = 0; i < nq; ++i) std::memcpy(&qp[i * 24], &q[i * d], d * 4);
  std::vector<int32_t> gi(nq * 7);
  std::vector<float> gd(nq * 7);
  orc_exact_knn(xp.data(), n, d, 24, qp.data(), nq, 24, 7, 0, gi.data(), gd.data());
  std::printf("sanitize run ok\n");
  return 0;
}

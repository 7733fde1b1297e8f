// CPU oracle for the EFANNA kNN-graph build and graph search path.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain C++ restatement of the
// reference algorithm (ZJULearning/efanna, header-only C++) used as the checker
// for the HIP library and as the CPU baseline that bench.py times.  Nothing in
// the product library links, loads or calls it.
//
// Every function cites the reference file:line it restates.  Where the
// reference depends on unspecified or thread-order-dependent state (libc rand(),
// std::random_shuffle, mt19937 seeded per OpenMP thread, lock order), the oracle
// replaces that state with the counter-based generator below, which the HIP
// library implements identically, so that GPU and oracle agree bit for bit.
//
// Build: see oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off; all fused
// multiply-adds are written explicitly as fmaf()).

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>
#include <omp.h>

namespace {

thread_local std::string g_err;

// ---------------------------------------------------------------------------
// Counter-based generator (replaces rand()/random_shuffle/mt19937 streams).
// splitmix64 finaliser applied to a (seed, stream, a, b) tuple.
// ---------------------------------------------------------------------------
inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
inline uint64_t hash4(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b) {
  return mix64(mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL)) ^ a) ^ b);
}
enum : uint64_t {
  ST_PERM = 1,      // initial leaf-list permutation per tree (kdtreeub_index.hpp:1313-1319)
  ST_SPLIT = 2,     // choice among the top-5 variance dims (kdtreeub_index.hpp:997)
  ST_INIT_TOP = 3,  // top-up of the merge heap to S (kdtreeub_index.hpp:1399-1405)
  ST_INIT_NEW = 4,  // GenRandom fill of nn_new (kdtreeub_index.hpp:1390-1393, 1121-1135)
  ST_RNN_NEW = 5,   // reverse-list sampling (base_index.hpp:426-429)
  ST_RNN_OLD = 6,   // reverse-list sampling (base_index.hpp:431-434)
  ST_SEARCH = 7,    // random top-up of the search seed pool (kdtreeub_index.hpp:411-417)
};

// Feistel bijection on [0, 2^(2h)) with cycle walking onto [0, n): the
// permutation that stands in for std::random_shuffle(iota(N)).
uint64_t perm_index(uint64_t seed, uint64_t tree, uint64_t n, uint64_t i) {
  unsigned bits = 2;
  while ((1ULL << bits) < n) ++bits;
  unsigned h = (bits + 1) / 2;
  uint64_t mask = (1ULL << h) - 1;
  uint64_t x = i;
  do {
    uint64_t L = x >> h, R = x & mask;
    for (uint64_t r = 0; r < 4; ++r) {
      uint64_t F = hash4(seed, ST_PERM, (tree << 8) | r, R) & mask;
      uint64_t nL = R, nR = L ^ F;
      L = nL;
      R = nR;
    }
    x = (L << h) | R;
  } while (x >= n);
  return x;
}

// ---------------------------------------------------------------------------
// Squared L2 in the accumulation order of L2DistanceAVX::compare
// (general/distance.hpp:299-323) as g++ -O3 -march=native compiles it: eight
// lane accumulators, each a chain of fused multiply-adds (the compiler fuses
// _mm256_mul_ps + _mm256_add_ps), the first tail chunk [DD, DD+8) first
// (distance.hpp:314), then elements i+j and i+8+j for i = 0, 16, ... (316-319),
// then the left-to-right sum u0+...+u7 (321).
// The reference drops the tail chunks [DD+8, D) when D % 32 is 16 or 24; here
// they are added after the first tail chunk, which leaves every other d
// bit-identical to the reference.
// ---------------------------------------------------------------------------
float l2_avx(const float* a, const float* b, size_t d) {
  size_t D = (d + 7) & ~size_t(7);
  size_t DR = D % 32, DD = D - DR;
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t c = DD; c < D; c += 8)
    for (int j = 0; j < 8; ++j) {
      float t = a[c + j] - b[c + j];
      acc[j] = fmaf(t, t, acc[j]);
    }
  for (size_t i = 0; i < DD; i += 8)
    for (int j = 0; j < 8; ++j) {
      float t = a[i + j] - b[i + j];
      acc[j] = fmaf(t, t, acc[j]);
    }
  float s = acc[0];
  for (int j = 1; j < 8; ++j) s = s + acc[j];
  return s;
}

inline uint64_t dkey(float d, uint32_t id) {
  uint32_t u;
  std::memcpy(&u, &d, 4);
  return (uint64_t(u) << 32) | id;
}
inline float key_dist(uint64_t k) {
  uint32_t u = uint32_t(k >> 32);
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// Point + InsertIntoKnn (base_index.hpp:24-62).
struct Point {
  uint32_t id;
  float dist;
  bool flag;
};
unsigned insert_into_knn(Point* addr, unsigned K, Point nn) {
  unsigned j, i = K;
  while (i > 0) {
    j = i - 1;
    if (addr[j].dist <= nn.dist) break;
    i = j;
  }
  unsigned l = i;
  while (l > 0) {
    j = l - 1;
    if (addr[j].dist < nn.dist) break;
    if (addr[j].id == nn.id) return K + 1;
    l = j;
  }
  j = K;
  while (j > i) {
    addr[j] = addr[j - 1];
    --j;
  }
  addr[i] = nn;
  return i;
}

struct Node {
  int32_t dim = -1;
  float val = 0.f;
  int32_t left = -1, right = -1, parent = -1;
  uint32_t start = 0, end = 0;
  int32_t depth = 0, tree = 0;
};

struct Params {
  int trees = 4, trees_build = 4, ml = 4, epochs = 4, check = 25, L = 30, K = 10, S = 10, tns = 10;
  uint64_t seed = 1998;
  int join_mode = 0;  // 0 = batched (the HIP library's semantics), 1 = sequential (reference order),
                      // 2 = parallel with the reference's locking (CPU baseline)
};

// Hashing index (IEH seeds, hashing_index.hpp): per table the members of
// each code in id order (BuildHashTable32/64, 323-347 / 629-655: points are
// appended in id order), the Hamming-ball masks in radius order
// (generateMask32/64, 349-627 / 657-872: radius r = the r-bit combinations of
// the code's bits in lexicographic order, i < j < k ...), the query codes.
struct HashState {
  int T = 0, bits = 0, radius = 0;
  std::vector<std::unordered_map<uint64_t, std::vector<uint32_t>>> tab;
  std::vector<uint64_t> masks;
  std::vector<uint64_t> qcodes;  // T x nq
  size_t nq = 0;
};

struct Index {
  HashState hs;
  bool hashing = false;
  size_t n = 0, d = 0, stride = 0;
  std::vector<float> data;  // n x stride, zero padded (samples/*.cc load_data)
  Params p;
  // trees
  int T = 0;
  std::vector<Node> nodes;
  std::vector<int32_t> roots;
  std::vector<uint32_t> leaf;  // T x n
  // build state
  std::vector<Point> pool;     // n x (L+1)
  std::vector<uint32_t> fill, range;
  std::vector<float> radius, radiusM;
  std::vector<uint8_t> found;
  std::vector<std::vector<uint32_t>> nn_new, nn_old;
  int iter = 0;
  bool error_flag = false;
  int max_depth = 0x0fffffff;
  // graph (CSR)
  std::vector<uint64_t> goff;
  std::vector<uint32_t> gids;
  // counters
  uint64_t n_dist_merge = 0, n_dist_init = 0, n_dist_join = 0, n_pairs_join = 0;
  std::vector<uint64_t> join_pairs_per_iter;
  // test hook: reverse lists of the last update before ([0] new, [1] old)
  // and after sampling ([2] new, [3] old), each sorted
  bool keep_rev = false;
  std::vector<std::vector<uint32_t>> rev[4];

  const float* row(size_t i) const { return &data[i * stride]; }
};

// ---------------------------------------------------------------------------
// Tree construction: kdtreeub_index.hpp:1295-1363 (= buildTrees 1192-1276),
// meanSplit/planeSplit/selectDivision 902-999, DFSbuild 1138-1167.
// Level-synchronous over all trees.  The BFS/DFS phase rule of the reference
// (BFS while 0 < |ActiveSet| < 1100 with children enqueued iff size > S,
// then DFSbuild splitting while count > TNS) is reproduced exactly; the
// partition is the stable three-way partition (same lim1/lim2 and the same
// sets on each side as planeSplit, order within a side preserved).
// ---------------------------------------------------------------------------
void split_node(Index& ix, int ni) {
  Node& nd = ix.nodes[ni];
  const size_t d = ix.d;
  uint32_t* ids = &ix.leaf[size_t(nd.tree) * ix.n + nd.start];
  unsigned count = nd.end - nd.start;
  unsigned cnt = std::min(101u, count);  // SAMPLE_NUM + 1 (kdtreeub_index.hpp:912)
  std::vector<float> mean(d, 0.f), var(d, 0.f);
  for (unsigned j = 0; j < cnt; ++j) {
    const float* v = ix.row(ids[j]);
    for (size_t k = 0; k < d; ++k) mean[k] = mean[k] + v[k];
  }
  float div = 1.0f / float(cnt);
  for (size_t k = 0; k < d; ++k) mean[k] = mean[k] * div;
  for (unsigned j = 0; j < cnt; ++j) {
    const float* v = ix.row(ids[j]);
    for (size_t k = 0; k < d; ++k) {
      float t = v[k] - mean[k];
      var[k] = fmaf(t, t, var[k]);
    }
  }
  // selectDivision (kdtreeub_index.hpp:974-999), RAND_DIM = 5
  int num = 0;
  size_t top[5];
  for (size_t i = 0; i < d; ++i) {
    if (num < 5 || var[i] > var[top[num - 1]]) {
      if (num < 5) top[num++] = i;
      else top[num - 1] = i;
      int j = num - 1;
      while (j > 0 && var[top[j]] > var[top[j - 1]]) {
        std::swap(top[j], top[j - 1]);
        --j;
      }
    }
  }
  uint64_t rnd = hash4(ix.p.seed, ST_SPLIT, uint64_t(nd.tree), (uint64_t(nd.start) << 32) | nd.end) % uint64_t(num);
  unsigned cutdim = unsigned(top[rnd]);
  float cutval = mean[cutdim];
  // stable three-way partition (planeSplit, kdtreeub_index.hpp:954-973)
  std::vector<uint32_t> lt, eq, gt;
  for (unsigned j = 0; j < count; ++j) {
    float v = ix.row(ids[j])[cutdim];
    if (v < cutval) lt.push_back(ids[j]);
    else if (v <= cutval) eq.push_back(ids[j]);
    else gt.push_back(ids[j]);
  }
  unsigned lim1 = unsigned(lt.size()), lim2 = unsigned(lt.size() + eq.size());
  unsigned pos = 0;
  for (uint32_t x : lt) ids[pos++] = x;
  for (uint32_t x : eq) ids[pos++] = x;
  for (uint32_t x : gt) ids[pos++] = x;
  // balancing rule (kdtreeub_index.hpp:943-950)
  unsigned index;
  if (lim1 > count / 2) index = lim1;
  else if (lim2 < count / 2) index = lim2;
  else index = count / 2;
  if (lim1 == count || lim2 == 0) index = count / 2;
  nd.dim = int32_t(cutdim);
  nd.val = cutval;
  Node& L = ix.nodes[nd.left];
  Node& R = ix.nodes[nd.right];
  L.tree = R.tree = nd.tree;
  L.parent = R.parent = ni;
  L.depth = R.depth = nd.depth + 1;
  L.start = nd.start;
  L.end = nd.start + index;
  R.start = L.end;
  R.end = nd.end;
}

int build_trees(Index& ix) {
  const Params& p = ix.p;
  if (p.trees <= 0) { g_err = "tree number must be positive"; return -2; }
  ix.T = p.trees;
  const size_t n = ix.n;
  ix.leaf.assign(size_t(ix.T) * n, 0);
#pragma omp parallel for schedule(static)
  for (long long t = 0; t < ix.T; ++t)
    for (size_t i = 0; i < n; ++i) ix.leaf[size_t(t) * n + i] = uint32_t(perm_index(p.seed, uint64_t(t), n, i));
  ix.nodes.assign(ix.T, Node());
  ix.roots.resize(ix.T);
  std::vector<int32_t> active;
  for (int t = 0; t < ix.T; ++t) {
    ix.nodes[t].tree = t;
    ix.nodes[t].start = 0;
    ix.nodes[t].end = uint32_t(n);
    ix.roots[t] = t;
    active.push_back(t);
  }
  bool bfs = true;
  const unsigned S = unsigned(p.S), TNS = unsigned(p.tns);
  while (!active.empty()) {
    if (bfs && !(active.size() < 1100)) {
      bfs = false;
      std::vector<int32_t> keep;
      for (int32_t a : active)
        if (ix.nodes[a].end - ix.nodes[a].start > TNS) keep.push_back(a);
      active.swap(keep);
      if (active.empty()) break;
    }
    size_t base = ix.nodes.size();
    ix.nodes.resize(base + 2 * active.size());
    for (size_t i = 0; i < active.size(); ++i) {
      ix.nodes[active[i]].left = int32_t(base + 2 * i);
      ix.nodes[active[i]].right = int32_t(base + 2 * i + 1);
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (long long i = 0; i < (long long)active.size(); ++i) split_node(ix, active[i]);
    std::vector<int32_t> next;
    unsigned thr = bfs ? S : TNS;
    for (size_t i = 0; i < active.size(); ++i)
      for (int c = 0; c < 2; ++c) {
        int32_t ci = int32_t(base + 2 * i + c);
        if (ix.nodes[ci].end - ix.nodes[ci].start > thr) next.push_back(ci);
      }
    active.swap(next);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Conquer-to-depth merge: getMergeLevelNodeList (kdtreeub_index.hpp:1000-1011),
// SearchToLeaf (1012-1021), mergeSubGraphs (1022-1081), in the order a single
// thread runs them (1367-1374): trees 0..TB-1, depth-ml nodes left to right,
// post-order inside each, probers in range order, leaf members in range order.
// Heap: std::set<Candidate, greater> bounded to S, rejected ids appended to
// nn_new while |nn_new| < 2S.
// ---------------------------------------------------------------------------
struct Heap {
  std::vector<std::pair<float, uint32_t>> v;  // ascending (dist, id)
};
inline void heap_offer(Heap& h, std::vector<uint32_t>& extra, unsigned S, uint32_t id, float dist) {
  if (h.v.size() < S || dist < h.v.back().first) {
    std::pair<float, uint32_t> c(dist, id);
    auto it = std::lower_bound(h.v.begin(), h.v.end(), c);
    if (it != h.v.end() && *it == c) return;  // set semantics
    h.v.insert(it, c);
    if (h.v.size() > S) h.v.pop_back();
  } else if (extra.size() < 2 * S) {
    extra.push_back(id);
  }
}

void merge_level_list(Index& ix, int32_t node, int depth, std::vector<int32_t>& out) {
  const Node& nd = ix.nodes[node];
  if (nd.left >= 0 && nd.right >= 0 && depth < ix.p.ml) {
    merge_level_list(ix, nd.left, depth + 1, out);
    merge_level_list(ix, nd.right, depth + 1, out);
  } else if (depth == ix.p.ml) {
    out.push_back(node);
  } else {
    ix.error_flag = true;
    if (depth < ix.max_depth) ix.max_depth = depth;
  }
}

int32_t search_to_leaf(const Index& ix, int32_t node, const float* q) {
  while (ix.nodes[node].left >= 0 && ix.nodes[node].right >= 0) {
    const Node& nd = ix.nodes[node];
    node = (q[nd.dim] < nd.val) ? nd.left : nd.right;
  }
  return node;
}

void merge_subgraphs(Index& ix, int t, int32_t node, std::vector<Heap>& heaps, uint64_t& nd_count) {
  const Node& nd = ix.nodes[node];
  if (!(nd.left >= 0 && nd.right >= 0)) return;
  merge_subgraphs(ix, t, nd.left, heaps, nd_count);
  merge_subgraphs(ix, t, nd.right, heaps, nd_count);
  const Node& L = ix.nodes[nd.left];
  const Node& R = ix.nodes[nd.right];
  size_t numL = L.end - L.start, numR = R.end - R.start;
  int32_t root;
  size_t s, e;
  if (numL < numR) { root = nd.right; s = L.start; e = L.end; }
  else { root = nd.left; s = R.start; e = R.end; }
  const uint32_t* ll = &ix.leaf[size_t(t) * ix.n];
  const unsigned S = unsigned(ix.p.S);
  for (; s < e; ++s) {
    uint32_t p = ll[s];
    int32_t lf = search_to_leaf(ix, root, ix.row(p));
    for (uint32_t i = ix.nodes[lf].start; i < ix.nodes[lf].end; ++i) {
      uint32_t q = ll[i];
      float dist = l2_avx(ix.row(q), ix.row(p), ix.d);
      ++nd_count;
      heap_offer(heaps[q], ix.nn_new[q], S, p, dist);
      heap_offer(heaps[p], ix.nn_new[p], S, q, dist);
    }
  }
}

// GenRandom (kdtreeub_index.hpp:1121-1135) with the counter-based generator.
void gen_random(uint64_t seed, uint32_t node, uint32_t* addr, unsigned size, unsigned N) {
  for (unsigned i = 0; i < size; ++i) addr[i] = uint32_t(hash4(seed, ST_INIT_NEW, node, i) % (N - size));
  std::sort(addr, addr + size);
  for (unsigned i = 1; i < size; ++i)
    if (addr[i] <= addr[i - 1]) addr[i] = addr[i - 1] + 1;
  unsigned off = unsigned(hash4(seed, ST_INIT_NEW, node, size) % N);
  for (unsigned i = 0; i < size; ++i) addr[i] = (addr[i] + off) % N;
}

// Merge + random init: kdtreeub_index.hpp:1367-1428.
int merge_and_init(Index& ix) {
  const Params& p = ix.p;
  const size_t n = ix.n;
  const unsigned S = unsigned(p.S), Lp = unsigned(p.L);
  if (p.trees_build > ix.T || p.trees_build <= 0) { g_err = "invalid number of trees for the graph"; return -2; }
  if (n <= 2 * S + 1) { g_err = "too few points for the requested S"; return -2; }
  ix.nn_new.assign(n, {});
  ix.nn_old.assign(n, {});
  std::vector<Heap> heaps(n);
  ix.error_flag = false;
  ix.max_depth = 0x0fffffff;
  // Depth-ml subtrees of one tree hold disjoint points, so running them in
  // parallel inside a tree (trees in order) gives every point the same
  // arrival sequence as the single-thread order.
  uint64_t cnt = 0;
  for (int t = 0; t < p.trees_build; ++t) {
    std::vector<int32_t> lst;
    merge_level_list(ix, ix.roots[t], 0, lst);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : cnt)
    for (long long i = 0; i < (long long)lst.size(); ++i) {
      uint64_t c = 0;
      merge_subgraphs(ix, t, lst[i], heaps, c);
      cnt += c;
    }
  }
  ix.n_dist_merge = cnt;
  // random init (kdtreeub_index.hpp:1377-1423)
  ix.pool.assign(n * (Lp + 1), Point{0, 0.f, false});
  ix.fill.assign(n, S);
  ix.range.assign(n, S);
  ix.radius.assign(n, FLT_MAX);
  ix.radiusM.assign(n, 0.f);
  ix.found.assign(n, 0);
  uint64_t icnt = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : icnt)
  for (long long ni = 0; ni < (long long)n; ++ni) {
    uint32_t nn = uint32_t(ni);
    std::vector<uint32_t>& nw = ix.nn_new[nn];
    if (nw.size() < 2 * S) {
      nw.assign(2 * S, 0);
      gen_random(p.seed, nn, nw.data(), 2 * S, unsigned(n));
    }
    Heap& h = heaps[nn];
    uint64_t r = 0;
    while (h.v.size() < S) {
      uint32_t rid = uint32_t(hash4(p.seed, ST_INIT_TOP, nn, r++) % n);
      float dist = l2_avx(ix.row(nn), ix.row(rid), ix.d);
      ++icnt;
      std::pair<float, uint32_t> c(dist, rid);
      auto it = std::lower_bound(h.v.begin(), h.v.end(), c);
      if (it != h.v.end() && *it == c) continue;
      h.v.insert(it, c);
    }
    Point* pl = &ix.pool[size_t(nn) * (Lp + 1)];
    for (unsigned l = 0; l < S; ++l) {
      pl[l] = Point{h.v[l].second, h.v[l].first, true};
      nw[l] = h.v[l].second;
    }
  }
  ix.n_dist_init = icnt;
  ix.iter = 0;
  if (ix.error_flag) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "merge level deeper than tree, max merge deepth is %d", ix.max_depth - 1);
    g_err = buf;
    return -3;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// NN-descent join (base_index.hpp:328-371, Neighbor::join 95-107,
// Neighbor::insert 80-93), batched form.  All candidates of one iteration are
// filtered against the pools as they stood when the iteration began (a
// candidate worse than a full pool's last entry cannot enter it), then each
// pool becomes the best L of (pool + candidates) by (distance, id), duplicates
// removed, existing entries keeping their flag.  This is the pool the
// sequential inserts reach, up to the order of equal distances.
// `found` (base_index.hpp:364-368): node n is found when a candidate from its
// join, inserted at the first endpoint of a pair, holds a rank < Check_K in
// the merged pool; among equal candidates the lowest joining node gets it.
// ---------------------------------------------------------------------------
struct Striped {
  std::vector<std::mutex> m;
  Striped() : m(4096) {}
  std::mutex& at(uint32_t x) { return m[x & 4095]; }
};

void join_batched(Index& ix) {
  const size_t n = ix.n;
  const unsigned Lp = unsigned(ix.p.L), CK = unsigned(ix.p.check);
  // bucket entry: (key, joining node << 1 | first endpoint)
  std::vector<std::vector<std::pair<uint64_t, uint64_t>>> bucket(n);
  std::vector<uint8_t> credit(n, 0);
  Striped locks;
  uint64_t pairs = 0;
  auto worst = [&](uint32_t x) -> uint64_t {
    if (ix.fill[x] < Lp) return ~0ULL;
    const Point& w = ix.pool[size_t(x) * (Lp + 1) + Lp - 1];
    return dkey(w.dist, w.id);
  };
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : pairs)
  for (long long ni = 0; ni < (long long)n; ++ni) {
    uint32_t nn = uint32_t(ni);
    auto pair = [&](uint32_t a, uint32_t b) {
      float dist = l2_avx(ix.row(a), ix.row(b), ix.d);
      ++pairs;
      uint64_t ka = dkey(dist, b), kb = dkey(dist, a);
      uint64_t src = uint64_t(nn) << 1;
      if (ka < worst(a)) { std::lock_guard<std::mutex> g(locks.at(a)); bucket[a].push_back({ka, src | 1}); }
      if (kb < worst(b)) { std::lock_guard<std::mutex> g(locks.at(b)); bucket[b].push_back({kb, src}); }
    };
    const auto& nw = ix.nn_new[nn];
    const auto& od = ix.nn_old[nn];
    for (uint32_t i : nw) {
      for (uint32_t j : nw)
        if (i < j) pair(i, j);
      for (uint32_t j : od) pair(i, j);
    }
  }
  ix.n_pairs_join += pairs;
  ix.n_dist_join += pairs;
  ix.join_pairs_per_iter.push_back(pairs);
#pragma omp parallel for schedule(dynamic, 256)
  for (long long xi = 0; xi < (long long)n; ++xi) {
    uint32_t x = uint32_t(xi);
    auto& bk = bucket[x];
    if (bk.empty()) continue;
    Point* pl = &ix.pool[size_t(x) * (Lp + 1)];
    unsigned Lx = ix.fill[x];
    // (key, src, flag): src 0 = existing entry (wins duplicates), 1 = candidate
    std::vector<std::pair<uint64_t, unsigned>> all;
    all.reserve(Lx + bk.size());
    for (unsigned l = 0; l < Lx; ++l) all.push_back({dkey(pl[l].dist, pl[l].id), pl[l].flag ? 1u : 0u});
    for (auto& k : bk) all.push_back({k.first, 2u | 1u});
    std::sort(all.begin(), all.end());
    std::sort(bk.begin(), bk.end());
    unsigned out = 0;
    uint64_t prev = ~0ULL;
    for (auto& e : all) {
      if (e.first == prev) continue;
      prev = e.first;
      if (out == Lp) break;
      if (e.second == 3u && out < CK) {
        auto it = std::lower_bound(bk.begin(), bk.end(), std::make_pair(e.first, uint64_t(0)));
        uint64_t best = ~0ULL;
        for (; it != bk.end() && it->first == e.first; ++it)
          if ((it->second & 1) && (it->second >> 1) < best) best = it->second >> 1;
        if (best != ~0ULL) { std::lock_guard<std::mutex> g(locks.at(uint32_t(best))); credit[best] = 1; }
      }
      pl[out++] = Point{uint32_t(e.first & 0xffffffffu), key_dist(e.first), (e.second & 1u) != 0};
    }
    ix.fill[x] = out;
    std::vector<std::pair<uint64_t, uint64_t>>().swap(bk);
  }
  for (size_t i = 0; i < n; ++i) ix.found[i] = credit[i];
}

// Sequential form in the reference's own order (single thread), used to pin
// the batched form's statistics to the reference semantics.
void join_sequential(Index& ix) {
  const size_t n = ix.n;
  const unsigned Lp = unsigned(ix.p.L), CK = unsigned(ix.p.check);
  uint64_t pairs = 0;
  auto insert = [&](uint32_t x, uint32_t id, float dist) -> unsigned {
    if (dist > ix.radius[x]) return Lp + 1;
    Point* pl = &ix.pool[size_t(x) * (Lp + 1)];
    unsigned& L = ix.fill[x];
    unsigned l = insert_into_knn(pl, L, Point{id, dist, true});
    if (l <= L) {
      if (L + 1 < Lp + 1) ++L;
      else ix.radius[x] = pl[L - 1].dist;
    }
    return l;
  };
  for (size_t ni = 0; ni < n; ++ni) {
    unsigned uu = 0;
    ix.found[ni] = 0;
    auto cb = [&](uint32_t i, uint32_t j) {
      float dist = l2_avx(ix.row(i), ix.row(j), ix.d);
      ++pairs;
      unsigned r = insert(i, j, dist);
      if (r < CK) ++uu;
      insert(j, i, dist);
      if (r < CK) ++uu;
    };
    const auto nw = ix.nn_new[ni];
    const auto od = ix.nn_old[ni];
    for (uint32_t i : nw) {
      for (uint32_t j : nw)
        if (i < j) cb(i, j);
      for (uint32_t j : od) cb(i, j);
    }
    ix.found[ni] = uu > 0;
  }
  ix.n_pairs_join += pairs;
  ix.n_dist_join += pairs;
  ix.join_pairs_per_iter.push_back(pairs);
}

// Parallel form with the reference's threading (join_mode 2): OpenMP over
// the joining nodes (dynamic schedule of 100, base_index.hpp:331), each
// insert an unlocked radius check and then InsertIntoKnn under the target's
// lock (Neighbor::insert 80-93).  Like the reference it is not deterministic
// under several threads; it is the CPU baseline's form.
void join_parallel(Index& ix) {
  const size_t n = ix.n;
  const unsigned Lp = unsigned(ix.p.L), CK = unsigned(ix.p.check);
  uint64_t pairs = 0;
  Striped locks;
  auto insert = [&](uint32_t x, uint32_t id, float dist) -> unsigned {
    float r;
    __atomic_load(&ix.radius[x], &r, __ATOMIC_RELAXED);
    if (dist > r) return Lp + 1;
    std::lock_guard<std::mutex> g(locks.at(x));
    Point* pl = &ix.pool[size_t(x) * (Lp + 1)];
    unsigned& L = ix.fill[x];
    unsigned l = insert_into_knn(pl, L, Point{id, dist, true});
    if (l <= L) {
      if (L + 1 < Lp + 1) ++L;
      else __atomic_store(&ix.radius[x], &pl[L - 1].dist, __ATOMIC_RELAXED);
    }
    return l;
  };
#pragma omp parallel for schedule(dynamic, 100) reduction(+ : pairs)
  for (long long ni = 0; ni < (long long)n; ++ni) {
    unsigned uu = 0;
    auto cb = [&](uint32_t i, uint32_t j) {
      float dist = l2_avx(ix.row(i), ix.row(j), ix.d);
      ++pairs;
      unsigned r = insert(i, j, dist);
      if (r < CK) ++uu;
      insert(j, i, dist);
    };
    const auto& nw = ix.nn_new[ni];
    const auto& od = ix.nn_old[ni];
    for (uint32_t i : nw) {
      for (uint32_t j : nw)
        if (i < j) cb(i, j);
      for (uint32_t j : od) cb(i, j);
    }
    ix.found[ni] = uu > 0;
  }
  ix.n_pairs_join += pairs;
  ix.n_dist_join += pairs;
  ix.join_pairs_per_iter.push_back(pairs);
}

// update(paramL = 100): base_index.hpp:373-437.  Reverse lists longer than
// 100 keep the 100 members with the smallest generator value (in place of
// random_shuffle + truncate); list order is immaterial to the join.
void update(Index& ix, unsigned paramL) {
  const size_t n = ix.n;
  const unsigned Lp = unsigned(ix.p.L), S = unsigned(ix.p.S);
  for (size_t i = 0; i < n; ++i) {
    ix.nn_new[i].clear();
    ix.nn_old[i].clear();
    ix.radius[i] = ix.pool[i * (Lp + 1) + Lp].dist;  // pool.back() (base_index.hpp:379)
  }
#pragma omp parallel for schedule(static)
  for (long long ii = 0; ii < (long long)n; ++ii) {
    size_t i = size_t(ii);
    const Point* pl = &ix.pool[i * (Lp + 1)];
    if (ix.found[i]) {
      unsigned maxl = std::min(ix.range[i] + S, ix.fill[i]);
      unsigned c = 0, l = 0;
      while (l < maxl && c < S) {
        if (pl[l].flag) ++c;
        ++l;
      }
      ix.range[i] = l;
    }
    ix.radiusM[i] = pl[ix.range[i] - 1].dist;
  }
  std::vector<std::vector<uint32_t>> rnew(n), rold(n);
  Striped locks;
#pragma omp parallel for schedule(static)
  for (long long ii = 0; ii < (long long)n; ++ii) {
    uint32_t nn = uint32_t(ii);
    Point* pl = &ix.pool[size_t(nn) * (Lp + 1)];
    for (unsigned l = 0; l < ix.range[nn]; ++l) {
      Point& e = pl[l];
      uint32_t o = e.id;
      if (e.flag) {
        ix.nn_new[nn].push_back(o);
        if (e.dist > ix.radiusM[o]) { std::lock_guard<std::mutex> g(locks.at(o)); rnew[o].push_back(nn); }
        e.flag = false;
      } else {
        ix.nn_old[nn].push_back(o);
        if (e.dist > ix.radiusM[o]) { std::lock_guard<std::mutex> g(locks.at(o)); rold[o].push_back(nn); }
      }
    }
  }
  const uint64_t seed = ix.p.seed, it = uint64_t(ix.iter);
  auto sample = [&](std::vector<uint32_t>& v, uint64_t stream, uint32_t o) {
    if (!paramL || v.size() <= paramL) return;
    std::vector<std::pair<uint64_t, uint32_t>> k;
    k.reserve(v.size());
    for (uint32_t x : v) k.push_back({hash4(seed, stream, (it << 32) | o, x), x});
    std::nth_element(k.begin(), k.begin() + paramL, k.end());
    v.resize(paramL);
    for (unsigned i = 0; i < paramL; ++i) v[i] = k[i].second;
  };
  if (ix.keep_rev) {
    ix.rev[0] = rnew;
    ix.rev[1] = rold;
    for (int w = 0; w < 2; ++w)
      for (auto& v : ix.rev[w]) std::sort(v.begin(), v.end());
  }
#pragma omp parallel for schedule(dynamic, 256)
  for (long long ii = 0; ii < (long long)n; ++ii) {
    uint32_t o = uint32_t(ii);
    sample(rnew[o], ST_RNN_NEW, o);
    std::sort(rnew[o].begin(), rnew[o].end());
    ix.nn_new[o].insert(ix.nn_new[o].end(), rnew[o].begin(), rnew[o].end());
    sample(rold[o], ST_RNN_OLD, o);
    std::sort(rold[o].begin(), rold[o].end());
    ix.nn_old[o].insert(ix.nn_old[o].end(), rold[o].begin(), rold[o].end());
  }
  if (ix.keep_rev) {
    ix.rev[2] = rnew;
    ix.rev[3] = rold;
  }
}

// Graph output: first K pool ids (base_index.hpp:469-479).
void finalize_graph(Index& ix) {
  const size_t n = ix.n;
  const unsigned Lp = unsigned(ix.p.L), K = unsigned(ix.p.K);
  ix.goff.resize(n + 1);
  ix.gids.assign(n * K, 0);
  for (size_t i = 0; i <= n; ++i) ix.goff[i] = i * K;
  for (size_t i = 0; i < n; ++i) {
    const Point* pl = &ix.pool[i * (Lp + 1)];
    unsigned f = ix.fill[i];
    for (unsigned j = 0; j < K; ++j) ix.gids[i * K + j] = pl[std::min(j, f - 1)].id;
  }
}

// ---------------------------------------------------------------------------
// Search, method 0: getSearchNodeList (kdtreeub_index.hpp:332-346) and
// getNeighbors_nnexp (363-494).  Seeds beyond the leaves come from the
// counter-based generator instead of rand() (411-417).
// ---------------------------------------------------------------------------
void search_node_list(const Index& ix, int32_t node, const float* q, unsigned lsize, std::vector<int32_t>& vn) {
  if (vn.size() >= lsize) return;
  const Node& nd = ix.nodes[node];
  if (nd.left >= 0 && nd.right >= 0) {
    if (q[nd.dim] < nd.val) {
      search_node_list(ix, nd.left, q, lsize, vn);
      search_node_list(ix, nd.right, q, lsize, vn);
    } else {
      search_node_list(ix, nd.right, q, lsize, vn);
      search_node_list(ix, nd.left, q, lsize, vn);
    }
  } else {
    vn.push_back(node);
  }
}

struct SearchArgs {
  int epochs, init_num, extend_to, trees, K;
};

// Seed pool common to both methods (kdtreeub_index.hpp:387-417 / 521-550).
// On a hashing index the seeds come from the buckets of the query's code
// XOR each mask, mask outer, table inner (hashing_index.hpp:1108-1130).  The
// random top-up runs to topup_to ids (init_num; K for the hash-only search,
// 1005-1014).
void seed_pool(const Index& ix, const float* q, unsigned lsize, int init_num, int trees, uint64_t qid,
               std::vector<uint8_t>& visited, std::vector<uint32_t>& pool, int topup_to = -1) {
  unsigned p = 0;
  const unsigned want = unsigned(init_num);
  pool.assign(want, 0);
  if (ix.hashing) {
    const HashState& hs = ix.hs;
    for (size_t j = 0; j < hs.masks.size() && p < want; ++j) {
      for (int h = 0; h < hs.T && p < want; ++h) {
        const uint64_t code = hs.qcodes[size_t(h) * hs.nq + qid] ^ hs.masks[j];
        auto it = hs.tab[h].find(code);
        if (it == hs.tab[h].end()) continue;
        for (size_t k = 0; k < it->second.size() && p < want; ++k) {
          const uint32_t nn = it->second[k];
          if (visited[nn]) continue;
          visited[nn] = 1;
          pool[p++] = nn;
        }
      }
    }
  } else {
    std::vector<std::vector<int32_t>> cand(trees);
    for (int i = 0; i < trees; ++i) search_node_list(ix, ix.roots[i], q, lsize, cand[i]);
    for (unsigned ni = 0; ni < lsize && p < want; ++ni) {
      for (int i = 0; i < trees && p < want; ++i) {
        if (ni >= cand[i].size()) continue;
        const Node& lf = ix.nodes[cand[i][ni]];
        const uint32_t* ll = &ix.leaf[size_t(i) * ix.n];
        for (uint32_t j = lf.start; j < lf.end && p < want; ++j) {
          uint32_t nn = ll[j];
          if (visited[nn]) continue;
          visited[nn] = 1;
          pool[p++] = nn;
        }
      }
    }
  }
  if (topup_to >= 0 && p < unsigned(topup_to)) {
    // hash-only search: top up to K and keep the p candidates
    uint64_t r = 0;
    while (p < unsigned(topup_to)) {
      uint32_t nn = uint32_t(hash4(ix.p.seed, ST_SEARCH, qid, r++) % ix.n);
      if (visited[nn]) continue;
      visited[nn] = 1;
      pool[p++] = nn;
    }
  }
  if (topup_to >= 0) {
    pool.resize(p);
    return;
  }
  uint64_t r = 0;
  while (p < want) {
    uint32_t nn = uint32_t(hash4(ix.p.seed, ST_SEARCH, qid, r++) % ix.n);
    if (visited[nn]) continue;
    visited[nn] = 1;
    pool[p++] = nn;
  }
}

void search_nnexp(const Index& ix, const float* q, uint64_t qid, const SearchArgs& a, int32_t* out_ids,
                  float* out_d, uint64_t& ndist) {
  const size_t n = ix.n;
  std::vector<uint8_t> visited(n, 0), expanded(n, 0);
  unsigned lsize = unsigned(a.init_num * 2 / (5 * a.trees) + 1);
  std::vector<uint32_t> pool;
  seed_pool(ix, q, lsize, a.init_num, a.trees, qid, visited, pool);
  size_t R = size_t(std::max(a.extend_to, a.K));
  std::vector<std::pair<float, uint32_t>> res;
  res.reserve(pool.size() + 4096);
  for (uint32_t id : pool) res.push_back({l2_avx(q, ix.row(id), ix.d), id});
  ndist += pool.size();
  std::partial_sort(res.begin(), res.begin() + R, res.end());
  res.resize(R);
  std::vector<uint32_t> ids;
  for (int it = 0; it < a.epochs; ++it) {
    ids.clear();
    for (int j = 0; j < a.extend_to; ++j) {
      uint32_t pj = res[j].second;
      if (expanded[pj]) continue;
      expanded[pj] = 1;
      for (uint64_t g = ix.goff[pj]; g < ix.goff[pj + 1]; ++g) {
        uint32_t id = ix.gids[g];
        if (visited[id]) continue;
        visited[id] = 1;
        ids.push_back(id);
      }
    }
    for (uint32_t id : ids) res.push_back({l2_avx(q, ix.row(id), ix.d), id});
    ndist += ids.size();
    std::partial_sort(res.begin(), res.begin() + R, res.end());
    res.resize(R);
  }
  for (int i = 0; i < a.K; ++i) {
    out_ids[i] = int32_t(res[i].second);
    if (out_d) out_d[i] = res[i].first;
  }
}

// ---------------------------------------------------------------------------
// Learned stopping (a form beside the plain search; no reference code).  A
// probe search (method 0 with a small pool) is checkpointed after the seeds
// and after every epoch that expands something.  At a checkpoint the
// pool-state features below are computed; in trace mode they are recorded
// with a label (top-k ids also in the plain search's top k), in learned mode a
// network 16 -> 64 -> 64 -> 1 (ReLU, BF16 parameters, FP32 fused
// multiply-adds in input order) decides: stop when its logit reaches zthr,
// otherwise the query runs the full plain search from scratch.
// ---------------------------------------------------------------------------
constexpr int kNF = 16, kNH = 64;
constexpr int kW1 = 0, kB1 = kW1 + kNF * kNH, kW2 = kB1 + kNH, kB2 = kW2 + kNH * kNH, kW3 = kB2 + kNH,
              kB3 = kW3 + kNH, kNP = kB3 + 1;

inline float round_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u = ((u + 0x7fffu + ((u >> 16) & 1u)) >> 16) << 16;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

// Features of a checkpoint: e; best and k-th distance; k-th/best ratios to the
// best, the extend_to-th and the last pool entry; relative gains of the k-th
// and best distances over the last 1..3 checkpoints; ids measured; ids
// expanded in the last epoch; top-k ids new since the last checkpoint;
// unexpanded ids among the first extend_to.
void checkpoint_features(int e, float d1, float dk, float dE, float dR, const float* dkh, const float* d1h,
                         float visited, float expanded, float changed, float frontier, float* f) {
  auto ratio = [](float a, float b) { return b > 0.f ? a / b : 1.f; };
  auto gain = [](float before, float now) { return before > 0.f ? (before - now) / before : 0.f; };
  f[0] = float(e);
  f[1] = d1;
  f[2] = dk;
  f[3] = ratio(d1, dk);
  f[4] = ratio(dk, dE);
  f[5] = ratio(dk, dR);
  for (int j = 0; j < 3; ++j) {
    f[6 + j] = e > j ? gain(dkh[j], dk) : 0.f;
    f[9 + j] = e > j ? gain(d1h[j], d1) : 0.f;
  }
  f[12] = visited;
  f[13] = expanded;
  f[14] = changed;
  f[15] = frontier;
}

// Logit on BF16-rounded parameters (w holds the rounded values).
float net_logit(const float* w, const float* norm, const float* f) {
  float x[kNF], h1[kNH];
  for (int i = 0; i < kNF; ++i) x[i] = (f[i] - norm[i]) * norm[kNF + i];
  for (int j = 0; j < kNH; ++j) {
    float a = w[kB1 + j];
    for (int i = 0; i < kNF; ++i) a = std::fmaf(w[kW1 + i * kNH + j], x[i], a);
    h1[j] = a > 0.f ? a : 0.f;
  }
  float z = w[kB3];
  for (int j = 0; j < kNH; ++j) {
    float a = w[kB2 + j];
    for (int i = 0; i < kNH; ++i) a = std::fmaf(w[kW2 + i * kNH + j], h1[i], a);
    z = std::fmaf(w[kW3 + j], a > 0.f ? a : 0.f, z);
  }
  return z;
}

struct ProbeArgs {
  SearchArgs full, probe;
  int mode;             // 1 learned (restart), 2 trace, 3 learned (continuation)
  const float* w;       // mode 1: rounded parameters
  const float* norm;    // mode 1
  float zthr;           // mode 1
  const int32_t* plain; // mode 2: this query's plain top k
};

// The leaf sequence of a query: the ids of the first lsize leaves of every
// tree, leaf rank outer, tree inner (the order seed_pool walks, without
// dedup or bound).
void leaf_sequence(const Index& ix, const float* q, unsigned lsize, int trees, std::vector<uint32_t>& seq) {
  std::vector<std::vector<int32_t>> cand(trees);
  for (int i = 0; i < trees; ++i) search_node_list(ix, ix.roots[i], q, lsize, cand[i]);
  seq.clear();
  for (unsigned ni = 0; ni < lsize; ++ni)
    for (int i = 0; i < trees; ++i) {
      if (ni >= cand[i].size()) continue;
      const Node& lf = ix.nodes[cand[i][ni]];
      const uint32_t* ll = &ix.leaf[size_t(i) * ix.n];
      for (uint32_t j = lf.start; j < lf.end; ++j) seq.push_back(ll[j]);
    }
}

// The next `want` ids not yet visited, from position `from` of the sequence,
// then from the generator's draws 0, 1, 2, ... (the search top-up); marks
// them visited.  *resume: the position after the last id taken from the
// sequence (its end when the sequence ran out).
void take_seeds(const Index& ix, const std::vector<uint32_t>& seq, size_t from, int want, uint64_t qid,
                std::vector<uint8_t>& visited, std::vector<uint32_t>& out, size_t* resume) {
  out.clear();
  size_t pos = from;
  for (; pos < seq.size() && int(out.size()) < want; ++pos) {
    const uint32_t nn = seq[pos];
    if (visited[nn]) continue;
    visited[nn] = 1;
    out.push_back(nn);
  }
  if (resume) *resume = pos;
  uint64_t r = 0;
  while (int(out.size()) < want) {
    const uint32_t nn = uint32_t(hash4(ix.p.seed, ST_SEARCH, qid, r++) % ix.n);
    if (visited[nn]) continue;
    visited[nn] = 1;
    out.push_back(nn);
  }
}

// Returns whether the network stopped the probe (modes 1 and 3).  feat/label
// (mode 2) hold probe.epochs + 1 checkpoints.
//
// Mode 3 (continuation): the probe takes its seeds from the full search's
// leaf sequence.  A query the network does not stop continues in place: the
// next min(full.init_num - probe.init_num, n - visited) ids not yet visited,
// from where the probe's scan stopped (then the top-up draws from 0), are
// measured and taken with the probe's pool into the best max(full.extend_to,
// K); the full search's epochs follow with the same visited and expanded
// flags.  When those ids and the probe's pool are fewer than that pool, the
// query runs the full search from scratch instead.
bool search_probe(const Index& ix, const float* q, uint64_t qid, const ProbeArgs& pa, int32_t* out_ids,
                  float* out_d, float* feat, uint8_t* label) {
  const SearchArgs& a = pa.probe;
  const size_t n = ix.n;
  const int K = a.K;
  std::vector<uint8_t> visited(n, 0), expanded(n, 0);
  std::vector<uint32_t> pool, seq;
  size_t resume = 0;
  if (pa.mode == 3) {
    const SearchArgs& f = pa.full;
    leaf_sequence(ix, q, unsigned(f.init_num * 2 / (5 * f.trees) + 1), f.trees, seq);
    take_seeds(ix, seq, 0, a.init_num, qid, visited, pool, &resume);
  } else {
    unsigned lsize = unsigned(a.init_num * 2 / (5 * a.trees) + 1);
    seed_pool(ix, q, lsize, a.init_num, a.trees, qid, visited, pool);
  }
  size_t R = size_t(std::max(a.extend_to, a.K));
  std::vector<std::pair<float, uint32_t>> res;
  for (uint32_t id : pool) res.push_back({l2_avx(q, ix.row(id), ix.d), id});
  std::partial_sort(res.begin(), res.begin() + R, res.end());
  res.resize(R);
  size_t nvis = pool.size();
  std::vector<int32_t> prevk(K, -1);
  float dkh[3] = {0, 0, 0}, d1h[3] = {0, 0, 0};
  auto ckpt = [&](int e, int nexp_last) -> bool {
    const size_t ER = std::min(size_t(a.extend_to), R);
    int fr = 0;
    for (size_t j = 0; j < ER; ++j) fr += expanded[res[j].second] ? 0 : 1;
    int ch = K;
    if (e > 0) {
      ch = 0;
      for (int j = 0; j < K; ++j) ch += std::find(prevk.begin(), prevk.end(), int32_t(res[j].second)) == prevk.end();
    }
    for (int j = 0; j < K; ++j) prevk[j] = int32_t(res[j].second);
    float f[kNF];
    checkpoint_features(e, res[0].first, res[K - 1].first, res[(ER > 0 ? ER : 1) - 1].first, res[R - 1].first,
                        dkh, d1h, float(nvis), float(nexp_last), float(ch), float(fr), f);
    dkh[2] = dkh[1]; dkh[1] = dkh[0]; dkh[0] = res[K - 1].first;
    d1h[2] = d1h[1]; d1h[1] = d1h[0]; d1h[0] = res[0].first;
    if (pa.mode == 2) {
      int match = 0;
      for (int j = 0; j < K; ++j) match += std::find(pa.plain, pa.plain + K, int32_t(res[j].second)) != pa.plain + K;
      std::memcpy(feat + size_t(e) * kNF, f, sizeof f);
      label[e] = uint8_t(std::min(match, 254));
      return false;
    }
    return net_logit(pa.w, pa.norm, f) >= pa.zthr;
  };
  bool stopped = ckpt(0, 0);
  std::vector<uint32_t> ids;
  for (int it = 0; it < a.epochs && !stopped; ++it) {
    ids.clear();
    int nexp = 0;
    for (int j = 0; j < a.extend_to; ++j) {
      uint32_t pj = res[j].second;
      if (expanded[pj]) continue;
      expanded[pj] = 1;
      ++nexp;
      for (uint64_t g = ix.goff[pj]; g < ix.goff[pj + 1]; ++g) {
        uint32_t id = ix.gids[g];
        if (visited[id]) continue;
        visited[id] = 1;
        ids.push_back(id);
      }
    }
    if (nexp == 0) break;
    for (uint32_t id : ids) res.push_back({l2_avx(q, ix.row(id), ix.d), id});
    nvis += ids.size();
    std::partial_sort(res.begin(), res.begin() + R, res.end());
    res.resize(R);
    stopped = ckpt(it + 1, nexp);
  }
  if (pa.mode == 1 && !stopped) {
    uint64_t nd = 0;
    search_nnexp(ix, q, qid, pa.full, out_ids, out_d, nd);
    return false;
  }
  if (pa.mode == 3 && !stopped) {
    const SearchArgs& f = pa.full;
    const size_t Rf = size_t(std::max(f.extend_to, f.K));
    const int want = std::min(f.init_num - a.init_num, int(n - nvis));
    if (size_t(want) + R < Rf) {
      uint64_t nd = 0;
      search_nnexp(ix, q, qid, f, out_ids, out_d, nd);
      return false;
    }
    std::vector<uint32_t> more;
    take_seeds(ix, seq, resume, want, qid, visited, more, nullptr);
    for (uint32_t id : more) res.push_back({l2_avx(q, ix.row(id), ix.d), id});
    std::partial_sort(res.begin(), res.begin() + Rf, res.end());
    res.resize(Rf);
    for (int it = 0; it < f.epochs; ++it) {
      ids.clear();
      for (int j = 0; j < f.extend_to; ++j) {
        uint32_t pj = res[j].second;
        if (expanded[pj]) continue;
        expanded[pj] = 1;
        for (uint64_t g = ix.goff[pj]; g < ix.goff[pj + 1]; ++g) {
          uint32_t id = ix.gids[g];
          if (visited[id]) continue;
          visited[id] = 1;
          ids.push_back(id);
        }
      }
      for (uint32_t id : ids) res.push_back({l2_avx(q, ix.row(id), ix.d), id});
      std::partial_sort(res.begin(), res.begin() + Rf, res.end());
      res.resize(Rf);
    }
    for (int i = 0; i < K; ++i) {
      out_ids[i] = int32_t(res[i].second);
      if (out_d) out_d[i] = res[i].first;
    }
    return false;
  }
  for (int i = 0; i < K; ++i) {
    out_ids[i] = int32_t(res[i].second);
    if (out_d) out_d[i] = res[i].first;
  }
  return stopped;
}

// Mean binary cross-entropy on the logit and its gradient, in double, with
// the parameters as given (the caller rounds them when it wants the BF16
// network); inputs standardised in float as the network does.
double net_loss_grad64(const float* w, const float* norm, const float* feat, const uint8_t* label, size_t n,
                       double* grad) {
  if (grad) std::fill(grad, grad + kNP, 0.0);
  double loss = 0.0;
  for (size_t r = 0; r < n; ++r) {
    double x[kNF], z1[kNH], h1[kNH], z2[kNH], h2[kNH];
    for (int i = 0; i < kNF; ++i) x[i] = double((feat[r * kNF + i] - norm[i]) * norm[kNF + i]);
    for (int j = 0; j < kNH; ++j) {
      double a = w[kB1 + j];
      for (int i = 0; i < kNF; ++i) a += double(w[kW1 + i * kNH + j]) * x[i];
      z1[j] = a;
      h1[j] = a > 0 ? a : 0;
    }
    for (int j = 0; j < kNH; ++j) {
      double a = w[kB2 + j];
      for (int i = 0; i < kNH; ++i) a += double(w[kW2 + i * kNH + j]) * h1[i];
      z2[j] = a;
      h2[j] = a > 0 ? a : 0;
    }
    double z = w[kB3];
    for (int j = 0; j < kNH; ++j) z += double(w[kW3 + j]) * h2[j];
    const double y = label[r] ? 1.0 : 0.0;
    loss += std::max(z, 0.0) - z * y + std::log1p(std::exp(-std::fabs(z)));
    if (!grad) continue;
    const double dz = (1.0 / (1.0 + std::exp(-z)) - y) / double(n);
    double d2[kNH], d1[kNH];
    grad[kB3] += dz;
    for (int j = 0; j < kNH; ++j) {
      grad[kW3 + j] += dz * h2[j];
      d2[j] = z2[j] > 0 ? dz * double(w[kW3 + j]) : 0.0;
      grad[kB2 + j] += d2[j];
    }
    for (int i = 0; i < kNH; ++i) {
      double a = 0;
      for (int j = 0; j < kNH; ++j) {
        grad[kW2 + i * kNH + j] += h1[i] * d2[j];
        a += double(w[kW2 + i * kNH + j]) * d2[j];
      }
      d1[i] = z1[i] > 0 ? a : 0.0;
    }
    for (int j = 0; j < kNH; ++j) {
      grad[kB1 + j] += d1[j];
      for (int i = 0; i < kNF; ++i) grad[kW1 + i * kNH + j] += x[i] * d1[j];
    }
  }
  return loss / double(n);
}

// Hash-only search (hashing_index.hpp:929-1005 getNeighbors32, 1007-1085
// getNeighbors64; no graph loaded): the first init_num distinct bucket
// members, a random top-up only below K, and the K nearest of them.
void search_hash_only(const Index& ix, const float* q, uint64_t qid, const SearchArgs& a, int32_t* out_ids,
                      float* out_d, uint64_t& ndist) {
  std::vector<uint8_t> visited(ix.n, 0);
  std::vector<uint32_t> pool;
  seed_pool(ix, q, 0, a.init_num, 0, qid, visited, pool, a.K);
  std::vector<std::pair<float, uint32_t>> res;
  for (uint32_t id : pool) res.push_back({l2_avx(q, ix.row(id), ix.d), id});
  ndist += pool.size();
  std::partial_sort(res.begin(), res.begin() + a.K, res.end());
  for (int i = 0; i < a.K; ++i) {
    out_ids[i] = int32_t(res[i].second);
    if (out_d) out_d[i] = res[i].first;
  }
}

// Hamming-ball masks (generateMask32/64): 0, then for r = 1..radius the
// r-bit combinations of bits 0..bits-1 in lexicographic order of (i < j <
// ...), as the reference's nested loops emit them.
void gen_masks(int bits, int radius, std::vector<uint64_t>& out) {
  out.assign(1, 0ull);
  for (int r = 1; r <= radius && r <= bits; ++r) {
    std::vector<int> idx(r);
    for (int i = 0; i < r; ++i) idx[i] = i;
    for (;;) {
      uint64_t m = 0;
      for (int i : idx) m |= 1ull << i;
      out.push_back(m);
      int i = r - 1;
      while (i >= 0 && idx[i] == bits - r + i) --i;
      if (i < 0) break;
      ++idx[i];
      for (int j = i + 1; j < r; ++j) idx[j] = idx[j - 1] + 1;
    }
  }
}

// Search, method 1: getNeighbors_kgraph (kdtreeub_index.hpp:496-654).  The
// seeds of one epoch are ordered by (distance, pool rank) where the reference
// calls the unstable std::sort on distance alone (588).
void search_kgraph(const Index& ix, const float* q, uint64_t qid, const SearchArgs& a, int init_num_eff,
                   bool sorted, unsigned lsize, int32_t* out_ids, float* out_d, uint64_t& ndist) {
  const size_t n = ix.n;
  std::vector<uint8_t> flags(n, 0);
  std::vector<uint32_t> pool;
  seed_pool(ix, q, lsize, init_num_eff, a.trees, qid, flags, pool);
  const unsigned E = unsigned(a.extend_to), K = unsigned(a.K);
  unsigned pool_size = unsigned(a.epochs) * E;
  std::vector<std::pair<float, uint32_t>> res;
  for (uint32_t id : pool) res.push_back({l2_avx(q, ix.row(id), ix.d), id});
  ndist += pool.size();
  if (sorted) {
    std::partial_sort(res.begin(), res.begin() + pool_size, res.end());
    res.resize(pool_size);
  }
  std::fill(flags.begin(), flags.end(), 0);
  std::vector<Point> knn(K + E + 1);
  std::vector<Point> results;
  std::vector<std::pair<std::pair<float, unsigned>, Point>> tmp;
  for (unsigned it = 0; it < unsigned(a.epochs); ++it) {
    unsigned L = 0;
    tmp.clear();
    for (unsigned j = 0; j < E; ++j) {
      const auto& e = res[it * E + j];
      if (flags[e.second]) continue;
      flags[e.second] = 1;
      tmp.push_back({{e.first, j}, Point{e.second, e.first, true}});
    }
    std::sort(tmp.begin(), tmp.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    for (auto& t : tmp) knn[L++] = t.second;
    unsigned k = 0;
    while (k < L) {
      unsigned nk = L;
      if (knn[k].flag) {
        knn[k].flag = false;
        uint32_t nd = knn[k].id;
        uint64_t deg = ix.goff[nd + 1] - ix.goff[nd];
        unsigned maxM = unsigned(std::min<uint64_t>(E, deg));
        for (unsigned m = 0; m < maxM; ++m) {
          uint32_t id = ix.gids[ix.goff[nd] + m];
          if (flags[id]) continue;
          flags[id] = 1;
          float dist = l2_avx(q, ix.row(id), ix.d);
          ++ndist;
          unsigned r = insert_into_knn(knn.data(), L, Point{id, dist, true});
          if (L + 1 < knn.size()) ++L;
          if (r < nk) nk = r;
        }
      }
      if (nk <= k) k = nk;
      else ++k;
    }
    if (L > K) L = K;
    if (results.empty()) {
      results.resize(L + 1, Point{0xffffffffu, FLT_MAX, false});
      std::copy(knn.begin(), knn.begin() + L, results.begin());
    } else {
      for (unsigned l = 0; l < L; ++l) {
        unsigned r = insert_into_knn(results.data(), unsigned(results.size() - 1), knn[l]);
        if (r < results.size() && results.size() < K + 1) results.resize(results.size() + 1, Point{0xffffffffu, FLT_MAX, false});
      }
    }
  }
  // Only valid entries are reported; the reference also emits the unused
  // spare slot when fewer than K were found (650-652).
  unsigned valid = results.empty() ? 0u : unsigned(results.size() - 1);
  for (unsigned i = 0; i < K; ++i) {
    bool ok = i < valid;
    out_ids[i] = ok ? int32_t(results[i].id) : -1;
    if (out_d) out_d[i] = ok ? results[i].dist : FLT_MAX;
  }
}

}  // namespace

// ===========================================================================
// C interface (ctypes) -------------------------------------------------------
// ===========================================================================
extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

uint64_t orc_hash4(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b) { return hash4(seed, stream, a, b); }
uint64_t orc_mix64(uint64_t x) { return mix64(x); }
uint64_t orc_perm(uint64_t seed, uint64_t tree, uint64_t n, uint64_t i) { return perm_index(seed, tree, n, i); }
float orc_l2(const float* a, const float* b, size_t d) { return l2_avx(a, b, d); }

unsigned orc_insert_into_knn(uint32_t* ids, float* dists, unsigned K, uint32_t id, float dist) {
  std::vector<Point> v(K + 1);
  for (unsigned i = 0; i <= K; ++i) v[i] = Point{ids[i], dists[i], false};
  unsigned r = insert_into_knn(v.data(), K, Point{id, dist, true});
  for (unsigned i = 0; i <= K; ++i) { ids[i] = v[i].id; dists[i] = v[i].dist; }
  return r;
}

void orc_set_threads(int t) { if (t > 0) omp_set_num_threads(t); }

// Exact kNN: full scan, k smallest by (distance, id) (matrix.hpp:57-68 with the
// self row excluded when exclude_self is set).
void orc_exact_knn(const float* base, size_t n, size_t d, size_t bstride, const float* q, size_t nq, size_t qstride,
                   int k, int exclude_self, int32_t* ids, float* dists) {
#pragma omp parallel for schedule(dynamic, 4)
  for (long long qi = 0; qi < (long long)nq; ++qi) {
    std::vector<std::pair<float, uint32_t>> all;
    all.reserve(n);
    const float* qr = q + size_t(qi) * qstride;
    for (size_t i = 0; i < n; ++i) {
      if (exclude_self && i == size_t(qi)) continue;
      all.push_back({l2_avx(qr, base + i * bstride, d), uint32_t(i)});
    }
    size_t kk = std::min<size_t>(size_t(k), all.size());
    std::partial_sort(all.begin(), all.begin() + kk, all.end());
    for (size_t j = 0; j < size_t(k); ++j) {
      ids[size_t(qi) * k + j] = j < kk ? int32_t(all[j].second) : -1;
      if (dists) dists[size_t(qi) * k + j] = j < kk ? all[j].first : FLT_MAX;
    }
  }
}

void* orc_new(const float* data, size_t n, size_t d, size_t stride) {
  Index* ix = new Index();
  ix->n = n;
  ix->d = d;
  ix->stride = (d + 7) & ~size_t(7);
  ix->data.assign(n * ix->stride, 0.f);
  for (size_t i = 0; i < n; ++i) std::memcpy(&ix->data[i * ix->stride], data + i * stride, d * sizeof(float));
  return ix;
}
void orc_free(void* h) { delete static_cast<Index*>(h); }

void orc_set_params(void* h, int trees, int trees_build, int ml, int epochs, int check, int L, int K, int S, int tns,
                    uint64_t seed, int join_mode) {
  Index* ix = static_cast<Index*>(h);
  ix->p.trees = trees;
  ix->p.trees_build = trees_build > 0 ? trees_build : trees;
  ix->p.ml = ml;
  ix->p.epochs = epochs;
  ix->p.check = check;
  ix->p.L = L;
  ix->p.K = K;
  ix->p.S = S;
  ix->p.tns = tns;
  ix->p.seed = seed;
  ix->p.join_mode = join_mode;
}

// Test hooks for the random-fill branches: GenRandom (kdtreeub_index.hpp:
// 1121-1135), the search seed pool with its top-up (387-417), and the
// reverse lists around the cap of 100 (base_index.hpp:421-436).
void orc_gen_random(uint64_t seed, uint32_t node, uint32_t* out, unsigned size, unsigned N) {
  gen_random(seed, node, out, size, N);
}
void orc_keep_rev(void* h, int on) { static_cast<Index*>(h)->keep_rev = on != 0; }
// which: 0 new before sampling, 1 old before, 2 new after, 3 old after.
size_t orc_rev_total(void* h, int which) {
  size_t t = 0;
  for (auto& v : static_cast<Index*>(h)->rev[which & 3]) t += v.size();
  return t;
}
void orc_get_rev(void* h, int which, uint64_t* off, uint32_t* ids) {
  Index* ix = static_cast<Index*>(h);
  const auto& r = ix->rev[which & 3];
  off[0] = 0;
  for (size_t i = 0; i < r.size(); ++i) {
    std::memcpy(ids + off[i], r[i].data(), r[i].size() * 4);
    off[i + 1] = off[i] + r[i].size();
  }
}

int orc_build_trees(void* h) { return build_trees(*static_cast<Index*>(h)); }
int orc_merge_init(void* h) { return merge_and_init(*static_cast<Index*>(h)); }
// One refine iteration (join + update(100)): base_index.hpp:443-448.
int orc_refine_iter(void* h) {
  Index* ix = static_cast<Index*>(h);
  if (ix->p.join_mode == 1) join_sequential(*ix);
  else if (ix->p.join_mode == 2) join_parallel(*ix);
  else join_batched(*ix);
  ix->iter++;
  update(*ix, 100);
  return 0;
}
int orc_finalize(void* h) { finalize_graph(*static_cast<Index*>(h)); return 0; }

// buildIndex: buildIndexImpl (kdtreeub_index.hpp:92-113) = initGraph + refineGraph.
int orc_build_index(void* h) {
  Index* ix = static_cast<Index*>(h);
  int rc = build_trees(*ix);
  if (rc) return rc;
  rc = merge_and_init(*ix);
  if (rc) return rc;
  for (int e = 0; e < ix->p.epochs; ++e) orc_refine_iter(h);
  finalize_graph(*ix);
  return 0;
}

size_t orc_node_count(void* h) { return static_cast<Index*>(h)->nodes.size(); }
int orc_tree_count(void* h) { return static_cast<Index*>(h)->T; }
void orc_get_nodes(void* h, int32_t* dim, float* val, int32_t* left, int32_t* right, uint32_t* start, uint32_t* end,
                   int32_t* roots) {
  Index* ix = static_cast<Index*>(h);
  for (size_t i = 0; i < ix->nodes.size(); ++i) {
    const Node& nd = ix->nodes[i];
    dim[i] = nd.dim; val[i] = nd.val; left[i] = nd.left; right[i] = nd.right; start[i] = nd.start; end[i] = nd.end;
  }
  for (int t = 0; t < ix->T; ++t) roots[t] = ix->roots[t];
}
void orc_get_leaf_lists(void* h, uint32_t* out) {
  Index* ix = static_cast<Index*>(h);
  std::memcpy(out, ix->leaf.data(), ix->leaf.size() * sizeof(uint32_t));
}
// Install trees (e.g. the GPU's or a loaded file's) for search.
void orc_set_trees(void* h, int T, size_t m, const int32_t* dim, const float* val, const int32_t* left,
                   const int32_t* right, const uint32_t* start, const uint32_t* end, const int32_t* roots,
                   const uint32_t* leaf) {
  Index* ix = static_cast<Index*>(h);
  ix->T = T;
  ix->nodes.assign(m, Node());
  for (size_t i = 0; i < m; ++i) {
    Node& nd = ix->nodes[i];
    nd.dim = dim[i]; nd.val = val[i]; nd.left = left[i]; nd.right = right[i]; nd.start = start[i]; nd.end = end[i];
  }
  ix->roots.assign(roots, roots + T);
  for (int t = 0; t < T; ++t) {  // recompute parent/depth/tree
    std::vector<int32_t> st{roots[t]};
    ix->nodes[roots[t]].depth = 0;
    while (!st.empty()) {
      int32_t x = st.back();
      st.pop_back();
      Node& nd = ix->nodes[x];
      nd.tree = t;
      for (int32_t c : {nd.left, nd.right})
        if (c >= 0) { ix->nodes[c].parent = x; ix->nodes[c].depth = nd.depth + 1; st.push_back(c); }
    }
  }
  ix->leaf.assign(leaf, leaf + size_t(T) * ix->n);
}

size_t orc_graph_total(void* h) { return static_cast<Index*>(h)->gids.size(); }
void orc_get_graph(void* h, uint64_t* off, uint32_t* ids) {
  Index* ix = static_cast<Index*>(h);
  std::memcpy(off, ix->goff.data(), ix->goff.size() * sizeof(uint64_t));
  std::memcpy(ids, ix->gids.data(), ix->gids.size() * sizeof(uint32_t));
}
void orc_set_graph(void* h, const uint64_t* off, const uint32_t* ids) {
  Index* ix = static_cast<Index*>(h);
  ix->goff.assign(off, off + ix->n + 1);
  ix->gids.assign(ids, ids + off[ix->n]);
}

// Build state inspection: pools (n x L keys (dist<<32|id), flags, fill) and the
// join lists (CSR).
void orc_get_pools(void* h, uint64_t* keys, uint8_t* flags, uint32_t* fill, uint32_t* range) {
  Index* ix = static_cast<Index*>(h);
  const unsigned Lp = unsigned(ix->p.L);
  for (size_t i = 0; i < ix->n; ++i) {
    for (unsigned l = 0; l < Lp; ++l) {
      const Point& e = ix->pool[i * (Lp + 1) + l];
      bool v = l < ix->fill[i];
      keys[i * Lp + l] = v ? dkey(e.dist, e.id) : ~0ULL;
      flags[i * Lp + l] = v && e.flag ? 1 : 0;
    }
    fill[i] = ix->fill[i];
    if (range) range[i] = ix->range[i];
  }
}
size_t orc_list_total(void* h, int which) {
  Index* ix = static_cast<Index*>(h);
  auto& v = which ? ix->nn_old : ix->nn_new;
  size_t s = 0;
  for (auto& x : v) s += x.size();
  return s;
}
void orc_get_lists(void* h, int which, uint64_t* off, uint32_t* ids) {
  Index* ix = static_cast<Index*>(h);
  auto& v = which ? ix->nn_old : ix->nn_new;
  off[0] = 0;
  for (size_t i = 0; i < ix->n; ++i) {
    std::copy(v[i].begin(), v[i].end(), ids + off[i]);
    off[i + 1] = off[i] + v[i].size();
  }
}
void orc_get_found(void* h, uint8_t* found, float* radiusM) {
  Index* ix = static_cast<Index*>(h);
  for (size_t i = 0; i < ix->n; ++i) { found[i] = ix->found[i]; if (radiusM) radiusM[i] = ix->radiusM[i]; }
}
// counters: [merge distances, init distances, join distances, error_flag]
void orc_get_counters(void* h, uint64_t* out) {
  Index* ix = static_cast<Index*>(h);
  out[0] = ix->n_dist_merge;
  out[1] = ix->n_dist_init;
  out[2] = ix->n_dist_join;
  out[3] = ix->error_flag ? 1 : 0;
}
size_t orc_join_pairs(void* h, uint64_t* out, size_t cap) {
  Index* ix = static_cast<Index*>(h);
  size_t m = std::min(cap, ix->join_pairs_per_iter.size());
  for (size_t i = 0; i < m; ++i) out[i] = ix->join_pairs_per_iter[i];
  return ix->join_pairs_per_iter.size();
}

// Probe search, trace mode (2) or learned mode (1), see search_probe.
int orc_search_probe(void* h, const float* q, size_t nq, size_t qstride, int k, int epochs, int init_num,
                     int extend_to, int search_trees, int p_pool, int p_extend, int p_epochs, int mode,
                     const float* params, const float* norm, float zthr, const int32_t* plain, int32_t* ids,
                     float* dists, float* feat, uint8_t* label, uint8_t* stopped) {
  Index* ix = static_cast<Index*>(h);
  if (extend_to > init_num) extend_to = init_num;
  if (p_extend > p_pool) p_extend = p_pool;
  if (search_trees <= 0 || search_trees > ix->T) { g_err = "wrong tree number"; return -4; }
  if (mode < 1 || mode > 3) { g_err = "mode must be 1, 2 or 3"; return -2; }
  if (mode == 3 && (p_pool > init_num || p_extend > extend_to || std::max(p_extend, k) > p_pool)) {
    g_err = "probe pool and extend_to must lie within the search's, max(extend_to, k) within the probe pool";
    return -2;
  }
  ProbeArgs pa{};
  pa.full = SearchArgs{epochs, init_num, extend_to, search_trees, k};
  pa.probe = SearchArgs{p_epochs, p_pool, p_extend, search_trees, k};
  pa.mode = mode;
  std::vector<float> w(kNP);
  if (mode != 2)
    for (int i = 0; i < kNP; ++i) w[i] = round_bf16(params[i]);
  pa.w = w.data();
  pa.norm = norm;
  pa.zthr = zthr;
  const size_t C = size_t(p_epochs) + 1;
  if (mode == 2) std::fill(label, label + nq * C, uint8_t(255));
  size_t qs = (ix->d + 7) & ~size_t(7);
  std::vector<float> qpad(nq * qs, 0.f);
  for (size_t i = 0; i < nq; ++i) std::memcpy(&qpad[i * qs], q + i * qstride, ix->d * sizeof(float));
#pragma omp parallel for schedule(dynamic, 8)
  for (long long qi = 0; qi < (long long)nq; ++qi) {
    ProbeArgs a2 = pa;
    a2.plain = plain ? plain + size_t(qi) * k : nullptr;
    bool st = search_probe(*ix, &qpad[size_t(qi) * qs], uint64_t(qi), a2, ids + size_t(qi) * k,
                           dists ? dists + size_t(qi) * k : nullptr, feat ? feat + size_t(qi) * C * kNF : nullptr,
                           label ? label + size_t(qi) * C : nullptr);
    if (stopped) stopped[qi] = st ? 1 : 0;
  }
  return 0;
}

// Network logits (FP32, BF16-rounded parameters) and the FP64 loss/gradient.
void orc_mlp_logits(const float* params, const float* norm, const float* feat, size_t n, float* z) {
  std::vector<float> w(kNP);
  for (int i = 0; i < kNP; ++i) w[i] = round_bf16(params[i]);
  for (size_t r = 0; r < n; ++r) z[r] = net_logit(w.data(), norm, feat + r * kNF);
}
double orc_mlp_grad64(const float* params, const float* norm, const float* feat, const uint8_t* label, size_t n,
                      int round_params, double* grad) {
  std::vector<float> w(params, params + kNP);
  if (round_params)
    for (auto& v : w) v = round_bf16(v);
  return net_loss_grad64(w.data(), norm, feat, label, n, grad);
}
int orc_mlp_param_count() { return kNP; }

// Hashing index: T tables of base codes (T x n) and query codes (T x nq),
// already shifted right by the code-length shift; bits = the code length
// used (codelength - lenshift), radius = the Hamming radius.
int orc_hash_set(void* h, int T, int bits, int radius, const uint64_t* base_codes, const uint64_t* query_codes,
                 size_t nq) {
  Index* ix = static_cast<Index*>(h);
  if (T <= 0 || bits <= 0 || bits > 64 || radius < 0) { g_err = "invalid hashing parameters"; return -2; }
  HashState& hs = ix->hs;
  hs.T = T;
  hs.bits = bits;
  hs.radius = radius;
  hs.nq = nq;
  hs.tab.assign(T, {});
  for (int t = 0; t < T; ++t)
    for (size_t i = 0; i < ix->n; ++i) hs.tab[t][base_codes[size_t(t) * ix->n + i]].push_back(uint32_t(i));
  hs.qcodes.assign(query_codes, query_codes + size_t(T) * nq);
  gen_masks(bits, radius, hs.masks);
  ix->hashing = true;
  return 0;
}
size_t orc_hash_masks(void* h, uint64_t* out, size_t cap) {
  const HashState& hs = static_cast<Index*>(h)->hs;
  for (size_t i = 0; i < hs.masks.size() && i < cap; ++i) out[i] = hs.masks[i];
  return hs.masks.size();
}
// The seed ids of query qi before the distances (the bucket sequence's
// first init_num distinct members, then the random top-up to topup_to ids
// when topup_to >= 0, else to init_num).
size_t orc_hash_seeds(void* h, size_t qi, int init_num, int topup_to, uint32_t* out) {
  Index* ix = static_cast<Index*>(h);
  std::vector<uint8_t> visited(ix->n, 0);
  std::vector<uint32_t> pool;
  seed_pool(*ix, nullptr, 0, init_num, 0, uint64_t(qi), visited, pool, topup_to);
  std::memcpy(out, pool.data(), pool.size() * 4);
  return pool.size();
}

// knnSearch on a hashing index (hashing_index.hpp:891-927): without a graph
// the hash-only search; with one, method 0 = kgraph and 1 = nnexp (the
// hashing index numbers them the other way round from the kd-tree index).
int orc_hash_search(void* h, const float* q, size_t nq, size_t qstride, int k, int epochs, int* init_num_io,
                    int extend_to, int method, int32_t* ids, float* dists, uint64_t* ndist) {
  Index* ix = static_cast<Index*>(h);
  if (!ix->hashing) { g_err = "no hash tables"; return -3; }
  if (nq > ix->hs.nq) { g_err = "more queries than query codes"; return -2; }
  int init_num = *init_num_io;
  if (extend_to > init_num) extend_to = init_num;
  if (init_num > int(ix->n) || k > std::max(extend_to, init_num)) { g_err = "pool larger than the data set"; return -2; }
  SearchArgs a{epochs, init_num, extend_to, ix->hs.T, k};
  std::vector<float> qpad;
  size_t qs = (ix->d + 7) & ~size_t(7);
  qpad.assign(nq * qs, 0.f);
  for (size_t i = 0; i < nq; ++i) std::memcpy(&qpad[i * qs], q + i * qstride, ix->d * sizeof(float));
  const bool graph = ix->goff.size() == ix->n + 1;
  if (!graph) {
#pragma omp parallel for schedule(dynamic, 8)
    for (long long qi = 0; qi < (long long)nq; ++qi) {
      uint64_t nd = 0;
      search_hash_only(*ix, &qpad[size_t(qi) * qs], uint64_t(qi), a, ids + size_t(qi) * k,
                       dists ? dists + size_t(qi) * k : nullptr, nd);
      if (ndist) ndist[qi] = nd;
    }
  } else if (method == 1) {
#pragma omp parallel for schedule(dynamic, 8)
    for (long long qi = 0; qi < (long long)nq; ++qi) {
      uint64_t nd = 0;
      search_nnexp(*ix, &qpad[size_t(qi) * qs], uint64_t(qi), a, ids + size_t(qi) * k,
                   dists ? dists + size_t(qi) * k : nullptr, nd);
      if (ndist) ndist[qi] = nd;
    }
  } else if (method == 0) {
    bool sorted = true;
    unsigned pool_size = unsigned(epochs) * unsigned(extend_to);
    if (pool_size >= unsigned(init_num)) { init_num = int(pool_size); sorted = false; }
    if (init_num > int(ix->n)) { g_err = "pool larger than the data set"; return -2; }
    *init_num_io = init_num;
    a.init_num = init_num;
#pragma omp parallel for schedule(dynamic, 8)
    for (long long qi = 0; qi < (long long)nq; ++qi) {
      uint64_t nd = 0;
      search_kgraph(*ix, &qpad[size_t(qi) * qs], uint64_t(qi), a, init_num, sorted, 0, ids + size_t(qi) * k,
                    dists ? dists + size_t(qi) * k : nullptr, nd);
      if (ndist) ndist[qi] = nd;
    }
  } else {
    g_err = "no such searching method";
    return -5;
  }
  return 0;
}

// Seed pool of one query in sequence order (method 0), init_num ids.
int orc_search_seeds(void* h, const float* q, uint64_t qid, int init_num, int search_trees, uint32_t* out) {
  Index* ix = static_cast<Index*>(h);
  if (search_trees <= 0 || search_trees > ix->T) { g_err = "wrong tree number"; return -4; }
  if (init_num <= 0 || init_num > int(ix->n)) { g_err = "pool larger than the data set"; return -2; }
  std::vector<float> qpad((ix->d + 7) & ~size_t(7), 0.f);
  std::memcpy(qpad.data(), q, ix->d * sizeof(float));
  std::vector<uint8_t> visited(ix->n, 0);
  std::vector<uint32_t> pool;
  unsigned lsize = unsigned(init_num * 2 / (5 * search_trees) + 1);
  seed_pool(*ix, qpad.data(), lsize, init_num, search_trees, qid, visited, pool);
  std::memcpy(out, pool.data(), pool.size() * 4);
  return 0;
}

// knnSearch with setSearchParams semantics (base_index.hpp:153-161):
// extend_to is clamped to init_num; method 0 = nnexp, 1 = kgraph (which
// raises init_num to epochs*extend_to for good, kdtreeub_index.hpp:510-514;
// the effective init_num is written back through *init_num_io).
int orc_search(void* h, const float* q, size_t nq, size_t qstride, int k, int epochs, int* init_num_io,
               int extend_to, int search_trees, int method, int32_t* ids, float* dists, uint64_t* ndist) {
  Index* ix = static_cast<Index*>(h);
  int init_num = *init_num_io;
  if (extend_to > init_num) extend_to = init_num;
  if (search_trees <= 0) { g_err = "search tree number must be positive"; return -2; }
  if (search_trees > ix->T) { g_err = "wrong tree number"; return -4; }
  if (init_num > int(ix->n) || k > std::max(extend_to, init_num)) { g_err = "pool larger than the data set"; return -2; }
  SearchArgs a{epochs, init_num, extend_to, search_trees, k};
  std::vector<float> qpad;
  size_t qs = (ix->d + 7) & ~size_t(7);
  qpad.assign(nq * qs, 0.f);
  for (size_t i = 0; i < nq; ++i) std::memcpy(&qpad[i * qs], q + i * qstride, ix->d * sizeof(float));
  if (method == 0) {
#pragma omp parallel for schedule(dynamic, 8)
    for (long long qi = 0; qi < (long long)nq; ++qi) {
      uint64_t nd = 0;
      search_nnexp(*ix, &qpad[size_t(qi) * qs], uint64_t(qi), a, ids + size_t(qi) * k, dists ? dists + size_t(qi) * k : nullptr, nd);
      if (ndist) ndist[qi] = nd;
    }
  } else if (method == 1) {
    unsigned lsize = unsigned(init_num * 2 / (5 * search_trees) + 1);
    bool sorted = true;
    unsigned pool_size = unsigned(epochs) * unsigned(extend_to);
    if (pool_size >= unsigned(init_num)) { init_num = int(pool_size); sorted = false; }
    if (init_num > int(ix->n)) { g_err = "pool larger than the data set"; return -2; }
    *init_num_io = init_num;
#pragma omp parallel for schedule(dynamic, 8)
    for (long long qi = 0; qi < (long long)nq; ++qi) {
      uint64_t nd = 0;
      search_kgraph(*ix, &qpad[size_t(qi) * qs], uint64_t(qi), a, init_num, sorted, lsize, ids + size_t(qi) * k,
                    dists ? dists + size_t(qi) * k : nullptr, nd);
      if (ndist) ndist[qi] = nd;
    }
  } else {
    g_err = "no such searching method";
    return -5;
  }
  return 0;
}

}  // extern "C"

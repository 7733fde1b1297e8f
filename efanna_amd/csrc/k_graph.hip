// kNN-graph construction after the trees: conquer-to-depth merge, random
// init, NN-descent refine, graph output.
//
// Merge (kdtreeub_index.hpp:1000-1081, run as one thread runs it, 1367-1374):
//   every point sees its candidates in the reference's single-thread order —
//   trees in order; inside a tree, internal nodes X at depth >= ml deepest
//   first (post-order along the point's root path); at X the point is either a
//   prober of the smaller child (it meets the members of the leaf it reaches
//   in the larger child, in range order) or a member of a leaf of the larger
//   child (it meets, in range order, the probers that reach its leaf).
//   Probe records (leaf reached, depth of X, prober position) are bucketed by
//   leaf and sorted; one thread per point then replays its arrivals into an
//   S-bounded (distance, id) set, appending rejected ids to nn_new while it
//   holds fewer than 2S (1051-1077).  The result equals the sequential merge.
// Init (1377-1423): nn_new below 2S is replaced by GenRandom's draw (counter-
//   based generator); the set is topped up to S with generator ids; pool =
//   the set in (distance, id) order, all flagged new; nn_new[0..S) = pool ids.
// Join (base_index.hpp:328-371), batched: every pair of every node's new x
//   new (i < j) and new x old lists is measured once; a candidate for a
//   target whose pool is full is kept only when it beats the pool's last
//   entry (as it stood when the iteration began).  Candidates are bucketed by
//   target; each target's pool becomes the best L of (pool + candidates) by
//   (distance, id), duplicates removed, existing entries keeping their flag.
//   `found` of a joining node: one of its first-endpoint candidates sits at
//   rank < Check_K in the merged pool (the lowest joining node gets it).
// Update (base_index.hpp:373-437): Range / radiusM, forward new/old lists,
//   reverse lists for edges longer than the other end's radiusM, reverse lists
//   above 100 keep the 100 ids with the smallest generator value.
#include <algorithm>
#include <chrono>
#include <type_traits>
#include <vector>

#include "efg_device.hpp"
#include "efg_internal.hpp"
#include "efg_mfma.hpp"

namespace efg {

namespace {

constexpr uint32_t kFirst = 0x80000000u;  // staging src flag: first endpoint of the pair

__device__ inline float l2_thread(const float* __restrict__ a, const float* __restrict__ b, int D) {
  // single-thread form of the reference's AVX order
  const int DR = D & 31, DD = D - DR;
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int c = DD; c < D; c += 8) {
    const float4 x0 = *reinterpret_cast<const float4*>(a + c), x1 = *reinterpret_cast<const float4*>(a + c + 4);
    const float4 y0 = *reinterpret_cast<const float4*>(b + c), y1 = *reinterpret_cast<const float4*>(b + c + 4);
    float t;
    t = x0.x - y0.x; acc[0] = __builtin_fmaf(t, t, acc[0]);
    t = x0.y - y0.y; acc[1] = __builtin_fmaf(t, t, acc[1]);
    t = x0.z - y0.z; acc[2] = __builtin_fmaf(t, t, acc[2]);
    t = x0.w - y0.w; acc[3] = __builtin_fmaf(t, t, acc[3]);
    t = x1.x - y1.x; acc[4] = __builtin_fmaf(t, t, acc[4]);
    t = x1.y - y1.y; acc[5] = __builtin_fmaf(t, t, acc[5]);
    t = x1.z - y1.z; acc[6] = __builtin_fmaf(t, t, acc[6]);
    t = x1.w - y1.w; acc[7] = __builtin_fmaf(t, t, acc[7]);
  }
  for (int c = 0; c < DD; c += 8) {
    const float4 x0 = *reinterpret_cast<const float4*>(a + c), x1 = *reinterpret_cast<const float4*>(a + c + 4);
    const float4 y0 = *reinterpret_cast<const float4*>(b + c), y1 = *reinterpret_cast<const float4*>(b + c + 4);
    float t;
    t = x0.x - y0.x; acc[0] = __builtin_fmaf(t, t, acc[0]);
    t = x0.y - y0.y; acc[1] = __builtin_fmaf(t, t, acc[1]);
    t = x0.z - y0.z; acc[2] = __builtin_fmaf(t, t, acc[2]);
    t = x0.w - y0.w; acc[3] = __builtin_fmaf(t, t, acc[3]);
    t = x1.x - y1.x; acc[4] = __builtin_fmaf(t, t, acc[4]);
    t = x1.y - y1.y; acc[5] = __builtin_fmaf(t, t, acc[5]);
    t = x1.z - y1.z; acc[6] = __builtin_fmaf(t, t, acc[6]);
    t = x1.w - y1.w; acc[7] = __builtin_fmaf(t, t, acc[7]);
  }
  float s = acc[0];
  for (int j = 1; j < 8; ++j) s = s + acc[j];
  return s;
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------
// merge
// ---------------------------------------------------------------------------
// Depth and tree of every node (walk to the root); the largest depth goes to
// *max_depth.
__global__ void k_node_depth(const int32_t* __restrict__ parent, int M, const int32_t* __restrict__ roots, int T,
                             int8_t* __restrict__ depth, int32_t* __restrict__ tree_of, int* __restrict__ max_depth) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  int dd = 0, x = i;
  while (parent[x] >= 0) {
    x = parent[x];
    ++dd;
  }
  depth[i] = int8_t(dd > 127 ? 127 : dd);
  int t = 0;
  while (t < T - 1 && roots[t] != x) ++t;
  tree_of[i] = t;
  atomicMax(max_depth, dd);
}

// leaf_at[t][pos] = leaf holding position pos; error flag for leaves shallower
// than ml in the first TB trees (getMergeLevelNodeList, 1000-1011).
__global__ void k_leaf_at(const int4* __restrict__ node4, const uint2* __restrict__ range,
                          const int8_t* __restrict__ depth, const int32_t* __restrict__ tree_of, int M, uint32_t n,
                          int TB, int ml, int32_t* __restrict__ leaf_at, int* __restrict__ err_depth,
                          int* __restrict__ max_leaf) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  if (node4[i].z >= 0) return;
  int t = tree_of[i];
  uint2 r = range[i];
  for (uint32_t p = r.x; p < r.y; ++p) leaf_at[size_t(t) * n + p] = i;
  if (t < TB && depth[i] < ml) atomicMin(err_depth, int(depth[i]));
  if (t < TB) atomicMax(max_leaf, int(r.y - r.x));
}

__device__ inline int32_t descend(const int4* __restrict__ node4, int32_t x, const float* __restrict__ q) {
  for (;;) {
    int4 nd = node4[x];
    if (nd.z < 0) return x;
    x = (q[nd.x] < __int_as_float(nd.y)) ? nd.z : nd.w;
  }
}

// descend over rows read through row_elem (FP32 or exact 16-bit rows)
template <int kRow>
__device__ inline int32_t descend_rows(const int4* __restrict__ node4, int32_t x, const void* __restrict__ X,
                                       size_t base) {
  for (;;) {
    int4 nd = node4[x];
    if (nd.z < 0) return x;
    x = (row_elem<kRow>(X, base + nd.x) < __int_as_float(nd.y)) ? nd.z : nd.w;
  }
}

// For each position: walk the ancestors at depth >= ml; where the position is
// in the smaller child, probe the larger one.  Pass 0 counts records per leaf
// reached and keeps the leaf of every (position, depth(X) - ml) slot in lam
// (-1: no probe there); pass 1 reads it back instead of descending again and
// writes (63 - depth(X)) << 32 | pos into the leaf's bucket.
template <int kRow>
__global__ void k_probe(const void* __restrict__ X, int D, const int4* __restrict__ node4,
                        const int32_t* __restrict__ parent, const uint2* __restrict__ range,
                        const int8_t* __restrict__ depth, const uint32_t* __restrict__ leaf_t,
                        const int32_t* __restrict__ leaf_at_t, uint32_t n, int ml, int pass,
                        uint32_t* __restrict__ cnt, const int32_t* __restrict__ boff,
                        uint32_t* __restrict__ cursor, uint64_t* __restrict__ rec, int32_t* __restrict__ lam_s,
                        int nlev) {
  uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= n) return;
  int32_t child = leaf_at_t[pos];
  int32_t x = parent[child];
  if (pass == 1) {
    while (x >= 0 && depth[x] >= ml) {
      const int dx = depth[x];
      const int32_t lam = lam_s[size_t(pos) * nlev + (dx - ml)];
      if (lam >= 0) {
        uint32_t at = atomicAdd(&cursor[lam], 1u);
        rec[size_t(boff[lam]) + at] = (uint64_t(63 - dx) << 32) | pos;
      }
      x = parent[x];
    }
    return;
  }
  const uint32_t p = leaf_t[pos];
  const size_t xb = size_t(p) * D;
  while (x >= 0 && depth[x] >= ml) {
    const int4 nd = node4[x];
    const uint2 rl = range[nd.z], rr = range[nd.w];
    const uint32_t numL = rl.y - rl.x, numR = rr.y - rr.x;
    const int32_t small = (numL < numR) ? nd.z : nd.w;
    int32_t lam = -1;
    if (child == small) {
      const int32_t big = (small == nd.z) ? nd.w : nd.z;
      lam = descend_rows<kRow>(node4, big, X, xb);
      atomicAdd(&cnt[lam], 1u);
    }
    lam_s[size_t(pos) * nlev + (depth[x] - ml)] = lam;
    child = x;
    x = parent[x];
  }
}

__device__ inline uint64_t shfl_xor_u64(uint64_t v, int m) {
  uint32_t lo = uint32_t(v), hi = uint32_t(v >> 32);
  lo = uint32_t(__shfl_xor(int(lo), m));
  hi = uint32_t(__shfl_xor(int(hi), m));
  return (uint64_t(hi) << 32) | lo;
}

// Sort every bucket ascending: one wavefront per bucket; up to 128 records
// in registers (bitonic network across lanes, two records per lane above
// 64), up to 1024 in LDS, longer ones by one lane.
__global__ void __launch_bounds__(kBlock) k_sort_buckets(const int32_t* __restrict__ boff, int M,
                                                         uint64_t* __restrict__ rec, int32_t* __restrict__ big,
                                                         int* __restrict__ nbig) {
  __shared__ uint64_t sh[kBlock / 64][1024];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int i = blockIdx.x * (kBlock / 64) + wv; i < M; i += gridDim.x * (kBlock / 64)) {
    const int b0 = boff[i], m = boff[i + 1] - b0;
    if (m <= 1) continue;
    if (m <= 64) {
      uint64_t v = lane < m ? rec[b0 + lane] : kKeyMax;
      for (int k = 2; k <= 64; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          uint64_t o = shfl_xor_u64(v, j);
          bool up = (lane & k) == 0, lower = (lane & j) == 0;
          uint64_t mn = v < o ? v : o, mx = v < o ? o : v;
          v = (lower == up) ? mn : mx;
        }
      }
      if (lane < m) rec[b0 + lane] = v;
    } else if (m <= 128) {  // two records per lane (e = lane + 64 h), registers only
      uint64_t v[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) v[h] = lane + 64 * h < m ? rec[b0 + lane + 64 * h] : kKeyMax;
      for (int k = 2; k <= 128; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          if (j == 64) {  // partner: the other record of this lane
            const bool up = (lane & k) == 0;
            const uint64_t mn = v[0] < v[1] ? v[0] : v[1], mx = v[0] < v[1] ? v[1] : v[0];
            v[0] = up ? mn : mx;
            v[1] = up ? mx : mn;
          } else {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const uint64_t o = shfl_xor_u64(v[h], j);
              const bool up = ((lane + 64 * h) & k) == 0, lower = (lane & j) == 0;
              const uint64_t mn = v[h] < o ? v[h] : o, mx = v[h] < o ? o : v[h];
              v[h] = (lower == up) ? mn : mx;
            }
          }
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (lane + 64 * h < m) rec[b0 + lane + 64 * h] = v[h];
    } else if (m <= 1024) {
      uint64_t* b = sh[wv];
      int p2 = next_pow2(m);
      for (int t = lane; t < p2; t += 64) b[t] = t < m ? rec[b0 + t] : kKeyMax;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      for (int k = 2; k <= p2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
          for (int t = lane; t < p2; t += 64) {
            int u = t ^ j;
            if (u > t) {
              uint64_t x = b[t], y = b[u];
              bool up = (t & k) == 0;
              if ((x > y) == up) {
                b[t] = y;
                b[u] = x;
              }
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
      }
      for (int t = lane; t < m; t += 64) rec[b0 + t] = b[t];
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    } else if (lane == 0) {
      big[atomicAdd(nbig, 1)] = i;  // sorted by k_sort_big
    }
  }
}

// Buckets above 1024 records: one workgroup each, bitonic sort in LDS (up to
// 8192 records); longer buckets are sorted by one thread.
__global__ void __launch_bounds__(kBlock) k_sort_big(const int32_t* __restrict__ boff, const int32_t* __restrict__ big,
                                                     const int* __restrict__ nbig, uint64_t* __restrict__ rec) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  uint64_t* b = reinterpret_cast<uint64_t*>(smem_raw);
  const int nb = *nbig;
  for (int w = blockIdx.x; w < nb; w += gridDim.x) {
    const int i = big[w];
    const int b0 = boff[i], m = boff[i + 1] - b0;
    if (m <= 8192) {
      const int p2 = next_pow2(m);
      for (int t = threadIdx.x; t < p2; t += kBlock) b[t] = t < m ? rec[b0 + t] : kKeyMax;
      __syncthreads();
      block_bitonic_sort(b, p2);
      for (int t = threadIdx.x; t < m; t += kBlock) rec[b0 + t] = b[t];
    } else if (threadIdx.x == 0) {
      for (int a = b0 + 1; a < b0 + m; ++a) {
        uint64_t v = rec[a];
        int j = a - 1;
        while (j >= b0 && rec[j] > v) {
          rec[j + 1] = rec[j];
          --j;
        }
        rec[j + 1] = v;
      }
    }
    __syncthreads();
  }
}

// bounded (distance, id) set of capacity S kept ascending in h[0..sz)
__device__ inline void heap_offer(uint64_t* h, int& sz, int S, uint32_t* extra, uint32_t& ecnt, uint32_t id,
                                  float dist) {
  if (sz < S || dist < key_dist(h[sz - 1])) {
    uint64_t k = make_key(dist, id);
    int i = sz;
    while (i > 0 && h[i - 1] > k) --i;
    if (i > 0 && h[i - 1] == k) return;  // already present
    int last = sz < S ? sz : S - 1;
    for (int j = last; j > i; --j) h[j] = h[j - 1];
    h[i] = k;
    if (sz < S) ++sz;
  } else if (ecnt < uint32_t(2 * S)) {
    extra[ecnt++] = id;
  }
}

__global__ void __launch_bounds__(kBlock) k_replay(const float* __restrict__ X, int D, const int4* __restrict__ node4,
                                                   const int32_t* __restrict__ parent, const uint2* __restrict__ range,
                                                   const int8_t* __restrict__ depth,
                                                   const uint32_t* __restrict__ leaf_t,
                                                   const int32_t* __restrict__ leaf_at_t, uint32_t n, int ml, int S,
                                                   const int32_t* __restrict__ boff, const uint64_t* __restrict__ rec,
                                                   uint64_t* __restrict__ hkey, uint32_t* __restrict__ hsize,
                                                   uint32_t* __restrict__ extra, uint32_t* __restrict__ ecount,
                                                   unsigned long long* __restrict__ ndist) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  uint64_t* h = reinterpret_cast<uint64_t*>(smem_raw) + size_t(threadIdx.x) * S;
  uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long nd_local = 0;
  if (pos < n) {
    const uint32_t p = leaf_t[pos];
    const float* xp = X + size_t(p) * D;
    int sz = int(hsize[p]);
    for (int j = 0; j < sz; ++j) h[j] = hkey[size_t(p) * S + j];
    uint32_t ecnt = ecount[p];
    uint32_t* ex = extra + size_t(p) * 2 * S;
    const int32_t lf = leaf_at_t[pos];
    int rb = boff[lf];
    const int re = boff[lf + 1];
    int32_t child = lf;
    int32_t x = parent[child];
    while (x >= 0 && depth[x] >= ml) {
      const int4 nd = node4[x];
      const uint2 rl = range[nd.z], rr = range[nd.w];
      const uint32_t numL = rl.y - rl.x, numR = rr.y - rr.x;
      const int32_t small = (numL < numR) ? nd.z : nd.w;
      if (child == small) {
        const int32_t big = (small == nd.z) ? nd.w : nd.z;
        const int32_t lam = descend(node4, big, xp);
        const uint2 r = range[lam];
        for (uint32_t i = r.x; i < r.y; ++i) {
          uint32_t q = leaf_t[i];
          float dist = l2_thread(X + size_t(q) * D, xp, D);
          ++nd_local;
          heap_offer(h, sz, S, ex, ecnt, q, dist);
        }
      } else {
        const uint64_t want = uint64_t(63 - depth[x]);
        while (rb < re && (rec[rb] >> 32) < want) ++rb;
        while (rb < re && (rec[rb] >> 32) == want) {
          uint32_t q = leaf_t[uint32_t(rec[rb] & 0xffffffffu)];
          float dist = l2_thread(xp, X + size_t(q) * D, D);
          heap_offer(h, sz, S, ex, ecnt, q, dist);
          ++rb;
        }
      }
      child = x;
      x = parent[x];
    }
    for (int j = 0; j < sz; ++j) hkey[size_t(p) * S + j] = h[j];
    hsize[p] = uint32_t(sz);
    ecount[p] = ecnt;
  }
  // per-block distance count
  __shared__ unsigned long long red[kBlock / 64];
  unsigned long long v = nd_local;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += red[w];
    atomicAdd(ndist, s);
  }
}

// Logical column l of the reference's accumulation order (the D % 32 tail
// first, then [0, D - D % 32)) -> physical column of a padded row.
__device__ inline int acc_col(int l, int D) {
  const int DR = D & 31;
  return l < DR ? D - DR + l : l - DR;
}

// Merge distances as leaf tiles: one wavefront per leaf bucket takes the
// bucket's records PB at a time (PB x members <= 256 pairs, four per lane)
// and, one chunk of DC logical columns after the other, stages the leaf's
// member rows and the batch's prober rows in LDS (row stride DC + 4); each
// lane keeps the reference's eight FMA chains for its pairs across the
// chunks.  It writes dist[record * LM + member] for every record of the
// bucket (sorted order), and pslot[pos * nlev + depth(X) - ml] = (record,
// leaf) for the prober side of the replay.  Squared L2 is symmetric bit for
// bit, so one value serves both endpoints.
__global__ void __launch_bounds__(kBlock) k_leaf_dist(const float* __restrict__ X, int D, int DC,
                                                      const uint2* __restrict__ range,
                                                      const uint32_t* __restrict__ leaf_t,
                                                      const int32_t* __restrict__ boff, int M,
                                                      const uint64_t* __restrict__ rec, int LM, int PB, int ml,
                                                      int nlev, float* __restrict__ dist, uint2* __restrict__ pslot) {
  extern __shared__ __align__(16) float4 tiles_s[];
  __shared__ uint32_t pids[kBlock / 64][64];  // prober ids of the batch (PB <= 64)
  const int RS4 = (DC + 4) >> 2, nch = (D + DC - 1) / DC;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  float4* rows_m = tiles_s + size_t(w) * (LM + PB) * RS4;
  float4* rows_p = rows_m + size_t(LM) * RS4;
  for (int l = blockIdx.x * nw + w; l < M; l += gridDim.x * nw) {
    const int b0 = boff[l], R = boff[l + 1] - b0;
    if (R == 0) continue;
    const uint2 r = range[l];
    const int m = int(r.y - r.x);
    for (int rb = 0; rb < R; rb += PB) {
      const int nb = min(PB, R - rb), cells = nb * m;
      if (lane < nb) {
        const uint64_t v = rec[b0 + rb + lane];
        const uint32_t pos = uint32_t(v & 0xffffffffu);
        const int dep = 63 - int(v >> 32);
        pids[w][lane] = leaf_t[pos];
        pslot[size_t(pos) * nlev + (dep - ml)] = make_uint2(uint32_t(b0 + rb + lane), uint32_t(l));
      }
      float acc[4][8];
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[j][e] = 0.f;
      for (int c = 0; c < nch; ++c) {
        const int l0 = c * DC, w4 = (min(DC, D - l0)) >> 2;
        wave_sync();
        for (int i = lane; i < (m + nb) * w4; i += 64) {
          const int rr = i / w4, c4 = i - rr * w4;
          const uint32_t id = rr < m ? leaf_t[r.x + rr] : pids[w][rr - m];
          rows_m[(rr < m ? rr : LM + rr - m) * RS4 + c4] =
              *reinterpret_cast<const float4*>(X + size_t(id) * D + acc_col(l0 + 4 * c4, D));
        }
        wave_sync();
        const int nb8 = w4 >> 1;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int t = lane + 64 * j;
          if (t < cells) {
            const int b = t / m, jj = t - b * m;
            const float4* x = rows_m + jj * RS4;
            const float4* y = rows_p + b * RS4;
#pragma unroll 1
            for (int e8 = 0; e8 < nb8; ++e8) {
              const float4 x0 = x[2 * e8], x1 = x[2 * e8 + 1], y0 = y[2 * e8], y1 = y[2 * e8 + 1];
              float d;
              d = x0.x - y0.x; acc[j][0] = __builtin_fmaf(d, d, acc[j][0]);
              d = x0.y - y0.y; acc[j][1] = __builtin_fmaf(d, d, acc[j][1]);
              d = x0.z - y0.z; acc[j][2] = __builtin_fmaf(d, d, acc[j][2]);
              d = x0.w - y0.w; acc[j][3] = __builtin_fmaf(d, d, acc[j][3]);
              d = x1.x - y1.x; acc[j][4] = __builtin_fmaf(d, d, acc[j][4]);
              d = x1.y - y1.y; acc[j][5] = __builtin_fmaf(d, d, acc[j][5]);
              d = x1.z - y1.z; acc[j][6] = __builtin_fmaf(d, d, acc[j][6]);
              d = x1.w - y1.w; acc[j][7] = __builtin_fmaf(d, d, acc[j][7]);
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int t = lane + 64 * j;
        if (t < cells) {
          const int b = t / m, jj = t - b * m;
          float sdist = acc[j][0];
#pragma unroll
          for (int e = 1; e < 8; ++e) sdist = sdist + acc[j][e];
          dist[size_t(b0 + rb + b) * LM + jj] = sdist;
        }
      }
      wave_sync();
    }
  }
}

// The replay of k_replay over the distances of k_leaf_dist.
__global__ void __launch_bounds__(kBlock) k_replay_dist(const int4* __restrict__ node4, const int32_t* __restrict__ parent,
                                                        const uint2* __restrict__ range, const int8_t* __restrict__ depth,
                                                        const uint32_t* __restrict__ leaf_t,
                                                        const int32_t* __restrict__ leaf_at_t, uint32_t n, int ml, int S,
                                                        const int32_t* __restrict__ boff,
                                                        const uint64_t* __restrict__ rec, const float* __restrict__ dist,
                                                        const uint2* __restrict__ pslot, int LM, int nlev,
                                                        uint64_t* __restrict__ hkey, uint32_t* __restrict__ hsize,
                                                        uint32_t* __restrict__ extra, uint32_t* __restrict__ ecount,
                                                        unsigned long long* __restrict__ ndist,
                                                        const int32_t* __restrict__ lam_s) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  uint64_t* h = reinterpret_cast<uint64_t*>(smem_raw) + size_t(threadIdx.x) * S;
  uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long nd_local = 0;
  if (pos < n) {
    const uint32_t p = leaf_t[pos];
    int sz = int(hsize[p]);
    for (int j = 0; j < sz; ++j) h[j] = hkey[size_t(p) * S + j];
    uint32_t ecnt = ecount[p];
    uint32_t* ex = extra + size_t(p) * 2 * S;
    const int32_t lf = leaf_at_t[pos];
    const int j0 = int(pos - range[lf].x);
    int rb = boff[lf];
    const int re = boff[lf + 1];
    // the ancestors X from the leaf's parent up to depth ml, deepest first;
    // the probe's counting pass recorded at each whether this position is in
    // X's smaller child (its landing leaf, else -1), so no node is read here
    for (int dx = int(depth[lf]) - 1; dx >= ml; --dx) {
      if (lam_s[size_t(pos) * nlev + (dx - ml)] >= 0) {
        const uint2 e = pslot[size_t(pos) * nlev + (dx - ml)];
        const uint2 r = range[e.y];
        const float* dr = dist + size_t(e.x) * LM;
        for (uint32_t i = r.x; i < r.y; ++i) {
          ++nd_local;
          heap_offer(h, sz, S, ex, ecnt, leaf_t[i], dr[i - r.x]);
        }
      } else {
        // this level's records of the leaf's bucket (sorted by level key,
        // then position): [lo, hi) by binary search from rb, then offered in
        // order, their loads issued eight at a time
        const uint64_t want = uint64_t(63 - dx);
        int lo = rb, hi = re;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if ((rec[mid] >> 32) < want) lo = mid + 1;
          else hi = mid;
        }
        int e1 = lo;
        hi = re;
        while (e1 < hi) {
          const int mid = (e1 + hi) >> 1;
          if ((rec[mid] >> 32) <= want) e1 = mid + 1;
          else hi = mid;
        }
        for (int r0 = lo; r0 < e1; r0 += 8) {
          uint32_t bid[8];
          float bd[8];
#pragma unroll
          for (int b = 0; b < 8; ++b) {
            const int r = r0 + b < e1 ? r0 + b : e1 - 1;
            bid[b] = leaf_t[uint32_t(rec[r] & 0xffffffffu)];
            bd[b] = dist[size_t(r) * LM + j0];
          }
#pragma unroll
          for (int b = 0; b < 8; ++b)
            if (r0 + b < e1) heap_offer(h, sz, S, ex, ecnt, bid[b], bd[b]);
        }
        rb = e1;
      }
    }
    for (int j = 0; j < sz; ++j) hkey[size_t(p) * S + j] = h[j];
    hsize[p] = uint32_t(sz);
    ecount[p] = ecnt;
  }
  __shared__ unsigned long long red[kBlock / 64];
  unsigned long long v = nd_local;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s2 = 0;
    for (int w = 0; w < kBlock / 64; ++w) s2 += red[w];
    atomicAdd(ndist, s2);
  }
}

// Random init (kdtreeub_index.hpp:1377-1423).
__global__ void k_init(const float* __restrict__ X, int D, uint32_t n, int S, int L, uint64_t seed,
                       uint64_t* __restrict__ hkey, uint32_t* __restrict__ hsize, const uint32_t* __restrict__ extra,
                       const uint32_t* __restrict__ ecount, uint64_t* __restrict__ pool, uint8_t* __restrict__ flag,
                       uint32_t* __restrict__ fill, uint32_t* __restrict__ rng, uint32_t* __restrict__ new_ids,
                       unsigned long long* __restrict__ ndist) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t* nw = new_ids + size_t(p) * 2 * S;
  const uint32_t S2 = 2u * S;
  if (ecount[p] < S2) {  // GenRandom(rng, nn_new, 2S, N) (1121-1135)
    for (uint32_t i = 0; i < S2; ++i) nw[i] = uint32_t(hash4(seed, ST_INIT_NEW, p, i) % uint64_t(n - S2));
    for (uint32_t i = 1; i < S2; ++i) {  // sort
      uint32_t v = nw[i];
      int j = int(i) - 1;
      while (j >= 0 && nw[j] > v) {
        nw[j + 1] = nw[j];
        --j;
      }
      nw[j + 1] = v;
    }
    for (uint32_t i = 1; i < S2; ++i)
      if (nw[i] <= nw[i - 1]) nw[i] = nw[i - 1] + 1;
    uint32_t off = uint32_t(hash4(seed, ST_INIT_NEW, p, S2) % uint64_t(n));
    for (uint32_t i = 0; i < S2; ++i) nw[i] = (nw[i] + off) % n;
  } else {
    for (uint32_t i = 0; i < S2; ++i) nw[i] = extra[size_t(p) * S2 + i];
  }
  uint64_t* h = hkey + size_t(p) * S;
  int sz = int(hsize[p]);
  uint64_t r = 0;
  unsigned long long cnt = 0;
  while (sz < S) {  // top up to S (1399-1405)
    uint32_t rid = uint32_t(hash4(seed, ST_INIT_TOP, p, r++) % uint64_t(n));
    float dist = l2_thread(X + size_t(p) * D, X + size_t(rid) * D, D);
    ++cnt;
    uint64_t k = make_key(dist, rid);
    int i = sz;
    while (i > 0 && h[i - 1] > k) --i;
    if (i > 0 && h[i - 1] == k) continue;
    for (int j = sz; j > i; --j) h[j] = h[j - 1];
    h[i] = k;
    ++sz;
  }
  for (int l = 0; l < L; ++l) {
    pool[size_t(p) * L + l] = l < S ? h[l] : kKeyMax;
    flag[size_t(p) * L + l] = l < S ? 1 : 0;
  }
  for (int l = 0; l < S; ++l) nw[l] = key_id(h[l]);
  fill[p] = uint32_t(S);
  rng[p] = uint32_t(S);
  if (cnt) atomicAdd(ndist, cnt);
}

// ---------------------------------------------------------------------------
// join
// ---------------------------------------------------------------------------
__global__ void k_worst(const uint64_t* __restrict__ pool, const uint32_t* __restrict__ fill, uint32_t n, int L,
                        uint64_t* __restrict__ worst) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  worst[p] = fill[p] < uint32_t(L) ? kKeyMax : pool[size_t(p) * L + L - 1];
}

// Bucket room for direct routing: the pairs each row is in, an upper bound
// of the candidates it can receive (one per pair).  A node's new row pairs
// with its other m - 1 rows, an old row with its na new ones; nodes without
// new rows join nothing.  A workgroup takes kRoomNodes consecutive nodes of
// `order` (tree 0's leaf list, or null: node order), sixteen lanes per node;
// nodes of one leaf share most of their list rows, so the room is summed per
// row in an LDS hash first (about a quarter as many global atomics, which
// are not waited for).  A row that finds no slot within kRoomProbe probes
// goes to the global sum directly.
constexpr int kRoomNodes = 256, kRoomSlots = 4096, kRoomProbe = 32;
constexpr uint32_t kRoomFree = 0xffffffffu;  // an empty LDS slot
__global__ void __launch_bounds__(kBlock) k_bucket_room(uint32_t n, const uint32_t* __restrict__ order,
                                                        const uint64_t* __restrict__ new_off,
                                                        const uint32_t* __restrict__ new_ids,
                                                        const uint64_t* __restrict__ old_off,
                                                        const uint32_t* __restrict__ old_ids,
                                                        uint32_t* __restrict__ room) {
  __shared__ uint32_t key_s[kRoomSlots], val_s[kRoomSlots];
  const int tid = threadIdx.x;
  for (int i = tid; i < kRoomSlots; i += kBlock) {
    key_s[i] = kRoomFree;
    val_s[i] = 0u;
  }
  __syncthreads();
  auto add = [&](uint32_t t, uint32_t w) {
    uint32_t h = (t * 2654435761u) >> 20;  // 12 bits: kRoomSlots
    for (int k = 0; k < kRoomProbe; ++k, h = (h + 1) & uint32_t(kRoomSlots - 1)) {
      const uint32_t prev = atomicCAS(&key_s[h], kRoomFree, t);
      if (prev == kRoomFree || prev == t) {
        atomicAdd(&val_s[h], w);
        return;
      }
    }
    atomicAdd(&room[t], w);
  };
  const uint32_t i0 = blockIdx.x * uint32_t(kRoomNodes);
  const int l = tid & 15;
  for (int j = tid >> 4; j < kRoomNodes; j += kBlock / 16) {
    const uint32_t i = i0 + uint32_t(j);
    if (i >= n) break;
    const uint32_t u = order ? order[i] : i;
    const uint64_t a0 = new_off[u], b0 = old_off[u];
    const uint32_t na = uint32_t(new_off[u + 1] - a0), nb = uint32_t(old_off[u + 1] - b0);
    if (na == 0) continue;
    if (na + nb > 1)
      for (uint32_t e = uint32_t(l); e < na; e += 16) add(new_ids[a0 + e], na + nb - 1);
    for (uint32_t e = uint32_t(l); e < nb; e += 16) add(old_ids[b0 + e], na);
  }
  __syncthreads();
  for (int i = tid; i < kRoomSlots; i += kBlock)
    if (key_s[i] != kRoomFree) atomicAdd(&room[key_s[i]], val_s[i]);
}

// A kept candidate staged for its target's bucket.
struct Stage {
  uint32_t target;  // kNoRow: an unused staging slot
  uint32_t src;     // joining node | kFirst
  uint64_t key;
};

// Join tasks.  The rows of a joining node are its new list followed by its
// old list, cut into blocks of kJBR rows.  Every pair the reference measures
// is a pair of positions p < q with p new (new x new pairs once per unordered
// position pair, new x old pairs always have p < q), so task (node, I, J),
// I <= J, I a block holding new rows, covers all pairs with p in block I and
// q in block J.
constexpr int kJBR = 32;
constexpr uint32_t kNoRow = 0xffffffffu;
__device__ inline uint32_t join_task_count(uint32_t na, uint32_t nb) {
  if (na == 0) return 0;
  const uint32_t nI = (na + kJBR - 1) / kJBR, nJ = (na + nb + kJBR - 1) / kJBR;
  return nI * nJ - nI * (nI - 1) / 2;
}

// Tasks per node, the nodes taken in the order of tree 0's leaf list
// (order[i] = node at position i; node i when order is null), so that nodes of one leaf, whose lists
// share rows, get neighbouring task ranges and run at about the same time on
// one XCD; nodes with at most node_rows rows get none (the whole-node
// kernel joins them).  sums[0] accumulates the pairs (na (na - 1) / 2 +
// na nb per node, two candidates each at most), sums[2] the rows joined
// (|new| + |old| over the nodes with a non-empty new list): a grid-stride
// loop, one atomic per workgroup and sum (atomics on one address from every
// wavefront of a million-thread grid serialise).
__global__ void __launch_bounds__(kBlock) k_join_count(uint32_t n, const uint32_t* __restrict__ order,
                                                       const uint64_t* __restrict__ new_off,
                                                       const uint64_t* __restrict__ old_off, uint32_t node_rows,
                                                       int32_t* __restrict__ cnt,
                                                       unsigned long long* __restrict__ sums) {
  __shared__ unsigned long long red_s[2][kBlock / 64];
  unsigned long long r = 0, v = 0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const uint32_t p = order ? order[i] : i;
    const uint32_t na = uint32_t(new_off[p + 1] - new_off[p]), nb = uint32_t(old_off[p + 1] - old_off[p]);
    cnt[i] = na + nb <= node_rows ? 0 : int32_t(join_task_count(na, nb));
    r += na ? na + nb : 0;
    v += (unsigned long long)na * (na > 0 ? na - 1 : 0) / 2 + (unsigned long long)na * nb;
  }
  for (int o = 32; o > 0; o >>= 1) {
    r += (unsigned long long)__shfl_down((long long)r, o);
    v += (unsigned long long)__shfl_down((long long)v, o);
  }
  if ((threadIdx.x & 63) == 0) {
    red_s[0][threadIdx.x >> 6] = v;
    red_s[1][threadIdx.x >> 6] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sv = 0, sr = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      sv += red_s[0][w];
      sr += red_s[1][w];
    }
    if (sv) atomicAdd(sums, sv);
    if (sr) atomicAdd(sums + 2, sr);
  }
}

// Task records of the nodes with tasks (position i of the order, tcnt[i] >
// 0: at least one new row and more than node_rows rows), one wavefront per
// node: header (node, sizeA | newA << 8 | sizeB << 16 | newB << 24, I == J)
// and the rows' ids, slots [0, 32) block I, [32, 64) block J (unused when
// I == J), kNoRow past the blocks' ends.  Each wavefront reads the task
// counts of 64 positions at a time and takes the nodes that have tasks from
// a ballot (few nodes have them).
__global__ void __launch_bounds__(kBlock) k_join_tasks(uint32_t n, const uint32_t* __restrict__ order,
                                                       const uint64_t* __restrict__ new_off,
                                                       const uint32_t* __restrict__ new_ids,
                                                       const uint64_t* __restrict__ old_off,
                                                       const uint32_t* __restrict__ old_ids,
                                                       const int32_t* __restrict__ tcnt,
                                                       const int32_t* __restrict__ toff, uint4* __restrict__ hdr,
                                                       uint32_t* __restrict__ tids) {
  const int lane = threadIdx.x & 63;
  const uint32_t w0 = (blockIdx.x * kBlock + threadIdx.x) >> 6, nw = gridDim.x * (kBlock / 64);
  for (uint32_t c0 = w0 * 64; c0 < n; c0 += nw * 64) {
    uint64_t todo = __ballot(c0 + lane < n && tcnt[c0 + lane] > 0);
    while (todo) {
      const uint32_t i = c0 + uint32_t(__ffsll((long long)todo) - 1);
      todo &= todo - 1;
      const uint32_t p = order ? order[i] : i;
      const uint64_t a0 = new_off[p], b0 = old_off[p];
      const int na = int(new_off[p + 1] - a0), m = na + int(old_off[p + 1] - b0);
      const int nI = (na + kJBR - 1) / kJBR, nJ = (m + kJBR - 1) / kJBR;
      size_t t = size_t(toff[i]);
      for (int I = 0; I < nI; ++I)
        for (int J = I; J < nJ; ++J, ++t) {
          const int sizeA = min(kJBR, m - I * kJBR), newA = min(kJBR, na - I * kJBR);
          const int sizeB = min(kJBR, m - J * kJBR), newB = max(0, min(sizeB, na - J * kJBR));
          const int side = lane >> 5, r = lane & 31;
          const int size = side ? sizeB : sizeA, pos = (side ? J : I) * kJBR + r;
          uint32_t id = kNoRow;
          if (r < size && !(side && I == J)) id = pos < na ? new_ids[a0 + pos] : old_ids[b0 + pos - na];
          tids[t * 64 + lane] = id;
          if (lane == 0)
            hdr[t] = make_uint4(p,
                                uint32_t(sizeA) | uint32_t(newA) << 8 | uint32_t(sizeB) << 16 | uint32_t(newB) << 24,
                                I == J ? 1u : 0u, 0u);
        }
    }
  }
}

// Pair k of a task -> positions (p, q) in blocks I and J.  I != J: all
// newA x sizeB; I == J: q > p, row p holding sizeA - 1 - p pairs.
__device__ inline void join_cell(int k, bool same, int sizeA, int sizeB, int& p, int& q) {
  if (!same) {
    p = k / sizeB;
    q = k - p * sizeB;
    return;
  }
  // largest p with start(p) = p (sizeA - 1) - p (p - 1) / 2 <= k
  int lo = 0, hi = sizeA - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mid * (sizeA - 1) - mid * (mid - 1) / 2 <= k) lo = mid;
    else hi = mid - 1;
  }
  p = lo;
  q = p + 1 + (k - (p * (sizeA - 1) - p * (p - 1) / 2));
}

// Endpoints of the pair at positions (p, q) of blocks I and J: slots and
// ids, the first endpoint (whose candidate carries the kFirst credit tag)
// being the lower id of a new x new pair.  False for a pair of one id: the
// reference skips it in new x new (base_index.hpp:99).  In new x old it
// cannot arise under a symmetric distance (an id within a node's Range is
// not beyond its radiusM, so it is not also on the node's reverse list), but
// under the precision form's approximate keys d(x, y) and d(y, x) may differ
// by a rounding, and the pair would put x into its own pool at distance 0.
__device__ inline bool pair_ends(int p, int q, bool same, int newA, int newB, const uint32_t* ids_s, int& sa,
                                 int& sb, uint32_t& a, uint32_t& b) {
  sa = p;
  sb = same ? q : 32 + q;
  a = ids_s[sa];
  b = ids_s[sb];
  if (a == b) return false;
  if (q < (same ? newA : newB)) {  // new x new: once, first endpoint = lower id
    if (b < a) {
      const uint32_t t2 = a;
      a = b;
      b = t2;
      const int s2 = sa;
      sa = sb;
      sb = s2;
    }
  }
  return true;
}
// Pair k of a task (join_cell order).
__device__ inline bool join_ends(int k, bool same, int sizeA, int newA, int sizeB, int newB,
                                 const uint32_t* ids_s, int& sa, int& sb, uint32_t& a, uint32_t& b) {
  int p, q;
  join_cell(k, same, sizeA, sizeB, p, q);
  return pair_ends(p, q, same, newA, newB, ids_s, sa, sb, a, b);
}
// Pair (p, q) of a full 32 x 32 block, if the reference measures it.
__device__ inline bool block_ends(int p, int q, bool same, int sizeA, int newA, int sizeB, int newB,
                                  const uint32_t* ids_s, int& sa, int& sb, uint32_t& a, uint32_t& b) {
  if (p >= newA || q >= (same ? sizeA : sizeB) || (same && q <= p)) return false;
  return pair_ends(p, q, same, newA, newB, ids_s, sa, sb, a, b);
}

// Positions P < Q of a node's rows (new list, then old list): the pair the
// reference measures when P is new; slots and ids of the two endpoints, the
// first endpoint being the lower id of a new x new pair (as pair_ends, and
// like it false for a pair of one id).
__device__ inline bool node_ends(int P, int Q, int na, int m, const uint32_t* ids, int& sa, int& sb, uint32_t& a,
                                 uint32_t& b) {
  if (P >= na || Q >= m || Q <= P) return false;
  sa = P;
  sb = Q;
  a = ids[P];
  b = ids[Q];
  if (a == b) return false;
  if (Q < na) {
    if (b < a) {
      const uint32_t t2 = a;
      a = b;
      b = t2;
      sa = Q;
      sb = P;
    }
  }
  return true;
}

// Load k of this thread for the rows of one chunk (DC logical columns of up
// to 64 rows, float4 i = threadIdx.x + k * kBlock is row i / (DC / 4)).
__device__ __forceinline__ bool join_slot(int k, int DC4, int w4, int sizeA, int sizeB, bool same, int& r, int& c4) {
  const int i = threadIdx.x + k * kBlock;
  r = i / DC4;
  c4 = i - r * DC4;
  return r < 64 && c4 < w4 && ((r < 32) ? r < sizeA : (!same && r - 32 < sizeB));
}
__device__ __forceinline__ float4 join_load(const float* __restrict__ X, int D, int DC, const uint4& h,
                                            const uint32_t* ids, int c, int k) {
  const int sizeA = int(h.y & 0xff), sizeB = int((h.y >> 16) & 0xff), l0 = c * DC;
  int r, c4;
  const bool in = join_slot(k, DC >> 2, (min(DC, D - l0)) >> 2, sizeA, sizeB, h.z != 0, r, c4);
  // out-of-range slots load row 0, column 0 (never stored)
  const uint32_t id = in ? ids[r] : 0u;
  return *reinterpret_cast<const float4*>(X + size_t(id) * D + (in ? acc_col(l0 + 4 * c4, D) : 0));
}

// Staging slots are reserved per workgroup in runs of at least kStageRun, so
// the shared counter sees one atomic per run; a task's candidates go to
// consecutive slots, grouped by target row.  The slots a workgroup leaves
// unused at its end get target kNoRow.
constexpr unsigned long long kStageRun = 4096;

// One workgroup walks a sequence of (task, chunk) steps.  The two blocks'
// rows of one chunk of DC = 4 kJPF logical columns... sit in LDS (row stride
// DC + 4); one thread per pair keeps the reference's eight FMA chains for up
// to four pairs.  The next step's rows are loaded into registers while this
// step computes, from row ids staged in LDS one task ahead (their global
// loads are issued two tasks ahead).  Tasks are split into eight contiguous
// ranges, one per XCD (the workgroup id modulo 8), so a node's tasks share
// that XCD's L2.  After a task's last chunk its kept candidates (key below
// the target's worst) are counted per target row in LDS; the counts go to the
// targets' totals with one atomic per (task, target row) whose result is not
// waited for, and the candidates to the workgroup's staging run.
constexpr int kJPairs = kJBR * kJBR / kBlock;  // pairs per thread at most

// Direct routing (kDirect): instead of staging, a row's kept candidates take
// consecutive slots of its target's bucket, which starts at boff[target]
// (room for every pair the target is in); the row's first slot is the
// bucket's fill before it, an atomic on tcount[target].
__device__ __forceinline__ void put_direct(uint64_t* __restrict__ bkey, uint32_t* __restrict__ bsrc, uint32_t o,
                                           uint64_t key, uint32_t src) {
  bkey[o] = key;
  bsrc[o] = src;
}
// Pairs measured, and with direct routing the candidates kept (the staging
// counter is free then): one atomic per workgroup each, at its end.
template <bool kDirect>
__device__ __forceinline__ void join_totals(unsigned long long pairs, unsigned long long kept,
                                            unsigned long long* __restrict__ npairs,
                                            unsigned long long* __restrict__ nkept) {
  __shared__ unsigned long long red_s[2][kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) {
    pairs += (unsigned long long)__shfl_down((long long)pairs, o);
    kept += (unsigned long long)__shfl_down((long long)kept, o);
  }
  if ((threadIdx.x & 63) == 0) {
    red_s[0][threadIdx.x >> 6] = pairs;
    red_s[1][threadIdx.x >> 6] = kept;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sp = 0, sk = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      sp += red_s[0][w];
      sk += red_s[1][w];
    }
    if (sp) atomicAdd(npairs, sp);
    if (kDirect && sk) atomicAdd(nkept, sk);
  }
}

template <int kJPF, bool kDirect>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_join_tiles(const float* __restrict__ X, int D, uint32_t ntask, const uint4* __restrict__ hdr,
             const uint32_t* __restrict__ tids, const uint64_t* __restrict__ worst, Stage* __restrict__ stage,
             unsigned long long cap, unsigned long long* __restrict__ nstage, uint32_t* __restrict__ tcount,
             unsigned long long* __restrict__ npairs, const int32_t* __restrict__ boff,
             uint64_t* __restrict__ bkey, uint32_t* __restrict__ bsrc) {
  constexpr int kDC = kJPF * 16;  // columns per chunk at most
  extern __shared__ __align__(16) float4 rows_s[];  // [64][(DC + 4) / 4]
  __shared__ uint32_t ids_s[2][64];
  __shared__ uint4 hdr_s[2];
  __shared__ uint64_t w_s[64];
  __shared__ uint32_t cnt_s[64], off_s[64];
  __shared__ unsigned long long run_s[2];  // this task's slots: i < run_s[1] at run_s[0] + i, else past a new run
  __shared__ unsigned long long newrun_s;
  const int tid = threadIdx.x, lane = tid & 63;
  const int DC = min(D, kDC), RS4 = (DC + 4) >> 2, DC4 = DC >> 2;
  const int nch = (D + DC - 1) / DC;
  const uint32_t xcd = blockIdx.x & 7u, ns = gridDim.x >> 3;
  const uint32_t lo = uint32_t(uint64_t(ntask) * xcd / 8), hi = uint32_t(uint64_t(ntask) * (xcd + 1) / 8);
  unsigned long long pairs_local = 0, kept_local = 0;
  unsigned long long wbase = 0, wend = 0;  // staging run (wave 0, lane 63)
  uint32_t t = lo + blockIdx.x / 8;
  // prologue: ids of the first task in LDS, of the second in registers
  uint32_t reg_id = kNoRow;
  uint4 reg_hdr = make_uint4(0, 0, 0, 0);
  if (t < hi && tid < 64) {
    ids_s[0][tid] = tids[size_t(t) * 64 + tid];
    if (tid == 0) hdr_s[0] = hdr[t];
    if (t + ns < hi) {
      reg_id = tids[size_t(t + ns) * 64 + tid];
      reg_hdr = hdr[t + ns];
    }
  }
  __syncthreads();
  float4 pf[kJPF];
  if (t < hi) {
    const uint4 h0 = hdr_s[0];
#pragma unroll
    for (int k = 0; k < kJPF; ++k) pf[k] = join_load(X, D, DC, h0, ids_s[0], 0, k);
  }
  float acc[kJPairs][8];
  int cur = 0, c = 0;
  while (t < hi) {
    const uint4 h = hdr_s[cur];
    const bool same = h.z != 0;
    const int sizeA = int(h.y & 0xff), newA = int((h.y >> 8) & 0xff), sizeB = int((h.y >> 16) & 0xff),
              newB = int(h.y >> 24);
    if (c == 0 && tid < 64) {
      const uint32_t id = ids_s[cur][tid];
      w_s[tid] = id != kNoRow ? worst[id] : 0ull;
      cnt_s[tid] = 0;
      // the next task's ids into LDS, the one after into registers
      ids_s[cur ^ 1][tid] = reg_id;
      if (tid == 0) hdr_s[cur ^ 1] = reg_hdr;
      const uint32_t t2 = t + 2 * ns;
      if (t2 < hi && t2 > t) {
        reg_id = tids[size_t(t2) * 64 + tid];
        reg_hdr = hdr[t2];
      }
    }
    {
      const int w4 = (min(DC, D - c * DC)) >> 2;
#pragma unroll
      for (int k = 0; k < kJPF; ++k) {
        int r, c4;
        if (join_slot(k, DC4, w4, sizeA, sizeB, same, r, c4)) rows_s[r * RS4 + c4] = pf[k];
      }
    }
    __syncthreads();
    // next step's rows
    const bool last = c + 1 == nch;
    const uint32_t tn = last ? t + ns : t;
    const int cn = last ? 0 : c + 1, nb = last ? cur ^ 1 : cur;
    if (tn < hi && tn >= t) {
      const uint4 hn = hdr_s[nb];
#pragma unroll
      for (int k = 0; k < kJPF; ++k) pf[k] = join_load(X, D, DC, hn, ids_s[nb], cn, k);
    }
    const int cells = same ? newA * sizeA - newA * (newA + 1) / 2 : newA * sizeB;
    if (c == 0) {
#pragma unroll
      for (int j = 0; j < kJPairs; ++j)
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[j][e] = 0.f;
    }
    const int nb8 = (min(DC, D - c * DC)) >> 3;
    const float4* rowsB = rows_s + (same ? 0 : 32 * RS4);
#pragma unroll
    for (int j = 0; j < kJPairs; ++j) {
      const int k = tid + j * kBlock;
      if (k < cells) {
        int p, q;
        join_cell(k, same, sizeA, sizeB, p, q);
        const float4* a = rows_s + p * RS4;
        const float4* b = rowsB + q * RS4;
#pragma unroll 1
        for (int e8 = 0; e8 < nb8; ++e8) {
          const float4 x0 = a[2 * e8], x1 = a[2 * e8 + 1];
          const float4 y0 = b[2 * e8], y1 = b[2 * e8 + 1];
          float d;
          d = x0.x - y0.x; acc[j][0] = __builtin_fmaf(d, d, acc[j][0]);
          d = x0.y - y0.y; acc[j][1] = __builtin_fmaf(d, d, acc[j][1]);
          d = x0.z - y0.z; acc[j][2] = __builtin_fmaf(d, d, acc[j][2]);
          d = x0.w - y0.w; acc[j][3] = __builtin_fmaf(d, d, acc[j][3]);
          d = x1.x - y1.x; acc[j][4] = __builtin_fmaf(d, d, acc[j][4]);
          d = x1.y - y1.y; acc[j][5] = __builtin_fmaf(d, d, acc[j][5]);
          d = x1.z - y1.z; acc[j][6] = __builtin_fmaf(d, d, acc[j][6]);
          d = x1.w - y1.w; acc[j][7] = __builtin_fmaf(d, d, acc[j][7]);
        }
      }
    }
    if (last) {
      // candidates, up to two per pair: the final distance and the
      // candidate's index among its target row's candidates (LDS count)
      const uint32_t* ids = ids_s[cur];
      float dist[kJPairs];
      uint32_t li[kJPairs];  // first endpoint's index | second's << 16; 0xffff = none
#pragma unroll
      for (int j = 0; j < kJPairs; ++j) {
        li[j] = 0xffffffffu;
        float sdist = acc[j][0];
#pragma unroll
        for (int e = 1; e < 8; ++e) sdist = sdist + acc[j][e];
        dist[j] = sdist;
        const int k = tid + j * kBlock;
        int sa, sb;
        uint32_t a, b;
        if (k < cells && join_ends(k, same, sizeA, newA, sizeB, newB, ids, sa, sb, a, b)) {
          ++pairs_local;
          uint32_t la = 0xffffu, lb = 0xffffu;
          if (make_key(sdist, b) < w_s[sa]) la = atomicAdd(&cnt_s[sa], 1u);
          if (make_key(sdist, a) < w_s[sb]) lb = atomicAdd(&cnt_s[sb], 1u);
          li[j] = la | lb << 16;
        }
      }
      __syncthreads();
      if (tid < 64 && kDirect) {
        // the row's first slot in its target's bucket
        const uint32_t cnt = cnt_s[tid];
        off_s[tid] = cnt ? uint32_t(boff[ids[tid]]) + atomicAdd(&tcount[ids[tid]], cnt) : 0u;
        kept_local += cnt;
      } else if (tid < 64) {
        // totals per target (result not waited for), offsets within the task
        const uint32_t cnt = cnt_s[tid];
        if (cnt) atomicAdd(&tcount[ids[tid]], cnt);
        uint32_t x = cnt;
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t y = uint32_t(__shfl_up(int(x), o));
          if (lane >= o) x += y;
        }
        off_s[tid] = x - cnt;
        // this task's slots: the rest of the run, then a new run if needed
        const unsigned long long tot = x, rem = wend - wbase;
        unsigned long long nr = 0;
        const bool grow = tot > rem;
        if (grow && tid == 63) nr = atomicAdd(nstage, tot - rem > kStageRun ? tot - rem : kStageRun);
        if (tid == 63) {
          run_s[0] = wbase;
          run_s[1] = rem;
          newrun_s = nr;
          if (grow) {
            wbase = nr + (tot - rem);
            wend = nr + (tot - rem > kStageRun ? tot - rem : kStageRun);
          } else {
            wbase += tot;
          }
        }
      }
      __syncthreads();
      const unsigned long long r0 = run_s[0], rem = run_s[1], nr = newrun_s;
      const uint32_t nn = h.x;
#pragma unroll
      for (int j = 0; j < kJPairs; ++j) {
        if (li[j] == 0xffffffffu) continue;
        int sa, sb;
        uint32_t a, b;
        join_ends(tid + j * kBlock, same, sizeA, newA, sizeB, newB, ids, sa, sb, a, b);
        const uint32_t la = li[j] & 0xffffu, lb = li[j] >> 16;
        if (kDirect) {
          if (la != 0xffffu) put_direct(bkey, bsrc, off_s[sa] + la, make_key(dist[j], b), nn | kFirst);
          if (lb != 0xffffu) put_direct(bkey, bsrc, off_s[sb] + lb, make_key(dist[j], a), nn);
          continue;
        }
        if (la != 0xffffu) {
          const unsigned long long i = off_s[sa] + la, at = i < rem ? r0 + i : nr + (i - rem);
          if (at < cap) stage[at] = Stage{a, nn | kFirst, make_key(dist[j], b)};
        }
        if (lb != 0xffffu) {
          const unsigned long long i = off_s[sb] + lb, at = i < rem ? r0 + i : nr + (i - rem);
          if (at < cap) stage[at] = Stage{b, nn, make_key(dist[j], a)};
        }
      }
    }
    __syncthreads();
    if (last) cur ^= 1;
    t = tn;
    c = cn;
  }
  // unused rest of the staging run
  if (tid < 64 && !kDirect) {
    const unsigned long long from = (unsigned long long)__shfl((long long)wbase, 63),
                             to = (unsigned long long)__shfl((long long)wend, 63);
    for (unsigned long long i = from + lane; i < to && i < cap; i += 64) stage[i].target = kNoRow;
  }
  join_totals<kDirect>(pairs_local, kept_local, npairs, nstage);
}

// Whole-node join in FP32 for the nodes with at most 32 rows: one wavefront
// per node (the block tasks give such a node a whole workgroup and leave
// most of its threads idle).  The node's rows go through the wavefront's own
// LDS slice one chunk of DC logical columns at a time (the next chunk loaded
// into registers meanwhile); each lane keeps the reference's eight FMA
// chains for four pairs (P < Q, P new, in join_cell order) at a time, a
// second pass over the chunks taking pairs 256 .. 495, so every distance
// equals the block tasks' and the reference's.  Nodes are taken in the
// order of tree 0's leaf list (order, null: node order), one contiguous
// range per XCD; candidates are counted, offset and staged as in
// k_join_node_mfma.
constexpr int kNodeRows = 32;
constexpr int kNodePairs = 8;  // per lane: ceil(32 * 31 / 2 / 64)
constexpr int kNodeDC = 64;
constexpr int kNodeRS = kNodeDC + 4;  // LDS row stride (floats)
constexpr int kNodeLd = kNodeRows * kNodeDC / 4 / 64;  // float4 loads per lane and chunk (8)
__device__ __forceinline__ float4 node_load(const float* __restrict__ X, int D, const uint32_t* ids, int m, int l0,
                                            int w4, int k) {
  const int i = (threadIdx.x & 63) + 64 * k, r = i / (kNodeDC / 4), c4 = i - r * (kNodeDC / 4);
  const bool in = r < m && c4 < w4;
  // out-of-range slots load row 0, column 0 (never stored)
  return *reinterpret_cast<const float4*>(X + size_t(in ? ids[r] : 0u) * D + (in ? acc_col(l0 + 4 * c4, D) : 0));
}
// Distances of this lane's pairs k0 + lane + 64 j, j < 4 (k < cells).
__device__ __forceinline__ void node_pairs4(const float* __restrict__ X, int D, const uint32_t* ids, int m, int cells,
                                            int k0, float* rows, const int* pq, float* out) {
  const int lane = threadIdx.x & 63;
  const int nch = (D + kNodeDC - 1) / kNodeDC;
  float acc[4][8];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[j][e] = 0.f;
  float4 pf[kNodeLd];
#pragma unroll
  for (int k = 0; k < kNodeLd; ++k) pf[k] = node_load(X, D, ids, m, 0, min(kNodeDC, D) >> 2, k);
#pragma unroll 1
  for (int c = 0; c < nch; ++c) {
    const int w4 = min(kNodeDC, D - c * kNodeDC) >> 2;
#pragma unroll
    for (int k = 0; k < kNodeLd; ++k) {
      const int ii = lane + 64 * k, r = ii / (kNodeDC / 4), c4 = ii - r * (kNodeDC / 4);
      if (r < m && c4 < w4) *reinterpret_cast<float4*>(rows + r * kNodeRS + 4 * c4) = pf[k];
    }
    wave_sync();
    if (c + 1 < nch) {
      const int w4n = min(kNodeDC, D - (c + 1) * kNodeDC) >> 2;
#pragma unroll
      for (int k = 0; k < kNodeLd; ++k) pf[k] = node_load(X, D, ids, m, (c + 1) * kNodeDC, w4n, k);
    }
    const int nb8 = w4 >> 1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (k0 + lane + 64 * j >= cells) continue;
      const float4* a = reinterpret_cast<const float4*>(rows + (pq[j] & 0xff) * kNodeRS);
      const float4* b = reinterpret_cast<const float4*>(rows + (pq[j] >> 8) * kNodeRS);
#pragma unroll 1
      for (int e8 = 0; e8 < nb8; ++e8) {
        const float4 x0 = a[2 * e8], x1 = a[2 * e8 + 1];
        const float4 y0 = b[2 * e8], y1 = b[2 * e8 + 1];
        float d;
        d = x0.x - y0.x; acc[j][0] = __builtin_fmaf(d, d, acc[j][0]);
        d = x0.y - y0.y; acc[j][1] = __builtin_fmaf(d, d, acc[j][1]);
        d = x0.z - y0.z; acc[j][2] = __builtin_fmaf(d, d, acc[j][2]);
        d = x0.w - y0.w; acc[j][3] = __builtin_fmaf(d, d, acc[j][3]);
        d = x1.x - y1.x; acc[j][4] = __builtin_fmaf(d, d, acc[j][4]);
        d = x1.y - y1.y; acc[j][5] = __builtin_fmaf(d, d, acc[j][5]);
        d = x1.z - y1.z; acc[j][6] = __builtin_fmaf(d, d, acc[j][6]);
        d = x1.w - y1.w; acc[j][7] = __builtin_fmaf(d, d, acc[j][7]);
      }
    }
    wave_sync();  // the slice is rewritten by the next chunk
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float sdist = acc[j][0];
#pragma unroll
    for (int e = 1; e < 8; ++e) sdist = sdist + acc[j][e];
    out[j] = sdist;
  }
}
template <bool kDirect>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_join_node_f32(const float* __restrict__ X, int D, uint32_t n, const uint32_t* __restrict__ order,
                const uint64_t* __restrict__ new_off, const uint32_t* __restrict__ new_ids,
                const uint64_t* __restrict__ old_off, const uint32_t* __restrict__ old_ids,
                const uint64_t* __restrict__ worst, Stage* __restrict__ stage, unsigned long long cap,
                unsigned long long* __restrict__ nstage, uint32_t* __restrict__ tcount,
                unsigned long long* __restrict__ npairs, const int32_t* __restrict__ boff,
                uint64_t* __restrict__ bkey, uint32_t* __restrict__ bsrc) {
  constexpr int kW = kBlock / 64;
  __shared__ __align__(16) float rows_s[kW][kNodeRows * kNodeRS];
  __shared__ uint32_t ids_s[kW][64];
  __shared__ uint64_t w_s[kW][64];
  __shared__ uint32_t cnt_s[kW][64];
  __shared__ uint32_t off_s[kW][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t xcd = blockIdx.x & 7u, nsl = (gridDim.x >> 3) * kW;
  const uint32_t lo = uint32_t(uint64_t(n) * xcd / 8), hi = uint32_t(uint64_t(n) * (xcd + 1) / 8);
  unsigned long long pairs_local = 0, kept_local = 0;
  unsigned long long wbase = 0, wend = 0;  // staging run
  uint32_t* ids = ids_s[wv];
  float* rows = rows_s[wv];
  // the wavefront's nodes i0, i0 + nsl, ... are read 64 at a time, one per
  // lane; the nodes taken here are then worked through in order
  for (uint32_t i0 = lo + (blockIdx.x >> 3) * kW + wv; i0 < hi; i0 += 64 * nsl) {
    const uint32_t ij = i0 + uint32_t(lane) * nsl;
    uint32_t ndj = 0;
    uint64_t a0j = 0, b0j = 0;
    int naj = 0, mj = 0;
    if (ij < hi) {
      ndj = order ? order[ij] : ij;
      a0j = new_off[ndj];
      b0j = old_off[ndj];
      naj = int(new_off[ndj + 1] - a0j);
      mj = naj + int(old_off[ndj + 1] - b0j);
    }
    // no pairs / the block tasks' nodes
    uint64_t todo = __ballot(ij < hi && naj > 0 && mj <= kNodeRows);
    while (todo) {
      const int jn = __ffsll((long long)todo) - 1;
      todo &= todo - 1;
      const uint32_t nd = uint32_t(__shfl(int(ndj), jn));
      const uint64_t a0 = uint64_t(__shfl((long long)a0j, jn)), b0 = uint64_t(__shfl((long long)b0j, jn));
      const int na = __shfl(naj, jn), m = __shfl(mj, jn);
      const uint32_t id = lane < m ? (lane < na ? new_ids[a0 + lane] : old_ids[b0 + lane - na]) : kNoRow;
      ids[lane] = id;
      w_s[wv][lane] = id != kNoRow ? worst[id] : 0ull;
      cnt_s[wv][lane] = 0;
      wave_sync();
      const int cells = na * m - na * (na + 1) / 2;
      int pq[kNodePairs];  // P | Q << 8 of this lane's pairs
#pragma unroll
      for (int j = 0; j < kNodePairs; ++j) {
        int P = 0, Q = 0;
        if (lane + 64 * j < cells) join_cell(lane + 64 * j, true, m, m, P, Q);
        pq[j] = P | Q << 8;
      }
      float dist[kNodePairs];
      node_pairs4(X, D, ids, m, cells, 0, rows, pq, dist);
      if (cells > 256) node_pairs4(X, D, ids, m, cells, 256, rows, pq + 4, dist + 4);
      uint32_t li[kNodePairs];  // first endpoint's index | second's << 16; ~0u: no pair
#pragma unroll
      for (int j = 0; j < kNodePairs; ++j) {
        li[j] = ~0u;
        int sa, sb;
        uint32_t x, y;
        if (lane + 64 * j < cells && node_ends(pq[j] & 0xff, pq[j] >> 8, na, m, ids, sa, sb, x, y)) {
          ++pairs_local;
          uint32_t la = 0xffffu, lb = 0xffffu;
          if (make_key(dist[j], y) < w_s[wv][sa]) la = atomicAdd(&cnt_s[wv][sa], 1u);
          if (make_key(dist[j], x) < w_s[wv][sb]) lb = atomicAdd(&cnt_s[wv][sb], 1u);
          li[j] = la | lb << 16;
        }
      }
      wave_sync();
      const uint32_t cnt = cnt_s[wv][lane];
      unsigned long long rem = 0, nr = 0, r0 = 0;
      if constexpr (kDirect) {
        off_s[wv][lane] = cnt ? uint32_t(boff[id]) + atomicAdd(&tcount[id], cnt) : 0u;
        kept_local += cnt;
      } else {
        // totals per target (result not waited for), offsets within the node
        if (cnt) atomicAdd(&tcount[id], cnt);
        uint32_t x = cnt;
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t y = uint32_t(__shfl_up(int(x), o));
          if (lane >= o) x += y;
        }
        off_s[wv][lane] = x - cnt;
        const unsigned long long tot = (unsigned long long)(uint32_t(__shfl(int(x), 63)));
        // this node's slots: the rest of the run, then a new run if needed
        rem = wend - wbase;
        if (tot > rem) {
          if (lane == 0) nr = atomicAdd(nstage, tot - rem > kStageRun ? tot - rem : kStageRun);
          nr = (unsigned long long)__shfl((long long)nr, 0);
        }
        r0 = wbase;
        if (tot > rem) {
          wbase = nr + (tot - rem);
          wend = nr + (tot - rem > kStageRun ? tot - rem : kStageRun);
        } else {
          wbase += tot;
        }
      }
      wave_sync();
#pragma unroll
      for (int j = 0; j < kNodePairs; ++j) {
        if (li[j] == ~0u) continue;
        int sa, sb;
        uint32_t xa, xb;
        node_ends(pq[j] & 0xff, pq[j] >> 8, na, m, ids, sa, sb, xa, xb);
        const uint32_t la = li[j] & 0xffffu, lb = li[j] >> 16;
        if (la != 0xffffu) {
          if constexpr (kDirect) {
            put_direct(bkey, bsrc, off_s[wv][sa] + la, make_key(dist[j], xb), nd | kFirst);
          } else {
            const unsigned long long q = off_s[wv][sa] + la, at = q < rem ? r0 + q : nr + (q - rem);
            if (at < cap) stage[at] = Stage{xa, nd | kFirst, make_key(dist[j], xb)};
          }
        }
        if (lb != 0xffffu) {
          if constexpr (kDirect) {
            put_direct(bkey, bsrc, off_s[wv][sb] + lb, make_key(dist[j], xa), nd);
          } else {
            const unsigned long long q = off_s[wv][sb] + lb, at = q < rem ? r0 + q : nr + (q - rem);
            if (at < cap) stage[at] = Stage{xb, nd, make_key(dist[j], xa)};
          }
        }
      }
      wave_sync();
    }
  }
  if (!kDirect)
    for (unsigned long long i = wbase + lane; i < wend && i < cap; i += 64) stage[i].target = kNoRow;
  join_totals<kDirect>(pairs_local, kept_local, npairs, nstage);
}

// Staged candidates into their targets' buckets.  Consecutive slots with one
// target (a task's candidates for one row) form a segment within each
// wavefront; the segment's first lane reserves its bucket slots with one
// atomic.
__global__ void k_join_scatter(const Stage* __restrict__ stage, unsigned long long m, const int32_t* __restrict__ toff,
                               uint32_t* __restrict__ cursor, uint64_t* __restrict__ bkey,
                               uint32_t* __restrict__ bsrc) {
  const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  Stage v{kNoRow, 0u, 0ull};
  if (i < m) v = stage[i];
  const bool valid = v.target != kNoRow;
  const uint32_t prev = uint32_t(__shfl_up(int(v.target), 1));
  const bool lead = valid && (lane == 0 || prev != v.target);
  const uint64_t leads = __ballot(lead), breaks = __ballot(lead || !valid);
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  const int my_lead = 63 - __clzll(leads & upto);  // valid lanes have a leader at or below
  uint32_t base = 0;
  if (lead) {
    const uint64_t after = lane == 63 ? 0ull : (breaks & ~((2ull << lane) - 1));
    const int end = after ? __ffsll((long long)after) - 1 : 64;
    base = atomicAdd(&cursor[v.target], uint32_t(end - lane));
  }
  base = uint32_t(__shfl(int(base), valid ? my_lead : lane));
  if (valid) {
    const size_t o = size_t(toff[v.target]) + base + uint32_t(lane - my_lead);
    bkey[o] = v.key;
    bsrc[o] = v.src;
  }
}

// ---------------------------------------------------------------------------
// Precision form of the build: the merge's leaf tiles and the join's pair
// tiles as |x|^2 + |y|^2 - 2 x.y with the dot products on 16-bit MFMA blocks
// and FP32 accumulation.  Every row is split into two 16-bit parts, x = hi +
// lo (hi = x rounded, lo = the rest rounded), and x.y is taken as hi.hi' +
// (hi.lo' + lo.hi'): about 22 significant bits with FP16 parts, 16 with BF16
// ones, instead of 11 or 8 for a single 16-bit copy, whose error at |x|^2 >>
// neighbour distances would reorder the neighbours.  A pair's approximate
// distance may differ between tasks (block roles, accumulation order) and
// from its FP32 distance, so the pool merge drops duplicates by id in this
// form.  The norms are FP32 sums of the FP32 rows.  The merged S-sets and the
// first 2K entries of the final pools are re-ranked by the direct FP32
// difference in the reference's order.  On integer rows 0..255 with d <= 128
// lo is zero and every sum is exact, so the build equals the FP32 build bit
// for bit.
// ---------------------------------------------------------------------------

// Split rows: hi[r][e], lo[r][e] (Dk a multiple of 64, zero padded) and the
// FP32 squared norm of row r; one wavefront per row at a time.  stat[0]: some low part
// is non-zero; stat[1]: some value is not an integer; stat[2]: the largest
// squared norm (float bits, non-negative, so they order as integers).
template <bool kF16>
__global__ void k_pack_split(const float* __restrict__ X, size_t n, int d, int D, int Dk, uint16_t* __restrict__ hi,
                             uint16_t* __restrict__ lo, float* __restrict__ norm, int* __restrict__ stat) {
  const int lane = threadIdx.x & 63;
  const size_t nw = size_t(gridDim.x) * (blockDim.x / 64);
  bool nz = false, frac = false;
  int nmax = 0;
  for (size_t r = size_t(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6); r < n; r += nw) {
    float acc = 0.f;
    for (int e = lane; e < Dk; e += 64) {
      const float v = e < d ? X[r * D + e] : 0.f;
      frac = frac || __builtin_truncf(v) != v;
      uint16_t h, l;
      if (kF16) {
        const _Float16 x = _Float16(v);
        const _Float16 y = _Float16(v - float(x));
        __builtin_memcpy(&h, &x, 2);
        __builtin_memcpy(&l, &y, 2);
      } else {
        const uint32_t u = __float_as_uint(v);
        h = uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
        const float rest = v - __uint_as_float(uint32_t(h) << 16);
        const uint32_t w = __float_as_uint(rest);
        l = uint16_t((w + 0x7fffu + ((w >> 16) & 1u)) >> 16);
      }
      hi[r * Dk + e] = h;
      lo[r * Dk + e] = l;
      nz = nz || (l & 0x7fffu) != 0;  // a low part of +-0 adds nothing
      acc = __builtin_fmaf(v, v, acc);
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) norm[r] = acc;
    nmax = max(nmax, __float_as_int(acc));
  }
  // one update per wavefront of each statistic
  const bool anz = __ballot(nz) != 0, afrac = __ballot(frac) != 0;
  if (lane == 0) {
    if (anz) atomicOr(&stat[0], 1);
    if (afrac) atomicOr(&stat[1], 1);
    atomicMax(&stat[2], nmax);
  }
}

template <bool kF16>
__device__ __forceinline__ f32x16_t mfma16(const uint4& a, const uint4& b, const f32x16_t& c) {
  if (kF16) {
    f16x8_t x, y;
    __builtin_memcpy(&x, &a, 16);
    __builtin_memcpy(&y, &b, 16);
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(x, y, c, 0, 0, 0);
  } else {
    bf16x8_t x, y;
    __builtin_memcpy(&x, &a, 16);
    __builtin_memcpy(&y, &b, 16);
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y, c, 0, 0, 0);
  }
}

// One 32 x 32 block of split dot products: lane (r, h) = (lane % 32,
// lane / 32) feeds A row r from row ra and B row r from row rb.  Columns go
// in chunks of 64; within a chunk the lane holds the 32 consecutive columns
// from 32 h as four 16-byte pieces, piece j feeding k-step j, so k-step j
// takes columns 8 j .. 8 j + 7 and 32 + 8 j .. 32 + 8 j + 7 of the chunk:
// one fixed order of the columns for both operands, which leaves every
// product and the dot product's terms unchanged, while each row is read
// in whole 128-byte lines.  The next chunk is loaded while this one is
// multiplied.  Element l of the result is A row 8 (l / 4) + 4 (lane / 32) +
// l % 4 against B row lane % 32.  Dk is a multiple of 64; with kSame the B
// rows are the A rows (rb is not read).
//
// kLo = false: every low part of the rows is zero (rows exact in the 16-bit
// type), so hi.lo' and lo.hi' are exact zeros; the low parts are neither
// read nor multiplied, which leaves every accumulator unchanged.
template <bool kF16, bool kSame, bool kLo = true>
__device__ __forceinline__ f32x16_t mfma_split(const uint16_t* __restrict__ hi, const uint16_t* __restrict__ lo,
                                               size_t ra, size_t rb, int Dk) {
  const int kh = (threadIdx.x & 32) ? 32 : 0;
  const uint4* ah = reinterpret_cast<const uint4*>(hi + ra * Dk + kh);
  const uint4* al = reinterpret_cast<const uint4*>(lo + ra * Dk + kh);
  const uint4* bh = reinterpret_cast<const uint4*>(hi + rb * Dk + kh);
  const uint4* bl = reinterpret_cast<const uint4*>(lo + rb * Dk + kh);
  const int nch = Dk >> 6;
  uint4 a0[4], a1[4], b0[4], b1[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    a0[j] = ah[j];
    if (kLo) a1[j] = al[j];
    if (!kSame) {
      b0[j] = bh[j];
      if (kLo) b1[j] = bl[j];
    }
  }
  f32x16_t acc = f32x16_t{};
#pragma unroll 1
  for (int c = 0; c < nch; ++c) {
    const int o = (c + 1 < nch ? c + 1 : c) * 8;  // next chunk (the last one again at the end)
    uint4 n0[4], n1[4], m0[4], m1[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      n0[j] = ah[o + j];
      if (kLo) n1[j] = al[o + j];
      if (!kSame) {
        m0[j] = bh[o + j];
        if (kLo) m1[j] = bl[o + j];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint4& y0 = kSame ? a0[j] : b0[j];
      const uint4& y1 = kSame ? a1[j] : b1[j];
      acc = mfma16<kF16>(a0[j], y0, acc);
      if (kLo) {
        acc = mfma16<kF16>(a0[j], y1, acc);
        acc = mfma16<kF16>(a1[j], y0, acc);
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a0[j] = n0[j];
      if (kLo) a1[j] = n1[j];
      if (!kSame) {
        b0[j] = m0[j];
        if (kLo) b1[j] = m1[j];
      }
    }
  }
  return acc;
}

// Merge leaf tiles on MFMA: one wavefront per (leaf bucket, 32 records),
// members in blocks of 32 (LM <= 64); writes the same dist / pslot entries
// as k_leaf_dist.
template <bool kF16, bool kLo>
__global__ void __launch_bounds__(kBlock) k_leaf_mfma(const uint16_t* __restrict__ xhi, const uint16_t* __restrict__ xlo,
                                                      const float* __restrict__ xn, int Dk,
                                                      const uint2* __restrict__ range,
                                                      const uint32_t* __restrict__ leaf_t,
                                                      const int32_t* __restrict__ boff, int M,
                                                      const uint64_t* __restrict__ rec, int LM, int ml, int nlev,
                                                      float* __restrict__ dist, uint2* __restrict__ pslot) {
  const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  const int r32 = lane & 31;
  for (int l = blockIdx.x * nw + (threadIdx.x >> 6); l < M; l += gridDim.x * nw) {
    const int b0 = boff[l], R = boff[l + 1] - b0;
    if (R == 0) continue;
    const uint2 r = range[l];
    const int m = int(r.y - r.x);
    for (int rb = 0; rb < R; rb += 32) {
      const int nb = min(32, R - rb);
      uint32_t pid = 0;  // prober of A row r32
      if (r32 < nb) {
        const uint64_t v = rec[b0 + rb + r32];
        const uint32_t pos = uint32_t(v & 0xffffffffu);
        pid = leaf_t[pos];
        if (lane < 32) pslot[size_t(pos) * nlev + (63 - int(v >> 32) - ml)] = make_uint2(uint32_t(b0 + rb + r32), uint32_t(l));
      }
      for (int mb = 0; mb < m; mb += 32) {
        const uint32_t mid = (mb + r32 < m) ? leaf_t[r.x + mb + r32] : 0u;  // member of B row r32
        const f32x16_t acc = mfma_split<kF16, false, kLo>(xhi, xlo, pid, mid, Dk);
        const float pn = xn[pid], mn = xn[mid];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          const int p = 8 * (i / 4) + 4 * (lane >> 5) + (i & 3);
          const float pnp = __shfl(pn, p);
          if (p < nb && mb + r32 < m) {
            const float dd = __builtin_fmaf(-2.f, acc[i], pnp + mn);
            dist[size_t(b0 + rb + p) * LM + mb + r32] = dd > 0.f ? dd : 0.f;
          }
        }
      }
    }
  }
}

// Join tasks on MFMA: one wavefront per task (the task records of
// k_join_tasks), the 32 x 32 block of dot products in one accumulator; then
// the candidates as in k_join_tiles, counted per target row in the
// wavefront's LDS slice, staged into per-wavefront runs.
template <bool kF16, bool kLo, bool kDirect>
__global__ void __launch_bounds__(kBlock) k_join_mfma(const uint16_t* __restrict__ xhi, const uint16_t* __restrict__ xlo,
                                                      const float* __restrict__ xn, int Dk, uint32_t ntask,
                                                      const uint4* __restrict__ hdr,
                                                      const uint32_t* __restrict__ tids,
                                                      const uint64_t* __restrict__ worst, Stage* __restrict__ stage,
                                                      unsigned long long cap, unsigned long long* __restrict__ nstage,
                                                      uint32_t* __restrict__ tcount,
                                                      unsigned long long* __restrict__ npairs,
                                                      const int32_t* __restrict__ boff, uint64_t* __restrict__ bkey,
                                                      uint32_t* __restrict__ bsrc) {
  constexpr int kW = kBlock / 64;
  __shared__ uint32_t ids_s[kW][64];
  __shared__ uint64_t w_s[kW][64];
  __shared__ uint32_t cnt_s[kW][64];
  __shared__ uint32_t off_s[kW][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r32 = lane & 31;
  const uint32_t xcd = blockIdx.x & 7u, nsl = (gridDim.x >> 3) * kW;
  const uint32_t lo = uint32_t(uint64_t(ntask) * xcd / 8), hi = uint32_t(uint64_t(ntask) * (xcd + 1) / 8);
  unsigned long long pairs_local = 0, kept_local = 0;
  unsigned long long wbase = 0, wend = 0;  // staging run (lane 63)
  uint32_t* ids = ids_s[wv];
  for (uint32_t t = lo + (blockIdx.x >> 3) * kW + wv; t < hi; t += nsl) {
    const uint4 h = hdr[t];
    const bool same = h.z != 0;
    const int sizeA = int(h.y & 0xff), newA = int((h.y >> 8) & 0xff), sizeB = int((h.y >> 16) & 0xff),
              newB = int(h.y >> 24);
    const uint32_t id = tids[size_t(t) * 64 + lane];
    ids[lane] = id;
    w_s[wv][lane] = id != kNoRow ? worst[id] : 0ull;
    cnt_s[wv][lane] = 0;
    const uint32_t ia = uint32_t(__shfl(int(id), r32)), ib = uint32_t(__shfl(int(id), same ? r32 : 32 + r32));
    const uint32_t ra = ia != kNoRow ? ia : 0u, rbw = ib != kNoRow ? ib : 0u;
    const f32x16_t acc = same ? mfma_split<kF16, true, kLo>(xhi, xlo, ra, ra, Dk)
                                : mfma_split<kF16, false, kLo>(xhi, xlo, ra, rbw, Dk);
    const float na = xn[ra], nb = xn[rbw];
    wave_sync();
    float dist[16];
    uint32_t li[16];  // first endpoint's index | second's << 16; ~0u: no pair
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int p = 8 * (i / 4) + 4 * (lane >> 5) + (i & 3), q = r32;
      const float nap = __shfl(na, p);
      const float dd = __builtin_fmaf(-2.f, acc[i], nap + nb);
      dist[i] = dd > 0.f ? dd : 0.f;
      li[i] = ~0u;
      int sa, sb;
      uint32_t a, b;
      if (block_ends(p, q, same, sizeA, newA, sizeB, newB, ids, sa, sb, a, b)) {
        ++pairs_local;
        uint32_t la = 0xffffu, lb = 0xffffu;
        if (make_key(dist[i], b) < w_s[wv][sa]) la = atomicAdd(&cnt_s[wv][sa], 1u);
        if (make_key(dist[i], a) < w_s[wv][sb]) lb = atomicAdd(&cnt_s[wv][sb], 1u);
        li[i] = la | lb << 16;
      }
    }
    wave_sync();
    const uint32_t cnt = cnt_s[wv][lane];
    unsigned long long rem = 0, nr = 0, r0 = 0;
    if constexpr (kDirect) {
      // the row's first slot in its target's bucket
      off_s[wv][lane] = cnt ? uint32_t(boff[id]) + atomicAdd(&tcount[id], cnt) : 0u;
      kept_local += cnt;
    } else {
      // totals per target (result not waited for), offsets within the task
      if (cnt) atomicAdd(&tcount[id], cnt);
      uint32_t x = cnt;
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = uint32_t(__shfl_up(int(x), o));
        if (lane >= o) x += y;
      }
      off_s[wv][lane] = x - cnt;
      const unsigned long long tot = (unsigned long long)(uint32_t(__shfl(int(x), 63)));
      // this task's slots: the rest of the run, then a new run if needed
      rem = wend - wbase;
      if (tot > rem) {
        if (lane == 0) nr = atomicAdd(nstage, tot - rem > kStageRun ? tot - rem : kStageRun);
        nr = (unsigned long long)__shfl((long long)nr, 0);
      }
      r0 = wbase;
      if (tot > rem) {
        wbase = nr + (tot - rem);
        wend = nr + (tot - rem > kStageRun ? tot - rem : kStageRun);
      } else {
        wbase += tot;
      }
    }
    wave_sync();
    const uint32_t nn = h.x;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      if (li[i] == ~0u) continue;
      const int p = 8 * (i / 4) + 4 * (lane >> 5) + (i & 3), q = r32;
      int sa, sb;
      uint32_t a, b;
      block_ends(p, q, same, sizeA, newA, sizeB, newB, ids, sa, sb, a, b);
      const uint32_t la = li[i] & 0xffffu, lb = li[i] >> 16;
      if constexpr (kDirect) {
        if (la != 0xffffu) put_direct(bkey, bsrc, off_s[wv][sa] + la, make_key(dist[i], b), nn | kFirst);
        if (lb != 0xffffu) put_direct(bkey, bsrc, off_s[wv][sb] + lb, make_key(dist[i], a), nn);
        continue;
      }
      if (la != 0xffffu) {
        const unsigned long long j = off_s[wv][sa] + la, at = j < rem ? r0 + j : nr + (j - rem);
        if (at < cap) stage[at] = Stage{a, nn | kFirst, make_key(dist[i], b)};
      }
      if (lb != 0xffffu) {
        const unsigned long long j = off_s[wv][sb] + lb, at = j < rem ? r0 + j : nr + (j - rem);
        if (at < cap) stage[at] = Stage{b, nn, make_key(dist[i], a)};
      }
    }
    wave_sync();
  }
  if (!kDirect)
    for (unsigned long long i = wbase + lane; i < wend && i < cap; i += 64) stage[i].target = kNoRow;
  for (int o = 32; o > 0; o >>= 1) {
    pairs_local += (unsigned long long)__shfl_down((long long)pairs_local, o);
    kept_local += (unsigned long long)__shfl_down((long long)kept_local, o);
  }
  if (lane == 0 && pairs_local) atomicAdd(npairs, pairs_local);
  if (kDirect && lane == 0 && kept_local) atomicAdd(nstage, kept_local);
}

// Three 32 x 32 blocks of split dot products of one node's rows, loaded once:
// A = rows ra (positions 0..31), B = rows rb (positions 32..63); t00 = A.A,
// t01 = A.B, t11 = B.B, columns in the order of mfma_split.
template <bool kF16, bool kLo>
__device__ __forceinline__ void mfma_split3(const uint16_t* __restrict__ hi, const uint16_t* __restrict__ lo,
                                            size_t ra, size_t rb, int Dk, f32x16_t& t00, f32x16_t& t01,
                                            f32x16_t& t11) {
  const int kh = (threadIdx.x & 32) ? 32 : 0;
  const uint4* ah = reinterpret_cast<const uint4*>(hi + ra * Dk + kh);
  const uint4* al = reinterpret_cast<const uint4*>(lo + ra * Dk + kh);
  const uint4* bh = reinterpret_cast<const uint4*>(hi + rb * Dk + kh);
  const uint4* bl = reinterpret_cast<const uint4*>(lo + rb * Dk + kh);
  const int nch = Dk >> 6;
  uint4 a0[4], a1[4], b0[4], b1[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    a0[j] = ah[j];
    b0[j] = bh[j];
    if (kLo) {
      a1[j] = al[j];
      b1[j] = bl[j];
    }
  }
  t00 = f32x16_t{};
  t01 = f32x16_t{};
  t11 = f32x16_t{};
#pragma unroll 1
  for (int c = 0; c < nch; ++c) {
    const int o = (c + 1 < nch ? c + 1 : c) * 8;  // next chunk (the last one again at the end)
    uint4 n0[4], n1[4], m0[4], m1[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      n0[j] = ah[o + j];
      m0[j] = bh[o + j];
      if (kLo) {
        n1[j] = al[o + j];
        m1[j] = bl[o + j];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      t00 = mfma16<kF16>(a0[j], a0[j], t00);
      t01 = mfma16<kF16>(a0[j], b0[j], t01);
      t11 = mfma16<kF16>(b0[j], b0[j], t11);
      if (kLo) {
        t00 = mfma16<kF16>(a0[j], a1[j], t00);
        t01 = mfma16<kF16>(a0[j], b1[j], t01);
        t11 = mfma16<kF16>(b0[j], b1[j], t11);
        t00 = mfma16<kF16>(a1[j], a0[j], t00);
        t01 = mfma16<kF16>(a1[j], b0[j], t01);
        t11 = mfma16<kF16>(b1[j], b0[j], t11);
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      a0[j] = n0[j];
      b0[j] = m0[j];
      if (kLo) {
        a1[j] = n1[j];
        b1[j] = m1[j];
      }
    }
  }
}

// Whole-node join on MFMA (precision form) for the nodes with at most 64
// rows: one wavefront per node, each row read once (the block tasks read a
// node's rows once per task), the node's pairs from the blocks (0,0), and
// for more than 32 rows (0,1) and (1,1).  kNT = 1 takes the nodes of at most
// 32 rows, kNT = 3 those of 33 to 64 (two launches: the one-block form needs
// a third of the registers, so more wavefronts stay resident).  Nodes are
// taken in the order of tree 0's leaf list (order, null: node order), one
// contiguous range per XCD.  Candidates are counted, offset and staged as in
// k_join_mfma.
template <bool kF16, bool kLo, int kNT, bool kDirect>
__global__ void __launch_bounds__(kBlock, kNT == 1 ? 4 : 3) k_join_node_mfma(const uint16_t* __restrict__ xhi,
                                                           const uint16_t* __restrict__ xlo,
                                                           const float* __restrict__ xn, int Dk, uint32_t n,
                                                           const uint32_t* __restrict__ order,
                                                           const uint64_t* __restrict__ new_off,
                                                           const uint32_t* __restrict__ new_ids,
                                                           const uint64_t* __restrict__ old_off,
                                                           const uint32_t* __restrict__ old_ids,
                                                           const uint64_t* __restrict__ worst, Stage* __restrict__ stage,
                                                           unsigned long long cap,
                                                           unsigned long long* __restrict__ nstage,
                                                           uint32_t* __restrict__ tcount,
                                                           unsigned long long* __restrict__ npairs,
                                                           const int32_t* __restrict__ boff,
                                                           uint64_t* __restrict__ bkey,
                                                           uint32_t* __restrict__ bsrc) {
  constexpr int kW = kBlock / 64;
  __shared__ uint32_t ids_s[kW][64];
  __shared__ uint64_t w_s[kW][64];
  __shared__ uint32_t cnt_s[kW][64];
  __shared__ uint32_t off_s[kW][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r32 = lane & 31;
  const uint32_t xcd = blockIdx.x & 7u, nsl = (gridDim.x >> 3) * kW;
  const uint32_t lo = uint32_t(uint64_t(n) * xcd / 8), hi = uint32_t(uint64_t(n) * (xcd + 1) / 8);
  unsigned long long pairs_local = 0, kept_local = 0;
  unsigned long long wbase = 0, wend = 0;  // staging run
  uint32_t* ids = ids_s[wv];
  // the wavefront's nodes i0, i0 + nsl, ... are read 64 at a time, one per
  // lane; the nodes this launch takes are then worked through in order
  for (uint32_t i0 = lo + (blockIdx.x >> 3) * kW + wv; i0 < hi; i0 += 64 * nsl) {
    const uint32_t ij = i0 + uint32_t(lane) * nsl;
    uint32_t ndj = 0;
    uint64_t a0j = 0, b0j = 0;
    int naj = 0, mj = 0;
    if (ij < hi) {
      ndj = order ? order[ij] : ij;
      a0j = new_off[ndj];
      b0j = old_off[ndj];
      naj = int(new_off[ndj + 1] - a0j);
      mj = naj + int(old_off[ndj + 1] - b0j);
    }
    // no pairs / the block tasks' nodes / the other launch's nodes
    uint64_t todo = __ballot(ij < hi && naj > 0 && mj <= 64 && (mj > 32) == (kNT == 3));
    while (todo) {
      const int jn = __ffsll((long long)todo) - 1;
      todo &= todo - 1;
      const uint32_t nd = uint32_t(__shfl(int(ndj), jn));
      const uint64_t a0 = uint64_t(__shfl((long long)a0j, jn)), b0 = uint64_t(__shfl((long long)b0j, jn));
      const int na = __shfl(naj, jn), m = __shfl(mj, jn);
      const uint32_t id = lane < m ? (lane < na ? new_ids[a0 + lane] : old_ids[b0 + lane - na]) : kNoRow;
      ids[lane] = id;
      w_s[wv][lane] = id != kNoRow ? worst[id] : 0ull;
      cnt_s[wv][lane] = 0;
      const uint32_t ia = uint32_t(__shfl(int(id), r32)), ib = uint32_t(__shfl(int(id), 32 + r32));
      const uint32_t ra = ia != kNoRow ? ia : 0u, rb = ib != kNoRow ? ib : 0u;
      f32x16_t t[kNT];  // blocks (0,0) [, (0,1), (1,1)]
      if constexpr (kNT == 3) mfma_split3<kF16, kLo>(xhi, xlo, ra, rb, Dk, t[0], t[1], t[2]);
      else t[0] = mfma_split<kF16, true, kLo>(xhi, xlo, ra, ra, Dk);
      const float nra = xn[ra], nrb = xn[rb];
      wave_sync();
      // per pair: slot of the first endpoint's candidate | second's << 8 (0xff:
      // none; a row meets at most 63 others), two pairs per register; the
      // distances are formed again from the accumulators when staging
      uint32_t li[kNT][8];
#pragma unroll
      for (int b = 0; b < kNT; ++b) {
        const int pb = b == 2 ? 32 : 0, qb = b == 0 ? 0 : 32;  // position offsets of the A and B rows
        const float nq = b == 0 ? nra : nrb;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int p = 8 * (e / 4) + 4 * (lane >> 5) + (e & 3);
          const float np = __shfl(b == 2 ? nrb : nra, p);
          const float dd = __builtin_fmaf(-2.f, t[b][e], np + nq);
          const float dist = dd > 0.f ? dd : 0.f;
          uint32_t v = 0xffffu;
          int sa, sb;
          uint32_t x, y;
          if (node_ends(pb + p, qb + r32, na, m, ids, sa, sb, x, y)) {
            ++pairs_local;
            uint32_t la = 0xffu, lb = 0xffu;
            if (make_key(dist, y) < w_s[wv][sa]) la = atomicAdd(&cnt_s[wv][sa], 1u);
            if (make_key(dist, x) < w_s[wv][sb]) lb = atomicAdd(&cnt_s[wv][sb], 1u);
            v = la | lb << 8;
          }
          if (e & 1) li[b][e >> 1] |= v << 16;
          else li[b][e >> 1] = v;
        }
      }
      wave_sync();
      const uint32_t cnt = cnt_s[wv][lane];
      unsigned long long rem = 0, nr = 0, r0 = 0;
      if constexpr (kDirect) {
        off_s[wv][lane] = cnt ? uint32_t(boff[id]) + atomicAdd(&tcount[id], cnt) : 0u;
        kept_local += cnt;
      } else {
        // totals per target (result not waited for), offsets within the node
        if (cnt) atomicAdd(&tcount[id], cnt);
        uint32_t x = cnt;
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t y = uint32_t(__shfl_up(int(x), o));
          if (lane >= o) x += y;
        }
        off_s[wv][lane] = x - cnt;
        const unsigned long long tot = (unsigned long long)(uint32_t(__shfl(int(x), 63)));
        // this node's slots: the rest of the run, then a new run if needed
        rem = wend - wbase;
        if (tot > rem) {
          if (lane == 0) nr = atomicAdd(nstage, tot - rem > kStageRun ? tot - rem : kStageRun);
          nr = (unsigned long long)__shfl((long long)nr, 0);
        }
        r0 = wbase;
        if (tot > rem) {
          wbase = nr + (tot - rem);
          wend = nr + (tot - rem > kStageRun ? tot - rem : kStageRun);
        } else {
          wbase += tot;
        }
      }
      wave_sync();
#pragma unroll
      for (int b = 0; b < kNT; ++b) {
        const int pb = b == 2 ? 32 : 0, qb = b == 0 ? 0 : 32;
        const float nq = b == 0 ? nra : nrb;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int p = 8 * (e / 4) + 4 * (lane >> 5) + (e & 3);
          const float np = __shfl(b == 2 ? nrb : nra, p);
          const uint32_t v = (li[b][e >> 1] >> (16 * (e & 1))) & 0xffffu;
          if (v == 0xffffu) continue;
          const float dd = __builtin_fmaf(-2.f, t[b][e], np + nq);
          const float dist = dd > 0.f ? dd : 0.f;
          int sa, sb;
          uint32_t xa, xb;
          node_ends(pb + p, qb + r32, na, m, ids, sa, sb, xa, xb);
          const uint32_t la = v & 0xffu, lb = v >> 8;
          if (la != 0xffu) {
            if constexpr (kDirect) {
              put_direct(bkey, bsrc, off_s[wv][sa] + la, make_key(dist, xb), nd | kFirst);
            } else {
              const unsigned long long j = off_s[wv][sa] + la, at = j < rem ? r0 + j : nr + (j - rem);
              if (at < cap) stage[at] = Stage{xa, nd | kFirst, make_key(dist, xb)};
            }
          }
          if (lb != 0xffu) {
            if constexpr (kDirect) {
              put_direct(bkey, bsrc, off_s[wv][sb] + lb, make_key(dist, xa), nd);
            } else {
              const unsigned long long j = off_s[wv][sb] + lb, at = j < rem ? r0 + j : nr + (j - rem);
              if (at < cap) stage[at] = Stage{xb, nd, make_key(dist, xa)};
            }
          }
        }
      }
      wave_sync();
    }
  }
  if (!kDirect)
    for (unsigned long long i = wbase + lane; i < wend && i < cap; i += 64) stage[i].target = kNoRow;
  join_totals<kDirect>(pairs_local, kept_local, npairs, nstage);
}

// FP32 re-rank of the final pools (the first min(fill, 2K) entries) and of
// the merged S-sets (all of them): each entry gets its direct FP32 distance
// (a lane pair per entry, the reference's order); k_sort_pools then orders
// that prefix (keys, with their flags for the pools).  Points are taken in
// `order` (tree 0's leaf list, or null: point order), so points re-ranked
// together are close and share most of their neighbours' rows in the caches.
__global__ void k_rerank_pool(const float* __restrict__ X, int D, uint32_t n, int L, int R,
                              const uint32_t* __restrict__ order, uint64_t* __restrict__ pool,
                              const uint32_t* __restrict__ fill) {
  const size_t g = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 1;
  const int h = threadIdx.x & 1;
  const uint32_t i = uint32_t(g / R), j = uint32_t(g % R);
  const uint32_t p = order && i < n ? order[i] : i;
  const bool in = g < size_t(n) * R && j < fill[p < n ? p : 0];
  const uint32_t id = in ? key_id(pool[size_t(p) * L + j]) : 0u;
  const float dd = l2_pair_lanes(X + size_t(in ? p : 0) * D, X + size_t(id) * D, D, h);
  if (in && h) pool[size_t(p) * L + j] = make_key(dd, id);
}
// One wavefront per point: the prefix, padded to a power of two with the
// largest key, is sorted in LDS by a bitonic network (flags move with their
// keys).  R is at most 1024.
template <bool kFlags>
__global__ void __launch_bounds__(kBlock) k_sort_pools(uint32_t n, int L, int R, uint64_t* __restrict__ pool,
                                                       uint8_t* __restrict__ flag, const uint32_t* __restrict__ fill) {
  __shared__ uint64_t ks[kBlock / 64][1024];
  __shared__ uint8_t fs[kBlock / 64][kFlags ? 1024 : 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint32_t p = blockIdx.x * (kBlock / 64) + w; p < n; p += gridDim.x * (kBlock / 64)) {
    uint64_t* k = pool + size_t(p) * L;
    const int m = min(int(fill[p]), R);
    if (m < 2) continue;
    const int p2 = next_pow2(m);
    for (int j = lane; j < p2; j += 64) {
      ks[w][j] = j < m ? k[j] : kKeyMax;
      if (kFlags) fs[w][j] = j < m ? flag[size_t(p) * L + j] : 0;
    }
    wave_sync();
    for (int kk = 2; kk <= p2; kk <<= 1) {
      for (int jj = kk >> 1; jj > 0; jj >>= 1) {
        for (int t = lane; t < p2; t += 64) {
          const int u = t ^ jj;
          if (u > t) {
            const uint64_t a = ks[w][t], b = ks[w][u];
            if ((a > b) == ((t & kk) == 0)) {
              ks[w][t] = b;
              ks[w][u] = a;
              if (kFlags) {
                const uint8_t fa = fs[w][t];
                fs[w][t] = fs[w][u];
                fs[w][u] = fa;
              }
            }
          }
        }
        wave_sync();
      }
    }
    for (int j = lane; j < m; j += 64) {
      k[j] = ks[w][j];
      if (kFlags) flag[size_t(p) * L + j] = fs[w][j];
    }
    wave_sync();
  }
}

// Candidate tags: existing entries carry their flag (0/1); candidates carry
// 1<<31 | (not first endpoint) << 30 | joining node, so duplicates resolve to
// the existing entry, else to the lowest first-endpoint joining node.
__device__ inline uint32_t cand_tag(uint32_t s) {
  return 0x80000000u | ((s & kFirst) ? 0u : 0x40000000u) | (s & 0x3fffffffu);
}

// Target x's candidates are bucket slots [toff[x], toff[x + 1]) after the
// scatter, or [toff[x], toff[x] + tcnt[x]) with direct routing (tcnt given).
//
// Per target: pool' = best L of (pool + candidates) by (key, tag), one
// wavefront per target, in rounds: the running list (at most L entries,
// sorted, unique) plus the next candidates at or below its worst key (every
// candidate while it holds fewer than L), up to kCap entries, are sorted by
// (key, tag) in the wavefront's LDS window, duplicates dropped, and the
// first L kept.  An entry beyond the L-th unique key of a subset is beyond it
// in the union, so rounds give the one-shot result, and a candidate above a
// full list's worst key can be skipped.
//
// kById (precision form of the build): a pair's approximate distance may
// differ from the FP32 distance its id already carries in the pool (from the
// merge's re-ranked sets or the initial top-up), so duplicates are dropped by
// id through a per-wavefront LDS hash of the window's ids: an entry already
// in the pool stays (as an equal key does in the FP32 form, where a
// duplicate's key is the pool entry's), else the first candidate in (key,
// tag) order.
template <int kCap, bool kById>
__global__ void __launch_bounds__(kBlock) k_merge_wave(uint32_t n, int L, int CK, const int32_t* __restrict__ toff,
                                                       const uint32_t* __restrict__ tcnt,
                                                       const uint64_t* __restrict__ bkey,
                                                       const uint32_t* __restrict__ bsrc,
                                                       uint64_t* __restrict__ pool, uint8_t* __restrict__ flag,
                                                       uint32_t* __restrict__ fill, uint8_t* __restrict__ credit) {
  constexpr int kW = kCap >= 1024 ? 2 : 4;  // wavefronts per workgroup
  constexpr int kHS = kById ? 2 * kCap : 1;  // id hash slots per wavefront
  __shared__ uint64_t ks[kW][kCap];
  __shared__ uint32_t ts[kW][kCap];
  __shared__ uint32_t hid[kW][kHS], hpos[kW][kHS];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t* k = ks[w];
  uint32_t* tg = ts[w];
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (uint32_t x = blockIdx.x * kW + w; x < n; x += gridDim.x * kW) {
    const int c0 = toff[x], c1 = tcnt ? c0 + int(tcnt[x]) : toff[x + 1];
    if (c0 == c1) continue;
    int cur = int(fill[x]);
    for (int i = lane; i < cur; i += 64) {
      k[i] = pool[size_t(x) * L + i];
      tg[i] = flag[size_t(x) * L + i];
    }
    for (int base = c0; base < c1;) {
      // the window: the pool, then the next candidates at or below the pool's
      // worst key (all of them while the pool has room); a candidate above a
      // full pool's worst key cannot be among the best L distinct keys
      const uint64_t wk = cur >= L ? k[L - 1] : kKeyMax;
      int m = cur;
      while (base < c1 && m < kCap) {
        const int i = base + lane;
        const uint64_t v = i < c1 ? bkey[i] : kKeyMax;
        const bool keep = i < c1 && v <= wk;
        const uint64_t km = __ballot(keep);
        const int room = kCap - m, rank = __popcll(km & lt), nk = __popcll(km);
        if (keep && rank < room) {
          k[m + rank] = v;
          tg[m + rank] = cand_tag(bsrc[i]);
        }
        if (nk > room) {  // the kept candidate of rank `room` opens the next window
          base += __ffsll((long long)__ballot(keep && rank == room)) - 1;
          m = kCap;
        } else {
          base += 64;
          m += nk;
        }
      }
      if (m == cur) break;  // no candidate left that could enter the pool
      const int p2 = next_pow2(m < 2 ? 2 : m);
      for (int i = m + lane; i < p2; i += 64) {
        k[i] = kKeyMax;
        tg[i] = 0xffffffffu;
      }
      wave_sync();
      for (int kk = 2; kk <= p2; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
          for (int t = lane; t < p2; t += 64) {
            const int u = t ^ j;
            if (u > t) {
              const uint64_t ka = k[t], kb = k[u];
              const uint32_t ta = tg[t], tb = tg[u];
              const bool gt = (ka > kb) || (ka == kb && ta > tb);
              if (gt == ((t & kk) == 0)) {
                k[t] = kb;
                k[u] = ka;
                tg[t] = tb;
                tg[u] = ta;
              }
            }
          }
          wave_sync();
        }
      }
      if (kById) {  // per id of the window: its pool entry, else its first candidate
        uint32_t* hi_ = hid[w];
        uint32_t* hp = hpos[w];
        for (int i = lane; i < kHS; i += 64) {
          hi_[i] = 0xffffffffu;
          hp[i] = 0xffffffffu;
        }
        wave_sync();
        for (int i = lane; i < m; i += 64) {
          const uint32_t id = key_id(k[i]);
          uint32_t h = (id * 2654435761u) & uint32_t(kHS - 1);
          for (;;) {
            const uint32_t prev = atomicCAS(&hi_[h], 0xffffffffu, id);
            if (prev == 0xffffffffu || prev == id) break;
            h = (h + 1) & uint32_t(kHS - 1);
          }
          atomicMin(&hp[h], (tg[i] & 0x80000000u) | uint32_t(i));  // a pool entry before any candidate
        }
        wave_sync();
      }
      // drop duplicates (the first of equal keys stays; of equal ids, the pool
      // entry or else the first), compact to the front
      int kept = 0;
      for (int b0 = 0; b0 < m && kept < L; b0 += 64) {
        const int i = b0 + lane;
        bool first = false;
        uint64_t kv = 0;
        uint32_t tv = 0;
        if (i < m) {
          kv = k[i];
          tv = tg[i];
          if (kById) {
            const uint32_t id = key_id(kv);
            uint32_t h = (id * 2654435761u) & uint32_t(kHS - 1);
            while (hid[w][h] != id) h = (h + 1) & uint32_t(kHS - 1);
            first = hpos[w][h] == ((tv & 0x80000000u) | uint32_t(i));
          } else {
            first = i == 0 || k[i - 1] != kv;
          }
        }
        const uint64_t fm = __ballot(first);
        const int at = kept + __popcll(fm & lt);
        wave_sync();
        if (first && at < L) {  // at <= i: only this lane's own slot or earlier ones are written
          k[at] = kv;
          tg[at] = tv;
        }
        wave_sync();
        kept += __popcll(fm);
      }
      cur = kept < L ? kept : L;
    }
    for (int i = lane; i < L; i += 64) {
      if (i < cur) {
        const uint32_t tv = tg[i];
        pool[size_t(x) * L + i] = k[i];
        flag[size_t(x) * L + i] = (tv & 0x80000000u) ? 1 : uint8_t(tv & 1u);
        if ((tv & 0xc0000000u) == 0x80000000u && i < CK) credit[tv & 0x3fffffffu] = 1;
      } else {
        pool[size_t(x) * L + i] = kKeyMax;
        flag[size_t(x) * L + i] = 0;
      }
    }
    if (lane == 0) fill[x] = uint32_t(cur);
    wave_sync();
  }
}

// (key, tag) order of two window entries
__device__ __forceinline__ bool kt_less(uint64_t ka, uint32_t ta, uint64_t kb, uint32_t tb) {
  return ka < kb || (ka == kb && ta < tb);
}

// Pool merge for pools of at most kLMax entries (FP32 build): the result of
// k_merge_wave<*, false> (best L distinct keys of pool + candidates, the
// lowest tag staying among equal keys), reached with less work.  Rounds take
// up to 64 candidates at or below the running list's worst key, one per lane;
// they are sorted by (key, tag) in registers (bitonic network over shuffles)
// and merged with the sorted running list by ranks into a second LDS buffer:
// a candidate's place is its sorted index plus the number of list entries at
// or before it (binary search), a list entry's place its index plus the
// number of candidates before it.  Exactly equal entries put the list's
// first.  Then duplicates are dropped and the first L copied back as the new
// running list.  One wavefront per target.  kById (precision form of the
// build): duplicates are dropped by id, an entry already in the pool staying
// (else the first in (key, tag) order), through an LDS hash of the window's
// ids (as in k_merge_wave).
template <int kLMax, bool kById>
__global__ void __launch_bounds__(kBlock) k_merge_reg(uint32_t n, int L, int CK, const int32_t* __restrict__ toff,
                                                      const uint32_t* __restrict__ tcnt,
                                                      const uint64_t* __restrict__ bkey,
                                                      const uint32_t* __restrict__ bsrc, uint64_t* __restrict__ pool,
                                                      uint8_t* __restrict__ flag, uint32_t* __restrict__ fill,
                                                      uint8_t* __restrict__ credit) {
  constexpr int kW = kBlock / 64;
  __shared__ uint64_t ls_k[kW][kLMax];  // running list
  __shared__ uint32_t ls_t[kW][kLMax];
  __shared__ uint64_t ms_k[kW][kLMax + 64];  // merged window
  __shared__ uint32_t ms_t[kW][kLMax + 64];
  __shared__ uint64_t cs_k[kW][64];  // candidates of the round
  __shared__ uint32_t cs_t[kW][64];
  // id hash slots: at least 1.6 (kLMax + 64), a power of two
  constexpr int kHS = kById ? (kLMax <= 64 ? 256 : (kLMax <= 256 ? 512 : 1024)) : 1;
  __shared__ uint32_t hid_s[kW][kHS], hpos_s[kW][kHS];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t* A = ls_k[w];
  uint32_t* AT = ls_t[w];
  uint64_t* B = ms_k[w];
  uint32_t* BT = ms_t[w];
  uint64_t* C = cs_k[w];
  uint32_t* CT = cs_t[w];
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (uint32_t x = blockIdx.x * kW + w; x < n; x += gridDim.x * kW) {
    const int c0 = toff[x], c1 = tcnt ? c0 + int(tcnt[x]) : toff[x + 1];
    if (c0 == c1) continue;
    int cur = int(fill[x]);
    // the first window's candidates are loaded with the pool (one latency
    // for both; most targets have at most one window)
    const int i0 = c0 + lane;
    const uint64_t v0 = i0 < c1 ? bkey[i0] : kKeyMax;
    const uint32_t s0 = i0 < c1 ? bsrc[i0] : 0u;
    for (int i = lane; i < cur; i += 64) {
      A[i] = pool[size_t(x) * L + i];
      AT[i] = flag[size_t(x) * L + i];
    }
    wave_sync();
    for (int base = c0; base < c1;) {
      const uint64_t wk = cur >= L ? A[L - 1] : kKeyMax;
      int m = 0;
      while (base < c1 && m < 64) {
        const bool pre = base == c0;  // the first window (base only grows)
        const int i = base + lane;
        const uint64_t v = pre ? v0 : (i < c1 ? bkey[i] : kKeyMax);
        const bool keep = i < c1 && v <= wk;
        const uint64_t km = __ballot(keep);
        const int room = 64 - m, rank = __popcll(km & lt), nk = __popcll(km);
        if (keep && rank < room) {
          C[m + rank] = v;
          CT[m + rank] = cand_tag(pre ? s0 : bsrc[i]);
        }
        if (nk > room) {  // the kept candidate of rank `room` opens the next round
          base += __ffsll((long long)__ballot(keep && rank == room)) - 1;
          m = 64;
        } else {
          base += 64;
          m += nk;
        }
      }
      if (m == 0) break;  // no candidate left that could enter the list
      wave_sync();
      uint64_t ck = lane < m ? C[lane] : kKeyMax;
      uint32_t ct = lane < m ? CT[lane] : 0xffffffffu;
      const int m2 = next_pow2(m < 2 ? 2 : m);
      for (int kk = 2; kk <= m2; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
          const uint32_t olo = uint32_t(__shfl_xor(int(uint32_t(ck)), j));
          const uint32_t ohi = uint32_t(__shfl_xor(int(uint32_t(ck >> 32)), j));
          const uint32_t ot = uint32_t(__shfl_xor(int(ct), j));
          const uint64_t ok = uint64_t(ohi) << 32 | olo;
          const bool want_min = ((lane & j) == 0) == ((lane & kk) == 0);
          const bool take = want_min ? kt_less(ok, ot, ck, ct) : kt_less(ck, ct, ok, ot);
          if (take) {
            ck = ok;
            ct = ot;
          }
        }
      }
      wave_sync();  // every lane has read C before the sorted copy replaces it
      C[lane] = ck;
      CT[lane] = ct;
      wave_sync();
      if (lane < m) {  // list entries at or before the candidate
        int lo = 0, hi = cur;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (!kt_less(ck, ct, A[mid], AT[mid])) lo = mid + 1;
          else hi = mid;
        }
        B[lane + lo] = ck;
        BT[lane + lo] = ct;
      }
      for (int i = lane; i < cur; i += 64) {  // candidates strictly before the list entry
        const uint64_t pk = A[i];
        const uint32_t pt = AT[i];
        int lo = 0, hi = m;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (kt_less(C[mid], CT[mid], pk, pt)) lo = mid + 1;
          else hi = mid;
        }
        B[i + lo] = pk;
        BT[i + lo] = pt;
      }
      wave_sync();
      const int mm = cur + m;
      if (kById) {  // per id of the window: its pool entry, else its first candidate
        uint32_t* hi_ = hid_s[w];
        uint32_t* hp = hpos_s[w];
        for (int i = lane; i < kHS; i += 64) {
          hi_[i] = 0xffffffffu;
          hp[i] = 0xffffffffu;
        }
        wave_sync();
        for (int i = lane; i < mm; i += 64) {
          const uint32_t id = key_id(B[i]);
          uint32_t h = (id * 2654435761u) & uint32_t(kHS - 1);
          for (;;) {
            const uint32_t prev = atomicCAS(&hi_[h], 0xffffffffu, id);
            if (prev == 0xffffffffu || prev == id) break;
            h = (h + 1) & uint32_t(kHS - 1);
          }
          atomicMin(&hp[h], (BT[i] & 0x80000000u) | uint32_t(i));  // a pool entry before any candidate
        }
        wave_sync();
      }
      // drop duplicates (the first of equal keys stays; of equal ids, the pool entry or else the first);
      // the first L kept become the running list
      int kept = 0;
      for (int b0 = 0; b0 < mm && kept < L; b0 += 64) {
        const int i = b0 + lane;
        bool first = false;
        uint64_t kv = 0;
        uint32_t tv = 0;
        if (i < mm) {
          kv = B[i];
          tv = BT[i];
          if (kById) {
            const uint32_t id = key_id(kv);
            uint32_t h = (id * 2654435761u) & uint32_t(kHS - 1);
            while (hid_s[w][h] != id) h = (h + 1) & uint32_t(kHS - 1);
            first = hpos_s[w][h] == ((tv & 0x80000000u) | uint32_t(i));
          } else {
            first = i == 0 || B[i - 1] != kv;
          }
        }
        const uint64_t fm = __ballot(first);
        const int at = kept + __popcll(fm & lt);
        if (first && at < L) {
          A[at] = kv;
          AT[at] = tv;
        }
        kept += __popcll(fm);
      }
      wave_sync();
      cur = kept < L ? kept : L;
    }
    for (int i = lane; i < L; i += 64) {
      if (i < cur) {
        const uint32_t tv = AT[i];
        pool[size_t(x) * L + i] = A[i];
        flag[size_t(x) * L + i] = (tv & 0x80000000u) ? 1 : uint8_t(tv & 1u);
        if ((tv & 0xc0000000u) == 0x80000000u && i < CK) credit[tv & 0x3fffffffu] = 1;
      } else {
        pool[size_t(x) * L + i] = kKeyMax;
        flag[size_t(x) * L + i] = 0;
      }
    }
    if (lane == 0) fill[x] = uint32_t(cur);
    wave_sync();
  }
}

// ---------------------------------------------------------------------------
// update
// ---------------------------------------------------------------------------
// The update kernels give each node a group of G lanes (G a power of two,
// at most 64; 64 / G nodes per wavefront), so that a node's pool entries are
// read in one coalesced pass and counted with ballots over the group.
template <int G>
__device__ __forceinline__ uint64_t group_mask() {
  const int lane = threadIdx.x & 63;
  return (G == 64 ? ~0ull : ((1ull << G) - 1)) << (lane & ~(G - 1));
}

// Range (base_index.hpp:383-393): for found nodes, advance over at most
// min(Range + S, fill) entries until S new ones are seen; then radiusM =
// distance of entry Range - 1 (394).
template <int G>
__global__ void k_range(const uint64_t* __restrict__ pool, const uint8_t* __restrict__ flag,
                        const uint32_t* __restrict__ fill, const uint8_t* __restrict__ found, uint32_t n, int L, int S,
                        uint32_t* __restrict__ rng, float* __restrict__ radiusM) {
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint32_t p = uint32_t(t / G);
  const int l0 = int(t % G), lane = threadIdx.x & 63;
  if (p >= n) return;  // whole groups leave together
  const uint64_t seg = group_mask<G>();
  uint32_t R = rng[p];
  if (found[p]) {
    const uint32_t maxl = R + uint32_t(S) < fill[p] ? R + uint32_t(S) : fill[p];
    uint32_t c = 0, nr = maxl;
    for (int lb = 0; lb < int(maxl); lb += G) {  // maxl is the same across the group
      const int l = lb + l0;
      const bool f = uint32_t(l) < maxl && flag[size_t(p) * L + l] != 0;
      uint64_t m = (__ballot(f) & seg) >> (lane & ~(G - 1));
      const uint32_t cnt = uint32_t(__popcll(m));
      if (c + cnt >= uint32_t(S)) {  // the S-th new entry lies in this block
        for (uint32_t i = 1; i < uint32_t(S) - c; ++i) m &= m - 1;
        nr = uint32_t(lb + __ffsll((long long)m));
        break;
      }
      c += cnt;
    }
    R = nr;
    if (l0 == 0) rng[p] = R;
  }
  if (l0 == 0) radiusM[p] = key_dist(pool[size_t(p) * L + R - 1]);
}

// pass 0: forward counts and reverse counts, each reverse entry's slot in
// its list (the count before it) kept in rslot (kNoSlot: no reverse entry);
// pass 1: reverse fill from the kept slots (base_index.hpp:396-420).  The
// order within a reverse list is the order of pass 0's atomics: lists above
// the cap are sampled by generator value, and the join's candidates are
// merged by (key, tag), neither depending on it.
constexpr uint32_t kNoSlot = 0xffffffffu, kOldSlot = 0x80000000u;
template <int G>
__global__ void k_lists(const uint64_t* __restrict__ pool, const uint8_t* __restrict__ flag,
                        const uint32_t* __restrict__ rng, const float* __restrict__ radiusM, uint32_t n, int L, int pass,
                        uint32_t* __restrict__ fcnt_new, uint32_t* __restrict__ fcnt_old,
                        uint32_t* __restrict__ rcnt_new, uint32_t* __restrict__ rcnt_old,
                        const int32_t* __restrict__ roff_new, const int32_t* __restrict__ roff_old,
                        uint32_t* __restrict__ rslot, uint32_t* __restrict__ rids_new,
                        uint32_t* __restrict__ rids_old) {
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint32_t p = uint32_t(t / G);
  const int l0 = int(t % G);
  if (p >= n) return;
  const uint64_t seg = group_mask<G>();
  const uint32_t R = rng[p];
  uint32_t fn = 0;
  for (int lb = 0; lb < int(R); lb += G) {
    const int l = lb + l0;
    const bool act = uint32_t(l) < R;
    if (pass == 1) {
      if (act) {
        const uint32_t sl = rslot[size_t(p) * L + l];
        if (sl != kNoSlot) {
          const uint32_t o = key_id(pool[size_t(p) * L + l]);
          if (sl & kOldSlot) rids_old[size_t(roff_old[o]) + (sl & ~kOldSlot)] = p;
          else rids_new[size_t(roff_new[o]) + sl] = p;
        }
      }
      continue;
    }
    bool isnew = false, rev = false;
    uint32_t o = 0;
    if (act) {
      const uint64_t e = pool[size_t(p) * L + l];
      o = key_id(e);
      isnew = flag[size_t(p) * L + l] != 0;
      rev = key_dist(e) > radiusM[o];
    }
    fn += uint32_t(__popcll(__ballot(act && isnew) & seg));
    if (act) {
      uint32_t sl = kNoSlot;
      if (rev) sl = isnew ? atomicAdd(&rcnt_new[o], 1u) : (atomicAdd(&rcnt_old[o], 1u) | kOldSlot);
      rslot[size_t(p) * L + l] = sl;
    }
  }
  if (pass == 0 && l0 == 0) {
    fcnt_new[p] = fn;
    fcnt_old[p] = R - fn;
  }
}

// Reverse-list sizes after sampling (min(size, cap)); the lists above cap
// are appended to big[] for k_sample.  One thread per node, both kinds.
__global__ void k_rev_sizes(uint32_t n, const int32_t* __restrict__ roffn, const int32_t* __restrict__ roffo, int cap,
                            uint32_t* __restrict__ rszn, uint32_t* __restrict__ rszo, int32_t* __restrict__ bign,
                            int32_t* __restrict__ bigo, int* __restrict__ nbig) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int c = roffn[p + 1] - roffn[p];
  rszn[p] = uint32_t(c < cap ? c : cap);
  if (c > cap) bign[atomicAdd(&nbig[0], 1)] = int32_t(p);
  c = roffo[p + 1] - roffo[p];
  rszo[p] = uint32_t(c < cap ? c : cap);
  if (c > cap) bigo[atomicAdd(&nbig[1], 1)] = int32_t(p);
}

// Reverse lists above `cap` (listed in big[0, *nbig)) keep the `cap` members
// with the smallest generator value (hash of iteration, owner, member); one
// workgroup per list.
__global__ void __launch_bounds__(kBlock) k_sample(const int32_t* __restrict__ big, const int* __restrict__ nbig,
                                                   const int32_t* __restrict__ roff, uint32_t* __restrict__ rids,
                                                   int cap, uint64_t seed, uint64_t stream, uint64_t iter, int chunk) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  uint64_t* k = reinterpret_cast<uint64_t*>(smem_raw);
  uint32_t* v = reinterpret_cast<uint32_t*>(k + chunk);
  const int tid = threadIdx.x;
  const int nb = *nbig;
  for (int bi = blockIdx.x; bi < nb; bi += gridDim.x) {
    const uint32_t o = uint32_t(big[bi]);
    const int c0 = roff[o], c1 = roff[o + 1];
    int cur = 0;
    for (int base = c0; base < c1;) {
      int take = c1 - base;
      if (take > chunk - cur) take = chunk - cur;
      for (int i = tid; i < take; i += kBlock) {
        uint32_t x = rids[base + i];
        k[cur + i] = hash4(seed, stream, (iter << 32) | o, x);
        v[cur + i] = x;
      }
      base += take;
      int m = cur + take;
      int p2 = next_pow2(m);
      for (int i = m + tid; i < p2; i += kBlock) {
        k[i] = kKeyMax;
        v[i] = 0xffffffffu;
      }
      __syncthreads();
      for (int kk = 2; kk <= p2; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
          for (int i = tid; i < p2; i += kBlock) {
            int ixj = i ^ j;
            if (ixj > i) {
              uint64_t ka = k[i], kb = k[ixj];
              uint32_t va = v[i], vb = v[ixj];
              bool gt = (ka > kb) || (ka == kb && va > vb);
              bool up = (i & kk) == 0;
              if (gt == up) {
                k[i] = kb;
                k[ixj] = ka;
                v[i] = vb;
                v[ixj] = va;
              }
            }
          }
          __syncthreads();
        }
      }
      cur = cap;
    }
    for (int i = tid; i < cap; i += kBlock) rids[c0 + i] = v[i];
    __syncthreads();
  }
}

__global__ void k_list_sizes(uint32_t n, const uint32_t* __restrict__ fnew, const uint32_t* __restrict__ rnew,
                             const uint32_t* __restrict__ fold, const uint32_t* __restrict__ rold,
                             int32_t* __restrict__ snew, int32_t* __restrict__ sold) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  snew[p] = int32_t(fnew[p] + rnew[p]);
  sold[p] = int32_t(fold[p] + rold[p]);
}

// Final join lists: forward part from the pool (new flags cleared) in pool
// order, then the sampled reverse part; a group of G lanes per node.
template <int G>
__global__ void k_list_fill(uint32_t n, int L, const uint64_t* __restrict__ pool, uint8_t* __restrict__ flag,
                            const uint32_t* __restrict__ rng, const int32_t* __restrict__ onew,
                            const int32_t* __restrict__ oold, const int32_t* __restrict__ roff_new,
                            const uint32_t* __restrict__ rnew_ids, const uint32_t* __restrict__ rnew_sz,
                            const int32_t* __restrict__ roff_old, const uint32_t* __restrict__ rold_ids,
                            const uint32_t* __restrict__ rold_sz, uint64_t* __restrict__ new_off,
                            uint32_t* __restrict__ new_ids, uint64_t* __restrict__ old_off,
                            uint32_t* __restrict__ old_ids) {
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint32_t p = uint32_t(t / G);
  const int l0 = int(t % G), lane = threadIdx.x & 63;
  if (p > n) return;
  if (l0 == 0) {
    new_off[p] = uint64_t(onew[p]);
    old_off[p] = uint64_t(oold[p]);
  }
  if (p == n) return;
  const uint64_t seg = group_mask<G>();
  const uint64_t below = seg & ((1ull << lane) - 1);
  uint32_t wn = uint32_t(onew[p]), wo = uint32_t(oold[p]);
  const uint32_t R = rng[p];
  for (int lb = 0; lb < int(R); lb += G) {
    const int l = lb + l0;
    const bool act = uint32_t(l) < R;
    bool nw = false;
    uint32_t o = 0;
    if (act) {
      o = key_id(pool[size_t(p) * L + l]);
      nw = flag[size_t(p) * L + l] != 0;
    }
    const uint64_t mn = __ballot(act && nw) & seg, mo = __ballot(act && !nw) & seg;
    if (act) {
      if (nw) {
        new_ids[wn + uint32_t(__popcll(mn & below))] = o;
        flag[size_t(p) * L + l] = 0;
      } else {
        old_ids[wo + uint32_t(__popcll(mo & below))] = o;
      }
    }
    wn += uint32_t(__popcll(mn));
    wo += uint32_t(__popcll(mo));
  }
  const uint32_t rn = rnew_sz[p], ro = rold_sz[p];
  for (uint32_t i = uint32_t(l0); i < rn; i += G) new_ids[wn + i] = rnew_ids[size_t(roff_new[p]) + i];
  for (uint32_t i = uint32_t(l0); i < ro; i += G) old_ids[wo + i] = rold_ids[size_t(roff_old[p]) + i];
}

__global__ void k_graph_out(uint32_t n, int L, int K, const uint64_t* __restrict__ pool,
                            const uint32_t* __restrict__ fill, uint64_t* __restrict__ goff,
                            uint32_t* __restrict__ gids) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > n) return;
  goff[p] = uint64_t(p) * K;
  if (p == n) return;
  uint32_t f = fill[p];
  for (int j = 0; j < K; ++j) {
    uint32_t jj = uint32_t(j) < f ? uint32_t(j) : f - 1;
    gids[size_t(p) * K + j] = key_id(pool[size_t(p) * L + jj]);
  }
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int grid_for(size_t n) { return int((n + 255) / 256); }

}  // namespace

// The split rows of the build's precision form (made once per build).
int ensure_split(efg_index* ix, hipStream_t st) {
  if (ix->b_precision == 0) return EFG_OK;
  if (ix->split_hi.p && ix->split_kind == ix->b_precision) return EFG_OK;
  const int Dk = int((ix->d + 63) / 64 * 64);  // whole 64-column chunks (zero padded)
  if (!ix->split_hi.alloc(ix->n * size_t(Dk)) || !ix->split_lo.alloc(ix->n * size_t(Dk)) ||
      !ix->split_norm.alloc(ix->n))
    return fail(EFG_ERR_MEMORY, "out of device memory (split rows)");
  int dev = 0, ncu = 256;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const unsigned grid = unsigned(std::min<size_t>((ix->n + 3) / 4, size_t(ncu) * 16));  // a wavefront per row
  DevBuf<int> stat;
  if (!stat.alloc(3)) return fail(EFG_ERR_MEMORY, "out of device memory (split rows)");
  hipMemsetAsync(stat.p, 0, 3 * sizeof(int), st);
  if (ix->b_precision == 2)
    k_pack_split<true><<<grid, 256, 0, st>>>(ix->data.p, ix->n, int(ix->d), int(ix->D), Dk, ix->split_hi.p,
                                             ix->split_lo.p, ix->split_norm.p, stat.p);
  else
    k_pack_split<false><<<grid, 256, 0, st>>>(ix->data.p, ix->n, int(ix->d), int(ix->D), Dk, ix->split_hi.p,
                                              ix->split_lo.p, ix->split_norm.p, stat.p);
  int h[3] = {1, 1, 0x7f800000};
  hipMemcpyAsync(h, stat.p, sizeof h, hipMemcpyDeviceToHost, st);
  if (int rc = check_hip(hipStreamSynchronize(st), "split rows")) return rc;
  // rows exact in the 16-bit type leave every low part zero: the tiles then
  // neither read nor multiply them
  ix->split_lo_used = h[0] != 0;
  // integer rows exact in the 16-bit type with squared norms below 2^22: every
  // partial sum of a dot product is an integer below 2^22 (Cauchy-Schwarz),
  // every distance an integer below (|x| + |y|)^2 < 2^24, so the tiles'
  // |x|^2 + |y|^2 - 2 x.y and the FP32 difference form (partial sums at most
  // the distance) are both exact: equal keys, no re-rank and no id dedup
  ix->split_exact = h[0] == 0 && h[1] == 0 && __builtin_bit_cast(float, h[2]) < 4194304.f;
  ix->split_Dk = Dk;
  ix->split_kind = ix->b_precision;
  return check_hip(hipGetLastError(), "split rows");
}

namespace {

void release_split(efg_index* ix) {
  ix->split_hi.reset();
  ix->split_lo.reset();
  ix->split_norm.reset();
  ix->split_kind = 0;
}

}  // namespace

int gpu_merge_init(efg_index* ix) {
  const efg_build_params& p = ix->bp;
  if (ix->T == 0) return fail(EFG_ERR_STATE, "no trees: build trees first");
  const int TB = p.trees_build > 0 ? p.trees_build : p.trees;
  if (TB > ix->T) return fail(EFG_ERR_ARG, "invalid number of trees for the graph");
  const uint32_t n = uint32_t(ix->n);
  const int S = p.S, L = p.L, D = int(ix->D);
  if (n <= uint32_t(2 * S + 1)) return fail(EFG_ERR_ARG, "too few points for the requested S");
  if (S > L) return fail(EFG_ERR_ARG, "S must not exceed L");
  if (S > 64) return fail(EFG_ERR_ARG, "S above 64 is not supported");
  if (L > 1024) return fail(EFG_ERR_ARG, "L above 1024 is not supported");
  hipStream_t st = ix->stream;
  const int M = int(ix->M);
  BuildState& bs = ix->bs;
  bs.ready = false;
  DevBuf<int8_t> depth;
  DevBuf<int32_t> tree_of, leaf_at, boff, sums, scal;
  DevBuf<uint32_t> cnt, cursor, hsize, ecount, extra;
  DevBuf<uint64_t> hkey, rec;
  DevBuf<unsigned long long> nd;
  if (!depth.alloc(M) || !tree_of.alloc(M) || !leaf_at.alloc(size_t(ix->T) * n) || !boff.alloc(M + 1) ||
      !scal.alloc(4) || !cnt.alloc(M) || !cursor.alloc(M) || !hsize.alloc(n) || !ecount.alloc(n) ||
      !extra.alloc(size_t(n) * 2 * S) || !hkey.alloc(size_t(n) * S) || !nd.alloc(2))
    return fail(EFG_ERR_MEMORY, "out of device memory (merge)");
  int init4[4] = {127, 0, 0, 0};
  hipMemcpyAsync(scal.p, init4, 16, hipMemcpyHostToDevice, st);
  k_node_depth<<<grid_for(M), 256, 0, st>>>(ix->parent.p, M, ix->roots.p, ix->T, depth.p, tree_of.p, scal.p + 2);
  k_leaf_at<<<grid_for(M), 256, 0, st>>>(ix->node4.p, ix->range.p, depth.p, tree_of.p, M, n, TB, p.merge_level,
                                         leaf_at.p, scal.p, scal.p + 3);
  hipMemsetAsync(hsize.p, 0, size_t(n) * 4, st);
  hipMemsetAsync(ecount.p, 0, size_t(n) * 4, st);
  hipMemsetAsync(nd.p, 0, 16, st);
  int h4[4] = {127, 0, 0, 0};
  hipMemcpyAsync(h4, scal.p, 16, hipMemcpyDeviceToHost, st);
  if (int rc = check_hip(hipStreamSynchronize(st), "merge")) return rc;
  const int err_depth = h4[0], max_depth = h4[2], LM = h4[3];
  // distance tiles: per wavefront (LM + PB) rows of one chunk of DC columns
  // in LDS, PB probers per batch with PB x LM <= 256 pairs
  const int DCm = std::min(D, 64);
  const int PB = std::max(1, std::min(64, 256 / std::max(1, LM)));
  const size_t wave_lds = size_t(LM + PB) * (DCm + 4) * 4;
  const int tile_waves = int(std::min<size_t>(4, (64 * 1024) / std::max<size_t>(wave_lds, 1)));
  // precision form: leaf tiles on MFMA (leaves of at most 64 members)
  const bool mfma_leaf = ix->b_precision > 0 && LM > 0 && LM <= 64;
  if (int rc = ensure_split(ix, st)) return rc;
  const bool tiles = LM > 0 && (tile_waves >= 1 || mfma_leaf);
  const int nlev = std::max(1, max_depth - p.merge_level + 1);
  DevBuf<float> mdist;
  DevBuf<uint2> pslot;
  if (tiles) {
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_leaf_dist), hipFuncAttributeMaxDynamicSharedMemorySize,
                        int(wave_lds * tile_waves));
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_replay_dist), hipFuncAttributeMaxDynamicSharedMemorySize,
                        int(size_t(kBlock) * S * 8));
    if (!pslot.alloc(size_t(n) * nlev)) return fail(EFG_ERR_MEMORY, "out of device memory (merge)");
  }
  const size_t lds = size_t(kBlock) * S * 8;
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_replay), hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sort_big), hipFuncAttributeMaxDynamicSharedMemorySize, 8192 * 8);
  DevBuf<int32_t> bigl, lam_s;
  DevBuf<int> nbig;
  if (!bigl.alloc(M) || !nbig.alloc(1) || !lam_s.alloc(size_t(n) * nlev))
    return fail(EFG_ERR_MEMORY, "out of device memory (merge)");
  int ncu_m = 256;
  {
    int dv = 0;
    hipGetDevice(&dv);
    hipDeviceGetAttribute(&ncu_m, hipDeviceAttributeMultiprocessorCount, dv);
  }
  // probes descend over the rows the trees were built from: the exact 16-bit
  // rows of the precision form, else FP32
  const bool rows16 = ix->b_precision > 0 && !ix->split_lo_used;
  const void* PX = rows16 ? static_cast<const void*>(ix->split_hi.p) : static_cast<const void*>(ix->data.p);
  const int PD = rows16 ? ix->split_Dk : D;
  const auto probe = !rows16 ? &k_probe<0> : (ix->b_precision == 2 ? &k_probe<2> : &k_probe<1>);
  for (int t = 0; t < TB; ++t) {
    const uint32_t* lt = ix->leaf.p + size_t(t) * n;
    const int32_t* la = leaf_at.p + size_t(t) * n;
    hipMemsetAsync(cnt.p, 0, size_t(M) * 4, st);
    probe<<<grid_for(n), 256, 0, st>>>(PX, PD, ix->node4.p, ix->parent.p, ix->range.p, depth.p, lt, la, n,
                                         p.merge_level, 0, cnt.p, nullptr, nullptr, nullptr, lam_s.p, nlev);
    if (int rc = device_exclusive_scan(reinterpret_cast<int32_t*>(cnt.p), M, boff.p, scal.p + 1, sums, st)) return rc;
    int total = 0;
    hipMemcpyAsync(&total, scal.p + 1, 4, hipMemcpyDeviceToHost, st);
    if (int rc = check_hip(hipStreamSynchronize(st), "merge")) return rc;
    if (!rec.ensure(size_t(total) + 1)) return fail(EFG_ERR_MEMORY, "out of device memory (merge records)");
    hipMemsetAsync(cursor.p, 0, size_t(M) * 4, st);
    probe<<<grid_for(n), 256, 0, st>>>(PX, PD, ix->node4.p, ix->parent.p, ix->range.p, depth.p, lt, la, n,
                                         p.merge_level, 1, cnt.p, boff.p, cursor.p, rec.p, lam_s.p, nlev);
    hipMemsetAsync(nbig.p, 0, sizeof(int), st);
    k_sort_buckets<<<unsigned(std::min<size_t>((M + 3) / 4, 65536)), kBlock, 0, st>>>(boff.p, M, rec.p, bigl.p,
                                                                                       nbig.p);
    k_sort_big<<<unsigned(ncu_m * 2), kBlock, size_t(8192) * 8, st>>>(boff.p, bigl.p, nbig.p, rec.p);
    if (tiles) {
      if (!mdist.ensure(size_t(total) * LM + 1)) return fail(EFG_ERR_MEMORY, "out of device memory (merge distances)");
      if (mfma_leaf) {
        const bool f16 = ix->b_precision == 2, lo = ix->split_lo_used;
        auto kern = f16 ? (lo ? k_leaf_mfma<true, true> : k_leaf_mfma<true, false>)
                        : (lo ? k_leaf_mfma<false, true> : k_leaf_mfma<false, false>);
        kern<<<unsigned(ncu_m * 8), kBlock, 0, st>>>(ix->split_hi.p, ix->split_lo.p, ix->split_norm.p, ix->split_Dk,
                                                     ix->range.p, lt, boff.p, M, rec.p, LM, p.merge_level, nlev,
                                                     mdist.p, pslot.p);
      } else {
        k_leaf_dist<<<unsigned(ncu_m * 8), 64 * tile_waves, wave_lds * tile_waves, st>>>(
            ix->data.p, D, DCm, ix->range.p, lt, boff.p, M, rec.p, LM, PB, p.merge_level, nlev, mdist.p, pslot.p);
      }
      k_replay_dist<<<grid_for(n), kBlock, lds, st>>>(ix->node4.p, ix->parent.p, ix->range.p, depth.p, lt, la, n,
                                                      p.merge_level, S, boff.p, rec.p, mdist.p, pslot.p, LM, nlev,
                                                      hkey.p, hsize.p, extra.p, ecount.p, nd.p, lam_s.p);
    } else {
      k_replay<<<grid_for(n), kBlock, lds, st>>>(ix->data.p, D, ix->node4.p, ix->parent.p, ix->range.p, depth.p, lt,
                                                 la, n, p.merge_level, S, boff.p, rec.p, hkey.p, hsize.p, extra.p,
                                                 ecount.p, nd.p);
    }
  }
  // precision form: the merged sets get their FP32 distances back
  if (mfma_leaf && !ix->split_exact) {
    const size_t m = size_t(n) * S * 2;
    k_rerank_pool<<<unsigned((m + 255) / 256), 256, 0, st>>>(ix->data.p, D, n, S, S,
                                                             ix->leaf0_perm ? ix->leaf.p : nullptr, hkey.p, hsize.p);
    k_sort_pools<false><<<unsigned(std::min<size_t>((n + 3) / 4, size_t(ncu_m) * 16)), kBlock, 0, st>>>(
        n, S, S, hkey.p, nullptr, hsize.p);
  }
  // init
  bs.L = L;
  if (!bs.pool.alloc(size_t(n) * L) || !bs.flag.alloc(size_t(n) * L) || !bs.fill.alloc(n) || !bs.range.alloc(n) ||
      !bs.radiusM.alloc(n) || !bs.found.alloc(n) || !bs.new_off.alloc(n + 1) || !bs.new_ids.alloc(size_t(n) * 2 * S) ||
      !bs.old_off.alloc(n + 1) || !bs.old_ids.alloc(1))
    return fail(EFG_ERR_MEMORY, "out of device memory (pools)");
  k_init<<<grid_for(n), 256, 0, st>>>(ix->data.p, D, n, S, L, p.seed, hkey.p, hsize.p, extra.p, ecount.p, bs.pool.p,
                                      bs.flag.p, bs.fill.p, bs.range.p, bs.new_ids.p, nd.p + 1);
  {
    std::vector<uint64_t> off(n + 1);
    for (uint32_t i = 0; i <= n; ++i) off[i] = uint64_t(i) * 2 * S;
    hipMemcpyAsync(bs.new_off.p, off.data(), off.size() * 8, hipMemcpyHostToDevice, st);
    hipMemsetAsync(bs.old_off.p, 0, size_t(n + 1) * 8, st);
    unsigned long long h_nd[2] = {0, 0};
    hipMemcpyAsync(h_nd, nd.p, 16, hipMemcpyDeviceToHost, st);
    if (int rc = check_hip(hipStreamSynchronize(st), "init")) return rc;
    ix->counters[1] = h_nd[0];
    ix->counters[2] = h_nd[1];
  }
  bs.new_n = size_t(n) * 2 * S;
  bs.old_n = 0;
  bs.iter = 0;
  bs.keep_frac = 1.0;
  bs.ready = true;
  ix->counters[3] = 0;
  ix->counters[6] = 0;
  ix->counters[7] = 0;
  ix->join_pairs.clear();
  if (err_depth < p.merge_level) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "merge level deeper than tree, max merge deepth is %d", err_depth - 1);
    bs.ready = false;
    return fail(EFG_ERR_STATE, buf);
  }
  return EFG_OK;
}

int gpu_refine_iter(efg_index* ix) {
  BuildState& bs = ix->bs;
  if (!bs.ready) return fail(EFG_ERR_STATE, "no build state: run the merge first");
  const efg_build_params& p = ix->bp;
  const uint32_t n = uint32_t(ix->n);
  const int L = bs.L, S = p.S, D = int(ix->D);
  hipStream_t st = ix->stream;
  double t0 = now_ms();
  BuildState::IterScratch& sc = bs.it;
  DevBuf<uint64_t>& worst = sc.worst;
  DevBuf<uint32_t>&tcount = sc.tcount, &cursor = sc.cursor;
  DevBuf<int32_t>&toff = sc.toff, &sums = sc.sums, &scal = sc.scal;
  DevBuf<unsigned long long>& ctr = sc.ctr;
  if (!worst.ensure(n) || !tcount.ensure(n) || !cursor.ensure(n) || !toff.ensure(n + 1) || !scal.ensure(4) ||
      !ctr.ensure(6))
    return fail(EFG_ERR_MEMORY, "out of device memory (join)");
  int dev = 0, ncu = 256;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  k_worst<<<grid_for(n), 256, 0, st>>>(bs.pool.p, bs.fill.p, n, L, worst.p);
  hipMemsetAsync(ctr.p, 0, 24, st);
  DevBuf<int32_t>&tcnt = sc.tcnt, &tskoff = sc.tskoff;
  if (!tcnt.ensure(n) || !tskoff.ensure(n + 1)) return fail(EFG_ERR_MEMORY, "out of device memory (join tasks)");
  // tree 0's leaf list (a permutation of the nodes) orders the tasks
  const uint32_t* order = ix->leaf0_perm ? ix->leaf.p : nullptr;
  // nodes of at most 64 (precision form) or kNodeRows (FP32) rows go to the
  // whole-node kernels, the others to the block tasks
  const uint32_t node_rows = ix->b_precision > 0 ? 64u : uint32_t(kNodeRows);
  k_join_count<<<unsigned(std::min<size_t>(grid_for(n), size_t(ncu) * 4)), kBlock, 0, st>>>(
      n, order, bs.new_off.p, bs.old_off.p, node_rows, tcnt.p, ctr.p);
  if (int rc = device_exclusive_scan(tcnt.p, int(n), tskoff.p, scal.p + 2, sums, st)) return rc;
  unsigned long long bound = 0, rows = 0;
  int ntask = 0;
  hipMemcpyAsync(&bound, ctr.p, 8, hipMemcpyDeviceToHost, st);
  hipMemcpyAsync(&rows, ctr.p + 2, 8, hipMemcpyDeviceToHost, st);
  hipMemcpyAsync(&ntask, scal.p + 2, 4, hipMemcpyDeviceToHost, st);
  if (int rc = check_hip(hipStreamSynchronize(st), "join tasks")) return rc;
  ix->counters[6] += rows;
  // eight task ranges, one per XCD: a multiple of eight workgroups
  const size_t join_grid = size_t((ncu * 4 + 7) / 8) * 8;
  const size_t node_grid = size_t(ncu) * 8;  // whole-node FP32 kernel: workgroups of four wavefronts
  if (!bs.jhdr.ensure(size_t(ntask) + 1) || !bs.jtids.ensure((size_t(ntask) + 1) * 64))
    return fail(EFG_ERR_MEMORY, "out of device memory (join tasks)");
  if (ntask)
    k_join_tasks<<<unsigned(std::min<size_t>((size_t(n) + 255) / 256, size_t(ncu) * 4)), kBlock, 0, st>>>(
        n, order, bs.new_off.p, bs.new_ids.p, bs.old_off.p, bs.old_ids.p, tcnt.p, tskoff.p, bs.jhdr.p, bs.jtids.p);
  // one chunk of at most 64 columns of both blocks' rows in LDS
  const int DC = std::min(D, 64);
  const size_t lds_join = size_t(64) * (DC + 4) * 4;
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_join_tiles<4, false>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_join));
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_join_tiles<4, true>),
                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_join));
  if (ix->b_precision > 0)
    if (int rc = ensure_split(ix, st)) return rc;
  // The join kernels: whole nodes, then block tasks.  kD: direct routing
  // (candidates straight into buckets at toff), else staged.  Counters:
  // nst staged slots (direct: candidates kept), cnt per-target counts
  // (direct: bucket fill), np pairs.
  auto run_nodes = [&](auto direct_c, Stage* stage, unsigned long long cap, unsigned long long* nst, uint32_t* cnt,
                       unsigned long long* np) {
    constexpr bool kD = decltype(direct_c)::value;
    const int32_t* boff = kD ? toff.p : nullptr;
    uint64_t* bk = kD ? bs.jbkey.p : nullptr;
    uint32_t* bsr = kD ? bs.jbsrc.p : nullptr;
    if (ix->b_precision > 0) {
      // precision form on MFMA: one wavefront per node of at most 32, then
      // of 33 to 64 rows
      const unsigned g16 = unsigned(size_t(ncu) * 8);
      const bool f16 = ix->b_precision == 2;
      const bool lo = ix->split_lo_used;
      (f16 ? (lo ? k_join_node_mfma<true, true, 1, kD> : k_join_node_mfma<true, false, 1, kD>)
           : (lo ? k_join_node_mfma<false, true, 1, kD> : k_join_node_mfma<false, false, 1, kD>))<<<g16, kBlock, 0,
                                                                                                   st>>>(
          ix->split_hi.p, ix->split_lo.p, ix->split_norm.p, ix->split_Dk, n, order, bs.new_off.p, bs.new_ids.p,
          bs.old_off.p, bs.old_ids.p, worst.p, stage, cap, nst, cnt, np, boff, bk, bsr);
      (f16 ? (lo ? k_join_node_mfma<true, true, 3, kD> : k_join_node_mfma<true, false, 3, kD>)
           : (lo ? k_join_node_mfma<false, true, 3, kD> : k_join_node_mfma<false, false, 3, kD>))<<<g16, kBlock, 0,
                                                                                                   st>>>(
          ix->split_hi.p, ix->split_lo.p, ix->split_norm.p, ix->split_Dk, n, order, bs.new_off.p, bs.new_ids.p,
          bs.old_off.p, bs.old_ids.p, worst.p, stage, cap, nst, cnt, np, boff, bk, bsr);
    } else {
      k_join_node_f32<kD><<<unsigned(node_grid), kBlock, 0, st>>>(ix->data.p, D, n, order, bs.new_off.p,
                                                                  bs.new_ids.p, bs.old_off.p, bs.old_ids.p, worst.p,
                                                                  stage, cap, nst, cnt, np, boff, bk, bsr);
    }
  };
  // block tasks of the larger nodes
  auto run_tasks = [&](auto direct_c, Stage* stage, unsigned long long cap, unsigned long long* nst, uint32_t* cnt,
                       unsigned long long* np) {
    constexpr bool kD = decltype(direct_c)::value;
    if (!ntask) return;
    if (ix->b_precision > 0) {
      const unsigned g16 = unsigned(size_t(ncu) * 8);
      const bool f16 = ix->b_precision == 2;
      const bool lo = ix->split_lo_used;
      (f16 ? (lo ? k_join_mfma<true, true, kD> : k_join_mfma<true, false, kD>)
           : (lo ? k_join_mfma<false, true, kD> : k_join_mfma<false, false, kD>))<<<g16, kBlock, 0, st>>>(
          ix->split_hi.p, ix->split_lo.p, ix->split_norm.p, ix->split_Dk, uint32_t(ntask), bs.jhdr.p, bs.jtids.p,
          worst.p, stage, cap, nst, cnt, np, kD ? toff.p : nullptr, kD ? bs.jbkey.p : nullptr,
          kD ? bs.jbsrc.p : nullptr);
    } else {
      k_join_tiles<4, kD><<<unsigned(join_grid), kBlock, lds_join, st>>>(
          ix->data.p, D, uint32_t(ntask), bs.jhdr.p, bs.jtids.p, worst.p, stage, cap, nst, cnt, np,
          kD ? toff.p : nullptr, kD ? bs.jbkey.p : nullptr, kD ? bs.jbsrc.p : nullptr);
    }
  };
  auto run_tasks_staged = [&](Stage* stage, unsigned long long c, unsigned long long* nst, uint32_t* cnt,
                              unsigned long long* np) { run_tasks(std::false_type{}, stage, c, nst, cnt, np); };
  // Direct routing when every target's room (the pairs it is in, 2 x bound
  // slots in all) fits 31-bit offsets and memory; else staging.  The MFMA
  // block tasks of rows above 128 columns stay staged (their candidates then
  // appended to the same buckets by the scatter): routed directly, their
  // scattered writes made that short kernel up to a third slower (1M x 960),
  // more than the staging round trip costs.
  const bool direct = ix->b_join_direct && 2 * bound + 1 < (1ull << 31) &&
                      bs.jbkey.ensure(size_t(2 * bound) + 1) && bs.jbsrc.ensure(size_t(2 * bound) + 1);
  const bool tasks_direct = ix->b_precision == 0 || ix->split_Dk <= 128;
  // staging: the kept candidates (two per pair at most) plus one partly used
  // run per workgroup (per wavefront on the MFMA path, ncu * 8 workgroups of
  // four: three kernels staged, or the block tasks' alone with direct
  // routing).  It is sized from the last iteration's kept fraction with a
  // margin; when the candidates do not fit, the staged kernels (which do not
  // change the pools) run again with room for every reservation counted.
  const size_t slack = ix->b_precision > 0
                           ? size_t(direct ? 1 : 3) * size_t(ncu) * 8 * (kBlock / 64) * kStageRun + 1024
                           : (join_grid + node_grid * (kBlock / 64)) * kStageRun + 1024;
  const size_t full = size_t(2 * bound) + slack;
  size_t cap = std::min(full, size_t(double(2 * bound) * std::min(1.0, ix->b_stage_margin * bs.keep_frac)) + slack);
  // Stage the candidates of `run` (the join kernels on the staging counters
  // sc[0], sc[1] and per-target counts cnt), joining again with room for
  // every reservation when they do not fit; h gets the slots used and pairs.
  auto staged = [&](auto run, unsigned long long* sc, uint32_t* cnt, unsigned long long* h) -> int {
    if (!bs.jstage.ensure(cap * sizeof(Stage))) return fail(EFG_ERR_MEMORY, "out of device memory (join staging)");
    for (int attempt = 0;; ++attempt) {
      hipMemsetAsync(sc, 0, 16, st);
      hipMemsetAsync(cnt, 0, size_t(n) * 4, st);
      run(reinterpret_cast<Stage*>(bs.jstage.p), (unsigned long long)cap, sc, cnt, sc + 1);
      hipMemcpyAsync(h, sc, 16, hipMemcpyDeviceToHost, st);
      if (int rc = check_hip(hipStreamSynchronize(st), "join")) return rc;
      // every candidate reserved a slot, written or not: a second run fits in
      // the reservations counted plus one partly used run per workgroup
      if (h[0] <= cap) return EFG_OK;
      if (attempt > 0 || cap >= full) return fail(EFG_ERR_STATE, "join staging overflow");
      cap = std::min(full, size_t(h[0]) + slack);
      ix->counters[7] += 1;
      if (!bs.jstage.ensure(cap * sizeof(Stage)))
        return fail(EFG_ERR_MEMORY, "out of device memory (join staging)");
    }
  };
  unsigned long long h_ctr[2] = {0, 0}, h_task[2] = {0, 0};
  if (direct) {
    hipMemsetAsync(cursor.p, 0, size_t(n) * 4, st);
    k_bucket_room<<<unsigned((size_t(n) + kRoomNodes - 1) / kRoomNodes), kBlock, 0, st>>>(
        n, order, bs.new_off.p, bs.new_ids.p, bs.old_off.p, bs.old_ids.p, cursor.p);
    if (int rc = device_exclusive_scan(reinterpret_cast<int32_t*>(cursor.p), int(n), toff.p, scal.p + 3, sums, st))
      return rc;
    hipMemsetAsync(ctr.p, 0, 16, st);
    hipMemsetAsync(tcount.p, 0, size_t(n) * 4, st);
    run_nodes(std::true_type{}, nullptr, 0ull, ctr.p, tcount.p, ctr.p + 1);
    if (tasks_direct) {
      run_tasks(std::true_type{}, nullptr, 0ull, ctr.p, tcount.p, ctr.p + 1);
    } else if (ntask) {
      // the staged tasks' own per-target counts go to `cursor` (unused here)
      if (int rc = staged(run_tasks_staged, ctr.p + 3, cursor.p, h_task)) return rc;
      if (h_task[0])
        k_join_scatter<<<unsigned((h_task[0] + 255) / 256), 256, 0, st>>>(
            reinterpret_cast<Stage*>(bs.jstage.p), h_task[0], toff.p, tcount.p, bs.jbkey.p, bs.jbsrc.p);
    }
  } else {
    auto run_all = [&](Stage* stage, unsigned long long c, unsigned long long* nst, uint32_t* cnt,
                       unsigned long long* np) {
      run_nodes(std::false_type{}, stage, c, nst, cnt, np);
      run_tasks(std::false_type{}, stage, c, nst, cnt, np);
    };
    if (int rc = staged(run_all, ctr.p, tcount.p, h_ctr)) return rc;
    const unsigned long long nst = h_ctr[0];
    Stage* stage = reinterpret_cast<Stage*>(bs.jstage.p);
    ix->join_pairs.push_back(h_ctr[1]);
    ix->counters[3] += h_ctr[1];
    if (nst >= (1ull << 31)) return fail(EFG_ERR_STATE, "too many join candidates for one iteration");
    if (int rc = device_exclusive_scan(reinterpret_cast<int32_t*>(tcount.p), int(n), toff.p, scal.p, sums, st))
      return rc;
    int nkept = 0;
    hipMemcpyAsync(&nkept, scal.p, 4, hipMemcpyDeviceToHost, st);
    if (int rc = check_hip(hipStreamSynchronize(st), "join")) return rc;
    bs.keep_frac = bound ? double(nkept) / double(2 * bound) : 1.0;
    if (!bs.jbkey.ensure(size_t(nkept) + 1) || !bs.jbsrc.ensure(size_t(nkept) + 1))
      return fail(EFG_ERR_MEMORY, "out of device memory (buckets)");
    hipMemsetAsync(cursor.p, 0, size_t(n) * 4, st);
    if (nst)
      k_join_scatter<<<unsigned((nst + 255) / 256), 256, 0, st>>>(stage, nst, toff.p, cursor.p, bs.jbkey.p,
                                                                  bs.jbsrc.p);
  }
  hipMemsetAsync(bs.found.p, 0, n, st);
  {
    // one wavefront per target; the LDS window holds at least 2L entries
    const bool byid = ix->b_precision > 0 && !ix->split_exact;  // approximate keys
    const unsigned grid = unsigned(std::min<size_t>((n + 3) / 4, size_t(ncu) * 8));
    const uint32_t* tc = direct ? tcount.p : nullptr;
    uint64_t* bkey = bs.jbkey.p;
    uint32_t* bsrc = bs.jbsrc.p;
    if (L <= 512)
      (L <= 64 ? (byid ? k_merge_reg<64, true> : k_merge_reg<64, false>)
               : L <= 256 ? (byid ? k_merge_reg<256, true> : k_merge_reg<256, false>)
                          : (byid ? k_merge_reg<512, true> : k_merge_reg<512, false>))<<<grid, kBlock, 0, st>>>(
          n, L, p.check, toff.p, tc, bkey, bsrc, bs.pool.p, bs.flag.p, bs.fill.p, bs.found.p);
    else if (2 * L <= 256)
      (byid ? k_merge_wave<256, true> : k_merge_wave<256, false>)<<<grid, kBlock, 0, st>>>(
          n, L, p.check, toff.p, tc, bkey, bsrc, bs.pool.p, bs.flag.p, bs.fill.p, bs.found.p);
    else if (2 * L <= 512)
      (byid ? k_merge_wave<512, true> : k_merge_wave<512, false>)<<<grid, kBlock, 0, st>>>(
          n, L, p.check, toff.p, tc, bkey, bsrc, bs.pool.p, bs.flag.p, bs.fill.p, bs.found.p);
    else if (2 * L <= 1024)
      (byid ? k_merge_wave<1024, true> : k_merge_wave<1024, false>)<<<grid * 2, 128, 0, st>>>(
          n, L, p.check, toff.p, tc, bkey, bsrc, bs.pool.p, bs.flag.p, bs.fill.p, bs.found.p);
    else if (byid)
      return fail(EFG_ERR_ARG, "the build's precision form supports pools of up to 512 entries");
    else
      k_merge_wave<2048, false><<<grid * 2, 128, 0, st>>>(n, L, p.check, toff.p, tc, bkey, bsrc, bs.pool.p,
                                                          bs.flag.p, bs.fill.p, bs.found.p);
    if (direct) hipMemcpyAsync(h_ctr, ctr.p, 16, hipMemcpyDeviceToHost, st);
    if (int rc = check_hip(hipStreamSynchronize(st), "join merge")) return rc;
    if (direct) {
      // kept fraction of the directly routed candidates (exact counts)
      ix->join_pairs.push_back(h_ctr[1] + h_task[1]);
      ix->counters[3] += h_ctr[1] + h_task[1];
      bs.keep_frac = h_ctr[1] ? double(h_ctr[0]) / double(2 * h_ctr[1]) : 1.0;
    }
  }
  if (int rc = check_hip(hipStreamSynchronize(st), "join merge")) return rc;
  double t1 = now_ms();
  // ---- update(100)
  bs.iter += 1;
  DevBuf<uint32_t>&fnew = sc.fnew, &fold = sc.fold, &rcn = sc.rcn, &rco = sc.rco, &rszn = sc.rszn, &rszo = sc.rszo,
                   &ridn = sc.ridn, &rido = sc.rido;
  DevBuf<int32_t>&roffn = sc.roffn, &roffo = sc.roffo, &snew = sc.snew, &sold = sc.sold, &onew = sc.onew,
                  &oold = sc.oold;
  if (!fnew.ensure(n) || !fold.ensure(n) || !rcn.ensure(n) || !rco.ensure(n) || !rszn.ensure(n) || !rszo.ensure(n) || !roffn.ensure(n + 1) || !roffo.ensure(n + 1) ||
      !sc.rslot.ensure(size_t(n) * L) ||
      !snew.ensure(n + 1) || !sold.ensure(n + 1) || !onew.ensure(n + 1) || !oold.ensure(n + 1))
    return fail(EFG_ERR_MEMORY, "out of device memory (update)");
  // a group of G lanes per node: the pool length rounded up to a power of
  // two, 16 to 64
  const int G = L <= 16 ? 16 : (L <= 32 ? 32 : 64);
  const unsigned ggrid = unsigned((size_t(n) * G + 255) / 256), ggrid1 = unsigned((size_t(n + 1) * G + 255) / 256);
#define EFG_G(kern) (G == 16 ? kern<16> : (G == 32 ? kern<32> : kern<64>))
  EFG_G(k_range)<<<ggrid, 256, 0, st>>>(bs.pool.p, bs.flag.p, bs.fill.p, bs.found.p, n, L, S, bs.range.p,
                                        bs.radiusM.p);
  hipMemsetAsync(rcn.p, 0, size_t(n) * 4, st);
  hipMemsetAsync(rco.p, 0, size_t(n) * 4, st);
  EFG_G(k_lists)<<<ggrid, 256, 0, st>>>(bs.pool.p, bs.flag.p, bs.range.p, bs.radiusM.p, n, L, 0, fnew.p, fold.p, rcn.p,
                                       rco.p, nullptr, nullptr, sc.rslot.p, nullptr, nullptr);
  if (int rc = device_exclusive_scan(reinterpret_cast<int32_t*>(rcn.p), int(n), roffn.p, scal.p, sums, st)) return rc;
  if (int rc = device_exclusive_scan(reinterpret_cast<int32_t*>(rco.p), int(n), roffo.p, scal.p + 1, sums, st)) return rc;
  int rtot[2] = {0, 0};
  hipMemcpyAsync(rtot, scal.p, 8, hipMemcpyDeviceToHost, st);
  if (int rc = check_hip(hipStreamSynchronize(st), "update")) return rc;
  if (!ridn.ensure(size_t(rtot[0]) + 1) || !rido.ensure(size_t(rtot[1]) + 1))
    return fail(EFG_ERR_MEMORY, "out of device memory (reverse lists)");
  EFG_G(k_lists)<<<ggrid, 256, 0, st>>>(bs.pool.p, bs.flag.p, bs.range.p, bs.radiusM.p, n, L, 1, nullptr, nullptr,
                                       nullptr, nullptr, roffn.p, roffo.p, sc.rslot.p, ridn.p, rido.p);
  const int rcap = 100;
  const int schunk = 4096;
  const size_t lds_s = size_t(schunk) * 12;
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sample), hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_s));
  {
    if (!sc.bign.ensure(n) || !sc.bigo.ensure(n) || !sc.nbig.ensure(2))
      return fail(EFG_ERR_MEMORY, "out of device memory (update)");
    hipMemsetAsync(sc.nbig.p, 0, 2 * sizeof(int), st);
    k_rev_sizes<<<grid_for(n), 256, 0, st>>>(n, roffn.p, roffo.p, rcap, rszn.p, rszo.p, sc.bign.p, sc.bigo.p,
                                             sc.nbig.p);
    unsigned grid = unsigned(std::min<size_t>(n, size_t(ncu) * 32));
    k_sample<<<grid, kBlock, lds_s, st>>>(sc.bign.p, sc.nbig.p, roffn.p, ridn.p, rcap, p.seed, ST_RNN_NEW,
                                          uint64_t(bs.iter), schunk);
    k_sample<<<grid, kBlock, lds_s, st>>>(sc.bigo.p, sc.nbig.p + 1, roffo.p, rido.p, rcap, p.seed, ST_RNN_OLD,
                                          uint64_t(bs.iter), schunk);
  }
  k_list_sizes<<<grid_for(n), 256, 0, st>>>(n, fnew.p, rszn.p, fold.p, rszo.p, snew.p, sold.p);
  if (int rc = device_exclusive_scan(snew.p, int(n), onew.p, scal.p, sums, st)) return rc;
  if (int rc = device_exclusive_scan(sold.p, int(n), oold.p, scal.p + 1, sums, st)) return rc;
  int ltot[2] = {0, 0};
  hipMemcpyAsync(ltot, scal.p, 8, hipMemcpyDeviceToHost, st);
  if (int rc = check_hip(hipStreamSynchronize(st), "update")) return rc;
  if (!bs.new_ids.ensure(size_t(ltot[0]) + 1) || !bs.old_ids.ensure(size_t(ltot[1]) + 1))
    return fail(EFG_ERR_MEMORY, "out of device memory (join lists)");
  EFG_G(k_list_fill)<<<ggrid1, 256, 0, st>>>(n, L, bs.pool.p, bs.flag.p, bs.range.p, onew.p, oold.p, roffn.p,
                                               ridn.p, rszn.p, roffo.p, rido.p, rszo.p, bs.new_off.p, bs.new_ids.p,
                                               bs.old_off.p, bs.old_ids.p);
#undef EFG_G
  bs.new_n = size_t(ltot[0]);
  bs.old_n = size_t(ltot[1]);
  if (int rc = check_hip(hipStreamSynchronize(st), "update")) return rc;
  double t2 = now_ms();
  ix->times[2] += t1 - t0;
  ix->times[3] += t2 - t1;
  return EFG_OK;
}

int gpu_finalize_graph(efg_index* ix) {
  BuildState& bs = ix->bs;
  if (!bs.ready) return fail(EFG_ERR_STATE, "no build state: run the merge first");
  const uint32_t n = uint32_t(ix->n);
  const int K = ix->bp.K;
  if (!ix->goff.alloc(n + 1) || !ix->gids.alloc(size_t(n) * K)) return fail(EFG_ERR_MEMORY, "out of device memory (graph)");
  if (ix->b_precision > 0 && !ix->split_exact) {  // precision form: FP32 re-rank of the first 2K entries of every pool
    const int R = std::min(bs.L, 2 * K);
    const size_t m = size_t(n) * R * 2;
    k_rerank_pool<<<unsigned((m + 255) / 256), 256, 0, ix->stream>>>(
        ix->data.p, int(ix->D), n, bs.L, R, ix->leaf0_perm ? ix->leaf.p : nullptr, bs.pool.p, bs.fill.p);
    k_sort_pools<true><<<unsigned(std::min<size_t>((n + 3) / 4, 256 * 16)), kBlock, 0, ix->stream>>>(
        n, bs.L, R, bs.pool.p, bs.flag.p, bs.fill.p);
  }
  k_graph_out<<<grid_for(n + 1), 256, 0, ix->stream>>>(n, bs.L, K, bs.pool.p, bs.fill.p, ix->goff.p, ix->gids.p);
  if (int rc = check_hip(hipStreamSynchronize(ix->stream), "graph")) return rc;
  release_split(ix);
  bs.jhdr.reset();
  bs.jtids.reset();
  bs.jstage.reset();
  bs.jbkey.reset();
  bs.jbsrc.reset();
  bs.it.release();
  ix->nnz = size_t(n) * K;
  ix->max_deg = K;
  ix->reg_deg = K;
  return EFG_OK;
}

}  // namespace efg

extern "C" {

int efg_get_pools(const efg_index* ix, uint64_t* keys, uint8_t* flags, uint32_t* fill, uint32_t* range) {
  if (!ix) return efg::fail(EFG_ERR_ARG, "null index");
  const efg::BuildState& bs = ix->bs;
  if (!bs.pool.p) return efg::fail(EFG_ERR_STATE, "no build state");
  const size_t n = ix->n, L = size_t(bs.L);
  hipStreamSynchronize(ix->stream);
  int rc = 0;
  if (keys) rc |= efg::check_hip(hipMemcpy(keys, bs.pool.p, n * L * 8, hipMemcpyDeviceToHost), "copy");
  if (flags) rc |= efg::check_hip(hipMemcpy(flags, bs.flag.p, n * L, hipMemcpyDeviceToHost), "copy");
  if (fill) rc |= efg::check_hip(hipMemcpy(fill, bs.fill.p, n * 4, hipMemcpyDeviceToHost), "copy");
  if (range) rc |= efg::check_hip(hipMemcpy(range, bs.range.p, n * 4, hipMemcpyDeviceToHost), "copy");
  return rc ? EFG_ERR_DEVICE : EFG_OK;
}

int efg_get_lists(const efg_index* ix, int which, uint64_t* off, uint32_t* ids, size_t cap) {
  if (!ix) return efg::fail(EFG_ERR_ARG, "null index");
  const efg::BuildState& bs = ix->bs;
  if (!bs.new_off.p) return efg::fail(EFG_ERR_STATE, "no build state");
  hipStreamSynchronize(ix->stream);
  const uint64_t* o = which ? bs.old_off.p : bs.new_off.p;
  const uint32_t* g = which ? bs.old_ids.p : bs.new_ids.p;
  size_t tot = which ? bs.old_n : bs.new_n;
  int rc = 0;
  if (off) rc |= efg::check_hip(hipMemcpy(off, o, (ix->n + 1) * 8, hipMemcpyDeviceToHost), "copy");
  if (ids && cap && tot) rc |= efg::check_hip(hipMemcpy(ids, g, std::min(cap, tot) * 4, hipMemcpyDeviceToHost), "copy");
  return rc ? EFG_ERR_DEVICE : EFG_OK;
}

}  // extern "C"

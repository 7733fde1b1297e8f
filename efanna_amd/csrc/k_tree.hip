// Truncated randomized KD-trees: kdtreeub_index.hpp:1192-1276 (buildTrees,
// identical to the tree part of initGraph 1295-1363), meanSplit /
// planeSplit / selectDivision (902-999), DFSbuild (1138-1167).
//
// Level-synchronous over all trees at once.  Level l holds the active nodes
// (to be split); children of active node i get indices base+2i, base+2i+1.
// The reference's phase rule is kept: while 0 < |active| < 1100 (BFS) every
// active node is split and a child stays active iff its size > S; the first
// level where |active| >= 1100 keeps only nodes larger than TNS and from then
// on (DFS) a node is split iff its size > TNS.
//
// Per level:
//   k_split_stats  a group of threads per active node: mean and
//                  unnormalised variance over the first min(101, count) ids
//                  of the node's range (sequential sums per dimension), top-5
//                  variance dimensions, the choice among them from the
//                  counter-based generator, cut value = mean; a node of at
//                  most one id per thread of its group is partitioned and
//                  split right there.
//   k_part_count / k_part_node / k_part_scatter / k_part_copy
//                  stable three-way partition (<, ==, > cut) of every larger
//                  active range in blocks of 1024 ids: per-block class
//                  counts, a per-node scan that also applies the balancing
//                  rule and creates the children, then the scatter and copy
//                  back.
//   k_next_flags + scan + k_next_compact
//                  children that stay active, in child order.
#include <algorithm>
#include <vector>

#include "efg_device.hpp"
#include "efg_internal.hpp"

namespace efg {

namespace {

constexpr int kPartB = 1024;  // ids per partition block (4 per thread)

__global__ void k_init_perm(uint32_t* __restrict__ leaf, int T, uint32_t n, uint64_t seed) {
  size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= size_t(T) * n) return;
  uint32_t t = uint32_t(gid / n), i = uint32_t(gid % n);
  leaf[gid] = uint32_t(feistel_perm(seed, ST_PERM, t, n, i));
}

__global__ void k_init_roots(int4* node4, int32_t* parent, uint2* range, int32_t* roots, int T, uint32_t n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  node4[t] = make_int4(-1, 0, -1, -1);
  parent[t] = -1;
  range[t] = make_uint2(0, n);
  roots[t] = t;
}

__device__ inline int classify(float v, float cut) { return v < cut ? 0 : (v <= cut ? 1 : 2); }

// Children of active node ni (slot a) after its partition into c0 ids below
// the cut, c1 equal to it and the rest above: the balancing rule
// (kdtreeub_index.hpp:943-950) and the two child nodes.
__device__ inline void make_children(int ni, int a, int32_t base, uint2 r, int4 nd, uint32_t c0, uint32_t c1,
                                     int4* __restrict__ node4, int32_t* __restrict__ parent,
                                     uint2* __restrict__ range, int32_t* __restrict__ tree_of) {
  const unsigned count = r.y - r.x;
  const unsigned lim1 = c0, lim2 = c0 + c1;
  unsigned index;
  if (lim1 > count / 2) index = lim1;
  else if (lim2 < count / 2) index = lim2;
  else index = count / 2;
  if (lim1 == count || lim2 == 0) index = count / 2;
  const int32_t L = base + 2 * a, R = base + 2 * a + 1;
  nd.z = L;
  nd.w = R;
  node4[ni] = nd;
  node4[L] = make_int4(-1, 0, -1, -1);
  node4[R] = make_int4(-1, 0, -1, -1);
  parent[L] = ni;
  parent[R] = ni;
  range[L] = make_uint2(r.x, r.x + index);
  range[R] = make_uint2(r.x + index, r.y);
  tree_of[L] = tree_of[ni];
  tree_of[R] = tree_of[ni];
}

// Split statistics: a group of `tpn` threads (a multiple of 64) per active
// node, `kBlock / tpn` nodes per workgroup.  A thread owns dimensions k,
// k + tpn, ...; the up to 101 sampled values of its dimension are loaded
// kChunk at a time into registers (the row index is clamped, values past the
// sample are never used) and summed in row order as the reference does; the
// variance pass loads them again (from cache).  Row offsets are 32-bit when
// the rows span fewer than 2^30 elements (fewer registers per load in flight;
// the kernel is bound by the latency of its short per-node steps, so it
// needs the occupancy).  The first wavefront of the group picks the top-5 variance
// dimensions by five arg-max rounds over (variance descending, dimension
// ascending), which is the order the reference's insertion keeps.
//
// A node of at most `small` ids (at most one per thread of its group) is
// also partitioned here, in the order k_part_count / k_part_scatter give:
// each thread classifies one id (the row was just read by the statistics),
// wavefront ballots and per-wavefront counts give every id its place in the
// stable three-way order, the ids are written back over the node's range
// and the children are made.  The block partition kernels skip such nodes.
constexpr int kSample = 101;
constexpr int kChunk = 32;
template <bool kOff32>
__device__ __forceinline__ auto row_off(uint32_t id, int D) {
  if constexpr (kOff32) return id * uint32_t(D);
  else return size_t(id) * size_t(D);
}
template <bool kOff32, int kRow>
__global__ void __launch_bounds__(kBlock, 5) k_split_stats(const void* __restrict__ X, int d, int D,
                                                        const int32_t* __restrict__ active, int na,
                                                        uint2* __restrict__ range, int32_t* __restrict__ tree_of,
                                                        uint32_t* __restrict__ leaf, uint32_t n, uint64_t seed,
                                                        int4* __restrict__ node4, int32_t* __restrict__ parent,
                                                        int tpn, int small, int32_t base) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int G = kBlock / tpn, g = threadIdx.x / tpn, lt = threadIdx.x - g * tpn;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  float* mean = reinterpret_cast<float*>(smem_raw) + size_t(g) * 2 * d;
  float* var = mean + d;
  __shared__ uint32_t ids_s[kBlock / 64][kSample];
  __shared__ int cd_s[kBlock / 64];
  __shared__ uint32_t wcnt[kBlock / 64][3];
  for (int a0 = blockIdx.x * G; a0 < na; a0 += gridDim.x * G) {
    const int a = a0 + g;
    const bool valid = a < na;
    int ni = 0, t = 0;
    uint2 r = make_uint2(0, 1);
    if (valid) {
      ni = active[a];
      r = range[ni];
      t = tree_of[ni];
    }
    const unsigned count = r.y - r.x;
    const int cnt = int(count < unsigned(kSample) ? count : unsigned(kSample));
    const bool part = valid && count <= unsigned(small);  // the same for the whole group
    uint32_t my_id = 0;
    if (part && unsigned(lt) < count) my_id = leaf[size_t(t) * n + r.x + lt];
    if (valid)
      for (int j = lt; j < cnt; j += tpn) ids_s[g][j] = leaf[size_t(t) * n + r.x + j];
    __syncthreads();
    if (valid) {
      const float div = 1.0f / float(cnt);
      for (int k = lt; k < d; k += tpn) {
        float m = 0.f;
        for (int j0 = 0; j0 < cnt; j0 += kChunk) {
          float xv[kChunk];
#pragma unroll
          for (int j = 0; j < kChunk; ++j) {
            const int jj = j0 + j < cnt ? j0 + j : cnt - 1;
            xv[j] = row_elem<kRow>(X, row_off<kOff32>(ids_s[g][jj], D) + k);
          }
#pragma unroll
          for (int j = 0; j < kChunk; ++j)
            if (j0 + j < cnt) m = m + xv[j];
        }
        m = m * div;
        float v = 0.f;
        for (int j0 = 0; j0 < cnt; j0 += kChunk) {
          float xv[kChunk];
#pragma unroll
          for (int j = 0; j < kChunk; ++j) {
            const int jj = j0 + j < cnt ? j0 + j : cnt - 1;
            xv[j] = row_elem<kRow>(X, row_off<kOff32>(ids_s[g][jj], D) + k);
          }
#pragma unroll
          for (int j = 0; j < kChunk; ++j)
            if (j0 + j < cnt) {
              const float tt = xv[j] - m;
              v = __builtin_fmaf(tt, tt, v);
            }
        }
        mean[k] = m;
        var[k] = v;
      }
    }
    __syncthreads();
    if (valid && lt < 64) {
      const int num = d < 5 ? d : 5;
      int top[5] = {0, 0, 0, 0, 0};
      for (int q = 0; q < num; ++q) {
        float bv = -1.f;
        int bi = 0x7fffffff;
        for (int k = lt; k < d; k += 64) {
          bool taken = false;
          for (int z = 0; z < q; ++z) taken = taken || top[z] == k;
          const float vk = var[k];
          if (!taken && (vk > bv || (vk == bv && k < bi))) {
            bv = vk;
            bi = k;
          }
        }
        for (int o = 32; o > 0; o >>= 1) {
          const float ov = __shfl_xor(bv, o);
          const int oi = __shfl_xor(bi, o);
          if (ov > bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
          }
        }
        top[q] = bi;
      }
      if (lt == 0) {
        const uint64_t rnd = hash4(seed, ST_SPLIT, uint64_t(t), (uint64_t(r.x) << 32) | r.y) % uint64_t(num);
        int cd = top[0];
        for (int z = 1; z < 5; ++z)
          if (uint64_t(z) == rnd) cd = top[z];
        cd_s[g] = cd;
        if (!part) {
          int4 nd = node4[ni];
          nd.x = cd;
          nd.y = __float_as_int(mean[cd]);
          node4[ni] = nd;
        }
      }
    }
    __syncthreads();
    // small nodes: stable three-way partition (<, ==, > cut) of the ids
    int c = 3;
    uint64_t m0 = 0, m1 = 0, m2 = 0;
    if (part) {
      const int cd = cd_s[g];
      const float cut = mean[cd];
      if (unsigned(lt) < count) c = classify(row_elem<kRow>(X, size_t(my_id) * D + cd), cut);
      m0 = __ballot(c == 0);
      m1 = __ballot(c == 1);
      m2 = __ballot(c == 2);
      if (lane == 0) {
        wcnt[wid][0] = uint32_t(__popcll(m0));
        wcnt[wid][1] = uint32_t(__popcll(m1));
        wcnt[wid][2] = uint32_t(__popcll(m2));
      }
    }
    __syncthreads();
    if (part) {
      const int w0 = g * (tpn >> 6);
      uint32_t p0 = 0, p1 = 0, p2 = 0, c0 = 0, c1 = 0;
      for (int w = w0; w < w0 + (tpn >> 6); ++w) {
        if (w < wid) {
          p0 += wcnt[w][0];
          p1 += wcnt[w][1];
          p2 += wcnt[w][2];
        }
        c0 += wcnt[w][0];
        c1 += wcnt[w][1];
      }
      const uint64_t below = (1ull << lane) - 1ull;
      if (c != 3) {
        const uint32_t dst = c == 0 ? p0 + uint32_t(__popcll(m0 & below))
                                    : (c == 1 ? c0 + p1 + uint32_t(__popcll(m1 & below))
                                              : c0 + c1 + p2 + uint32_t(__popcll(m2 & below)));
        leaf[size_t(t) * n + r.x + dst] = my_id;
      }
      if (lt == 0) {
        const int cd = cd_s[g];
        make_children(ni, a, base, r, make_int4(cd, __float_as_int(mean[cd]), -1, -1), c0, c1, node4, parent,
                      range, tree_of);
      }
    }
    __syncthreads();
  }
}

// block -> (active slot, block index within the node) via the block prefix
__device__ inline int find_node(const int32_t* __restrict__ bstart, int na, int b) {
  int lo = 0, hi = na;  // largest i with bstart[i] <= b
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (bstart[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

template <int kRow>
__global__ void __launch_bounds__(kBlock) k_part_count(const void* __restrict__ X, int D,
                                                       const int32_t* __restrict__ active, int na,
                                                       const int32_t* __restrict__ bstart, const int* __restrict__ nblk,
                                                       const uint2* __restrict__ range,
                                                       const int32_t* __restrict__ tree_of,
                                                       const int4* __restrict__ node4,
                                                       const uint32_t* __restrict__ leaf, uint32_t n,
                                                       uint8_t* __restrict__ cls, uint32_t* __restrict__ bcnt) {
  __shared__ int tmp[kBlock / 64];
  const int total = *nblk;
  for (int b = blockIdx.x; b < total; b += gridDim.x) {
    const int a = find_node(bstart, na, b);
    const int ni = active[a];
    const uint2 r = range[ni];
    const int t = tree_of[ni];
    const int4 nd = node4[ni];
    const float cut = __int_as_float(nd.y);
    const uint32_t s0 = r.x + uint32_t(b - bstart[a]) * kPartB;
    // counts of '<' and '==' packed in 16-bit fields; '>' follows from the
    // number of valid ids (a prefix of the block)
    const uint32_t nvalid = (r.y - s0) < uint32_t(kPartB) ? (r.y - s0) : uint32_t(kPartB);
    uint32_t packed = 0;
    for (int e = 0; e < 4; ++e) {
      uint32_t pos = s0 + threadIdx.x * 4 + e;
      if (pos < r.y) {
        uint32_t id = leaf[size_t(t) * n + pos];
        int c = classify(row_elem<kRow>(X, size_t(id) * D + nd.x), cut);
        cls[size_t(t) * n + pos] = uint8_t(c);
        if (c < 2) packed += 1u << (16 * c);
      }
    }
    int tot;
    block_exclusive_scan(int(packed), tmp, &tot);
    if (threadIdx.x == 0) {
      uint32_t u = uint32_t(tot);
      uint32_t c0 = u & 0xffffu, c1 = u >> 16;
      bcnt[size_t(b) * 3 + 0] = c0;
      bcnt[size_t(b) * 3 + 1] = c1;
      bcnt[size_t(b) * 3 + 2] = nvalid - c0 - c1;
    }
  }
}

// One wavefront per active node of more than `small` ids: class totals, then
// per-block exclusive offsets (class bases [<][==][>] added) in rounds of 64
// blocks, lim1/lim2, the balancing rule, children.
__global__ void k_part_node(const int32_t* __restrict__ active, int na, const int32_t* __restrict__ bstart,
                            uint32_t* __restrict__ bcnt, int4* __restrict__ node4, int32_t* __restrict__ parent,
                            uint2* __restrict__ range, int32_t* __restrict__ tree_of, int32_t base, int small) {
  const int a = int((size_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (a >= na) return;  // whole wavefronts
  const int ni = active[a];
  const uint2 r = range[ni];
  if (r.y - r.x <= unsigned(small)) return;  // partitioned with its statistics
  const int b0 = bstart[a], b1 = bstart[a + 1];
  uint32_t c0 = 0, c1 = 0;
  for (int b = b0 + lane; b < b1; b += 64) {
    c0 += bcnt[size_t(b) * 3];
    c1 += bcnt[size_t(b) * 3 + 1];
  }
  for (int o = 32; o > 0; o >>= 1) {
    c0 += __shfl_xor(c0, o);
    c1 += __shfl_xor(c1, o);
  }
  uint32_t run0 = 0, run1 = 0, run2 = 0;
  for (int bb = b0; bb < b1; bb += 64) {
    const int b = bb + lane;
    uint32_t x0 = 0, x1 = 0, x2 = 0;
    if (b < b1) {
      x0 = bcnt[size_t(b) * 3];
      x1 = bcnt[size_t(b) * 3 + 1];
      x2 = bcnt[size_t(b) * 3 + 2];
    }
    uint32_t s0 = x0, s1 = x1, s2 = x2;  // inclusive scans over the round
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y0 = __shfl_up(s0, o), y1 = __shfl_up(s1, o), y2 = __shfl_up(s2, o);
      if (lane >= o) {
        s0 += y0;
        s1 += y1;
        s2 += y2;
      }
    }
    if (b < b1) {
      bcnt[size_t(b) * 3] = run0 + s0 - x0;
      bcnt[size_t(b) * 3 + 1] = c0 + run1 + s1 - x1;
      bcnt[size_t(b) * 3 + 2] = c0 + c1 + run2 + s2 - x2;
    }
    run0 += __shfl(s0, 63);
    run1 += __shfl(s1, 63);
    run2 += __shfl(s2, 63);
  }
  if (lane == 0) make_children(ni, a, base, r, node4[ni], c0, c1, node4, parent, range, tree_of);
}

__global__ void __launch_bounds__(kBlock) k_part_scatter(const int32_t* __restrict__ active, int na,
                                                         const int32_t* __restrict__ bstart,
                                                         const int* __restrict__ nblk, const uint2* __restrict__ range,
                                                         const int32_t* __restrict__ tree_of,
                                                         const uint32_t* __restrict__ bcnt,
                                                         const uint8_t* __restrict__ cls,
                                                         const uint32_t* __restrict__ leaf, uint32_t n,
                                                         uint32_t* __restrict__ tmpbuf) {
  __shared__ int tmp[kBlock / 64];
  const int total = *nblk;
  for (int b = blockIdx.x; b < total; b += gridDim.x) {
    const int a = find_node(bstart, na, b);
    const int ni = active[a];
    const uint2 r = range[ni];
    const int t = tree_of[ni];
    const uint32_t s0 = r.x + uint32_t(b - bstart[a]) * kPartB;
    const uint32_t nvalid = (r.y - s0) < uint32_t(kPartB) ? (r.y - s0) : uint32_t(kPartB);
    int c[4];
    uint32_t packed = 0;
    for (int e = 0; e < 4; ++e) {
      uint32_t pos = s0 + threadIdx.x * 4 + e;
      c[e] = pos < r.y ? int(cls[size_t(t) * n + pos]) : 3;
      if (c[e] < 2) packed += 1u << (16 * c[e]);
    }
    int tot;
    uint32_t ex = uint32_t(block_exclusive_scan(int(packed), tmp, &tot));
    uint32_t before = threadIdx.x * 4u < nvalid ? threadIdx.x * 4u : nvalid;
    uint32_t r0 = ex & 0xffffu, r1 = ex >> 16, r2 = before - r0 - r1;
    for (int e = 0; e < 4; ++e) {
      if (c[e] == 3) continue;
      uint32_t pos = s0 + threadIdx.x * 4 + e;
      uint32_t rk = c[e] == 0 ? r0++ : (c[e] == 1 ? r1++ : r2++);
      uint32_t dst = r.x + bcnt[size_t(b) * 3 + c[e]] + rk;
      tmpbuf[size_t(t) * n + dst] = leaf[size_t(t) * n + pos];
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_part_copy(const int32_t* __restrict__ active, int na,
                                                      const int32_t* __restrict__ bstart, const int* __restrict__ nblk,
                                                      const uint2* __restrict__ range,
                                                      const int32_t* __restrict__ tree_of,
                                                      const uint32_t* __restrict__ tmpbuf, uint32_t n,
                                                      uint32_t* __restrict__ leaf) {
  const int total = *nblk;
  for (int b = blockIdx.x; b < total; b += gridDim.x) {
    const int a = find_node(bstart, na, b);
    const int ni = active[a];
    const uint2 r = range[ni];
    const int t = tree_of[ni];
    const uint32_t s0 = r.x + uint32_t(b - bstart[a]) * kPartB;
    for (int e = threadIdx.x; e < kPartB; e += kBlock) {
      uint32_t pos = s0 + e;
      if (pos < r.y) leaf[size_t(t) * n + pos] = tmpbuf[size_t(t) * n + pos];
    }
  }
}

// partition blocks per active node (none for a node of at most `small` ids)
__global__ void k_blocks_per_node(const int32_t* __restrict__ active, int na, const uint2* __restrict__ range,
                                  int small, int32_t* __restrict__ nb) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= na) return;
  uint2 r = range[active[a]];
  const uint32_t c = r.y - r.x;
  nb[a] = c <= uint32_t(small) ? 0 : int((c + kPartB - 1) / kPartB);
}

// flags[i] = node (cands[i]) stays active: size > thr
__global__ void k_flags(const int32_t* __restrict__ cands, int m, const uint2* __restrict__ range, uint32_t thr,
                        int base_is_children, int32_t base, int32_t* __restrict__ flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int32_t ni = base_is_children ? base + i : cands[i];
  uint2 r = range[ni];
  flags[i] = (r.y - r.x > thr) ? 1 : 0;
}

__global__ void k_compact(const int32_t* __restrict__ cands, int m, int base_is_children, int32_t base,
                          const int32_t* __restrict__ flags, const int32_t* __restrict__ pos,
                          int32_t* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  if (flags[i]) out[pos[i]] = base_is_children ? base + i : cands[i];
}

// ---- device exclusive scan (int32) -------------------------------------------
__global__ void __launch_bounds__(kBlock) k_scan_blocks(const int32_t* __restrict__ in, int m,
                                                        int32_t* __restrict__ out, int32_t* __restrict__ sums) {
  __shared__ int tmp[kBlock / 64];
  int i = blockIdx.x * kBlock + threadIdx.x;
  int v = i < m ? in[i] : 0;
  int tot;
  int ex = block_exclusive_scan(v, tmp, &tot);
  if (i < m) out[i] = ex;
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}
__global__ void __launch_bounds__(kBlock) k_scan_sums(int32_t* __restrict__ sums, int nb, int32_t* __restrict__ total) {
  __shared__ int tmp[kBlock / 64];
  int run = 0;
  for (int b0 = 0; b0 < nb; b0 += kBlock) {
    int i = b0 + threadIdx.x;
    int v = i < nb ? sums[i] : 0;
    int tot;
    int ex = block_exclusive_scan(v, tmp, &tot);
    if (i < nb) sums[i] = run + ex;
    run += tot;
  }
  if (threadIdx.x == 0) *total = run;
}
__global__ void k_scan_add(int32_t* __restrict__ out, int m, const int32_t* __restrict__ sums,
                           const int32_t* __restrict__ total) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < m) out[i] += sums[blockIdx.x];
  if (i == m) out[m] = *total;
}

}  // namespace

// Exclusive scan of m ints; out[m] = total; *total_dev = total.
int device_exclusive_scan(const int32_t* in, int m, int32_t* out, int32_t* total_dev, DevBuf<int32_t>& sums,
                          hipStream_t st) {
  int nb = (m + kBlock - 1) / kBlock;
  if (nb < 1) nb = 1;
  if (!sums.ensure(nb)) return fail(EFG_ERR_MEMORY, "out of device memory (scan)");
  k_scan_blocks<<<nb, kBlock, 0, st>>>(in, m, out, sums.p);
  k_scan_sums<<<1, kBlock, 0, st>>>(sums.p, nb, total_dev);
  k_scan_add<<<nb + ((m % kBlock) == 0 ? 1 : 0), kBlock, 0, st>>>(out, m, sums.p, total_dev);
  return EFG_OK;
}

int gpu_build_trees(efg_index* ix) {
  const efg_build_params& p = ix->bp;
  if (p.trees <= 0) return fail(EFG_ERR_ARG, "tree number must be positive");
  if (p.S <= 0 || p.tns <= 0) return fail(EFG_ERR_ARG, "invalid S or TNS");
  const int T = p.trees;
  const uint32_t n = uint32_t(ix->n);
  hipStream_t st = ix->stream;
  const size_t cap = size_t(T) * (2 * size_t(n) + 1);
  if (!ix->node4.alloc(cap) || !ix->parent.alloc(cap) || !ix->range.alloc(cap) || !ix->roots.alloc(T) ||
      !ix->leaf.alloc(size_t(T) * n))
    return fail(EFG_ERR_MEMORY, "out of device memory (trees)");
  DevBuf<int32_t> tree_of, act, nxt, flags, pos, nb, bstart, sums, scal;
  DevBuf<uint32_t> tmpbuf, bcnt;
  DevBuf<uint8_t> cls;
  const size_t act_cap = cap;
  if (!tree_of.alloc(cap) || !act.alloc(act_cap) || !nxt.alloc(act_cap) || !flags.alloc(act_cap + 1) ||
      !pos.alloc(act_cap + 1) || !nb.alloc(act_cap + 1) || !bstart.alloc(act_cap + 1) || !scal.alloc(4) ||
      !tmpbuf.alloc(size_t(T) * n) || !cls.alloc(size_t(T) * n))
    return fail(EFG_ERR_MEMORY, "out of device memory (tree build)");
  const size_t blk_cap = size_t(T) * n / kPartB + act_cap + 1;
  if (!bcnt.alloc(blk_cap * 3)) return fail(EFG_ERR_MEMORY, "out of device memory (tree build)");
  {
    size_t tot = size_t(T) * n;
    k_init_perm<<<unsigned((tot + 255) / 256), 256, 0, st>>>(ix->leaf.p, T, n, p.seed);
    k_init_roots<<<(T + 255) / 256, 256, 0, st>>>(ix->node4.p, ix->parent.p, ix->range.p, ix->roots.p, T, n);
    std::vector<int32_t> h_act(T);
    for (int t = 0; t < T; ++t) h_act[t] = t;
    hipMemcpyAsync(act.p, h_act.data(), T * 4, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(tree_of.p, h_act.data(), T * 4, hipMemcpyHostToDevice, st);
    if (int rc = check_hip(hipStreamSynchronize(st), "tree build")) return rc;
  }
  int dev = 0, ncu = 256;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  std::vector<int8_t> level_of;  // per node level
  level_of.assign(T, 0);
  int na = T;
  int32_t M = T;
  bool bfs = true;
  int Lb = -1, level = 0;
  // threads per node: the dimensions rounded up to wavefronts, at most a workgroup
  const int tpn = int(std::min<size_t>(kBlock, (ix->d + 63) / 64 * 64));
  const int small = tpn;  // nodes of at most this many ids are partitioned with their statistics
  const size_t lds_stats = size_t(kBlock / tpn) * 2 * ix->d * sizeof(float);
  if (lds_stats > 150 * 1024) return fail(EFG_ERR_ARG, "dimension too large for the split statistics");
  // rows read by the statistics and partitions: FP32, or in the build's
  // precision form the 16-bit high parts when every row is exact in them
  // (the same values in the same order of operations: the same trees, from
  // half the bytes)
  const void* X = ix->data.p;
  size_t xs = ix->D;
  int xk = 0;
  if (ix->b_precision > 0) {
    if (int rc = ensure_split(ix, st)) return rc;
    if (!ix->split_lo_used) {
      X = ix->split_hi.p;
      xs = size_t(ix->split_Dk);
      xk = ix->b_precision;
    }
  }
  const bool off32 = size_t(n) * xs < (size_t(1) << 30);
  using StatsFn = decltype(&k_split_stats<true, 0>);
  using CountFn = decltype(&k_part_count<0>);
  const StatsFn stats_fns[2][3] = {{&k_split_stats<false, 0>, &k_split_stats<false, 1>, &k_split_stats<false, 2>},
                                   {&k_split_stats<true, 0>, &k_split_stats<true, 1>, &k_split_stats<true, 2>}};
  const CountFn count_fns[3] = {&k_part_count<0>, &k_part_count<1>, &k_part_count<2>};
  const StatsFn stats_fn = stats_fns[off32 ? 1 : 0][xk];
  const CountFn count_fn = count_fns[xk];
  hipFuncSetAttribute(reinterpret_cast<const void*>(stats_fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                      int(lds_stats));
  int32_t h_total = 0;
  while (na > 0) {
    if (bfs && !(na < 1100)) {  // switch to DFS: keep nodes larger than TNS
      bfs = false;
      Lb = level;
      k_flags<<<(na + 255) / 256, 256, 0, st>>>(act.p, na, ix->range.p, uint32_t(p.tns), 0, 0, flags.p);
      if (int rc = device_exclusive_scan(flags.p, na, pos.p, scal.p, sums, st)) return rc;
      k_compact<<<(na + 255) / 256, 256, 0, st>>>(act.p, na, 0, 0, flags.p, pos.p, nxt.p);
      hipMemcpyAsync(&h_total, scal.p, 4, hipMemcpyDeviceToHost, st);
      if (int rc = check_hip(hipStreamSynchronize(st), "tree build")) return rc;
      std::swap(act.p, nxt.p);
      std::swap(act.n, nxt.n);
      na = h_total;
      if (na == 0) break;
    }
    const int32_t base = M;
    if (size_t(base) + 2 * size_t(na) > cap) return fail(EFG_ERR_STATE, "tree node capacity exceeded");
    // split statistics
    {
      const int per = kBlock / tpn;
      unsigned g = unsigned(std::min<int>((na + per - 1) / per, 65535 * 4));
      stats_fn<<<g, kBlock, lds_stats, st>>>(
          X, int(ix->d), int(xs), act.p, na, ix->range.p, tree_of.p, ix->leaf.p, n, p.seed, ix->node4.p,
          ix->parent.p, tpn, small, base);
    }
    // blocks of the nodes larger than `small`
    k_blocks_per_node<<<(na + 255) / 256, 256, 0, st>>>(act.p, na, ix->range.p, small, nb.p);
    if (int rc = device_exclusive_scan(nb.p, na, bstart.p, scal.p + 1, sums, st)) return rc;
    {
      unsigned g = unsigned(size_t(ncu) * 16);
      count_fn<<<g, kBlock, 0, st>>>(X, int(xs), act.p, na, bstart.p, scal.p + 1, ix->range.p,
                                         tree_of.p, ix->node4.p, ix->leaf.p, n, cls.p, bcnt.p);
      k_part_node<<<unsigned((size_t(na) * 64 + 255) / 256), 256, 0, st>>>(
          act.p, na, bstart.p, bcnt.p, ix->node4.p, ix->parent.p, ix->range.p, tree_of.p, base, small);
      k_part_scatter<<<g, kBlock, 0, st>>>(act.p, na, bstart.p, scal.p + 1, ix->range.p, tree_of.p, bcnt.p, cls.p,
                                           ix->leaf.p, n, tmpbuf.p);
      k_part_copy<<<g, kBlock, 0, st>>>(act.p, na, bstart.p, scal.p + 1, ix->range.p, tree_of.p, tmpbuf.p, n,
                                        ix->leaf.p);
    }
    M = base + 2 * na;
    level_of.resize(M, int8_t(level + 1));
    // next active list: children larger than S (BFS) or TNS (DFS)
    const int m = 2 * na;
    k_flags<<<(m + 255) / 256, 256, 0, st>>>(nullptr, m, ix->range.p, uint32_t(bfs ? p.S : p.tns), 1, base, flags.p);
    if (int rc = device_exclusive_scan(flags.p, m, pos.p, scal.p, sums, st)) return rc;
    k_compact<<<(m + 255) / 256, 256, 0, st>>>(nullptr, m, 1, base, flags.p, pos.p, nxt.p);
    hipMemcpyAsync(&h_total, scal.p, 4, hipMemcpyDeviceToHost, st);
    if (int rc = check_hip(hipStreamSynchronize(st), "tree build")) return rc;
    std::swap(act.p, nxt.p);
    std::swap(act.n, nxt.n);
    na = h_total;
    ++level;
  }
  if (int rc = check_hip(hipStreamSynchronize(st), "tree build")) return rc;
  ix->T = T;
  ix->M = size_t(M);
  ix->snode_valid = false;
  ix->leaf0_perm = true;  // every leaf list is a permutation of the rows
  // DivDim written for leaves in the trees file: leaves the DFS phase made
  // (and roots) carry -1, leaves left behind by the BFS phase carry 0
  // (zero-initialised nodes, kdtreeub_index.hpp:1249, 1142-1143).
  // (derived from the node levels when the trees are read back, so the
  // build does not copy the node ranges to the host)
  ix->leaf_dim.clear();
  ix->leaf_dim_lazy = Lb != 0;
  ix->node_level = std::move(level_of);
  ix->tree_lb = Lb;
  ix->tree_S = p.S;
  ix->s_trees = T;
  return EFG_OK;
}

}  // namespace efg

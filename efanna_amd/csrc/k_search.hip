// Graph ANN search, method 0 (getNeighbors_nnexp, kdtreeub_index.hpp:363-494).
//
// Stage 1, k_descend: one thread per (query, tree) walks its tree depth first,
// query side first, and records the first `lsize` leaves (getSearchNodeList,
// kdtreeub_index.hpp:332-346), with the pending far children on a short
// LDS stack and parent links past its depth.
//
// Stage 2, query order: queries are ordered by the position of their first
// leaf in tree 0 (a counting sort), so queries that read the same base rows
// run close together; the search kernel hands each XCD (workgroup id mod 8)
// one contiguous slice of that order, keeping the shared rows in one L2.
// Results do not depend on the order.
//
// Stage 3, k_search_nnexp: one workgroup of 256 threads per query.
//   * seeds: leaves visited round robin (leaf rank outer, tree inner,
//     396-410); the first init_num distinct ids are kept.  Candidates are
//     taken 256 at a time; a candidate is kept when it is not yet visited and
//     is the first occurrence of its id in the chunk (a 512-slot LDS table
//     keyed by id holds the minimum sequence position).  Missing seeds come
//     from the counter-based generator (stands in for rand(), 411-417).
//   * visited set: exact open-addressing hash set in LDS (bit 31 of a slot =
//     expanded, i.e. newflag cleared, 450); a query whose set would pass 3/4
//     load is redone by a second launch with a per-workgroup table in global
//     memory.
//   * distances: two lanes per row, float4 loads, the reference's AVX
//     accumulation order (bit-identical FP32).
//   * pool = the best R = max(extend_to, K) by (distance, id) (partial_sort
//     on pair<float, size_t>, 438-442): a block radix select over the seed
//     keys finds the R-th key, the R keys are sorted.  Each epoch expands the
//     unexpanded ids among the first extend_to, collects unvisited
//     neighbours, and merges them into the sorted pool by rank (446-484).  An
//     epoch that expands nothing ends the loop (later ones could not change
//     the pool).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "efg_device.hpp"
#include "efg_internal.hpp"
#include "efg_mlp.hpp"

namespace efg {

// The far children passed on the way down are kept on a per-thread stack in
// LDS (the last kDescStack of them), so the next leaf is reached from the
// deepest pending far child without walking back up; past the stack's depth
// the walk climbs the parent links.  Nodes are read from the 16-byte descent
// copy: the left child of an internal node is the next node, and the node
// says which of its children are leaves, so a leaf is recorded from its
// parent (or its stack entry, bit 31) without being read.  Each trip round
// the loop reads one internal node for every lane, so the lanes of a
// wavefront, which walk different trees, advance together.  With kLdsQ the
// workgroup's query rows are staged in LDS first, so each step waits on one
// global load (the node), not two.
constexpr int kDescStack = 16;
constexpr int kDescThreads = 256;
constexpr size_t kDescLdsQ = 64 * 1024;  // largest staged query block
constexpr uint32_t kDescLeaf = 0x80000000u;

// dim word of an internal descent node: split dimension in bits 0-23, left
// child is a leaf in bit 24, right child in bit 25; -1 marks a leaf
__device__ __forceinline__ uint32_t desc_left(int32_t node, int32_t dimw) {
  return uint32_t(node + 1) | (uint32_t(dimw >> 24 & 1) << 31);
}
__device__ __forceinline__ uint32_t desc_right(const DNode& c) {
  return uint32_t(c.right) | (uint32_t(c.dim >> 25 & 1) << 31);
}

template <bool kLdsQ>
__global__ void __launch_bounds__(kDescThreads)
    k_descend(const float* __restrict__ Q, size_t q_stride, int nq, int d, int T, int lsize,
              const int32_t* __restrict__ roots, const DNode* __restrict__ dn, int32_t* __restrict__ leaves,
              const int32_t* __restrict__ qlist, const int* __restrict__ nlist, const int* __restrict__ qlsize,
              size_t lstride) {
  __shared__ uint32_t stk[kDescStack][kDescThreads];
  extern __shared__ float qrows[];  // kLdsQ: rows of list positions [p0, p1]
  const int tid = threadIdx.x;
  const size_t gid = size_t(blockIdx.x) * kDescThreads + tid;
  const int nw = qlist ? *nlist : nq;  // queries of the list only, when given
  const size_t nthr = size_t(nw) * T;
  const size_t p0 = size_t(blockIdx.x) * kDescThreads / T;
  if (kLdsQ) {
    if (size_t(blockIdx.x) * kDescThreads >= nthr) return;  // whole workgroup idle
    const size_t last = std::min(nthr, size_t(blockIdx.x + 1) * kDescThreads) - 1;
    const int rows = int(last / T - p0 + 1);
    for (int e = tid; e < rows * d; e += kDescThreads) {
      const int r = e / d, j = e - r * d;
      const size_t p = p0 + r;
      const int qr = qlist ? qlist[p] : int(p);
      qrows[e] = Q[size_t(qr) * q_stride + j];
    }
    __syncthreads();
  }
  if (gid >= nthr) return;
  const int qi = qlist ? qlist[gid / T] : int(gid / T), t = int(gid % T);
  const float* q = kLdsQ ? qrows + (gid / T - p0) * d : Q + size_t(qi) * q_stride;
  // a batch of classes: leaves per tree per query (qlsize), rows of lstride slots
  if (qlsize) lsize = qlsize[qi];
  int32_t* out = leaves + size_t(qi) * (qlsize ? lstride : size_t(lsize) * T);  // [ni][t]
  int cnt = 0;
  int node = roots[t];
  DNode cur = dn[node];
  uint32_t e = cur.dim < 0 ? uint32_t(node) | kDescLeaf : uint32_t(node);
  int top = 0, bottom = 0;  // stack entries [bottom, top), slot i % kDescStack
  for (;;) {
    if (!(e & kDescLeaf)) {  // internal node (read): one step down, query side first
      const bool left = q[cur.dim & 0xFFFFFF] < cur.val;
      const uint32_t l = desc_left(node, cur.dim), r = desc_right(cur);
      stk[top % kDescStack][tid] = left ? r : l;
      if (++top - bottom > kDescStack) ++bottom;
      e = left ? l : r;
    }
    while (e & kDescLeaf) {  // leaf: record it, then the next entry
      node = int(e & ~kDescLeaf);
      out[size_t(cnt) * T + t] = node;
      if (++cnt == lsize) break;
      if (top > bottom) {  // deepest pending far child
        --top;
        e = stk[top % kDescStack][tid];
        continue;
      }
      // stack exhausted: up to the first ancestor whose second child is unvisited
      bool more = false;
      int p = dn[node].parent;
      while (p >= 0) {
        const DNode pd = dn[p];
        const bool pleft = q[pd.dim & 0xFFFFFF] < pd.val;
        if ((node == p + 1) == pleft) {  // node was the side taken first
          e = pleft ? desc_right(pd) : desc_left(p, pd.dim);
          more = true;
          break;
        }
        node = p;
        p = pd.parent;
      }
      if (!more) break;
    }
    if (cnt == lsize || (e & kDescLeaf)) break;
    node = int(e);
    cur = dn[node];
  }
  for (; cnt < lsize; ++cnt) out[size_t(cnt) * T + t] = -1;
}

// query order: bucket = start of the first leaf in tree 0 >> shift
__global__ void k_qkey(const int32_t* __restrict__ leaves, int nleaf, int nq, const SNode* __restrict__ sn,
                       int shift, int32_t* __restrict__ key, uint32_t* __restrict__ hist) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  int nd = leaves[size_t(q) * nleaf];
  int k = nd >= 0 ? int(sn[nd].start >> shift) : 0;
  key[q] = k;
  atomicAdd(&hist[k], 1u);
}
__global__ void k_qscatter(const int32_t* __restrict__ key, int nq, const int32_t* __restrict__ off,
                           uint32_t* __restrict__ cursor, int32_t* __restrict__ order) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  int k = key[q];
  order[off[k] + int(atomicAdd(&cursor[k], 1u))] = q;
}

namespace {

constexpr uint32_t kEmpty = 0xffffffffu;
constexpr uint32_t kExp = 0x80000000u;
// Per query, with T threads: the seed chunk is T positions, deduplicated in a
// 2T-slot hash table; new candidates gather in a 2T-key buffer between merges.
__host__ __device__ constexpr int ctab_slots(int threads) { return 2 * threads; }
__host__ __device__ constexpr int new_buf(int threads) { return 2 * threads; }
constexpr int kKPT = 16;   // seed keys per thread held in registers (init_num <= 4096)
constexpr int kMiscInts = 12;  // per-query LDS scalars (Smem::misc)

__device__ inline int tab_find(const uint32_t* tab, int logcap, uint32_t id, bool* present) {
  const uint32_t mask = (1u << logcap) - 1;
  uint32_t s = hslot(id, logcap);
  for (;;) {
    uint32_t v = tab[s];
    if (v == kEmpty) { *present = false; return int(s); }
    if ((v & ~kExp) == id) { *present = true; return int(s); }
    s = (s + 1) & mask;
  }
}
__device__ inline bool tab_insert(uint32_t* tab, int logcap, uint32_t id) {
  const uint32_t mask = (1u << logcap) - 1;
  uint32_t s = hslot(id, logcap);
  for (;;) {
    uint32_t v = tab[s];
    if (v == kEmpty) {
      uint32_t old = atomicCAS(&tab[s], kEmpty, id);
      if (old == kEmpty) return true;
      v = old;
    }
    if ((v & ~kExp) == id) return false;
    s = (s + 1) & mask;
  }
}

// Insert a table entry as it is (id with its expanded flag), for moving a
// visited set into a larger table; ids are unique, so no duplicate check.
__device__ inline void tab_put_raw(uint32_t* tab, int logcap, uint32_t v) {
  const uint32_t mask = (1u << logcap) - 1;
  uint32_t s = hslot(v & ~kExp, logcap);
  while (atomicCAS(&tab[s], kEmpty, v) != kEmpty) s = (s + 1) & mask;
}

struct Smem {
  float* q;
  uint64_t* buf;  // seeds; then pool P[Rp] | new N[new_buf] | merge output O[Rp]
  uint32_t* tab;
  uint32_t* ctab_id;
  uint32_t* ctab_pos;
  int32_t* lpos;
  int32_t* lstart;
  uint32_t* explist;
  int32_t* expoff;
  uint32_t* hist;
  int* scan;
  int* misc;
};

__device__ inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

// Block radix select over unique keys buf[0, m): returns T with exactly R
// keys <= T.  Any block of at least 64 threads.
__device__ uint64_t block_select(const uint64_t* buf, int m, int R, uint32_t* hist, int* misc) {
  const int tid = threadIdx.x;
  uint64_t prefix = 0, mask = 0;
  int need = R;
  for (int shift = 56; shift >= 0; shift -= 8) {
    for (int i = tid; i < 256; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    for (int i = tid; i < m; i += blockDim.x) {
      uint64_t k = buf[i];
      if ((k & mask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (tid < 64) {
      uint32_t h0 = hist[4 * tid], h1 = hist[4 * tid + 1], h2 = hist[4 * tid + 2], h3 = hist[4 * tid + 3];
      int loc = int(h0 + h1 + h2 + h3);
      int x = loc;
      for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o);
        if (tid >= o) x += y;
      }
      int before = x - loc;  // count in bins < 4*tid
      if (before < need && need <= x) {
        int c = before;
        int b = 4 * tid;
        uint32_t hh[4] = {h0, h1, h2, h3};
        int j = 0;
        while (c + int(hh[j]) < need) {
          c += int(hh[j]);
          ++j;
        }
        misc[4] = b + j;
        misc[5] = c;
        misc[6] = int(hh[j]);
      }
    }
    __syncthreads();
    const int b = misc[4], c = misc[5], cnt = misc[6];
    need -= c;
    prefix |= uint64_t(b) << shift;
    mask |= uint64_t(0xff) << shift;
    __syncthreads();
    if (cnt == need) return prefix | ~mask;
  }
  return prefix;
}

// Wavefront-wide InsertIntoKnn (base_index.hpp:37-62) on LDS arrays sorted by
// distance: insert after every entry with distance <= nd, duplicate check only
// among equal distances (returns K + 1), entries [i, K) shift up one slot
// (the slot K is written).  Executed by one full wavefront.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
__device__ int wave_insert(float* dist, uint32_t* ids, uint8_t* fl, int K, float nd, uint32_t nid) {
  const int lane = threadIdx.x & 63;
  int cnt = 0;
  bool dup = false;
  for (int b = 0; b < K; b += 64) {
    const int j = b + lane;
    const bool in = j < K;
    const float dj = in ? dist[j] : 0.f;
    cnt += __popcll(__ballot(in && dj <= nd));
    dup = dup || (in && dj == nd && ids[j] == nid);
  }
  if (__ballot(dup)) return K + 1;
  const int i = cnt;
  wave_sync();
  for (int b = ((K - 1) >> 6) << 6; b >= 0 && K > 0; b -= 64) {
    const int j = b + lane;
    const bool act = j >= i && j < K;
    float dv = 0.f;
    uint32_t iv = 0;
    uint8_t fv = 0;
    if (act) {
      dv = dist[j];
      iv = ids[j];
      fv = fl[j];
    }
    wave_sync();
    if (act) {
      dist[j + 1] = dv;
      ids[j + 1] = iv;
      fl[j + 1] = fv;
    }
    wave_sync();
  }
  if (lane == 0) {
    dist[i] = nd;
    ids[i] = nid;
    fl[i] = 1;
  }
  wave_sync();
  return i;
}

// wavefront bitonic sort of n (power of two) keys in LDS
__device__ void wave_bitonic_sort(uint64_t* b, int n) {
  const int lane = threadIdx.x & 63;
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = lane; t < n; t += 64) {
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
      wave_sync();
    }
  }
}

// Squared L2 of an FP32 query (LDS) against a BF16 row stored lane-split:
// for lane h of a pair the row holds, chunk by chunk, the four elements
// 8m + 4h .. 8m + 4h + 3 (so each lane reads its part contiguously).  Same
// accumulator chains and order as l2_pair_lanes, on the widened values.
__device__ inline float l2_pair_lanes_bf16(const float* __restrict__ q, const uint16_t* __restrict__ row, int D,
                                           int h) {
  const int DR = D & 31, DD = D - DR;
  const uint16_t* part = row + size_t(h) * (D / 2);  // this lane's D/2 values
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  auto step = [&](int m) {  // chunk m: elements 8m + 4h + 0..3
    const uint2 raw = *reinterpret_cast<const uint2*>(part + 4 * m);
    const float4 x = *reinterpret_cast<const float4*>(q + 8 * m + 4 * h);
    const float y0 = __uint_as_float(raw.x << 16), y1 = __uint_as_float(raw.x & 0xffff0000u);
    const float y2 = __uint_as_float(raw.y << 16), y3 = __uint_as_float(raw.y & 0xffff0000u);
    float t0 = x.x - y0, t1 = x.y - y1, t2 = x.z - y2, t3 = x.w - y3;
    acc.x = __builtin_fmaf(t0, t0, acc.x);
    acc.y = __builtin_fmaf(t1, t1, acc.y);
    acc.z = __builtin_fmaf(t2, t2, acc.z);
    acc.w = __builtin_fmaf(t3, t3, acc.w);
  };
  for (int m = DD / 8; m < D / 8; ++m) step(m);
#pragma unroll 8
  for (int m = 0; m < DD / 8; ++m) step(m);
  float s0 = ((acc.x + acc.y) + acc.z) + acc.w;
  float p = __shfl_xor(s0, 1);
  return (((p + acc.x) + acc.y) + acc.z) + acc.w;
}

}  // namespace

// BF16 copy of the base rows in the lane-split layout of l2_pair_lanes_bf16
// (round to nearest even).
__global__ void k_pack_bf16(const float* __restrict__ X, int D, size_t n, uint16_t* __restrict__ out) {
  const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n * size_t(D)) return;
  const size_t r = gid / D;
  const int e = int(gid % D);
  const int m = e / 8, w = e % 8, h = w / 4, i = w % 4;
  const uint32_t u = __float_as_uint(X[gid]);
  const uint32_t rounded = u + 0x7fffu + ((u >> 16) & 1u);
  out[r * D + size_t(h) * (D / 2) + 4 * m + i] = uint16_t(rounded >> 16);
}

// Dynamic LDS layout for the search kernel; also used by the host to size it.
// LDS of one query's common state (the layout of Smem), before the arrays
// of the learned and method 1 searches
__host__ __device__ inline size_t search_lds_common(const SearchCfg& c, bool global_table) {
  size_t b = 0;
  auto a16 = [](size_t x) { return (x + 15) & ~size_t(15); };
  b += a16(sizeof(float) * c.D);
  b += a16(sizeof(uint64_t) * c.kb);
  if (!global_table) b += a16(sizeof(uint32_t) << c.logcap);
  b += a16(sizeof(uint32_t) * ctab_slots(c.threads)) * 2;
  b += a16(sizeof(int32_t) * (c.nleaf + 1)) + a16(sizeof(int32_t) * c.nleaf);
  if (c.method != 1) b += a16(sizeof(uint32_t) * c.E) + a16(sizeof(int32_t) * (c.E + 1));  // method 0 epochs
  b += a16(sizeof(uint32_t) * 256) + a16(sizeof(int) * 4) + a16(sizeof(int) * kMiscInts);
  return b;
}

__host__ __device__ inline size_t search_lds_bytes(const SearchCfg& c, bool global_table) {
  size_t b = search_lds_common(c, global_table);
  auto a16 = [](size_t x) { return (x + 15) & ~size_t(15); };
  if (c.stop) {
    b += a16(sizeof(int32_t) * c.K);            // top k ids of the last checkpoint
    b += a16(sizeof(float) * (kFeat + 2 * kHid + 8 + kHid + 1 + 2 * kFeat));  // features, hidden units,
                                                                             // history, W3 | b3, norm
  }
  if (c.method == 1) {
    const int kc = c.K + c.E + 1;
    b += a16(4 * size_t(kc)) * 2 + a16(size_t(kc));            // knn: dist, id, flag
    b += a16(4 * size_t(c.K + 1)) * 2 + a16(size_t(c.K + 1));  // results
    b += a16(8 * size_t(next_pow2(c.E < 2 ? 2 : c.E)));        // epoch seed order
    b += a16(8 * 64);                                           // new neighbours
  }
  return b;
}

// Probe search state for the learned stopping rule (c.stop != 0).
struct LearnArgs {
  const uint32_t* wpack;     // network: packed BF16 columns, then W3 | b3 as floats
  const float* norm;         // feature mean | 1 / deviation
  float zthr;                // stop when the logit reaches zthr
  int32_t* fb_list;          // stop 1: queries left to the full search
  int* fb_count;             // stop 1: [count, running total of fallbacks, running total of queries]
  float* tr_feat;            // stop 2: [query][checkpoint][kFeat]
  uint8_t* tr_label;         // stop 2: [query][checkpoint], top k ids found in plain_ids (254 at most), 255 = not reached
  const int32_t* plain_ids;  // stop 2: [query][K]
  int p_pool, p_ext, p_epochs;  // stop 3: the probe phase's init_num, extend_to and epochs
  // batched search in one pass (kBatch): the configuration of every class and
  // the class of every query
  const SearchCfg* ccfg;
  const uint8_t* qcls;
};

// Probe-search checkpoint: features of the sorted pool P (R entries); stop 2
// writes them with the label, stop 1 and 3 evaluate the network and return
// true when its logit reaches la.zthr.  e: checkpoint index, nexp_last: ids
// expanded in the last epoch, visited: ids measured so far.  Called by every
// thread of the workgroup after a barrier; wavefront 0 does all the work
// (ballot counts, one lane per hidden unit, layer inputs passed between lanes
// with readlane, sums in the same order as the training forward pass), and
// one barrier publishes the decision.  Kept out of line so that the search
// loop around it is register-allocated as in the plain kernel.
template <int kStop>
__device__ __noinline__ bool ckpt_eval(int cE, int cK, int cepochs, int clogcap, const uint32_t* __restrict__ wpack,
                                       float zthr, float* tr_feat, uint8_t* tr_label, const int32_t* plain_ids,
                                       const uint32_t* tab, int* misc, const uint64_t* P, int R, int32_t* prevk,
                                       float* ckf, int qi, int e, int nexp_last, int visited) {
  const int tid = threadIdx.x;
  auto rl = [](float x, int j) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j)); };
  if (tid < 64) {
    const int lane = tid;
    // stop 1 and 3: lane j loads the packed W1 column and biases of hidden
    // unit j now; the loads complete while the pool statistics are gathered
    constexpr int kW1Words = kFeat / 2;
    uint32_t wreg[kW1Words + 1];
    if (kStop != 2) {
#pragma unroll
      for (int i = 0; i < kW1Words; ++i) wreg[i] = wpack[i * kHid + lane];
      wreg[kW1Words] = wpack[(kPackWords - 1) * kHid + lane];
    }
    const int ER = cE < R ? cE : R;
    int fr = 0;  // unexpanded ids among the first extend_to
    for (int b0 = 0; b0 < ER; b0 += 64) {
      const int j = b0 + lane;
      bool v = false;
      if (j < ER) {
        bool pres;
        const int sl = tab_find(tab, clogcap, key_id(P[j]), &pres);
        v = (tab[sl] & kExp) == 0;
      }
      fr += __popcll(__ballot(v));
    }
    int ch = cK;  // top k ids that were not there at the last checkpoint
    if (e > 0) {
      ch = 0;
      for (int b0 = 0; b0 < cK; b0 += 64) {
        const int j = b0 + lane;
        bool v = false;
        if (j < cK) {
          const int32_t id = int32_t(key_id(P[j]));
          v = true;
          for (int t2 = 0; t2 < cK; ++t2)
            if (prevk[t2] == id) {
              v = false;
              break;
            }
        }
        ch += __popcll(__ballot(v));
      }
    }
    int match = 0;  // stop 2: top k ids that are in the plain result's top k
    if (kStop == 2) {
      const int32_t* pl = plain_ids + size_t(qi) * cK;
      for (int b0 = 0; b0 < cK; b0 += 64) {
        const int j = b0 + lane;
        bool v = false;
        if (j < cK) {
          const int32_t id = int32_t(key_id(P[j]));
          for (int t2 = 0; t2 < cK; ++t2)
            if (pl[t2] == id) {
              v = true;
              break;
            }
        }
        match += __popcll(__ballot(v));
      }
    }
    wave_sync();  // prevk is read before it is replaced
    for (int j = lane; j < cK; j += 64) prevk[j] = int32_t(key_id(P[j]));
    float* hist = ckf + kFeat + 2 * kHid;  // k-th [0..2] and best [3..5] distance, latest first
    if (lane == 0) {
      CkptState st;
      st.e = e;
      st.d1 = key_dist(P[0]);
      st.dk = key_dist(P[cK - 1]);
      st.dE = key_dist(P[(ER > 0 ? ER : 1) - 1]);
      st.dR = key_dist(P[R - 1]);
      for (int j = 0; j < 3; ++j) {
        st.dk_hist[j] = hist[j];
        st.d1_hist[j] = hist[3 + j];
      }
      st.visited = float(visited);
      st.expanded = float(nexp_last);
      st.changed = float(ch);
      st.frontier = float(fr);
      ckpt_features(st, ckf);
      hist[2] = hist[1];
      hist[1] = hist[0];
      hist[0] = st.dk;
      hist[5] = hist[4];
      hist[4] = hist[3];
      hist[3] = st.d1;
    }
    wave_sync();
    if (kStop == 2) {
      const size_t slot = size_t(qi) * (cepochs + 1) + e;
      if (lane < kFeat) tr_feat[slot * kFeat + lane] = ckf[lane];
      if (lane == 0) tr_label[slot] = uint8_t(match < 254 ? match : 254);
    } else {
      const float* w3s = ckf + kFeat + 2 * kHid + 8;  // W3 | b3
      const float* nrm = w3s + kHid + 1;               // feature mean | 1 / deviation
      const uint32_t* __restrict__ w2 = wpack + (kFeat / 2) * kHid;
      auto lo = [](uint32_t v) { return __uint_as_float(v << 16); };
      auto hi = [](uint32_t v) { return __uint_as_float(v & 0xffff0000u); };
      float a = lo(wreg[kW1Words]);
#pragma unroll
      for (int i = 0; i < kFeat; ++i)
        a = __builtin_fmaf((i & 1) ? hi(wreg[i / 2]) : lo(wreg[i / 2]), mlp_input(ckf, nrm, i), a);
      const float h1 = a > 0.f ? a : 0.f;
      a = hi(wreg[kW1Words]);
#pragma unroll 8
      for (int i = 0; i < kHid; i += 2) {
        const uint32_t v = w2[(i / 2) * kHid + lane];
        a = __builtin_fmaf(lo(v), rl(h1, i), a);
        a = __builtin_fmaf(hi(v), rl(h1, i + 1), a);
      }
      const float h2 = a > 0.f ? a : 0.f;
      const float w3 = w3s[lane];
      float z = w3s[kHid];
#pragma unroll 8
      for (int j = 0; j < kHid; ++j)
        z = __builtin_fmaf(rl(w3, j), rl(h2, j), z);
      if (lane == 0) misc[7] = z >= zthr ? 1 : 0;
    }
  }
  __syncthreads();  // wavefront 0 has read P and the table before they change
  if (kStop == 2) return false;
  const bool stop = misc[7] != 0;
  __syncthreads();  // misc[7] is read before it is reused
  return stop;
}

// The BF16-row variant is held to 128 VGPRs (four wavefronts per SIMD, which
// its LDS allows); the others keep their registers for loads in flight (a cap
// of four wavefronts made the plain kernel 40 % slower at pool 500).
template <bool kGlobal, int kMethod, bool kBf16, int kStop, int kThreads, bool kSpill = false, bool kBatch = false>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kBf16 || kStop == 3 ? 4 : 1)))
k_search_nnexp(SearchCfg c, LearnArgs la, const float* __restrict__ Q,
                                                         size_t q_stride,
                                                         const float* __restrict__ X, const uint16_t* __restrict__ Xb,
                                                         const int32_t* __restrict__ leaves,
                                                         const SNode* __restrict__ sn,
                                                         const uint32_t* __restrict__ leaf_lists,
                                                         const uint64_t* __restrict__ goff,
                                                         const uint32_t* __restrict__ gids, int32_t* __restrict__ out_ids,
                                                         float* __restrict__ out_d, uint64_t* __restrict__ ndist_q,
                                                         unsigned long long* __restrict__ ndist_total,
                                                         const int32_t* __restrict__ qlist, int nlist_host,
                                                         const int* __restrict__ nlist_dev, int xcd_slices,
                                                         uint32_t* __restrict__ gtab, int32_t* __restrict__ overflow,
                                                         int* __restrict__ noverflow,
                                                         unsigned long long* __restrict__ work,
                                                         uint32_t* __restrict__ spill, int spill_logcap,
                                                         int* __restrict__ qctr) {
  const int nlist = nlist_dev ? *nlist_dev : nlist_host;
  constexpr int kCtab = ctab_slots(kThreads), kNB = new_buf(kThreads);
  constexpr int kCtabLog = kThreads == 64 ? 7 : (kThreads == 128 ? 8 : 9);
  static_assert((1 << kCtabLog) == kCtab, "chunk table size");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Smem s;
  {
    size_t o = 0;
    auto take = [&](size_t bytes) {
      unsigned char* p = smem_raw + o;
      o += align16(bytes);
      return p;
    };
    s.q = reinterpret_cast<float*>(take(sizeof(float) * c.D));
    s.buf = reinterpret_cast<uint64_t*>(take(sizeof(uint64_t) * c.kb));
    if (!kGlobal) s.tab = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) << c.logcap));
    else s.tab = gtab + (size_t(blockIdx.x) << c.logcap);
    s.ctab_id = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * kCtab));
    s.ctab_pos = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * kCtab));
    s.lpos = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * (c.nleaf + 1)));
    s.lstart = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * c.nleaf));
    s.explist = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * (kMethod == 1 ? 0 : c.E)));
    s.expoff = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * (kMethod == 1 ? 0 : c.E + 1)));
    s.hist = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * 256));
    s.scan = reinterpret_cast<int*>(take(sizeof(int) * 4));
    s.misc = reinterpret_cast<int*>(take(sizeof(int) * kMiscInts));
  }
  // learned stopping working arrays (after the common layout)
  int32_t* prevk = nullptr;
  float* ckf = nullptr;  // features [kFeat] | h1 [kHid] | h2 [kHid] | history [8] | W3 | b3 | norm
  if (kStop) {
    SearchCfg c0 = c;
    c0.stop = 0;
    c0.method = 0;
    size_t o = search_lds_bytes(c0, kGlobal);
    auto take = [&](size_t bytes) {
      unsigned char* p = smem_raw + o;
      o += align16(bytes);
      return p;
    };
    prevk = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * c.K));
    ckf = reinterpret_cast<float*>(take(sizeof(float) * (kFeat + 2 * kHid + 8 + kHid + 1 + 2 * kFeat)));
  }
  float* w3s = ckf ? ckf + kFeat + 2 * kHid + 8 : nullptr;  // W3 | b3
  float* nrm = w3s ? w3s + kHid + 1 : nullptr;              // feature mean | 1 / deviation
  if (kStop == 1 || kStop == 3) {
    const float* w3 = reinterpret_cast<const float*>(la.wpack + kPackLen);
    for (int i = threadIdx.x; i <= kHid; i += kThreads) w3s[i] = w3[i];
    for (int i = threadIdx.x; i < 2 * kFeat; i += kThreads) nrm[i] = la.norm[i];
    __syncthreads();
  }
  // method 1 (kgraph) working arrays
  float* kd = nullptr;
  uint32_t* ki = nullptr;
  uint8_t* kf = nullptr;
  float* rd = nullptr;
  uint32_t* ri = nullptr;
  uint8_t* rf = nullptr;
  uint64_t* eord = nullptr;
  uint64_t* nwl = nullptr;
  if (kMethod == 1) {
    size_t o = search_lds_common(c, kGlobal);
    auto take = [&](size_t bytes) {
      unsigned char* p = smem_raw + o;
      o += align16(bytes);
      return p;
    };
    const int kc = c.K + c.E + 1;
    kd = reinterpret_cast<float*>(take(4 * size_t(kc)));
    ki = reinterpret_cast<uint32_t*>(take(4 * size_t(kc)));
    kf = reinterpret_cast<uint8_t*>(take(size_t(kc)));
    rd = reinterpret_cast<float*>(take(4 * size_t(c.K + 1)));
    ri = reinterpret_cast<uint32_t*>(take(4 * size_t(c.K + 1)));
    rf = reinterpret_cast<uint8_t*>(take(size_t(c.K + 1)));
    eord = reinterpret_cast<uint64_t*>(take(8 * size_t(next_pow2(c.E < 2 ? 2 : c.E))));
    nwl = reinterpret_cast<uint64_t*>(take(8 * 64));
  }
  const int tid = threadIdx.x;
  // the visited table of the current query: the LDS (or global-pass)
  // table; method 0 moves a query that is about to outgrow the LDS table
  // into this workgroup's spill table in global memory (when given) and
  // continues there instead of handing it to the global-table pass
  uint32_t* const tab0 = s.tab;
  int logcap = c.logcap;
  int cap = 1 << logcap;
  int limit = cap - cap / 4;  // keep the table under 3/4 load
  unsigned long long acc_spill = 0;
  const int pr = tid >> 1, h = tid & 1;  // lane pair for distances
  // the pool: R entries sorted in Rs slots.  The continuation form (stop 3)
  // runs its probe phase with the probe's smaller pool in the same layout
  // (N and O after the full Rp slots); the other forms keep them constant.
  const int Rp = c.Rp;
  const int lstride = c.nleaf;  // kBatch: leaf slots per query (the classes' largest)
  std::conditional_t<kStop == 3 || kBatch, int, const int> R = c.R;
  int Rs = c.Rp;  // stop 3 only
  uint64_t* P = s.buf;
  uint64_t* N = s.buf + Rp;
  uint64_t* O = s.buf + Rp + kNB;

  // distances for keys a[from, to) whose low 32 bits hold ids
  auto distances = [&](uint64_t* a, int from, int to) {
    if (kBf16) {
      for (int i = from + pr; i < to; i += kThreads / 2) {
        const uint32_t id = key_id(a[i]);
        const float dd = l2_pair_lanes_bf16(s.q, Xb + size_t(id) * c.D, c.D, h);
        if (h) a[i] = make_key(dd, id);
      }
    } else {
#pragma unroll 2
      for (int i = from + pr; i < to; i += kThreads / 2) {  // two rows in flight per lane pair
        const uint32_t id = key_id(a[i]);
        const float dd = l2_pair_lanes(s.q, X + size_t(id) * c.D, c.D, h);
        if (h) a[i] = make_key(dd, id);
      }
    }
  };
  // merge N[0, m) (unique, unsorted, disjoint from P) into the sorted pool P[0, R)
  auto merge_new = [&](int m) {
    const int mp = next_pow2(m < 2 ? 2 : m);
    if (mp <= kThreads || kStop == 3) {
      // sort N, then every key's place is its index plus its rank in the
      // other sorted array (binary search); the continuation form also
      // sorts a batch above the block this way (in LDS): its probe phase
      // meets such batches at 256 threads
      for (int i = m + tid; i < mp; i += kThreads) N[i] = kKeyMax;
      if constexpr (kStop == 3) {
        if (mp <= kThreads) {
          block_bitonic_sort_reg(N, mp);
        } else {
          __syncthreads();
          block_bitonic_sort(N, mp);
        }
      } else {
        block_bitonic_sort_reg(N, mp);
      }
      for (int i = tid; i < m; i += kThreads) {
        const uint64_t key = N[i];
        int lo = 0, hi = R;  // pool keys below key
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (P[mid] < key) lo = mid + 1; else hi = mid;
        }
        if (lo + i < R) O[lo + i] = key;
      }
      for (int i = tid; i < R; i += kThreads) {
        const uint64_t key = P[i];
        int lo = 0, hi = m;  // new keys below key
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (N[mid] < key) lo = mid + 1; else hi = mid;
        }
        if (i + lo < R) O[i + lo] = key;
      }
      __syncthreads();
      for (int i = tid; i < R; i += kThreads) P[i] = O[i];
      __syncthreads();
      return;
    }
    for (int i = tid; i < m; i += kThreads) {
      uint64_t key = N[i];
      int lo = 0, hi = R;  // number of pool keys below key
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (P[mid] < key) lo = mid + 1; else hi = mid;
      }
      int rn = 0;
      for (int j = 0; j < m; ++j) rn += (N[j] < key) ? 1 : 0;
      if (lo + rn < R) O[lo + rn] = key;
    }
    for (int i = tid; i < R; i += kThreads) {
      uint64_t key = P[i];
      int rn = 0;
      for (int j = 0; j < m; ++j) rn += (N[j] < key) ? 1 : 0;
      if (i + rn < R) O[i + rn] = key;
    }
    __syncthreads();
    for (int i = tid; i < R; i += kThreads) P[i] = O[i];
    __syncthreads();
  };

  // Probe-search checkpoint (all threads), out of line (see ckpt_eval).
  auto checkpoint = [&](int qi, int e, int nexp_last, int visited) -> bool {
    return ckpt_eval<kStop>(kStop == 3 ? la.p_ext : c.E, c.K, c.epochs, c.logcap, la.wpack, la.zthr, la.tr_feat,
                            la.tr_label, la.plain_ids, s.tab, s.misc, P, R, prevk, ckf, qi, e, nexp_last, visited);
  };

  // work distribution: the query order in slices, one per XCD (so queries
  // that share rows share that XCD's L2).  A workgroup's first query is its
  // index in the slice.  Workgroups of 256 threads draw every later one
  // from the slice's counter, so those that drew cheap queries take more
  // (pool 800, 2 epochs: 1.34 -> 1.02 ms per 10K queries); smaller ones step
  // through the slice, since the counter's atomic, which a wavefront's
  // loads wait behind, costs them more than the balance gains.
  constexpr bool kDyn = kThreads == 256;
  __shared__ int wnext_s;
  int wbase = 0, wend = nlist, nwg = gridDim.x, wfirst = blockIdx.x;
  int* wctr = qctr;
  if (xcd_slices && (gridDim.x & 7) == 0) {
    const int x = blockIdx.x & 7;
    wbase = int((long long)nlist * x / 8);
    wend = int((long long)nlist * (x + 1) / 8);
    nwg = gridDim.x >> 3;
    wfirst = blockIdx.x >> 3;
    wctr = qctr + x;
  }
  int wn = 0;  // thread 0: the next query, drawn while this one runs
  unsigned long long acc_nd = 0, acc_seed = 0, acc_exp = 0;  // thread 0: this workgroup's counters
  int acc_q = 0, acc_fb = 0;
  for (int w = wbase + wfirst; w < wend; w = kDyn ? wnext_s : w + nwg) {
    const int qi = qlist ? qlist[w] : w;
    if constexpr (kBatch) {  // this query's class: pool, extend_to, k, leaves, table size
      // read through the constant address space: the configuration is
      // uniform and read-only, so it lands in scalar registers as a kernel
      // argument does
      // (only the fields that differ between classes: the others stay the
      // kernel argument's, loop invariant)
      using ConstU8 = const __attribute__((address_space(4))) uint8_t;
      using ConstI32 = const __attribute__((address_space(4))) int32_t;
      const int cls = __builtin_amdgcn_readfirstlane(int(((ConstU8*)la.qcls)[__builtin_amdgcn_readfirstlane(qi)]));
      const SearchCfg* cc = la.ccfg + cls;
      auto field = [&](const int* f) { return *(ConstI32*)f; };
      c.lsize = field(&cc->lsize);
      c.nleaf = field(&cc->nleaf);
      c.init_num = field(&cc->init_num);
      c.E = field(&cc->E);
      c.K = field(&cc->K);
      c.R = field(&cc->R);
      c.Rp = field(&cc->Rp);
      c.logcap = field(&cc->logcap);
      R = c.R;
      Rs = c.Rp;
      logcap = c.logcap;
      cap = 1 << logcap;
      limit = cap - cap / 4;
    }
    const float* qg = Q + size_t(qi) * q_stride;
    for (int i = tid; i < c.D; i += kThreads) s.q[i] = (i < c.d) ? qg[i] : 0.f;
    if (kSpill && s.tab != tab0) {  // the last query moved to the spill table
      s.tab = tab0;
      logcap = c.logcap;
      cap = 1 << logcap;
      limit = cap - cap / 4;
    }
    for (int i = tid; i < cap; i += kThreads) s.tab[i] = kEmpty;
    // leaf sequence: sizes and prefix in (rank, tree) order
    int run = 0;
    for (int b0 = 0; b0 < c.nleaf; b0 += kThreads) {
      int li = b0 + tid;
      int sz = 0;
      if (li < c.nleaf) {
        int nd = leaves[size_t(qi) * (kBatch ? lstride : c.nleaf) + li];
        int st = 0;
        if (nd >= 0) {
          st = int(sn[nd].start);
          sz = int(sn[nd].end - sn[nd].start);
        }
        s.lstart[li] = st;
      }
      int tot;
      int ex = block_exclusive_scan(sz, s.scan, &tot);
      if (li < c.nleaf) s.lpos[li] = run + ex;
      run += tot;
    }
    if (tid == 0) s.lpos[c.nleaf] = run;
    __syncthreads();
    const int total = run;

    // ---- seed pool -------------------------------------------------------
    // The continuation form (stop 3) seeds in two phases: the probe's pool
    // now; for a query the network does not stop, the rest of the full pool
    // later (from the sequence position where this scan stopped, into the
    // area after the pool).  Seeds go to sb[0, want).
    int p = 0;  // kept so far (uniform)
    int want = kStop == 3 ? la.p_pool : 0;  // stop 3 only
    uint64_t* sb = s.buf;                     // stop 3 only
    if constexpr (kStop == 3) {  // the probe phase's pool
      R = la.p_ext > c.K ? la.p_ext : c.K;
      Rs = next_pow2(R < 2 ? 2 : R);
    }
    auto seed_chunk = [&](bool valid, uint32_t id, uint32_t sq) {
      for (int i = tid; i < kCtab; i += kThreads) { s.ctab_id[i] = kEmpty; s.ctab_pos[i] = kEmpty; }
      __syncthreads();
      bool present = false;
      int cs = -1;
      if (valid) {
        tab_find(s.tab, logcap, id, &present);
        if (!present) {
          uint32_t sl = hslot(id, kCtabLog);
          for (;;) {
            uint32_t v = s.ctab_id[sl];
            if (v == kEmpty) {
              uint32_t old = atomicCAS(&s.ctab_id[sl], kEmpty, id);
              if (old == kEmpty || old == id) break;
              v = old;
            }
            if (v == id) break;
            sl = (sl + 1) & (kCtab - 1);
          }
          cs = int(sl);
          atomicMin(&s.ctab_pos[sl], sq);
        }
      }
      __syncthreads();
      int first = (valid && !present && s.ctab_pos[cs] == sq) ? 1 : 0;
      int cnt;
      int rank = block_exclusive_scan(first, s.scan, &cnt);
      if (first && p + rank < (kStop == 3 ? want : c.init_num)) {
        tab_insert(s.tab, logcap, id);
        (kStop == 3 ? sb : s.buf)[p + rank] = uint64_t(id);
        if constexpr (kStop == 3)
          if (p + rank == want - 1) s.misc[8] = int(sq) + 1;  // the position after the last seed
      }
      __syncthreads();
      p = min(kStop == 3 ? want : c.init_num, p + cnt);
    };
    int seq_read = 0;
    auto seed_id = [&](int sq) -> uint32_t {  // id at position sq of the leaf sequence
      int lo = 0, hi = c.nleaf;  // largest li with lpos[li] <= sq
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (s.lpos[mid] <= sq) lo = mid; else hi = mid;
      }
      const int t = lo % c.T;
      return leaf_lists[size_t(t) * c.n + s.lstart[lo] + (sq - s.lpos[lo])];
    };
    // A chunk that cannot fill the pool admits every id not yet in it, in
    // any order: method 0 ranks the pool by key, so only the chunk that
    // fills it needs the first occurrences in sequence order (seed_chunk).
    // Its count alternates between misc[6] and misc[7], each cleared while
    // the other is in use.
    auto seed_fast = [&](bool valid, uint32_t id, int par) {
      if (valid && tab_insert(s.tab, logcap, id)) (kStop == 3 ? sb : s.buf)[p + atomicAdd(&s.misc[6 + par], 1)] = uint64_t(id);
      __syncthreads();
      const int cnt = s.misc[6 + par];
      if (tid == 0) s.misc[7 - par] = 0;
      p += cnt;
    };
    if (tid == 0) {
      s.misc[6] = 0;
      s.misc[7] = 0;
    }
    __syncthreads();
    int par = 0;
    int resume = 0;  // stop 3: the position after the last id taken or passed over
    // the next chunk's ids are loaded before this chunk is processed
    uint32_t id_next = tid < total ? seed_id(tid) : 0u;
    for (int b0 = 0; b0 < total && p < (kStop == 3 ? want : c.init_num); b0 += kThreads) {
      seq_read = min(total, b0 + kThreads);
      const int sq = b0 + tid;
      const bool valid = sq < total;
      const uint32_t id = id_next;
      if (sq + kThreads < total) id_next = seed_id(sq + kThreads);
      if constexpr (kStop == 3) resume = seq_read;
      if (kMethod != 1 && p + kThreads <= (kStop == 3 ? want : c.init_num)) {
        seed_fast(valid, id, par);
        par ^= 1;
      } else {
        seed_chunk(valid, id, uint32_t(sq));
        if constexpr (kStop == 3)
          if (p >= want) resume = s.misc[8];
      }
    }
    // the hash-only search tops up to K only (hashing_index.hpp:983-991)
    // and measures the candidates it has: the ids a chunk adds past K are
    // left out of the count
    const int pseq = p;
    for (uint32_t r0 = 0; p < (kMethod == 2 ? min(c.K, c.init_num) : (kStop == 3 ? want : c.init_num)); r0 += kThreads) {
      uint32_t id = uint32_t(hash4(c.seed, ST_SEARCH, uint64_t(qi), r0 + tid) % uint64_t(c.n));
      seed_chunk(true, id, r0 + tid);
    }
    const int ninit = kMethod == 2 ? max(pseq, min(c.K, c.init_num)) : (kStop == 3 ? want : c.init_num);
    int inserted = ninit;
    uint64_t nexp_total = 0;
    const uint64_t seeds_read = uint64_t(seq_read);
    // the next query: the atomic is issued here so that the seed
    // distances' loads, which complete after it, cover its latency
    if (kDyn && tid == 0) wn = wbase + nwg + atomicAdd(wctr, 1);
    distances(s.buf, 0, ninit);
    if (kMethod == 2)  // slots past the candidates sort last
      for (int i = ninit + tid; i < c.init_num; i += kThreads) s.buf[i] = kKeyMax;
    __syncthreads();
    // best R seeds: select the R-th key, gather the R keys, sort them
    // (method 1 without sorting keeps the seed order, kdtreeub_index.hpp:567-570)
    if (kMethod == 0 || kMethod == 2 || c.sorted) {
      const uint64_t thr = (R < (kStop == 3 ? want : c.init_num))
                               ? block_select(s.buf, kStop == 3 ? want : c.init_num, R, s.hist, s.misc)
                               : kKeyMax;
      uint64_t mine[kKPT];
      int nm = 0;
#pragma unroll
      for (int j = 0; j < kKPT; ++j) {
        const int i = tid + j * kThreads;
        mine[j] = (i < (kStop == 3 ? want : c.init_num)) ? s.buf[i] : kKeyMax;
        nm += (i < (kStop == 3 ? want : c.init_num) && mine[j] <= thr) ? 1 : 0;
      }
      int tot;
      int o = block_exclusive_scan(nm, s.scan, &tot);  // barriers inside: all reads precede the writes
#pragma unroll
      for (int j = 0; j < kKPT; ++j) {
        const int i = tid + j * kThreads;
        if (i < (kStop == 3 ? want : c.init_num) && mine[j] <= thr) P[o++] = mine[j];
      }
      for (int i = R + tid; i < (kStop == 3 || kBatch ? Rs : Rp); i += kThreads) P[i] = kKeyMax;
      __syncthreads();
      if ((kStop == 3 || kBatch ? Rs : Rp) <= kThreads) block_bitonic_sort_reg(P, kStop == 3 || kBatch ? Rs : Rp);
      else block_bitonic_sort(P, kStop == 3 || kBatch ? Rs : Rp);
    }
    uint64_t ndq = uint64_t(ninit);
    bool ovf = false;
    bool stopped = false;
    if (kStop) stopped = checkpoint(qi, 0, 0, inserted);

    int kg_valid = 0;  // method 1: valid result entries
    if (kMethod == 1) {
      // ---- method 1: KGraph-style expansion (kdtreeub_index.hpp:572-648) -----
      for (int i = tid; i < cap; i += kThreads) s.tab[i] = kEmpty;  // flags.reset() (572)
      __syncthreads();
      const uint64_t* pool = (c.sorted) ? P : s.buf;
      if (tid < 64) {
        const int lane = tid;
        const int kcap = c.K + c.E + 1;
        int rsize = 0;  // results.size() (spare included)
        int ins = 0;
        for (int it = 0; it < c.epochs && !ovf; ++it) {
          // this epoch's seeds not yet flagged, ordered by (distance, pool rank)
          int L = 0;
          for (int b = 0; b < c.E; b += 64) {
            const int j = b + lane;
            bool take = false;
            uint64_t e = 0;
            if (j < c.E) {
              e = pool[size_t(it) * c.E + j];
              take = tab_insert(s.tab, logcap, key_id(e));
            }
            const uint64_t mk = __ballot(take);
            if (take) {
              const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
              eord[L + __popcll(mk & lt)] = (e & 0xffffffff00000000ull) | uint32_t(it * c.E + j);
            }
            L += __popcll(mk);
            ins += __popcll(mk);
          }
          const int p2 = next_pow2(L < 2 ? 2 : L);
          for (int t = L + lane; t < p2; t += 64) eord[t] = kKeyMax;
          wave_sync();
          wave_bitonic_sort(eord, p2);
          for (int t = lane; t < L; t += 64) {
            const uint64_t e = eord[t];
            kd[t] = key_dist(e);
            ki[t] = key_id(pool[uint32_t(e)]);
            kf[t] = 1;
          }
          wave_sync();
          int k = 0;
          // on a graph whose rows all have one degree (at most 64), the
          // neighbour ids of the next node to expand (the first flagged entry
          // after this one as the list stands) are loaded, one per lane,
          // while this node's distances are measured; pf_node is the node
          // they belong to, so a list changed by the inserts only costs the
          // load
          const bool pf_on = c.deg > 0 && c.deg <= 64;
          uint32_t pf_node = 0xffffffffu, pf_id = 0xffffffffu;
          while (k < L) {
            // an entry without its flag only moves k on: go straight to the
            // first flagged entry at or after k (64 entries per ballot)
            if (!kf[k]) {
              int nxt = L;
              for (int b = k; b < L; b += 64) {
                const int j = b + lane;
                const uint64_t mk = __ballot(j < L && kf[j] != 0);
                if (mk) {
                  nxt = b + __ffsll((unsigned long long)mk) - 1;
                  break;
                }
              }
              k = nxt;
              continue;
            }
            int nk = L;
            if (kf[k]) {
              wave_sync();
              if (lane == 0) kf[k] = 0;
              const uint32_t nd = ki[k];
              uint64_t g0, deg;
              if (c.deg > 0) {
                g0 = uint64_t(nd) * uint64_t(c.deg);
                deg = uint64_t(c.deg);
              } else {
                g0 = goff[nd];
                deg = goff[nd + 1] - g0;
              }
              const int maxM = int(deg < uint64_t(c.E) ? deg : uint64_t(c.E));
              const bool have = pf_on && nd == pf_node;
              const uint32_t id0 = have ? pf_id : 0xffffffffu;
              if (pf_on) {
                int nxt = -1;
                for (int b = k + 1; b < L && nxt < 0; b += 64) {
                  const int j = b + lane;
                  const uint64_t mk = __ballot(j < L && kf[j] != 0);
                  if (mk) nxt = b + __ffsll((unsigned long long)mk) - 1;
                }
                pf_node = nxt >= 0 ? ki[nxt] : 0xffffffffu;
                pf_id = (nxt >= 0 && lane < maxM) ? gids[uint64_t(pf_node) * uint64_t(c.deg) + lane] : 0xffffffffu;
              }
              for (int b = 0; b < maxM; b += 64) {
                if (ins + 64 > limit) {
                  ovf = true;
                  break;
                }
                const int m = b + lane;
                uint32_t id = m < maxM ? (have ? id0 : gids[g0 + m]) : 0xffffffffu;
                // first occurrence in this slice and not flagged
                bool dupl = false;
                for (int o2 = 0; o2 < 64 && b + o2 < maxM; ++o2) {
                  const uint32_t other = uint32_t(__shfl(int(id), o2));
                  dupl = dupl || (o2 < lane && other == id);
                }
                bool isnew = false;
                if (m < maxM && !dupl) isnew = tab_insert(s.tab, logcap, id);
                const uint64_t mk = __ballot(isnew);
                const int nn = __popcll(mk);
                ins += nn;
                if (isnew) {
                  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
                  nwl[__popcll(mk & lt)] = uint64_t(id);
                }
                wave_sync();
                // distances, two lanes per row
                for (int r = (lane >> 1); r < nn; r += 32) {
                  const uint32_t rid = uint32_t(nwl[r]);
                  const float dd = l2_pair_lanes(s.q, X + size_t(rid) * c.D, c.D, lane & 1);
                  if (lane & 1) nwl[r] = make_key(dd, rid);
                }
                wave_sync();
                ndq += uint64_t(nn);
                for (int r = 0; r < nn; ++r) {  // inserts in neighbour order
                  const uint64_t e = nwl[r];
                  const int rr = wave_insert(kd, ki, kf, L, key_dist(e), key_id(e));
                  if (L + 1 < kcap) ++L;
                  if (rr < nk) nk = rr;
                }
              }
              if (ovf) break;
            }
            if (nk <= k) k = nk;
            else ++k;
          }
          if (ovf) break;
          if (L > c.K) L = c.K;
          if (rsize == 0) {  // results = first L of knn (+ spare)
            for (int t = lane; t < L; t += 64) {
              rd[t] = kd[t];
              ri[t] = ki[t];
              rf[t] = kf[t];
            }
            rsize = L + 1;
            wave_sync();
          } else {
            for (int l = 0; l < L; ++l) {
              const int rr = wave_insert(rd, ri, rf, rsize - 1, kd[l], ki[l]);
              if (rr < rsize && rsize < c.K + 1) ++rsize;
            }
          }
        }
        if (lane == 0) {
          s.misc[1] = ovf ? 1 : 0;
          s.misc[2] = rsize > 0 ? rsize - 1 : 0;
          s.misc[3] = int(ndq - uint64_t(c.init_num));
        }
      }
      __syncthreads();
      ovf = s.misc[1] != 0;
      kg_valid = s.misc[2];
      ndq = uint64_t(c.init_num) + uint64_t(s.misc[3]);
    }
    // ---- epochs (method 0) ---------------------------------------------------
    // stop 3: the probe phase runs with the probe's extend_to and epochs
    int cE = kStop == 3 ? la.p_ext : 0, cEp = kStop == 3 ? la.p_epochs : 0;  // stop 3 only
    for (int ep = 0; kMethod == 0 && ep < (kStop == 3 ? cEp : c.epochs) && !stopped; ++ep) {
      // expansion list: unexpanded ids among the first extend_to (448-451)
      int nexp = 0;
      for (int b0 = 0; b0 < (kStop == 3 ? cE : c.E); b0 += kThreads) {
        int j = b0 + tid;
        int take = 0;
        uint32_t id = 0;
        int sl = -1;
        if (j < (kStop == 3 ? cE : c.E)) {
          id = key_id(P[j]);
          bool pres;
          sl = tab_find(s.tab, logcap, id, &pres);
          take = (s.tab[sl] & kExp) ? 0 : 1;
        }
        int tot;
        int rk = block_exclusive_scan(take, s.scan, &tot);
        if (take) {
          s.tab[sl] |= kExp;
          s.explist[nexp + rk] = id;
        }
        nexp += tot;
      }
      __syncthreads();
      if (nexp == 0) break;
      nexp_total += uint64_t(nexp);
      int nnb = 0;
      if (c.deg > 0) {  // regular graph: the neighbours of expansion j are at (explist[j] * deg, + deg)
        nnb = nexp * c.deg;
      } else {
        for (int b0 = 0; b0 < nexp; b0 += kThreads) {
          int j = b0 + tid;
          int deg = 0;
          if (j < nexp) {
            uint32_t e = s.explist[j];
            deg = int(goff[e + 1] - goff[e]);
          }
          int tot;
          int ex = block_exclusive_scan(deg, s.scan, &tot);
          if (j < nexp) s.expoff[j] = nnb + ex;
          nnb += tot;
        }
        if (tid == 0) s.expoff[nexp] = nnb;
      }
      if (tid == 0) s.misc[0] = 0;
      __syncthreads();
      int m = 0;
      for (int b0 = 0; b0 < nnb; b0 += kThreads) {
        int sq = b0 + tid;
        if (sq < nnb) {
          uint32_t id;
          if (c.deg > 0) {
            const int j = sq / c.deg;
            id = gids[size_t(s.explist[j]) * c.deg + (sq - j * c.deg)];
          } else {
            int lo = 0, hi = nexp;
            while (hi - lo > 1) {
              int mid = (lo + hi) >> 1;
              if (s.expoff[mid] <= sq) lo = mid; else hi = mid;
            }
            uint32_t e = s.explist[lo];
            id = gids[goff[e] + uint64_t(sq - s.expoff[lo])];
          }
          if (tab_insert(s.tab, logcap, id)) N[atomicAdd(&s.misc[0], 1)] = uint64_t(id);
        }
        __syncthreads();
        const int m2 = s.misc[0];
        inserted += m2 - m;
        m = m2;
        if (kSpill && inserted + kThreads > limit && !kGlobal && kStop == 0 && spill && logcap < spill_logcap) {
          // move the visited set into this workgroup's spill table and go on
          uint32_t* g = spill + (size_t(blockIdx.x) << spill_logcap);
          const int gcap = 1 << spill_logcap;
          for (int i = tid; i < gcap; i += kThreads) g[i] = kEmpty;
          __syncthreads();
          for (int i = tid; i < cap; i += kThreads) {
            const uint32_t v = s.tab[i];
            if (v != kEmpty) tab_put_raw(g, spill_logcap, v);
          }
          __syncthreads();
          s.tab = g;
          logcap = spill_logcap;
          cap = gcap;
          limit = cap - cap / 4;
          if (tid == 0) ++acc_spill;
        }
        if (inserted + kThreads > limit) {
          ovf = true;
          break;
        }
        const bool last = b0 + kThreads >= nnb;
        if (m > 0 && (last || m + kThreads > kNB)) {
          distances(N, 0, m);
          ndq += uint64_t(m);
          __syncthreads();
          merge_new(m);
          m = 0;
          if (tid == 0) s.misc[0] = 0;
          __syncthreads();
        }
      }
      if (ovf) break;
      if (kStop) stopped = checkpoint(qi, ep + 1, nexp, inserted);
    }
    // stop 3: a query the network did not stop continues in place with the
    // rest of the full pool (ids from the sequence position where the probe's
    // scan stopped, then the random top-up, measured and taken with the
    // probe's pool into the best R) and the full search's extend_to and
    // epochs, keeping its visited table (with the expanded flags)
    bool restart = false;  // stop 3: no room to continue in place; the full search runs instead
    if constexpr (kStop == 3) {
      if (!stopped && !ovf) {
        const int R1 = R;
        p = 0;
        want = min(c.init_num - la.p_pool, c.n - inserted);
        if (want + R1 < c.R || inserted + want + kThreads > limit) restart = true;
      }
      if (!stopped && !ovf && !restart) {
        const int R1 = R;
        sb = s.buf + Rp;
        if (tid == 0) {
          s.misc[6] = 0;
          s.misc[7] = 0;
        }
        __syncthreads();
        int par2 = 0;
        uint32_t nxt = resume + tid < total ? seed_id(resume + tid) : 0u;
        for (int b0 = resume; b0 < total && p < want; b0 += kThreads) {
          const int sq = b0 + tid;
          const bool valid = sq < total;
          const uint32_t id = nxt;
          if (sq + kThreads < total) nxt = seed_id(sq + kThreads);
          if (p + kThreads <= want) {
            seed_fast(valid, id, par2);
            par2 ^= 1;
          } else {
            seed_chunk(valid, id, uint32_t(sq));
          }
        }
        for (uint32_t r0 = 0; p < want; r0 += kThreads) {
          uint32_t id = uint32_t(hash4(c.seed, ST_SEARCH, uint64_t(qi), r0 + tid) % uint64_t(c.n));
          seed_chunk(true, id, r0 + tid);
        }
        inserted += want;
        ndq += uint64_t(want);
        distances(sb, 0, want);
        for (int i = tid; i < R1; i += kThreads) sb[want + i] = P[i];
        __syncthreads();
        // best R of the new seeds and the probe's pool
        R = c.R;
        Rs = c.Rp;
        const int cnt = want + R1;
        const uint64_t thr = (R < cnt) ? block_select(sb, cnt, R, s.hist, s.misc) : kKeyMax;
        uint64_t mine[kKPT];
        int nm = 0;
#pragma unroll
        for (int j = 0; j < kKPT; ++j) {
          const int i = tid + j * kThreads;
          mine[j] = (i < cnt) ? sb[i] : kKeyMax;
          nm += (i < cnt && mine[j] <= thr) ? 1 : 0;
        }
        int tot;
        int o = block_exclusive_scan(nm, s.scan, &tot);
#pragma unroll
        for (int j = 0; j < kKPT; ++j) {
          const int i = tid + j * kThreads;
          if (i < cnt && mine[j] <= thr) P[o++] = mine[j];
        }
        for (int i = R + tid; i < Rs; i += kThreads) P[i] = kKeyMax;
        __syncthreads();
        if (Rs <= kThreads) block_bitonic_sort_reg(P, Rs);
        else block_bitonic_sort(P, Rs);
        for (int ep = 0; ep < c.epochs; ++ep) {
          // expansion list: unexpanded ids among the first extend_to (448-451)
          int nexp = 0;
          for (int b0 = 0; b0 < c.E; b0 += kThreads) {
            int j = b0 + tid;
            int take = 0;
            uint32_t id = 0;
            int sl = -1;
            if (j < c.E) {
              id = key_id(P[j]);
              bool pres;
              sl = tab_find(s.tab, logcap, id, &pres);
              take = (s.tab[sl] & kExp) ? 0 : 1;
            }
            int tot;
            int rk = block_exclusive_scan(take, s.scan, &tot);
            if (take) {
              s.tab[sl] |= kExp;
              s.explist[nexp + rk] = id;
            }
            nexp += tot;
          }
          __syncthreads();
          if (nexp == 0) break;
          nexp_total += uint64_t(nexp);
          int nnb = 0;
          if (c.deg > 0) {  // regular graph: the neighbours of expansion j are at (explist[j] * deg, + deg)
            nnb = nexp * c.deg;
          } else {
            for (int b0 = 0; b0 < nexp; b0 += kThreads) {
              int j = b0 + tid;
              int deg = 0;
              if (j < nexp) {
                uint32_t e = s.explist[j];
                deg = int(goff[e + 1] - goff[e]);
              }
              int tot;
              int ex = block_exclusive_scan(deg, s.scan, &tot);
              if (j < nexp) s.expoff[j] = nnb + ex;
              nnb += tot;
            }
            if (tid == 0) s.expoff[nexp] = nnb;
          }
          if (tid == 0) s.misc[0] = 0;
          __syncthreads();
          int m = 0;
          for (int b0 = 0; b0 < nnb; b0 += kThreads) {
            int sq = b0 + tid;
            if (sq < nnb) {
              uint32_t id;
              if (c.deg > 0) {
                const int j = sq / c.deg;
                id = gids[size_t(s.explist[j]) * c.deg + (sq - j * c.deg)];
              } else {
                int lo = 0, hi = nexp;
                while (hi - lo > 1) {
                  int mid = (lo + hi) >> 1;
                  if (s.expoff[mid] <= sq) lo = mid; else hi = mid;
                }
                uint32_t e = s.explist[lo];
                id = gids[goff[e] + uint64_t(sq - s.expoff[lo])];
              }
              if (tab_insert(s.tab, logcap, id)) N[atomicAdd(&s.misc[0], 1)] = uint64_t(id);
            }
            __syncthreads();
            const int m2 = s.misc[0];
            inserted += m2 - m;
            m = m2;
            if (kSpill && inserted + kThreads > limit && !kGlobal && kStop == 0 && spill && logcap < spill_logcap) {
              // move the visited set into this workgroup's spill table and go on
              uint32_t* g = spill + (size_t(blockIdx.x) << spill_logcap);
              const int gcap = 1 << spill_logcap;
              for (int i = tid; i < gcap; i += kThreads) g[i] = kEmpty;
              __syncthreads();
              for (int i = tid; i < cap; i += kThreads) {
                const uint32_t v = s.tab[i];
                if (v != kEmpty) tab_put_raw(g, spill_logcap, v);
              }
              __syncthreads();
              s.tab = g;
              logcap = spill_logcap;
              cap = gcap;
              limit = cap - cap / 4;
              if (tid == 0) ++acc_spill;
            }
            if (inserted + kThreads > limit) {
              ovf = true;
              break;
            }
            const bool last = b0 + kThreads >= nnb;
            if (m > 0 && (last || m + kThreads > kNB)) {
              distances(N, 0, m);
              ndq += uint64_t(m);
              __syncthreads();
              merge_new(m);
              m = 0;
              if (tid == 0) s.misc[0] = 0;
              __syncthreads();
            }
          }
          if (ovf) break;
        }
      }
    }
    if (kBf16 && kMethod == 0 && !ovf) {
      // re-rank the best min(2K, R) by direct FP32 difference
      const int rr = (2 * c.K < c.R) ? 2 * c.K : c.R;
      const int p2 = next_pow2(rr < 2 ? 2 : rr);
      for (int i = pr; i < rr; i += kThreads / 2) {
        const uint32_t id = key_id(P[i]);
        const float dd = l2_pair_lanes(s.q, X + size_t(id) * c.D, c.D, h);
        if (h) O[i] = make_key(dd, id);
      }
      for (int i = rr + tid; i < p2; i += kThreads) O[i] = kKeyMax;
      __syncthreads();
      if (p2 <= kThreads) block_bitonic_sort_reg(O, p2);
      else block_bitonic_sort(O, p2);
      for (int i = tid; i < rr; i += kThreads) P[i] = O[i];
      __syncthreads();
      ndq += uint64_t(rr);
    }
    // stop 1: queries the network did not stop (or whose visited set
    // overflowed) go to the full search, which writes their results
    // stop 3: those that overflowed or found no room to continue
    const bool to_full = (kStop == 1 && (ovf || !stopped)) || (kStop == 3 && (ovf || restart));
    if (kStop == 1 && tid == 0) {
      ++acc_q;
      if (to_full) {
        la.fb_list[atomicAdd(la.fb_count, 1)] = qi;
        ++acc_fb;
      }
    }
    if (kStop == 3 && tid == 0) {
      ++acc_q;
      if (to_full) la.fb_list[atomicAdd(la.fb_count, 1)] = qi;
      if (!stopped) ++acc_fb;  // not stopped: continued (or run from scratch)
    }
    if (to_full) {
      acc_nd += ndq;
    } else if (ovf) {
      // kBatch: the query outgrew its spill table, sized as the global
      // pass's would be, so it has no result either (the row says so)
      if (kStop == 0 && !kBatch && tid == 0) overflow[atomicAdd(noverflow, 1)] = qi;
      if (kBatch && tid == 0) atomicAdd(noverflow + 1, 1);
      if (kGlobal || kBatch) {
        // outgrew the largest global table as well: no result, the row says so
        for (int j = tid; j < c.K; j += kThreads) {
          out_ids[size_t(qi) * c.out_stride + j] = -1;
          if (out_d) out_d[size_t(qi) * c.out_stride + j] = 3.402823466e38f;
        }
      }
    } else {
      for (int j = tid; j < c.K; j += kThreads) {
        if (kMethod != 1) {
          uint64_t k = P[j];
          out_ids[size_t(qi) * c.out_stride + j] = int32_t(key_id(k));
          if (out_d) out_d[size_t(qi) * c.out_stride + j] = key_dist(k);
        } else {
          const bool ok = j < kg_valid;
          out_ids[size_t(qi) * c.out_stride + j] = ok ? int32_t(ri[j]) : -1;
          if (out_d) out_d[size_t(qi) * c.out_stride + j] = ok ? rd[j] : 3.402823466e38f;
        }
      }
      if (tid == 0) {
        if (ndist_q) ndist_q[qi] = ndq;
        acc_nd += ndq;
        acc_seed += seeds_read;
        acc_exp += nexp_total;
      }
    }
    // read after the barrier; written again only after the next query's
    // barriers
    if (kDyn && tid == 0) wnext_s = wn;
    __syncthreads();
  }
  // counters: one atomic per workgroup (same-address atomics per query would
  // serialise the whole grid)
  if (tid == 0 && acc_spill) atomicAdd(noverflow + 2, int(acc_spill));
  if (tid == 0) {
    if (acc_nd | acc_seed | acc_exp) {
      atomicAdd(ndist_total, acc_nd);
      atomicAdd(ndist_total + 1, acc_seed);
      atomicAdd(ndist_total + 2, acc_exp);
      if (work) {
        atomicAdd(work, acc_nd);
        atomicAdd(work + 1, acc_seed);
        atomicAdd(work + 2, acc_exp);
      }
    }
    if ((kStop == 1 || kStop == 3) && acc_q) {
      atomicAdd(la.fb_count + 1, acc_fb);
      atomicAdd(la.fb_count + 2, acc_q);
      atomicAdd(la.fb_count + 3, acc_fb);
    }
  }
}

namespace {
int ceil_log2(size_t x) {
  int l = 0;
  while ((size_t(1) << l) < x) ++l;
  return l;
}
}  // namespace

// Depth-first pre-order copy of the trees for the search descent.
int build_search_layout(efg_index* ix) {
  const size_t M = ix->M;
  std::vector<int4> n4(M);
  std::vector<uint2> rg(M);
  std::vector<int32_t> roots(ix->T);
  int rc = 0;
  rc |= check_hip(hipMemcpy(n4.data(), ix->node4.p, M * sizeof(int4), hipMemcpyDeviceToHost), "copy");
  rc |= check_hip(hipMemcpy(rg.data(), ix->range.p, M * sizeof(uint2), hipMemcpyDeviceToHost), "copy");
  rc |= check_hip(hipMemcpy(roots.data(), ix->roots.p, ix->T * 4, hipMemcpyDeviceToHost), "copy");
  if (rc) return EFG_ERR_DEVICE;
  std::vector<int32_t> nid(M, -1), order;
  order.reserve(M);
  std::vector<int32_t> st;
  for (int t = 0; t < ix->T; ++t) {
    st.push_back(roots[t]);
    while (!st.empty()) {
      int32_t x = st.back();
      st.pop_back();
      nid[x] = int32_t(order.size());
      order.push_back(x);
      if (n4[x].z >= 0) {
        st.push_back(n4[x].w);
        st.push_back(n4[x].z);  // left first
      }
    }
  }
  std::vector<SNode> out(order.size());
  for (size_t i = 0; i < order.size(); ++i) {
    const int32_t x = order[i];
    SNode& o = out[i];
    o.dim = n4[x].x;
    std::memcpy(&o.val, &n4[x].y, 4);
    o.left = n4[x].z >= 0 ? nid[n4[x].z] : -1;
    o.right = n4[x].w >= 0 ? nid[n4[x].w] : -1;
    o.parent = -1;
    o.start = rg[x].x;
    o.end = rg[x].y;
    o.pad = 0;
  }
  for (size_t i = 0; i < out.size(); ++i) {
    if (out[i].left >= 0) {
      out[out[i].left].parent = int32_t(i);
      out[out[i].right].parent = int32_t(i);
    }
  }
  if (ix->d >= (size_t(1) << 24)) return fail(EFG_ERR_ARG, "dimension too large for the search trees");
  std::vector<DNode> dn(out.size());
  for (size_t i = 0; i < out.size(); ++i) {
    const SNode& o = out[i];
    dn[i].dim = o.left >= 0 ? o.dim | int32_t(out[o.left].left < 0) << 24 | int32_t(out[o.right].left < 0) << 25 : -1;
    dn[i].val = out[i].val;
    dn[i].right = out[i].right;
    dn[i].parent = out[i].parent;
  }
  std::vector<int32_t> sroots(ix->T);
  for (int t = 0; t < ix->T; ++t) sroots[t] = nid[roots[t]];
  if (!ix->snode.alloc(out.size()) || !ix->dnode.alloc(out.size()) || !ix->sroots.alloc(ix->T))
    return fail(EFG_ERR_MEMORY, "out of device memory (search trees)");
  rc |= check_hip(hipMemcpy(ix->snode.p, out.data(), out.size() * sizeof(SNode), hipMemcpyHostToDevice), "copy");
  rc |= check_hip(hipMemcpy(ix->dnode.p, dn.data(), dn.size() * sizeof(DNode), hipMemcpyHostToDevice), "copy");
  rc |= check_hip(hipMemcpy(ix->sroots.p, sroots.data(), ix->T * 4, hipMemcpyHostToDevice), "copy");
  if (rc) return EFG_ERR_DEVICE;
  ix->snode_valid = true;
  return EFG_OK;
}

// knnSearch on device buffers: method 0 (getNeighbors_nnexp,
// kdtreeub_index.hpp:363-494) or method 1 (getNeighbors_kgraph, 496-654).
namespace {

// Threads per query and the key buffer that depends on it (seeds, or pool +
// new-candidate buffer + merge output).
void set_threads(SearchCfg& c, int t) {
  c.threads = t;
  // method 1 keeps the seeds and the sorted pool only (no merge buffers)
  c.kb = c.method == 1 ? std::max(c.init_num, c.Rp) : std::max(c.init_num, 2 * c.Rp + new_buf(t));
}

// Search configuration from (k, init_num, extend_to, epochs, trees, method)
// with the reference's checks (kdtreeub_index.hpp:363-392, 496-514).
int make_cfg(efg_index* ix, int k, int init, int E, int epochs, int trees, int method, int stop, SearchCfg* out,
             int* init_eff) {
  if (init <= 0 || E < 0 || epochs < 0) return fail(EFG_ERR_ARG, "invalid search parameters");
  const int lsize = init * 2 / (5 * trees) + 1;  // before method 1 raises init_num (507)
  int sorted = 1, R = E > k ? E : k;
  if (method == 1) {
    if (E <= 0) return fail(EFG_ERR_ARG, "method 1 needs extend_to > 0");
    const int pool_size = epochs * E;  // kdtreeub_index.hpp:510-514: init_num raised for good
    if (pool_size >= init) {
      init = pool_size;
      sorted = 0;
    }
    R = pool_size;
    if (k <= 0 || k > 4096) return fail(EFG_ERR_ARG, "invalid k");
  } else if (k <= 0 || R > init) {
    return fail(EFG_ERR_ARG, "k must be positive and at most the pool size");
  }
  if (size_t(init) > ix->n) return fail(EFG_ERR_ARG, "pool size larger than the data set");
  if (init > kKPT * 256) return fail(EFG_ERR_ARG, "pool size above 4096 is not supported");
  SearchCfg c{};
  c.n = int(ix->n);
  c.d = int(ix->d);
  c.D = int(ix->D);
  c.T = trees;
  c.lsize = lsize;
  c.nleaf = c.T * c.lsize;
  c.init_num = init;
  c.E = E;
  c.epochs = epochs;
  c.K = k;
  c.R = R;
  c.Rp = next_pow2(R < 2 ? 2 : R);
  c.method = method;
  c.sorted = sorted;
  set_threads(c, 256);
  c.logcap = ceil_log2(size_t(init + 512) * 4 / 3 + 1);
  if (c.logcap < 10) c.logcap = 10;
  c.seed = ix->bp.seed;
  c.deg = ix->reg_deg;
  c.stop = stop;
  c.out_stride = k;
  *out = c;
  if (init_eff) *init_eff = init;
  return EFG_OK;
}

using SearchKernel = void (*)(SearchCfg, LearnArgs, const float*, size_t, const float*, const uint16_t*,
                              const int32_t*, const SNode*, const uint32_t*, const uint64_t*, const uint32_t*,
                              int32_t*, float*, uint64_t*, unsigned long long*, const int32_t*, int, const int*, int,
                              uint32_t*, int32_t*, int*, unsigned long long*, uint32_t*, int, int*);

// Capacity (log2 entries) of a global-memory visited table: the worst-case
// visited set of a query (seeds plus every expansion), at most 2^26.
int global_logcap(const efg_index* ix, const SearchCfg& c) {
  size_t worst = size_t(c.init_num) + size_t(c.epochs) * size_t(c.E + c.K + 1) *
                                          size_t(ix->max_deg > 0 ? ix->max_deg : 1) + 512;
  if (worst > ix->n + 512) worst = ix->n + 512;
  int l = ceil_log2(worst * 4 / 3 + 1);
  if (l < c.logcap) l = c.logcap;
  if (l > 26) l = 26;
  return l;
}

// One search pass over the queries of qlist (all nq, in leaf order, when
// qlist is null; *nlist_dev of them otherwise): descent, then the search
// kernel with the visited table in LDS, then the queries whose table
// overflowed again with a table in global memory.
__global__ void k_zero_i32(int* p, int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) p[i] = 0;
}

// acc[0] += queries that outgrew the LDS table (re-run in the global pass,
// v[0], or moved to a spill table, v[2]); acc[1] += those that outgrew the
// global table too (v[1]).  Atomic: batch classes on parallel streams share
// the sum.
__global__ void k_add_over(int* __restrict__ acc, const int* __restrict__ v) {
  if (threadIdx.x == 0) {
    atomicAdd(acc, v[0] + v[2]);
    atomicAdd(acc + 1, v[1]);
  }
}

// Counters of one pass (w.nover): [0] queries that outgrew the LDS table,
// [1] those that outgrew the global table too, [2] queries moved to a spill
// table, [4, 12) the query counters of the eight XCD slices, [12] the query
// counter of the global-table pass.  (The pass clears them with a kernel,
// not a memset, so that a captured batch replays correctly.)
constexpr int kNoverInts = 16;
int run_pass_ws(efg_index* ix, efg::Workspace& w, const SearchCfg& c, const LearnArgs& la, bool bf16,
                const float* dq, size_t nq, size_t q_stride, int32_t* d_ids, float* d_dists, uint64_t* d_ndist,
                const int32_t* qlist, const int* nlist_dev, hipEvent_t* ev, hipStream_t st, bool count_work) {
  unsigned long long* work = count_work ? ix->work.p : nullptr;
  const int nb_q = 4096;
  if (!w.leaves.ensure(nq * size_t(c.nleaf)) || !w.overflow.ensure(2 * nq) || !w.nover.ensure(kNoverInts) ||
      !w.nover_sum.ensure(2) || !w.ndist.ensure(3) || !w.qorder.ensure(nq) || !w.qkey.ensure(nq) ||
      !w.qhist.ensure(nb_q + 1) || !w.qoff.ensure(nb_q + 1))
    return fail(EFG_ERR_MEMORY, "out of device memory (search workspace)");
  size_t lds = search_lds_bytes(c, false);
  const size_t lds_cap = 150 * 1024;
  const bool table_in_lds = lds <= lds_cap;
  if (!table_in_lds && search_lds_bytes(c, true) > lds_cap) return fail(EFG_ERR_ARG, "pool too large for one workgroup");
  k_zero_i32<<<1, 64, 0, st>>>(w.nover.p, kNoverInts);
  if (ev) hipEventRecord(ev[0], st);
  // seed slots: bucket probes on a hashing index, tree descent otherwise
  const SNode* snp = ix->hashing ? ix->h_nodes.p : ix->snode.p;
  const uint32_t* lists = ix->hashing ? ix->h_lists.p : ix->leaf.p;
  {
    if (ix->hashing) {
      hash_probe(ix, c.nleaf, nq, w.leaves.p, qlist, nlist_dev, st);
    } else {
      size_t threads = nq * size_t(c.T);
      const unsigned grid = unsigned((threads + kDescThreads - 1) / kDescThreads);
      const size_t qbytes = size_t((kDescThreads - 1) / c.T + 2) * size_t(c.d) * sizeof(float);
      if (qbytes <= kDescLdsQ)
        k_descend<true><<<grid, kDescThreads, qbytes, st>>>(dq, q_stride, int(nq), c.d, c.T, c.lsize, ix->sroots.p,
                                                            ix->dnode.p, w.leaves.p, qlist, nlist_dev, nullptr, 0);
      else
        k_descend<false><<<grid, kDescThreads, 0, st>>>(dq, q_stride, int(nq), c.d, c.T, c.lsize, ix->sroots.p,
                                                        ix->dnode.p, w.leaves.p, qlist, nlist_dev, nullptr, 0);
    }
    if (!qlist) {  // query order by first leaf (bucket)
      int shift = std::max(0, ceil_log2(ix->n + 1) - 12);
      k_zero_i32<<<(nb_q + 1 + 255) / 256, 256, 0, st>>>(reinterpret_cast<int*>(w.qhist.p), nb_q + 1);
      k_qkey<<<unsigned((nq + 255) / 256), 256, 0, st>>>(w.leaves.p, c.nleaf, int(nq), snp, shift, w.qkey.p,
                                                         w.qhist.p);
      if (int rc = device_exclusive_scan(reinterpret_cast<int32_t*>(w.qhist.p), nb_q, w.qoff.p, w.qoff.p + nb_q,
                                         w.scan_sums, st))
        return rc;
      k_zero_i32<<<(nb_q + 1 + 255) / 256, 256, 0, st>>>(reinterpret_cast<int*>(w.qhist.p), nb_q + 1);
      k_qscatter<<<unsigned((nq + 255) / 256), 256, 0, st>>>(w.qkey.p, int(nq), w.qoff.p, w.qhist.p, w.qorder.p);
    }
  }
  if (ev) hipEventRecord(ev[1], st);
  int dev = 0, ncu = 256;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const int32_t* order = qlist ? qlist : w.qorder.p;
  const int tq = c.threads;
  // passes over query lists (batch classes, the learned form's full
  // searches) move queries that outgrow the LDS table into a per-workgroup
  // spill table and continue; the plain pass over all queries keeps its
  // registers (the spill path costs a wave per SIMD) and hands such queries
  // to the global-table pass.  Workspaces of fixed size never spill.
  const bool use_spill = qlist != nullptr && c.method == 0 && c.stop == 0 && !bf16 && w.spill.owned;
  auto pick = [&](bool global) -> SearchKernel {
#define EFG_K(G, M, B, S, T) &k_search_nnexp<G, M, B, S, T>
#define EFG_KT(G, M, B, S) (tq == 64 ? EFG_K(G, M, B, S, 64) : tq == 128 ? EFG_K(G, M, B, S, 128) : EFG_K(G, M, B, S, 256))
    if (c.method == 1) return global ? EFG_K(true, 1, false, 0, 256) : EFG_K(false, 1, false, 0, 256);
    if (c.method == 2) return global ? EFG_K(true, 2, false, 0, 256) : EFG_K(false, 2, false, 0, 256);
    if (c.stop == 1) return global ? nullptr : EFG_KT(false, 0, false, 1);
    if (c.stop == 2) return global ? nullptr : EFG_KT(false, 0, false, 2);
    if (c.stop == 3) return global ? nullptr : EFG_KT(false, 0, false, 3);
    if (bf16) return global ? EFG_K(true, 0, true, 0, 256) : EFG_K(false, 0, true, 0, 256);
    if (!global && use_spill)
      return tq == 64 ? &k_search_nnexp<false, 0, false, 0, 64, true>
                      : tq == 128 ? &k_search_nnexp<false, 0, false, 0, 128, true>
                                  : &k_search_nnexp<false, 0, false, 0, 256, true>;
    return global ? EFG_K(true, 0, false, 0, 256) : EFG_KT(false, 0, false, 0);
#undef EFG_KT
#undef EFG_K
  };
  if (table_in_lds) {
    // workgroups per CU: as many as are resident at once (LDS and
    // registers), at most 32 wavefronts; the workgroups draw queries from a
    // counter, so a larger grid would only add workgroups that find none
    auto kern = pick(false);
    hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    int per_cu = int(lds_cap / lds), resident = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, reinterpret_cast<const void*>(kern), tq, lds) ==
            hipSuccess &&
        resident > 0)
      per_cu = std::min(per_cu, resident);
    per_cu = std::max(1, std::min(per_cu, 32 / (tq / 64)));
    size_t grid = size_t(ncu) * per_cu;
    if (grid > nq) grid = (nq + 7) & ~size_t(7);
    // method 0 (plain): a spill table per workgroup, so that queries
    // outgrowing the LDS table continue in place (not with a fixed-size
    // workspace, which keeps the global-table pass)
    uint32_t* spill = nullptr;
    int spill_logcap = 0;
    if (use_spill) {
      spill_logcap = global_logcap(ix, c);
      if (spill_logcap > c.logcap) {
        if (!w.spill.ensure(grid << spill_logcap)) return fail(EFG_ERR_MEMORY, "out of device memory (spill tables)");
        spill = w.spill.p;
      }
    }
    kern<<<unsigned(grid), unsigned(tq), lds, st>>>(c, la, dq, q_stride, ix->data.p, ix->data_bf16.p, w.leaves.p,
                                                    snp, lists, ix->goff.p, ix->gids.p, d_ids, d_dists,
                                                    d_ndist, w.ndist.p, order, int(nq), nlist_dev, qlist ? 0 : 1,
                                                    nullptr, w.overflow.p, w.nover.p, work, spill, spill_logcap,
                                                    w.nover.p + 4);
  } else if (c.stop) {
    return fail(EFG_ERR_ARG, "probe pool too large for one workgroup");
  }
  if (ev) hipEventRecord(ev[2], st);
  if (c.stop == 0) {
    // queries whose visited set outgrew the LDS table (or all of them when
    // the table never fits) run again with a per-workgroup table in global
    // memory
    SearchCfg c2 = c;
    set_threads(c2, 256);
    c2.logcap = global_logcap(ix, c);
    size_t lds2 = search_lds_bytes(c2, true);
    // as many workgroups as are resident at once (LDS and registers), with
    // at most 1 GiB of tables
    auto kern2 = pick(true);
    hipFuncSetAttribute(reinterpret_cast<const void*>(kern2), hipFuncAttributeMaxDynamicSharedMemorySize, int(lds2));
    int per_cu2 = std::max(1, std::min(8, int(lds_cap / std::max<size_t>(lds2, 1)))), resident2 = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident2, reinterpret_cast<const void*>(kern2), kBlock, lds2) ==
            hipSuccess &&
        resident2 > 0)
      per_cu2 = std::min(per_cu2, resident2);
    const size_t tab_bytes = size_t(4) << c2.logcap;
    unsigned grid2 = unsigned(std::min<size_t>(size_t(ncu) * per_cu2,
                                               std::max<size_t>(128, (size_t(1) << 30) / tab_bytes)));
    if (!w.gtab.owned) grid2 = unsigned(std::min<size_t>(grid2, w.gtab.n >> c2.logcap));  // tables that fit a view
    if (grid2 == 0 || !w.gtab.ensure(size_t(grid2) << c2.logcap))
      return fail(EFG_ERR_MEMORY, "out of device memory (search table)");
    if (table_in_lds) {
      kern2<<<grid2, kBlock, lds2, st>>>(c2, la, dq, q_stride, ix->data.p, ix->data_bf16.p, w.leaves.p, snp,
                                         lists, ix->goff.p, ix->gids.p, d_ids, d_dists, d_ndist, w.ndist.p,
                                         w.overflow.p, 0, w.nover.p, 0, w.gtab.p, w.overflow.p + nq, w.nover.p + 1,
                                         work, nullptr, 0, w.nover.p + 12);
    } else {
      kern2<<<grid2, kBlock, lds2, st>>>(c2, la, dq, q_stride, ix->data.p, ix->data_bf16.p, w.leaves.p, snp,
                                         lists, ix->goff.p, ix->gids.p, d_ids, d_dists, d_ndist, w.ndist.p,
                                         order, int(nq), nlist_dev, 0, w.gtab.p, w.overflow.p + nq, w.nover.p + 1,
                                         work, nullptr, 0, w.nover.p + 12);
    }
  }
  if (ev) hipEventRecord(ev[3], st);
  k_add_over<<<1, 64, 0, st>>>(w.nover_sum.p, w.nover.p);
  return check_hip(hipGetLastError(), "search launch");
}

int run_pass(efg_index* ix, const SearchCfg& c, const LearnArgs& la, bool bf16, const float* dq, size_t nq,
             size_t q_stride, int32_t* d_ids, float* d_dists, uint64_t* d_ndist, const int32_t* qlist,
             const int* nlist_dev, hipEvent_t* ev, hipStream_t st, bool count_work = true) {
  return run_pass_ws(ix, ix->ws, c, la, bf16, dq, nq, q_stride, d_ids, d_dists, d_ndist, qlist, nlist_dev, ev, st,
                     count_work);
}

// Probe search parameters: set explicitly, else pool = max(k, init_num / 6),
// extend_to = max(1, extend_to / 4), the same epochs.
void probe_params(const efg_index* ix, int k, int* pool, int* ext, int* ep) {
  *pool = ix->lp_pool > 0 ? ix->lp_pool : std::max(k, ix->s_init / 6);
  *ext = ix->lp_extend > 0 ? ix->lp_extend : std::max(1, ix->s_extend / 4);
  if (*ext > *pool) *ext = *pool;
  *ep = ix->lp_epochs > 0 ? ix->lp_epochs : ix->s_epochs;
}

// The continuation form's probe must fit inside the full search: its pool
// and extend_to at most the search's, and max(extend_to, k) at most its pool.
int probe_within(const SearchCfg& c, int pp, int pe) {
  if (pp > c.init_num || pe > c.E || std::max(pe, c.K) > pp)
    return fail(EFG_ERR_ARG, "probe pool and extend_to must lie within the search's, max(extend_to, k) within the probe pool");
  return EFG_OK;
}

// Key buffer of the continuation form: the pool's Rp slots, then the rest of
// the full pool's seeds and the probe's pool together.
void cfg_continue(SearchCfg& c, int pp, int pe) {
  const int r1 = std::max(pe, c.K);
  c.kb = std::max(c.kb, c.Rp + (c.init_num - pp) + r1);
}

// Threads per query for a configuration: forced, tuned, or by pool size
// (a wavefront for pools up to 256, two up to 1024, else four); the seed
// keys held in registers need init_num <= 16 * threads.
int choose_threads(const efg_index* ix, const SearchCfg& c) {
  if (c.method != 0) return 256;  // kgraph and the hash-only search run 256 threads per query
  int t = 0;
  if (ix->s_threads) {
    t = ix->s_threads;
  } else {
    auto it = ix->tuned.find({c.init_num, c.E, c.epochs, c.K, c.T, c.stop});
    if (it != ix->tuned.end()) t = it->second;
    else t = c.init_num <= 256 ? 64 : (c.init_num <= 1024 ? 128 : 256);
  }
  while (t < 256 && c.init_num > kKPT * t) t *= 2;
  return t;
}

int search_prepare(efg_index* ix, size_t nq, size_t q_stride, hipStream_t st) {
  if (!ix->s_set) return fail(EFG_ERR_STATE, "search parameters not set (setSearchParams)");
  if (ix->s_method != 0 && ix->s_method != 1) return fail(EFG_ERR_METHOD, "no such searching method");
  if (ix->hashing) {  // hash tables instead of trees; the graph is optional
    if (!ix->h_built) return fail(EFG_ERR_STATE, "no hash tables: build the index first");
    if (nq > ix->h_nq) return fail(EFG_ERR_ARG, "more queries than query codes");
  } else {
    if (ix->T == 0) return fail(EFG_ERR_STATE, "no trees: build or load trees first");
    if (ix->nnz == 0) return fail(EFG_ERR_STATE, "no graph: build or load a graph first");
    if (ix->s_trees <= 0) return fail(EFG_ERR_ARG, "search tree number must be positive");
    if (ix->s_trees > ix->T) return fail(EFG_ERR_TREES, "wrong tree number");
  }
  if (nq > (size_t(1) << 31)) return fail(EFG_ERR_ARG, "too many queries");
  if (q_stride < ix->d) return fail(EFG_ERR_ARG, "query stride below the dimension");
  if (!ix->hashing && !ix->snode_valid) {
    if (int rc = check_hip(hipStreamSynchronize(st), "sync")) return rc;
    if (int rc = build_search_layout(ix)) return rc;
  }
  efg::Workspace& w = ix->ws;
  if (!w.ndist.ensure(3) || !w.nover_sum.ensure(2))
    return fail(EFG_ERR_MEMORY, "out of device memory (search workspace)");
  if (int rc = check_hip(hipMemsetAsync(w.ndist.p, 0, 3 * sizeof(unsigned long long), st), "memset")) return rc;
  k_zero_i32<<<1, 64, 0, st>>>(w.nover_sum.p, 2);
  if (!ix->work.p) {
    if (!ix->work.ensure(3)) return fail(EFG_ERR_MEMORY, "out of device memory (counters)");
    if (int rc = check_hip(hipMemsetAsync(ix->work.p, 0, 3 * sizeof(unsigned long long), st), "memset")) return rc;
  }
  return EFG_OK;
}

}  // namespace

int search_device(efg_index* ix, int k, const float* dq, size_t nq, size_t q_stride, int32_t* d_ids, float* d_dists,
                  uint64_t* d_ndist, hipStream_t st) {
  if (int rc = search_prepare(ix, nq, q_stride, st)) return rc;
  if (nq == 0) return EFG_OK;
  // a hashing index numbers the methods the other way round (0 = kgraph,
  // 1 = nnexp, hashing_index.hpp:902-924) and without a graph runs the
  // hash-only search (method 2 here)
  const int method = !ix->hashing ? ix->s_method : (ix->nnz == 0 ? 2 : 1 - ix->s_method);
  SearchCfg c;
  int init_eff = 0;
  if (int rc = make_cfg(ix, k, ix->s_init, method == 2 ? 0 : ix->s_extend, method == 2 ? 0 : ix->s_epochs,
                        ix->hashing ? ix->h_T : ix->s_trees, method, 0, &c, &init_eff))
    return rc;
  if (ix->hashing) {  // one seed slot per (mask, table)
    c.lsize = ix->h_nmask;
    c.nleaf = c.T * c.lsize;
  }
  if (method == 1) ix->s_init = init_eff;
  const bool bf16 = method == 0 && ix->s_precision == 1;
  set_threads(c, bf16 ? 256 : choose_threads(ix, c));
  if (bf16 && !ix->data_bf16.p) {
    if (!ix->data_bf16.alloc(ix->n * ix->D)) return fail(EFG_ERR_MEMORY, "out of device memory (bf16 rows)");
    const size_t tot = ix->n * ix->D;
    k_pack_bf16<<<unsigned((tot + 255) / 256), 256, 0, st>>>(ix->data.p, int(ix->D), ix->n, ix->data_bf16.p);
  }
  const bool learned = method == 0 && ix->s_mode != 0 && !ix->hashing;
  if (learned && !ix->net_set) return fail(EFG_ERR_STATE, "no stopping network set");
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  const bool prof = ix->profiling && ix->prof_events.size() < 4 * 4096;
  if (prof)
    for (auto& e : ev) hipEventCreate(&e);
  LearnArgs la{};
  if (!learned) {
    if (int rc = run_pass(ix, c, la, bf16, dq, nq, q_stride, d_ids, d_dists, d_ndist, nullptr, nullptr,
                          prof ? ev : nullptr, st))
      return rc;
  } else {
    // probe search stopped by the network; mode 1: the rest runs the full
    // search from scratch; mode 2: the rest continues in place into the
    // full search (one pass), and only queries without room to continue
    // run the full search
    int pp, pe, pep;
    probe_params(ix, k, &pp, &pe, &pep);
    SearchCfg cp;
    if (ix->s_mode == 2) {
      if (int rc = probe_within(c, pp, pe)) return rc;
      cp = c;
      cp.stop = 3;
      set_threads(cp, choose_threads(ix, cp));
      cfg_continue(cp, pp, pe);
      la.p_pool = pp;
      la.p_ext = pe;
      la.p_epochs = pep;
    } else {
      if (int rc = make_cfg(ix, k, pp, pe, pep, ix->s_trees, 0, 1, &cp, nullptr)) return rc;
      set_threads(cp, choose_threads(ix, cp));
    }
    if (!ix->fb_count.ensure(4)) return fail(EFG_ERR_MEMORY, "out of device memory");
    if (!ix->fb_count_init) {
      hipMemsetAsync(ix->fb_count.p, 0, 4 * sizeof(int), st);
      ix->fb_count_init = true;
    }
    if (!ix->fb_list.ensure(nq)) return fail(EFG_ERR_MEMORY, "out of device memory");
    hipMemsetAsync(ix->fb_count.p, 0, sizeof(int), st);
    hipMemsetAsync(ix->fb_count.p + 3, 0, sizeof(int), st);
    la.wpack = ix->net_pack.p;
    la.norm = ix->net_norm.p;
    la.zthr = ix->net_zthr;
    la.fb_list = ix->fb_list.p;
    la.fb_count = ix->fb_count.p;
    if (int rc = run_pass(ix, cp, la, false, dq, nq, q_stride, d_ids, d_dists, d_ndist, nullptr, nullptr,
                          ix->s_mode == 2 && prof ? ev : nullptr, st))
      return rc;
    if (int rc = run_pass(ix, c, LearnArgs{}, false, dq, nq, q_stride, d_ids, d_dists, d_ndist, ix->fb_list.p,
                          ix->fb_count.p, ix->s_mode == 1 && prof ? ev : nullptr, st))
      return rc;
  }
  if (prof)
    for (auto& e : ev) ix->prof_events.push_back(e);
  return EFG_OK;
}

// ---------------------------------------------------------------------------
// Search without device allocation (efg_knn_search_device_noalloc).  Every
// scratch buffer of the pass (leaf slots, query order, overflow lists,
// counters, global visited tables) is cut from one file-scope device array.
// Host code never names the array itself: its address on the current device
// comes from hipGetSymbolAddress, taken once per device and kept.  Calls on
// one GPU share the array, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(EFG_NOALLOC_SCRATCH_BYTES);
__device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

int scratch_address(unsigned char** out) {
  int dev = 0;
  if (int rc = check_hip(hipGetDevice(&dev), "current device")) return rc;
  if (dev < 0 || dev >= kMaxDevices) return fail(EFG_ERR_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    if (int rc = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "scratch address")) return rc;
    g_scratch[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch[dev];
  return EFG_OK;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
}  // namespace

int search_device_noalloc(efg_index* ix, int k, const float* dq, size_t nq, size_t q_stride, int32_t* d_ids,
                          float* d_dists, hipStream_t st) {
  // the checks of search_prepare, without its allocations
  if (!ix->s_set) return fail(EFG_ERR_STATE, "search parameters not set (setSearchParams)");
  if (ix->s_method != 0 && ix->s_method != 1) return fail(EFG_ERR_METHOD, "no such searching method");
  if (ix->hashing) return fail(EFG_ERR_STATE, "the allocation-free search runs on kd-tree indexes");
  if (ix->T == 0) return fail(EFG_ERR_STATE, "no trees: build or load trees first");
  if (ix->nnz == 0) return fail(EFG_ERR_STATE, "no graph: build or load a graph first");
  if (ix->s_trees <= 0) return fail(EFG_ERR_ARG, "search tree number must be positive");
  if (ix->s_trees > ix->T) return fail(EFG_ERR_TREES, "wrong tree number");
  if (ix->s_mode != 0) return fail(EFG_ERR_STATE, "learned stopping allocates: use efg_knn_search_device");
  if (nq > (size_t(1) << 31)) return fail(EFG_ERR_ARG, "too many queries");
  if (q_stride < ix->d) return fail(EFG_ERR_ARG, "query stride below the dimension");
  if (nq && (!dq || !d_ids)) return fail(EFG_ERR_ARG, "null query or result buffer");
  if (!ix->snode_valid)
    return fail(EFG_ERR_STATE, "search trees not laid out yet: run one efg_knn_search(_device) on the index first");
  SearchCfg c;
  int init_eff = 0;
  if (int rc = make_cfg(ix, k, ix->s_init, ix->s_extend, ix->s_epochs, ix->s_trees, ix->s_method, 0, &c, &init_eff))
    return rc;
  if (ix->s_method == 1) ix->s_init = init_eff;  // as the allocating search (kdtreeub_index.hpp:510-514)
  const bool bf16 = ix->s_method == 0 && ix->s_precision == 1;
  if (bf16 && !ix->data_bf16.p)
    return fail(EFG_ERR_STATE, "BF16 rows not made yet: run one search with this precision on the index first");
  set_threads(c, bf16 ? 256 : choose_threads(ix, c));
  if (nq == 0) return EFG_OK;
  // scratch layout (256-byte aligned pieces); the global visited tables take the rest
  const int nb_q = 4096;
  const size_t o_ndist = 0, o_nover = o_ndist + 256, o_nsum = o_nover + 256, o_sums = o_nsum + 256;
  const size_t o_hist = o_sums + align256(64 * 4), o_qoff = o_hist + align256((nb_q + 1) * 4);
  const size_t o_qorder = o_qoff + align256((nb_q + 1) * 4), o_qkey = o_qorder + align256(nq * 4);
  const size_t o_over = o_qkey + align256(nq * 4), o_leaves = o_over + align256(2 * nq * 4);
  const size_t o_gtab = o_leaves + align256(nq * size_t(c.nleaf) * 4);
  const size_t table = size_t(4) << global_logcap(ix, c);
  if (o_gtab > kScratchBytes || kScratchBytes - o_gtab < table) {
    const size_t per_q = 4 * (size_t(c.nleaf) + 4);      // leaf slots, order, key, two overflow entries
    const size_t fixed = o_qorder + 4 * 256 + table;   // counters, histogram, alignment, one table
    const size_t room = kScratchBytes > fixed ? (kScratchBytes - fixed) / per_q : 0;
    return fail(EFG_ERR_ARG, "batch too large for the fixed search scratch of " + std::to_string(kScratchBytes >> 20) +
                                 " MiB: at most " + std::to_string(room) +
                                 " queries per call at these search parameters");
  }
  unsigned char* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  efg::Workspace w;  // views of the scratch: nothing below allocates
  w.ndist.view(reinterpret_cast<unsigned long long*>(base + o_ndist), 3);
  w.nover.view(reinterpret_cast<int*>(base + o_nover), kNoverInts);
  w.nover_sum.view(reinterpret_cast<int*>(base + o_nsum), 2);
  w.scan_sums.view(reinterpret_cast<int32_t*>(base + o_sums), 64);
  w.qhist.view(reinterpret_cast<uint32_t*>(base + o_hist), nb_q + 1);
  w.qoff.view(reinterpret_cast<int32_t*>(base + o_qoff), nb_q + 1);
  w.qorder.view(reinterpret_cast<int32_t*>(base + o_qorder), nq);
  w.qkey.view(reinterpret_cast<int32_t*>(base + o_qkey), nq);
  w.overflow.view(reinterpret_cast<int32_t*>(base + o_over), 2 * nq);
  w.leaves.view(reinterpret_cast<int32_t*>(base + o_leaves), nq * size_t(c.nleaf));
  w.gtab.view(reinterpret_cast<uint32_t*>(base + o_gtab), (kScratchBytes - o_gtab) / 4);
  // counters start at zero (a fill kernel, not a memset, so that callers may capture the call)
  k_zero_i32<<<1, 64, 0, st>>>(reinterpret_cast<int*>(base + o_ndist), 6);
  k_zero_i32<<<1, 64, 0, st>>>(w.nover_sum.p, 2);
  return run_pass_ws(ix, w, c, LearnArgs{}, bf16, dq, nq, q_stride, d_ids, d_dists, nullptr, nullptr, nullptr,
                     nullptr, st, false);
}

int search_autotune(efg_index* ix, int k, const float* dq, size_t nq, size_t q_stride, int* chosen,
                    hipStream_t st) {
  if (ix->hashing) return fail(EFG_ERR_STATE, "the tuner runs on kd-tree indexes");
  if (int rc = search_prepare(ix, nq, q_stride, st)) return rc;
  if (nq == 0) return fail(EFG_ERR_ARG, "no queries to tune on");
  if (ix->s_method != 0) {
    if (chosen) chosen[0] = chosen[1] = 256;
    return EFG_OK;
  }
  const bool learned = ix->s_mode != 0;
  if (learned && !ix->net_set) return fail(EFG_ERR_STATE, "no stopping network set");
  // the configurations this mode runs: the full search, and the probe (mode
  // 1) or the probe with its continuation (mode 2) when learned
  SearchCfg cfg[2];
  int ncfg = 0;
  if (int rc = make_cfg(ix, k, ix->s_init, ix->s_extend, ix->s_epochs, ix->s_trees, 0, 0, &cfg[ncfg++], nullptr))
    return rc;
  int pp = 0, pe = 0, pep = 0;
  if (learned) {
    probe_params(ix, k, &pp, &pe, &pep);
    if (ix->s_mode == 2) {
      if (int rc = probe_within(cfg[0], pp, pe)) return rc;
      cfg[ncfg] = cfg[0];
      cfg[ncfg++].stop = 3;
    } else if (int rc = make_cfg(ix, k, pp, pe, pep, ix->s_trees, 0, 1, &cfg[ncfg++], nullptr)) {
      return rc;
    }
    if (!ix->fb_list.ensure(nq)) return fail(EFG_ERR_MEMORY, "out of device memory");
  }
  DevBuf<int32_t> ids;
  DevBuf<float> dd;
  DevBuf<int> tcount;  // fallback counters of the tuning runs (zeroed before each)
  if (!ids.alloc(nq * size_t(k)) || !dd.alloc(nq * size_t(k)) || !tcount.alloc(4))
    return fail(EFG_ERR_MEMORY, "out of device memory");
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  int rc = EFG_OK;
  for (int ci = 0; ci < ncfg && rc == EFG_OK; ++ci) {
    SearchCfg c = cfg[ci];
    LearnArgs la{};
    if (c.stop != 0) {
      la.wpack = ix->net_pack.p;
      la.norm = ix->net_norm.p;
      la.zthr = ix->net_zthr;
      la.fb_list = ix->fb_list.p;
      la.fb_count = tcount.p;
      la.p_pool = pp;
      la.p_ext = pe;
      la.p_epochs = pep;
    }
    float best = 0.f;
    int best_t = 0;
    for (int t : {64, 128, 256}) {
      if (c.init_num > kKPT * t) continue;
      set_threads(c, t);
      if (c.stop == 3) cfg_continue(c, pp, pe);
      float ms_sum = 0.f;
      for (int rep = 0; rep < 3 && rc == EFG_OK; ++rep) {
        hipMemsetAsync(tcount.p, 0, 4 * sizeof(int), st);
        hipEventRecord(e0, st);
        rc = run_pass(ix, c, la, false, dq, nq, q_stride, ids.p, dd.p, nullptr, nullptr, nullptr, nullptr, st, false);
        hipEventRecord(e1, st);
        if (rc == EFG_OK) rc = check_hip(hipEventSynchronize(e1), "autotune");
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0) ms_sum += ms;  // the first run warms up
      }
      if (rc == EFG_OK && (best_t == 0 || ms_sum < best)) {
        best = ms_sum;
        best_t = t;
      }
    }
    if (rc == EFG_OK) {
      ix->tuned[{c.init_num, c.E, c.epochs, c.K, c.T, c.stop}] = best_t;
      if (chosen) chosen[ci] = best_t;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (ncfg == 1 && chosen) chosen[1] = 0;
  return rc;
}

namespace {
// Query classes of a heterogeneous batch: (pool, extend_to, k) -> queries.
struct BatchClass {
  int pool, ext, k;
  std::vector<int32_t> q;
};
int batch_classes(const efg_index* ix, size_t nq, const int* pool, const int* ext, const int* kq, int out_stride,
                  std::vector<BatchClass>* out) {
  std::map<std::array<int, 3>, size_t> where;
  for (size_t i = 0; i < nq; ++i) {
    const int p = pool[i], k = kq[i];
    const int e = ext ? ext[i] : std::max(1, p / 6);
    if (p <= 0 || k <= 0 || e < 0) return fail(EFG_ERR_ARG, "invalid per-query parameters");
    if (k > out_stride) return fail(EFG_ERR_ARG, "k above the output stride");
    auto key = std::array<int, 3>{p, std::min(e, p), k};
    auto it = where.find(key);
    if (it == where.end()) {
      it = where.emplace(key, out->size()).first;
      out->push_back(BatchClass{key[0], key[1], key[2], {}});
    }
    (*out)[it->second].q.push_back(int32_t(i));
  }
  (void)ix;
  return EFG_OK;
}
}  // namespace

__global__ void k_fill_results(int32_t* __restrict__ ids, float* __restrict__ d, size_t m) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  ids[i] = -1;
  if (d) d[i] = 3.402823466e38f;
}

// One persistent pass over a group of a batch's classes: the largest sizes
// over the group (its LDS layout), its spill-table size and its queries'
// range in the draw order.
struct FusedGroup {
  SearchCfg c0{};
  int spill_logcap = 0;
  int off = 0, count = 0;
  float share = 1.f;  // the part of the resident workgroups its grid takes
};

// A heterogeneous batch after its classes are formed, their query lists are
// on the device and their threads per query are chosen.  One-pass form
// (fused): the classes' configurations, each query's class and leaves per
// tree, and the draw order are on the device; the classes run in one or two
// groups, each one persistent pass.
struct BatchSetup {
  std::vector<SearchCfg> cfg;
  std::vector<int> offs, counts;
  size_t nq = 0;
  int out_stride = 0;
  bool fused = false;
  std::vector<FusedGroup> groups;
  int lstride = 0;  // leaf slots per query (the classes' largest)
};

// Shape of the one-pass form: the classes with pools up to `split` form one
// group at `t_small` threads per query, the others another at `t_large`
// (split 0: all in one group at t_large).  Two groups run at once, the small
// one's grid taking `share` of its resident workgroups and the large one's
// the rest of theirs, so that both persistent kernels are resident together.
struct FusedShape {
  int split = 0, t_small = 128, t_large = 256;
  float share = 0.5f;
};

SearchKernel fused_kernel(int tq) {
  return tq == 64 ? &k_search_nnexp<false, 0, false, 0, 64, true, true>
                  : tq == 128 ? &k_search_nnexp<false, 0, false, 0, 128, true, true>
                              : &k_search_nnexp<false, 0, false, 0, 256, true, true>;
}

// One group's launches (no host synchronisation, no memset, so a plan can
// capture them): a descent of the group's queries with each query's leaves
// per tree, then one persistent search kernel whose workgroups draw the
// group's queries from the order in w.border, each query with its class's
// configuration; a query about to outgrow the LDS visited table continues in
// the workgroup's spill table.  wg holds the group's counters and spill
// tables, w the batch's shared arrays.
int fused_launch_group(efg_index* ix, efg::Workspace& w, efg::Workspace& wg, const BatchSetup& b, size_t g,
                       const float* dq, size_t q_stride, int32_t* d_ids, float* d_dists, hipStream_t st) {
  const FusedGroup& fg = b.groups[g];
  const SearchCfg& c0 = fg.c0;
  if (fg.count == 0) return EFG_OK;
  if (!wg.nover.ensure(kNoverInts) || !wg.ndist.ensure(3) || !wg.overflow.ensure(1))
    return fail(EFG_ERR_MEMORY, "out of device memory (batch)");
  k_zero_i32<<<1, 64, 0, st>>>(wg.nover.p, kNoverInts);
  const int32_t* order = w.border.p + fg.off;
  const size_t threads = size_t(fg.count) * size_t(c0.T);
  const unsigned dgrid = unsigned((threads + kDescThreads - 1) / kDescThreads);
  const size_t qbytes = size_t((kDescThreads - 1) / c0.T + 2) * size_t(c0.d) * sizeof(float);
  if (qbytes <= kDescLdsQ)
    k_descend<true><<<dgrid, kDescThreads, qbytes, st>>>(dq, q_stride, int(b.nq), c0.d, c0.T, c0.lsize, ix->sroots.p,
                                                         ix->dnode.p, w.leaves.p, order, w.bcount.p + g, w.blsize.p,
                                                         size_t(b.lstride));
  else
    k_descend<false><<<dgrid, kDescThreads, 0, st>>>(dq, q_stride, int(b.nq), c0.d, c0.T, c0.lsize, ix->sroots.p,
                                                     ix->dnode.p, w.leaves.p, order, w.bcount.p + g, w.blsize.p,
                                                     size_t(b.lstride));
  const int tq = c0.threads;
  SearchKernel kern = fused_kernel(tq);
  const size_t lds = search_lds_bytes(c0, false);
  hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  int dev = 0, ncu = 256;
  hipGetDevice(&dev);
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  int per_cu = int(150 * 1024 / lds), resident = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, reinterpret_cast<const void*>(kern), tq, lds) ==
          hipSuccess &&
      resident > 0)
    per_cu = std::min(per_cu, resident);
  per_cu = std::max(1, std::min(per_cu, 32 / (tq / 64)));
  size_t grid = std::max<size_t>(8, size_t(double(size_t(ncu) * per_cu) * fg.share) & ~size_t(7));
  if (grid > size_t(fg.count)) grid = (size_t(fg.count) + 7) & ~size_t(7);
  if (!wg.spill.ensure(grid << fg.spill_logcap)) return fail(EFG_ERR_MEMORY, "out of device memory (spill tables)");
  LearnArgs la{};
  la.ccfg = w.bcfg.p;
  la.qcls = w.bcls.p;
  kern<<<unsigned(grid), unsigned(tq), lds, st>>>(c0, la, dq, q_stride, ix->data.p, ix->data_bf16.p, w.leaves.p,
                                                  ix->snode.p, ix->leaf.p, ix->goff.p, ix->gids.p, d_ids, d_dists,
                                                  nullptr, wg.ndist.p, order, fg.count, nullptr, 0, nullptr,
                                                  wg.overflow.p, wg.nover.p, ix->work.p, wg.spill.p,
                                                  fg.spill_logcap, wg.nover.p + 4);
  k_add_over<<<1, 64, 0, st>>>(w.nover_sum.p, wg.nover.p);
  return check_hip(hipGetLastError(), "batch launch");
}

// The one-pass form's configurations and device arrays for the classes
// `cls` in shape `sh`: each class at its group's threads per query; per
// group the largest sizes, the spill-table size and the draw order (the
// eight XCD slices of a group each take every eighth query of every class,
// classes by pool size, largest first, so that every slice ends on small
// queries).  Fails (EFG_ERR_ARG) when a class does not fit its group's
// threads or a group's largest configuration one workgroup's LDS.
int fused_setup(efg_index* ix, efg::Workspace& w, const std::vector<BatchClass>& cls, const FusedShape& sh,
                BatchSetup* b, hipStream_t st) {
  std::vector<SearchCfg> cc = b->cfg;
  std::vector<int> grp(cc.size());
  b->lstride = 0;
  for (size_t i = 0; i < cc.size(); ++i) {
    grp[i] = (sh.split > 0 && cc[i].init_num <= sh.split) ? 0 : 1;
    const int t = grp[i] == 0 ? sh.t_small : sh.t_large;
    if (cc[i].init_num > kKPT * t) return fail(EFG_ERR_ARG, "pool too large for the threads per query");
    set_threads(cc[i], t);
    b->lstride = std::max(b->lstride, cc[i].nleaf);
  }
  b->groups.clear();
  std::vector<int32_t> order;
  order.reserve(b->nq);
  std::vector<int> counts;
  for (int g = 0; g < 2; ++g) {
    std::vector<size_t> members;
    for (size_t i = 0; i < cc.size(); ++i)
      if (grp[i] == g) members.push_back(i);
    if (members.empty()) continue;
    std::stable_sort(members.begin(), members.end(),
                     [&](size_t a, size_t c) { return cc[a].init_num > cc[c].init_num; });
    FusedGroup fg;
    fg.c0 = cc[members[0]];
    for (size_t i : members) {
      const SearchCfg& c = cc[i];
      fg.c0.kb = std::max(fg.c0.kb, c.kb);
      fg.c0.logcap = std::max(fg.c0.logcap, c.logcap);
      fg.c0.lsize = std::max(fg.c0.lsize, c.lsize);
      fg.c0.E = std::max(fg.c0.E, c.E);
      fg.c0.K = std::max(fg.c0.K, c.K);
      fg.c0.R = std::max(fg.c0.R, c.R);
      fg.c0.Rp = std::max(fg.c0.Rp, c.Rp);
      fg.c0.init_num = std::max(fg.c0.init_num, c.init_num);
      fg.spill_logcap = std::max(fg.spill_logcap, global_logcap(ix, c));
    }
    fg.c0.nleaf = b->lstride;  // the leaf-slot row of every query
    if (search_lds_bytes(fg.c0, false) > 150 * 1024)
      return fail(EFG_ERR_ARG, "batch classes too large for one workgroup");
    fg.off = int(order.size());
    for (int x = 0; x < 8; ++x)
      for (size_t r : members)
        for (size_t j = size_t(x); j < cls[r].q.size(); j += 8) order.push_back(cls[r].q[j]);
    fg.count = int(order.size()) - fg.off;
    counts.push_back(fg.count);
    b->groups.push_back(fg);
  }
  if (b->groups.size() == 2) {
    b->groups[0].share = sh.share;
    b->groups[1].share = 1.f - sh.share;
  }
  const size_t nq = b->nq;
  std::vector<uint8_t> qc(nq);
  std::vector<int> ql(nq);
  for (size_t i = 0; i < cls.size(); ++i)
    for (int32_t qi : cls[i].q) {
      qc[size_t(qi)] = uint8_t(i);
      ql[size_t(qi)] = cc[i].lsize;
    }
  if (!w.bcfg.ensure(cc.size()) || !w.bcls.ensure(nq) || !w.blsize.ensure(nq) || !w.border.ensure(nq) ||
      !w.bcount.ensure(std::max<size_t>(counts.size(), b->counts.size())) ||
      !w.leaves.ensure(nq * size_t(b->lstride)) || !w.nover_sum.ensure(2))
    return fail(EFG_ERR_MEMORY, "out of device memory (batch)");
  // pageable sources: the calls return once the host data is staged
  if (int rc = check_hip(hipMemcpyAsync(w.bcfg.p, cc.data(), cc.size() * sizeof(SearchCfg), hipMemcpyHostToDevice,
                                        st), "copy classes"))
    return rc;
  if (int rc = check_hip(hipMemcpyAsync(w.bcls.p, qc.data(), nq, hipMemcpyHostToDevice, st), "copy classes"))
    return rc;
  if (int rc = check_hip(hipMemcpyAsync(w.blsize.p, ql.data(), nq * 4, hipMemcpyHostToDevice, st), "copy leaves"))
    return rc;
  if (int rc = check_hip(hipMemcpyAsync(w.border.p, order.data(), nq * 4, hipMemcpyHostToDevice, st), "copy order"))
    return rc;
  if (int rc = check_hip(hipMemcpyAsync(w.bcount.p, counts.data(), counts.size() * 4, hipMemcpyHostToDevice, st),
                         "copy counts"))
    return rc;
  b->fused = true;
  return EFG_OK;
}

int batch_streams(efg_index* ix);

// The one-pass launches of a batch: one group on the caller's stream, two
// forked onto two of the parallel streams (each with a workspace of its own
// for counters and spill tables) and joined back.
int fused_launch(efg_index* ix, efg::Workspace& w, std::vector<std::unique_ptr<efg::Workspace>>& cws,
                 const BatchSetup& b, const float* dq, size_t q_stride, int32_t* d_ids, float* d_dists,
                 hipStream_t st) {
  if (b.groups.size() == 1) return fused_launch_group(ix, w, w, b, 0, dq, q_stride, d_ids, d_dists, st);
  if (int rc = batch_streams(ix)) return rc;
  while (cws.size() < b.groups.size()) cws.emplace_back(new efg::Workspace());
  if (int rc = check_hip(hipEventRecord(ix->bfork, st), "batch fork")) return rc;
  for (size_t g = 0; g < b.groups.size(); ++g) {
    if (int rc = check_hip(hipStreamWaitEvent(ix->bstreams[g], ix->bfork, 0), "batch fork")) return rc;
    if (int rc = fused_launch_group(ix, w, *cws[g], b, g, dq, q_stride, d_ids, d_dists, ix->bstreams[g])) return rc;
    if (int rc = check_hip(hipEventRecord(ix->bjoin[g], ix->bstreams[g]), "batch join")) return rc;
    if (int rc = check_hip(hipStreamWaitEvent(st, ix->bjoin[g], 0), "batch join")) return rc;
  }
  return EFG_OK;
}

int batch_setup(efg_index* ix, efg::Workspace& w, const float* dq, size_t nq, size_t q_stride, const int* pool,
                const int* ext, const int* kq, int out_stride, int32_t* d_ids, float* d_dists, bool tune,
                hipStream_t st, BatchSetup* b) {
  if (ix->hashing) return fail(EFG_ERR_STATE, "batched search runs on kd-tree indexes");
  if (ix->s_method != 0) return fail(EFG_ERR_METHOD, "batched search runs method 0");
  if (out_stride <= 0) return fail(EFG_ERR_ARG, "invalid output stride");
  std::vector<BatchClass> cls;
  if (int rc = batch_classes(ix, nq, pool, ext, kq, out_stride, &cls)) return rc;
  // class configurations and their query lists (concatenated on the device)
  b->cfg.assign(cls.size(), SearchCfg{});
  b->offs.clear();
  b->nq = nq;
  b->out_stride = out_stride;
  std::vector<int32_t> lists;
  std::vector<int> counts;
  int max_nleaf = 0;
  for (size_t i = 0; i < cls.size(); ++i) {
    if (int rc = make_cfg(ix, cls[i].k, cls[i].pool, cls[i].ext, ix->s_epochs, ix->s_trees, 0, 0, &b->cfg[i], nullptr))
      return rc;
    b->cfg[i].out_stride = out_stride;
    max_nleaf = std::max(max_nleaf, b->cfg[i].nleaf);
    b->offs.push_back(int(lists.size()));
    counts.push_back(int(cls[i].q.size()));
    b->counts.push_back(int(cls[i].q.size()));
    lists.insert(lists.end(), cls[i].q.begin(), cls[i].q.end());
  }
  if (!w.leaves.ensure(nq * size_t(max_nleaf)) || !w.blist.ensure(lists.size()) || !w.bcount.ensure(counts.size()))
    return fail(EFG_ERR_MEMORY, "out of device memory (batch)");
  // pageable sources: the calls return once the host data is staged
  if (int rc = check_hip(hipMemcpyAsync(w.blist.p, lists.data(), lists.size() * 4, hipMemcpyHostToDevice, st),
                         "copy query lists"))
    return rc;
  if (int rc = check_hip(hipMemcpyAsync(w.bcount.p, counts.data(), counts.size() * 4, hipMemcpyHostToDevice, st),
                         "copy query counts"))
    return rc;
  int rc = EFG_OK;
  if (ix->b_mode == 1 && cls.size() <= 255) {
    // one pass over all classes, in one group or two split by pool size:
    // the shape (split, threads per query of each group) tuned on this
    // batch (tune) or kept from the last tuning
    FusedShape sh{ix->b_split, ix->b_tsmall, ix->b_tlarge, ix->b_share};
    if (tune) {
      std::vector<FusedShape> shapes;
      std::vector<int> pools;
      for (const auto& c : b->cfg) pools.push_back(c.init_num);
      std::sort(pools.begin(), pools.end());
      pools.erase(std::unique(pools.begin(), pools.end()), pools.end());
      for (int tl : {128, 256}) shapes.push_back(FusedShape{0, 128, tl, 1.f});
      for (size_t i = 0; i + 1 < pools.size(); ++i)
        for (int ts : {64, 128})
          for (float sh2 : {0.35f, 0.5f, 0.65f}) shapes.push_back(FusedShape{pools[i], ts, 256, sh2});
      hipEvent_t e0 = nullptr, e1 = nullptr;
      hipEventCreate(&e0);
      hipEventCreate(&e1);
      float best = 0.f;
      FusedShape best_sh{};
      bool found = false;
      for (const FusedShape& cand : shapes) {
        if (rc != EFG_OK) break;
        BatchSetup tb = *b;
        if (fused_setup(ix, w, cls, cand, &tb, st) != EFG_OK) continue;  // does not fit
        float sum = 0.f;
        for (int rep = 0; rep < 3 && rc == EFG_OK; ++rep) {
          hipEventRecord(e0, st);
          rc = fused_launch(ix, w, ix->bws, tb, dq, q_stride, d_ids, d_dists, st);
          hipEventRecord(e1, st);
          if (rc == EFG_OK) rc = check_hip(hipEventSynchronize(e1), "batch tuning");
          float ms = 0.f;
          hipEventElapsedTime(&ms, e0, e1);
          if (rep > 0) sum += ms;
        }
        if (rc == EFG_OK && (!found || sum < best)) {
          best = sum;
          best_sh = cand;
          found = true;
        }
      }
      hipEventDestroy(e0);
      hipEventDestroy(e1);
      if (rc == EFG_OK && found) {
        sh = best_sh;
        ix->b_split = sh.split;
        ix->b_tsmall = sh.t_small;
        ix->b_tlarge = sh.t_large;
        ix->b_share = sh.share;
      }
    }
    if (rc == EFG_OK) {
      rc = fused_setup(ix, w, cls, sh, b, st);
      if (rc != EFG_OK) rc = fused_setup(ix, w, cls, FusedShape{0, 128, 256, 1.f}, b, st);
      if (rc != EFG_OK) {  // too large for one pass: one pass per class
        b->fused = false;
        rc = EFG_OK;
      }
    }
    if (b->fused) return rc;
  }
  if (tune) {  // time 64/128/256 threads per query on each class, keep the fastest
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    for (size_t i = 0; i < b->cfg.size() && rc == EFG_OK; ++i) {
      SearchCfg c = b->cfg[i];
      float best = 0.f;
      int best_t = 0;
      for (int t : {64, 128, 256}) {
        if (c.init_num > kKPT * t) continue;
        set_threads(c, t);
        float sum = 0.f;
        for (int rep = 0; rep < 3 && rc == EFG_OK; ++rep) {
          hipEventRecord(e0, st);
          rc = run_pass_ws(ix, w, c, LearnArgs{}, false, dq, nq, q_stride, d_ids, d_dists, nullptr,
                           w.blist.p + b->offs[i], w.bcount.p + i, nullptr, st, false);
          hipEventRecord(e1, st);
          if (rc == EFG_OK) rc = check_hip(hipEventSynchronize(e1), "batch tuning");
          float ms = 0.f;
          hipEventElapsedTime(&ms, e0, e1);
          if (rep > 0) sum += ms;
        }
        if (rc == EFG_OK && (best_t == 0 || sum < best)) {
          best = sum;
          best_t = t;
        }
      }
      if (rc == EFG_OK) ix->tuned[{c.init_num, c.E, c.epochs, c.K, c.T, 0}] = best_t;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  }
  for (auto& c : b->cfg) set_threads(c, choose_threads(ix, c));
  return rc;
}

// Parallel streams of the batched search (default priority), created once
// per index, and the events that fork them from and join them to the
// caller's stream.
constexpr int kBatchStreams = 4;
int batch_streams(efg_index* ix) {
  if (!ix->bstreams.empty()) return EFG_OK;
  for (int i = 0; i < kBatchStreams; ++i) {
    hipStream_t s = nullptr;
    if (int rc = check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "batch stream")) return rc;
    ix->bstreams.push_back(s);
    hipEvent_t e = nullptr;
    if (int rc = check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "batch event")) return rc;
    ix->bjoin.push_back(e);
  }
  return check_hip(hipEventCreateWithFlags(&ix->bfork, hipEventDisableTiming), "batch event");
}

// The batch's launches only (no host synchronisation, no memset), so the
// sequence can be captured into a graph.  Each (pool, extend_to, k) class
// runs its passes with a workspace of its own (cws[i]) on one of the
// parallel streams, so the classes' passes overlap instead of each draining
// the GPU before the next begins; in a captured graph they become parallel
// branches between the fork and the join.
int batch_launch(efg_index* ix, efg::Workspace& w, std::vector<std::unique_ptr<efg::Workspace>>& cws,
                 const BatchSetup& b, const float* dq, size_t q_stride, int32_t* d_ids, float* d_dists,
                 hipStream_t st) {
  const size_t m = b.nq * size_t(b.out_stride);
  k_fill_results<<<unsigned((m + 255) / 256), 256, 0, st>>>(d_ids, d_dists, m);
  if (!w.nover_sum.ensure(2)) return fail(EFG_ERR_MEMORY, "out of device memory (batch)");
  k_zero_i32<<<1, 64, 0, st>>>(w.nover_sum.p, 2);
  if (b.fused) return fused_launch(ix, w, cws, b, dq, q_stride, d_ids, d_dists, st);
  const size_t nc = b.cfg.size();
  if (nc == 1) {
    return run_pass_ws(ix, w, b.cfg[0], LearnArgs{}, false, dq, b.nq, q_stride, d_ids, d_dists, nullptr, w.blist.p,
                       w.bcount.p, nullptr, st, true);
  }
  if (int rc = batch_streams(ix)) return rc;
  while (cws.size() < nc) cws.emplace_back(new efg::Workspace());
  if (int rc = check_hip(hipEventRecord(ix->bfork, st), "batch fork")) return rc;
  const int ns = int(std::min<size_t>(nc, ix->bstreams.size()));
  for (int j = 0; j < ns; ++j)
    if (int rc = check_hip(hipStreamWaitEvent(ix->bstreams[j], ix->bfork, 0), "batch fork")) return rc;
  // largest classes first, dealt round-robin over the streams
  std::vector<size_t> order(nc);
  for (size_t i = 0; i < nc; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) {
    return size_t(b.cfg[a].init_num) * size_t(b.counts[a]) > size_t(b.cfg[c].init_num) * size_t(b.counts[c]);
  });
  for (size_t r = 0; r < nc; ++r) {
    const size_t i = order[r];
    efg::Workspace& wi = *cws[i];
    wi.nover_sum.view(w.nover_sum.p, 2);  // the classes add into the batch's sum
    if (int rc = run_pass_ws(ix, wi, b.cfg[i], LearnArgs{}, false, dq, b.nq, q_stride, d_ids, d_dists, nullptr,
                             w.blist.p + b.offs[i], w.bcount.p + i, nullptr, ix->bstreams[r % ns], true))
      return rc;
  }
  for (int j = 0; j < ns; ++j) {
    if (int rc = check_hip(hipEventRecord(ix->bjoin[j], ix->bstreams[j]), "batch join")) return rc;
    if (int rc = check_hip(hipStreamWaitEvent(st, ix->bjoin[j], 0), "batch join")) return rc;
  }
  return EFG_OK;
}

int search_batch(efg_index* ix, const float* dq, size_t nq, size_t q_stride, const int* pool, const int* ext,
                 const int* kq, int out_stride, int32_t* d_ids, float* d_dists, bool tune, hipStream_t st) {
  if (int rc = search_prepare(ix, nq, q_stride, st)) return rc;
  if (nq == 0) return EFG_OK;
  BatchSetup b;
  if (int rc = batch_setup(ix, ix->ws, dq, nq, q_stride, pool, ext, kq, out_stride, d_ids, d_dists, tune, st, &b))
    return rc;
  return batch_launch(ix, ix->ws, ix->bws, b, dq, q_stride, d_ids, d_dists, st);
}

}  // namespace efg

// A batch plan: the launch sequence of one heterogeneous batch captured into
// a HIP graph, with a workspace of its own so that other searches on the
// index cannot move the buffers the graph refers to.
struct efg_batch_plan {
  efg_index* ix = nullptr;
  efg::Workspace ws;
  std::vector<std::unique_ptr<efg::Workspace>> cws;  // one per class
  efg::BatchSetup setup;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace efg {

int batch_plan_create(efg_index* ix, const float* dq, size_t nq, size_t q_stride, const int* pool, const int* ext,
                      const int* kq, int out_stride, int32_t* d_ids, float* d_dists, bool tune, hipStream_t st,
                      efg_batch_plan** out) {
  *out = nullptr;
  if (nq == 0) return fail(EFG_ERR_ARG, "empty batch");
  if (int rc = search_prepare(ix, nq, q_stride, st)) return rc;
  auto* pl = new efg_batch_plan();
  pl->ix = ix;
  int rc = batch_setup(ix, pl->ws, dq, nq, q_stride, pool, ext, kq, out_stride, d_ids, d_dists, tune, st, &pl->setup);
  // one eager run sizes the plan's workspace before the capture
  if (rc == EFG_OK) rc = batch_launch(ix, pl->ws, pl->cws, pl->setup, dq, q_stride, d_ids, d_dists, st);
  if (rc == EFG_OK) rc = check_hip(hipStreamSynchronize(st), "batch plan");
  hipStream_t cs = nullptr;
  if (rc == EFG_OK) rc = check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "capture stream");
  if (rc == EFG_OK) {
    rc = check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc == EFG_OK) {
      const int lrc = batch_launch(ix, pl->ws, pl->cws, pl->setup, dq, q_stride, d_ids, d_dists, cs);
      const int erc = check_hip(hipStreamEndCapture(cs, &pl->graph), "end capture");
      rc = lrc ? lrc : erc;
    }
  }
  if (rc == EFG_OK) rc = check_hip(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0), "instantiate");
  if (cs) hipStreamDestroy(cs);
  if (rc != EFG_OK) {
    batch_plan_destroy(pl);
    return rc;
  }
  *out = pl;
  return EFG_OK;
}

int batch_plan_run(efg_batch_plan* pl, hipStream_t st) {
  return check_hip(hipGraphLaunch(pl->exec, st), "batch plan launch");
}

void batch_plan_destroy(efg_batch_plan* pl) {
  if (!pl) return;
  if (pl->exec) hipGraphExecDestroy(pl->exec);
  if (pl->graph) hipGraphDestroy(pl->graph);
  delete pl;
}

int search_traces(efg_index* ix, int k, const float* dq, size_t nq, size_t q_stride, const int32_t* d_plain,
                  float* d_feat, uint8_t* d_label, hipStream_t st) {
  if (ix->hashing) return fail(EFG_ERR_STATE, "learned stopping runs on kd-tree indexes");
  if (int rc = search_prepare(ix, nq, q_stride, st)) return rc;
  if (nq == 0) return EFG_OK;
  int pp, pe, pep;
  probe_params(ix, k, &pp, &pe, &pep);
  SearchCfg cp;
  if (int rc = make_cfg(ix, k, pp, pe, pep, ix->s_trees, 0, 2, &cp, nullptr)) return rc;
  set_threads(cp, choose_threads(ix, cp));
  DevBuf<int32_t> ids;
  if (!ids.alloc(nq * size_t(k))) return fail(EFG_ERR_MEMORY, "out of device memory");
  hipMemsetAsync(d_label, 0xff, nq * size_t(pep + 1), st);
  LearnArgs la{};
  la.tr_feat = d_feat;
  la.tr_label = d_label;
  la.plain_ids = d_plain;
  if (int rc = run_pass(ix, cp, la, false, dq, nq, q_stride, ids.p, nullptr, nullptr, nullptr, nullptr, nullptr, st))
    return rc;
  return check_hip(hipStreamSynchronize(st), "traces");
}

}  // namespace efg

// Exact k nearest neighbours by brute force: the k smallest (distance, id)
// over all base rows (Matrix::brute_force_search, general/matrix.hpp:57-68;
// ties to the lower id), optionally excluding the query's own row.
//
// k_exact_tile: one workgroup per (block of QT queries, chunk of base rows).
// Query and base tiles are staged in LDS transposed ([element][row]); each
// thread owns an MQ x MP micro-tile and runs the reference's eight AVX lane
// chains one after another (chain j covers elements j mod 8, tail chunks
// first), adding each finished chain to the running total — the same
// operations and order as L2DistanceAVX::compare, so distances are
// bit-identical.  Candidates under the per-query k-th key go through a small
// LDS buffer into a sorted per-query top-k list.
// k_exact_merge: one workgroup per query merges the per-chunk lists.
#include <algorithm>

#include "efg_device.hpp"
#include "efg_internal.hpp"
#include "efg_mfma.hpp"

namespace efg {

template <int MQ, int MP>
__global__ void __launch_bounds__(kBlock) k_exact_tile(const float* __restrict__ X, size_t n, int d, int D,
                                                       size_t x_stride, const float* __restrict__ Q, int nq,
                                                       size_t q_stride, int k, int exclude_self, int chunk,
                                                       uint64_t* __restrict__ partial, int nchunks) {
  constexpr int QT = 16 * MQ, PT = 16 * MP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float* qs = reinterpret_cast<float*>(smem_raw);          // [D][QT]
  float* ps = qs + size_t(D) * QT;                          // [D][PT]
  uint64_t* topk = reinterpret_cast<uint64_t*>(ps + size_t(D) * PT);  // [QT][k]
  uint64_t* cand = topk + size_t(QT) * k;                   // [QT][PT]
  int* ccount = reinterpret_cast<int*>(cand + size_t(QT) * PT);  // [QT]

  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int q0 = blockIdx.x * QT;
  const size_t p0 = size_t(blockIdx.y) * chunk;
  const size_t p1 = (p0 + chunk < n) ? p0 + chunk : n;

  for (int i = tid; i < QT * D; i += kBlock) {
    int r = i / D, e = i % D;
    int qg = q0 + r;
    qs[size_t(e) * QT + r] = (qg < nq && e < d) ? Q[size_t(qg) * q_stride + e] : 0.f;
  }
  for (int i = tid; i < QT * k; i += kBlock) topk[i] = kKeyMax;
  for (int i = tid; i < QT; i += kBlock) ccount[i] = 0;
  __syncthreads();

  const int DR = D & 31, DD = D - DR;
  for (size_t pb = p0; pb < p1; pb += PT) {
    for (int i = tid; i < PT * D; i += kBlock) {
      int r = i / D, e = i % D;
      size_t pg = pb + r;
      ps[size_t(e) * PT + r] = (pg < p1) ? X[pg * x_stride + e] : 0.f;
    }
    __syncthreads();
    float tot[MQ][MP];
#pragma unroll
    for (int j8 = 0; j8 < 8; ++j8) {
      float acc[MQ][MP];
#pragma unroll
      for (int a = 0; a < MQ; ++a)
#pragma unroll
        for (int b = 0; b < MP; ++b) acc[a][b] = 0.f;
      auto step = [&](int e) {
        float qv[MQ], pv[MP];
#pragma unroll
        for (int a = 0; a < MQ; ++a) qv[a] = qs[size_t(e) * QT + ty + 16 * a];
#pragma unroll
        for (int b = 0; b < MP; ++b) pv[b] = ps[size_t(e) * PT + tx + 16 * b];
#pragma unroll
        for (int a = 0; a < MQ; ++a)
#pragma unroll
          for (int b = 0; b < MP; ++b) {
            float t = qv[a] - pv[b];
            acc[a][b] = __builtin_fmaf(t, t, acc[a][b]);
          }
      };
      for (int cc = DD; cc < D; cc += 8) step(cc + j8);
      for (int cc = 0; cc < DD; cc += 8) step(cc + j8);
#pragma unroll
      for (int a = 0; a < MQ; ++a)
#pragma unroll
        for (int b = 0; b < MP; ++b) tot[a][b] = (j8 == 0) ? acc[a][b] : tot[a][b] + acc[a][b];
    }
    // candidates under the current k-th key
#pragma unroll
    for (int a = 0; a < MQ; ++a) {
      int ql = ty + 16 * a;
      int qg = q0 + ql;
      uint64_t thr = topk[size_t(ql) * k + (k - 1)];
#pragma unroll
      for (int b = 0; b < MP; ++b) {
        size_t pg = pb + tx + 16 * b;
        if (qg < nq && pg < p1 && !(exclude_self && pg == size_t(qg))) {
          uint64_t key = make_key(tot[a][b], uint32_t(pg));
          if (key < thr) cand[size_t(ql) * PT + atomicAdd(&ccount[ql], 1)] = key;
        }
      }
    }
    __syncthreads();
    for (int ql = tid; ql < QT; ql += kBlock) {
      int m = ccount[ql];
      uint64_t* tk = topk + size_t(ql) * k;
      for (int c = 0; c < m; ++c) {
        uint64_t key = cand[size_t(ql) * PT + c];
        if (key >= tk[k - 1]) continue;
        int j = k - 1;
        while (j > 0 && tk[j - 1] > key) {
          tk[j] = tk[j - 1];
          --j;
        }
        tk[j] = key;
      }
      ccount[ql] = 0;
    }
    __syncthreads();
  }
  for (int i = tid; i < QT * k; i += kBlock) {
    int ql = i / k, j = i % k;
    int qg = q0 + ql;
    if (qg < nq) partial[(size_t(qg) * nchunks + blockIdx.y) * k + j] = topk[i];
  }
}

__global__ void __launch_bounds__(kBlock) k_exact_merge(const uint64_t* __restrict__ partial, int nq, int nchunks,
                                                        int k, int32_t* __restrict__ ids, float* __restrict__ dists) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  uint64_t* buf = reinterpret_cast<uint64_t*>(smem_raw);
  const int m = nchunks * k;
  const int p2 = next_pow2(m < 2 ? 2 : m);
  for (int q = blockIdx.x; q < nq; q += gridDim.x) {
    for (int i = threadIdx.x; i < p2; i += kBlock) buf[i] = (i < m) ? partial[size_t(q) * m + i] : kKeyMax;
    __syncthreads();
    block_bitonic_sort(buf, p2);
    for (int j = threadIdx.x; j < k; j += kBlock) {
      uint64_t key = buf[j];
      bool ok = key != kKeyMax;
      ids[size_t(q) * k + j] = ok ? int32_t(key_id(key)) : -1;
      if (dists) dists[size_t(q) * k + j] = ok ? key_dist(key) : 3.402823466e38f;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Precision form: distances as ||q||^2 + ||x||^2 - 2 q.x with the dot products
// on BF16 (or FP16) MFMA tiles, FP32 accumulation; the best kc = 2k per query
// are then re-ranked by the direct FP32 difference in the reference's order.
// On data whose values are exact in the 16-bit type and whose sums stay below
// 2^24 (e.g. integers 0..255 with d <= 128) every step is exact and the
// result equals the FP32 search.
// ---------------------------------------------------------------------------
constexpr int kMP = 128, kMK = 64;  // base rows per tile, elements per LDS stage
constexpr int kMKP = kMK + 8;        // padded row (halves): 16 B shift per row
constexpr int kDS = kMP + 1;         // padded distance-tile row (floats)

// One workgroup per (32 kQB queries, chunk of base rows).  For every kMP-row
// base tile the (32 kQB) x 128 dot products are 32 x 32 MFMA blocks, kQB per
// wavefront (block b = w kQB + a: query block b / 4, base block b % 4), over
// the elements in stages of kMK through LDS; distances go to an LDS tile.
// All threads then mark the tile's entries below their query's kc-th key
// (kBlock / (32 kQB) threads per query), and one thread per query inserts
// only those into a max-heap of its kc smallest (distance, id) keys.
template <bool kF16, int kQB>
__global__ void __launch_bounds__(kBlock) k_exact_mfma(const uint16_t* __restrict__ Xh, const float* __restrict__ xn,
                                                       size_t n, const uint16_t* __restrict__ Qh,
                                                       const float* __restrict__ qn, int nq, int Dk, int kc,
                                                       int exclude_self, size_t chunk, uint64_t* __restrict__ partial,
                                                       int nchunks) {
  constexpr int kMQ = 32 * kQB;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  uint16_t* qs = reinterpret_cast<uint16_t*>(smem_raw);  // [kMQ][kMKP]
  uint16_t* xs = qs + kMQ * kMKP;                         // [kMP][kMKP]
  float* dt = reinterpret_cast<float*>(xs + kMP * kMKP);  // [kMQ][kDS]
  uint64_t* topk = reinterpret_cast<uint64_t*>(dt + kMQ * kDS);  // [kMQ][kc]
  uint32_t* smask = reinterpret_cast<uint32_t*>(topk + kMQ * kc);  // [kBlock] survivor masks
  constexpr int kSub = kBlock / kMQ;                                // threads per query in the survivor scan
  static_assert(kMP / kSub <= 32, "one 32-bit mask per thread");
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int q0 = blockIdx.x * kMQ;
  const size_t p0 = size_t(blockIdx.y) * chunk;
  const size_t p1 = (p0 + chunk < n) ? p0 + chunk : n;
  for (int i = tid; i < kMQ * kc; i += kBlock) topk[i] = kKeyMax;
  const int qb = (w * kQB) / 4, pb = (w * kQB) % 4;  // query block; first base block (kQB consecutive)
  const int r32 = lane & 31, kh = (lane >> 5) * 8;
  // A stage is (kMQ query rows + kMP base rows) x kMK elements, kLoads 16-byte
  // pieces per thread.  The next stage's pieces are loaded into registers
  // while this stage's MFMAs (and, at a tile's end, the selection) run.
  constexpr int kPieces = (kMQ + kMP) * (kMK / 8), kLoads = (kPieces + kBlock - 1) / kBlock;
  uint4 pre[kLoads];
  auto load_stage = [&](size_t pt_, int k0_) {
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
      const int i = tid + u * kBlock;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (i < kPieces) {
        const int row = i / (kMK / 8), c8 = (i % (kMK / 8)) * 8;
        if (row < kMQ) {
          const int qg = q0 + row;
          if (qg < nq) v = *reinterpret_cast<const uint4*>(Qh + size_t(qg) * Dk + k0_ + c8);
        } else {
          const size_t pg = pt_ + (row - kMQ);
          if (pg < p1) v = *reinterpret_cast<const uint4*>(Xh + pg * Dk + k0_ + c8);
        }
      }
      pre[u] = v;
    }
  };
  if (p0 < p1) load_stage(p0, 0);
  for (size_t pt = p0; pt < p1; pt += kMP) {
    f32x16_t acc[kQB];
#pragma unroll
    for (int a = 0; a < kQB; ++a) acc[a] = f32x16_t{};
    for (int k0 = 0; k0 < Dk; k0 += kMK) {
      __syncthreads();  // the previous stage (or the last tile's selection) is done with LDS
#pragma unroll
      for (int u = 0; u < kLoads; ++u) {
        const int i = tid + u * kBlock;
        if (i < kPieces) {
          const int row = i / (kMK / 8), c8 = (i % (kMK / 8)) * 8;
          uint16_t* dst = row < kMQ ? qs + row * kMKP + c8 : xs + (row - kMQ) * kMKP + c8;
          *reinterpret_cast<uint4*>(dst) = pre[u];
        }
      }
      {  // the next stage: the next elements of this tile, else the next tile's first
        const bool last = k0 + kMK >= Dk;
        const size_t npt = last ? pt + kMP : pt;
        if (npt < p1) load_stage(npt, last ? 0 : k0 + kMK);
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < kMK; ks += 16) {
        const uint4 av = *reinterpret_cast<const uint4*>(qs + (qb * 32 + r32) * kMKP + ks + kh);
#pragma unroll
        for (int a = 0; a < kQB; ++a) {
          const uint4 bv = *reinterpret_cast<const uint4*>(xs + ((pb + a) * 32 + r32) * kMKP + ks + kh);
          if (kF16) {
            f16x8_t x, y;
            __builtin_memcpy(&x, &av, 16);
            __builtin_memcpy(&y, &bv, 16);
            acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_f16(x, y, acc[a], 0, 0, 0);
          } else {
            bf16x8_t x, y;
            __builtin_memcpy(&x, &av, 16);
            __builtin_memcpy(&y, &bv, 16);
            acc[a] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y, acc[a], 0, 0, 0);
          }
        }
      }
    }
    // element i of lane l: query row 8 (i / 4) + 4 (l / 32) + i % 4 of the
    // block, base column l % 32
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int ql = qb * 32 + 8 * (i / 4) + 4 * (lane >> 5) + (i & 3);
      const int qg = q0 + ql;
      const float qv = qg < nq ? qn[qg] : 0.f;
#pragma unroll
      for (int a = 0; a < kQB; ++a) {
        const int pl = (pb + a) * 32 + r32;
        const size_t pg = pt + pl;
        const float xv = pg < p1 ? xn[pg] : 0.f;
        const float dd = __builtin_fmaf(-2.f, acc[a][i], qv + xv);
        dt[ql * kDS + pl] = dd > 0.f ? dd : 0.f;
      }
    }
    __syncthreads();
    const int m = int(p1 - pt < size_t(kMP) ? p1 - pt : size_t(kMP));
    {
      // survivors: the kSub threads of a query check the columns sub, sub +
      // kSub, ... against the query's threshold at the start of the tile
      // and record them in bit masks (bit t: column sub + t kSub)
      const int ql = tid / kSub, sub = tid % kSub;
      const int qg = q0 + ql;
      const uint64_t thr = topk[size_t(ql) * kc];  // the heap's largest key
      uint32_t mask = 0;
      if (qg < nq) {
        for (int t = 0; sub + t * kSub < m; ++t) {
          const int j = sub + t * kSub;
          const size_t pg = pt + j;
          if (exclude_self && pg == size_t(qg)) continue;
          if (make_key(dt[ql * kDS + j], uint32_t(pg)) < thr) mask |= 1u << t;
        }
      }
      smask[tid] = mask;
    }
    __syncthreads();
    // one thread per query inserts its survivors into the query's max-heap
    // of its kc smallest keys (the re-rank sorts them; the kc smallest keys
    // do not depend on the order of insertion, keys being distinct)
    if (tid < kMQ && q0 + tid < nq) {
      uint64_t* tk = topk + size_t(tid) * kc;
      uint64_t thr = tk[0];
      for (int sub = 0; sub < kSub; ++sub) {
        uint32_t mask = smask[tid * kSub + sub];
        while (mask) {
          const int j = sub + (__ffs(int(mask)) - 1) * kSub;
          mask &= mask - 1;
          const uint64_t key = make_key(dt[tid * kDS + j], uint32_t(pt + j));
          if (key >= thr) continue;
          int z = 0;  // the root is replaced and sifted down
          for (;;) {
            const int l = 2 * z + 1;
            if (l >= kc) break;
            const int cz = (l + 1 < kc && tk[l + 1] > tk[l]) ? l + 1 : l;
            if (tk[cz] <= key) break;
            tk[z] = tk[cz];
            z = cz;
          }
          tk[z] = key;
          thr = tk[0];
        }
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < kMQ * kc; i += kBlock) {
    const int ql = i / kc, j = i % kc;
    const int qg = q0 + ql;
    if (qg < nq) partial[(size_t(qg) * nchunks + blockIdx.y) * kc + j] = topk[i];
  }
}

// One workgroup per query: the kc best keys over the chunks, re-measured by
// the direct FP32 difference (reference order), sorted, first k written.
__global__ void __launch_bounds__(kBlock) k_exact_rerank(const uint64_t* __restrict__ partial, int nchunks, int kc,
                                                         const float* __restrict__ X, size_t x_stride,
                                                         const float* __restrict__ Q, size_t q_stride, int d, int D,
                                                         int nq, int k, int32_t* __restrict__ ids,
                                                         float* __restrict__ dists) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  uint64_t* buf = reinterpret_cast<uint64_t*>(smem_raw);
  float* qrow = reinterpret_cast<float*>(buf + next_pow2(nchunks * kc < 2 ? 2 : nchunks * kc));
  const int m = nchunks * kc;
  const int p2 = next_pow2(m < 2 ? 2 : m);
  const int pc = next_pow2(kc < 2 ? 2 : kc);
  const int pr = threadIdx.x >> 1, h = threadIdx.x & 1;
  for (int q = blockIdx.x; q < nq; q += gridDim.x) {
    for (int i = threadIdx.x; i < p2; i += kBlock) buf[i] = (i < m) ? partial[size_t(q) * m + i] : kKeyMax;
    for (int i = threadIdx.x; i < D; i += kBlock) qrow[i] = i < d ? Q[size_t(q) * q_stride + i] : 0.f;
    __syncthreads();
    block_bitonic_sort(buf, p2);
    for (int i = pr; i < kc; i += kBlock / 2) {
      const uint64_t key = buf[i];
      float dd = 3.402823466e38f;
      if (key != kKeyMax) dd = l2_pair_lanes(qrow, X + size_t(key_id(key)) * x_stride, D, h);
      if (h) buf[i] = key != kKeyMax ? make_key(dd, key_id(key)) : kKeyMax;
    }
    for (int i = kc + threadIdx.x; i < pc; i += kBlock) buf[i] = kKeyMax;
    __syncthreads();
    block_bitonic_sort(buf, pc);
    for (int j = threadIdx.x; j < k; j += kBlock) {
      const uint64_t key = buf[j];
      const bool ok = key != kKeyMax;
      ids[size_t(q) * k + j] = ok ? int32_t(key_id(key)) : -1;
      if (dists) dists[size_t(q) * k + j] = ok ? key_dist(key) : 3.402823466e38f;
    }
    __syncthreads();
  }
}

namespace {
template <int MQ, int MP>
size_t tile_lds(int D, int k) {
  constexpr int QT = 16 * MQ, PT = 16 * MP;
  return size_t(D) * (QT + PT) * 4 + size_t(QT) * k * 8 + size_t(QT) * PT * 8 + QT * 4;
}
}  // namespace

int exact_knn_mfma(const float* X, size_t n, int d, size_t x_stride, const float* Q, int nq, size_t q_stride, int k,
                   int exclude_self, int precision, int32_t* ids, float* dists, hipStream_t st, std::string* err) {
  if (k <= 0 || k > 512 || nq <= 0 || n == 0 || d <= 0) { *err = "exact_knn: bad arguments"; return -2; }
  if (n >= (size_t(1) << 32)) { *err = "exact_knn: too many base rows"; return -2; }
  const int D = (d + 7) & ~7;
  if (x_stride < size_t(D)) { *err = "exact_knn: base rows must be padded to a multiple of 8"; return -2; }
  const int Dk = (d + kMK - 1) / kMK * kMK;
  const int kc = int(std::min<size_t>(n, size_t(2 * k)));  // candidates re-ranked per query
  auto lds_for = [&](int qb) {
    return size_t(32 * qb + kMP) * kMKP * 2 + size_t(32 * qb) * kDS * 4 + size_t(32 * qb) * kc * 8 +
           size_t(kBlock) * 4;
  };
  const size_t lds_cap = 150 * 1024;
  const int qb = lds_for(2) <= lds_cap ? 2 : 1;
  const size_t lds = lds_for(qb);
  if (lds > lds_cap) { *err = "exact_knn: k too large for the MFMA tiles"; return -2; }
  // base chunks per query tile: enough workgroups to fill the GPU, no more
  // (each chunk adds kc candidates per query to the re-rank)
  const size_t qtiles = (size_t(nq) + 32 * qb - 1) / (32 * qb);
  const size_t want = std::max<size_t>(1, std::min<size_t>((2048 + qtiles - 1) / qtiles, 16384 / size_t(kc)));
  const size_t chunk = std::max<size_t>(kMP, ((n + want - 1) / want + kMP - 1) / kMP * kMP);
  const int nchunks = int((n + chunk - 1) / chunk);
  const int m = nchunks * kc;
  const size_t lds_r = size_t(next_pow2(m < 2 ? 2 : m)) * 8 + size_t(D) * 4;
  if (lds_r > lds_cap) { *err = "exact_knn: too many candidates for the re-rank"; return -2; }
  uint16_t *xh = nullptr, *qh = nullptr;
  float *xn = nullptr, *qn = nullptr;
  uint64_t* partial = nullptr;
  if (hipMallocAsync(reinterpret_cast<void**>(&xh), n * Dk * 2, st) != hipSuccess ||
      hipMallocAsync(reinterpret_cast<void**>(&qh), size_t(nq) * Dk * 2, st) != hipSuccess ||
      hipMallocAsync(reinterpret_cast<void**>(&xn), n * 4, st) != hipSuccess ||
      hipMallocAsync(reinterpret_cast<void**>(&qn), size_t(nq) * 4, st) != hipSuccess ||
      hipMallocAsync(reinterpret_cast<void**>(&partial), size_t(nq) * m * 8, st) != hipSuccess) {
    *err = "exact_knn: out of device memory";
    return -9;
  }
  const bool f16 = precision == 2;
  auto pack = f16 ? &k_pack16<true> : &k_pack16<false>;
  pack<<<unsigned((n + 3) / 4), 256, 0, st>>>(X, n, d, x_stride, Dk, xh, xn);
  pack<<<unsigned((nq + 3) / 4), 256, 0, st>>>(Q, size_t(nq), d, q_stride, Dk, qh, qn);
  auto kern = f16 ? (qb == 2 ? &k_exact_mfma<true, 2> : &k_exact_mfma<true, 1>)
                  : (qb == 2 ? &k_exact_mfma<false, 2> : &k_exact_mfma<false, 1>);
  hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  dim3 grid(unsigned((nq + 32 * qb - 1) / (32 * qb)), unsigned(nchunks));
  kern<<<grid, kBlock, lds, st>>>(xh, xn, n, qh, qn, nq, Dk, kc, exclude_self, chunk, partial, nchunks);
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_exact_rerank), hipFuncAttributeMaxDynamicSharedMemorySize,
                      int(lds_r));
  k_exact_rerank<<<unsigned(nq < 4096 ? nq : 4096), kBlock, lds_r, st>>>(partial, nchunks, kc, X, x_stride, Q,
                                                                        q_stride, d, D, nq, k, ids, dists);
  hipFreeAsync(partial, st);
  hipFreeAsync(qn, st);
  hipFreeAsync(xn, st);
  hipFreeAsync(qh, st);
  hipFreeAsync(xh, st);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { *err = std::string("exact_knn: ") + hipGetErrorString(e); return -8; }
  return 0;
}

int exact_knn(const float* X, size_t n, int d, size_t x_stride, const float* Q, int nq, size_t q_stride, int k,
              int exclude_self, int32_t* ids, float* dists, hipStream_t st, std::string* err) {
  if (k <= 0 || k > 1024 || nq <= 0 || n == 0 || d <= 0) { *err = "exact_knn: bad arguments"; return -2; }
  const int D = (d + 7) & ~7;
  if (x_stride < size_t(D) && x_stride != size_t(d)) { *err = "exact_knn: base stride"; return -2; }
  // tiles read elements up to D; pad handled by masking only for queries, so
  // the base stride must cover D unless D == d
  if (x_stride < size_t(D)) { *err = "exact_knn: base rows must be padded to a multiple of 8"; return -2; }
  const size_t lds_cap = 150 * 1024;
  bool big = tile_lds<4, 2>(D, k) <= lds_cap;
  size_t lds = big ? tile_lds<4, 2>(D, k) : tile_lds<1, 1>(D, k);
  if (lds > lds_cap) { *err = "exact_knn: dimension or k too large for LDS tiles"; return -2; }
  const int QT = big ? 64 : 16;
  const int chunk = 16384;
  const int nchunks = int((n + chunk - 1) / chunk);
  uint64_t* partial = nullptr;
  if (hipMallocAsync(reinterpret_cast<void**>(&partial), size_t(nq) * nchunks * k * 8, st) != hipSuccess) {
    *err = "exact_knn: out of device memory";
    return -9;
  }
  dim3 grid((nq + QT - 1) / QT, nchunks);
  if (big) {
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_exact_tile<4, 2>), hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    k_exact_tile<4, 2><<<grid, kBlock, lds, st>>>(X, n, d, D, x_stride, Q, nq, q_stride, k, exclude_self, chunk, partial, nchunks);
  } else {
    hipFuncSetAttribute(reinterpret_cast<const void*>(&k_exact_tile<1, 1>), hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    k_exact_tile<1, 1><<<grid, kBlock, lds, st>>>(X, n, d, D, x_stride, Q, nq, q_stride, k, exclude_self, chunk, partial, nchunks);
  }
  int m = nchunks * k;
  size_t lds2 = size_t(next_pow2(m < 2 ? 2 : m)) * 8;
  if (lds2 > lds_cap) { hipFreeAsync(partial, st); *err = "exact_knn: too many chunks for the merge"; return -2; }
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_exact_merge), hipFuncAttributeMaxDynamicSharedMemorySize, int(lds2));
  int g2 = nq < 4096 ? nq : 4096;
  k_exact_merge<<<g2, kBlock, lds2, st>>>(partial, nq, nchunks, k, ids, dists);
  hipFreeAsync(partial, st);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) { *err = std::string("exact_knn: ") + hipGetErrorString(e); return -8; }
  return 0;
}

}  // namespace efg

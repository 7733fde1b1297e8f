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
#include "efg_device.hpp"
#include "efg_internal.hpp"

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

namespace {
template <int MQ, int MP>
size_t tile_lds(int D, int k) {
  constexpr int QT = 16 * MQ, PT = 16 * MP;
  return size_t(D) * (QT + PT) * 4 + size_t(QT) * k * 8 + size_t(QT) * PT * 8 + QT * 4;
}
}  // namespace

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

// Device-side building blocks shared by the EFANNA HIP kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace efg {

constexpr int kBlock = 256;  // threads per workgroup for the block-level kernels

// ---------------------------------------------------------------------------
// Counter-based generator: splitmix64 finaliser over (seed, stream, a, b).
// Streams name the reference's random draws this generator replaces.
// ---------------------------------------------------------------------------
enum : uint64_t {
  ST_PERM = 1,      // leaf-list shuffle (kdtreeub_index.hpp:1313-1319)
  ST_SPLIT = 2,     // top-5 variance pick (kdtreeub_index.hpp:997)
  ST_INIT_TOP = 3,  // heap top-up to S (kdtreeub_index.hpp:1399-1405)
  ST_INIT_NEW = 4,  // GenRandom fill of nn_new (kdtreeub_index.hpp:1390-1393)
  ST_RNN_NEW = 5,   // reverse-list sampling (base_index.hpp:426-429)
  ST_RNN_OLD = 6,   // reverse-list sampling (base_index.hpp:431-434)
  ST_SEARCH = 7,    // search seed top-up (kdtreeub_index.hpp:411-417)
  ST_MLP_INIT = 8,  // learned stopping network: initial weights
  ST_MLP_PERM = 9,  // learned stopping network: example order per training epoch
};

__host__ __device__ inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
__host__ __device__ inline uint64_t hash4(uint64_t seed, uint64_t stream, uint64_t a, uint64_t b) {
  return mix64(mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL)) ^ a) ^ b);
}

// Bijection of [0, n) keyed by (seed, stream, key): 4-round Feistel network on
// the next power-of-two domain with cycle walking.
__host__ __device__ inline uint64_t feistel_perm(uint64_t seed, uint64_t stream, uint64_t key, uint64_t n,
                                                 uint64_t i) {
  unsigned bits = 2;
  while ((1ULL << bits) < n) ++bits;
  const unsigned h = (bits + 1) / 2;
  const uint64_t mask = (1ULL << h) - 1;
  uint64_t x = i;
  do {
    uint64_t L = x >> h, R = x & mask;
    for (uint64_t r = 0; r < 4; ++r) {
      const uint64_t F = hash4(seed, stream, (key << 8) | r, R) & mask;
      const uint64_t nL = R, nR = L ^ F;
      L = nL;
      R = nR;
    }
    x = (L << h) | R;
  } while (x >= n);
  return x;
}

// One element of a row as FP32: kRow 0 reads FP32 rows, 1 the high BF16
// parts and 2 the high FP16 parts of the build's split rows.  The 16-bit
// forms are used only when every row is exact in them, so every kernel
// reading them sees the same values as from the FP32 rows.
template <int kRow>
__device__ __forceinline__ float row_elem(const void* __restrict__ X, size_t off) {
  if constexpr (kRow == 0) return static_cast<const float*>(X)[off];
  else if constexpr (kRow == 1) return __uint_as_float(uint32_t(static_cast<const uint16_t*>(X)[off]) << 16);
  else return float(static_cast<const _Float16*>(X)[off]);
}

// (distance, id) packed so that unsigned order = (distance, id) order for
// non-negative distances.
__host__ __device__ inline uint64_t make_key(float d, uint32_t id) {
  return (uint64_t(__float_as_uint(d)) << 32) | id;
}
__host__ __device__ inline float key_dist(uint64_t k) { return __uint_as_float(uint32_t(k >> 32)); }
__host__ __device__ inline uint32_t key_id(uint64_t k) { return uint32_t(k & 0xffffffffu); }
constexpr uint64_t kKeyMax = ~0ULL;

// ---------------------------------------------------------------------------
// Squared L2 in the accumulation order of the reference's AVX kernel
// (general/distance.hpp:299-323, fused multiply-add per lane as g++ -O3
// -march=native emits it): eight lane accumulators over elements j mod 8, the
// tail chunks [DD, D) first, then [0, DD) in order, then u0+u1+...+u7 left to
// right.  Two cooperating lanes per row: lane h (0/1) owns accumulators
// 4h..4h+3 and loads the float4 at offset 8m+4h of every 8-float chunk m.
// `a` may live in LDS or global memory; `b` is a data row padded to D with
// zeros.  The result is valid in the h == 1 lane of the pair.
// ---------------------------------------------------------------------------
__device__ inline float l2_pair_lanes(const float* __restrict__ a, const float* __restrict__ b, int D, int h) {
  const int DR = D & 31, DD = D - DR;
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#define EFG_L2_STEP(off)                                                   \
  {                                                                        \
    const float4 x = *reinterpret_cast<const float4*>(a + (off));          \
    const float4 y = *reinterpret_cast<const float4*>(b + (off));          \
    float t0 = x.x - y.x, t1 = x.y - y.y, t2 = x.z - y.z, t3 = x.w - y.w; \
    acc.x = __builtin_fmaf(t0, t0, acc.x);                                 \
    acc.y = __builtin_fmaf(t1, t1, acc.y);                                 \
    acc.z = __builtin_fmaf(t2, t2, acc.z);                                 \
    acc.w = __builtin_fmaf(t3, t3, acc.w);                                 \
  }
  for (int c = DD; c < D; c += 8) EFG_L2_STEP(c + 4 * h)
#pragma unroll 8
  for (int c = 0; c < DD; c += 8) EFG_L2_STEP(c + 4 * h)
#undef EFG_L2_STEP
  // lane 0: ((u0+u1)+u2)+u3 ; lane 1 continues with +u4+u5+u6+u7
  float s0 = ((acc.x + acc.y) + acc.z) + acc.w;
  float p = __shfl_xor(s0, 1);
  float s1 = (((p + acc.x) + acc.y) + acc.z) + acc.w;
  return s1;
}

// Block-wide bitonic sort of buf[0, n), n a power of two no larger than the
// block: one key per thread in registers; partners within a wavefront are
// exchanged by shuffles, only steps across wavefronts (j >= 64) go through
// LDS.  Ascending; threads >= n hold the largest key.
__device__ inline void block_bitonic_sort_reg(uint64_t* buf, int n) {
  const int tid = threadIdx.x, nt = blockDim.x;
  __syncthreads();  // the callers' writes to buf are visible
  uint64_t v = tid < n ? buf[tid] : ~0ull;
  const int m = n < 64 ? 64 : n;  // sort at least a wavefront's worth (padding sorts last)
  for (int k = 2; k <= m; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      uint64_t w;
      if (j >= 64) {
        if (tid < m) buf[tid] = v;
        __syncthreads();
        w = (tid < m) ? buf[tid ^ j] : v;
        __syncthreads();
      } else {
        const uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(v)), j)), hi = uint32_t(__shfl_xor(int(uint32_t(v >> 32)), j));
        w = uint64_t(hi) << 32 | lo;
      }
      const bool lower = (tid & j) == 0, asc = (tid & k) == 0;
      v = (lower == asc) ? (v < w ? v : w) : (v < w ? w : v);
    }
  }
  (void)nt;
  if (tid < n) buf[tid] = v;
  __syncthreads();
}

// Block-wide bitonic sort of n (power of two) keys in LDS, ascending.
__device__ inline void block_bitonic_sort(uint64_t* buf, int n) {
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int ixj = i ^ j;
        if (ixj > i) {
          uint64_t a = buf[i], b = buf[ixj];
          bool up = (i & k) == 0;
          if ((a > b) == up) {
            buf[i] = b;
            buf[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

__host__ __device__ inline int next_pow2(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

__device__ inline uint32_t hslot(uint32_t id, int logcap) { return (id * 0x9E3779B1u) >> (32 - logcap); }

// Block exclusive scan of one int per thread (blockDim == kBlock); returns the
// exclusive prefix and writes the block total to *total.  `tmp` holds
// kBlock/64 ints in LDS.
__device__ inline int block_exclusive_scan(int v, int* tmp, int* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) tmp[wid] = x;
  __syncthreads();
  int base = 0, tot = 0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) {
    int t = tmp[w];
    if (w < wid) base += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

}  // namespace efg

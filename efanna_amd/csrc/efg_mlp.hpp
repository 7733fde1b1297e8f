// Learned stopping network shared by the search kernel and the training
// kernels: 16 pool-state features -> 64 -> 64 -> 1 logit, ReLU hidden units.
// Parameters (5,313) in one flat array:
//   W1[16][64] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3
// Inference and the training forward pass use the parameters rounded to
// BF16 and accumulate in FP32 with fused multiply-adds in input order, so a
// CPU restatement with the same order reproduces every logit bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace efg {

constexpr int kFeat = 16, kHid = 64;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kFeat * kHid, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + kHid * kHid,
              kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + kHid, kMlpParams = kOffB3 + 1;
static_assert(kMlpParams == 5313, "parameter count");

__host__ __device__ inline float bf16_to_float(uint16_t b) {
  union {
    uint32_t u;
    float f;
  } x;
  x.u = uint32_t(b) << 16;
  return x.f;
}
__host__ __device__ inline uint16_t float_to_bf16(float f) {  // round to nearest even
  union {
    float f;
    uint32_t u;
  } x;
  x.f = f;
  return uint16_t((x.u + 0x7fffu + ((x.u >> 16) & 1u)) >> 16);
}

// Network layout for the search kernel: hidden unit j's BF16 weights packed
// two per word so that thread j reads word i at pack[i * kHid + j]:
// words [0, 8) W1 rows (2i, 2i+1), [8, 40) W2 rows (2i, 2i+1), 40 b1 | b2.
// Then W3[64] and b3 as floats.
constexpr int kPackWords = kFeat / 2 + kHid / 2 + 1;
constexpr int kPackLen = kPackWords * kHid;  // uint32 words, followed by kHid + 1 floats

// Raw features of one checkpoint of a probe search (pool sorted ascending).
struct CkptState {
  int e;              // checkpoint index: 0 after the seeds, e after epoch e
  float d1, dk;       // best and k-th pool distance
  float dE, dR;       // distance at extend_to and at the pool end
  float dk_hist[3];   // k-th distance at checkpoints e-1, e-2, e-3 (valid when e > j)
  float d1_hist[3];   // best distance at checkpoints e-1, e-2, e-3
  float visited;      // ids measured so far
  float expanded;     // ids expanded in the last epoch
  float changed;      // ids in the top k that were not there at the last checkpoint
  float frontier;     // unexpanded ids among the first extend_to
};

__host__ __device__ inline float safe_ratio(float a, float b) { return b > 0.f ? a / b : 1.f; }
__host__ __device__ inline float rel_gain(float before, float now) {
  return before > 0.f ? (before - now) / before : 0.f;
}

__host__ __device__ inline void ckpt_features(const CkptState& s, float* f) {
  f[0] = float(s.e);
  f[1] = s.d1;
  f[2] = s.dk;
  f[3] = safe_ratio(s.d1, s.dk);
  f[4] = safe_ratio(s.dk, s.dE);
  f[5] = safe_ratio(s.dk, s.dR);
  for (int j = 0; j < 3; ++j) {
    f[6 + j] = s.e > j ? rel_gain(s.dk_hist[j], s.dk) : 0.f;
    f[9 + j] = s.e > j ? rel_gain(s.d1_hist[j], s.d1) : 0.f;
  }
  f[12] = s.visited;
  f[13] = s.expanded;
  f[14] = s.changed;
  f[15] = s.frontier;
}

// Standardised input: (f - mu) * inv_sd, per feature (norm = mu[16] | inv_sd[16]).
__host__ __device__ inline float mlp_input(const float* f, const float* norm, int i) {
  return (f[i] - norm[i]) * norm[kFeat + i];
}

}  // namespace efg

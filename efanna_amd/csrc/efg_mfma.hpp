// MFMA operand types and the 16-bit row copy shared by the exact-kNN and
// build kernels of the precision form.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace efg {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

namespace {

// 16-bit copy of rows (n x D floats, stride ld) padded to Dk (a multiple of
// 16) with zeros, and the FP32 squared norm of the rounded row.
template <bool kF16>
__global__ void k_pack16(const float* __restrict__ X, size_t n, int d, size_t ld, int Dk,
                         uint16_t* __restrict__ out, float* __restrict__ norm) {
  const size_t r = size_t(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= n) return;
  float acc = 0.f;
  for (int e = lane; e < Dk; e += 64) {
    const float v = e < d ? X[r * ld + e] : 0.f;
    uint16_t h;
    float back;
    if (kF16) {
      const _Float16 x = _Float16(v);
      __builtin_memcpy(&h, &x, 2);
      back = float(x);
    } else {
      const uint32_t u = __float_as_uint(v);
      h = uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
      back = __uint_as_float(uint32_t(h) << 16);
    }
    out[r * Dk + e] = h;
    acc = __builtin_fmaf(back, back, acc);
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) norm[r] = acc;
}

}  // namespace

}  // namespace efg

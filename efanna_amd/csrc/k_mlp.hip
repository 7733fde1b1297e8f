// Learned stopping network: batch inference, training forward + backward
// (mean binary cross-entropy on the logit), and Adam, all on the GPU.
//
// Inference and the training forward pass evaluate the network on BF16
// parameters with FP32 fused multiply-adds in input order (efg_mlp.hpp).
// Backward uses the same BF16 values (straight-through to the FP32 master
// copy).  Gradients are summed per workgroup over fixed example slices and
// the workgroup partials are added in workgroup order, so a training run is
// deterministic.
#include <cmath>
#include <vector>

#include "efg_device.hpp"
#include "efg_internal.hpp"
#include "efg_mlp.hpp"

namespace efg {

namespace {

constexpr int kCB = 64;        // examples per chunk in the gradient kernel
constexpr int kHS = kHid + 1;  // padded row stride of the activation tiles in LDS

__global__ void k_mlp_pack(const float* __restrict__ master, uint16_t* __restrict__ wb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kMlpParams) wb[p] = float_to_bf16(master[p]);
}

__device__ inline void load_params(const uint16_t* __restrict__ wb, float* w) {
  for (int p = threadIdx.x; p < kMlpParams; p += blockDim.x) w[p] = bf16_to_float(wb[p]);
}

// Logit of one standardised input, one thread (x holds kFeat inputs).
__device__ inline float mlp_logit_thread(const float* w, const float* x) {
  float h1[kHid];
#pragma unroll
  for (int j = 0; j < kHid; ++j) {
    float a = w[kOffB1 + j];
#pragma unroll
    for (int i = 0; i < kFeat; ++i) a = __builtin_fmaf(w[kOffW1 + i * kHid + j], x[i], a);
    h1[j] = a > 0.f ? a : 0.f;
  }
  float z = w[kOffB3];
  for (int j = 0; j < kHid; ++j) {
    float a = w[kOffB2 + j];
#pragma unroll
    for (int i = 0; i < kHid; ++i) a = __builtin_fmaf(w[kOffW2 + i * kHid + j], h1[i], a);
    z = __builtin_fmaf(w[kOffW3 + j], a > 0.f ? a : 0.f, z);
  }
  return z;
}

// Logits of n raw feature rows: one thread per row.
__global__ void __launch_bounds__(256) k_mlp_forward(const uint16_t* __restrict__ wb, const float* __restrict__ norm,
                                                     const float* __restrict__ feat, size_t n,
                                                     float* __restrict__ z) {
  __shared__ float w[kMlpParams];
  __shared__ float nm[2 * kFeat];
  load_params(wb, w);
  if (threadIdx.x < 2 * kFeat) nm[threadIdx.x] = norm[threadIdx.x];
  __syncthreads();
  for (size_t r = size_t(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += size_t(gridDim.x) * blockDim.x) {
    float x[kFeat];
#pragma unroll
    for (int i = 0; i < kFeat; ++i) x[i] = mlp_input(feat + r * kFeat, nm, i);
    z[r] = mlp_logit_thread(w, x);
  }
}

// Gradient of mean BCE over examples [base, base + count) of the order `idx`
// (identity when null): workgroup g takes the g-th contiguous slice, in
// chunks of kCB, and writes its partial sums to part[g][0, kMlpParams) and
// its loss sum to part[g][kMlpParams].  inv_n = 1 / (examples in the step).
__global__ void __launch_bounds__(256) k_mlp_grad(const uint16_t* __restrict__ wb, const float* __restrict__ norm,
                                                  const float* __restrict__ feat, const uint8_t* __restrict__ label,
                                                  const uint32_t* __restrict__ idx, size_t base, int count,
                                                  float inv_n, float* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float* w = reinterpret_cast<float*>(smem_raw);  // kMlpParams
  float* w2t = w + kMlpParams;                    // W2 transposed [j][i]
  float* X = w2t + kHid * kHid;                   // [kCB][kFeat]
  float* H1 = X + kCB * kFeat;                    // [kCB][kHS] post-ReLU
  float* H2 = H1 + kCB * kHS;
  float* D2 = H2 + kCB * kHS;  // dloss/dz2
  float* D1 = D2 + kCB * kHS;  // dloss/dz1
  float* dz = D1 + kCB * kHS;  // [kCB]
  float* ls = dz + kCB;        // [kCB] loss per example
  __shared__ float nm[2 * kFeat];
  const int tid = threadIdx.x;
  load_params(wb, w);
  if (tid < 2 * kFeat) nm[tid] = norm[tid];
  __syncthreads();
  for (int o = tid; o < kHid * kHid; o += blockDim.x) {
    const int i = o / kHid, j = o % kHid;
    w2t[j * kHid + i] = w[kOffW2 + o];
  }
  constexpr int kPer = (kMlpParams + 255) / 256;
  float acc[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r) acc[r] = 0.f;
  float loss_acc = 0.f;
  const int per_wg = (count + int(gridDim.x) - 1) / int(gridDim.x);
  const int s0 = min(count, int(blockIdx.x) * per_wg), s1 = min(count, s0 + per_wg);
  for (int c0 = s0; c0 < s1; c0 += kCB) {
    const int cb = min(kCB, s1 - c0);
    __syncthreads();
    for (int o = tid; o < kCB * kFeat; o += blockDim.x) {
      const int b = o / kFeat, i = o % kFeat;
      float v = 0.f;
      if (b < cb) {
        const size_t e = idx ? size_t(idx[base + c0 + b]) : base + c0 + b;
        v = mlp_input(feat + e * kFeat, nm, i);
      }
      X[o] = v;
    }
    if (tid < kCB) {
      float y = 0.f;
      if (tid < cb) {
        const size_t e = idx ? size_t(idx[base + c0 + tid]) : base + c0 + tid;
        y = label[e] ? 1.f : 0.f;
      }
      dz[tid] = y;  // label, replaced by the logit gradient below
    }
    __syncthreads();
    for (int o = tid; o < kCB * kHid; o += blockDim.x) {
      const int b = o / kHid, j = o % kHid;
      float a = w[kOffB1 + j];
      for (int i = 0; i < kFeat; ++i) a = __builtin_fmaf(w[kOffW1 + i * kHid + j], X[b * kFeat + i], a);
      H1[b * kHS + j] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    for (int o = tid; o < kCB * kHid; o += blockDim.x) {
      const int b = o / kHid, j = o % kHid;
      float a = w[kOffB2 + j];
      for (int i = 0; i < kHid; ++i) a = __builtin_fmaf(w[kOffW2 + i * kHid + j], H1[b * kHS + i], a);
      H2[b * kHS + j] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (tid < kCB) {
      float z = w[kOffB3];
      for (int i = 0; i < kHid; ++i) z = __builtin_fmaf(w[kOffW3 + i], H2[tid * kHS + i], z);
      const float y = dz[tid];
      const float pz = 1.f / (1.f + expf(-z));
      const bool in = tid < cb;
      dz[tid] = in ? (pz - y) * inv_n : 0.f;
      ls[tid] = in ? fmaxf(z, 0.f) - z * y + log1pf(expf(-fabsf(z))) : 0.f;
    }
    __syncthreads();
    for (int o = tid; o < kCB * kHid; o += blockDim.x) {
      const int b = o / kHid, i = o % kHid;
      D2[b * kHS + i] = H2[b * kHS + i] > 0.f ? dz[b] * w[kOffW3 + i] : 0.f;
    }
    __syncthreads();
    for (int o = tid; o < kCB * kHid; o += blockDim.x) {
      const int b = o / kHid, i = o % kHid;
      float a = 0.f;
      for (int j = 0; j < kHid; ++j) a = __builtin_fmaf(w2t[j * kHid + i], D2[b * kHS + j], a);
      D1[b * kHS + i] = H1[b * kHS + i] > 0.f ? a : 0.f;
    }
    __syncthreads();
    // parameter sums over the chunk's examples, in example order
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int p = tid + 256 * r;
      if (p >= kMlpParams) continue;
      float sm = 0.f;
      if (p < kOffB1) {
        const int i = (p - kOffW1) / kHid, j = (p - kOffW1) % kHid;
        for (int b = 0; b < kCB; ++b) sm = __builtin_fmaf(X[b * kFeat + i], D1[b * kHS + j], sm);
      } else if (p < kOffW2) {
        const int j = p - kOffB1;
        for (int b = 0; b < kCB; ++b) sm += D1[b * kHS + j];
      } else if (p < kOffB2) {
        const int i = (p - kOffW2) / kHid, j = (p - kOffW2) % kHid;
        for (int b = 0; b < kCB; ++b) sm = __builtin_fmaf(H1[b * kHS + i], D2[b * kHS + j], sm);
      } else if (p < kOffW3) {
        const int j = p - kOffB2;
        for (int b = 0; b < kCB; ++b) sm += D2[b * kHS + j];
      } else if (p < kOffB3) {
        const int i = p - kOffW3;
        for (int b = 0; b < kCB; ++b) sm = __builtin_fmaf(dz[b], H2[b * kHS + i], sm);
      } else {
        for (int b = 0; b < kCB; ++b) sm += dz[b];
      }
      acc[r] += sm;
    }
    if (tid == 0)
      for (int b = 0; b < kCB; ++b) loss_acc += ls[b];
  }
  float* out = part + size_t(blockIdx.x) * (kMlpParams + 1);
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const int p = tid + 256 * r;
    if (p < kMlpParams) out[p] = acc[r];
  }
  if (tid == 0) out[kMlpParams] = loss_acc;
}

// Sum of the workgroup partials (in workgroup order) into grad[0, kMlpParams]
// (the last entry is the loss sum); with lr > 0 also one Adam step on the
// FP32 master copy and its BF16 image.
__global__ void k_mlp_adam(const float* __restrict__ part, int G, float* __restrict__ grad,
                           float* __restrict__ master, float* __restrict__ m1, float* __restrict__ m2,
                           uint16_t* __restrict__ wb, float lr, float c1, float c2, float* __restrict__ loss_acc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kMlpParams) return;
  float g = 0.f;
  for (int w = 0; w < G; ++w) g += part[size_t(w) * (kMlpParams + 1) + p];
  if (grad) grad[p] = g;
  if (p == kMlpParams) {
    if (loss_acc) *loss_acc += g;
    return;
  }
  if (lr > 0.f) {
    const float a = 0.9f * m1[p] + 0.1f * g;
    const float b = 0.999f * m2[p] + 0.001f * g * g;
    m1[p] = a;
    m2[p] = b;
    const float mh = a / c1, vh = b / c2;
    const float v = master[p] - lr * mh / (sqrtf(vh) + 1e-8f);
    master[p] = v;
    wb[p] = float_to_bf16(v);
  }
}

constexpr int kGradG = 16;  // workgroups of the gradient kernel

size_t grad_lds() {
  return sizeof(float) * (size_t(kMlpParams) + kHid * kHid + kCB * kFeat + 4 * size_t(kCB) * kHS + 2 * kCB);
}

}  // namespace

int mlp_logits(const float* params, const float* norm, const float* feat, size_t n, float* z, hipStream_t st) {
  DevBuf<float> dm, dn, df, dz;
  DevBuf<uint16_t> wb;
  if (!dm.alloc(kMlpParams) || !wb.alloc(kMlpParams) || !dn.alloc(2 * kFeat) || !df.alloc(n * kFeat + 1) ||
      !dz.alloc(n + 1))
    return fail(EFG_ERR_MEMORY, "out of device memory (network)");
  hipMemcpyAsync(dm.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st);
  hipMemcpyAsync(dn.p, norm, sizeof(float) * 2 * kFeat, hipMemcpyHostToDevice, st);
  if (n) hipMemcpyAsync(df.p, feat, sizeof(float) * n * kFeat, hipMemcpyHostToDevice, st);
  k_mlp_pack<<<(kMlpParams + 255) / 256, 256, 0, st>>>(dm.p, wb.p);
  if (n) {
    const unsigned grid = unsigned(std::min<size_t>((n + 255) / 256, 1024));
    k_mlp_forward<<<grid, 256, 0, st>>>(wb.p, dn.p, df.p, n, dz.p);
    hipMemcpyAsync(z, dz.p, sizeof(float) * n, hipMemcpyDeviceToHost, st);
  }
  return check_hip(hipStreamSynchronize(st), "network forward");
}

int mlp_gradient(const float* params, const float* norm, const float* feat, const uint8_t* label, size_t n,
                 float* grad, float* loss, hipStream_t st) {
  if (n == 0 || n > (size_t(1) << 30)) return fail(EFG_ERR_ARG, "invalid example count");
  DevBuf<float> dm, dn, df, part, dg;
  DevBuf<uint16_t> wb;
  DevBuf<uint8_t> dl;
  if (!dm.alloc(kMlpParams) || !wb.alloc(kMlpParams) || !dn.alloc(2 * kFeat) || !df.alloc(n * kFeat) ||
      !dl.alloc(n) || !part.alloc(size_t(kGradG) * (kMlpParams + 1)) || !dg.alloc(kMlpParams + 1))
    return fail(EFG_ERR_MEMORY, "out of device memory (network)");
  hipMemcpyAsync(dm.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st);
  hipMemcpyAsync(dn.p, norm, sizeof(float) * 2 * kFeat, hipMemcpyHostToDevice, st);
  hipMemcpyAsync(df.p, feat, sizeof(float) * n * kFeat, hipMemcpyHostToDevice, st);
  hipMemcpyAsync(dl.p, label, n, hipMemcpyHostToDevice, st);
  k_mlp_pack<<<(kMlpParams + 255) / 256, 256, 0, st>>>(dm.p, wb.p);
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mlp_grad), hipFuncAttributeMaxDynamicSharedMemorySize,
                      int(grad_lds()));
  k_mlp_grad<<<kGradG, 256, grad_lds(), st>>>(wb.p, dn.p, df.p, dl.p, nullptr, 0, int(n), 1.f / float(n), part.p);
  k_mlp_adam<<<(kMlpParams + 256) / 256, 256, 0, st>>>(part.p, kGradG, dg.p, nullptr, nullptr, nullptr, nullptr, 0.f,
                                                       1.f, 1.f, nullptr);
  std::vector<float> h(kMlpParams + 1);
  hipMemcpyAsync(h.data(), dg.p, sizeof(float) * (kMlpParams + 1), hipMemcpyDeviceToHost, st);
  if (int rc = check_hip(hipStreamSynchronize(st), "network gradient")) return rc;
  std::copy(h.begin(), h.begin() + kMlpParams, grad);
  if (loss) *loss = h[kMlpParams] / float(n);
  return EFG_OK;
}

// Adam on mean BCE: `epochs` passes over the examples in an order drawn per
// epoch from the generator, `batch` examples per step.  params holds the
// initial FP32 weights on entry and the trained ones on return; loss_hist
// (nullable) receives the mean loss of every epoch.
int mlp_train(const float* feat, const uint8_t* label, size_t n, const float* norm, float* params, int epochs,
              int batch, float lr, uint64_t seed, float* loss_hist, hipStream_t st) {
  if (n == 0 || n > (size_t(1) << 30)) return fail(EFG_ERR_ARG, "invalid example count");
  if (epochs < 0 || batch <= 0 || !(lr > 0.f)) return fail(EFG_ERR_ARG, "invalid training parameters");
  DevBuf<float> dm, m1, m2, dn, df, part, lacc;
  DevBuf<uint16_t> wb;
  DevBuf<uint8_t> dl;
  DevBuf<uint32_t> order;
  if (!dm.alloc(kMlpParams) || !m1.alloc(kMlpParams) || !m2.alloc(kMlpParams) || !wb.alloc(kMlpParams) ||
      !dn.alloc(2 * kFeat) || !df.alloc(n * kFeat) || !dl.alloc(n) || !order.alloc(n) ||
      !part.alloc(size_t(kGradG) * (kMlpParams + 1)) || !lacc.alloc(size_t(epochs) + 1))
    return fail(EFG_ERR_MEMORY, "out of device memory (training)");
  hipMemcpyAsync(dm.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st);
  hipMemsetAsync(m1.p, 0, sizeof(float) * kMlpParams, st);
  hipMemsetAsync(m2.p, 0, sizeof(float) * kMlpParams, st);
  hipMemsetAsync(lacc.p, 0, sizeof(float) * (size_t(epochs) + 1), st);
  hipMemcpyAsync(dn.p, norm, sizeof(float) * 2 * kFeat, hipMemcpyHostToDevice, st);
  hipMemcpyAsync(df.p, feat, sizeof(float) * n * kFeat, hipMemcpyHostToDevice, st);
  hipMemcpyAsync(dl.p, label, n, hipMemcpyHostToDevice, st);
  k_mlp_pack<<<(kMlpParams + 255) / 256, 256, 0, st>>>(dm.p, wb.p);
  hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mlp_grad), hipFuncAttributeMaxDynamicSharedMemorySize,
                      int(grad_lds()));
  std::vector<uint32_t> h_order(n);
  long long t = 0;
  double b1t = 1.0, b2t = 1.0;
  for (int ep = 0; ep < epochs; ++ep) {
    for (size_t i = 0; i < n; ++i) h_order[i] = uint32_t(feistel_perm(seed, ST_MLP_PERM, uint64_t(ep), n, i));
    hipMemcpyAsync(order.p, h_order.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, st);
    for (size_t b0 = 0; b0 < n; b0 += size_t(batch)) {
      const int cnt = int(std::min<size_t>(size_t(batch), n - b0));
      ++t;
      b1t *= 0.9;
      b2t *= 0.999;
      k_mlp_grad<<<kGradG, 256, grad_lds(), st>>>(wb.p, dn.p, df.p, dl.p, order.p, b0, cnt, 1.f / float(cnt),
                                                  part.p);
      k_mlp_adam<<<(kMlpParams + 256) / 256, 256, 0, st>>>(part.p, kGradG, nullptr, dm.p, m1.p, m2.p, wb.p, lr,
                                                           float(1.0 - b1t), float(1.0 - b2t), lacc.p + ep);
    }
    // the hipMemcpyAsync of the next epoch's order reads a host buffer this
    // loop rewrites: finish the epoch first
    if (int rc = check_hip(hipStreamSynchronize(st), "training")) return rc;
  }
  hipMemcpyAsync(params, dm.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st);
  std::vector<float> h_loss(size_t(epochs) + 1);
  hipMemcpyAsync(h_loss.data(), lacc.p, sizeof(float) * (size_t(epochs) + 1), hipMemcpyDeviceToHost, st);
  if (int rc = check_hip(hipStreamSynchronize(st), "training")) return rc;
  if (loss_hist)
    for (int ep = 0; ep < epochs; ++ep) loss_hist[ep] = h_loss[ep] / float(n);
  return EFG_OK;
}

// Initial weights: uniform in +-sqrt(6 / (fan_in + fan_out)) from the
// generator, zero biases.
void mlp_init(float* params, uint64_t seed) {
  auto fill = [&](int off, int fin, int fout) {
    const double a = std::sqrt(6.0 / double(fin + fout));
    for (int i = 0; i < fin * fout; ++i) {
      const double u = double(hash4(seed, ST_MLP_INIT, uint64_t(off), uint64_t(i)) >> 11) * (1.0 / 9007199254740992.0);
      params[off + i] = float((2.0 * u - 1.0) * a);
    }
  };
  for (int p = 0; p < kMlpParams; ++p) params[p] = 0.f;
  fill(kOffW1, kFeat, kHid);
  fill(kOffW2, kHid, kHid);
  fill(kOffW3, kHid, 1);
}

}  // namespace efg

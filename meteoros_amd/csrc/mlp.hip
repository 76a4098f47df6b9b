// mlp.hip -- training and batch inference of the learned-light network.
//
// No reference counterpart.  The network (mlp_common.h) learns the cone
// light march's log transmittance from samples the march itself produces.
// Training is plain minibatch Adam on the MSE loss, run entirely on the GPU
// by this file's kernels:
//   train_step_kernel  one wave = 64 samples: forward on BF16 MFMA (FP32
//                      accumulation), loss gradient, backward (dW on MFMA,
//                      dH = dZ W^T on MFMA), FP32 gradient atomics
//   adam_kernel        FP32 Adam update of the master weights, BF16 copies
//   pack_kernel        BF16 MFMA fragments for inference
//   forward_kernel     batch inference through infer_wave (the routine the
//                      learned march kernel calls)
#include "mlp_common.h"

namespace mtc {
namespace mlp {

// BF16 row-major copies W[k][n] of the three hidden layers (layer 1 padded
// to 32 rows), written by adam_kernel / pack_kernel, read by training.
constexpr int kWb1 = 0, kWb2 = 32 * 64, kWb3 = 32 * 64 + 64 * 64, kWbTotal = 32 * 64 + 2 * 64 * 64;

__global__ __launch_bounds__(256) void pack_kernel(const float* __restrict__ p, bf16x8* __restrict__ frag,
                                                   uint16_t* __restrict__ wb) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < kFrags) {
    // fragment f, lane l: A(m, slot) with m = l % 16 the output unit in
    // tile nt and slot 8 (l / 16) + e the input (permuted for layers 2, 3)
    const int f = t / 64, lane = t % 64, g = lane / 16, m = lane % 16;
    bf16x8 v;
    if (f < 4) {
      for (int e = 0; e < 8; ++e) {
        const int k = 8 * g + e;
        v[e] = bfv(k < kIn ? f2bf(p[OW1 + k * 64 + 16 * f + m]) : (uint16_t)0);
      }
    } else {
      const int local = f < 12 ? f - 4 : f - 12, ow = f < 12 ? OW2 : OW3;
      const int nt = local / 2, c = local % 2;
      for (int e = 0; e < 8; ++e) v[e] = bfv(f2bf(p[ow + hidden_unit(c, g, e) * 64 + 16 * nt + m]));
    }
    frag[t] = v;
  }
  if (t < kWbTotal) {
    int layer = t < kWb2 ? 0 : (t < kWb3 ? 1 : 2);
    int local = t - (layer == 0 ? kWb1 : (layer == 1 ? kWb2 : kWb3));
    int k = local / 64, n = local % 64;
    int ow = layer == 0 ? OW1 : (layer == 1 ? OW2 : OW3);
    wb[t] = (layer == 0 && k >= kIn) ? (uint16_t)0 : f2bf(p[ow + k * 64 + n]);
  }
}

struct TrainLds {
  uint16_t X[64 * 32];   // inputs, zero padded
  uint16_t H1[64 * 64];  // activations (bf16); H3 / H2 are reused for dZ3 / dZ1
  uint16_t H2[64 * 64];
  uint16_t H3[64 * 64];
  uint16_t S1[64 * 64];  // silu'(z) of each layer (bf16)
  uint16_t S2[64 * 64];
  uint16_t S3[64 * 64];
  uint16_t dZ[64 * 64];  // dZ2
  float dy[64];
};

// dW (k x n tiles) of W[k][n]: A = H^T with H(q, k) at h[q * hs + k],
// B = dZ(q, n).  Tiles cover krows rows; only rows k < kvalid exist in W.
__device__ __forceinline__ void accumulate_dw(const uint16_t* h, int hs, int krows, int kvalid, const uint16_t* dz,
                                              float* __restrict__ grad) {
  const int lane = threadIdx.x & 63;
  for (int m0 = 0; m0 < krows; m0 += 16)
    for (int n0 = 0; n0 < 64; n0 += 16) {
      floatx4 acc = tile_strided(h, 1, hs, dz, 64, 1, m0, n0, 64);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int k = m0 + 4 * (lane >> 4) + q, n = n0 + (lane & 15);
        if (k < kvalid) atomicAdd(&grad[k * 64 + n], acc[q]);
      }
    }
}

// db[n] = sum_q dZ(q, n) from the bf16 dZ
__device__ __forceinline__ void accumulate_db(const uint16_t* dz, float* __restrict__ grad) {
  const int lane = threadIdx.x & 63;
  float s = 0.0f;
  for (int q = 0; q < 64; ++q) s += bf2f(dz[q * 64 + lane]);
  atomicAdd(&grad[lane], s);
}

// dZ_prev(q, k) = (dZ W^T)(q, k) * S(q, k); W(k, n) row-major in global bf16
__device__ __forceinline__ void backprop(const uint16_t* dz, const uint16_t* __restrict__ w, const uint16_t* s,
                                         uint16_t* out) {
  const int lane = threadIdx.x & 63;
  floatx4 acc[4][4];
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt)
      // A = dZ(q, n) [q x n], B = W^T(n, k) = w[k * 64 + n]
      acc[mt][nt] = tile_strided(dz, 64, 1, w, 1, 64, mt * 16, nt * 16, 64);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int row = mt * 16 + 4 * (lane >> 4) + q, col = nt * 16 + (lane & 15);
        out[row * 64 + col] = f2bf(acc[mt][nt][q] * bf2f(s[row * 64 + col]));
      }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// forward of one layer: out = silu(A W + b), S = silu'(A W + b)
__device__ __forceinline__ void forward_layer(const uint16_t* a, int as, int K, const uint16_t* __restrict__ w,
                                              const float* __restrict__ b, uint16_t* out, uint16_t* s) {
  const int lane = threadIdx.x & 63;
  for (int mt = 0; mt < 4; ++mt)
    for (int nt = 0; nt < 4; ++nt) {
      floatx4 acc = tile_strided(a, as, 1, w, 64, 1, mt * 16, nt * 16, K);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int row = mt * 16 + 4 * (lane >> 4) + q, col = nt * 16 + (lane & 15);
        float z = acc[q] + b[col];
        out[row * 64 + col] = f2bf(silu(z));
        s[row * 64 + col] = f2bf(silu_grad(z));
      }
    }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sample index of minibatch slot i: a hash of (seed, i) over the n_total
// stored samples, or slot i itself when seed == 0.
__device__ __forceinline__ int minibatch_index(uint32_t seed, int i, int n_total) {
  if (seed == 0u) return i;
  uint32_t v = (uint32_t)i * 747796405u + seed * 2891336453u + 0x9e3779b9u;
  v = ((v >> ((v >> 28u) + 4u)) ^ v) * 277803737u;
  v = (v >> 22u) ^ v;
  return (int)(((uint64_t)v * (uint64_t)n_total) >> 32);
}

// One training step over a minibatch of `n` samples drawn from the n_total
// stored ones: every 64-thread workgroup takes 64 slots (slots past n are
// zero-weighted), adds its gradient of loss = mean((f(x) - y)^2) into grad,
// and its part of the loss into loss.
__global__ __launch_bounds__(64) void train_step_kernel(const float* __restrict__ xs, const float* __restrict__ ys,
                                                        int n_total, uint32_t seed, int n,
                                                        const float* __restrict__ p, const uint16_t* __restrict__ wb,
                                                        float* __restrict__ grad, float* __restrict__ loss) {
  __shared__ __align__(16) TrainLds L;
  const int lane = threadIdx.x;
  const int sidx = blockIdx.x * 64 + lane;
  const bool live = sidx < n;
  const int src = live ? minibatch_index(seed, sidx, n_total) : 0;
  for (int k = 0; k < 32; ++k) L.X[lane * 32 + k] = (live && k < kIn) ? f2bf(xs[(size_t)src * kIn + k]) : (uint16_t)0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  forward_layer(L.X, 32, 32, wb + kWb1, p + OB1, L.H1, L.S1);
  forward_layer(L.H1, 64, 64, wb + kWb2, p + OB2, L.H2, L.S2);
  forward_layer(L.H2, 64, 64, wb + kWb3, p + OB3, L.H3, L.S3);
  float yhat = p[OB4];
  for (int k = 0; k < 64; ++k) yhat += bf2f(L.H3[lane * 64 + k]) * p[OW4 + k];
  const float err = live ? yhat - (float)ys[src] : 0.0f;
  const float dy = 2.0f * err / (float)n;
  L.dy[lane] = dy;
  float se = err * err;
  for (int off = 32; off > 0; off >>= 1) se += __shfl_xor(se, off, 64);
  if (lane == 0) atomicAdd(loss, se / (float)n);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // output layer: dw4[k] = sum_q H3(q, k) dy_q, db4 = sum_q dy_q
  {
    float s = 0.0f;
    for (int q = 0; q < 64; ++q) s += bf2f(L.H3[q * 64 + lane]) * L.dy[q];
    atomicAdd(&grad[OW4 + lane], s);
    float sb = dy;
    for (int off = 32; off > 0; off >>= 1) sb += __shfl_xor(sb, off, 64);
    if (lane == 0) atomicAdd(&grad[OB4], sb);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // dZ3(q, k) = dy_q w4[k] silu'(z3) -> into the H3 buffer
  for (int k = 0; k < 64; ++k) {
    float g = dy * p[OW4 + k] * bf2f(L.S3[lane * 64 + k]);
    L.H3[lane * 64 + k] = f2bf(g);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  accumulate_dw(L.H2, 64, 64, 64, L.H3, grad + OW3);
  accumulate_db(L.H3, grad + OB3);
  backprop(L.H3, wb + kWb3, L.S2, L.dZ);  // dZ2
  accumulate_dw(L.H1, 64, 64, 64, L.dZ, grad + OW2);
  accumulate_db(L.dZ, grad + OB2);
  backprop(L.dZ, wb + kWb2, L.S1, L.H2);  // dZ1 into the H2 buffer
  accumulate_dw(L.X, 32, 16, kIn, L.H2, grad + OW1);  // one 16-row tile, rows k < 9
  accumulate_db(L.H2, grad + OB1);
}

// Adam on the FP32 master weights; clears the gradient.  The BF16 copies
// are refreshed by pack_kernel afterwards.
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ p, float* __restrict__ grad,
                                                   float* __restrict__ m, float* __restrict__ v, float lr,
                                                   float b1, float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= kParams) return;
  float g = grad[t];
  grad[t] = 0.0f;
  float mm = b1 * m[t] + (1.0f - b1) * g;
  float vv = b2 * v[t] + (1.0f - b2) * g * g;
  m[t] = mm;
  v[t] = vv;
  p[t] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

__global__ __launch_bounds__(256) void zero_kernel(float* __restrict__ a, int n) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < n) a[t] = 0.0f;
}

// Batch inference: one wave per 64 queries, through infer_wave (the
// routine the learned march kernel calls).
__global__ __launch_bounds__(64) void forward_kernel(const float* __restrict__ xs, int n,
                                                     const bf16x8* __restrict__ frag, const float* __restrict__ p,
                                                     float* __restrict__ ys) {
  __shared__ __align__(16) uint16_t X[64 * 16];
  const int lane = threadIdx.x;
  const int q = blockIdx.x * 64 + lane;
  for (int k = 0; k < 16; ++k) X[lane * 16 + k] = (q < n && k < kIn) ? f2bf(xs[(size_t)q * kIn + k]) : (uint16_t)0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int nq = min(64, n - blockIdx.x * 64);
  float y = infer_wave(X, frag, p, nq);
  if (q < n) ys[q] = y;
}

}  // namespace mlp
}  // namespace mtc

hipError_t mlp_pack(const float* p, void* frag, uint16_t* wb, hipStream_t s) {
  hipLaunchKernelGGL(mtc::mlp::pack_kernel, dim3((mtc::mlp::kWbTotal + 255) / 256), dim3(256), 0, s, p,
                     (bf16x8*)frag, wb);
  return hipGetLastError();
}

hipError_t mlp_train_step(const float* xs, const float* ys, int n_total, uint32_t seed, int n, const float* p,
                          const uint16_t* wb, float* grad, float* loss, hipStream_t s) {
  hipLaunchKernelGGL(mtc::mlp::train_step_kernel, dim3((n + 63) / 64), dim3(64), 0, s, xs, ys, n_total, seed, n, p,
                     wb, grad, loss);
  return hipGetLastError();
}

hipError_t mlp_adam(float* p, float* grad, float* m, float* v, float lr, float b1, float b2, float eps, int step,
                    hipStream_t s) {
  const float c1 = 1.0f / (1.0f - powf(b1, (float)step));
  const float c2 = 1.0f / (1.0f - powf(b2, (float)step));
  hipLaunchKernelGGL(mtc::mlp::adam_kernel, dim3((mtc::mlp::kParams + 255) / 256), dim3(256), 0, s, p, grad, m, v, lr,
                     b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t mlp_zero(float* a, int n, hipStream_t s) {
  hipLaunchKernelGGL(mtc::mlp::zero_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a, n);
  return hipGetLastError();
}

hipError_t mlp_forward(const float* xs, int n, const void* frag, const float* p, float* ys, hipStream_t s) {
  hipLaunchKernelGGL(mtc::mlp::forward_kernel, dim3((n + 63) / 64), dim3(64), 0, s, xs, n, (const bf16x8*)frag, p, ys);
  return hipGetLastError();
}

// mlp_common.h -- the learned-light network: 9 -> 64 -> 64 -> 64 -> 1, SiLU.
//
// It predicts the log transmittance of the cone light march (-dl of
// GetLightEnergy, cloudRayMarch.comp:331-388) from nine inputs: position in
// the noise tile (fract of the relative position), height fraction, local
// density, sun direction, and cos(view, sun).  Hidden layers run on
// gfx950's v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32 accumulation);
// biases, SiLU and the output dot product are FP32 on the VALU.
//
// MFMA 16x16x32 operand layout (wave64), used by every helper here:
//   A (16 x 32):  lane l holds A[l % 16][8 * (l / 16) + e], e = 0..7
//   B (32 x 16):  lane l holds B[8 * (l / 16) + e][l % 16]
//   D (16 x 16):  lane l holds D[4 * (l / 16) + r][l % 16], r = 0..3
//
// Inference runs transposed, H^T = W^T X^T: weights are the A operand
// (rows = output units), the 16 queries of a block are the B operand's
// columns.  A layer's D tiles are then already in B-operand position for
// the next layer once the K order inside each 32-wide chunk is permuted:
// chunk c, slot 8g + e holds hidden unit 32c + 16(e / 4) + 4g + e % 4, the
// units lane group g owns in D tiles 2c and 2c + 1.  The weight fragments
// are packed in that order, so hidden activations never leave registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace mtc {
namespace mlp {

constexpr int kIn = 9;       // inputs
constexpr int kInPad = 32;   // layer-1 K padded to one MFMA k-chunk
constexpr int kHid = 64;     // hidden width
constexpr int kParams = 9025;
// offsets of the FP32 parameter vector: W1[9][64], b1, W2[64][64], b2,
// W3[64][64], b3, w4[64], b4 (W[k][n]: input k, output n)
constexpr int OW1 = 0, OB1 = 576, OW2 = 640, OB2 = 4736, OW3 = 4800, OB3 = 8896, OW4 = 8960, OB4 = 9024;
// packed BF16 weight fragments: layer 1 (1 k-chunk), layers 2 and 3 (2
// k-chunks), 4 n-tiles each, 64 lanes x 8 values per (chunk, tile)
// packed BF16 weight fragments, 64 lanes x 8 values each: layer 1 has one
// per output tile (nt), layers 2 and 3 one per (nt, k-chunk)
constexpr int kFragL1 = 0, kFragL2 = 4 * 64, kFragL3 = 12 * 64, kFrags = 20 * 64;

// FP32 -> BF16, round to nearest even (one v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }
__device__ __forceinline__ __bf16 bfv(uint16_t h) { return __builtin_bit_cast(__bf16, h); }

__device__ __forceinline__ float sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.4426950408889634f));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
  float s = sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

// Hidden unit held in slot e of lane group g's operand for k-chunk c.
__host__ __device__ constexpr int hidden_unit(int c, int g, int e) { return 32 * c + 16 * (e / 4) + 4 * g + e % 4; }
static_assert(OB1 % 4 == 0 && OB2 % 4 == 0 && OB3 % 4 == 0 && OW4 % 4 == 0, "bias rows are read as float4");

// One 16x16 output tile of A[M x K] * B[K x N] from LDS operands addressed
// with arbitrary strides (so transposes cost nothing): A(m, k) at
// a[m * sam + k * sak], B(k, n) at b[k * sbk + n * sbn].  K is a multiple
// of 32.
__device__ __forceinline__ floatx4 tile_strided(const uint16_t* a, int sam, int sak, const uint16_t* b, int sbk,
                                                int sbn, int m0, int n0, int K) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, kb = 8 * (lane >> 4);
  floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int kc = 0; kc < K; kc += 32) {
    bf16x8 av, bv;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      av[e] = bfv(a[(m0 + r) * sam + (kc + kb + e) * sak]);
      bv[e] = bfv(b[(kc + kb + e) * sbk + (n0 + r) * sbn]);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
  }
  return acc;
}

// A zero the compiler cannot see through: offsetting the weight pointers by
// it keeps their loads inside the query loop (next to their MFMAs) instead
// of hoisting all twenty fragments and the biases into registers for the
// whole march.
__device__ __forceinline__ int opaque_zero() {
  int z;
  asm volatile("s_mov_b32 %0, 0" : "=s"(z));
  return z;
}

// D tile initialised with the bias of its output units (C operand)
__device__ __forceinline__ floatx4 bias_tile(const float* __restrict__ b, int nt, int g) {
  const float4 v = *(const float4*)(b + 16 * nt + 4 * g);
  return floatx4{v.x, v.y, v.z, v.w};
}

// SiLU of two D tiles (units 32c .. 32c + 31) into one B operand
__device__ __forceinline__ bf16x8 activate(floatx4 z0, floatx4 z1) {
  bf16x8 h;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    h[r] = (__bf16)silu(z0[r]);
    h[4 + r] = (__bf16)silu(z1[r]);
  }
  return h;
}

// Forward pass of one block of 16 queries.  `x` is the B operand of layer
// 1: lane l holds inputs 8 (l / 16) .. + 7 of query l % 16 (zero beyond the
// ninth).  Every lane returns the output of query l % 16.
__device__ __forceinline__ float infer16(bf16x8 x, const bf16x8* __restrict__ frag_in, const float* __restrict__ p_in) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int oz = opaque_zero();
  const bf16x8* frag = frag_in + oz;
  const float* p = p_in + oz;
  floatx4 z[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
    z[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(frag[(kFragL1 / 64 + nt) * 64 + lane], x, bias_tile(p + OB1, nt, g),
                                                    0, 0, 0);
  bf16x8 h0 = activate(z[0], z[1]), h1 = activate(z[2], z[3]);
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const bf16x8* fl = frag + (layer == 0 ? kFragL2 : kFragL3);
    const float* bl = p + (layer == 0 ? OB2 : OB3);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      z[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fl[(nt * 2) * 64 + lane], h0, bias_tile(bl, nt, g), 0, 0, 0);
      z[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fl[(nt * 2 + 1) * 64 + lane], h1, z[nt], 0, 0, 0);
    }
    if (layer == 0) {
      h0 = activate(z[0], z[1]);
      h1 = activate(z[2], z[3]);
    }
  }
  // output unit: BF16-rounded H3 (as in training) dotted with w4 in FP32
  float y = 0.0f;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const float4 w = *(const float4*)(p + OW4 + 16 * nt + 4 * g);
    y += bf2f(f2bf(silu(z[nt][0]))) * w.x;
    y += bf2f(f2bf(silu(z[nt][1]))) * w.y;
    y += bf2f(f2bf(silu(z[nt][2]))) * w.z;
    y += bf2f(f2bf(silu(z[nt][3]))) * w.w;
  }
  y += __shfl_xor(y, 16, 64);
  y += __shfl_xor(y, 32, 64);
  return y + p[OB4];
}

// Inference of up to 64 queries of one wave.  `x` is the wave's LDS area
// of 64 rows x 16 bf16; row q holds query q's nine inputs, zero padded.
// Only the first `nq` rows are read (ceil(nq / 16) blocks).  Lane q
// returns the output of query q.
__device__ __forceinline__ float infer_wave(const uint16_t* x, const bf16x8* __restrict__ frag,
                                            const float* __restrict__ p, int nq) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  float mine = 0.0f;
  for (int qb = 0; qb * 16 < nq; ++qb) {
    bf16x8 xin;
    if (g < 2) {
      xin = *(const bf16x8*)(x + (qb * 16 + (lane & 15)) * 16 + 8 * g);
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) xin[e] = bfv((uint16_t)0);
    }
    const float y = infer16(xin, frag, p);
    if (qb == g) mine = y;
  }
  return mine;
}

}  // namespace mlp
}  // namespace mtc

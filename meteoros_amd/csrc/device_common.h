// device_common.h -- FP32 device math shared by the cloud-stage kernels.
//
// Written for gfx950 (wave64).  The "exact" helpers keep GLSL source order
// and one rounding per operation (the files including this header are
// compiled with -ffp-contract=off); explicit fma appears only where the
// sampler model of the stage defines it (lerp = fma(w, q - p, p)).
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "runtime_internal.h"

namespace mtc {

struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float length(f3 a) { return __builtin_sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalize(f3 a) {
  float l = length(a);
  return a / l;
}
__device__ __forceinline__ f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ float smoothstepf(float e0, float e1, float x) {
  float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }
// cloudRayMarch.comp:153-169
__device__ __forceinline__ float remap(float v, float a, float b, float c, float d) {
  return c + (((v - a) / (b - a)) * (d - c));
}
__device__ __forceinline__ float remapClamped(float v, float a, float b, float c, float d) {
  float t = c + (((v - a) / (b - a)) * (d - c));
  return clampf(t, c, d);
}
__device__ __forceinline__ float remapCBA(float v, float a, float b, float c, float d) {
  v = clampf(v, a, b);
  float t = c + (((v - a) / (b - a)) * (d - c));
  return clampf(t, c, d);
}

// ---------------------------------------------------------------------------
// hand-written filtering (the stage's sampler model)
//   u = s*N - 0.5, i0 = floor(u), a = u - floor(u), REPEAT = i & (N-1)
//   lerp(p, q, w) = fma(w, q - p, p); x, then y, then z
//   UNORM8: integer codes filtered, result * fl(1/255)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float lerpf(float p, float q, float w) { return __fmaf_rn(w, q - p, p); }

struct c4 {
  float r, g, b, a;
};
__device__ __forceinline__ c4 unpack8(uint32_t t) {
  // each is a single v_cvt_f32_ubyteN
  return c4{(float)(t & 0xffu), (float)((t >> 8) & 0xffu), (float)((t >> 16) & 0xffu), (float)(t >> 24)};
}
__device__ __forceinline__ c4 lerp4(c4 p, c4 q, float w) {
  return c4{lerpf(p.r, q.r, w), lerpf(p.g, q.g, w), lerpf(p.b, q.b, w), lerpf(p.a, q.a, w)};
}
__device__ __forceinline__ c4 scale4(c4 c, float k) { return c4{c.r * k, c.g * k, c.b * k, c.a * k}; }

#define MTC_INV255 ((float)(1.0 / 255.0))

// 3D RGBA8 volume, N = 1 << sh, x fastest.
__device__ __forceinline__ c4 sample3d_u8(const uint32_t* __restrict__ vol, int sh, f3 s) {
  const float N = (float)(1 << sh);
  const int m = (1 << sh) - 1;
  float ux = s.x * N - 0.5f, uy = s.y * N - 0.5f, uz = s.z * N - 0.5f;
  float fx = floorf(ux), fy = floorf(uy), fz = floorf(uz);
  float ax = ux - fx, ay = uy - fy, az = uz - fz;
  int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  int x0 = ix & m, x1 = (ix + 1) & m;
  int y0 = iy & m, y1 = (iy + 1) & m;
  int z0 = iz & m, z1 = (iz + 1) & m;
  int r00 = ((z0 << sh) + y0) << sh, r10 = ((z0 << sh) + y1) << sh;
  int r01 = ((z1 << sh) + y0) << sh, r11 = ((z1 << sh) + y1) << sh;
  uint32_t t000 = vol[r00 + x0], t100 = vol[r00 + x1];
  uint32_t t010 = vol[r10 + x0], t110 = vol[r10 + x1];
  uint32_t t001 = vol[r01 + x0], t101 = vol[r01 + x1];
  uint32_t t011 = vol[r11 + x0], t111 = vol[r11 + x1];
  c4 c00 = lerp4(unpack8(t000), unpack8(t100), ax);
  c4 c10 = lerp4(unpack8(t010), unpack8(t110), ax);
  c4 c01 = lerp4(unpack8(t001), unpack8(t101), ax);
  c4 c11 = lerp4(unpack8(t011), unpack8(t111), ax);
  c4 c0 = lerp4(c00, c10, ay);
  c4 c1 = lerp4(c01, c11, ay);
  return scale4(lerp4(c0, c1, az), MTC_INV255);
}

__device__ __forceinline__ c4 unpack16(uint2 t) {
  return c4{__half2float(__ushort_as_half((unsigned short)(t.x & 0xffffu))),
            __half2float(__ushort_as_half((unsigned short)(t.x >> 16))),
            __half2float(__ushort_as_half((unsigned short)(t.y & 0xffffu))),
            __half2float(__ushort_as_half((unsigned short)(t.y >> 16)))};
}
// 3D RGBA16F volume (4 halves per texel)
__device__ __forceinline__ c4 sample3d_f16(const uint2* __restrict__ vol, int sh, f3 s) {
  const float N = (float)(1 << sh);
  const int m = (1 << sh) - 1;
  float ux = s.x * N - 0.5f, uy = s.y * N - 0.5f, uz = s.z * N - 0.5f;
  float fx = floorf(ux), fy = floorf(uy), fz = floorf(uz);
  float ax = ux - fx, ay = uy - fy, az = uz - fz;
  int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  int x0 = ix & m, x1 = (ix + 1) & m;
  int y0 = iy & m, y1 = (iy + 1) & m;
  int z0 = iz & m, z1 = (iz + 1) & m;
  int r00 = ((z0 << sh) + y0) << sh, r10 = ((z0 << sh) + y1) << sh;
  int r01 = ((z1 << sh) + y0) << sh, r11 = ((z1 << sh) + y1) << sh;
  c4 c00 = lerp4(unpack16(vol[r00 + x0]), unpack16(vol[r00 + x1]), ax);
  c4 c10 = lerp4(unpack16(vol[r10 + x0]), unpack16(vol[r10 + x1]), ax);
  c4 c01 = lerp4(unpack16(vol[r01 + x0]), unpack16(vol[r01 + x1]), ax);
  c4 c11 = lerp4(unpack16(vol[r11 + x0]), unpack16(vol[r11 + x1]), ax);
  c4 c0 = lerp4(c00, c10, ay);
  c4 c1 = lerp4(c01, c11, ay);
  return lerp4(c0, c1, az);
}

// 2D RGBA8 map, N = 1 << sh
__device__ __forceinline__ c4 sample2d_u8(const uint32_t* __restrict__ map, int sh, float sx, float sy) {
  const float N = (float)(1 << sh);
  const int m = (1 << sh) - 1;
  float ux = sx * N - 0.5f, uy = sy * N - 0.5f;
  float fx = floorf(ux), fy = floorf(uy);
  float ax = ux - fx, ay = uy - fy;
  int ix = (int)fx, iy = (int)fy;
  int x0 = ix & m, x1 = (ix + 1) & m;
  int y0 = (iy & m) << sh, y1 = ((iy + 1) & m) << sh;
  c4 c0 = lerp4(unpack8(map[y0 + x0]), unpack8(map[y0 + x1]), ax);
  c4 c1 = lerp4(unpack8(map[y1 + x0]), unpack8(map[y1 + x1]), ax);
  return scale4(lerp4(c0, c1, ay), MTC_INV255);
}

// EncodeFloatRGBA, cloudRayMarch.comp:105-112
__device__ __forceinline__ float4 encode_float_rgba(float v) {
  float e0 = fractf(1.0f * v), e1 = fractf(255.0f * v), e2 = fractf(65025.0f * v), e3 = fractf(16581375.0f * v);
  const float k = MTC_INV255;
  return make_float4(e0 - e1 * k, e1 - e2 * k, e2 - e3 * k, e3 - e3 * 0.0f);
}

// Wave-level sum of a counter, one atomic per wave.
__device__ __forceinline__ void wave_count(unsigned long long* ctr, unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(ctr, v);
}

}  // namespace mtc

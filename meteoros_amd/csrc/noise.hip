// noise.hip -- noise-volume generation (kernel a).
//
// The reference loads Houdini-made TGA slices (Sky.cpp:28-57) and has no
// generator; this kernel produces the same four resources procedurally
// (tileable Perlin-Worley / Worley fBm volumes, curl map, weather map) in
// the layout the sampler reads: x fastest, then y, then z, 4 channels per
// texel (ImageLoadingUtility.cpp:75-139).  Integer hashing plus FP32
// add/mul/sqrt/floor in a fixed order (-ffp-contract=off), so the output is
// bit-identical to the CPU generator the tests hold.  One lane per texel;
// a wave covers 64 consecutive x of one row, so the 27 feature-cell hashes
// of a Worley octave are shared by neighbouring lanes through the L1/LDS
// path of the cache and the stores are fully coalesced.
#include <hip/hip_fp16.h>

#include "device_common.h"

namespace mtc {
namespace noise {

__device__ __forceinline__ uint32_t pcg(uint32_t v) {
  uint32_t state = v * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
__device__ __forceinline__ uint32_t hash3(uint32_t x, uint32_t y, uint32_t z, uint32_t s) {
  return pcg(x + pcg(y + pcg(z + pcg(s))));
}
__device__ __forceinline__ float u01(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }
// wrap for power-of-two and general periods alike
__device__ __forceinline__ int wrapi(int i, int n) { return ((i % n) + n) % n; }

__device__ float worley(float px, float py, float pz, int cells, uint32_t seed) {
  float qx = px * (float)cells, qy = py * (float)cells, qz = pz * (float)cells;
  float cx = floorf(qx), cy = floorf(qy), cz = floorf(qz);
  float fx = qx - cx, fy = qy - cy, fz = qz - cz;
  int ix = (int)cx, iy = (int)cy, iz = (int)cz;
  float best = 16.0f;
  for (int dz = -1; dz <= 1; ++dz) {
    uint32_t hz = (uint32_t)wrapi(iz + dz, cells);
    for (int dy = -1; dy <= 1; ++dy) {
      uint32_t hy = (uint32_t)wrapi(iy + dy, cells);
      for (int dx = -1; dx <= 1; ++dx) {
        uint32_t h = hash3((uint32_t)wrapi(ix + dx, cells), hy, hz, seed);
        float ox = u01(pcg(h)), oy = u01(pcg(h ^ 0x9e3779b9u)), oz = u01(pcg(h ^ 0x7f4a7c15u));
        float ex = ((float)dx + ox) - fx;
        float ey = ((float)dy + oy) - fy;
        float ez = ((float)dz + oz) - fz;
        float d2 = ex * ex + ey * ey + ez * ez;
        best = best < d2 ? best : d2;
      }
    }
  }
  float d = __builtin_sqrtf(best);
  float w = 1.0f - d * 1.1f;
  return w < 0.0f ? 0.0f : (w > 1.0f ? 1.0f : w);
}

__device__ __forceinline__ float worley_fbm_from(float a, float b, float c) { return (a * 0.625f + b * 0.25f) + c * 0.125f; }

__device__ __forceinline__ float grad_dot(uint32_t h, float x, float y, float z) {
  switch (h % 12u) {
    case 0: return x + y;
    case 1: return -x + y;
    case 2: return x - y;
    case 3: return -x - y;
    case 4: return x + z;
    case 5: return -x + z;
    case 6: return x - z;
    case 7: return -x - z;
    case 8: return y + z;
    case 9: return -y + z;
    case 10: return y - z;
    default: return -y - z;
  }
}
__device__ __forceinline__ float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
__device__ __forceinline__ float lerpn(float a, float b, float w) { return a + w * (b - a); }

__device__ float perlin(float px, float py, float pz, int cells, uint32_t seed) {
  float qx = px * (float)cells, qy = py * (float)cells, qz = pz * (float)cells;
  float cx = floorf(qx), cy = floorf(qy), cz = floorf(qz);
  float fx = qx - cx, fy = qy - cy, fz = qz - cz;
  int ix = (int)cx, iy = (int)cy, iz = (int)cz;
  float v[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    int dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
    uint32_t h = hash3((uint32_t)wrapi(ix + dx, cells), (uint32_t)wrapi(iy + dy, cells),
                       (uint32_t)wrapi(iz + dz, cells), seed);
    v[k] = grad_dot(h, fx - (float)dx, fy - (float)dy, fz - (float)dz);
  }
  float ux = fade(fx), uy = fade(fy), uz = fade(fz);
  float x00 = lerpn(v[0], v[1], ux), x10 = lerpn(v[2], v[3], ux);
  float x01 = lerpn(v[4], v[5], ux), x11 = lerpn(v[6], v[7], ux);
  float y0 = lerpn(x00, x10, uy), y1 = lerpn(x01, x11, uy);
  return lerpn(y0, y1, uz);
}
__device__ float perlin_fbm(float px, float py, float pz, int cells, int octaves, uint32_t seed) {
  float sum = 0.0f, amp = 1.0f, norm = 0.0f;
  for (int o = 0; o < octaves; ++o) {
    sum = sum + amp * perlin(px, py, pz, cells << o, seed + 101u * (uint32_t)o);
    norm = norm + amp;
    amp = amp * 0.5f;
  }
  return sum / norm;
}

__device__ __forceinline__ float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }
__device__ __forceinline__ uint32_t to_u8(float v) { return (uint32_t)(int)floorf(clamp01(v) * 255.0f + 0.5f); }

struct Out4 {
  float v[4];
};

__device__ __forceinline__ void store_texel(void* out, size_t idx, int fmt, const Out4& t) {
  if (fmt == MTC_FMT_FP16) {
    uint32_t lo = (uint32_t)__half_as_ushort(__float2half_rn(clamp01(t.v[0]))) |
                  ((uint32_t)__half_as_ushort(__float2half_rn(clamp01(t.v[1]))) << 16);
    uint32_t hi = (uint32_t)__half_as_ushort(__float2half_rn(clamp01(t.v[2]))) |
                  ((uint32_t)__half_as_ushort(__float2half_rn(clamp01(t.v[3]))) << 16);
    ((uint2*)out)[idx] = make_uint2(lo, hi);
  } else {
    ((uint32_t*)out)[idx] = to_u8(t.v[0]) | (to_u8(t.v[1]) << 8) | (to_u8(t.v[2]) << 16) | (to_u8(t.v[3]) << 24);
  }
}

// kind 0: low-frequency volume
__global__ __launch_bounds__(256) void lf_kernel(void* out, int N, int fmt, uint32_t seed) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t total = (size_t)N * N * N;
  if (idx >= total) return;
  const int x = (int)(idx % N), y = (int)((idx / N) % N), z = (int)(idx / ((size_t)N * N));
  float px = ((float)x + 0.5f) / (float)N, py = ((float)y + 0.5f) / (float)N, pz = ((float)z + 0.5f) / (float)N;
  // octaves 4..64; worley_fbm(f) = w(f), w(2f), w(4f) with seed + cells
  float w4 = worley(px, py, pz, 4, seed + 4u);
  float w8 = worley(px, py, pz, 8, seed + 8u);
  float w16 = worley(px, py, pz, 16, seed + 16u);
  float w32 = worley(px, py, pz, 32, seed + 32u);
  float w64 = worley(px, py, pz, 64, seed + 64u);
  float f4 = worley_fbm_from(w4, w8, w16);
  float f8 = worley_fbm_from(w8, w16, w32);
  float f16 = worley_fbm_from(w16, w32, w64);
  float p = perlin_fbm(px, py, pz, 4, 4, seed ^ 0x51ed270bu);
  float p01 = clamp01(p * 0.9f + 0.5f);
  float pw = f4 + p01 * (1.0f - f4);
  Out4 t;
  t.v[0] = clamp01((pw - 0.7128f) * 0.958f + 0.504f);
  t.v[1] = clamp01((f4 - 0.426f) * 0.77f + 0.687f);
  t.v[2] = clamp01((f8 - 0.426f) * 0.77f + 0.687f);
  t.v[3] = clamp01((f16 - 0.426f) * 0.77f + 0.687f);
  store_texel(out, idx, fmt, t);
}

// kind 1: high-frequency volume
__global__ __launch_bounds__(256) void hf_kernel(void* out, int N, int fmt, uint32_t seed) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t total = (size_t)N * N * N;
  if (idx >= total) return;
  const int x = (int)(idx % N), y = (int)((idx / N) % N), z = (int)(idx / ((size_t)N * N));
  float px = ((float)x + 0.5f) / (float)N, py = ((float)y + 0.5f) / (float)N, pz = ((float)z + 0.5f) / (float)N;
  uint32_t s = seed ^ 0x2545f491u;
  float w2 = worley(px, py, pz, 2, s + 2u);
  float w4 = worley(px, py, pz, 4, s + 4u);
  float w8 = worley(px, py, pz, 8, s + 8u);
  float w16 = worley(px, py, pz, 16, s + 16u);
  float w32 = worley(px, py, pz, 32, s + 32u);
  Out4 t;
  t.v[0] = clamp01((worley_fbm_from(w2, w4, w8) - 0.428f) * 0.77f + 0.687f);
  t.v[1] = clamp01((worley_fbm_from(w4, w8, w16) - 0.428f) * 0.77f + 0.687f);
  t.v[2] = clamp01((worley_fbm_from(w8, w16, w32) - 0.428f) * 0.77f + 0.687f);
  t.v[3] = 0.0f;
  store_texel(out, idx, fmt, t);
}

__device__ __forceinline__ float curl_potential(int x, int y, int N, uint32_t seed) {
  int xx = wrapi(x, N), yy = wrapi(y, N);
  float px = ((float)xx + 0.5f) / (float)N, py = ((float)yy + 0.5f) / (float)N;
  return perlin_fbm(px, py, 0.5f, 4, 3, seed ^ 0x68e31da4u);
}

// kind 2: curl map (RGBA8)
__global__ __launch_bounds__(256) void curl_kernel(uint32_t* out, int N, uint32_t seed) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= N * N) return;
  const int x = idx % N, y = idx / N;
  float dpdy = (curl_potential(x, y + 1, N, seed) - curl_potential(x, y - 1, N, seed)) * (float)N * 0.5f;
  float dpdx = (curl_potential(x + 1, y, N, seed) - curl_potential(x - 1, y, N, seed)) * (float)N * 0.5f;
  float cx = dpdy * 0.05f, cy = -dpdx * 0.05f;
  Out4 t;
  t.v[0] = clamp01(cx * 0.5f + 0.5f);
  t.v[1] = clamp01(cy * 0.5f + 0.5f);
  t.v[2] = clamp01(__builtin_sqrtf(cx * cx + cy * cy));
  t.v[3] = 1.0f;
  store_texel(out, (size_t)idx, MTC_FMT_UNORM8, t);
}

// kind 3: weather map (RGBA8)
__global__ __launch_bounds__(256) void weather_kernel(uint32_t* out, int N, uint32_t seed) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= N * N) return;
  const int x = idx % N, y = idx / N;
  const int L = 16;
  float qx = (((float)x + 0.5f) / (float)N) * (float)L, qy = (((float)y + 0.5f) / (float)N) * (float)L;
  float cx = floorf(qx), cy = floorf(qy);
  float fx = qx - cx, fy = qy - cy;
  int ix = (int)cx, iy = (int)cy;
  float v00 = u01(hash3((uint32_t)wrapi(ix, L), (uint32_t)wrapi(iy, L), 0u, seed));
  float v10 = u01(hash3((uint32_t)wrapi(ix + 1, L), (uint32_t)wrapi(iy, L), 0u, seed));
  float v01 = u01(hash3((uint32_t)wrapi(ix, L), (uint32_t)wrapi(iy + 1, L), 0u, seed));
  float v11 = u01(hash3((uint32_t)wrapi(ix + 1, L), (uint32_t)wrapi(iy + 1, L), 0u, seed));
  float sx = fx * fx * (3.0f - 2.0f * fx), sy = fy * fy * (3.0f - 2.0f * fy);
  float cov = lerpn(lerpn(v00, v10, sx), lerpn(v01, v11, sx), sy);
  int cell = N / 8 > 0 ? N / 8 : 1;
  uint32_t th = hash3((uint32_t)(x / cell), (uint32_t)(y / cell), 1u, seed);
  Out4 t;
  t.v[0] = cov;
  t.v[1] = (float)(th % 3u) * 0.5f;
  t.v[2] = 0.0f;
  t.v[3] = 1.0f;
  store_texel(out, (size_t)idx, MTC_FMT_UNORM8, t);
}

}  // namespace noise
}  // namespace mtc

hipError_t launch_noise(int kind, void* out, int N, int fmt, uint32_t seed, hipStream_t stream) {
  size_t total = (kind <= 1) ? (size_t)N * N * N : (size_t)N * N;
  unsigned blocks = (unsigned)((total + 255) / 256);
  switch (kind) {
    case 0: hipLaunchKernelGGL(mtc::noise::lf_kernel, dim3(blocks), dim3(256), 0, stream, out, N, fmt, seed); break;
    case 1: hipLaunchKernelGGL(mtc::noise::hf_kernel, dim3(blocks), dim3(256), 0, stream, out, N, fmt, seed); break;
    case 2: hipLaunchKernelGGL(mtc::noise::curl_kernel, dim3(blocks), dim3(256), 0, stream, (uint32_t*)out, N, seed); break;
    default: hipLaunchKernelGGL(mtc::noise::weather_kernel, dim3(blocks), dim3(256), 0, stream, (uint32_t*)out, N, seed); break;
  }
  return hipGetLastError();
}

// noise_oracle.hpp -- CPU specification of the noise-volume generator.
//
// TEST INFRASTRUCTURE ONLY (see cloud_oracle.hpp).  The reference ships its
// noise as pre-authored TGA slices (Sky.cpp:28-57,
// ImageLoadingUtility.cpp:75-139; README.md:302 says they came from
// Houdini) and has no generator, so this file *is* the specification of the
// HIP generator kernel: the kernel must reproduce it bit for bit.  All
// arithmetic is integer hashing plus FP32 add/mul/sqrt/floor, compiled with
// -ffp-contract=off, so every value is reproducible on any IEEE machine.
//
// Volumes (x fastest, then y, then z; 4 channels per texel):
//   LF  R = Perlin-Worley: periodic Perlin fBm (4 octaves from freq 4)
//           dilated by the inverted-Worley fBm at freq 4
//       G,B,A = inverted-Worley fBm at freq 4, 8, 16 (3 octaves each,
//           weights 0.625/0.25/0.125, the same weights the shader uses to
//           recombine them at cloudRayMarch.comp:505-507)
//   HF  R,G,B = inverted-Worley fBm at freq 2, 4, 8; A = 0
//   curl  R,G = curl of a periodic 2D Perlin-fBm potential, B = |curl|, A = 255
//   weather R = coverage, a smoothed uniform[0,1] field on a 16^2 lattice,
//           G = cloud type in {0, 0.5, 1} per 64^2-texel cell (scaled to the
//           map size), B = 0, A = 255
// Statistics are tuned to the shipped volumes (tests/golden/texture_stats.json).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc_noise {

static inline uint32_t pcg(uint32_t v) {
  uint32_t state = v * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
static inline uint32_t hash3(uint32_t x, uint32_t y, uint32_t z, uint32_t s) {
  return pcg(x + pcg(y + pcg(z + pcg(s))));
}
// 24 random bits -> [0, 1) exactly representable
static inline float u01(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

static inline int wrapi(int i, int n) { return ((i % n) + n) % n; }

// Inverted Worley F1 at point p in [0,1)^3, `cells` feature cells per tile.
static inline float worley(float px, float py, float pz, int cells, uint32_t seed) {
  float qx = px * (float)cells, qy = py * (float)cells, qz = pz * (float)cells;
  float cx = std::floor(qx), cy = std::floor(qy), cz = std::floor(qz);
  float fx = qx - cx, fy = qy - cy, fz = qz - cz;
  int ix = (int)cx, iy = (int)cy, iz = (int)cz;
  float best = 16.0f;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        uint32_t h = hash3((uint32_t)wrapi(ix + dx, cells), (uint32_t)wrapi(iy + dy, cells),
                           (uint32_t)wrapi(iz + dz, cells), seed);
        float ox = u01(pcg(h)), oy = u01(pcg(h ^ 0x9e3779b9u)), oz = u01(pcg(h ^ 0x7f4a7c15u));
        float ex = ((float)dx + ox) - fx;
        float ey = ((float)dy + oy) - fy;
        float ez = ((float)dz + oz) - fz;
        float d2 = ex * ex + ey * ey + ez * ez;
        best = best < d2 ? best : d2;
      }
  float d = std::sqrt(best);
  float w = 1.0f - d * 1.1f;
  return w < 0.0f ? 0.0f : (w > 1.0f ? 1.0f : w);
}
static inline float worley_fbm(float px, float py, float pz, int cells, uint32_t seed) {
  float a = worley(px, py, pz, cells, seed + (uint32_t)cells);
  float b = worley(px, py, pz, cells * 2, seed + (uint32_t)(cells * 2));
  float c = worley(px, py, pz, cells * 4, seed + (uint32_t)(cells * 4));
  return (a * 0.625f + b * 0.25f) + c * 0.125f;
}

// 12 cube-edge gradient directions
static inline float grad_dot(uint32_t h, float x, float y, float z) {
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
static inline float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
static inline float lerpf(float a, float b, float w) { return a + w * (b - a); }

// Periodic Perlin gradient noise, period `cells`, roughly in [-1, 1].
static inline float perlin(float px, float py, float pz, int cells, uint32_t seed) {
  float qx = px * (float)cells, qy = py * (float)cells, qz = pz * (float)cells;
  float cx = std::floor(qx), cy = std::floor(qy), cz = std::floor(qz);
  float fx = qx - cx, fy = qy - cy, fz = qz - cz;
  int ix = (int)cx, iy = (int)cy, iz = (int)cz;
  float v[8];
  for (int k = 0; k < 8; ++k) {
    int dx = k & 1, dy = (k >> 1) & 1, dz = (k >> 2) & 1;
    uint32_t h = hash3((uint32_t)wrapi(ix + dx, cells), (uint32_t)wrapi(iy + dy, cells),
                       (uint32_t)wrapi(iz + dz, cells), seed);
    v[k] = grad_dot(h, fx - (float)dx, fy - (float)dy, fz - (float)dz);
  }
  float ux = fade(fx), uy = fade(fy), uz = fade(fz);
  float x00 = lerpf(v[0], v[1], ux), x10 = lerpf(v[2], v[3], ux);
  float x01 = lerpf(v[4], v[5], ux), x11 = lerpf(v[6], v[7], ux);
  float y0 = lerpf(x00, x10, uy), y1 = lerpf(x01, x11, uy);
  return lerpf(y0, y1, uz);
}
static inline float perlin_fbm(float px, float py, float pz, int cells, int octaves, uint32_t seed) {
  float sum = 0.0f, amp = 1.0f, norm = 0.0f;
  for (int o = 0; o < octaves; ++o) {
    sum = sum + amp * perlin(px, py, pz, cells << o, seed + 101u * (uint32_t)o);
    norm = norm + amp;
    amp = amp * 0.5f;
  }
  return sum / norm;
}

static inline float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }
static inline uint8_t to_unorm8(float v) { return (uint8_t)(int)std::floor(clamp01(v) * 255.0f + 0.5f); }

// IEEE binary16 round-to-nearest-even of a finite float in [0, 1].
static inline uint16_t to_half(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  int e = (int)((x >> 23) & 0xff) - 127 + 15;
  uint32_t m = x & 0x7fffffu;
  if (((x >> 23) & 0xff) == 0) return (uint16_t)sign;  // zero / float denormal -> 0
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    m |= 0x800000u;
    int shift = 14 - e;
    uint32_t hm = m >> shift;
    uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (hm & 1u))) hm++;
    return (uint16_t)(sign | hm);
  }
  uint32_t hm = m >> 13, rem = m & 0x1fffu;
  uint32_t h = sign | ((uint32_t)e << 10) | hm;
  if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) h++;
  return (uint16_t)h;
}

// --- per-texel channel values (the specification) -------------------------

// Low-frequency volume texel (x, y, z) of an N^3 volume -> 4 floats in [0, 1]
static inline void lf_texel(int x, int y, int z, int N, uint32_t seed, float out[4]) {
  float px = ((float)x + 0.5f) / (float)N, py = ((float)y + 0.5f) / (float)N, pz = ((float)z + 0.5f) / (float)N;
  float w4 = worley_fbm(px, py, pz, 4, seed);
  float w8 = worley_fbm(px, py, pz, 8, seed);
  float w16 = worley_fbm(px, py, pz, 16, seed);
  float p = perlin_fbm(px, py, pz, 4, 4, seed ^ 0x51ed270bu);
  float p01 = clamp01(p * 0.9f + 0.5f);
  // Perlin-Worley: remap(perlin, 0, 1, worley, 1) then re-centred
  float pw = w4 + p01 * (1.0f - w4);
  float r = (pw - 0.7128f) * 0.958f + 0.504f;
  out[0] = clamp01(r);
  out[1] = clamp01((w4 - 0.426f) * 0.77f + 0.687f);
  out[2] = clamp01((w8 - 0.426f) * 0.77f + 0.687f);
  out[3] = clamp01((w16 - 0.426f) * 0.77f + 0.687f);
}
// High-frequency volume texel
static inline void hf_texel(int x, int y, int z, int N, uint32_t seed, float out[4]) {
  float px = ((float)x + 0.5f) / (float)N, py = ((float)y + 0.5f) / (float)N, pz = ((float)z + 0.5f) / (float)N;
  uint32_t s = seed ^ 0x2545f491u;
  out[0] = clamp01((worley_fbm(px, py, pz, 2, s) - 0.428f) * 0.77f + 0.687f);
  out[1] = clamp01((worley_fbm(px, py, pz, 4, s) - 0.428f) * 0.77f + 0.687f);
  out[2] = clamp01((worley_fbm(px, py, pz, 8, s) - 0.428f) * 0.77f + 0.687f);
  out[3] = 0.0f;
}
static inline float curl_potential(int x, int y, int N, uint32_t seed) {
  int xx = wrapi(x, N), yy = wrapi(y, N);
  float px = ((float)xx + 0.5f) / (float)N, py = ((float)yy + 0.5f) / (float)N;
  return perlin_fbm(px, py, 0.5f, 4, 3, seed ^ 0x68e31da4u);
}
// Curl-noise texel (x, y) of an N^2 map
static inline void curl_texel(int x, int y, int N, uint32_t seed, float out[4]) {
  float dpdy = (curl_potential(x, y + 1, N, seed) - curl_potential(x, y - 1, N, seed)) * (float)N * 0.5f;
  float dpdx = (curl_potential(x + 1, y, N, seed) - curl_potential(x - 1, y, N, seed)) * (float)N * 0.5f;
  float cx = dpdy * 0.05f, cy = -dpdx * 0.05f;
  out[0] = clamp01(cx * 0.5f + 0.5f);
  out[1] = clamp01(cy * 0.5f + 0.5f);
  out[2] = clamp01(std::sqrt(cx * cx + cy * cy));
  out[3] = 1.0f;
}
// Weather-map texel (x, y) of an N^2 map
static inline void weather_texel(int x, int y, int N, uint32_t seed, float out[4]) {
  const int L = 16;  // coverage lattice
  float qx = (((float)x + 0.5f) / (float)N) * (float)L, qy = (((float)y + 0.5f) / (float)N) * (float)L;
  float cx = std::floor(qx), cy = std::floor(qy);
  float fx = qx - cx, fy = qy - cy;
  int ix = (int)cx, iy = (int)cy;
  float v00 = u01(hash3((uint32_t)wrapi(ix, L), (uint32_t)wrapi(iy, L), 0u, seed));
  float v10 = u01(hash3((uint32_t)wrapi(ix + 1, L), (uint32_t)wrapi(iy, L), 0u, seed));
  float v01 = u01(hash3((uint32_t)wrapi(ix, L), (uint32_t)wrapi(iy + 1, L), 0u, seed));
  float v11 = u01(hash3((uint32_t)wrapi(ix + 1, L), (uint32_t)wrapi(iy + 1, L), 0u, seed));
  float sx = fx * fx * (3.0f - 2.0f * fx), sy = fy * fy * (3.0f - 2.0f * fy);
  float cov = lerpf(lerpf(v00, v10, sx), lerpf(v01, v11, sx), sy);
  int cell = N / 8 > 0 ? N / 8 : 1;  // 64 texels at N = 512
  uint32_t th = hash3((uint32_t)(x / cell), (uint32_t)(y / cell), 1u, seed);
  float type = (float)(th % 3u) * 0.5f;
  out[0] = cov;
  out[1] = type;
  out[2] = 0.0f;
  out[3] = 1.0f;
}

}  // namespace orc_noise

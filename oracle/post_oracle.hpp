// post_oracle.hpp -- CPU restatement of the two post-process passes that
// consume the cloud stage's output.
//
// TEST INFRASTRUCTURE ONLY (the checker of meteoros_amd/csrc/post.hip).
//
//   god_rays   postProcess_GodRays.frag:56-150
//   tone_map   postProcess_ToneMap.frag:23-84
//
// One template, two twins as in cloud_oracle.hpp: T=float (compiled with
// -ffp-contract=off) follows the shader's operation order; T=double is the
// accuracy twin.  The fragment's uv is its pixel centre (p + 0.5) / dim and
// it writes its own pixel (the shaders' round(dim * uv) lands on a .5 tie
// there, an implementation-defined choice).  Samplers: linear filtering,
// CLAMP_TO_BORDER with opaque black (0, 0, 0, 1) (Texture2D.cpp:75,
// Image.cpp:305-345), u = uv * dim - 0.5.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../include/meteoros_cloud.h"

namespace orc_post {

template <class T> struct P4 { T x, y, z, w; };

template <class T> static inline P4<T> texel(const float* img, int W, int H, int x, int y) {
  if (x < 0 || y < 0 || x >= W || y >= H) return P4<T>{T(0), T(0), T(0), T(1)};
  const float* p = img + ((size_t)y * W + x) * 4;
  return P4<T>{(T)p[0], (T)p[1], (T)p[2], (T)p[3]};
}
template <class T> static inline T lerp(T p, T q, T w) { return std::fma(w, q - p, p); }
template <class T> static inline P4<T> lerp4(P4<T> a, P4<T> b, T w) {
  return P4<T>{lerp(a.x, b.x, w), lerp(a.y, b.y, w), lerp(a.z, b.z, w), lerp(a.w, b.w, w)};
}
// texture(sampler2D, uv) of an RGBA32F image
template <class T> static inline P4<T> bilinear(const float* img, int W, int H, T u, T v) {
  T x = u * (T)W - T(0.5), y = v * (T)H - T(0.5);
  T fx = std::floor(x), fy = std::floor(y);
  T wx = x - fx, wy = y - fy;
  int ix = (int)fx, iy = (int)fy;
  P4<T> a = lerp4(texel<T>(img, W, H, ix, iy), texel<T>(img, W, H, ix + 1, iy), wx);
  P4<T> b = lerp4(texel<T>(img, W, H, ix, iy + 1), texel<T>(img, W, H, ix + 1, iy + 1), wx);
  return lerp4(a, b, wy);
}

// ---------------------------------------------------------------------------
// god rays
// ---------------------------------------------------------------------------
template <class T> struct GodRayFrame {
  T sun_ss[2];
  T blend;
  T light[3];
};

// postProcess_GodRays.frag:72-98: static sun at (0, 1, 0), camera.eye as
// stored, camForward = -normalize(view[0][2], view[1][2], view[2][2]),
// sunPos_ss = clamp(((proj * view) * (0, 1, 0, 1)).xy + 1) / 2, 0, 1).
template <class T> static GodRayFrame<T> god_ray_frame(const mtc_camera& cam, const mtc_sun& sun) {
  GodRayFrame<T> f;
  T ts[3] = {T(0) - (T)cam.eye[0], T(1) - (T)cam.eye[1], T(0) - (T)cam.eye[2]};
  T l = std::sqrt(ts[0] * ts[0] + ts[1] * ts[1] + ts[2] * ts[2]);
  for (T& v : ts) v = v / l;
  T fw[3] = {(T)cam.view[2], (T)cam.view[6], (T)cam.view[10]};
  l = std::sqrt(fw[0] * fw[0] + fw[1] * fw[1] + fw[2] * fw[2]);
  for (T& v : fw) v = -(v / l);
  f.blend = (ts[0] * fw[0] + ts[1] * fw[1]) + ts[2] * fw[2];
  T pv[16];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      T acc = (T)cam.proj[r] * (T)cam.view[c * 4];
      for (int k = 1; k < 4; ++k) acc = acc + (T)cam.proj[k * 4 + r] * (T)cam.view[c * 4 + k];
      pv[c * 4 + r] = acc;
    }
  const T p[4] = {T(0), T(1), T(0), T(1)};
  for (int r = 0; r < 2; ++r) {
    T acc = pv[r] * p[0];
    for (int c = 1; c < 4; ++c) acc = acc + pv[c * 4 + r] * p[c];
    T ss = (acc + T(1)) / T(2);
    f.sun_ss[r] = std::fmin(std::fmax(ss, T(0)), T(1));
  }
  for (int k = 0; k < 3; ++k) f.light[k] = (T)sun.light_color[k];
  return f;
}

// dot(x, bitDec) of extract32fFromRGBA8f (:39-43)
template <class T> static inline T decode(P4<T> c) {
  const T d1 = T(1) / T(255), d2 = T(1) / T(65025), d3 = T(1) / T(16581375);
  return ((c.x * T(1) + c.y * d1) + c.z * d2) + c.w * d3;
}

// One pixel of main (:66-150); the first branch (:105-127) is dead because
// sunPos_ss is clamped into [0, 1] before the frustum test.
template <class T>
static void god_ray_pixel(const GodRayFrame<T>& f, const float* rgba, const float* mask, int W, int H, int px, int py,
                          float out[4]) {
  const float* src = rgba + ((size_t)py * W + px) * 4;
  if (f.blend < T(0)) {
    std::memcpy(out, src, 4 * sizeof(float));
    return;
  }
  T u = ((T)px + T(0.5)) / (T)W, v = ((T)py + T(0.5)) / (T)H;
  const T du = ((u - f.sun_ss[0]) / T(100)) * T(1), dv = ((v - f.sun_ss[1]) / T(100)) * T(1);
  T acc[3] = {T(0), T(0), T(0)};
  T decay = T(1);
  for (int i = 0; i < 100; ++i) {
    const T a = decode(bilinear<T>(mask, W, H, u, v));
    for (int k = 0; k < 3; ++k) acc[k] = acc[k] + (f.light[k] * a) * (decay * (T)0.001f);
    decay = decay * T(1);
    u = u - du;
    v = v - dv;
  }
  for (int k = 0; k < 3; ++k) out[k] = (float)((T)src[k] + (acc[k] * T(1)) * f.blend);
  out[3] = (float)((T)src[3] + T(1) * f.blend);
}

// ---------------------------------------------------------------------------
// tone map
// ---------------------------------------------------------------------------
template <class T> static inline T uncharted2(T x) {
  const T A = (T)0.15f, B = (T)0.50f, C = (T)0.10f, D = (T)0.20f, E = (T)0.02f, F = (T)0.30f;
  return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}

// WangHashNoise (:52-66) before the float conversion
static inline uint32_t wang_hash(uint32_t u, uint32_t v, uint32_t s) {
  uint32_t seed = (u * 1664525u + v) + s;
  seed = (seed ^ 61u) ^ (seed >> 16u);
  seed *= 9u;
  seed = seed ^ (seed >> 4u);
  seed *= 0x27d4eb2du;
  seed = seed ^ (seed >> 15u);
  return seed;
}

// main (:68-84) for one pixel: RGB before quantisation
template <class T>
static void tone_map_pixel(const float* rgba, int W, int H, int px, int py, uint32_t time_s, T out[3]) {
  T u = ((T)px + T(0.5)) / (T)W, v = ((T)py + T(0.5)) / (T)H;
  P4<T> c = bilinear<T>(rgba, W, H, u, v);
  const T in[3] = {c.x, c.y, c.z};
  const T whitemap = T(1) / uncharted2<T>((T)100.0f);
  T noise = (T)(float)wang_hash((uint32_t)px, (uint32_t)py, time_s) * (T)(1.0 / 4294967296.0);
  noise = noise * (T)0.01f;
  for (int k = 0; k < 3; ++k) {
    T col = uncharted2<T>((T)2.5f * in[k]);
    col = col * whitemap;
    out[k] = std::pow(col, (T)(1.0f / 2.2f)) + noise;
  }
}

}  // namespace orc_post

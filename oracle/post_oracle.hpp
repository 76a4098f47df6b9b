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

// ---------------------------------------------------------------------------
// TXAA (postProcess_TXAA.frag:58-270)
// ---------------------------------------------------------------------------
template <class T> struct V3 { T x, y, z; };
template <class T> static inline V3<T> v3n(T x, T y, T z) {
  T l = std::sqrt(x * x + y * y + z * z);
  return V3<T>{x / l, y / l, z / l};
}

// imageLoad of an RGBA8 image as UNORM; 0 outside
template <class T> static inline P4<T> load8(const uint8_t* img, int W, int H, int x, int y) {
  if (x < 0 || y < 0 || x >= W || y >= H) return P4<T>{T(0), T(0), T(0), T(0)};
  const uint8_t* p = img + ((size_t)y * W + x) * 4;
  const T k = (T)(float)(1.0 / 255.0);
  return P4<T>{(T)p[0] * k, (T)p[1] * k, (T)p[2] * k, (T)p[3] * k};
}
// texture() of an RGBA8 image: codes filtered, then scaled; border (0, 0, 0, 1)
template <class T> static inline P4<T> sample8(const uint8_t* img, int W, int H, T u, T v) {
  T x = u * (T)W - T(0.5), y = v * (T)H - T(0.5);
  T fx = std::floor(x), fy = std::floor(y);
  T wx = x - fx, wy = y - fy;
  int ix = (int)std::fmin(std::fmax(fx, T(-2)), (T)W), iy = (int)std::fmin(std::fmax(fy, T(-2)), (T)H);
  P4<T> t[4];
  for (int k = 0; k < 4; ++k) {
    int xx = ix + (k & 1), yy = iy + (k >> 1);
    if (xx < 0 || yy < 0 || xx >= W || yy >= H) {
      t[k] = P4<T>{T(0), T(0), T(0), T(255)};
    } else {
      const uint8_t* p = img + ((size_t)yy * W + xx) * 4;
      t[k] = P4<T>{(T)p[0], (T)p[1], (T)p[2], (T)p[3]};
    }
  }
  P4<T> r = lerp4(lerp4(t[0], t[1], wx), lerp4(t[2], t[3], wx), wy);
  const T k = (T)(float)(1.0 / 255.0);
  return P4<T>{r.x * k, r.y * k, r.z * k, r.w * k};
}
template <class T> static inline P4<T> pmin(P4<T> a, P4<T> b) {
  return P4<T>{std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z), std::fmin(a.w, b.w)};
}
template <class T> static inline P4<T> pmax(P4<T> a, P4<T> b) {
  return P4<T>{std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z), std::fmax(a.w, b.w)};
}
template <class T> static inline P4<T> padd(P4<T> a, P4<T> b) { return P4<T>{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
template <class T> static inline P4<T> pmul(P4<T> a, T s) { return P4<T>{a.x * s, a.y * s, a.z * s, a.w * s}; }
template <class T> static inline P4<T> pdiv(P4<T> a, T s) { return P4<T>{a.x / s, a.y / s, a.z / s, a.w / s}; }

// main (:222-270) for one pixel; returns the blended colour before the
// RGBA8 store.  Also reports the reprojected uv.
template <class T>
static P4<T> txaa_pixel(const mtc_camera& cam, const mtc_camera& old, const mtc_time& tm, const uint8_t* cur,
                        const uint8_t* prev, int W, int H, int px, int py, T* old_uv) {
  const T u = ((T)px + T(0.5)) / (T)W, v = ((T)py + T(0.5)) / (T)H;
  const T ex = -(T)cam.eye[0], ey = -(T)cam.eye[1], ez = -(T)cam.eye[2];
  const float* m = cam.view;
  V3<T> right = v3n<T>(m[0], m[4], m[8]), up = v3n<T>(m[1], m[5], m[9]), lk = v3n<T>(m[2], m[6], m[10]);
  // getJitterOffset(frameCountMod16, dim) (:58-76)
  const int idx = tm.frame_count_mod16 / 2;
  const T jx = (T)(idx < 4 ? tm.halton1[idx] : tm.halton3[idx - 4]) / (T)W;
  const T jy = (T)(idx < 4 ? tm.halton2[idx] : tm.halton4[idx - 4]) / (T)H;
  T nx = u * T(2) - T(1) + jx, ny = v * T(2) - T(1) + jy;
  const T tfx = (T)cam.tan_fov_by2[0], tfy = (T)cam.tan_fov_by2[1];
  const T cxs = nx * tfx, cys = ny * tfy;
  const T p0 = ((ex - lk.x) + cxs * right.x) + cys * up.x, p1 = ((ey - lk.y) + cxs * right.y) + cys * up.y,
          p2 = ((ez - lk.z) + cxs * right.z) + cys * up.z;
  V3<T> d = v3n<T>(p0 - ex, p1 - ey, p2 - ez);
  // raySphereIntersection with EARTH_RADIUS + 7500, centre (eye.x, -EARTH_RADIUS, eye.z) (:119-165)
  const T R = (T)6371000.0f + (T)7500.0f;
  const T c0 = ex, c1 = -(T)6371000.0f, c2 = ez;
  const T r0 = (ex - c0) / R, r1 = (ey - c1) / R, r2 = (ez - c2) / R;
  const T A = (d.x * d.x + d.y * d.y) + d.z * d.z, B = T(2) * ((d.x * r0 + d.y * r1) + d.z * r2);
  const T C = ((r0 * r0 + r1 * r1) + r2 * r2) - T(1);
  const T disc = B * B - T(4) * A * C;
  T P[3] = {T(0), T(0), T(0)};
  if (!(disc < T(0))) {
    T t = (-B - std::sqrt(disc)) / (T(2) * A);
    if (t < T(0)) t = (-B + std::sqrt(disc)) / (T(2) * A);
    if (t >= T(0)) {
      P[0] = (r0 + t * d.x) * R + c0;
      P[1] = (r1 + t * d.y) * R + c1;
      P[2] = (r2 + t * d.z) * R + c2;
    }
  }
  const float* o = old.view;
  T q[3];
  for (int r = 0; r < 3; ++r) q[r] = (((T)o[r] * P[0] + (T)o[4 + r] * P[1]) + (T)o[8 + r] * P[2]) + (T)o[12 + r];
  V3<T> od = v3n<T>(q[0], q[1], q[2]);
  const T nz = -od.z;
  const T ou = ((od.x / nz) / tfx) * T(0.5) + T(0.5), ov = ((od.y / nz) / tfy) * T(0.5) + T(0.5);
  if (old_uv) {
    old_uv[0] = ou;
    old_uv[1] = ov;
  }
  // neighbourHoodClamping (:191-220)
  P4<T> n[9];
  for (int k = 0; k < 9; ++k) n[k] = load8<T>(cur, W, H, px + (k % 3) - 1, py + (k / 3) - 1);
  P4<T> cmin = n[8], cmax = n[8];
  for (int k = 7; k >= 0; --k) {
    cmin = pmin(n[k], cmin);
    cmax = pmax(n[k], cmax);
  }
  P4<T> cavg = n[0];
  for (int k = 1; k < 9; ++k) cavg = padd(cavg, n[k]);
  cavg = pdiv(cavg, T(9));
  P4<T> cmin5 = pmin(n[1], pmin(n[3], pmin(n[4], pmin(n[5], n[7]))));
  P4<T> cmax5 = pmax(n[1], pmax(n[3], pmax(n[4], pmax(n[5], n[7]))));
  P4<T> cavg5 = pdiv(padd(padd(padd(padd(n[1], n[3]), n[4]), n[5]), n[7]), T(5));
  cmin = pmul(padd(cmin, cmin5), T(0.5));
  cmax = pmul(padd(cmax, cmax5), T(0.5));
  cavg = pmul(padd(cavg, cavg5), T(0.5));
  const P4<T> curr = n[4];
  P4<T> pr = sample8<T>(prev, W, H, ou, ov);
  // clip_aabb (:170-189)
  const T pw = std::fmin(std::fmax(cavg.w, cmin.w), cmax.w);
  const T eps = (T)1e-10f;
  const T pcx = T(0.5) * (cmax.x + cmin.x), pcy = T(0.5) * (cmax.y + cmin.y), pcz = T(0.5) * (cmax.z + cmin.z);
  const T ecx = T(0.5) * (cmax.x - cmin.x) + eps, ecy = T(0.5) * (cmax.y - cmin.y) + eps,
          ecz = T(0.5) * (cmax.z - cmin.z) + eps;
  const T vx = pr.x - pcx, vy = pr.y - pcy, vz = pr.z - pcz, vw = pr.w - pw;
  const T ma = std::fmax(std::fabs(vx / ecx), std::fmax(std::fabs(vy / ecy), std::fabs(vz / ecz)));
  if (ma > T(1)) pr = P4<T>{pcx + vx / ma, pcy + vy / ma, pcz + vz / ma, pw + vw / ma};
  // Luminance and feedback (:246-258)
  const T l0 = (curr.x * (T)0.2125f + curr.y * (T)0.7154f) + curr.z * (T)0.0721f;
  const T l1 = (pr.x * (T)0.2125f + pr.y * (T)0.7154f) + pr.z * (T)0.0721f;
  const T diff = std::fabs(l0 - l1) / std::fmax(l0, std::fmax(l1, (T)0.2f));
  const T w = T(1) - diff;
  const T k = T(0) * (T(1) - w * w) + T(0.5) * (w * w);
  auto mix = [&](T a, T b) { return a * (T(1) - k) + b * k; };
  return P4<T>{mix(pr.x, curr.x), mix(pr.y, curr.y), mix(pr.z, curr.z), mix(pr.w, curr.w)};
}

}  // namespace orc_post

// cloud_oracle.hpp -- CPU restatement of Meteoros' cloud compute stage.
//
// TEST INFRASTRUCTURE ONLY.  This is the correctness oracle for the HIP
// kernels of meteoros_amd; nothing in the product links or calls it.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// and only as the checker or the timed CPU baseline.
//
// One template, two twins: T=float is the FP32 parity oracle (compile with
// -ffp-contract=off, no fast-math: every operation rounds once, in GLSL
// source order); T=double is the FP64 accuracy oracle (same code, same
// texels, double arithmetic).
//
// Followed sources (text only; the GLSL itself cannot run here):
//   src/CloudScapes/shaders/cloudRayMarch.comp   (all functions cited below)
//   src/CloudScapes/shaders/reprojection.comp    (:72-244)
//   src/CloudScapes/Texture3D.cpp:92-134, Image.cpp:305-346 (sampler state)
//
// Sampler model (the Vulkan driver's texture() restated; the driver's own
// filtering precision is not reproducible here):
//   u = s*N - 0.5 ; i0 = floor(u) ; a = u - floor(u) ; REPEAT wrap i & (N-1)
//   lerp(p, q, w) = fma(w, q - p, p), x first, then y, then z
//   UNORM8 texels enter as their integer code 0..255 and the filtered value
//   is multiplied by fl(1/255); FP16 texels enter as their exact value.
//
// Pinning (tests/test_oracle_*.py, tests/test_lighting_pins.py): closed
// forms for Halton, remap, smoothstep, EncodeFloatRGBA, HG normalisation
// and the sphere quadratic; brute-force Vulkan-formula trilinear on tiny
// volumes; invariants of the march (empty volume => background, ocean
// colour, accum in range, trip counts); the sky's scattering constants
// recomputed from the physics, its symmetries and the Rayleigh phase
// normalisation; GetLightEnergy's factor structure; FP32 vs FP64
// agreement; the air-mass term by its physical limits (1 at the zenith,
// sec z while the atmosphere looks plane-parallel, ~36 at the horizon).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/meteoros_cloud.h"

namespace orc {

// <cmath> through these names (called unqualified), so that the counting
// scalar of flop_count.hpp can overload them by argument-dependent lookup.
template <class T> static inline T m_floor(T x) { return std::floor(x); }
template <class T> static inline T m_sqrt(T x) { return std::sqrt(x); }
template <class T> static inline T m_exp(T x) { return std::exp(x); }
template <class T> static inline T m_pow(T x, T y) { return std::pow(x, y); }
template <class T> static inline T m_acos(T x) { return std::acos(x); }
template <class T> static inline T m_cos(T x) { return std::cos(x); }
template <class T> static inline T m_fabs(T x) { return std::fabs(x); }
template <class T> static inline T m_fma(T a, T b, T c) { return std::fma(a, b, c); }
template <class T> static inline T m_fmax(T a, T b) { return std::fmax(a, b); }
template <class T> static inline T m_fmin(T a, T b) { return std::fmin(a, b); }
template <class T> static inline T m_nearbyint(T x) { return std::nearbyint(x); }

// Work regions of the march, for the operation counter of flop_count.hpp
// (a no-op for float and double): per-ray set-up and composite, primary
// step, in-cloud step, cone sample, cone erosion, reprojected pixel.
enum FlopRegion { FR_RAY, FR_PRIMARY, FR_CLOUD, FR_CONE, FR_CONE_EROSION, FR_REPROJ, FR_COUNT };
template <class T> struct Region {
  static void set(int) {}
};

template <class T> struct v3 {
  T x, y, z;
};
template <class T> static inline v3<T> V(T x, T y, T z) { return v3<T>{x, y, z}; }
template <class T> static inline v3<T> operator+(v3<T> a, v3<T> b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> static inline v3<T> operator-(v3<T> a, v3<T> b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> static inline v3<T> operator-(v3<T> a) { return V(-a.x, -a.y, -a.z); }
template <class T> static inline v3<T> operator*(v3<T> a, T s) { return V(a.x * s, a.y * s, a.z * s); }
template <class T> static inline v3<T> operator*(T s, v3<T> a) { return V(s * a.x, s * a.y, s * a.z); }
template <class T> static inline v3<T> operator*(v3<T> a, v3<T> b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T> static inline v3<T> operator/(v3<T> a, T s) { return V(a.x / s, a.y / s, a.z / s); }
template <class T> static inline v3<T> operator/(v3<T> a, v3<T> b) { return V(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class T> static inline T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> static inline T length(v3<T> a) { return m_sqrt(dot(a, a)); }
// GLSL normalize(v) = v / length(v)
template <class T> static inline v3<T> normalize(v3<T> a) {
  T l = length(a);
  return a / l;
}
template <class T> static inline v3<T> cross(v3<T> a, v3<T> b) {
  return V(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
template <class T> static inline T clampT(T x, T lo, T hi) { return m_fmin(m_fmax(x, lo), hi); }
// GLSL mix(x, y, a) = x*(1-a) + y*a
template <class T> static inline T mixT(T x, T y, T a) { return x * (T(1) - a) + y * a; }
template <class T> static inline T fractT(T x) { return x - m_floor(x); }
template <class T> static inline T smoothstepT(T e0, T e1, T x) {
  T t = clampT((x - e0) / (e1 - e0), T(0), T(1));
  return t * t * (T(3) - T(2) * t);
}
template <class T> static inline T fmaT(T a, T b, T c) { return m_fma(a, b, c); }

// cloudRayMarch.comp:153-169
template <class T> static inline T remap(T v, T omin, T omax, T nmin, T nmax) {
  return nmin + (((v - omin) / (omax - omin)) * (nmax - nmin));
}
template <class T> static inline T remapClamped(T v, T omin, T omax, T nmin, T nmax) {
  T t = nmin + (((v - omin) / (omax - omin)) * (nmax - nmin));
  return clampT(t, nmin, nmax);
}
template <class T> static inline T remapClampedBeforeAndAfter(T v, T omin, T omax, T nmin, T nmax) {
  v = clampT(v, omin, omax);
  T t = nmin + (((v - omin) / (omax - omin)) * (nmax - nmin));
  return clampT(t, nmin, nmax);
}

// IEEE binary16 -> value (exact).
static inline double half_to_double(uint16_t h) {
  int s = (h >> 15) & 1, e = (h >> 10) & 31, m = h & 1023;
  double v;
  if (e == 0) v = std::ldexp((double)m, -24);
  else if (e == 31) v = m ? NAN : INFINITY;
  else v = std::ldexp((double)(1024 + m), e - 25);
  return s ? -v : v;
}

struct Volumes {
  const void* lf; int lf_dim;
  const void* hf; int hf_dim;
  const uint8_t* curl; int curl_dim;
  const uint8_t* weather; int weather_dim;
  int format;  // MTC_FMT_*
};

template <class T> struct C4 { T r, g, b, a; };

template <class T> static inline C4<T> texel(const void* base, int fmt, size_t idx) {
  if (fmt == MTC_FMT_FP16) {
    const uint16_t* p = (const uint16_t*)base + idx * 4;
    return C4<T>{(T)half_to_double(p[0]), (T)half_to_double(p[1]), (T)half_to_double(p[2]),
                 (T)half_to_double(p[3])};
  }
  const uint8_t* p = (const uint8_t*)base + idx * 4;
  return C4<T>{(T)p[0], (T)p[1], (T)p[2], (T)p[3]};
}
template <class T> static inline T lerp(T p, T q, T w) { return fmaT(w, q - p, p); }
template <class T> static inline C4<T> lerp4(C4<T> p, C4<T> q, T w) {
  return C4<T>{lerp(p.r, q.r, w), lerp(p.g, q.g, w), lerp(p.b, q.b, w), lerp(p.a, q.a, w)};
}
template <class T> static inline C4<T> unorm_scale(C4<T> c, int fmt) {
  if (fmt == MTC_FMT_FP16) return c;
  const T k = (T)(1.0 / 255.0);
  return C4<T>{c.r * k, c.g * k, c.b * k, c.a * k};
}

// texture(sampler3D, p): linear filter, REPEAT, LOD 0 (Texture3D.cpp:92-134)
template <class T> static C4<T> sample3d(const void* base, int fmt, int N, v3<T> s) {
  T ux = s.x * (T)N - T(0.5), uy = s.y * (T)N - T(0.5), uz = s.z * (T)N - T(0.5);
  T fx = m_floor(ux), fy = m_floor(uy), fz = m_floor(uz);
  T ax = ux - fx, ay = uy - fy, az = uz - fz;
  long long ix = (long long)fx, iy = (long long)fy, iz = (long long)fz;
  const long long m = N - 1;
  size_t x0 = (size_t)(ix & m), x1 = (size_t)((ix + 1) & m);
  size_t y0 = (size_t)(iy & m), y1 = (size_t)((iy + 1) & m);
  size_t z0 = (size_t)(iz & m), z1 = (size_t)((iz + 1) & m);
  size_t n = (size_t)N;
  auto at = [&](size_t x, size_t y, size_t z) { return texel<T>(base, fmt, (z * n + y) * n + x); };
  C4<T> c00 = lerp4(at(x0, y0, z0), at(x1, y0, z0), ax);
  C4<T> c10 = lerp4(at(x0, y1, z0), at(x1, y1, z0), ax);
  C4<T> c01 = lerp4(at(x0, y0, z1), at(x1, y0, z1), ax);
  C4<T> c11 = lerp4(at(x0, y1, z1), at(x1, y1, z1), ax);
  C4<T> c0 = lerp4(c00, c10, ay);
  C4<T> c1 = lerp4(c01, c11, ay);
  return unorm_scale(lerp4(c0, c1, az), fmt);
}
// texture(sampler2D, p): linear filter, REPEAT, UNORM8 (Image.cpp:305-346)
template <class T> static C4<T> sample2d(const uint8_t* base, int N, T sx, T sy) {
  T ux = sx * (T)N - T(0.5), uy = sy * (T)N - T(0.5);
  T fx = m_floor(ux), fy = m_floor(uy);
  T ax = ux - fx, ay = uy - fy;
  long long ix = (long long)fx, iy = (long long)fy;
  const long long m = N - 1;
  size_t x0 = (size_t)(ix & m), x1 = (size_t)((ix + 1) & m);
  size_t y0 = (size_t)(iy & m), y1 = (size_t)((iy + 1) & m);
  size_t n = (size_t)N;
  auto at = [&](size_t x, size_t y) { return texel<T>(base, MTC_FMT_UNORM8, y * n + x); };
  C4<T> c0 = lerp4(at(x0, y0), at(x1, y0), ax);
  C4<T> c1 = lerp4(at(x0, y1), at(x1, y1), ax);
  return unorm_scale(lerp4(c0, c1, ay), MTC_FMT_UNORM8);
}

struct Counters {
  uint64_t rays_marched = 0, rays_ocean = 0, rays_sky = 0, primary_steps = 0, cloud_steps = 0,
           cone_samples = 0, cone_erosions = 0, early_exits = 0;
  void add(const Counters& o) {
    rays_marched += o.rays_marched; rays_ocean += o.rays_ocean; rays_sky += o.rays_sky;
    primary_steps += o.primary_steps; cloud_steps += o.cloud_steps;
    cone_samples += o.cone_samples; cone_erosions += o.cone_erosions; early_exits += o.early_exits;
  }
};

// Everything a ray needs from the uniforms, converted to T.
template <class T> struct Frame {
  // camera (cloudRayMarch.comp:15-21)
  T view[16];
  T eye_ubo[3];
  T tfx, tfy;
  // time (:23-31)
  T h1[4], h2[4], h3[4], h4[4];
  T total_time;
  int frame;  // frameCountMod16
  // sun
  v3<T> sun_loc, sky_sun;
  T sun_intensity;
  // params
  T earth_r, r_in, r_out, thickness, inner_off;
  int step_mode, steps_min, steps_max, light_samples, coverage_mode;
  T base_density_factor, jitter_factor, jitter_dim, sample_divisor, erosion_scale, accum_scale,
      cone_erosion_scale, brightness, hg_g, hg_si, hg_ss, wind_bias, cloud_speed, top_offset,
      wind_skew, coverage_const, weather_scale, bg_min, fade_start, fade_end;
  v3<T> wind_dir;
  int blur_samples;
  int W, H;
};

template <class T>
static Frame<T> make_frame(const mtc_camera& cam, const mtc_time& tm, const mtc_sun& sun,
                           const mtc_cloud_params& p, int W, int H) {
  Frame<T> f;
  for (int i = 0; i < 16; ++i) f.view[i] = (T)cam.view[i];
  f.eye_ubo[0] = (T)cam.eye[0]; f.eye_ubo[1] = (T)cam.eye[1]; f.eye_ubo[2] = (T)cam.eye[2];
  f.tfx = (T)cam.tan_fov_by2[0]; f.tfy = (T)cam.tan_fov_by2[1];
  for (int i = 0; i < 4; ++i) {
    f.h1[i] = (T)tm.halton1[i]; f.h2[i] = (T)tm.halton2[i];
    f.h3[i] = (T)tm.halton3[i]; f.h4[i] = (T)tm.halton4[i];
  }
  f.total_time = (T)tm.time[1];
  f.frame = tm.frame_count_mod16;
  f.sun_loc = V((T)sun.sun_location[0], (T)sun.sun_location[1], (T)sun.sun_location[2]);
  f.sky_sun = V((T)p.sky_sun_location[0], (T)p.sky_sun_location[1], (T)p.sky_sun_location[2]);
  f.sun_intensity = (T)sun.sun_intensity;
  f.earth_r = (T)p.earth_radius; f.r_in = (T)p.shell_inner_radius; f.r_out = (T)p.shell_outer_radius;
  f.thickness = f.r_out - f.r_in;       // ATMOSPHERE_THICKNESS
  f.inner_off = f.r_in - f.earth_r;     // ATMOSPHERE_RADIUS_INNER - EARTH_RADIUS
  f.step_mode = p.step_mode; f.steps_min = p.steps_min; f.steps_max = p.steps_max;
  f.light_samples = p.light_samples; f.coverage_mode = p.coverage_mode;
  f.base_density_factor = (T)p.base_density_factor; f.jitter_factor = (T)p.jitter_factor;
  f.jitter_dim = (T)p.jitter_dim; f.sample_divisor = (T)p.sample_divisor;
  f.erosion_scale = (T)p.erosion_scale; f.accum_scale = (T)p.accum_scale;
  f.cone_erosion_scale = (T)p.cone_erosion_scale; f.brightness = (T)p.brightness;
  f.hg_g = (T)p.hg_eccentricity; f.hg_si = (T)p.hg_silver_intensity; f.hg_ss = (T)p.hg_silver_spread;
  f.wind_dir = V((T)p.wind_direction[0], (T)p.wind_direction[1], (T)p.wind_direction[2]);
  f.wind_bias = (T)p.wind_upward_bias; f.cloud_speed = (T)p.cloud_speed;
  f.top_offset = (T)p.cloud_top_offset; f.wind_skew = (T)p.wind_skew;
  f.coverage_const = (T)p.coverage_constant; f.weather_scale = (T)p.weather_scale;
  f.bg_min = (T)p.background_min; f.fade_start = (T)p.cloud_fade_start; f.fade_end = (T)p.cloud_fade_end;
  f.blur_samples = p.motion_blur_samples;
  f.W = W; f.H = H;
  return f;
}

// ---------------------------------------------------------------------------
// cloudRayMarch.comp
// ---------------------------------------------------------------------------

// EncodeFloatRGBA, :105-112
template <class T> static inline void encode_float_rgba(T v, T out[4]) {
  const T e[4] = {T(1), T(255), T(65025), T(16581375)};
  T enc[4];
  for (int i = 0; i < 4; ++i) enc[i] = fractT(e[i] * v);
  const T k = (T)(1.0 / 255.0);
  // enc -= enc.yzww * vec2(1./255., 0.).xxxy
  T s0 = enc[1] * k, s1 = enc[2] * k, s2 = enc[3] * k, s3 = enc[3] * T(0);
  out[0] = enc[0] - s0; out[1] = enc[1] - s1; out[2] = enc[2] - s2; out[3] = enc[3] - s3;
}

// getJitterOffset(int, ivec2) of the march, :114-132 (seq3/seq4 for index >= 4)
template <class T> static inline void jitter_march(const Frame<T>& f, int index, T dimx, T dimy, T& jx, T& jy) {
  index = index / 2;
  if (index < 4) { jx = f.h1[index]; jy = f.h2[index]; }
  else { index -= 4; jx = f.h3[index]; jy = f.h4[index]; }
  jx = jx / dimx; jy = jy / dimy;
}
// getJitterOffset(int, ivec2) of the reprojection, reprojection.comp:72-90
// (seq1/seq2 in both halves)
template <class T> static inline void jitter_reproj(const Frame<T>& f, int index, T dimx, T dimy, T& jx, T& jy) {
  index = index / 2;
  if (index < 4) { jx = f.h1[index]; jy = f.h2[index]; }
  else { index -= 4; jx = f.h1[index]; jy = f.h2[index]; }
  jx = jx / dimx; jy = jy / dimy;
}

template <class T> struct Ray { v3<T> o, d; };

// castRay, :194-226 and reprojection.comp:111-144
template <class T>
static Ray<T> cast_ray(const T* view, T tfx, T tfy, T ux, T uy, v3<T> eye, T jx, T jy) {
  v3<T> right = normalize(V(view[0], view[4], view[8]));   // view[0][0], view[1][0], view[2][0]
  v3<T> up = normalize(V(view[1], view[5], view[9]));      // view[0][1], view[1][1], view[2][1]
  v3<T> look = -normalize(V(view[2], view[6], view[10]));  // view[0][2], view[1][2], view[2][2]
  T nx = ux * T(2) - T(1), ny = uy * T(2) - T(1);
  nx = nx + jx; ny = ny + jy;
  v3<T> cx = (nx * tfx) * right;
  v3<T> cy = (ny * tfy) * up;
  v3<T> ref = eye + look;
  v3<T> p = (ref + cx) + cy;
  Ray<T> r;
  r.o = eye;
  r.d = normalize(p - eye);
  return r;
}

template <class T> struct Isect { v3<T> normal, point; bool valid; T t; };

// raySphereIntersection, :229-273 (t = length(point_world - rO_normalised))
template <class T> static Isect<T> ray_sphere(v3<T> rO, v3<T> rD, v3<T> c, T R) {
  Isect<T> is;
  is.valid = false; is.point = V(T(0), T(0), T(0)); is.normal = V(T(0), T(1), T(0)); is.t = T(0);
  rO = rO - c;
  rO = rO / R;
  T A = dot(rD, rD);
  T B = T(2) * dot(rD, rO);
  T C = dot(rO, rO) - T(1);
  T disc = B * B - T(4) * A * C;
  if (disc < T(0)) return is;
  T t = (-B - m_sqrt(disc)) / (T(2) * A);
  if (t < T(0)) t = (-B + m_sqrt(disc)) / (T(2) * A);
  if (t >= T(0)) {
    v3<T> p = rO + t * rD;
    is.valid = true;
    is.normal = normalize(p);
    p = p * R;
    p = p + c;
    is.point = p;
    is.t = length(p - rO);
  }
  return is;
}

// HenyeyGreenstein / HGModified, :295-306
template <class T> static inline T hg(T c, T g) {
  T num = T(1) - g * g;
  T den = m_pow((T(1) + g * g - T(2) * g * c), T(1.5));
  return (num / den) * (T)0.07957747154594767;
}
template <class T> static inline T hg_modified(T c, T g, T si, T ss) {
  return m_fmax(hg(c, g), si * hg(c, T(0.99) - ss));
}

// GetLightEnergy, :331-388 (third instance, "MEDIUM")
template <class T>
static inline T light_energy(T h, T dl, T ds, T phase, T cosang, T /*step*/, T brightness) {
  T prim = m_exp(-dl);
  T sec = m_exp(-dl);
  T att = m_fmax(remap(cosang, T(0.7), T(1.0), sec, sec * T(0.25)), prim);
  T depth = T(0.05) + m_pow(ds, clampT(remap(h * T(0.125), T(0.3), T(0.85), T(0.5), T(2.0)), T(0.5), T(2.0)));
  T vert = m_pow(clampT(remap(h * T(1.5), T(0.07), T(0.34), T(0.1), T(1.0)), T(0.1), T(1.0)), T(0.8));
  T insc = depth * vert;
  return att * prim * insc * phase * brightness;
}

// Preetham sky, :401-467
template <class T> static inline T rayleigh_phase(T c) { return (T)0.05968310365946075 * (T(1) + c * c); }
template <class T> static inline v3<T> rayleigh_total() {
  return V((T)5.804542996261093E-6, (T)1.3562911419845635E-5, (T)3.0265902468824876E-5);
}
template <class T> static inline v3<T> calc_beta_r(const Frame<T>& f) {
  T rayleigh = T(2);
  T sunFade = T(1) - clampT(T(1) - m_exp(f.sky_sun.y / T(450000.0)), T(0), T(1));
  return rayleigh_total<T>() * (rayleigh - T(1) + sunFade);
}
// MIE_CONST (:85): pi (2 pi / lambda)^(v - 2) K at lambda = 680/550/450 nm,
// v = 4, K = (0.686, 0.678, 0.666), scaled by 1e-14 (the shader's value)
template <class T> static inline v3<T> mie_const() {
  return V((T)1.839991851443397, (T)2.779802391966052, (T)4.079047954386109);
}
template <class T> static inline v3<T> calc_beta_v() {
  T turbidity = T(10);
  T mie = T(0.005);
  T c = (T(0.2) * turbidity) * T(10E-18);
  v3<T> mc = mie_const<T>();
  return ((T(0.434) * c) * mc) * mie;
}
template <class T> static inline T calc_sun_intensity(const Frame<T>& f) {
  T zc = clampT(normalize(f.sky_sun).y, T(-1), T(1));
  return T(1000.0) * m_fmax(T(0), T(1) - m_pow((T)2.718281828459,
                                                     -(((T)1.6110731557 - m_acos(zc)) / T(1.5))));
}
// the optical length's "inverse" (:438): Kasten's relative optical air mass
// 1 / (cos z + 0.15 (93.885 - z_deg)^-1.253)
template <class T> static inline T relative_air_mass(T zenith) {
  return T(1.0) / (m_cos(zenith) + T(0.15) * m_pow(T(93.885) - ((zenith * T(180.0)) / (T)3.14159265), T(-1.253)));
}
template <class T> static v3<T> atmosphere_color(const Frame<T>& f, v3<T> dir, v3<T> sunDir, T sunIntensity) {
  sunDir = normalize(sunDir);
  T sunE = sunIntensity * calc_sun_intensity(f);
  v3<T> BR = calc_beta_r(f);
  v3<T> BM = calc_beta_v<T>();
  T zenith = m_acos(m_fmax(T(0), dir.y));
  T inv = relative_air_mass(zenith);
  T sR = T(8.4E3) * inv;
  T sM = T(1.25E3) * inv;
  v3<T> ex = (-BR) * sR + BM * sM;
  v3<T> fex = V(m_exp(ex.x), m_exp(ex.y), m_exp(ex.z));
  T cosTheta = dot(sunDir, dir);
  T rPhase = rayleigh_phase(cosTheta * T(0.5) + T(0.5));
  v3<T> bRT = BR * rPhase;
  T mPhase = hg(cosTheta, T(0.8));
  v3<T> bMT = BM * mPhase;
  T yDot = T(1) - sunDir.y;
  yDot = yDot * (yDot * yDot * yDot * yDot);
  v3<T> betas = (bRT + bMT) / (BR + BM);
  v3<T> a = (sunE * betas) * (V(T(1), T(1), T(1)) - fex);
  v3<T> Lin = V(m_pow(a.x, T(1.5)), m_pow(a.y, T(1.5)), m_pow(a.z, T(1.5)));
  v3<T> b = (sunE * betas) * fex;
  v3<T> pb = V(m_pow(b.x, T(0.5)), m_pow(b.y, T(0.5)), m_pow(b.z, T(0.5)));
  T yc = clampT(yDot, T(0), T(1));
  v3<T> m = V(mixT(T(1), pb.x, yc), mixT(T(1), pb.y, yc), mixT(T(1), pb.z, yc));
  Lin = Lin * m;
  v3<T> L0 = T(0.1) * fex;
  const T sac = (T)0.999956676946448443553574619906976478926848692873900859324;
  T sunDisk = smoothstepT(sac, sac + T(0.00002), cosTheta);
  L0 = L0 + ((sunE * T(15000.0)) * fex) * sunDisk;
  return (Lin + L0) * T(0.04) + V(T(0), T(0.0003), T(0.00075));
}

// getRelativeHeightInAtmosphere, :171-186
template <class T>
static inline T rel_height(const Frame<T>& f, v3<T> point, v3<T> c, v3<T> start, v3<T> dir, v3<T> eye) {
  T lr = length(point - eye);
  T ls = length(start - eye);
  v3<T> n = normalize(point - c);
  T cosT = dot(dir, n);
  T num = m_fabs(cosT * (lr - ls));
  return num / f.thickness;
}
// getRelativePositionInAtmosphere, :188-191
template <class T> static inline v3<T> rel_pos(const Frame<T>& f, v3<T> pos, v3<T> c) {
  return (pos - V(c.x, f.inner_off, c.z)) / f.thickness;
}
// skewSamplePointWithWind, :489-497
template <class T> static inline v3<T> skew(const Frame<T>& f, v3<T> p, T h) {
  p = p + ((h * f.wind_dir) * f.top_offset) * f.wind_skew;
  p = p + ((f.wind_dir + V(T(0), f.wind_bias, T(0))) * f.cloud_speed) * f.total_time;
  return p;
}
// getDensityHeightGradientForPoint, :475-487 (weather-coverage mode only)
template <class T> static inline T height_gradient(T h, T type) {
  h = clampT(h, T(0), T(1));
  T cu = m_fmax(T(0), remap(h, T(0.01), T(0.3), T(0), T(1)) * remap(h, T(0.6), T(0.95), T(1), T(0)));
  T sc = m_fmax(T(0), remap(h, T(0), T(0.25), T(0), T(1)) * remap(h, T(0.3), T(0.65), T(1), T(0)));
  T st = m_fmax(T(0), remap(h, T(0), T(0.1), T(0), T(1)) * remap(h, T(0.2), T(0.3), T(1), T(0)));
  T a = mixT(st, sc, clampT(type * T(2), T(0), T(1)));
  T b = mixT(sc, st, clampT((type - T(0.5)) * T(2), T(0), T(1)));
  (void)cu;
  return mixT(a, b, type);
}
// sampleLowFrequency, :499-540
template <class T>
static inline T sample_low(const Frame<T>& f, const Volumes& vol, v3<T> p, v3<T> unskewed, T h,
                           C4<T>* lf_out = nullptr) {
  C4<T> n = sample3d<T>(vol.lf, vol.format, vol.lf_dim, p);
  if (lf_out) *lf_out = n;
  T fbm = (n.g * T(0.625)) + (n.b * T(0.25)) + (n.a * T(0.125));
  fbm = clampT(fbm, T(0), T(1));
  T base = remapClamped(n.r, (fbm - T(0.9)), T(1.0), T(0), T(1));
  T cov = f.coverage_const;
  if (f.coverage_mode == MTC_COVERAGE_WEATHER) {
    C4<T> w = sample2d<T>(vol.weather, vol.weather_dim, unskewed.x * f.weather_scale, unskewed.z * f.weather_scale);
    base = base * (height_gradient(h, w.g) * T(0.5));
    cov = w.r;
  }
  T bc = remapClampedBeforeAndAfter(base, cov, T(1.0), T(0), T(1));
  bc = bc * cov;
  return bc;
}
// erodeCloudWithHighFrequency, :542-563
template <class T>
static inline T erode(const Frame<T>& f, const Volumes& vol, T base, v3<T> p, T h, C4<T>* curl_out = nullptr,
                      C4<T>* hf_out = nullptr, v3<T>* p_out = nullptr) {
  (void)f;
  C4<T> cn = sample2d<T>(vol.curl, vol.curl_dim, p.x, p.y);
  // point.xy += curlNoise.xy * (1.0 - height_fraction) * 0.5
  T k = T(1) - h;
  p.x = p.x + (cn.r * k) * T(0.5);
  p.y = p.y + (cn.g * k) * T(0.5);
  C4<T> hn = sample3d<T>(vol.hf, vol.format, vol.hf_dim, p);
  if (curl_out) *curl_out = cn;
  if (hf_out) *hf_out = hn;
  if (p_out) *p_out = p;
  T fbm = (hn.r * T(0.625)) + (hn.g * T(0.25)) + (hn.b * T(0.125));
  T mod = clampT(mixT(fbm, T(1) - fbm, clampT(h * T(2), T(0), T(1))), T(0), T(1));
  return remap(base, mod * T(0.005), T(1.0), T(0), T(1));
}

// One in-cloud step as the learned light sees it: its step index, the nine
// network inputs (the fifth is the cone's first density term) and the
// training target, the log transmittance of the rest of the cone,
// -(terms 1 .. n-1).  No GLSL counterpart; recorded by ray_march when asked.
struct LightRecord {
  int k;
  double x[9];
  double y;
  double c[6];     // the cone's density terms (0 where the sample is empty)
  double w[2];     // weather (coverage, type) at the cone's first sample
};

// The values of one executed primary step, recorded when ray_march is given
// a trace list, so that tests can pin each stage of the step against closed
// forms.  All doubles (k and in_cloud too) so the record is one flat array.
// No GLSL counterpart.
struct StepTrace {
  double k, in_cloud, t;
  double pos[3];   // ray.o + t * (ray.d + jitter)
  double sp[3];    // getRelativePositionInAtmosphere(pos) / sample_divisor
  double sk[3];    // skewSamplePointWithWind(sp, h)
  double h;        // getRelativeHeightInAtmosphere
  double lf[4];    // the primary LF sample (texture(LF, sk))
  double base;     // sampleLowFrequency(...) * base_density_factor
  double curl[2];  // curl sample of the primary erosion (in-cloud steps)
  double hf[3];    // HF sample of the primary erosion
  double ero[3];   // the point the HF volume is sampled at
  double hfd, accum, dal, energy, transmittance;  // after the step
  double cone_lf_r[6];   // r of the LF sample of cone sample i
  double cone_pos[6][3]; // the point cone sample i reads the LF at
  double cone_e[6];      // cone sample i's eroded density (0 when empty)
};

// rayMarch, :565-688
template <class T>
static v3<T> ray_march(const Frame<T>& f, const Volumes& vol, const Ray<T>& ray, v3<T> c, v3<T> startPos,
                       T start_t, T end_t, int pixelID, T& accum, Counters& cnt,
                       std::vector<LightRecord>* rec = nullptr, std::vector<StepTrace>* trace = nullptr) {
  T d = dot(ray.d, V(T(0), T(1), T(0)));
  T maxSteps;
  if (f.step_mode == MTC_STEPS_FIXED) maxSteps = (T)f.steps_max;
  else maxSteps = m_floor(mixT((T)f.steps_min, (T)f.steps_max, T(1) - d));
  T span = end_t - start_t;
  T stepSize = span / maxSteps;
  T transmittance = T(1);
  v3<T> ret = V(T(0), T(0), T(0));

  v3<T> lightDir = normalize(f.sun_loc - ray.o);
  T cos_angle = dot(normalize(ray.d), lightDir);
  T HG_light = hg_modified(cos_angle, f.hg_g, f.hg_si, f.hg_ss);

  v3<T> mc;
  T ax = m_fabs(lightDir.x), ay = m_fabs(lightDir.y), az = m_fabs(lightDir.z);
  if (ax > ay && ax > az) mc = V(ax, T(0), T(0));
  else if (ay > ax && ay > az) mc = V(T(0), ay, T(0));
  else mc = V(T(0), T(0), az);
  v3<T> zc = cross(lightDir, mc);
  v3<T> xc = cross(zc, lightDir);
  // sunRotMatrix = mat3(xc, lightDir, zc); M*v = xc*v.x + lightDir*v.y + zc*v.z
  const T kv[6][3] = {{T(0.1), T(0.25), T(-0.15)}, {T(0.2), T(0.5), T(0.2)}, {T(-0.2), T(0.1), T(-0.1)},
                      {T(-0.05), T(0.75), T(0.05)}, {T(-0.1), T(1.0), T(0.0)}, {T(0.0), T(3.0), T(0.0)}};
  v3<T> kern[6];
  for (int i = 0; i < 6; ++i) kern[i] = (xc * kv[i][0] + lightDir * kv[i][1]) + zc * kv[i][2];

  // Iteration cap: the GLSL loop has none, but t + stepSize == t would spin
  // forever; the cap is far above any trip count of a valid scene.
  const int cap = 4 * (f.steps_min > f.steps_max ? f.steps_min : f.steps_max) + 64;
  int iter = 0;
  for (T t = start_t; t < end_t && iter < cap; t += stepSize, ++iter) {
    Region<T>::set(FR_PRIMARY);
    cnt.primary_steps++;
    int idx = pixelID + (int)t;
    T fm = (T)idx;
    int _index = (int)(fm - T(16) * m_floor(fm / T(16)));  // GLSL mod(x, 16.0)
    T jx, jy;
    jitter_march(f, _index, f.jitter_dim, f.jitter_dim, jx, jy);
    v3<T> jd = V(jx, (jx + jy) * f.jitter_factor, jy);
    v3<T> pos = ray.o + t * (ray.d + jd);
    v3<T> sp = rel_pos(f, pos, c);
    sp = sp / f.sample_divisor;
    T h = rel_height(f, pos, c, startPos, ray.d, ray.o);
    v3<T> sk = skew(f, sp, h);
    StepTrace st{};
    C4<T> lfs{}, cls{}, hfs{};
    v3<T> ep{};
    T base = sample_low(f, vol, sk, sp, h, &lfs) * f.base_density_factor;
    if (trace) {
      st.k = iter; st.t = (double)t; st.h = (double)h; st.base = (double)base;
      const v3<T> vs[3] = {pos, sp, sk};
      double* ds[3] = {st.pos, st.sp, st.sk};
      for (int j = 0; j < 3; ++j) { ds[j][0] = (double)vs[j].x; ds[j][1] = (double)vs[j].y; ds[j][2] = (double)vs[j].z; }
      st.lf[0] = (double)lfs.r; st.lf[1] = (double)lfs.g; st.lf[2] = (double)lfs.b; st.lf[3] = (double)lfs.a;
    }
    if (base > T(0)) {
      Region<T>::set(FR_CLOUD);
      cnt.cloud_steps++;
      T hfd = erode(f, vol, base * f.erosion_scale, sk, h, &cls, &hfs, &ep);
      accum += hfd * f.accum_scale;
      T dal = T(0);
      T terms[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
      for (int i = 0; i < f.light_samples; ++i) {
        Region<T>::set(FR_CONE);
        cnt.cone_samples++;
        v3<T> lp = pos + ((stepSize * kern[i]) * (T)i);
        v3<T> slp = rel_pos(f, lp, c);
        C4<T> cls_i{};
        T cb = sample_low(f, vol, slp, slp, h, &cls_i);
        if (trace && i < 6) {
          st.cone_lf_r[i] = (double)cls_i.r;
          st.cone_pos[i][0] = (double)slp.x; st.cone_pos[i][1] = (double)slp.y; st.cone_pos[i][2] = (double)slp.z;
        }
        if (cb > T(0)) {
          Region<T>::set(FR_CONE_EROSION);
          cnt.cone_erosions++;
          T e = erode(f, vol, f.cone_erosion_scale * cb, sk, h);
          if (i < 6) terms[i] = e;
          dal += e;
        }
      }
      Region<T>::set(FR_CLOUD);
      if (rec) {
        // inputs: position in the shell (x, z modulo the weather tile, y
        // modulo the LF tile), height, the cone's first term, sun direction,
        // cos of the ray's zenith angle
        v3<T> rel = rel_pos(f, pos, c);
        LightRecord r;
        r.k = iter;
        r.x[0] = (double)fractT(rel.x * f.weather_scale);
        r.x[1] = (double)fractT(rel.y);
        r.x[2] = (double)fractT(rel.z * f.weather_scale);
        r.x[3] = (double)h;
        r.x[4] = (double)terms[0];  // local density as the light march sees it: its first term
        r.x[5] = (double)lightDir.x;
        r.x[6] = (double)lightDir.y;
        r.x[7] = (double)lightDir.z;
        r.x[8] = (double)ray.d.y;
        // target: the rest of the cone, -(terms 1 .. n-1)
        T rest = T(0);
        for (int i = 1; i < f.light_samples && i < 6; ++i) rest = rest + terms[i];
        r.y = -(double)rest;
        for (int i = 0; i < 6; ++i) r.c[i] = (double)terms[i];
        C4<T> w0 = sample2d<T>(vol.weather, vol.weather_dim, rel.x * f.weather_scale, rel.z * f.weather_scale);
        r.w[0] = (double)w0.r;
        r.w[1] = (double)w0.g;
        rec->push_back(r);
      }
      T E = light_energy(h, dal, base, HG_light, cos_angle, stepSize, f.brightness);
      transmittance = mixT(transmittance, E, T(1) - accum);
      ret = ret + V(transmittance, transmittance, transmittance);
      if (trace) {
        st.in_cloud = 1.0;
        st.curl[0] = (double)cls.r; st.curl[1] = (double)cls.g;
        st.hf[0] = (double)hfs.r; st.hf[1] = (double)hfs.g; st.hf[2] = (double)hfs.b;
        st.ero[0] = (double)ep.x; st.ero[1] = (double)ep.y; st.ero[2] = (double)ep.z;
        st.hfd = (double)hfd; st.dal = (double)dal; st.energy = (double)E;
        for (int i = 0; i < 6; ++i) st.cone_e[i] = (double)terms[i];
      }
    }
    Region<T>::set(FR_PRIMARY);
    if (trace) {
      st.accum = (double)accum;
      st.transmittance = (double)transmittance;
      trace->push_back(st);
    }
    if (accum >= T(1)) {
      accum = T(1);
      cnt.early_exits++;
      break;
    }
  }
  Region<T>::set(FR_RAY);
  return ret;
}

// main of cloudRayMarch.comp, :690-826, for one chosen pixel.
template <class T>
static void march_pixel(const Frame<T>& f, const Volumes& vol, int px, int py, int pixelID, T rgba[4], T mask[4],
                        Counters& cnt, std::vector<LightRecord>* rec = nullptr,
                        std::vector<StepTrace>* trace = nullptr) {
  Region<T>::set(FR_RAY);
  T ux = (T)px / (T)f.W, uy = (T)py / (T)f.H;
  uy = T(1) - uy;
  v3<T> eye = V(-f.eye_ubo[0], -f.eye_ubo[1], -f.eye_ubo[2]);
  T jx, jy;
  jitter_march(f, pixelID, (T)f.W, (T)f.H, jx, jy);
  Ray<T> ray = cast_ray(f.view, f.tfx, f.tfy, ux, uy, eye, jx, jy);
  T d = dot(V(T(0), T(1), T(0)), ray.d);
  T bgMul = m_fmax(f.bg_min, d);
  for (int i = 0; i < 4; ++i) mask[i] = T(0);
  if (d < T(0)) {
    v3<T> c1 = V(T(0.0), T(0.16), T(0.51)) * T(0.4);
    v3<T> c2 = V(T(0.0), T(0.73), T(0.95)) * T(0.5);
    T a = -ray.d.y * T(5.5);
    rgba[0] = mixT(c1.x, c2.x, a); rgba[1] = mixT(c1.y, c2.y, a); rgba[2] = mixT(c1.z, c2.z, a);
    rgba[3] = T(1);
    cnt.rays_ocean++;
    return;
  }
  v3<T> bg = atmosphere_color(f, ray.d, f.sky_sun - ray.o, f.sun_intensity);
  bg = bg * bgMul;
  if (d < f.fade_start) {
    rgba[0] = bg.x; rgba[1] = bg.y; rgba[2] = bg.z; rgba[3] = T(1);
    cnt.rays_sky++;
    return;
  }
  cnt.rays_marched++;
  v3<T> c = eye;
  c.y = -f.earth_r;
  Isect<T> in = ray_sphere(ray.o, ray.d, c, f.r_in);
  Isect<T> out = ray_sphere(ray.o, ray.d, c, f.r_out);
  T accum = T(0);
  v3<T> col = ray_march(f, vol, ray, c, in.point, in.t, out.t, pixelID, accum, cnt, rec, trace);
  T godAccum = accum;
  accum = accum * smoothstepT(T(0), T(1), m_fmin(T(1), remap(ray.d.y, f.fade_start, f.fade_end, T(0), T(1))));
  rgba[0] = mixT(bg.x, col.x, accum);
  rgba[1] = mixT(bg.y, col.y, accum);
  rgba[2] = mixT(bg.z, col.z, accum);
  rgba[3] = T(1);
  T grey = T(25.0) * m_fmin(T(0.05), T(1) - godAccum);
  encode_float_rgba(grey, mask);
  // :775-779, reachable only when cloud_fade_start < 0.05
  if (ray.d.y < T(0.05)) {
    for (int i = 0; i < 4; ++i) mask[i] = mask[i] * m_fmax(T(5.0), ray.d.y);
  }
}

// main of reprojection.comp, :193-244, for one pixel.  idx (optional)
// receives the linear source index of every tap.
template <class T>
static void reproject_pixel(const Frame<T>& f, const T* oldview, const float* prev, int x, int y, float out[4],
                            int32_t* idx) {
  Region<T>::set(FR_REPROJ);
  T ux = (T)x / (T)f.W, uy = (T)y / (T)f.H;
  int pixelID = f.frame;
  v3<T> eye = V(-f.eye_ubo[0], -f.eye_ubo[1], -f.eye_ubo[2]);
  T jx, jy;
  jitter_reproj(f, pixelID, (T)f.W, (T)f.H, jx, jy);
  Ray<T> ray = cast_ray(f.view, f.tfx, f.tfy, ux, uy, eye, jx, jy);
  v3<T> c = eye;
  c.y = -f.earth_r;
  Isect<T> in = ray_sphere(ray.o, ray.d, c, f.r_in);
  const T* m = oldview;
  v3<T> P = in.point;
  // cameraOld.view * vec4(P, 1): sum of columns left to right
  v3<T> q = V(((m[0] * P.x + m[4] * P.y) + m[8] * P.z) + m[12] * T(1),
              ((m[1] * P.x + m[5] * P.y) + m[9] * P.z) + m[13] * T(1),
              ((m[2] * P.x + m[6] * P.y) + m[10] * P.z) + m[14] * T(1));
  v3<T> od = normalize(q);
  T nz = -od.z;
  od = od / nz;
  T ou = (od.x / f.tfx) * T(0.5) + T(0.5);
  T ov = (od.y / f.tfy) * T(0.5) + T(0.5);
  T mx = ou - ux, my = ov - uy;
  T accT[4] = {T(0), T(0), T(0), T(0)};
  const int N = f.blur_samples;
  for (int i = 0; i < N; ++i) {
    T fr = (T)i / (T)N;
    T bx = mx * fr, by = my * fr;
    // round(): ties to even (GLSL leaves ties to the implementation);
    // clamp(ivec2(...), 0, dim-1) done before the conversion so that
    // out-of-range values saturate instead of being undefined.
    T iu = m_nearbyint((ou - bx) * (T)f.W);
    T iv = m_nearbyint((ov - by) * (T)f.H);
    long long sx = (long long)m_fmin(m_fmax(iu, T(0)), (T)(f.W - 1));
    long long sy = (long long)m_fmin(m_fmax(iv, T(0)), (T)(f.H - 1));
    size_t li = (size_t)sy * (size_t)f.W + (size_t)sx;
    if (idx) idx[i] = (int32_t)li;
    if (prev) {
      for (int ch = 0; ch < 4; ++ch) accT[ch] += (T)prev[li * 4 + ch];
    }
  }
  for (int ch = 0; ch < 4; ++ch) out[ch] = (float)(accT[ch] / (T)N);
}

}  // namespace orc

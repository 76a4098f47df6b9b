// march_fast.hip -- optimised cloud ray-march kernels (kernel b, fast and
// precision variants) and the footprint-packed resources they read.
//
// Same algorithm as cloudRayMarch.comp (:565-826) and the plain kernel; the
// differences are all in how the work maps onto gfx950:
//
//  * Footprint-packed texels.  Every trilinear (bilinear) fetch of the
//    plain kernel issues 8 (4) dword loads.  Here each texel of a derived
//    volume stores its whole filter footprint in one entry, so every sample
//    touches one cache line: an LF sample is two dwordx4 loads from one
//    32-byte entry (the 2x2x2 footprint), a curl / weather / HF sample one
//    16-byte load.  Only the channels the density needs are kept, already
//    combined: LF -> (r, fbm = .625g + .25b + .125a) as half2 (filtering is
//    linear, so filtering the combined channel equals combining the
//    filtered channels, :505-507); HF -> fbm (:552-554); curl -> (r, g);
//    weather -> (coverage, type).  Built once per volume upload.
//  * Erosion hoisted out of the cone.  The six cone erosions of a step
//    read the curl and HF volumes at the primary step's skewed point with
//    the primary height (:665), so their high-frequency modifier equals
//    the primary erosion's; it is fetched once per cloud step and each
//    erosion is one remap.
//  * Loop invariants (length(start - eye), the step's normalisation,
//    reciprocals of the constant denominators) are hoisted; the single
//    exp of GetLightEnergy is shared by its two identical terms; pow runs
//    as exp2(y * log2(x)) on the transcendental unit.
//  * t_k = start + k * step is evaluated directly (one rounding) instead of
//    by repeated addition.
//  * Variant 1 (precision): filtering in packed FP16 (v_pk_fma_f16), FP16
//    volumes, the shell quadratic in FP64; transmittance and radiance stay
//    FP32.
// Multiply-adds within one expression contract to fma in this file (the
// library builds with -ffp-contract=off so that the plain kernel keeps the
// shader's one rounding per operation); the optimised kernels are held to
// the FP64 oracle, which fma approaches more closely, not to bit equality.
// Contraction stays within an expression ("on", not "fast"): across
// statements it would depend on what the compiler sees between them (a
// lane shuffle, or none when one lane marches a ray), and a ray's bits
// must not depend on how many lanes share it.
#pragma clang fp contract(on)
#include "march_common.h"
#include "mlp_common.h"
#include "reproject_pixel.h"

typedef _Float16 h2v __attribute__((ext_vector_type(2)));

namespace mtc {
namespace fast {

__device__ __forceinline__ h2v as_h2(uint32_t v) { return __builtin_bit_cast(h2v, v); }
__device__ __forceinline__ uint32_t pack_h2(float a, float b) {
  h2v h = {(_Float16)a, (_Float16)b};
  return __builtin_bit_cast(uint32_t, h);
}
__device__ __forceinline__ float2 f2(uint32_t v) {
  h2v h = as_h2(v);
  return make_float2((float)h.x, (float)h.y);
}
__device__ __forceinline__ float2 lerp2(float2 p, float2 q, float w) {
  return make_float2(__fmaf_rn(w, q.x - p.x, p.x), __fmaf_rn(w, q.y - p.y, p.y));
}
__device__ __forceinline__ h2v lerp2h(h2v p, h2v q, h2v w) { return __builtin_elementwise_fma(w, q - p, p); }

// 16-byte entry i of an array as a 32-bit unsigned byte offset from the
// array base, so the load takes the base from scalar registers and needs
// no per-lane 64-bit address arithmetic.  Footprint arrays stay within
// 4 GiB: volumes up to 512^3 (2^27 LF entries of 32 bytes; mtc_create
// rejects larger ones for these kernels), maps up to 8192^2.
__device__ __forceinline__ uint4 entry(const uint4* __restrict__ q, uint32_t i) {
  return *(const uint4*)((const char*)q + (i << 4));
}

struct Coord {
  uint32_t i0;   // linear index of the (x0, y0, z0) texel
  float ax, ay, az;
  uint32_t x0, y0, z0;
};

// s * N - 0.5 as one fma: s * N is exact (N a power of two), so this is
// the same single rounding as the multiply-then-subtract form.
__device__ __forceinline__ Coord coord3(int sh, f3 s) {
  const float N = (float)(1 << sh);
  const uint32_t m = (1u << sh) - 1u;
  float ux = __fmaf_rn(s.x, N, -0.5f), uy = __fmaf_rn(s.y, N, -0.5f), uz = __fmaf_rn(s.z, N, -0.5f);
  float fx = floorf(ux), fy = floorf(uy), fz = floorf(uz);
  Coord c;
  c.ax = ux - fx;
  c.ay = uy - fy;
  c.az = uz - fz;
  const uint32_t x0 = (uint32_t)(int)fx & m, y0 = (uint32_t)(int)fy & m, z0 = (uint32_t)(int)fz & m;
  c.i0 = (((z0 << sh) + y0) << sh) + x0;
  c.x0 = x0;
  c.y0 = y0;
  c.z0 = z0;
  return c;
}

// LF: (r, fbm) trilinear from the texel's 32-byte footprint entry (two
// 16-byte halves: the z0 and z1 slices of the 2x2x2 footprint).  One cache
// line per sample instead of two measured 17% faster for the whole march
// than the 16-byte 2x2 (x, y) entries, although the table doubles (64 MiB
// at 128^3): the march is limited by L2 request throughput more than by
// capacity.
//
// filter_lf: the trilinear of one footprint, a = the z0 slice's texels
// (x0y0, x1y0, x0y1, x1y1), b = the z1 slice's, each a half2 (r, fbm).
template <bool F16>
__device__ __forceinline__ float2 filter_lf(const Coord& c, uint4 a, uint4 b) {
  if (F16) {
    h2v wx = {(_Float16)c.ax, (_Float16)c.ax}, wy = {(_Float16)c.ay, (_Float16)c.ay},
        wz = {(_Float16)c.az, (_Float16)c.az};
    h2v a0 = lerp2h(as_h2(a.x), as_h2(a.y), wx), a1 = lerp2h(as_h2(a.z), as_h2(a.w), wx);
    h2v b0 = lerp2h(as_h2(b.x), as_h2(b.y), wx), b1 = lerp2h(as_h2(b.z), as_h2(b.w), wx);
    h2v r = lerp2h(lerp2h(a0, a1, wy), lerp2h(b0, b1, wy), wz);
    return make_float2((float)r.x, (float)r.y);
  }
  float2 a0 = lerp2(f2(a.x), f2(a.y), c.ax), a1 = lerp2(f2(a.z), f2(a.w), c.ax);
  float2 b0 = lerp2(f2(b.x), f2(b.y), c.ax), b1 = lerp2(f2(b.z), f2(b.w), c.ax);
  return lerp2(lerp2(a0, a1, c.ay), lerp2(b0, b1, c.ay), c.az);
}

template <bool F16>
__device__ __forceinline__ float2 sample_lf(const uint4* __restrict__ q, int sh, f3 s) {
  const Coord c = coord3(sh, s);
  return filter_lf<F16>(c, entry(q, 2u * c.i0), entry(q, 2u * c.i0 + 1u));
}

// HF: single channel, the full 2x2x2 footprint in one entry
template <bool F16>
__device__ __forceinline__ float sample_hf(const uint4* __restrict__ q, int sh, f3 s) {
  const Coord c = coord3(sh, s);
  const float ax = c.ax, ay = c.ay, az = c.az;
  uint4 e = entry(q, c.i0);  // (x0y0z0,x1y0z0) (x0y1z0,x1y1z0) (x0y0z1,x1y0z1) (x0y1z1,x1y1z1)
  float2 p0 = f2(e.x), p1 = f2(e.y), p2 = f2(e.z), p3 = f2(e.w);
  float c0 = __fmaf_rn(ax, p0.y - p0.x, p0.x), c1 = __fmaf_rn(ax, p1.y - p1.x, p1.x);
  float c2 = __fmaf_rn(ax, p2.y - p2.x, p2.x), c3 = __fmaf_rn(ax, p3.y - p3.x, p3.x);
  float d0 = __fmaf_rn(ay, c1 - c0, c0), d1 = __fmaf_rn(ay, c3 - c2, c2);
  return __fmaf_rn(az, d1 - d0, d0);
}

// 2D maps: two channels, the 2x2 footprint in one entry, sampled at
// (sx, sy) * scale (maps are at most 8192^2 entries, 1 GiB)
template <bool F16>
__device__ __forceinline__ float2 sample_map(const uint4* __restrict__ q, int sh, float sx, float sy, float scale) {
  const float sN = scale * (float)(1 << sh);
  const uint32_t m = (1u << sh) - 1u;
  float ux = __fmaf_rn(sx, sN, -0.5f), uy = __fmaf_rn(sy, sN, -0.5f);
  float fx = floorf(ux), fy = floorf(uy);
  float ax = ux - fx, ay = uy - fy;
  uint4 e = entry(q, ((((uint32_t)(int)fy & m) << sh) + ((uint32_t)(int)fx & m)));
  if (F16) {
    h2v wx = {(_Float16)ax, (_Float16)ax}, wy = {(_Float16)ay, (_Float16)ay};
    h2v r = lerp2h(lerp2h(as_h2(e.x), as_h2(e.y), wx), lerp2h(as_h2(e.z), as_h2(e.w), wx), wy);
    return make_float2((float)r.x, (float)r.y);
  }
  return lerp2(lerp2(f2(e.x), f2(e.y), ax), lerp2(f2(e.z), f2(e.w), ax), ay);
}

__device__ __forceinline__ float fast_pow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
__device__ __forceinline__ float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// castRay (:194-226) with the camera basis and eye + look taken from the
// per-frame constants (the host computes them with the same FP32
// operations, so the ray is bit-identical to the per-ray form).
__device__ __forceinline__ Ray cast_ray_folded(const FrameConsts& fc, float ux, float uy, float jx, float jy) {
  float nx = ux * 2.0f - 1.0f, ny = uy * 2.0f - 1.0f;
  nx = nx + jx;
  ny = ny + jy;
  const f3 cx = (nx * fc.tfx) * ld3(fc.rp_right);
  const f3 cy = (ny * fc.tfy) * ld3(fc.rp_up);
  const f3 eye = ld3(fc.eye);
  const f3 p = (ld3(fc.rp_ref) + cx) + cy;
  Ray r;
  r.o = eye;
  r.d = normalize(p - eye);
  return r;
}

// HenyeyGreenstein (:295-300) with b^1.5 as b sqrt(b) and a reciprocal.
__device__ __forceinline__ float hg_fast(float c, float g) {
  const float b = 1.0f + g * g - 2.0f * g * c;
  return (1.0f - g * g) * rcp(b * __builtin_amdgcn_sqrtf(b)) * 0.07957747154594767f;
}

// getAtmosphereColorPhysical (:427-467) on the transcendental unit:
// cos(acos(y)) = y, pow(x, 1.5) = x sqrt(x), pow(x, 0.5) = sqrt(x), the
// zenith pow and the three exps as exp2/log2, divisions by per-frame
// constants as reciprocals; the shader's operations otherwise.  The IEEE
// library pow alone costs ~165 instructions and the sky runs once per ray,
// i.e. once per wave for all of its rays.
__device__ __forceinline__ f3 atmosphere_fast(const FrameConsts& fc, f3 dir) {
  const f3 BR = ld3(fc.beta_r), BM = ld3(fc.beta_m), BS = ld3(fc.beta_sum);
  const f3 sunDir = ld3(fc.sky_sun_dir);
  const float sunE = fc.sunE;
  const float y = fmaxf(0.0f, dir.y);
  const float zenith = acosf(y);
  const float inv = rcp(y + 0.15f * fast_pow(93.885f - zenith * (180.0f / 3.14159265f), -1.253f));
  const float sR = 8.4E3f * inv, sM = 1.25E3f * inv;
  const float cosTheta = dot(sunDir, dir);
  const float cc = cosTheta * 0.5f + 0.5f;
  const float rPhase = 0.05968310365946075f * (1.0f + cc * cc);
  const float mPhase = hg_fast(cosTheta, 0.8f);
  const float yc = fc.ydot_clamped;
  const float disk = sunE * 15000.0f * smoothstepf(fc.sun_disk_lo, fc.sun_disk_hi, cosTheta);
  const float br[3] = {BR.x, BR.y, BR.z}, bm[3] = {BM.x, BM.y, BM.z}, bs[3] = {BS.x, BS.y, BS.z};
  const float add[3] = {0.0f, 0.0003f, 0.00075f};
  float o[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float fex = fast_exp(-br[c] * sR + bm[c] * sM);
    const float sb = sunE * ((br[c] * rPhase + bm[c] * mPhase) * rcp(bs[c]));
    const float a = sb * (1.0f - fex), b = sb * fex;
    const float lin = a * __builtin_amdgcn_sqrtf(a) * mixf(1.0f, __builtin_amdgcn_sqrtf(b), yc);
    o[c] = (lin + (0.1f * fex + disk * fex)) * 0.04f + add[c];
  }
  return mk(o[0], o[1], o[2]);
}

// Weather mode: the height-gradient weight of a sample, mix(a, b, type)
// * 0.5 (getDensityHeightGradientForPoint :475-487 and the x0.5 of the
// weather path), from the step's two profiles sc, st.
//
// For type t in [0, 1] the shader's three mixes reduce to one quadratic
// per half, with D = sc - st:
//   t <= 0.5:  0.5 st + D t (1.5 - t)
//   t >  0.5:  0.5 sc + D t (0.5 - t)
// (both 0.5 sc at t = 0.5; expanding mix(mix(st, sc, 2t), sc, t) and
// mix(sc, mix(sc, st, 2t - 1), t)).  Seven operations instead of fifteen.
__device__ __forceinline__ float gradient_weight(float type, float sc, float st) {
  const bool lo = type <= 0.5f;
  const float c = (lo ? st : sc) * 0.5f, k = lo ? 1.5f : 0.5f;
  return __fmaf_rn((sc - st) * type, k - type, c);
}

// Density of one LF sample (sampleLowFrequency, :499-540) with coverage
// `cov`; in weather mode `gw` is the sample's gradient_weight.
//
// Exact empty-space test: remapClamped(...) is at most 1, so the weighted
// base is at most gw, and remapClampedBeforeAndAfter(base, cov, 1, 0, 1)
// is 0 whenever base <= cov.  So gw <= cov means density 0 without the LF
// sample; the callers skip the fetch then (bit-identical result).
template <int COV>
__device__ __forceinline__ float density(float2 lf, float cov, float gw) {
  float fbm = clampf(lf.y, 0.0f, 1.0f);
  // remapClamped(r, fbm - 0.9, 1, 0, 1)
  float base = clampf((lf.x - (fbm - 0.9f)) * rcp(1.9f - fbm), 0.0f, 1.0f);
  if (COV == MTC_COVERAGE_WEATHER) base = base * gw;
  // remapClampedBeforeAndAfter(base, cov, 1, 0, 1) * cov
  float v = clampf(base, cov, 1.0f);
  return clampf((v - cov) * rcp(1.0f - cov), 0.0f, 1.0f) * cov;
}

struct FastViews {
  const uint4* lf;
  const uint4* hf;
  const uint4* curl;
  const uint4* weather;
  int lf_sh, hf_sh, curl_sh, weather_sh;
  // learned light (MTC_KERNEL_LEARNED): network and its error calibration
  const bf16x8* mlp_frag;
  const float* mlp_params;
  float calib_hi, calib_inv;  // calib_bin(y, hi, inv): 16 bins
  unsigned int calib_mask;    // bins whose held-out error is too high
  float fallback_density;     // local density above which the cone runs
  int audit;                  // also cone-march one in `audit` answered steps (power of two, 0 = off)
  // LDS-staged LF bricks (MTC_FLAG_LDS_BRICKS): the LF's half2 (r, fbm)
  // texels in 4^3 bricks of 256 bytes (pack_lf_brick_kernel)
  const uint4* lf_brick;
};

// Shell entry/exit of a ray (raySphereIntersection, :229-273, with the
// t = length(point_world - rO_normalised) quirk of :269).
template <bool FP64SHELL>
__device__ __forceinline__ void shell_span(const FrameConsts& fc, const Ray& ray, f3 c, float& start_t, float& end_t,
                                           f3& startPos) {
  if (FP64SHELL) {
    // FP64: C = dot(rO, rO) - 1 cancels in FP32 at |rO| ~ 0.9998
    // (radius ~6.4e6 m)
    double ox = ray.o.x, oy = ray.o.y, oz = ray.o.z;
    double dx = ray.d.x, dy = ray.d.y, dz = ray.d.z;
    double cx = c.x, cy = -fc.earth_r_d, cz = c.z;
    double ts[2] = {0.0, 0.0};
    double sp[3] = {0.0, 0.0, 0.0};
    // divisions as products with one reciprocal each (1 ulp in FP64)
    const double A = dx * dx + dy * dy + dz * dz;
    const double inv2A = 1.0 / (2.0 * A);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const double R = k == 0 ? fc.r_in_d : fc.r_out_d;
      const double iR = 1.0 / R;
      double rx0 = (ox - cx) * iR, ry0 = (oy - cy) * iR, rz0 = (oz - cz) * iR;
      double B = 2.0 * (dx * rx0 + dy * ry0 + dz * rz0);
      double C = (rx0 * rx0 + ry0 * ry0 + rz0 * rz0) - 1.0;
      double disc = B * B - 4.0 * A * C;
      double t = -1.0;
      if (disc >= 0.0) {
        double sq = sqrt(disc);
        t = (-B - sq) * inv2A;
        if (t < 0.0) t = (-B + sq) * inv2A;
      }
      if (t >= 0.0) {
        double pxw = (rx0 + t * dx) * R + cx, pyw = (ry0 + t * dy) * R + cy, pzw = (rz0 + t * dz) * R + cz;
        double ex = pxw - rx0, ey = pyw - ry0, ez = pzw - rz0;
        ts[k] = sqrt(ex * ex + ey * ey + ez * ez);
        if (k == 0) {
          sp[0] = pxw;
          sp[1] = pyw;
          sp[2] = pzw;
        }
      }
    }
    start_t = (float)ts[0];
    end_t = (float)ts[1];
    startPos = mk((float)sp[0], (float)sp[1], (float)sp[2]);
  } else {
    Isect in = ray_sphere(ray.o, ray.d, c, fc.r_in);
    Isect out = ray_sphere(ray.o, ray.d, c, fc.r_out);
    start_t = in.t;
    end_t = out.t;
    startPos = in.point;
  }
}

// Everything a step needs that is constant along the ray.
// Per-ray state of rayMarch.  Everything the same for all rays of a frame
// (eye, planet centre, shell offsets, wind, 1 / thickness, ...) is read
// from the frame constants where it is used, so it stays in scalar
// registers; values assigned in the divergent per-ray set-up would each
// take a vector register.
struct RayMarch {
  f3 d;             // ray direction (the origin is the eye)
  float start_t, end_t, stepSize, HG_light, ls, cone_scale;
  float cos_ratio;  // (cos_angle - 0.7) / (1 - 0.7) of GetLightEnergy's first remap
};

// One primary step as seen by its lane, before the cone march.
struct Prim {
  bool valid;  // t_k < end_t and k < cap
  bool cloud;  // base density > 0
  f3 rel;      // getRelativePositionInAtmosphere(pos) (cone samples start here)
  float h;     // relative height
  float base;  // base density * base_density_factor
  float sc, st;     // height-gradient profiles of the step
  float m, inv1m;   // erosion: high_freq_modifier * 0.005 and 1 / (1 - m)
  float da;         // accumulated-density increment of the step
};

// What eval_primary_pre hands to the LF fetch and eval_primary_post: the
// skewed sample point and the coverage terms of a step that needs its LF
// sample (need_lf false: the step's density is known to be 0 without it).
struct PrimLf {
  bool need_lf;
  f3 sk;
  float cov, gw;
};

// A primary step (:627-650) up to its LF fetch.
template <int COV, bool F16>
__device__ __forceinline__ Prim eval_primary_pre(const FrameConsts& fc, const FastViews& fv, const float* sjit,
                                                 const RayMarch& rm, int k, PrimLf& pl) {
  Prim o;
  o.cloud = false;
  o.da = 0.0f;
  o.m = 0.0f;
  o.inv1m = 1.0f;
  o.sc = o.st = 0.0f;
  o.h = 0.0f;
  o.base = 0.0f;
  o.rel = mk(0.0f, 0.0f, 0.0f);
  pl.need_lf = false;
  pl.sk = mk(0.0f, 0.0f, 0.0f);
  const float t = __fmaf_rn((float)k, rm.stepSize, rm.start_t);
  o.valid = (t < rm.end_t) && (k < fc.mk_cap);
  if (!o.valid) return o;
  const int ji = ((fc.frame + (int)t) & 15) * 3;
  const f3 dj = mk(rm.d.x + sjit[ji], rm.d.y + sjit[ji + 1], rm.d.z + sjit[ji + 2]);
  // pos = eye + t dj; the shader's differences pos - eye, pos - inner
  // (inner = (eye.x, inner_off, eye.z)) and pos - c (c = (eye.x, -R, eye.z))
  // from t dj and two per-frame offsets
  const f3 pe = t * dj;
  const float it = fc.mk_inv_thick;
  const f3 rel = mk(pe.x * it, (pe.y + fc.mk_rel_y0) * it, pe.z * it);
  const f3 sp = rel * fc.mk_inv_div;
  const f3 pc = mk(pe.x, pe.y + fc.mk_c_y0, pe.z);
  const float lr = __builtin_amdgcn_sqrtf(dot(pe, pe));  // v_sqrt_f32 (1 ulp); the plain kernel keeps IEEE
  const float cosT = dot(rm.d, pc) * __builtin_amdgcn_rsqf(dot(pc, pc));
  const float h = fabsf(cosT * (lr - rm.ls)) * it;
  const f3 sk = sp + h * ld3(fc.mk_wk) + ld3(fc.wind_offset);
  float sc = 0.0f, st = 0.0f;
  if (COV == MTC_COVERAGE_WEATHER) {
    // getDensityHeightGradientForPoint profiles (:479-481)
    float hc = clampf(h, 0.0f, 1.0f);
    sc = fmaxf(0.0f, (hc * 4.0f) * ((hc - 0.65f) * (1.0f / -0.35f)));
    st = fmaxf(0.0f, (hc * 10.0f) * ((hc - 0.3f) * (1.0f / -0.1f)));
  }
  o.rel = rel;
  o.h = h;
  o.sc = sc;
  o.st = st;
  // density, skipping the fetches whose result is known to be 0: outside
  // the gradient's support (h >= 0.65 or h <= 0, no weather fetch either)
  // and where the coverage bounds the weighted base (no LF fetch)
  float cov = fc.coverage_const, gw = 1.0f;
  if (COV == MTC_COVERAGE_WEATHER) {
    if (sc == 0.0f && st == 0.0f) return o;
    const float2 w = sample_map<F16>(fv.weather, fv.weather_sh, sp.x, sp.z, fc.weather_scale);
    cov = w.x;
    gw = gradient_weight(w.y, sc, st);
    if (gw <= cov) return o;
  }
  pl.need_lf = true;
  pl.sk = sk;
  pl.cov = cov;
  pl.gw = gw;
  return o;
}

// The rest of a primary step that needed its LF sample `lf`: density,
// and the erosion of an in-cloud step.
template <int COV, bool F16>
__device__ __forceinline__ void eval_primary_post(const FrameConsts& fc, const FastViews& fv, const PrimLf& pl,
                                                  float2 lf, Prim& o) {
  const float base = density<COV>(lf, pl.cov, pl.gw) * fc.base_density_factor;
  o.base = base;
  if (base > 0.0f) {
    o.cloud = true;
    // erodeCloudWithHighFrequency at (sk, h): one curl + HF fetch serves the
    // primary erosion and all cone erosions of the step (:645, :665)
    const f3 sk = pl.sk;
    const float h = o.h;
    float2 cn = sample_map<F16>(fv.curl, fv.curl_sh, sk.x, sk.y, 1.0f);
    float kk = (1.0f - h) * 0.5f;
    f3 p2 = mk(sk.x + cn.x * kk, sk.y + cn.y * kk, sk.z);
    float fh = sample_hf<F16>(fv.hf, fv.hf_sh, p2);
    float mod = clampf(mixf(fh, 1.0f - fh, clampf(h * 2.0f, 0.0f, 1.0f)), 0.0f, 1.0f);
    o.m = mod * 0.005f;
    o.inv1m = rcp(1.0f - o.m);
    o.da = ((base * fc.erosion_scale - o.m) * o.inv1m) * fc.accum_scale;
  }
}

template <int COV, bool F16>
__device__ __forceinline__ Prim eval_primary(const FrameConsts& fc, const FastViews& fv, const float* sjit,
                                             const RayMarch& rm, int k) {
  PrimLf pl;
  Prim o = eval_primary_pre<COV, F16>(fc, fv, sjit, rm, k, pl);
  if (pl.need_lf) eval_primary_post<COV, F16>(fc, fv, pl, sample_lf<F16>(fv.lf, fv.lf_sh, pl.sk), o);
  return o;
}

// ---------------------------------------------------------------------------
// LDS-staged LF bricks (MTC_FLAG_LDS_BRICKS)
//
// The primary steps of one wave (8 rays x 8 consecutive steps at lanes per
// ray 8) sample the LF at the skewed point, one eighth of the cone's
// frequency (:635), so on the steeper rays they stay within a few texels of
// each other for several chunks.  Each wave keeps an 8^3-texel box of the
// LF's half2 (r, fbm) texels in LDS (2 KiB), origin aligned to 4 texels,
// staged from the 4^3-bricked copy of the volume: the box is two bricks on
// each axis, and each of the 64 lanes loads 32 contiguous bytes of it.  A
// sample whose 2x2x2 footprint lies in the box (base texel within 0..6 of
// the origin on every axis, REPEAT wrap as a mask) reads its eight texels
// from LDS (ds_read2 pairs); the others read the footprint table.  Both
// hold the same half2 values and feed the same filter_lf, so the results
// are bit-identical to the table path.  The box moves (re-staged around the
// first sample that misses it) only when that serves at least
// kBrickMinGain more of the chunk's samples than the current box does.
// Cone samples (full frequency, scattered) stay on the table path.
// ---------------------------------------------------------------------------
constexpr int kBrickMinGain = 8;

struct BrickBox {
  int ox, oy, oz;  // texel origin of the staged box (multiples of 4); wave-uniform
  bool valid;
};

// Stage the box at (ox, oy, oz) into `box` (512 half2 texels, x fastest,
// then y, then z): lane l loads piece l & 7 (two 4-texel rows) of brick
// l >> 3 of the 2 x 2 x 2 bricks the box spans.
__device__ __forceinline__ void stage_brick_box(const uint4* __restrict__ bricks, int sh, int ox, int oy, int oz,
                                                uint32_t* __restrict__ box) {
  const int lane = threadIdx.x & 63;
  const int nbsh = sh - 2;
  const uint32_t nbm = (1u << nbsh) - 1u;
  const int b = lane >> 3, p = lane & 7;
  const uint32_t bx = ((uint32_t)(ox >> 2) + (uint32_t)(b & 1)) & nbm;
  const uint32_t by = ((uint32_t)(oy >> 2) + (uint32_t)((b >> 1) & 1)) & nbm;
  const uint32_t bz = ((uint32_t)(oz >> 2) + (uint32_t)(b >> 2)) & nbm;
  const uint32_t brick = (((bz << nbsh) + by) << nbsh) + bx;
  // brick texel (x, y, z) at x + 4 y + 16 z: piece p = rows (y, y + 1) of
  // slice z, y = 2 (p & 1), z = p >> 1
  const uint4 r0 = entry(bricks, brick * 16u + 2u * (uint32_t)p);
  const uint4 r1 = entry(bricks, brick * 16u + 2u * (uint32_t)p + 1u);
  const int zl = 4 * (b >> 2) + (p >> 1), yl = 4 * ((b >> 1) & 1) + 2 * (p & 1), xl = 4 * (b & 1);
  *(uint4*)(box + xl + 8 * yl + 64 * zl) = r0;
  *(uint4*)(box + xl + 8 * (yl + 1) + 64 * zl) = r1;
}

// The LF sample of every lane with `need` set, through the wave's LDS box
// where the footprint fits, else through the footprint table.  Called by
// all 64 lanes of the wave together (ballots, staging); lanes without
// `need` return an unused value.  `staged` / `served` count box stagings
// and samples read from LDS (counting builds).
template <bool F16>
__device__ __forceinline__ float2 brick_sample_lf(const FastViews& fv, uint32_t* __restrict__ box, BrickBox& bb,
                                                 bool need, f3 sk, unsigned long long& staged,
                                                 unsigned long long& served) {
  const unsigned long long needb = __ballot(need);
  if (!needb) return make_float2(0.0f, 0.0f);  // no LF sample in this chunk (most chunks)
  const int sh = fv.lf_sh;
  const uint32_t m = (1u << sh) - 1u;
  const Coord c = coord3(sh, sk);
  // offset of the footprint's base texel in a box at (ox, oy, oz); true when
  // the whole 2x2x2 footprint lies in the box
  auto in_box = [&](int ox, int oy, int oz, uint32_t& off) {
    const uint32_t dx = (c.x0 - (uint32_t)ox) & m, dy = (c.y0 - (uint32_t)oy) & m, dz = (c.z0 - (uint32_t)oz) & m;
    off = dx + 8u * dy + 64u * dz;
    return dx < 7u && dy < 7u && dz < 7u;
  };
  uint32_t off = 0u;
  bool inb = need && bb.valid && in_box(bb.ox, bb.oy, bb.oz, off);
  const unsigned long long inbb = __ballot(inb);
  const unsigned long long miss = needb & ~inbb;
  if (miss) {
    // candidate box: one texel behind the first missing sample, aligned to 4
    const int l = __builtin_ctzll(miss);
    const uint32_t rx = (uint32_t)__builtin_amdgcn_readlane((int)c.x0, l);
    const uint32_t ry = (uint32_t)__builtin_amdgcn_readlane((int)c.y0, l);
    const uint32_t rz = (uint32_t)__builtin_amdgcn_readlane((int)c.z0, l);
    const int cx = (int)(((rx - 1u) & m) & ~3u), cy = (int)(((ry - 1u) & m) & ~3u), cz = (int)(((rz - 1u) & m) & ~3u);
    uint32_t off2 = 0u;
    const bool fit = need && in_box(cx, cy, cz, off2);
    const unsigned long long fitb = __ballot(fit);
    if (__popcll(fitb) >= __popcll(inbb) + kBrickMinGain) {
      // the previous chunk's reads of the box are done before it is rewritten
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      stage_brick_box(fv.lf_brick, sh, cx, cy, cz, box);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      bb.ox = cx;
      bb.oy = cy;
      bb.oz = cz;
      bb.valid = true;
      off = off2;
      inb = fit;
      if ((threadIdx.x & 63) == 0) staged++;
    }
  }
  uint4 a = make_uint4(0u, 0u, 0u, 0u), b = a;
  if (inb) {
    const uint32_t* L = box + off;
    a = make_uint4(L[0], L[1], L[8], L[9]);
    b = make_uint4(L[64], L[65], L[72], L[73]);
    served++;
  } else if (need) {
    a = entry(fv.lf, 2u * c.i0);
    b = entry(fv.lf, 2u * c.i0 + 1u);
  }
  return filter_lf<F16>(c, a, b);
}

// Per-wave LDS work area of the compacted cone march.
struct ConeLds {
  float4 own[64][2];     // per in-cloud step (by rank): rel.xyz, cone_scale | sc, st, m, inv1m
  float contrib[6 * 64]; // one erosion result per cone sample
  unsigned char hit[6 * 64];
};

// G lanes march one ray: lane j of the group evaluates step k0 + j of each
// chunk of G steps.  The cone samples of all in-cloud steps of the wave are
// then compacted: a ballot ranks the in-cloud lanes, each publishes its
// step's data to LDS, and all 64 lanes process the (steps x samples) work
// list in ceil(n * samples / 64) rounds; each owner sums its samples back.
// Finally the group replays the accumulation / early-exit recurrence of
// :674-684 over the chunk in step order (identical values in all G lanes),
// so the result equals the sequential march.  The loop is wave-uniform: a
// lane whose ray has finished still takes its share of the cone work.
// Per-ray constants of rayMarch (:565-624) for a ray that is marched.
template <bool FP64SHELL>
__device__ __forceinline__ void setup_march(const FrameConsts& fc, const Ray& ray, RayMarch& rm) {
  const float d = ray.d.y;
  rm.d = ray.d;
  const f3 c = mk(ray.o.x, -fc.earth_r, ray.o.z);
  f3 startPos;
  shell_span<FP64SHELL>(fc, ray, c, rm.start_t, rm.end_t, startPos);
  float maxSteps = (fc.step_mode == MTC_STEPS_FIXED)
                       ? (float)fc.steps_max
                       : floorf(mixf((float)fc.steps_min, (float)fc.steps_max, 1.0f - d));
  rm.stepSize = (rm.end_t - rm.start_t) / maxSteps;
  const float cos_angle = dot(ray.d * __builtin_amdgcn_rsqf(dot(ray.d, ray.d)), ld3(fc.light_dir));
  rm.HG_light = fmaxf(hg_fast(cos_angle, fc.hg_g), fc.hg_si * hg_fast(cos_angle, fc.hg_g2));
  rm.cos_ratio = (cos_angle - 0.7f) / (1.0f - 0.7f);
  rm.ls = length(startPos - ray.o);
  rm.cone_scale = rm.stepSize * fc.mk_inv_thick;
}

// One density term of the cone light march (:348-369) for sample i of an
// in-cloud step: density at rel + kernel_i * i * cone_scale (unskewed,
// undivided), eroded with the step's own high-frequency modifier.
template <int COV>
__device__ __forceinline__ float cone_term(const FrameConsts& fc, const FastViews& fv, const float* skern,
                                          const Prim& pr, float cone_scale, int i) {
  const float s = (float)i * cone_scale;
  const f3 slp = mk(pr.rel.x + skern[3 * i] * s, pr.rel.y + skern[3 * i + 1] * s, pr.rel.z + skern[3 * i + 2] * s);
  float cov = fc.coverage_const, gw = 1.0f;
  if (COV == MTC_COVERAGE_WEATHER) {
    const float2 cw = sample_map<false>(fv.weather, fv.weather_sh, slp.x, slp.z, fc.weather_scale);
    cov = cw.x;
    gw = gradient_weight(cw.y, pr.sc, pr.st);
  }
  const float cb = gw <= cov ? 0.0f : density<COV>(sample_lf<false>(fv.lf, fv.lf_sh, slp), cov, gw);
  return cb > 0.0f ? (fc.cone_erosion_scale * cb - pr.m) * pr.inv1m : 0.0f;
}

// The network's nine inputs for one in-cloud step (learned light):
//   0-2  position in the shell, as the cone sees it: the cone samples the
//        LF volume at rel + offsets (period 1) and the weather map at
//        (rel.x, rel.z) * weather_scale, so x and z are taken modulo the
//        weather tile, y modulo the LF tile
//   3    height fraction
//   4    local density as the light march sees it: the cone's first term
//        (its sample at the step itself), computed exactly
//   5-7  sun direction
//   8    cos of the view ray's zenith angle: with the camera in the
//        atmosphere it fixes the shell span, hence the primary step and the
//        cone's length (step / thickness)
// The network answers the rest of the cone: -(terms 1 .. n-1).
__device__ __forceinline__ void mlp_features(const FrameConsts& fc, const Prim& pr, const RayMarch& rm, float c0,
                                             float f[9]) {
  const float ax = pr.rel.x * fc.weather_scale, az = pr.rel.z * fc.weather_scale;
  f[0] = ax - floorf(ax);
  f[1] = pr.rel.y - floorf(pr.rel.y);
  f[2] = az - floorf(az);
  f[3] = pr.h;
  f[4] = c0;
  f[5] = fc.light_dir[0];
  f[6] = fc.light_dir[1];
  f[7] = fc.light_dir[2];
  f[8] = rm.d.y;
}

__device__ __forceinline__ uint32_t step_hash(uint32_t a, uint32_t b, uint32_t c, uint32_t seed) {
  uint32_t h = a * 0x9E3779B1u ^ b * 0x85EBCA77u ^ c * 0xC2B2AE3Du ^ seed * 0x27D4EB2Fu;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return h;
}

// PAIR: cone rounds two at a time (both rounds' weather fetches, then both
// LF fetches): two dependent round trips per pair instead of four, for ~18
// more VGPRs (the frame kernel runs 5 waves per SIMD with it, faster than
// 6 without; the batch and learned kernels keep single rounds).
// BRICK: primary LF samples through the wave's LDS box (brick_sample_lf).
template <int COV, bool F16, bool FP64SHELL, bool COUNT, int G, bool LEARN = false, bool PAIR = false,
          bool BRICK = false>
__device__ __forceinline__ void march_tile(const FrameConsts& fc, const FastViews& fv, float4* __restrict__ out_color,
                                           float4* __restrict__ out_mask, unsigned long long* counters, int full,
                                           int rx0, int ry0, int tile_x, const float* sjit, const float* skern,
                                           ConeLds* cone_lds, uint16_t* mlp_lds = nullptr,
                                           uint32_t* lf_box = nullptr) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int slot = threadIdx.x / G;  // ray slot in the tile
  const int j = threadIdx.x % G;     // step offset within a chunk
  const int gbase = lane - j;        // first lane of this ray's group
  const int rx = rx0 + (slot % tile_x);
  const int ry = ry0 + (slot / tile_x);
  const int W = fc.width, H = fc.height;
  const int pixelID = fc.frame;
  const int px = full ? rx : rx * 4 + pixelID / 4;
  const int py = full ? ry : ry * 4 + pixelID % 4;
  const bool active = px < W && py < H;
  const int nls = fc.light_samples;
  ConeLds& L = cone_lds[wave];
  // cone work map: each round covers cgroups = 64 / nls steps; lane l does
  // sample ci = l % nls of slot cs = l / nls, its kernel vector in registers
  const int cgroups = nls > 0 ? 64 / nls : 1;
  const int ci = nls > 0 ? lane % nls : 0, cs = nls > 0 ? lane / nls : 64;
  const bool cone_lane = nls > 0 && cs < cgroups;
  const float fci = (float)ci;
  const float kx = cone_lane ? skern[3 * ci] : 0.0f, ky = cone_lane ? skern[3 * ci + 1] : 0.0f,
              kz = cone_lane ? skern[3 * ci + 2] : 0.0f;
  unsigned long long n_steps = 0, n_cloud = 0, n_cone = 0, n_cone_er = 0, n_early = 0;
  unsigned long long n_ocean = 0, n_sky = 0, n_marched = 0, n_mlp = 0, n_mlp_fb = 0, n_audit = 0;
  unsigned long long n_bstage = 0, n_bserved = 0;
  float audit_sq = 0.0f;
  BrickBox bb{0, 0, 0, false};

  // --- per-ray set-up (main, :690-757) --------------------------------------
  RayMarch rm;
  bool gdone = true;
  f3 bg = mk(0.0f, 0.0f, 0.0f);
  float4 color = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  Ray ray;
  ray.o = ld3(fc.eye);
  ray.d = mk(0.0f, 1.0f, 0.0f);
  if (active) {
    float ux = (float)px / (float)W, uy = 1.0f - (float)py / (float)H;
    ray = cast_ray_folded(fc, ux, uy, fc.jit_march[0], fc.jit_march[1]);
    const float d = ray.d.y;
    if (d < 0.0f) {
      float a = -ray.d.y * 5.5f;
      color = make_float4(mixf(0.0f * 0.4f, 0.0f * 0.5f, a), mixf(0.16f * 0.4f, 0.73f * 0.5f, a),
                          mixf(0.51f * 0.4f, 0.95f * 0.5f, a), 1.0f);
      n_ocean = 1;
    } else {
      bg = atmosphere_fast(fc, ray.d) * fmaxf(fc.bg_min, d);
      if (d < fc.fade_start) {
        color = make_float4(bg.x, bg.y, bg.z, 1.0f);
        n_sky = 1;
      } else {
        n_marched = 1;
        gdone = false;
        setup_march<FP64SHELL>(fc, ray, rm);
      }
    }
  }

  // --- rayMarch (:565-688), wave-uniform chunk loop ----------------------------
  float accum = 0.0f, transmittance = 1.0f, col = 0.0f;
  for (int k0 = 0; __any(!gdone); k0 += G) {
    Prim pr;
    pr.valid = false;
    pr.cloud = false;
    pr.da = 0.0f;
    if (BRICK) {
      // the LF fetch runs with all 64 lanes (ballots, box staging)
      PrimLf pl;
      pl.need_lf = false;
      pl.sk = mk(0.0f, 0.0f, 0.0f);
      pl.cov = pl.gw = 0.0f;
      if (!gdone) pr = eval_primary_pre<COV, F16>(fc, fv, sjit, rm, k0 + j, pl);
      const float2 lf = brick_sample_lf<F16>(fv, lf_box, bb, pl.need_lf, pl.sk, n_bstage, n_bserved);
      if (pl.need_lf) eval_primary_post<COV, F16>(fc, fv, pl, lf, pr);
    } else if (!gdone) {
      pr = eval_primary<COV, F16>(fc, fv, sjit, rm, k0 + j);
    }

    // compacted cone march of this chunk's in-cloud steps
    float E = 0.0f;
    int cone_hits = 0;
    float dal_mlp = 0.0f;
    bool need_cone = pr.cloud;
    bool audit = false;
    float c0 = 0.0f;
    if (LEARN && nls > 0) {
      // in-cloud steps at or below the fallback density query the network
      // (one row per step, ranked by ballot); the others, and those whose
      // prediction falls in an untrusted bin, go to the cone
      const bool query = pr.cloud && !(pr.base / fc.base_density_factor > fv.fallback_density);
      const unsigned long long qbal = __ballot(query);
      if (qbal) {
        uint16_t* X = mlp_lds + wave * 64 * 16;
        const int nq = __popcll(qbal);
        const int qrank = __popcll(qbal & ((1ull << lane) - 1ull));
        if (query) {
          c0 = cone_term<COV>(fc, fv, skern, pr, rm.cone_scale, 0);
          float f[9];
          mlp_features(fc, pr, rm, c0, f);
          uint32_t w[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const float lo = 2 * k < 9 ? f[2 * k] : 0.0f, hi = 2 * k + 1 < 9 ? f[2 * k + 1] : 0.0f;
            w[k] = (uint32_t)mlp::f2bf(lo) | ((uint32_t)mlp::f2bf(hi) << 16);
          }
          uint4* row = (uint4*)(X + qrank * 16);
          row[0] = make_uint4(w[0], w[1], w[2], w[3]);
          row[1] = make_uint4(w[4], w[5], w[6], w[7]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const float y = mlp::infer_wave(X, fv.mlp_frag, fv.mlp_params, nq);
        const float ym = __shfl(y, qrank, 64);
        if (query) {
          const int bin = calib_bin(ym, fv.calib_hi, fv.calib_inv);
          if (!((fv.calib_mask >> bin) & 1u)) {
            dal_mlp = c0 - ym;
            need_cone = false;
            audit = fv.audit && (step_hash((uint32_t)px, (uint32_t)py, (uint32_t)(k0 + j), 0u) & (uint32_t)(fv.audit - 1)) == 0u;
          }
        }
        // the next chunk rewrites the query rows
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    }
    const bool run_cone = need_cone || audit;
    const unsigned long long bal = __ballot(run_cone);
    const bool any_cloud = __ballot(pr.cloud) != 0ull;
    if (any_cloud) {
      const int ncl = __popcll(bal);
      const int rank = __popcll(bal & ((1ull << lane) - 1ull));
      if (run_cone) {
        L.own[rank][0] = make_float4(pr.rel.x, pr.rel.y, pr.rel.z, rm.cone_scale);
        L.own[rank][1] = make_float4(pr.sc, pr.st, pr.m, pr.inv1m);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // round q: lane l takes sample ci of step q * cgroups + cs
      for (int r0 = 0; r0 < ncl && !PAIR; r0 += cgroups) {
        const int r = r0 + cs;
        if (!cone_lane || r >= ncl) continue;
        const int w = r * nls + ci;
        const float4 a = L.own[r][0], b = L.own[r][1];
        const float s = fci * a.w;
        const f3 slp = mk(a.x + kx * s, a.y + ky * s, a.z + kz * s);
        float cov = fc.coverage_const, gw = 1.0f;
        if (COV == MTC_COVERAGE_WEATHER) {
          const float2 cw = sample_map<F16>(fv.weather, fv.weather_sh, slp.x, slp.z, fc.weather_scale);
          cov = cw.x;
          gw = gradient_weight(cw.y, b.x, b.y);
        }
        const float cb = gw <= cov ? 0.0f : density<COV>(sample_lf<F16>(fv.lf, fv.lf_sh, slp), cov, gw);
        L.contrib[w] = cb > 0.0f ? (fc.cone_erosion_scale * cb - b.z) * b.w : 0.0f;
        if (COUNT) L.hit[w] = cb > 0.0f ? 1 : 0;
      }
      // PAIR: rounds q and q + 1 together
      for (int r0 = 0; r0 < ncl && PAIR; r0 += 2 * cgroups) {
        const int rA = r0 + cs, rB = r0 + cgroups + cs;
        const bool vA = cone_lane && rA < ncl, vB = cone_lane && rB < ncl;
        if (!vA) continue;
        const float4 aA = L.own[rA][0], bA = L.own[rA][1];
        float4 aB = aA, bB = bA;
        if (vB) {
          aB = L.own[rB][0];
          bB = L.own[rB][1];
        }
        const float sA = fci * aA.w, sB = fci * aB.w;
        const f3 pA = mk(aA.x + kx * sA, aA.y + ky * sA, aA.z + kz * sA);
        const f3 pB = mk(aB.x + kx * sB, aB.y + ky * sB, aB.z + kz * sB);
        float covA = fc.coverage_const, gwA = 1.0f, covB = fc.coverage_const, gwB = 1.0f;
        if (COV == MTC_COVERAGE_WEATHER) {
          const float2 cwA = sample_map<F16>(fv.weather, fv.weather_sh, pA.x, pA.z, fc.weather_scale);
          float2 cwB = make_float2(0.0f, 0.0f);
          if (vB) cwB = sample_map<F16>(fv.weather, fv.weather_sh, pB.x, pB.z, fc.weather_scale);
          covA = cwA.x;
          gwA = gradient_weight(cwA.y, bA.x, bA.y);
          covB = cwB.x;
          gwB = vB ? gradient_weight(cwB.y, bB.x, bB.y) : 0.0f;
        }
        const bool lA = gwA > covA, lB = vB && gwB > covB;
        float2 fA = make_float2(0.0f, 0.0f), fB = make_float2(0.0f, 0.0f);
        if (lA) fA = sample_lf<F16>(fv.lf, fv.lf_sh, pA);
        if (lB) fB = sample_lf<F16>(fv.lf, fv.lf_sh, pB);
        const float cbA = lA ? density<COV>(fA, covA, gwA) : 0.0f;
        const int wA = rA * nls + ci;
        L.contrib[wA] = cbA > 0.0f ? (fc.cone_erosion_scale * cbA - bA.z) * bA.w : 0.0f;
        if (COUNT) L.hit[wA] = cbA > 0.0f ? 1 : 0;
        if (vB) {
          const float cbB = lB ? density<COV>(fB, covB, gwB) : 0.0f;
          const int wB = rB * nls + ci;
          L.contrib[wB] = cbB > 0.0f ? (fc.cone_erosion_scale * cbB - bB.z) * bB.w : 0.0f;
          if (COUNT) L.hit[wB] = cbB > 0.0f ? 1 : 0;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (pr.cloud) {
        float dal = dal_mlp;
        if (run_cone) {
          float cone = 0.0f;
          for (int i = 0; i < nls; ++i) {
            cone += L.contrib[rank * nls + i];
            if (COUNT) cone_hits += L.hit[rank * nls + i];
          }
          if (need_cone) {
            dal = cone;
          } else if (COUNT) {
            // the cone judges the network's answer for this step
            const float err = dal_mlp - cone;
            n_audit++;
            audit_sq += err * err;
          }
        }
        if (COUNT && !need_cone) cone_hits = -1;  // marks a step answered by the network
        // GetLightEnergy (:331-388): exp(-dl) once for both identical terms
        const float h = pr.h;
        float prim = fast_exp(-dal);
        // remap(cos, 0.7, 1, prim, 0.25 prim) with its per-ray ratio hoisted;
        // the two height remaps multiply by the reciprocal of their constant span
        float att = fmaxf(prim + rm.cos_ratio * (prim * 0.25f - prim), prim);
        float depth = 0.05f + fast_pow(pr.base, clampf(0.5f + ((h * 0.125f - 0.3f) * (1.0f / 0.55f)) * 1.5f, 0.5f, 2.0f));
        float vert = fast_pow(clampf(0.1f + ((h * 1.5f - 0.07f) * (1.0f / 0.27f)) * 0.9f, 0.1f, 1.0f), 0.8f);
        E = att * prim * (depth * vert) * rm.HG_light * fc.brightness;
      }
      // the next chunk rewrites the work area
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    // accumulation / early exit in step order (:674-684)
    if (!gdone && !any_cloud) {
      // no in-cloud step in the wave: nothing accumulates, so only the
      // step count and the validity of the group's last step change (valid
      // steps form a prefix of the chunk: t_k grows with k)
      const unsigned long long vb = __ballot(pr.valid);
      const unsigned int gm = (unsigned int)((vb >> gbase) & ((1ull << G) - 1ull));
      if (COUNT && j == 0) n_steps += (unsigned)__popc(gm);
      gdone = ((gm >> (G - 1)) & 1u) == 0u;
    } else if (!gdone) {
      // the group's valid / in-cloud bits from two wave ballots (no shuffle)
      const unsigned long long vball = __ballot(pr.valid), cball = __ballot(pr.cloud);
      const unsigned int vbits = (unsigned int)(vball >> gbase), cbits = (unsigned int)(cball >> gbase);
      bool last_valid = false, exited = false;
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const int src = gbase + i;
        const bool v = ((vbits >> i) & 1u) != 0u, cl = ((cbits >> i) & 1u) != 0u;
        const float da = G == 1 ? pr.da : __shfl(pr.da, src, 64);
        const float e = G == 1 ? E : __shfl(E, src, 64);
        const int ch = COUNT ? (G == 1 ? cone_hits : __shfl(cone_hits, src, 64)) : 0;
        last_valid = v;
        if (v && !exited) {
          if (COUNT && j == 0) n_steps++;
          if (cl) {
            if (COUNT && j == 0) {
              n_cloud++;
              if (ch >= 0) {
                n_cone += (unsigned)nls;
                n_cone_er += (unsigned)ch;
              } else if (LEARN) {
                n_cone += 1u;  // the network's input term, the cone's first sample
              }
              if (LEARN) {
                n_mlp++;
                if (ch >= 0) n_mlp_fb++;
              }
            }
            accum += da;
            transmittance = mixf(transmittance, e, 1.0f - accum);
            col += transmittance;
          }
          if (accum >= 1.0f) {
            accum = 1.0f;
            exited = true;
            if (COUNT && j == 0) n_early++;
          }
        }
      }
      gdone = exited || !last_valid;
    }
  }

  // --- composite and god-ray mask (:759-825) ---------------------------------
  if (active) {
    float4 mask = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (n_marched) {
      float god = accum;
      float a = accum * smoothstepf(0.0f, 1.0f, fminf(1.0f, remap(ray.d.y, fc.fade_start, fc.fade_end, 0.0f, 1.0f)));
      color = make_float4(mixf(bg.x, col, a), mixf(bg.y, col, a), mixf(bg.z, col, a), 1.0f);
      mask = encode_float_rgba(25.0f * fminf(0.05f, 1.0f - god));
      // :775-779, reachable only when cloud_fade_start < 0.05
      if (ray.d.y < 0.05f) {
        const float s = fmaxf(5.0f, ray.d.y);
        mask = make_float4(mask.x * s, mask.y * s, mask.z * s, mask.w * s);
      }
    }
    if (j == 0) {
      const size_t o = (size_t)py * (size_t)W + (size_t)px;
      out_color[o] = color;
      out_mask[o] = mask;
    }
  }
  if (COUNT) {
    if (j != 0) n_ocean = n_sky = n_marched = 0;
    wave_count(counters + 0, n_marched);
    wave_count(counters + 1, n_ocean);
    wave_count(counters + 2, n_sky);
    wave_count(counters + 3, n_steps);
    wave_count(counters + 4, n_cloud);
    wave_count(counters + 5, n_cone);
    wave_count(counters + 6, n_cone_er);
    wave_count(counters + 7, n_early);
    if (LEARN) {
      wave_count(counters + 8, n_mlp);
      wave_count(counters + 9, n_mlp_fb);
      wave_count(counters + 10, n_audit);
      wave_count(counters + 11, (unsigned long long)rintf(audit_sq * 1.0e6f));
    }
    if (BRICK) {
      wave_count(counters + 12, n_bstage);
      wave_count(counters + 13, n_bserved);
    }
  }
}

// The tile that workgroup r of an n-tile launch marches: the r-th of the
// previous march's tiles taken longest level first, or tile r when slot
// `rd` does not hold exactly one filing of every tile of this tiling.  The
// counts and the tile are wave-uniform (scalar loads).
__device__ __forceinline__ uint32_t ordered_tile(const TileOrder& to, uint32_t r, uint32_t n) {
  if (!to.rd || n > to.cap) return r;
  uint32_t total = 0;
#pragma unroll
  for (int l = 0; l < kOrderLevels; ++l) total += to.rd[l];
  if (total != n) return r;
  uint32_t base = 0;
  for (int l = kOrderLevels - 1; l >= 0; --l) {
    const uint32_t cnt = to.rd[l];
    if (r < base + cnt) return to.rd[kOrderLevels + (size_t)l * to.cap + (r - base)];
    base += cnt;
  }
  return r;
}

// Files tile `tile` under the level of the workgroup's duration since t0
// (100 MHz ticks; level 2 log2(ticks / 64), so 0.64 us .. 116 us over the
// 16 levels) and clears the counts of the slot after next (workgroup 0).
// One lane per workgroup writes, with vector atomics and stores.
__device__ __forceinline__ void file_tile(const TileOrder& to, uint32_t tile, uint32_t n, uint64_t t0) {
  if (!to.wr) return;
  if (blockIdx.x == 0 && threadIdx.x < kOrderLevels) to.clr[threadIdx.x] = 0u;
  if (threadIdx.x != 0 || n > to.cap) return;
  const float ticks = (float)(__builtin_amdgcn_s_memrealtime() - t0);
  const int level = ticks < 64.0f ? 0 : min(kOrderLevels - 1, (int)(2.0f * __log2f(ticks * (1.0f / 64.0f))));
  const uint32_t pos = atomicAdd(to.wr + level, 1u);
  if (pos < to.cap) to.wr[kOrderLevels + (size_t)level * to.cap + pos] = tile;
}

// One frame: a one-wave workgroup marches an 8 x (8 / G) tile of rays (a
// finer grain than four-wave groups: the launch drains in smaller steps),
// in strip order (above).  PAIR: cone rounds two at a time (95 VGPRs, 5
// waves per SIMD), for launches of a few tiles per wave slot, whose time
// is set by how soon the last tiles finish; single rounds (78 VGPRs, 6
// waves per SIMD) for longer launches, whose time is set by throughput.  Dealing tiles longest first (as the learned
// kernel does) scatters the resident waves over the image, and the lost
// texel locality costs more than the shorter tail gains.
//
// Workgroups n_tiles.. of the launch carry the frame's reprojection of the
// other 15/16 of the pixels instead (64 x kReprojStripRows pixel strips;
// reproject_pixel.h, bit-exact with the reprojection kernel): dealt after
// every march tile, they run in the wave slots the draining march leaves
// free, and the frame needs no second stream and no fork / join events.
// The two parts write disjoint pixels and the march reads nothing the
// reprojection writes.  Two rows per strip (each lane's pixels are serial
// chains of dependent tap loads; shorter strips pack the drain finer):
// measured on MI355X, 4K 4x4 frames 0.770 ms with 4 rows, 0.760 with 2,
// 0.763 with 1; 1080p frames within 0.5% for all three.
constexpr int kReprojStripRows = 2;
template <int COV, bool F16, bool FP64SHELL, bool COUNT, int G, bool BRICK, bool PAIR = true>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(COUNT ? 1 : (PAIR ? 5 : 6)))) void march_kernel(
    FrameConsts fc, FastViews fv, OutputViews ov, int full, int tiles_x, int n_tiles, int nstrip, int strip_k) {
  const int b = (int)blockIdx.x;
  if (b >= n_tiles) {
    const int strips_x = (fc.width + 63) / 64, r = b - n_tiles;
    const int x = (r % strips_x) * 64 + (int)threadIdx.x, y0 = (r / strips_x) * kReprojStripRows;
    for (int y = y0; y < y0 + kReprojStripRows; ++y)
      if (x < fc.width && y < fc.height && !mtc::reproj::marched_this_frame(fc, x, y))
        mtc::reproj::reproject_pixel<false>(fc, ov, x, y);
    return;
  }
  __shared__ float sjit[48];
  __shared__ float skern[18];
  __shared__ ConeLds cone_lds[1];
  __shared__ __align__(16) uint32_t lf_box[BRICK ? 512 : 4];
  if (threadIdx.x < 48) sjit[threadIdx.x] = (&fc.step_jitter[0][0])[threadIdx.x];
  if (threadIdx.x < 18) skern[threadIdx.x] = (&fc.kern[0][0])[threadIdx.x];
  __syncthreads();
  constexpr int TX = 8, TY = 64 / G / 8;
  int tx, ty;
  strip_tile(b, tiles_x, n_tiles / tiles_x, nstrip, strip_k, tx, ty);
  march_tile<COV, F16, FP64SHELL, COUNT, G, false, PAIR, BRICK>(fc, fv, ov.color, ov.mask, ov.counters, full,
                                                                tx * TX, ty * TY, TX, sjit, skern, cone_lds, nullptr,
                                                                lf_box);
}

// One frame with learned light (fast variant's sampling): the in-cloud
// steps of each chunk query the network on MFMA before the cone march.
// Tiles are dealt longest first (ordered_tile / file_tile); workgroups
// n_tiles.. carry the frame's reprojection as 64 x 4 pixel strips, one row
// per wave, as in the fast launch.
template <int COV, bool COUNT, int G>
__global__ __launch_bounds__(256) void march_learned_kernel(FrameConsts fc, FastViews fv, OutputViews ov, int full,
                                                            TileOrder to, int tiles_x, int n_tiles) {
  const int b = (int)blockIdx.x;
  if (b >= n_tiles) {
    const int strips_x = (fc.width + 63) / 64, r = b - n_tiles;
    const int x = (r % strips_x) * 64 + (int)(threadIdx.x & 63), y = (r / strips_x) * 4 + (int)(threadIdx.x >> 6);
    if (x < fc.width && y < fc.height && !mtc::reproj::marched_this_frame(fc, x, y))
      mtc::reproj::reproject_pixel<false>(fc, ov, x, y);
    return;
  }
  __shared__ float sjit[48];
  __shared__ float skern[18];
  __shared__ ConeLds cone_lds[4];
  __shared__ __align__(16) uint16_t mlp_lds[4 * 64 * 16];
  const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
  if (threadIdx.x < 48) sjit[threadIdx.x] = (&fc.step_jitter[0][0])[threadIdx.x];
  if (threadIdx.x < 18) skern[threadIdx.x] = (&fc.kern[0][0])[threadIdx.x];
  __syncthreads();
  constexpr int TX = 8, TY = 256 / G / 8;
  const uint32_t tile = ordered_tile(to, (uint32_t)b, (uint32_t)n_tiles);
  march_tile<COV, false, true, COUNT, G, true>(fc, fv, ov.color, ov.mask, ov.counters, full,
                                               (int)(tile % (uint32_t)tiles_x) * TX,
                                               (int)(tile / (uint32_t)tiles_x) * TY, TX, sjit, skern, cone_lds,
                                               mlp_lds);
  __syncthreads();  // the workgroup's duration is its slowest wave's
  file_tile(to, tile, (uint32_t)n_tiles, t0);
}

// Training samples for the learned light: every pixel of the view is
// marched (one lane per ray, the same steps and early exit as the frame
// kernels); each in-cloud step is kept with probability keep / 2^32 and
// written as (nine inputs, -(cone terms 1 .. n-1), key = py << 40 | px << 20
// | step) at an atomically claimed slot.  Slots past `cap` are counted but
// not written.
template <int COV>
__global__ __launch_bounds__(256) void light_samples_kernel(FrameConsts fc, FastViews fv, uint32_t seed, uint32_t keep,
                                                            float* __restrict__ xs, float* __restrict__ ys,
                                                            unsigned long long* __restrict__ keys,
                                                            unsigned long long* __restrict__ count,
                                                            unsigned long long cap) {
  __shared__ float sjit[48];
  __shared__ float skern[18];
  if (threadIdx.x < 48) sjit[threadIdx.x] = (&fc.step_jitter[0][0])[threadIdx.x];
  if (threadIdx.x < 18) skern[threadIdx.x] = (&fc.kern[0][0])[threadIdx.x];
  __syncthreads();
  const int px = blockIdx.x * 16 + (threadIdx.x & 15);
  const int py = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (px >= fc.width || py >= fc.height) return;
  const float ux = (float)px / (float)fc.width, uy = 1.0f - (float)py / (float)fc.height;
  const Ray ray = cast_ray_folded(fc, ux, uy, fc.jit_march[0], fc.jit_march[1]);
  if (ray.d.y < 0.0f || ray.d.y < fc.fade_start) return;
  RayMarch rm;
  setup_march<true>(fc, ray, rm);
  float accum = 0.0f;
  for (int k = 0;; ++k) {
    const Prim pr = eval_primary<COV, false>(fc, fv, sjit, rm, k);
    if (!pr.valid) break;
    if (pr.cloud) {
      accum += pr.da;
      if (step_hash((uint32_t)px, (uint32_t)py, (uint32_t)k, seed) < keep) {
        const float c0 = fc.light_samples > 0 ? cone_term<COV>(fc, fv, skern, pr, rm.cone_scale, 0) : 0.0f;
        float rest = 0.0f;
        for (int i = 1; i < fc.light_samples; ++i) rest += cone_term<COV>(fc, fv, skern, pr, rm.cone_scale, i);
        float f[9];
        mlp_features(fc, pr, rm, c0, f);
        const unsigned long long slot = atomicAdd(count, 1ull);
        if (slot < cap) {
          for (int i = 0; i < 9; ++i) xs[slot * 9 + i] = f[i];
          ys[slot] = -rest;
          keys[slot] = ((unsigned long long)py << 40) | ((unsigned long long)px << 20) | (unsigned long long)k;
        }
      }
    }
    if (accum >= 1.0f) break;
  }
}

// Many views in one launch: a persistent grid pulls (view, tile) work items
// from 8 lists, one per XCD (each view's tiles dealt to the XCDs in column
// strips, so that an XCD's L2 serves one eighth of each view; items sorted
// by estimated cost when the batch was built).  A workgroup takes
// kBatchClaim consecutive items per atomic add from the list of the XCD it
// runs on (read from the hardware register), and once that list is empty
// from the next lists in turn; it remembers the first list that was not
// empty (counters only grow, so an empty list stays empty) and stops when
// all 8 are.  Every add is a device-scope atomic that the memory side
// serialises, and its round trip is exposed before the tile starts;
// measured on MI355X (96-view set, precision, one list): 1 item per add
// 18.36 ms, 2 18.04, 4 17.97, 8 17.95.  Each view carries its own
// FrameConsts, size, output offset and steps.  Six waves per SIMD (80
// VGPRs; four values spilled, reloaded once per item): 16.93 -> 16.09 ms
// on the 96-view set against five (95 VGPRs); seven spill 17-20 values.
constexpr int kBatchClaim = 4;
constexpr int kHwRegXccId = 20 | (0 << 6) | (3 << 11);  // s_getreg operand: HW_REG_XCC_ID bits [3:0]

template <int COV, bool F16, int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void march_batch_kernel(const BatchView* __restrict__ views,
                                                          const BatchItem* __restrict__ items, BatchQueues bq,
                                                          unsigned int* __restrict__ queue, FastViews fv,
                                                          float4* __restrict__ color, float4* __restrict__ mask,
                                                          int tile_x) {
  __shared__ float sjit[48];
  __shared__ float skern[18];
  __shared__ ConeLds cone_lds[4];
  __shared__ int s_item, s_n;
  const int tile_y = 256 / G / tile_x;
  const int xcd = __builtin_amdgcn_s_getreg(kHwRegXccId) & 7;
  int from = 0;               // lists (xcd + from) .. are the ones not yet seen empty
  int it = 0, left = 0;       // next item of the current claim, items left in it
  for (;;) {
    if (left == 0) {
      if (threadIdx.x == 0) {
        int got = -1, n = 0;
        for (; from < 8; ++from) {
          const int q = (xcd + from) & 7;
          const unsigned p = atomicAdd(queue + kBatchQueueStride * q, (unsigned)kBatchClaim);
          if (p < (unsigned)bq.len[q]) {
            got = bq.off[q] + (int)p;
            n = min(kBatchClaim, bq.len[q] - (int)p);
            break;
          }
        }
        s_item = got;
        s_n = n;
      }
      __syncthreads();
      it = s_item;
      left = s_n;
      if (it < 0) break;
    }
    const BatchItem wi = items[it];
    const BatchView& bv = views[__builtin_amdgcn_readfirstlane(wi.view)];
    if (threadIdx.x < 48) sjit[threadIdx.x] = (&bv.fc.step_jitter[0][0])[threadIdx.x];
    if (threadIdx.x < 18) skern[threadIdx.x] = (&bv.fc.kern[0][0])[threadIdx.x];
    __syncthreads();
    const int tx = wi.tile % bv.tiles_x, ty = wi.tile / bv.tiles_x;
    march_tile<COV, F16, true, false, G>(bv.fc, fv, color + bv.offset, mask + bv.offset, nullptr, 1, tx * tile_x,
                                         ty * tile_y, tile_x, sjit, skern, cone_lds);
    ++it;
    --left;
    // the next item rewrites the tables and the queue slot
    __syncthreads();
  }
}

__global__ void reset_queue_kernel(unsigned int* queue) {
  if (threadIdx.x < 8) queue[kBatchQueueStride * threadIdx.x] = 0u;
}

// ---------------------------------------------------------------------------
// footprint packing (run once per volume upload)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void texel4(const void* src, int fmt, size_t idx, float v[4]) {
  if (fmt == MTC_FMT_FP16) {
    uint2 t = ((const uint2*)src)[idx];
    float2 a = f2(t.x), b = f2(t.y);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
  } else {
    uint32_t t = ((const uint32_t*)src)[idx];
    const float k = MTC_INV255;
    v[0] = (float)(t & 0xffu) * k;
    v[1] = (float)((t >> 8) & 0xffu) * k;
    v[2] = (float)((t >> 16) & 0xffu) * k;
    v[3] = (float)(t >> 24) * k;
  }
}

// One LF texel as half2 (r, .625g + .25b + .125a); both LF copies (the
// footprint table and the bricked copy) hold exactly these values.
__device__ __forceinline__ uint32_t lf_texel_h2(const void* src, int fmt, int sh, int x, int y, int z) {
  float v[4];
  texel4(src, fmt, (((size_t)z << sh) + y) * ((size_t)1 << sh) + x, v);
  return pack_h2(v[0], (v[1] * 0.625f + v[2] * 0.25f) + v[3] * 0.125f);
}

// LF -> (r, .625g + .25b + .125a) at the 2x2x2 footprint of each texel:
// entry 2 idx holds the z0 slice's 2x2, entry 2 idx + 1 the z1 slice's
__global__ __launch_bounds__(256) void pack_lf_kernel(const void* src, int fmt, int sh, uint4* out) {
  const int N = 1 << sh, m = N - 1;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)N * N * N) return;
  int x = (int)(idx & m), y = (int)((idx >> sh) & m), z = (int)(idx >> (2 * sh));
  uint32_t e[8];
  for (int k = 0; k < 8; ++k) e[k] = lf_texel_h2(src, fmt, sh, (x + (k & 1)) & m, (y + ((k >> 1) & 1)) & m, (z + (k >> 2)) & m);
  out[2 * idx] = make_uint4(e[0], e[1], e[2], e[3]);
  out[2 * idx + 1] = make_uint4(e[4], e[5], e[6], e[7]);
}

// LF -> the same half2 texels in 4^3 bricks (N >= 8): brick (bx, by, bz)
// of N/4 per axis at ((bz N/4 + by) N/4 + bx) * 64 texels, texel (x, y, z)
// of a brick at x + 4 y + 16 z.  One lane per texel of the output.
__global__ __launch_bounds__(256) void pack_lf_brick_kernel(const void* src, int fmt, int sh, uint32_t* out) {
  const int N = 1 << sh, nbsh = sh - 2, nbm = (N >> 2) - 1;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)N * N * N) return;
  const size_t brick = idx >> 6;
  const int t = (int)(idx & 63);
  const int bx = (int)(brick & nbm), by = (int)((brick >> nbsh) & nbm), bz = (int)(brick >> (2 * nbsh));
  out[idx] = lf_texel_h2(src, fmt, sh, 4 * bx + (t & 3), 4 * by + ((t >> 2) & 3), 4 * bz + (t >> 4));
}

// HF -> .625r + .25g + .125b at the 2x2x2 footprint
__global__ __launch_bounds__(256) void pack_hf_kernel(const void* src, int fmt, int sh, uint4* out) {
  const int N = 1 << sh, m = N - 1;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)N * N * N) return;
  int x = (int)(idx & m), y = (int)((idx >> sh) & m), z = (int)(idx >> (2 * sh));
  float f[8];
  for (int k = 0; k < 8; ++k) {
    int xx = (x + (k & 1)) & m, yy = (y + ((k >> 1) & 1)) & m, zz = (z + (k >> 2)) & m;
    float v[4];
    texel4(src, fmt, (((size_t)zz << sh) + yy) * N + xx, v);
    f[k] = (v[0] * 0.625f + v[1] * 0.25f) + v[2] * 0.125f;
  }
  out[idx] = make_uint4(pack_h2(f[0], f[1]), pack_h2(f[2], f[3]), pack_h2(f[4], f[5]), pack_h2(f[6], f[7]));
}

// 2D RGBA8 map -> (r, g) at the 2x2 footprint
__global__ __launch_bounds__(256) void pack_map_kernel(const uint32_t* src, int sh, uint4* out) {
  const int N = 1 << sh, m = N - 1;
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= N * N) return;
  int x = idx & m, y = idx >> sh;
  uint32_t e[4];
  for (int k = 0; k < 4; ++k) {
    int xx = (x + (k & 1)) & m, yy = (y + (k >> 1)) & m;
    float v[4];
    texel4(src, MTC_FMT_UNORM8, ((size_t)yy << sh) + xx, v);
    e[k] = pack_h2(v[0], v[1]);
  }
  out[idx] = make_uint4(e[0], e[1], e[2], e[3]);
}

}  // namespace fast
}  // namespace mtc

// lf_brick: the bricked LF copy of MTC_FLAG_LDS_BRICKS contexts (N^3
// half2 texels, lf_dim >= 8), or null
hipError_t pack_fast_resources(const VolumeViews& vv, uint4* lf, uint4* hf, uint4* curl, uint4* weather,
                               uint32_t* lf_brick, hipStream_t s) {
  size_t n = (size_t)vv.lf_dim * vv.lf_dim * vv.lf_dim;
  hipLaunchKernelGGL(mtc::fast::pack_lf_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, vv.lf, vv.format,
                     vv.lf_shift, lf);
  if (lf_brick) {
    if (vv.lf_dim < 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mtc::fast::pack_lf_brick_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, vv.lf,
                       vv.format, vv.lf_shift, lf_brick);
  }
  n = (size_t)vv.hf_dim * vv.hf_dim * vv.hf_dim;
  hipLaunchKernelGGL(mtc::fast::pack_hf_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, vv.hf, vv.format,
                     vv.hf_shift, hf);
  n = (size_t)vv.curl_dim * vv.curl_dim;
  hipLaunchKernelGGL(mtc::fast::pack_map_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, vv.curl,
                     vv.curl_shift, curl);
  n = (size_t)vv.weather_dim * vv.weather_dim;
  hipLaunchKernelGGL(mtc::fast::pack_map_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, vv.weather,
                     vv.weather_shift, weather);
  return hipGetLastError();
}

// Launches with at least this many march tiles per compute unit use the
// single-round kernel.  Measured on MI355X (precision march, tiles per CU:
// paired-round time -> single-round time): 1080p 4x4 frames (63) 0.2029 ->
// 0.2066 ms; 2560x1440 4x4 (112) 0.346 -> 0.342 ms; 3200x1800 4x4 (176)
// 0.516 -> 0.504 ms; 4K 4x4 (253) 0.875 -> 0.847 ms; 1080p full (1013) 2.16
// -> 2.12 ms.
constexpr int kSingleRoundTilesPerCu = 96;

// Compute units of the current device (cached per device).
static int compute_units() {
  static int cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cached[dev]) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    cached[dev] = cus;
  }
  return cached[dev];
}
template <int COV, bool F16, bool FP64, int G, bool BRICK = false>
static void launch_v(hipStream_t s, const FrameConsts& fc, const mtc::fast::FastViews& fv, const OutputViews& ov,
                     int full, bool with_reproj, const KernelEvents& ke) {
  const int rays_x = full ? fc.width : (fc.width + 3) / 4;
  const int rays_y = full ? fc.height : (fc.height + 3) / 4;
  constexpr int TX = 8, TY = 64 / G / 8;
  const int tiles_x = (rays_x + TX - 1) / TX, tiles_y = (rays_y + TY - 1) / TY, n_tiles = tiles_x * tiles_y;
  const int strips =
      with_reproj ? ((fc.width + 63) / 64) * ((fc.height + mtc::fast::kReprojStripRows - 1) / mtc::fast::kReprojStripRows)
                  : 0;
  const StripOrder so = strip_order(tiles_x);
  dim3 grid((unsigned)(n_tiles + strips)), block(64);
  constexpr bool kSingle = !BRICK && (G == 4 || G == 8);  // single-round kernels built for these
  if (ov.counters)
    launch_timed(mtc::fast::march_kernel<COV, F16, FP64, true, G, BRICK>, grid, block, s, ke, fc, fv, ov, full,
                 tiles_x, n_tiles, so.nstrip, so.k);
  else if (kSingle && n_tiles >= kSingleRoundTilesPerCu * compute_units())
    launch_timed(mtc::fast::march_kernel<COV, F16, FP64, false, G, BRICK, !kSingle>, grid, block, s, ke, fc, fv, ov,
                 full, tiles_x, n_tiles, so.nstrip, so.k);
  else
    launch_timed(mtc::fast::march_kernel<COV, F16, FP64, false, G, BRICK>, grid, block, s, ke, fc, fv, ov, full,
                 tiles_x, n_tiles, so.nstrip, so.k);
}

template <bool F16, bool FP64, int G>
static void launch_cov(hipStream_t s, const FrameConsts& fc, const mtc::fast::FastViews& fv, const OutputViews& ov,
                       int full, bool rp, const KernelEvents& ke) {
  if (fc.coverage_mode == MTC_COVERAGE_WEATHER) launch_v<MTC_COVERAGE_WEATHER, F16, FP64, G>(s, fc, fv, ov, full, rp, ke);
  else launch_v<MTC_COVERAGE_CONSTANT, F16, FP64, G>(s, fc, fv, ov, full, rp, ke);
}

template <bool F16, bool FP64>
static bool launch_g(int lanes_per_ray, hipStream_t s, const FrameConsts& fc, const mtc::fast::FastViews& fv,
                     const OutputViews& ov, int full, bool rp, const KernelEvents& ke) {
  switch (lanes_per_ray) {
    case 1: launch_cov<F16, FP64, 1>(s, fc, fv, ov, full, rp, ke); return true;
    case 2: launch_cov<F16, FP64, 2>(s, fc, fv, ov, full, rp, ke); return true;
    case 4: launch_cov<F16, FP64, 4>(s, fc, fv, ov, full, rp, ke); return true;
    case 8: launch_cov<F16, FP64, 8>(s, fc, fv, ov, full, rp, ke); return true;
    default: return false;
  }
}

// variant 0 = fast (FP32 filtering), 1 = precision (packed FP16
// filtering); lanes_per_ray in {1, 2, 4, 8}; with_reproj (4x4 frames
// only): the launch also carries the reprojection of the other pixels
// (ov.prev -> ov.color).  lf_brick (lanes_per_ray 8 only): the bricked LF
// copy, and the primary LF samples go through the waves' LDS boxes.
hipError_t launch_march_fast(const FrameConsts& fc, const uint4* lf, const uint4* hf, const uint4* curl,
                             const uint4* weather, const VolumeViews& vv, const OutputViews& ov, int full, int variant,
                             int lanes_per_ray, bool with_reproj, hipStream_t stream, const KernelEvents& ke,
                             const uint32_t* lf_brick) {
  if (with_reproj && (full || !ov.prev)) return hipErrorInvalidValue;
  mtc::fast::FastViews fv{lf, hf, curl, weather, vv.lf_shift, vv.hf_shift, vv.curl_shift, vv.weather_shift};
  if (lf_brick) {
    if (lanes_per_ray != 8 || vv.lf_dim < 8) return hipErrorInvalidValue;
    fv.lf_brick = (const uint4*)lf_brick;
    const bool w = fc.coverage_mode == MTC_COVERAGE_WEATHER, rp = with_reproj;
    if (variant == 0) {
      if (w) launch_v<MTC_COVERAGE_WEATHER, false, true, 8, true>(stream, fc, fv, ov, full, rp, ke);
      else launch_v<MTC_COVERAGE_CONSTANT, false, true, 8, true>(stream, fc, fv, ov, full, rp, ke);
    } else {
      if (w) launch_v<MTC_COVERAGE_WEATHER, true, true, 8, true>(stream, fc, fv, ov, full, rp, ke);
      else launch_v<MTC_COVERAGE_CONSTANT, true, true, 8, true>(stream, fc, fv, ov, full, rp, ke);
    }
    return hipGetLastError();
  }
  // Both variants keep the shell quadratic in FP64: in FP32 its hit points
  // move by 0.3-1 m (C = dot(rO, rO) - 1 cancels), which shifts every t_k and
  // flips the int(t) jitter index of most steps (:631).
  bool ok = variant == 0 ? launch_g<false, true>(lanes_per_ray, stream, fc, fv, ov, full, with_reproj, ke)
                         : launch_g<true, true>(lanes_per_ray, stream, fc, fv, ov, full, with_reproj, ke);
  if (!ok) return hipErrorInvalidValue;
  return hipGetLastError();
}

static mtc::fast::FastViews learned_views(const uint4* lf, const uint4* hf, const uint4* curl, const uint4* weather,
                                          const VolumeViews& vv, const LearnedLight& ll) {
  mtc::fast::FastViews fv{lf, hf, curl, weather, vv.lf_shift, vv.hf_shift, vv.curl_shift, vv.weather_shift};
  fv.mlp_frag = (const bf16x8*)ll.frag;
  fv.mlp_params = ll.params;
  fv.calib_hi = ll.calib_hi;
  fv.calib_inv = ll.calib_inv;
  fv.calib_mask = ll.calib_mask;
  fv.fallback_density = ll.fallback_density;
  fv.audit = ll.audit;
  return fv;
}

template <int COV, int G>
static void launch_learned_v(hipStream_t s, const FrameConsts& fc, const mtc::fast::FastViews& fv,
                             const OutputViews& ov, int full, bool with_reproj, const TileOrder& to,
                             const KernelEvents& ke) {
  const int rays_x = full ? fc.width : (fc.width + 3) / 4;
  const int rays_y = full ? fc.height : (fc.height + 3) / 4;
  constexpr int TX = 8, TY = 256 / G / 8;
  const int tiles_x = (rays_x + TX - 1) / TX, n_tiles = tiles_x * ((rays_y + TY - 1) / TY);
  const int strips = with_reproj ? ((fc.width + 63) / 64) * ((fc.height + 3) / 4) : 0;
  dim3 grid((unsigned)(n_tiles + strips)), block(256);
  if (ov.counters)
    launch_timed(mtc::fast::march_learned_kernel<COV, true, G>, grid, block, s, ke, fc, fv, ov, full, to, tiles_x,
                 n_tiles);
  else
    launch_timed(mtc::fast::march_learned_kernel<COV, false, G>, grid, block, s, ke, fc, fv, ov, full, to, tiles_x,
                 n_tiles);
}

// Learned-light frame (fast sampling, FP64 shell); lanes_per_ray 4 or 8;
// with_reproj as for launch_march_fast.
hipError_t launch_march_learned(const FrameConsts& fc, const uint4* lf, const uint4* hf, const uint4* curl,
                                const uint4* weather, const VolumeViews& vv, const OutputViews& ov, int full,
                                int lanes_per_ray, const LearnedLight& ll, bool with_reproj, const TileOrder& to,
                                hipStream_t s, const KernelEvents& ke) {
  if (with_reproj && (full || !ov.prev)) return hipErrorInvalidValue;
  const mtc::fast::FastViews fv = learned_views(lf, hf, curl, weather, vv, ll);
  const bool wx = fc.coverage_mode == MTC_COVERAGE_WEATHER;
  const bool rp = with_reproj;
  if (lanes_per_ray == 8) {
    if (wx) launch_learned_v<MTC_COVERAGE_WEATHER, 8>(s, fc, fv, ov, full, rp, to, ke);
    else launch_learned_v<MTC_COVERAGE_CONSTANT, 8>(s, fc, fv, ov, full, rp, to, ke);
  } else if (lanes_per_ray == 4) {
    if (wx) launch_learned_v<MTC_COVERAGE_WEATHER, 4>(s, fc, fv, ov, full, rp, to, ke);
    else launch_learned_v<MTC_COVERAGE_CONSTANT, 4>(s, fc, fv, ov, full, rp, to, ke);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_light_samples(const FrameConsts& fc, const uint4* lf, const uint4* hf, const uint4* curl,
                                const uint4* weather, const VolumeViews& vv, uint32_t seed, uint32_t keep, float* xs,
                                float* ys, unsigned long long* keys, unsigned long long* count, unsigned long long cap,
                                hipStream_t s) {
  mtc::fast::FastViews fv{lf, hf, curl, weather, vv.lf_shift, vv.hf_shift, vv.curl_shift, vv.weather_shift};
  dim3 grid((fc.width + 15) / 16, (fc.height + 15) / 16), block(256);
  if (fc.coverage_mode == MTC_COVERAGE_WEATHER)
    hipLaunchKernelGGL((mtc::fast::light_samples_kernel<MTC_COVERAGE_WEATHER>), grid, block, 0, s, fc, fv, seed, keep,
                       xs, ys, keys, count, cap);
  else
    hipLaunchKernelGGL((mtc::fast::light_samples_kernel<MTC_COVERAGE_CONSTANT>), grid, block, 0, s, fc, fv, seed,
                       keep, xs, ys, keys, count, cap);
  return hipGetLastError();
}

template <int COV, bool F16, int G>
static void launch_batch_g(const BatchView* views, const BatchItem* items, const BatchQueues& bq,
                           unsigned int* queue, const mtc::fast::FastViews& fv, float4* color, float4* mask,
                           int tile_x, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(mtc::fast::reset_queue_kernel, dim3(1), dim3(64), 0, s, queue);
  hipLaunchKernelGGL((mtc::fast::march_batch_kernel<COV, F16, G>), dim3(blocks), dim3(256), 0, s, views, items, bq,
                     queue, fv, color, mask, tile_x);
}

// Enqueue one batched render: queue reset + persistent march over the items
// (the 8 lists of bq; queue holds 8 * kBatchQueueStride words).  variant 0 =
// fast, 1 = precision; lanes_per_ray 4 or 8.
hipError_t launch_march_batch(const BatchView* views, const BatchItem* items, const BatchQueues& bq, unsigned int* queue,
                              const uint4* lf, const uint4* hf, const uint4* curl, const uint4* weather,
                              const VolumeViews& vv, float4* color, float4* mask, int coverage_mode, int variant,
                              int lanes_per_ray, int tile_x, int blocks, hipStream_t s) {
  mtc::fast::FastViews fv{lf, hf, curl, weather, vv.lf_shift, vv.hf_shift, vv.curl_shift, vv.weather_shift};
  const bool w = coverage_mode == MTC_COVERAGE_WEATHER;
#define MTC_BATCH(COVV, F16V, GV) launch_batch_g<COVV, F16V, GV>(views, items, bq, queue, fv, color, mask, tile_x, blocks, s)
  if (lanes_per_ray == 8) {
    if (variant == 0) { if (w) MTC_BATCH(MTC_COVERAGE_WEATHER, false, 8); else MTC_BATCH(MTC_COVERAGE_CONSTANT, false, 8); }
    else { if (w) MTC_BATCH(MTC_COVERAGE_WEATHER, true, 8); else MTC_BATCH(MTC_COVERAGE_CONSTANT, true, 8); }
  } else if (lanes_per_ray == 4) {
    if (variant == 0) { if (w) MTC_BATCH(MTC_COVERAGE_WEATHER, false, 4); else MTC_BATCH(MTC_COVERAGE_CONSTANT, false, 4); }
    else { if (w) MTC_BATCH(MTC_COVERAGE_WEATHER, true, 4); else MTC_BATCH(MTC_COVERAGE_CONSTANT, true, 4); }
  } else {
    return hipErrorInvalidValue;
  }
#undef MTC_BATCH
  return hipGetLastError();
}

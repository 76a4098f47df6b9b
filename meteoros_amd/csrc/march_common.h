// march_common.h -- per-ray pieces of cloudRayMarch.comp shared by the
// march kernels: castRay (:194-226), raySphereIntersection (:229-273),
// HenyeyGreenstein (:295-300), the Preetham background (:401-467) and
// getDensityHeightGradientForPoint (:475-487).  FP32 in GLSL source order
// (-ffp-contract=off), identical for every kernel variant; they run once
// per ray and are a few percent of a marched ray's work.
#pragma once
#include "device_common.h"

namespace mtc {

struct Ray {
  f3 o, d;
};

// castRay, cloudRayMarch.comp:194-226
__device__ __forceinline__ Ray cast_ray(const FrameConsts& fc, const float* view, float ux, float uy, float jx,
                                        float jy) {
  f3 right = normalize(mk(view[0], view[4], view[8]));
  f3 up = normalize(mk(view[1], view[5], view[9]));
  f3 lk = normalize(mk(view[2], view[6], view[10]));
  f3 look = mk(-lk.x, -lk.y, -lk.z);
  float nx = ux * 2.0f - 1.0f, ny = uy * 2.0f - 1.0f;
  nx = nx + jx;
  ny = ny + jy;
  f3 cx = (nx * fc.tfx) * right;
  f3 cy = (ny * fc.tfy) * up;
  f3 eye = ld3(fc.eye);
  f3 ref = eye + look;
  f3 p = (ref + cx) + cy;
  Ray r;
  r.o = eye;
  r.d = normalize(p - eye);
  return r;
}

struct Isect {
  f3 point;
  float t;
};
// raySphereIntersection, :229-273
__device__ __forceinline__ Isect ray_sphere(f3 rO, f3 rD, f3 c, float R) {
  Isect is;
  is.point = mk(0.0f, 0.0f, 0.0f);
  is.t = 0.0f;
  rO = rO - c;
  rO = rO / R;
  float A = dot(rD, rD);
  float B = 2.0f * dot(rD, rO);
  float C = dot(rO, rO) - 1.0f;
  float disc = B * B - 4.0f * A * C;
  if (disc < 0.0f) return is;
  float t = (-B - __builtin_sqrtf(disc)) / (2.0f * A);
  if (t < 0.0f) t = (-B + __builtin_sqrtf(disc)) / (2.0f * A);
  if (t >= 0.0f) {
    f3 p = rO + t * rD;
    p = p * R;
    p = p + c;
    is.point = p;
    is.t = length(p - rO);
  }
  return is;
}

// HenyeyGreenstein, :295-300
__device__ __forceinline__ float hg(float c, float g) {
  float num = 1.0f - g * g;
  float den = powf((1.0f + g * g - 2.0f * g * c), 1.5f);
  return (num / den) * 0.07957747154594767f;
}

// getAtmosphereColorPhysical, :427-467 (per-frame parts in FrameConsts)
__device__ f3 atmosphere(const FrameConsts& fc, f3 dir) {
  f3 BR = ld3(fc.beta_r), BM = ld3(fc.beta_m), BS = ld3(fc.beta_sum);
  f3 sunDir = ld3(fc.sky_sun_dir);
  float sunE = fc.sunE;
  float zenith = acosf(fmaxf(0.0f, dir.y));
  float inv = 1.0f / (cosf(zenith) + 0.15f * powf(93.885f - ((zenith * 180.0f) / 3.14159265f), -1.253f));
  float sR = 8.4E3f * inv;
  float sM = 1.25E3f * inv;
  f3 ex = mk(-BR.x * sR + BM.x * sM, -BR.y * sR + BM.y * sM, -BR.z * sR + BM.z * sM);
  f3 fex = mk(expf(ex.x), expf(ex.y), expf(ex.z));
  float cosTheta = dot(sunDir, dir);
  float cc = cosTheta * 0.5f + 0.5f;
  float rPhase = 0.05968310365946075f * (1.0f + cc * cc);
  f3 bRT = BR * rPhase;
  float mPhase = hg(cosTheta, 0.8f);
  f3 bMT = BM * mPhase;
  f3 betas = mk((bRT.x + bMT.x) / BS.x, (bRT.y + bMT.y) / BS.y, (bRT.z + bMT.z) / BS.z);
  f3 sb = sunE * betas;
  f3 a = mk(sb.x * (1.0f - fex.x), sb.y * (1.0f - fex.y), sb.z * (1.0f - fex.z));
  f3 b = mk(sb.x * fex.x, sb.y * fex.y, sb.z * fex.z);
  float yc = fc.ydot_clamped;
  f3 Lin = mk(powf(a.x, 1.5f), powf(a.y, 1.5f), powf(a.z, 1.5f));
  Lin = mk(Lin.x * mixf(1.0f, powf(b.x, 0.5f), yc), Lin.y * mixf(1.0f, powf(b.y, 0.5f), yc),
           Lin.z * mixf(1.0f, powf(b.z, 0.5f), yc));
  f3 L0 = 0.1f * fex;
  float sunDisk = smoothstepf(fc.sun_disk_lo, fc.sun_disk_hi, cosTheta);
  f3 sfx = (sunE * 15000.0f) * fex;
  L0 = L0 + sfx * sunDisk;
  f3 col = (Lin + L0) * 0.04f;
  return col + mk(0.0f, 0.0003f, 0.00075f);
}

// getDensityHeightGradientForPoint, :475-487
__device__ __forceinline__ float height_gradient(float h, float type) {
  h = clampf(h, 0.0f, 1.0f);
  float sc = fmaxf(0.0f, remap(h, 0.0f, 0.25f, 0.0f, 1.0f) * remap(h, 0.3f, 0.65f, 1.0f, 0.0f));
  float st = fmaxf(0.0f, remap(h, 0.0f, 0.1f, 0.0f, 1.0f) * remap(h, 0.2f, 0.3f, 1.0f, 0.0f));
  float a = mixf(st, sc, clampf(type * 2.0f, 0.0f, 1.0f));
  float b = mixf(sc, st, clampf((type - 0.5f) * 2.0f, 0.0f, 1.0f));
  return mixf(a, b, type);
}


}  // namespace mtc

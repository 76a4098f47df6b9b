// runtime_frame.cpp -- per-frame constants of the cloud stage.
//
// Every ray of a frame starts at the eye, so the parts of
// cloudRayMarch.comp that depend only on uniforms are evaluated once here,
// in FP32, with the same operation order the shader uses per ray:
//   lightDir, cone kernel        cloudRayMarch.comp:585-624
//   per-step jitter table        :631-633 with getJitterOffset :114-132
//   wind translation             :495
//   Preetham per-frame factors   :406-425, :431-434, :452-453, :460
// Compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "runtime_internal.h"

namespace {
struct f3 {
  float x, y, z;
};
inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 mul(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline f3 mulv(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// GLSL normalize = v / length(v)
inline f3 normalize(f3 a) {
  float l = std::sqrt(dot(a, a));
  return mk(a.x / l, a.y / l, a.z / l);
}
inline f3 cross(f3 a, f3 b) { return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
inline float clampf(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }
void put(float* d, f3 v) {
  d[0] = v.x;
  d[1] = v.y;
  d[2] = v.z;
}
bool finite_all(const float* p, int n) {
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}
}  // namespace

int build_frame_consts(const mtc_camera* cur, const mtc_camera* old, const mtc_time* tm, const mtc_sun* sun,
                       const mtc_cloud_params* p, int W, int H, FrameConsts* fc) {
  if (!cur || !tm || !sun || !p || !fc) return mtc_fail(MTC_ERR_INVALID_ARG, "frame: null uniform pointer");
  if (!finite_all(cur->view, 16) || !finite_all(cur->eye, 3) || !finite_all(cur->tan_fov_by2, 2))
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: camera uniform holds NaN/Inf");
  if (old && !finite_all(old->view, 16)) return mtc_fail(MTC_ERR_INVALID_ARG, "frame: old camera holds NaN/Inf");
  if (tm->frame_count_mod16 < 0 || tm->frame_count_mod16 > 15)
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: frame_count_mod16 must be in 0..15");
  if (!std::isfinite(tm->time[1])) return mtc_fail(MTC_ERR_INVALID_ARG, "frame: time is not finite");
  if (!finite_all(tm->halton1, 4) || !finite_all(tm->halton2, 4) || !finite_all(tm->halton3, 4) ||
      !finite_all(tm->halton4, 4))
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: Halton table holds NaN/Inf");
  if (!finite_all(sun->sun_location, 3) || !std::isfinite(sun->sun_intensity) || !finite_all(p->sky_sun_location, 3))
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: sun uniform holds NaN/Inf");
  if (!std::isfinite(p->earth_radius) || !std::isfinite(p->shell_inner_radius) || !std::isfinite(p->shell_outer_radius))
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: shell radii must be finite");
  if (!finite_all(p->wind_direction, 3) || !std::isfinite(p->cloud_speed) || !std::isfinite(p->weather_scale))
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: wind/weather parameters must be finite");
  if (!(p->shell_inner_radius > p->earth_radius) || !(p->shell_outer_radius > p->shell_inner_radius))
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: need earth_radius < shell_inner_radius < shell_outer_radius");
  if (p->step_mode != MTC_STEPS_REFERENCE && p->step_mode != MTC_STEPS_FIXED)
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: unknown step_mode");
  if (p->steps_max < 1 || p->steps_min < 1 || p->steps_max > 4096 || p->steps_min > 4096)
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: steps must be in 1..4096");
  if (p->light_samples < 0 || p->light_samples > 6)
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: light_samples must be in 0..6");
  if (p->coverage_mode != MTC_COVERAGE_CONSTANT && p->coverage_mode != MTC_COVERAGE_WEATHER)
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: unknown coverage_mode");
  if (p->motion_blur_samples < 1 || p->motion_blur_samples > 64)
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: motion_blur_samples must be in 1..64");
  if (!(p->jitter_dim > 0.0f) || !(p->sample_divisor > 0.0f))
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: jitter_dim and sample_divisor must be > 0");
  if (!(p->cloud_fade_end > p->cloud_fade_start))
    return mtc_fail(MTC_ERR_INVALID_ARG, "frame: cloud_fade_end must exceed cloud_fade_start");

  std::memset(fc, 0, sizeof(*fc));
  std::memcpy(fc->view, cur->view, sizeof(fc->view));
  if (old) std::memcpy(fc->old_view, old->view, sizeof(fc->old_view));
  else std::memcpy(fc->old_view, cur->view, sizeof(fc->old_view));
  f3 eye = mk(-cur->eye[0], -cur->eye[1], -cur->eye[2]);
  put(fc->eye, eye);
  fc->tfx = cur->tan_fov_by2[0];
  fc->tfy = cur->tan_fov_by2[1];
  fc->frame = tm->frame_count_mod16;
  fc->width = W;
  fc->height = H;
  const float* h1 = tm->halton1;
  const float* h2 = tm->halton2;
  const float* h3 = tm->halton3;
  const float* h4 = tm->halton4;
  {
    int idx = fc->frame / 2;
    float jx, jy;
    if (idx < 4) { jx = h1[idx]; jy = h2[idx]; }
    else { jx = h3[idx - 4]; jy = h4[idx - 4]; }
    fc->jit_march[0] = jx / (float)W;
    fc->jit_march[1] = jy / (float)H;
    if (idx < 4) { jx = h1[idx]; jy = h2[idx]; }
    else { jx = h1[idx - 4]; jy = h2[idx - 4]; }
    fc->jit_reproj[0] = jx / (float)W;
    fc->jit_reproj[1] = jy / (float)H;
  }
  for (int k = 0; k < 16; ++k) {
    int idx = k / 2;
    float jx, jy;
    if (idx < 4) { jx = h1[idx]; jy = h2[idx]; }
    else { jx = h3[idx - 4]; jy = h4[idx - 4]; }
    jx = jx / p->jitter_dim;
    jy = jy / p->jitter_dim;
    fc->step_jitter[k][0] = jx;
    fc->step_jitter[k][1] = (jx + jy) * p->jitter_factor;
    fc->step_jitter[k][2] = jy;
  }
  fc->earth_r = (float)p->earth_radius;
  fc->r_in = (float)p->shell_inner_radius;
  fc->r_out = (float)p->shell_outer_radius;
  fc->thickness = fc->r_out - fc->r_in;
  fc->inner_off = fc->r_in - fc->earth_r;
  fc->earth_r_d = p->earth_radius;
  fc->r_in_d = p->shell_inner_radius;
  fc->r_out_d = p->shell_outer_radius;
  fc->step_mode = p->step_mode;
  fc->steps_min = p->steps_min;
  fc->steps_max = p->steps_max;
  fc->light_samples = p->light_samples;
  fc->coverage_mode = p->coverage_mode;
  fc->blur_samples = p->motion_blur_samples;
  {
    // reprojection.comp castRay (:111-144) and the inner-shell hit's
    // rO = (rO - c) / R (:147-155): the same FP32 operations the shader
    // runs per pixel, done once
    const float* v = cur->view;
    f3 right = normalize(mk(v[0], v[4], v[8]));
    f3 up = normalize(mk(v[1], v[5], v[9]));
    f3 lk = normalize(mk(v[2], v[6], v[10]));
    f3 look = mk(-lk.x, -lk.y, -lk.z);
    put(fc->rp_right, right);
    put(fc->rp_up, up);
    put(fc->rp_ref, add(eye, look));
    const float Rin = (float)p->shell_inner_radius;
    f3 c = mk(eye.x, -(float)p->earth_radius, eye.z);
    f3 ro = sub(eye, c);
    ro = mk(ro.x / Rin, ro.y / Rin, ro.z / Rin);
    put(fc->rp_ro, ro);
    fc->rp_C = dot(ro, ro) - 1.0f;
    for (int i = 0; i < p->motion_blur_samples; ++i) fc->rp_frac[i] = (float)i / (float)p->motion_blur_samples;
  }
  fc->base_density_factor = p->base_density_factor;
  fc->sample_divisor = p->sample_divisor;
  fc->erosion_scale = p->erosion_scale;
  fc->accum_scale = p->accum_scale;
  fc->cone_erosion_scale = p->cone_erosion_scale;
  fc->brightness = p->brightness;
  fc->hg_g = p->hg_eccentricity;
  fc->hg_si = p->hg_silver_intensity;
  fc->hg_g2 = 0.99f - p->hg_silver_spread;
  f3 wd = mk(p->wind_direction[0], p->wind_direction[1], p->wind_direction[2]);
  put(fc->wind_dir, wd);
  fc->top_offset = p->cloud_top_offset;
  fc->wind_skew = p->wind_skew;
  put(fc->wind_offset, mul(mul(add(wd, mk(0.0f, p->wind_upward_bias, 0.0f)), p->cloud_speed), tm->time[1]));
  fc->coverage_const = p->coverage_constant;
  fc->weather_scale = p->weather_scale;
  fc->bg_min = p->background_min;
  fc->fade_start = p->cloud_fade_start;
  fc->fade_end = p->cloud_fade_end;

  // lighting (cloudRayMarch.comp:585-624)
  f3 sunloc = mk(sun->sun_location[0], sun->sun_location[1], sun->sun_location[2]);
  f3 ld = normalize(sub(sunloc, eye));
  put(fc->light_dir, ld);
  float ax = std::fabs(ld.x), ay = std::fabs(ld.y), az = std::fabs(ld.z);
  f3 mc;
  if (ax > ay && ax > az) mc = mk(ax, 0.0f, 0.0f);
  else if (ay > ax && ay > az) mc = mk(0.0f, ay, 0.0f);
  else mc = mk(0.0f, 0.0f, az);
  f3 zc = cross(ld, mc);
  f3 xc = cross(zc, ld);
  static const float kv[6][3] = {{0.1f, 0.25f, -0.15f}, {0.2f, 0.5f, 0.2f},   {-0.2f, 0.1f, -0.1f},
                                 {-0.05f, 0.75f, 0.05f}, {-0.1f, 1.0f, 0.0f}, {0.0f, 3.0f, 0.0f}};
  for (int i = 0; i < 6; ++i) put(fc->kern[i], add(add(mul(xc, kv[i][0]), mul(ld, kv[i][1])), mul(zc, kv[i][2])));

  // Preetham per-frame parts (cloudRayMarch.comp:406-467)
  f3 sky = mk(p->sky_sun_location[0], p->sky_sun_location[1], p->sky_sun_location[2]);
  f3 sdir = normalize(sub(sky, eye));
  put(fc->sky_sun_dir, sdir);
  float zcos = clampf(normalize(sky).y, -1.0f, 1.0f);
  float sunI = 1000.0f * std::fmax(0.0f, 1.0f - std::pow(2.718281828459f, -((1.6110731557f - std::acos(zcos)) / 1.5f)));
  fc->sunE = sun->sun_intensity * sunI;
  float sunFade = 1.0f - clampf(1.0f - std::exp(sky.y / 450000.0f), 0.0f, 1.0f);
  f3 rt = mk(5.804542996261093E-6f, 1.3562911419845635E-5f, 3.0265902468824876E-5f);
  f3 br = mul(rt, 2.0f - 1.0f + sunFade);
  float c = (0.2f * 10.0f) * 10E-18f;
  f3 mcst = mk(1.839991851443397f, 2.779802391966052f, 4.079047954386109f);
  f3 bm = mul(mul(mcst, 0.434f * c), 0.005f);
  put(fc->beta_r, br);
  put(fc->beta_m, bm);
  put(fc->beta_sum, add(br, bm));
  float yd = 1.0f - sdir.y;
  yd = yd * (yd * yd * yd * yd);
  fc->ydot_clamped = clampf(yd, 0.0f, 1.0f);
  const float sac = (float)0.999956676946448443553574619906976478926848692873900859324;
  fc->sun_disk_lo = sac;
  fc->sun_disk_hi = sac + 0.00002f;
  // per-frame values of the optimised marches (their per-ray set-up did
  // the same FP32 operations)
  fc->mk_inv_thick = 1.0f / fc->thickness;
  fc->mk_inv_div = 1.0f / fc->sample_divisor;
  fc->mk_rel_y0 = fc->eye[1] - fc->inner_off;
  fc->mk_c_y0 = fc->eye[1] + fc->earth_r;
  const float wsk = fc->top_offset * fc->wind_skew;
  for (int i = 0; i < 3; ++i) fc->mk_wk[i] = fc->wind_dir[i] * wsk;
  fc->mk_cap = 4 * std::max(fc->steps_min, fc->steps_max) + 64;
  (void)mulv;
  return MTC_OK;
}

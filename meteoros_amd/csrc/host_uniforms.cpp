// host_uniforms.cpp -- host-side producers of the stage's uniform blocks.
//
// The reference builds its uniforms with glm on the host every frame:
//   Camera      camera.cpp:3-117 (lookAt, perspective ZO, y-flip, tanFovBy2)
//   Time        Scene.cpp:65-138 (Halton base 3, frameCount mod 16)
//   SunAndSky   Sky.cpp:64-74
// These functions restate that arithmetic in float with glm's operation
// order (glm 0.9.9.0: normalize = v * inversesqrt(dot), dot = (x+y)+z,
// mat4*vec4 = (c0*x + c1*y) + (c2*z + c3*w)); this file is compiled with
// -ffp-contract=off so each operation rounds once.
#include <cmath>
#include <cstring>

#include "meteoros_cloud.h"
#include "runtime_internal.h"

namespace {

struct f3 {
  float x, y, z;
};
inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 mul(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline float dot(f3 a, f3 b) {
  float tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
  return (tx + ty) + tz;
}
inline f3 cross(f3 x, f3 y) {
  return mk(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
// glm::normalize(v) = v * inversesqrt(dot(v, v)), inversesqrt = 1 / sqrt
inline f3 normalize(f3 v) {
  float inv = 1.0f / std::sqrt(dot(v, v));
  return mul(v, inv);
}
inline float radians(float deg) { return deg * 0.01745329251994329576923690768489f; }

f3 load3(const float* p) { return mk(p[0], p[1], p[2]); }
void store3(float* p, f3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}

// Camera::RecomputeAttributes (camera.cpp:68-77)
void recompute(mtc_camera_state* c) {
  f3 eye = load3(c->eye), ref = load3(c->ref), wu = load3(c->world_up);
  f3 fwd = normalize(sub(ref, eye));
  f3 right = normalize(cross(fwd, wu));
  f3 up = cross(right, fwd);
  store3(c->forward, fwd);
  store3(c->right, right);
  store3(c->up, up);
  c->aspect = (float)c->width / (float)c->height;
}

// glm::rotate(mat4(1), a, axis) applied to vec4(v, 1) (camera.cpp:79-96)
f3 rotate_point(float a, f3 axis_in, f3 v) {
  float c = std::cos(a), s = std::sin(a);
  f3 axis = normalize(axis_in);
  f3 temp = mul(axis, 1.0f - c);
  float R[3][3];
  R[0][0] = c + temp.x * axis.x;
  R[0][1] = temp.x * axis.y + s * axis.z;
  R[0][2] = temp.x * axis.z - s * axis.y;
  R[1][0] = temp.y * axis.x - s * axis.z;
  R[1][1] = c + temp.y * axis.y;
  R[1][2] = temp.y * axis.z + s * axis.x;
  R[2][0] = temp.z * axis.x + s * axis.y;
  R[2][1] = temp.z * axis.y - s * axis.x;
  R[2][2] = c + temp.z * axis.z;
  // Result[i] = m[0]*R[i][0] + m[1]*R[i][1] + m[2]*R[i][2] with m = I,
  // column i = (R[i][0], R[i][1], R[i][2], 0); column 3 = (0, 0, 0, 1).
  // rotation * (v, 1) = (c0*v.x + c1*v.y) + (c2*v.z + c3*1)
  f3 out;
  float a0 = R[0][0] * v.x + R[1][0] * v.y, b0 = R[2][0] * v.z + 0.0f * 1.0f;
  float a1 = R[0][1] * v.x + R[1][1] * v.y, b1 = R[2][1] * v.z + 0.0f * 1.0f;
  float a2 = R[0][2] * v.x + R[1][2] * v.y, b2 = R[2][2] * v.z + 0.0f * 1.0f;
  out.x = a0 + b0;
  out.y = a1 + b1;
  out.z = a2 + b2;
  return out;
}

void rotate_about(mtc_camera_state* c, float deg, f3 axis) {
  float a = radians(deg);
  f3 eye = load3(c->eye), ref = load3(c->ref);
  f3 r = sub(ref, eye);
  r = rotate_point(a, axis, r);
  r = add(r, eye);
  store3(c->ref, r);
  recompute(c);
}

}  // namespace

extern "C" {

int mtc_camera_init(mtc_camera_state* cam, const float eye[3], const float ref[3], int32_t width, int32_t height,
                    float fovy_deg, float near_clip, float far_clip) {
  if (!cam || !eye || !ref) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_camera_init: null pointer");
  if (width <= 0 || height <= 0) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_camera_init: width/height must be > 0");
  if (!(fovy_deg > 0.0f && fovy_deg < 180.0f))
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_camera_init: fovy must be in (0, 180)");
  if (eye[0] == ref[0] && eye[1] == ref[1] && eye[2] == ref[2])
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_camera_init: eye and ref coincide");
  std::memset(cam, 0, sizeof(*cam));
  std::memcpy(cam->eye, eye, 3 * sizeof(float));
  std::memcpy(cam->ref, ref, 3 * sizeof(float));
  cam->world_up[1] = 1.0f;
  cam->width = width;
  cam->height = height;
  cam->fovy_deg = fovy_deg;
  cam->near_clip = near_clip;
  cam->far_clip = far_clip;
  recompute(cam);
  return MTC_OK;
}

int mtc_camera_rotate_about_up(mtc_camera_state* cam, float deg) {
  if (!cam) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_camera_rotate_about_up: null camera");
  rotate_about(cam, deg, load3(cam->up));
  return MTC_OK;
}

int mtc_camera_rotate_about_right(mtc_camera_state* cam, float deg) {
  if (!cam) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_camera_rotate_about_right: null camera");
  rotate_about(cam, deg, load3(cam->right));
  return MTC_OK;
}

int mtc_camera_translate(mtc_camera_state* cam, float look, float right, float up) {
  if (!cam) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_camera_translate: null camera");
  // TranslateAlongLook / Right / Up in that order (camera.cpp:98-117)
  const float amt[3] = {look, right, up};
  for (int k = 0; k < 3; ++k) {
    if (amt[k] == 0.0f) continue;
    f3 dir = load3(k == 0 ? cam->forward : (k == 1 ? cam->right : cam->up));
    f3 t = mul(dir, amt[k]);
    store3(cam->eye, add(load3(cam->eye), t));
    store3(cam->ref, add(load3(cam->ref), t));
    recompute(cam);
  }
  return MTC_OK;
}

int mtc_camera_ubo(const mtc_camera_state* cam, mtc_camera* out) {
  if (!cam || !out) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_camera_ubo: null pointer");
  std::memset(out, 0, sizeof(*out));
  f3 eye = load3(cam->eye), center = load3(cam->ref), up = load3(cam->up);
  // glm::lookAtRH (matrix_transform.inl:754-775)
  f3 f = normalize(sub(center, eye));
  f3 s = normalize(cross(f, up));
  f3 u = cross(s, f);
  float* V = out->view;  // column-major V[c*4 + r]
  V[0] = 1; V[5] = 1; V[10] = 1; V[15] = 1;
  V[0 * 4 + 0] = s.x;
  V[1 * 4 + 0] = s.y;
  V[2 * 4 + 0] = s.z;
  V[0 * 4 + 1] = u.x;
  V[1 * 4 + 1] = u.y;
  V[2 * 4 + 1] = u.z;
  V[0 * 4 + 2] = -f.x;
  V[1 * 4 + 2] = -f.y;
  V[2 * 4 + 2] = -f.z;
  V[3 * 4 + 0] = -dot(s, eye);
  V[3 * 4 + 1] = -dot(u, eye);
  V[3 * 4 + 2] = dot(f, eye);
  // glm::perspectiveRH_ZO (matrix_transform.inl:327-340) with
  // fovy = radians(fovy), aspect = width / (float)height
  float fovy = radians(cam->fovy_deg);
  float aspect = (float)cam->width / (float)cam->height;
  float tanHalf = std::tan(fovy / 2.0f);
  float* P = out->proj;
  P[0 * 4 + 0] = 1.0f / (aspect * tanHalf);
  P[1 * 4 + 1] = 1.0f / tanHalf;
  P[2 * 4 + 2] = cam->far_clip / (cam->near_clip - cam->far_clip);
  P[2 * 4 + 3] = -1.0f;
  P[3 * 4 + 2] = -(cam->far_clip * cam->near_clip) / (cam->far_clip - cam->near_clip);
  P[1 * 4 + 1] *= -1.0f;  // y flip (camera.cpp:36)
  out->eye[0] = eye.x;
  out->eye[1] = eye.y;
  out->eye[2] = eye.z;
  out->eye[3] = 1.0f;
  // tanFovBy2 with PI = 3.14159 in double (camera.cpp:40-41, camera.h:10)
  out->tan_fov_by2[1] = (float)std::fabs(std::tan((double)cam->fovy_deg * 0.5 * (3.14159 / 180.0)));
  out->tan_fov_by2[0] = cam->aspect * out->tan_fov_by2[1];
  return MTC_OK;
}

float mtc_halton(int32_t index, int32_t base) {
  // Scene::HaltonSequenceAt (Scene.cpp:125-138)
  float f = 1.0f, r = 0.0f;
  if (base < 2) return 0.0f;
  while (index > 0) {
    f = f / (float)base;
    r += f * (float)(index % base);
    index = index / base;
  }
  return r;
}

int mtc_time_init(mtc_time* t) {
  if (!t) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_time_init: null pointer");
  std::memset(t, 0, sizeof(*t));
  for (int i = 0; i < 4; ++i) {
    t->halton1[i] = mtc_halton(1 + i, 3);
    t->halton2[i] = mtc_halton(5 + i, 3);
    t->halton3[i] = mtc_halton(9 + i, 3);
    t->halton4[i] = mtc_halton(13 + i, 3);
  }
  t->frame_count_mod16 = 0;
  return MTC_OK;
}

int mtc_time_advance(mtc_time* t, float dt) {
  if (!t) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_time_advance: null pointer");
  if (!(dt >= 0.0f) || !std::isfinite(dt)) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_time_advance: dt must be finite and >= 0");
  t->time[0] = dt;
  t->time[1] += dt;
  t->frame_count_mod16 = (t->frame_count_mod16 + 1) % 16;
  return MTC_OK;
}

int mtc_params_reference(mtc_cloud_params* p, mtc_sun* sun) {
  if (!p) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_params_reference: null params");
  std::memset(p, 0, sizeof(*p));
  p->earth_radius = 6371000.0;
  p->shell_inner_radius = 6371000.0 + 7500.0;
  p->shell_outer_radius = 6371000.0 + 20000.0;
  p->sky_sun_location[0] = 0.0f;
  p->sky_sun_location[1] = (float)(6371000.0 * 2.0);
  p->sky_sun_location[2] = (float)(-6371000.0 * 10.0);
  p->step_mode = MTC_STEPS_REFERENCE;
  p->steps_min = 35;
  p->steps_max = 60;
  p->light_samples = 6;
  p->base_density_factor = 0.380f;
  p->jitter_factor = 1.180f;
  p->jitter_dim = 75.0f;
  p->sample_divisor = 8.0f;
  p->erosion_scale = 1.4f;
  p->accum_scale = 0.5f;
  p->cone_erosion_scale = 1.5f;
  p->brightness = 5.0f;
  p->hg_eccentricity = 0.6f;
  p->hg_silver_intensity = 0.7f;
  p->hg_silver_spread = 0.1f;
  p->wind_direction[0] = 1.0f;
  p->wind_upward_bias = 0.1f;
  p->cloud_speed = 0.080f;
  p->cloud_top_offset = 1.0f;
  p->wind_skew = 0.009f;
  p->coverage_mode = MTC_COVERAGE_CONSTANT;
  p->coverage_constant = 0.6f;
  p->weather_scale = 1.0f;
  p->background_min = 0.620f;
  p->cloud_fade_start = 0.06f;
  p->cloud_fade_end = 0.2f;
  p->motion_blur_samples = 10;
  if (sun) {
    std::memset(sun, 0, sizeof(*sun));
    // SUN_LOCATION (cloudRayMarch.comp:94), SUN_INTENSITY (:99)
    sun->sun_location[0] = 0.0f;
    sun->sun_location[1] = (float)((6371000.0 + 20000.0) * 0.9);
    sun->sun_location[2] = (float)(-(6371000.0 + 20000.0) * 0.9);
    sun->sun_direction[0] = sun->sun_direction[1] = sun->sun_direction[2] = 1.0f;  // Sky.cpp:71
    sun->light_color[0] = 1.0f;
    sun->light_color[1] = 1.0f;
    sun->light_color[2] = 0.57f;
    sun->light_color[3] = 1.0f;
    sun->sun_intensity = 0.780f;
  }
  return MTC_OK;
}

int mtc_sun_from_angles(float elevation_deg, float azimuth_deg, float distance_m, mtc_sun* sun, mtc_cloud_params* p) {
  if (!sun) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_sun_from_angles: null sun");
  if (!(distance_m > 0.0f)) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_sun_from_angles: distance must be > 0");
  double el = elevation_deg * (M_PI / 180.0), az = azimuth_deg * (M_PI / 180.0);
  double dx = std::cos(el) * std::sin(az), dy = std::sin(el), dz = -std::cos(el) * std::cos(az);
  sun->sun_location[0] = (float)(dx * distance_m);
  sun->sun_location[1] = (float)(dy * distance_m);
  sun->sun_location[2] = (float)(dz * distance_m);
  sun->sun_location[3] = 0.0f;
  sun->sun_direction[0] = (float)dx;
  sun->sun_direction[1] = (float)dy;
  sun->sun_direction[2] = (float)dz;
  if (p) {
    p->sky_sun_location[0] = (float)(dx * distance_m);
    p->sky_sun_location[1] = (float)(dy * distance_m);
    p->sky_sun_location[2] = (float)(dz * distance_m);
  }
  return MTC_OK;
}

float mtc_wind_speed_units(const mtc_cloud_params* p, float mps) {
  if (!p) return 0.0f;
  double thick = p->shell_outer_radius - p->shell_inner_radius;
  return (float)(mps / (thick * (double)p->sample_divisor));
}

int mtc_params_preset(int32_t preset, mtc_cloud_params* p, mtc_sun* sun) {
  if (!p || !sun) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_params_preset: null pointer");
  if (preset < 1 || preset > 5) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_params_preset: preset must be 1..5");
  mtc_params_reference(p, sun);
  // Cloud shell R + 1500 m .. R + 4000 m, weather-map coverage.
  p->shell_inner_radius = 6371000.0 + 1500.0;
  p->shell_outer_radius = 6371000.0 + 4000.0;
  p->coverage_mode = MTC_COVERAGE_WEATHER;
  p->weather_scale = 0.5f;
  p->step_mode = MTC_STEPS_FIXED;
  switch (preset) {
    case 1: p->steps_max = 64; break;   // 480x270 CPU case
    case 2: p->steps_max = 128; break;  // 1920x1080 frame sequence
    case 3: p->steps_max = 192; break;  // 3840x2160 precision case
    case 4: p->steps_max = 64; break;   // batched views (per-view steps vary)
    case 5: p->steps_max = 128; break;  // learned light
  }
  p->steps_min = p->steps_max;
  mtc_sun_from_angles(30.0f, 0.0f, 1.0e8f, sun, p);
  sun->sun_intensity = 0.780f;
  // 10 m/s wind along +x
  p->cloud_speed = mtc_wind_speed_units(p, 10.0f);
  return MTC_OK;
}

}  // extern "C"

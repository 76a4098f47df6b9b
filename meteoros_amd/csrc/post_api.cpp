// post_api.cpp -- C ABI of the post-process passes that consume the cloud
// stage's output (include/meteoros_cloud.h):
//   mtc_god_rays   postProcess_GodRays.frag:66-150 (disabled in the
//                  reference's command buffer, Renderer.cpp:826-832; run
//                  here as a compute pass on the latest result image)
//   mtc_tone_map   postProcess_ToneMap.frag:68-84 (Renderer.cpp:834-838)
// The host folds the per-frame constants in FP32 with GLSL's operation
// order (compiled with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "runtime_internal.h"

hipError_t launch_god_rays(const PostConsts& pc, const float4* mask, float4* footprint, float4* img, hipStream_t s);
hipError_t launch_tone_map(const float4* img, int W, int H, uint32_t seed_s, float whitemap, int f32, void* out,
                           hipStream_t s);

namespace {

#define POST_OK(expr, what)                                                                       \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) {                                                                       \
      return mtc_fail(_e == hipErrorOutOfMemory ? MTC_ERR_OUT_OF_MEMORY : MTC_ERR_HIP,            \
                      std::string(what) + ": " + hipGetErrorName(_e));                            \
    }                                                                                             \
  } while (0)

bool finite_all(const float* v, int n) {
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

float norm3(float v[3]) {
  const float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] = v[0] / l;
  v[1] = v[1] / l;
  v[2] = v[2] / l;
  return l;
}

// (A x + C B) ... of postProcess_ToneMap.frag:32-35 in FP32
float uncharted2(float x) {
  const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
  return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}

}  // namespace

// God-ray constants (postProcess_GodRays.frag:72-98).  Returns false when
// the sun is behind the camera (the pass then leaves the image alone).
bool god_ray_consts(const mtc_camera* cam, const mtc_sun* sun, int W, int H, PostConsts* pc) {
  // static god rays: sunLocation = (0, 1, 0); cam_to_sun uses camera.eye as stored
  float to_sun[3] = {0.0f - cam->eye[0], 1.0f - cam->eye[1], 0.0f - cam->eye[2]};
  norm3(to_sun);
  // camForward = -normalize(view[0][2], view[1][2], view[2][2]) (column-major view[c * 4 + r])
  float fwd[3] = {cam->view[0 * 4 + 2], cam->view[1 * 4 + 2], cam->view[2 * 4 + 2]};
  norm3(fwd);
  fwd[0] = -fwd[0];
  fwd[1] = -fwd[1];
  fwd[2] = -fwd[2];
  pc->blend = (to_sun[0] * fwd[0] + to_sun[1] * fwd[1]) + to_sun[2] * fwd[2];
  // sunPos_ndc = (proj * view) * vec4(0, 1, 0, 1), no perspective divide
  float pv[16];
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      float acc = cam->proj[0 * 4 + r] * cam->view[c * 4 + 0];
      for (int k = 1; k < 4; ++k) acc = acc + cam->proj[k * 4 + r] * cam->view[c * 4 + k];
      pv[c * 4 + r] = acc;
    }
  const float p[4] = {0.0f, 1.0f, 0.0f, 1.0f};
  for (int r = 0; r < 2; ++r) {
    float acc = pv[0 * 4 + r] * p[0];
    for (int c = 1; c < 4; ++c) acc = acc + pv[c * 4 + r] * p[c];
    const float ss = (acc + 1.0f) / 2.0f;
    pc->sun_ss[r] = std::fmin(std::fmax(ss, 0.0f), 1.0f);
  }
  pc->light[0] = sun->light_color[0];
  pc->light[1] = sun->light_color[1];
  pc->light[2] = sun->light_color[2];
  pc->width = W;
  pc->height = H;
  return !(pc->blend < 0.0f);
}

extern "C" {

int mtc_god_rays(mtc_ctx* c, const mtc_camera* cam, const mtc_sun* sun, void* stream) {
  if (!c || !cam || !sun) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_god_rays: null pointer");
  if (hipSetDevice(c->desc.device) != hipSuccess) return mtc_fail(MTC_ERR_HIP, "mtc_god_rays: hipSetDevice failed");
  if (!finite_all(cam->view, 16) || !finite_all(cam->proj, 16) || !finite_all(cam->eye, 4) ||
      !finite_all(sun->light_color, 4))
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_god_rays: non-finite camera or light colour");
  const int W = c->desc.width, H = c->desc.height;
  PostConsts pc;
  if (!god_ray_consts(cam, sun, W, H, &pc)) return MTC_OK;  // sun behind the camera: nothing to add
  if (!c->ray_fp)
    POST_OK(hipMalloc((void**)&c->ray_fp, (size_t)(W + 1) * (H + 1) * sizeof(float4)), "mtc_god_rays: footprint map");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  POST_OK(launch_god_rays(pc, c->mask, c->ray_fp, c->img[c->latest], s), "mtc_god_rays: launch");
  return MTC_OK;
}

int mtc_tone_map(mtc_ctx* c, const mtc_time* time, int32_t format, void* stream) {
  if (!c || !time) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_tone_map: null pointer");
  if (format != MTC_DISPLAY_RGBA8 && format != MTC_DISPLAY_RGBA32F)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_tone_map: unknown display format");
  if (!std::isfinite(time->time[1]) || time->time[1] < 0.0f || time->time[1] >= 4294967296.0f)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_tone_map: total time must be finite and in [0, 2^32)");
  if (hipSetDevice(c->desc.device) != hipSuccess) return mtc_fail(MTC_ERR_HIP, "mtc_tone_map: hipSetDevice failed");
  const int W = c->desc.width, H = c->desc.height;
  if (c->display_format != format) {
    (void)hipDeviceSynchronize();
    (void)hipFree(c->display);
    c->display = nullptr;
    c->display_format = -1;
    const size_t bytes = (size_t)W * H * (format == MTC_DISPLAY_RGBA8 ? 4 : 16);
    POST_OK(hipMalloc(&c->display, bytes), "mtc_tone_map: display image");
    c->display_format = format;
  }
  // whitemap = 1 / Uncharted2Tonemap(vec3(100)); uint(time.y) truncates
  const float whitemap = 1.0f / uncharted2(100.0f);
  const uint32_t seed_s = (uint32_t)time->time[1];
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  POST_OK(launch_tone_map(c->img[c->latest], W, H, seed_s, whitemap, format == MTC_DISPLAY_RGBA32F, c->display, s),
          "mtc_tone_map: launch");
  return MTC_OK;
}

int mtc_read_display(mtc_ctx* c, void* out, size_t n_bytes) {
  if (!c || !out) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_read_display: null pointer");
  if (!c->display) return mtc_fail(MTC_ERR_STATE, "mtc_read_display: no tone-mapped image yet");
  const size_t need = (size_t)c->desc.width * c->desc.height * (c->display_format == MTC_DISPLAY_RGBA8 ? 4 : 16);
  if (n_bytes < need) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_read_display: buffer too small");
  if (hipSetDevice(c->desc.device) != hipSuccess) return mtc_fail(MTC_ERR_HIP, "mtc_read_display: hipSetDevice failed");
  POST_OK(hipDeviceSynchronize(), "mtc_read_display: sync");
  POST_OK(hipMemcpy(out, c->display, need, hipMemcpyDeviceToHost), "mtc_read_display: copy");
  return MTC_OK;
}

}  // extern "C"

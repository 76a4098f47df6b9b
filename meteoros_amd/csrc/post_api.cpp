// post_api.cpp -- C ABI of the post-process passes that consume the cloud
// stage's output (include/meteoros_cloud.h):
//   mtc_god_rays   postProcess_GodRays.frag:66-150 (disabled in the
//                  reference's command buffer, Renderer.cpp:826-832; run
//                  here as a compute pass on the latest result image)
//   mtc_tone_map   postProcess_ToneMap.frag:68-84 (Renderer.cpp:834-838)
//   mtc_txaa       postProcess_TXAA.frag:222-270 (Renderer.cpp:840-846)
// The host folds the per-frame constants in FP32 with GLSL's operation
// order (compiled with -ffp-contract=off).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "runtime_internal.h"

hipError_t launch_god_rays(const PostConsts& pc, const float4* mask, void* scratch, float4* img, hipStream_t s);
hipError_t god_rays_scratch(int device, void** out);
hipError_t launch_tone_map(const float4* img, int W, int H, uint32_t seed_s, float whitemap, int f32, void* out,
                           hipStream_t s);
hipError_t launch_txaa(const TxaaConsts& tc, const uint32_t* cur, const uint32_t* prev, uint32_t* out, hipStream_t s);

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

// The two RGBA8 TXAA buffers, allocated and cleared (black history) on
// first use.
hipError_t ensure_txaa(mtc_ctx* c, hipStream_t s) {
  if (c->txaa[0] && c->txaa[1]) return hipSuccess;
  const size_t bytes = (size_t)c->desc.width * c->desc.height * 4;
  for (int k = 0; k < 2; ++k) {
    if (!c->txaa[k]) {
      hipError_t e = hipMalloc((void**)&c->txaa[k], bytes);
      if (e == hipSuccess) e = hipMemsetAsync(c->txaa[k], 0, bytes, s);
      if (e != hipSuccess) return e;
    }
  }
  c->txaa_latest = 0;
  return hipStreamSynchronize(s);
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

// TXAA constants (postProcess_TXAA.frag:87-116, :225-246): castRay's basis,
// the jitter of getJitterOffset(frameCountMod16, dim) (:58-76, the march's
// table), and the inner atmosphere of the TXAA shader, EARTH_RADIUS + 7500
// with the earth centre below the eye (:229-233).
void txaa_consts(const mtc_camera* cur, const mtc_camera* old, const mtc_time* tm, int W, int H, TxaaConsts* tc) {
  const float eye[3] = {-cur->eye[0], -cur->eye[1], -cur->eye[2]};
  const float* v = cur->view;
  float right[3] = {v[0], v[4], v[8]}, up[3] = {v[1], v[5], v[9]}, lk[3] = {v[2], v[6], v[10]};
  norm3(right);
  norm3(up);
  norm3(lk);
  for (int k = 0; k < 3; ++k) {
    tc->eye[k] = eye[k];
    tc->right[k] = right[k];
    tc->up[k] = up[k];
    tc->ref[k] = eye[k] + (-lk[k]);
  }
  const int idx = tm->frame_count_mod16 / 2;
  const float jx = idx < 4 ? tm->halton1[idx] : tm->halton3[idx - 4];
  const float jy = idx < 4 ? tm->halton2[idx] : tm->halton4[idx - 4];
  tc->jit[0] = jx / (float)W;
  tc->jit[1] = jy / (float)H;
  tc->tfx = cur->tan_fov_by2[0];
  tc->tfy = cur->tan_fov_by2[1];
  std::memcpy(tc->old_view, old->view, sizeof(tc->old_view));
  const float R = 6371000.0f + 7500.0f;
  const float center[3] = {eye[0], -6371000.0f, eye[2]};
  float ro[3];
  for (int k = 0; k < 3; ++k) {
    ro[k] = (eye[k] - center[k]) / R;
    tc->ro[k] = ro[k];
    tc->center[k] = center[k];
  }
  tc->C = ((ro[0] * ro[0] + ro[1] * ro[1]) + ro[2] * ro[2]) - 1.0f;
  tc->R = R;
  tc->width = W;
  tc->height = H;
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
  if (!c->ray_scratch)
    POST_OK(hipMalloc(&c->ray_scratch, god_rays_scratch_bytes(W, H)), "mtc_god_rays: decoded mask and blur table");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  POST_OK(launch_god_rays(pc, c->mask, c->ray_scratch, c->img[c->latest], s), "mtc_god_rays: launch");
  return MTC_OK;
}

int mtc_god_rays_noalloc(int32_t width, int32_t height, const mtc_camera* cam, const mtc_sun* sun, const float* mask,
                         float* image, void* stream) {
  if (!cam || !sun || !mask || !image) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_god_rays_noalloc: null pointer");
  if (width <= 0 || height <= 0)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_god_rays_noalloc: width and height must be positive");
  if ((int64_t)(width + 1) * (int64_t)(height + 1) > (int64_t)MTC_GOD_RAYS_NOALLOC_MAX_ENTRIES)
    return mtc_fail(MTC_ERR_INVALID_ARG,
                    "mtc_god_rays_noalloc: (width + 1) * (height + 1) exceeds MTC_GOD_RAYS_NOALLOC_MAX_ENTRIES (" +
                        std::to_string(MTC_GOD_RAYS_NOALLOC_MAX_ENTRIES) + ")");
  if (height > MTC_GOD_RAYS_NOALLOC_MAX_HEIGHT)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_god_rays_noalloc: height exceeds MTC_GOD_RAYS_NOALLOC_MAX_HEIGHT (" +
                                             std::to_string(MTC_GOD_RAYS_NOALLOC_MAX_HEIGHT) + ")");
  if (width > MTC_GOD_RAYS_NOALLOC_MAX_WIDTH)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_god_rays_noalloc: width exceeds MTC_GOD_RAYS_NOALLOC_MAX_WIDTH (" +
                                             std::to_string(MTC_GOD_RAYS_NOALLOC_MAX_WIDTH) + ")");
  // the three limits above bound the scratch the pass needs by the static
  // array's size; checked here so that a change of layout cannot overrun it
  if (god_rays_scratch_bytes(width, height) > (size_t)MTC_GOD_RAYS_NOALLOC_SCRATCH_BYTES)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_god_rays_noalloc: " + std::to_string(width) + "x" +
                                             std::to_string(height) + " needs more than MTC_GOD_RAYS_NOALLOC_SCRATCH_BYTES");
  if (!finite_all(cam->view, 16) || !finite_all(cam->proj, 16) || !finite_all(cam->eye, 4) ||
      !finite_all(sun->light_color, 4))
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_god_rays_noalloc: non-finite camera or light colour");
  PostConsts pc;
  if (!god_ray_consts(cam, sun, width, height, &pc)) return MTC_OK;  // sun behind the camera: nothing to add
  int device = -1;
  POST_OK(hipGetDevice(&device), "mtc_god_rays_noalloc: hipGetDevice");
  void* scratch = nullptr;
  POST_OK(god_rays_scratch(device, &scratch), "mtc_god_rays_noalloc: scratch address");
  POST_OK(launch_god_rays(pc, (const float4*)mask, scratch, (float4*)image, (hipStream_t)stream),
          "mtc_god_rays_noalloc: launch");
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

int mtc_txaa(mtc_ctx* c, const mtc_camera* cur, const mtc_camera* old, const mtc_time* time, void* stream) {
  if (!c || !cur || !old || !time) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_txaa: null pointer");
  if (!c->display || c->display_format != MTC_DISPLAY_RGBA8)
    return mtc_fail(MTC_ERR_STATE, "mtc_txaa: needs an RGBA8 tone-mapped image (mtc_tone_map) first");
  if (!finite_all(cur->view, 16) || !finite_all(cur->eye, 4) || !finite_all(cur->tan_fov_by2, 2) ||
      !finite_all(old->view, 16))
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_txaa: non-finite camera");
  if (time->frame_count_mod16 < 0 || time->frame_count_mod16 > 15 || !finite_all(time->halton1, 4) ||
      !finite_all(time->halton2, 4) || !finite_all(time->halton3, 4) || !finite_all(time->halton4, 4))
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_txaa: bad time uniform");
  if (hipSetDevice(c->desc.device) != hipSuccess) return mtc_fail(MTC_ERR_HIP, "mtc_txaa: hipSetDevice failed");
  const int W = c->desc.width, H = c->desc.height;
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  POST_OK(ensure_txaa(c, s), "mtc_txaa: history");
  TxaaConsts tc;
  txaa_consts(cur, old, time, W, H, &tc);
  const int dst = 1 - c->txaa_latest;
  POST_OK(launch_txaa(tc, (const uint32_t*)c->display, c->txaa[c->txaa_latest], c->txaa[dst], s), "mtc_txaa: launch");
  c->txaa_latest = dst;
  return MTC_OK;
}

int mtc_read_txaa(mtc_ctx* c, uint8_t* rgba8, size_t n_bytes) {
  if (!c || !rgba8) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_read_txaa: null pointer");
  if (!c->txaa[0]) return mtc_fail(MTC_ERR_STATE, "mtc_read_txaa: no TXAA output yet");
  const size_t need = (size_t)c->desc.width * c->desc.height * 4;
  if (n_bytes < need) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_read_txaa: buffer too small");
  if (hipSetDevice(c->desc.device) != hipSuccess) return mtc_fail(MTC_ERR_HIP, "mtc_read_txaa: hipSetDevice failed");
  POST_OK(hipDeviceSynchronize(), "mtc_read_txaa: sync");
  POST_OK(hipMemcpy(rgba8, c->txaa[c->txaa_latest], need, hipMemcpyDeviceToHost), "mtc_read_txaa: copy");
  return MTC_OK;
}

int mtc_set_txaa_history(mtc_ctx* c, const uint8_t* rgba8, size_t n_bytes) {
  if (!c || !rgba8) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_set_txaa_history: null pointer");
  const size_t need = (size_t)c->desc.width * c->desc.height * 4;
  if (n_bytes != need) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_set_txaa_history: need exactly W*H*4 bytes");
  if (hipSetDevice(c->desc.device) != hipSuccess) return mtc_fail(MTC_ERR_HIP, "mtc_set_txaa_history: hipSetDevice failed");
  POST_OK(hipDeviceSynchronize(), "mtc_set_txaa_history: sync");
  POST_OK(ensure_txaa(c, c->stream), "mtc_set_txaa_history: alloc");
  POST_OK(hipMemcpy(c->txaa[c->txaa_latest], rgba8, need, hipMemcpyHostToDevice), "mtc_set_txaa_history: copy");
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

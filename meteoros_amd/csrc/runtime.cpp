// runtime.cpp -- the C ABI of the cloud compute stage (include/meteoros_cloud.h).
//
// Replaces the compute half of Renderer (Renderer.cpp:653-722, :1034-1086,
// :1428-1447) and the texture side of Sky (Sky.cpp:25-58):
//   * the context owns every device buffer: the four noise resources, the
//     two RGBA32F result images (ping-pong) and the RGBA32F god-ray mask;
//   * a frame enqueues reprojection(prev -> cur) and the march of the
//     selected pixels into cur, then swaps the images.  The reference
//     recorded both dispatches with no barrier between them, a
//     write-after-write race; here the order is explicit: plain frames run
//     the two kernels in stream order, the optimised frames one launch
//     whose reprojection leaves the marched pixels out;
//   * uniforms are taken by value at enqueue time (kernel arguments), so
//     the caller may overwrite its structs as soon as the call returns.
// Errors are status codes plus a per-thread message; nothing throws across
// the boundary.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <array>
#include <new>
#include <string>
#include <vector>

#include "meteoros_cloud.h"
#include "runtime_internal.h"

hipError_t launch_march_plain(const FrameConsts& fc, const VolumeViews& vv, const OutputViews& ov, int full,
                              hipStream_t stream, const KernelEvents& ke);
hipError_t launch_march_fast(const FrameConsts& fc, const uint4* lf, const uint4* hf, const uint4* curl,
                             const uint4* weather, const VolumeViews& vv, const OutputViews& ov, int full, int variant,
                             int lanes_per_ray, bool with_reproj, hipStream_t stream, const KernelEvents& ke,
                             const uint32_t* lf_brick);
hipError_t launch_march_learned(const FrameConsts& fc, const uint4* lf, const uint4* hf, const uint4* curl,
                                const uint4* weather, const VolumeViews& vv, const OutputViews& ov, int full,
                                int lanes_per_ray, const LearnedLight& ll, bool with_reproj, const TileOrder& to,
                                hipStream_t s, const KernelEvents& ke);
hipError_t pack_fast_resources(const VolumeViews& vv, uint4* lf, uint4* hf, uint4* curl, uint4* weather,
                               uint32_t* lf_brick, hipStream_t s);
hipError_t launch_reproject(const FrameConsts& fc, const OutputViews& ov, int mode, hipStream_t stream,
                            const KernelEvents& ke = KernelEvents());
hipError_t launch_noise(int kind, void* out, int N, int fmt, uint32_t seed, hipStream_t stream);

namespace {
thread_local std::string g_last_error = "no error";

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
int log2i(int v) {
  int s = 0;
  while ((1 << s) < v) ++s;
  return s;
}
}  // namespace

int mtc_fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr, what)                                                                            \
  do {                                                                                                 \
    hipError_t _e = (expr);                                                                            \
    if (_e != hipSuccess) {                                                                            \
      int _code = (_e == hipErrorOutOfMemory) ? MTC_ERR_OUT_OF_MEMORY : MTC_ERR_HIP;                   \
      return mtc_fail(_code, std::string(what) + ": " + hipGetErrorName(_e) + " (" + hipGetErrorString(_e) + ")"); \
    }                                                                                                  \
  } while (0)

namespace {

size_t texel_bytes(int fmt) { return fmt == MTC_FMT_FP16 ? 8 : 4; }

hipEvent_t take_event(mtc_ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Fold the oldest `keep`-exceeding pending frames into the totals.  A
// frame's events: reprojection start / end (on the stream it ran on, null
// for a full march), march start / end (on the launch stream).
hipError_t fold_timings(mtc_ctx* c, size_t keep) {
  while (c->pending.size() > keep) {
    auto ev = c->pending.front();
    c->pending.erase(c->pending.begin());
    hipError_t e = hipEventSynchronize(ev[3]);
    if (e == hipSuccess && ev[1]) e = hipEventSynchronize(ev[1]);
    if (e != hipSuccess) return e;
    float a = 0.0f, b = 0.0f;
    if (ev[0] && ev[1]) (void)hipEventElapsedTime(&a, ev[0], ev[1]);
    (void)hipEventElapsedTime(&b, ev[2], ev[3]);
    c->totals.frames += 1;
    c->totals.reproject_ms += a;
    c->totals.march_ms += b;
    for (hipEvent_t x : ev)
      if (x) c->event_pool.push_back(x);
  }
  return hipSuccess;
}

void release(mtc_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->desc.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)hipDeviceSynchronize();
  for (auto& ev : c->pending)
    for (hipEvent_t x : ev)
      if (x) (void)hipEventDestroy(x);
  for (hipEvent_t x : c->event_pool) (void)hipEventDestroy(x);
  (void)hipFree(c->lf);
  (void)hipFree(c->hf);
  (void)hipFree(c->curl);
  (void)hipFree(c->weather);
  (void)hipFree(c->img[0]);
  (void)hipFree(c->img[1]);
  (void)hipFree(c->mask);
  (void)hipFree(c->counters);
  (void)hipFree(c->q_lf);
  (void)hipFree(c->q_hf);
  (void)hipFree(c->q_curl);
  (void)hipFree(c->q_weather);
  (void)hipFree(c->q_lf_brick);
  (void)hipFree(c->order);
  (void)hipFree(c->ray_scratch);
  (void)hipFree(c->display);
  (void)hipFree(c->txaa[0]);
  (void)hipFree(c->txaa[1]);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}


hipStream_t pick_stream(mtc_ctx* c, void* s) { return s ? (hipStream_t)s : c->stream; }

int check_ctx(mtc_ctx* c, const char* fn) {
  if (!c) return mtc_fail(MTC_ERR_INVALID_ARG, std::string(fn) + ": null context");
  hipError_t e = hipSetDevice(c->desc.device);
  if (e != hipSuccess) return mtc_fail(MTC_ERR_HIP, std::string(fn) + ": hipSetDevice failed");
  return MTC_OK;
}

// The tile-order slots of the next learned march (read the previous
// march's filing, file into the next slot, clear the one after).
TileOrder next_tile_order(mtc_ctx* c) {
  TileOrder to;
  if (!c->order) return to;
  const size_t words = order_slot_words(c->order_cap);
  const uint64_t k = c->order_seq++;
  to.rd = c->order + ((k + 2) % 3) * words;
  to.wr = c->order + (k % 3) * words;
  to.clr = c->order + ((k + 1) % 3) * words;
  to.cap = c->order_cap;
  return to;
}

// with_reproj: the optimised march launch also carries the reprojection
// of the frame's other pixels
int launch_march(mtc_ctx* c, const FrameConsts& fc, const OutputViews& ov, int full, hipStream_t s,
                 const KernelEvents& ke = KernelEvents(), bool with_reproj = false) {
  if (with_reproj && c->desc.kernel == MTC_KERNEL_PLAIN)
    return mtc_fail(MTC_ERR_UNSUPPORTED, "march: no reprojection in the plain kernel's launch");
  VolumeViews vv = volume_views(c);
  hipError_t e;
  switch (c->desc.kernel) {
    case MTC_KERNEL_PLAIN: e = launch_march_plain(fc, vv, ov, full, s, ke); break;
    case MTC_KERNEL_FAST:
      e = launch_march_fast(fc, c->q_lf, c->q_hf, c->q_curl, c->q_weather, vv, ov, full, 0, c->lanes_per_ray,
                            with_reproj, s, ke, c->q_lf_brick);
      break;
    case MTC_KERNEL_PRECISION:
      e = launch_march_fast(fc, c->q_lf, c->q_hf, c->q_curl, c->q_weather, vv, ov, full, 1, c->lanes_per_ray,
                            with_reproj, s, ke, c->q_lf_brick);
      break;
    case MTC_KERNEL_LEARNED: {
      LearnedLight ll;
      int rc = learned_light_of(c, &ll);
      if (rc) return rc;
      e = launch_march_learned(fc, c->q_lf, c->q_hf, c->q_curl, c->q_weather, vv, ov, full, c->lanes_per_ray, ll,
                               with_reproj, next_tile_order(c), s, ke);
      break;
    }
    default: return mtc_fail(MTC_ERR_UNSUPPORTED, "march: kernel variant not available in this build");
  }
  HIP_TRY(e, "march kernel launch");
  return MTC_OK;
}

}  // namespace

VolumeViews volume_views(const mtc_ctx* c) {
  VolumeViews v;
  v.lf = c->lf;
  v.hf = c->hf;
  v.curl = c->curl;
  v.weather = c->weather;
  v.lf_dim = c->desc.volumes.lf_dim;
  v.hf_dim = c->desc.volumes.hf_dim;
  v.curl_dim = c->desc.volumes.curl_dim;
  v.weather_dim = c->desc.volumes.weather_dim;
  v.lf_shift = log2i(v.lf_dim);
  v.hf_shift = log2i(v.hf_dim);
  v.curl_shift = log2i(v.curl_dim);
  v.weather_shift = log2i(v.weather_dim);
  v.format = c->desc.volumes.format;
  return v;
}

extern "C" {

const char* mtc_last_error(void) { return g_last_error.c_str(); }

const char* mtc_version(void) { return "meteoros_cloud 0.1 (gfx950)"; }

int mtc_device_count(int32_t* count) {
  if (!count) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_device_count: null pointer");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return MTC_OK;
}

int mtc_create(const mtc_desc* d, mtc_ctx** out) {
  if (!d || !out) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_create: null pointer");
  *out = nullptr;
  if (d->width <= 0 || d->height <= 0 || d->width > 16384 || d->height > 16384)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_create: width/height must be in 1..16384");
  const mtc_volume_desc& v = d->volumes;
  if (!is_pow2(v.lf_dim) || !is_pow2(v.hf_dim) || !is_pow2(v.curl_dim) || !is_pow2(v.weather_dim))
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_create: volume and map dimensions must be powers of two");
  if (v.lf_dim > 1024 || v.hf_dim > 1024 || v.curl_dim > 8192 || v.weather_dim > 8192)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_create: volume dimension too large");
  if (v.format != MTC_FMT_UNORM8 && v.format != MTC_FMT_FP16)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_create: unknown volume format");
  if (d->kernel < MTC_KERNEL_PLAIN || d->kernel > MTC_KERNEL_LEARNED)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_create: unknown kernel variant");
  if (d->kernel != MTC_KERNEL_PLAIN && (v.lf_dim > 512 || v.hf_dim > 512))
    return mtc_fail(MTC_ERR_INVALID_ARG,
                    "mtc_create: the fast, precision and learned kernels take volumes up to 512^3 "
                    "(32-bit offsets into their footprint copies); MTC_KERNEL_PLAIN takes 1024^3");
  if (d->kernel == MTC_KERNEL_LEARNED && d->lanes_per_ray != 0 && d->lanes_per_ray != 4 && d->lanes_per_ray != 8)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_create: the learned-light kernel takes lanes_per_ray 0, 4 or 8");
  if (d->lanes_per_ray != 0 && d->lanes_per_ray != 1 && d->lanes_per_ray != 2 && d->lanes_per_ray != 4 &&
      d->lanes_per_ray != 8)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_create: lanes_per_ray must be 0, 1, 2, 4 or 8");
  if (d->flags & ~(MTC_FLAG_COUNTERS | MTC_FLAG_FULL_MARCH | MTC_FLAG_TIMING | MTC_FLAG_LDS_BRICKS))
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_create: unknown flags");
  if ((d->flags & MTC_FLAG_LDS_BRICKS) &&
      ((d->kernel != MTC_KERNEL_FAST && d->kernel != MTC_KERNEL_PRECISION) ||
       (d->lanes_per_ray != 0 && d->lanes_per_ray != 8) || v.lf_dim < 8))
    return mtc_fail(MTC_ERR_INVALID_ARG,
                    "mtc_create: MTC_FLAG_LDS_BRICKS needs the fast or precision kernel, 8 lanes per ray and lf_dim >= 8");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return mtc_fail(MTC_ERR_HIP, "mtc_create: no HIP device available");
  if (d->device < 0 || d->device >= ndev) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_create: device ordinal out of range");
  HIP_TRY(hipSetDevice(d->device), "mtc_create: hipSetDevice");

  mtc_ctx* c = new (std::nothrow) mtc_ctx;
  if (!c) return mtc_fail(MTC_ERR_OUT_OF_MEMORY, "mtc_create: host allocation failed");
  c->desc = *d;
  c->lanes_per_ray = d->lanes_per_ray ? d->lanes_per_ray : 8;
  c->lf_bytes = (size_t)v.lf_dim * v.lf_dim * v.lf_dim * texel_bytes(v.format);
  c->hf_bytes = (size_t)v.hf_dim * v.hf_dim * v.hf_dim * texel_bytes(v.format);
  c->curl_bytes = (size_t)v.curl_dim * v.curl_dim * 4;
  c->weather_bytes = (size_t)v.weather_dim * v.weather_dim * 4;
  const size_t px = (size_t)d->width * d->height;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&c->lf, c->lf_bytes);
  if (e == hipSuccess) e = hipMalloc(&c->hf, c->hf_bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&c->curl, c->curl_bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&c->weather, c->weather_bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&c->img[0], px * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc((void**)&c->img[1], px * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc((void**)&c->mask, px * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc((void**)&c->counters, 16 * sizeof(unsigned long long));
  if (d->kernel != MTC_KERNEL_PLAIN) {
    const size_t lf3 = (size_t)v.lf_dim * v.lf_dim * v.lf_dim, hf3 = (size_t)v.hf_dim * v.hf_dim * v.hf_dim;
    if (e == hipSuccess) e = hipMalloc((void**)&c->q_lf, 2 * lf3 * sizeof(uint4));  // 32-byte 2x2x2 LF entries
    if (e == hipSuccess) e = hipMalloc((void**)&c->q_hf, hf3 * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc((void**)&c->q_curl, (size_t)v.curl_dim * v.curl_dim * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc((void**)&c->q_weather, (size_t)v.weather_dim * v.weather_dim * sizeof(uint4));
    if (e == hipSuccess && (d->flags & MTC_FLAG_LDS_BRICKS)) e = hipMalloc((void**)&c->q_lf_brick, lf3 * sizeof(uint32_t));
    if (d->kernel == MTC_KERNEL_LEARNED) {
      // tile order of the learned march, sized for the 4x4 frames' tiles of
      // 8 x 32/G rays (full frames have more tiles and go unordered)
      const uint32_t rays_x = (uint32_t)(d->width + 3) / 4, rays_y = (uint32_t)(d->height + 3) / 4;
      const uint32_t ty = 32u / (uint32_t)c->lanes_per_ray;
      c->order_cap = ((rays_x + 7u) / 8u) * ((rays_y + ty - 1u) / ty);
      const size_t order_bytes = 3 * order_slot_words(c->order_cap) * sizeof(uint32_t);
      if (e == hipSuccess) e = hipMalloc((void**)&c->order, order_bytes);
      if (e == hipSuccess) e = hipMemsetAsync(c->order, 0, order_bytes, c->stream);
    }
  }
  if (e == hipSuccess) e = hipMemsetAsync(c->img[0], 0, px * sizeof(float4), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->img[1], 0, px * sizeof(float4), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->mask, 0, px * sizeof(float4), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->counters, 0, 16 * sizeof(unsigned long long), c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    int code = e == hipErrorOutOfMemory ? MTC_ERR_OUT_OF_MEMORY : MTC_ERR_HIP;
    std::string msg = std::string("mtc_create: ") + hipGetErrorName(e);
    release(c);
    return mtc_fail(code, msg);
  }
  *out = c;
  return MTC_OK;
}

int mtc_destroy(mtc_ctx* c) {
  if (!c) return MTC_OK;
  release(c);
  return MTC_OK;
}

int mtc_set_volumes(mtc_ctx* c, const void* lf, const void* hf, const void* curl, const void* weather) {
  int rc = check_ctx(c, "mtc_set_volumes");
  if (rc) return rc;
  if (!lf || !hf || !curl || !weather) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_set_volumes: null volume pointer");
  HIP_TRY(hipMemcpyAsync(c->lf, lf, c->lf_bytes, hipMemcpyHostToDevice, c->stream), "mtc_set_volumes: lf");
  HIP_TRY(hipMemcpyAsync(c->hf, hf, c->hf_bytes, hipMemcpyHostToDevice, c->stream), "mtc_set_volumes: hf");
  HIP_TRY(hipMemcpyAsync(c->curl, curl, c->curl_bytes, hipMemcpyHostToDevice, c->stream), "mtc_set_volumes: curl");
  HIP_TRY(hipMemcpyAsync(c->weather, weather, c->weather_bytes, hipMemcpyHostToDevice, c->stream),
          "mtc_set_volumes: weather");
  if (c->q_lf) {
    HIP_TRY(pack_fast_resources(volume_views(c), c->q_lf, c->q_hf, c->q_curl, c->q_weather, c->q_lf_brick,
                                c->stream),
            "mtc_set_volumes: footprint packing");
  }
  HIP_TRY(hipStreamSynchronize(c->stream), "mtc_set_volumes: sync");
  c->volumes_ready = true;
  return MTC_OK;
}

int mtc_generate_volumes(mtc_ctx* c, const mtc_noise_desc* nd) {
  int rc = check_ctx(c, "mtc_generate_volumes");
  if (rc) return rc;
  if (!nd) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_generate_volumes: null noise desc");
  const mtc_volume_desc& v = c->desc.volumes;
  HIP_TRY(launch_noise(0, c->lf, v.lf_dim, v.format, nd->seed_volumes, c->stream), "noise lf");
  HIP_TRY(launch_noise(1, c->hf, v.hf_dim, v.format, nd->seed_volumes, c->stream), "noise hf");
  HIP_TRY(launch_noise(2, c->curl, v.curl_dim, MTC_FMT_UNORM8, nd->seed_volumes, c->stream), "noise curl");
  HIP_TRY(launch_noise(3, c->weather, v.weather_dim, MTC_FMT_UNORM8, nd->seed_weather, c->stream), "noise weather");
  if (c->q_lf) {
    HIP_TRY(pack_fast_resources(volume_views(c), c->q_lf, c->q_hf, c->q_curl, c->q_weather, c->q_lf_brick,
                                c->stream),
            "mtc_generate_volumes: footprint packing");
  }
  HIP_TRY(hipStreamSynchronize(c->stream), "mtc_generate_volumes: sync");
  c->volumes_ready = true;
  return MTC_OK;
}

int mtc_get_volumes(mtc_ctx* c, void* lf, void* hf, void* curl, void* weather) {
  int rc = check_ctx(c, "mtc_get_volumes");
  if (rc) return rc;
  if (!c->volumes_ready) return mtc_fail(MTC_ERR_STATE, "mtc_get_volumes: no volumes resident");
  HIP_TRY(hipStreamSynchronize(c->stream), "mtc_get_volumes: sync");
  if (lf) HIP_TRY(hipMemcpy(lf, c->lf, c->lf_bytes, hipMemcpyDeviceToHost), "mtc_get_volumes: lf");
  if (hf) HIP_TRY(hipMemcpy(hf, c->hf, c->hf_bytes, hipMemcpyDeviceToHost), "mtc_get_volumes: hf");
  if (curl) HIP_TRY(hipMemcpy(curl, c->curl, c->curl_bytes, hipMemcpyDeviceToHost), "mtc_get_volumes: curl");
  if (weather) HIP_TRY(hipMemcpy(weather, c->weather, c->weather_bytes, hipMemcpyDeviceToHost), "mtc_get_volumes: weather");
  return MTC_OK;
}

int mtc_frame(mtc_ctx* c, const mtc_camera* cur, const mtc_camera* old, const mtc_time* time, const mtc_sun* sun,
              const mtc_cloud_params* params, void* stream) {
  int rc = check_ctx(c, "mtc_frame");
  if (rc) return rc;
  if (!c->volumes_ready) return mtc_fail(MTC_ERR_STATE, "mtc_frame: volumes not set or generated");
  if (!old) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_frame: null old camera");
  FrameConsts fc;
  rc = build_frame_consts(cur, old, time, sun, params, c->desc.width, c->desc.height, &fc);
  if (rc) return rc;
  hipStream_t s = pick_stream(c, stream);
  const int dst = 1 - c->latest;
  OutputViews ov;
  ov.color = c->img[dst];
  ov.prev = c->img[c->latest];
  ov.mask = c->mask;
  ov.counters = (c->desc.flags & MTC_FLAG_COUNTERS) ? c->counters : nullptr;
  ov.reproj_idx = nullptr;
  const int full = (c->desc.flags & MTC_FLAG_FULL_MARCH) ? 1 : 0;
  if (c->desc.kernel == MTC_KERNEL_LEARNED) {  // fail before any side effect
    LearnedLight ll;
    rc = learned_light_of(c, &ll);
    if (rc) return rc;
  }
  const bool timing = (c->desc.flags & MTC_FLAG_TIMING) != 0 && (c->frames_seen++ % (uint64_t)c->timing_stride) == 0;
  // optimised 4x4 frames: one launch, the reprojection strips dealt after
  // the march tiles (march_fast.hip)
  const bool fused = !full && c->desc.kernel != MTC_KERNEL_PLAIN;
  // timing: start / end events attached to each kernel's dispatch; the
  // frame keeps its structure and no marker sits between the kernels
  std::array<hipEvent_t, 4> ev = {nullptr, nullptr, nullptr, nullptr};
  // events taken for a frame that then fails go back to the pool
  auto give_back = [&]() {
    for (hipEvent_t& x : ev)
      if (x) c->event_pool.push_back(x), x = nullptr;
  };
  if (timing) {
    HIP_TRY(fold_timings(c, 255), "mtc_frame: timing events");
    for (int k = (full || fused) ? 2 : 0; k < 4; ++k) {
      ev[k] = take_event(c);
      if (!ev[k]) {
        give_back();
        return mtc_fail(MTC_ERR_HIP, "mtc_frame: hipEventCreate failed");
      }
    }
  }
  if (!full && !fused) {
    const hipError_t e = launch_reproject(fc, ov, 0, s, KernelEvents{ev[0], ev[1]});
    if (e != hipSuccess) {
      give_back();
      return mtc_fail(MTC_ERR_HIP, std::string("reprojection kernel launch: ") + hipGetErrorName(e));
    }
  }
  rc = launch_march(c, fc, ov, full, s, KernelEvents{ev[2], ev[3]}, fused);
  if (rc) {
    give_back();
    return rc;
  }
  if (timing) c->pending.push_back(ev);
  c->latest = dst;
  return MTC_OK;
}

int mtc_march_full(mtc_ctx* c, const mtc_camera* cur, const mtc_time* time, const mtc_sun* sun,
                   const mtc_cloud_params* params, void* stream) {
  int rc = check_ctx(c, "mtc_march_full");
  if (rc) return rc;
  if (!c->volumes_ready) return mtc_fail(MTC_ERR_STATE, "mtc_march_full: volumes not set or generated");
  FrameConsts fc;
  rc = build_frame_consts(cur, nullptr, time, sun, params, c->desc.width, c->desc.height, &fc);
  if (rc) return rc;
  OutputViews ov;
  ov.color = c->img[c->latest];
  ov.prev = nullptr;
  ov.mask = c->mask;
  ov.counters = (c->desc.flags & MTC_FLAG_COUNTERS) ? c->counters : nullptr;
  ov.reproj_idx = nullptr;
  return launch_march(c, fc, ov, 1, pick_stream(c, stream));
}

int mtc_synchronize(mtc_ctx* c) {
  int rc = check_ctx(c, "mtc_synchronize");
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->stream), "mtc_synchronize");
  HIP_TRY(hipDeviceSynchronize(), "mtc_synchronize: device");
  return MTC_OK;
}

int mtc_read_output(mtc_ctx* c, float* rgba, float* mask, size_t n_floats) {
  int rc = check_ctx(c, "mtc_read_output");
  if (rc) return rc;
  const size_t need = (size_t)c->desc.width * c->desc.height * 4;
  if (n_floats < need) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_read_output: buffer smaller than W*H*4 floats");
  HIP_TRY(hipDeviceSynchronize(), "mtc_read_output: sync");
  if (rgba) HIP_TRY(hipMemcpy(rgba, c->img[c->latest], need * sizeof(float), hipMemcpyDeviceToHost), "mtc_read_output: rgba");
  if (mask) HIP_TRY(hipMemcpy(mask, c->mask, need * sizeof(float), hipMemcpyDeviceToHost), "mtc_read_output: mask");
  return MTC_OK;
}

int mtc_output_device_ptrs(mtc_ctx* c, void** rgba, void** mask) {
  if (!c) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_output_device_ptrs: null context");
  if (rgba) *rgba = c->img[c->latest];
  if (mask) *mask = c->mask;
  return MTC_OK;
}

int mtc_set_history(mtc_ctx* c, const float* rgba, size_t n_floats) {
  int rc = check_ctx(c, "mtc_set_history");
  if (rc) return rc;
  const size_t need = (size_t)c->desc.width * c->desc.height * 4;
  if (!rgba || n_floats != need) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_set_history: need exactly W*H*4 floats");
  HIP_TRY(hipDeviceSynchronize(), "mtc_set_history: sync");
  HIP_TRY(hipMemcpy(c->img[c->latest], rgba, need * sizeof(float), hipMemcpyHostToDevice), "mtc_set_history");
  return MTC_OK;
}

int mtc_reset_history(mtc_ctx* c) {
  int rc = check_ctx(c, "mtc_reset_history");
  if (rc) return rc;
  const size_t px = (size_t)c->desc.width * c->desc.height;
  HIP_TRY(hipMemsetAsync(c->img[0], 0, px * sizeof(float4), c->stream), "mtc_reset_history");
  HIP_TRY(hipMemsetAsync(c->img[1], 0, px * sizeof(float4), c->stream), "mtc_reset_history");
  HIP_TRY(hipMemsetAsync(c->mask, 0, px * sizeof(float4), c->stream), "mtc_reset_history");
  HIP_TRY(hipStreamSynchronize(c->stream), "mtc_reset_history: sync");
  c->latest = 0;
  return MTC_OK;
}

int mtc_get_counters(mtc_ctx* c, mtc_counters* out) {
  int rc = check_ctx(c, "mtc_get_counters");
  if (rc) return rc;
  if (!out) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_get_counters: null pointer");
  if (!(c->desc.flags & MTC_FLAG_COUNTERS)) return mtc_fail(MTC_ERR_STATE, "mtc_get_counters: context built without MTC_FLAG_COUNTERS");
  unsigned long long h[16];
  HIP_TRY(hipDeviceSynchronize(), "mtc_get_counters: sync");
  HIP_TRY(hipMemcpy(h, c->counters, sizeof(h), hipMemcpyDeviceToHost), "mtc_get_counters");
  out->rays_marched = h[0];
  out->rays_ocean = h[1];
  out->rays_sky = h[2];
  out->primary_steps = h[3];
  out->cloud_steps = h[4];
  out->cone_samples = h[5];
  out->cone_erosions = h[6];
  out->early_exits = h[7];
  out->mlp_queries = h[8];
  out->mlp_fallbacks = h[9];
  out->mlp_audits = h[10];
  out->mlp_audit_sq_err = h[11];
  out->lf_brick_stages = h[12];
  out->lf_brick_samples = h[13];
  return MTC_OK;
}

int mtc_reset_counters(mtc_ctx* c) {
  int rc = check_ctx(c, "mtc_reset_counters");
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(c->counters, 0, 16 * sizeof(unsigned long long), c->stream), "mtc_reset_counters");
  HIP_TRY(hipStreamSynchronize(c->stream), "mtc_reset_counters: sync");
  return MTC_OK;
}

int mtc_get_timings(mtc_ctx* c, mtc_timings* out) {
  int rc = check_ctx(c, "mtc_get_timings");
  if (rc) return rc;
  if (!out) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_get_timings: null pointer");
  if (!(c->desc.flags & MTC_FLAG_TIMING)) return mtc_fail(MTC_ERR_STATE, "mtc_get_timings: context built without MTC_FLAG_TIMING");
  HIP_TRY(fold_timings(c, 0), "mtc_get_timings");
  *out = c->totals;
  return MTC_OK;
}

int mtc_reset_timings(mtc_ctx* c) {
  int rc = check_ctx(c, "mtc_reset_timings");
  if (rc) return rc;
  HIP_TRY(fold_timings(c, 0), "mtc_reset_timings");
  c->totals = mtc_timings{0, 0.0, 0.0};
  c->frames_seen = 0;
  return MTC_OK;
}

int mtc_set_timing_stride(mtc_ctx* c, int32_t stride) {
  if (!c) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_set_timing_stride: null context");
  if (stride < 1) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_set_timing_stride: stride must be >= 1");
  c->timing_stride = stride;
  c->frames_seen = 0;
  return MTC_OK;
}

int mtc_reprojection_indices(mtc_ctx* c, const mtc_camera* cur, const mtc_camera* old, const mtc_time* time,
                             const mtc_cloud_params* params, int32_t* idx, size_t n_idx) {
  int rc = check_ctx(c, "mtc_reprojection_indices");
  if (rc) return rc;
  if (!idx || !old || !params) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_reprojection_indices: null pointer");
  const size_t need = (size_t)c->desc.width * c->desc.height * (size_t)params->motion_blur_samples;
  if (n_idx != need) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_reprojection_indices: need W*H*motion_blur_samples entries");
  mtc_sun sun;
  std::memset(&sun, 0, sizeof(sun));
  sun.sun_location[1] = 1.0f;
  FrameConsts fc;
  rc = build_frame_consts(cur, old, time, &sun, params, c->desc.width, c->desc.height, &fc);
  if (rc) return rc;
  int32_t* d = nullptr;
  HIP_TRY(hipMalloc((void**)&d, need * sizeof(int32_t)), "mtc_reprojection_indices: alloc");
  OutputViews ov;
  std::memset(&ov, 0, sizeof(ov));
  ov.reproj_idx = d;
  hipError_t e = launch_reproject(fc, ov, 1, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(idx, d, need * sizeof(int32_t), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  HIP_TRY(e, "mtc_reprojection_indices");
  return MTC_OK;
}

}  // extern "C"

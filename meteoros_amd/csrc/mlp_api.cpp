// mlp_api.cpp -- C ABI of the learned light (include/meteoros_cloud.h).
//
// No reference counterpart: the network stands in for the optical depth of
// GetLightEnergy's cone march (cloudRayMarch.comp:338-373).  A mtc_mlp owns
// the FP32 master parameters, the Adam state, the BF16 copies the kernels
// read (MFMA fragments for inference, row-major for training) and a device
// store of training samples that the march itself writes.  Training,
// inference and sample generation all run on the GPU; the host only
// orders launches, draws the initial parameters and bins the held-out
// error for the fallback rule.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "runtime_internal.h"

hipError_t mlp_pack(const float* p, void* frag, uint16_t* wb, hipStream_t s);
hipError_t mlp_train_step(const float* xs, const float* ys, int n_total, uint32_t seed, int n, const float* p,
                          const uint16_t* wb, float* grad, float* loss, hipStream_t s);
hipError_t mlp_adam(float* p, float* grad, float* m, float* v, float lr, float b1, float b2, float eps, int step,
                    hipStream_t s);
hipError_t mlp_zero(float* a, int n, hipStream_t s);
hipError_t mlp_forward(const float* xs, int n, const void* frag, const float* p, float* ys, hipStream_t s);
hipError_t launch_light_samples(const FrameConsts& fc, const uint4* lf, const uint4* hf, const uint4* curl,
                                const uint4* weather, const VolumeViews& vv, uint32_t seed, uint32_t keep, float* xs,
                                float* ys, unsigned long long* keys, unsigned long long* count, unsigned long long cap,
                                hipStream_t s);

namespace {
constexpr int kParams = MTC_MLP_PARAMS;
constexpr int kIn = MTC_MLP_INPUTS;
constexpr size_t kFragBytes = 20 * 64 * 16;   // 20 fragments x 64 lanes x 8 bf16
constexpr size_t kWbElems = 32 * 64 + 2 * 64 * 64;
constexpr int kMaxDirect = 65536;              // samples per gradient / predict call (bounded staging)
}  // namespace

struct mtc_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t capacity = 0;
  int64_t n = 0;  // stored samples
  float* xs = nullptr;
  float* ys = nullptr;
  unsigned long long* keys = nullptr;  // py << 40 | px << 20 | step of each sample
  unsigned long long* count = nullptr;
  float* p = nullptr;
  float* grad = nullptr;
  float* m1 = nullptr;
  float* m2 = nullptr;
  void* frag = nullptr;
  uint16_t* wb = nullptr;
  float* loss = nullptr;
  int loss_cap = 0;
  float* sx = nullptr;  // staging for gradient / predict calls
  float* sy = nullptr;
  int adam_step = 0;
  mtc_mlp_calibration calib;
  bool calibrated = false;
};

namespace {

#define MLP_OK(expr, what)                                                                        \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) {                                                                       \
      return mtc_fail(_e == hipErrorOutOfMemory ? MTC_ERR_OUT_OF_MEMORY : MTC_ERR_HIP,            \
                      std::string(what) + ": " + hipGetErrorName(_e));                            \
    }                                                                                             \
  } while (0)

int use(mtc_mlp* m, const char* fn) {
  if (!m) return mtc_fail(MTC_ERR_INVALID_ARG, std::string(fn) + ": null network");
  if (hipSetDevice(m->device) != hipSuccess) return mtc_fail(MTC_ERR_HIP, std::string(fn) + ": hipSetDevice failed");
  return MTC_OK;
}

// splitmix64: the parameter draw of mtc_mlp_init
uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void release(mtc_mlp* m) {
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  (void)hipFree(m->xs);
  (void)hipFree(m->ys);
  (void)hipFree(m->keys);
  (void)hipFree(m->count);
  (void)hipFree(m->p);
  (void)hipFree(m->grad);
  (void)hipFree(m->m1);
  (void)hipFree(m->m2);
  (void)hipFree(m->frag);
  (void)hipFree(m->wb);
  (void)hipFree(m->loss);
  (void)hipFree(m->sx);
  (void)hipFree(m->sy);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

void default_calibration(mtc_mlp_calibration* c) {
  std::memset(c, 0, sizeof(*c));
  c->lo = 0.0f;
  c->hi = 0.0f;
  c->mask = 0u;
  c->max_rms = INFINITY;
  c->fallback_density = 0.8f;
  c->rms_all = NAN;
  c->rms_mean_predictor = NAN;
}

// Refresh the BF16 copies after the FP32 parameters changed.
hipError_t repack(mtc_mlp* m) { return mlp_pack(m->p, m->frag, m->wb, m->stream); }

}  // namespace

int learned_light_of(const mtc_ctx* c, LearnedLight* out) {
  const mtc_mlp* m = c->mlp;
  if (!m) return mtc_fail(MTC_ERR_STATE, "learned frame: no network attached (mtc_set_learned_light)");
  out->frag = m->frag;
  out->params = m->p;
  const mtc_mlp_calibration& k = m->calib;
  out->calib_hi = k.hi;
  out->calib_inv = k.hi > k.lo ? 1.0f / (k.hi - k.lo) : 0.0f;
  out->calib_mask = k.mask;
  out->fallback_density = k.fallback_density;
  out->audit = c->mlp_audit;
  return MTC_OK;
}

extern "C" {

int mtc_mlp_create(int32_t device, int64_t capacity, mtc_mlp** out) {
  if (!out) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_create: null pointer");
  *out = nullptr;
  if (capacity < 1 || capacity > ((int64_t)1 << 28))
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_create: capacity must be in 1..2^28 samples");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return mtc_fail(MTC_ERR_HIP, "mtc_mlp_create: no HIP device");
  if (device < 0 || device >= ndev) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_create: device ordinal out of range");
  MLP_OK(hipSetDevice(device), "mtc_mlp_create: hipSetDevice");
  mtc_mlp* m = new (std::nothrow) mtc_mlp;
  if (!m) return mtc_fail(MTC_ERR_OUT_OF_MEMORY, "mtc_mlp_create: host allocation failed");
  m->device = device;
  m->capacity = capacity;
  default_calibration(&m->calib);
  hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&m->xs, (size_t)capacity * kIn * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&m->ys, (size_t)capacity * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&m->keys, (size_t)capacity * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc((void**)&m->count, sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc((void**)&m->p, kParams * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&m->grad, kParams * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&m->m1, kParams * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&m->m2, kParams * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->frag, kFragBytes);
  if (e == hipSuccess) e = hipMalloc((void**)&m->wb, kWbElems * sizeof(uint16_t));
  if (e == hipSuccess) e = hipMalloc((void**)&m->sx, (size_t)kMaxDirect * kIn * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&m->sy, (size_t)kMaxDirect * sizeof(float));
  if (e == hipSuccess) e = hipMemsetAsync(m->count, 0, sizeof(unsigned long long), m->stream);
  if (e == hipSuccess) e = mlp_zero(m->grad, kParams, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  if (e != hipSuccess) {
    int code = e == hipErrorOutOfMemory ? MTC_ERR_OUT_OF_MEMORY : MTC_ERR_HIP;
    release(m);
    return mtc_fail(code, std::string("mtc_mlp_create: ") + hipGetErrorName(e));
  }
  int rc = mtc_mlp_init(m, 1u);
  if (rc) {
    release(m);
    return rc;
  }
  *out = m;
  return MTC_OK;
}

int mtc_mlp_destroy(mtc_mlp* m) {
  if (!m) return MTC_OK;
  release(m);
  return MTC_OK;
}

int mtc_mlp_init(mtc_mlp* m, uint32_t seed) {
  int rc = use(m, "mtc_mlp_init");
  if (rc) return rc;
  // LeCun uniform: U(-sqrt(3 / fan_in), +sqrt(3 / fan_in)), zero biases
  std::vector<float> p(kParams, 0.0f);
  uint64_t s = 0x6D6C7000ull ^ ((uint64_t)seed << 17);
  auto fill = [&](int off, int count, int fan_in) {
    const double a = std::sqrt(3.0 / fan_in);
    for (int i = 0; i < count; ++i) {
      const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
      p[off + i] = (float)((2.0 * u - 1.0) * a);
    }
  };
  fill(0, 9 * 64, 9);
  fill(640, 64 * 64, 64);
  fill(4800, 64 * 64, 64);
  fill(8960, 64, 64);
  MLP_OK(hipMemcpy(m->p, p.data(), kParams * sizeof(float), hipMemcpyHostToDevice), "mtc_mlp_init: upload");
  MLP_OK(mlp_zero(m->m1, kParams, m->stream), "mtc_mlp_init: zero");
  MLP_OK(mlp_zero(m->m2, kParams, m->stream), "mtc_mlp_init: zero");
  MLP_OK(mlp_zero(m->grad, kParams, m->stream), "mtc_mlp_init: zero");
  MLP_OK(repack(m), "mtc_mlp_init: pack");
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_init: sync");
  m->adam_step = 0;
  m->calibrated = false;
  default_calibration(&m->calib);
  return MTC_OK;
}

int mtc_mlp_set_params(mtc_mlp* m, const float* params, size_t n) {
  int rc = use(m, "mtc_mlp_set_params");
  if (rc) return rc;
  if (!params || n != (size_t)kParams) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_set_params: need MTC_MLP_PARAMS floats");
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(params[i])) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_set_params: non-finite parameter");
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_set_params: sync");
  MLP_OK(hipMemcpy(m->p, params, n * sizeof(float), hipMemcpyHostToDevice), "mtc_mlp_set_params: upload");
  MLP_OK(repack(m), "mtc_mlp_set_params: pack");
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_set_params: sync");
  return MTC_OK;
}

int mtc_mlp_get_params(mtc_mlp* m, float* params, size_t n) {
  int rc = use(m, "mtc_mlp_get_params");
  if (rc) return rc;
  if (!params || n != (size_t)kParams) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_get_params: need MTC_MLP_PARAMS floats");
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_get_params: sync");
  MLP_OK(hipMemcpy(params, m->p, n * sizeof(float), hipMemcpyDeviceToHost), "mtc_mlp_get_params: download");
  return MTC_OK;
}

int mtc_mlp_collect(mtc_mlp* m, mtc_ctx* ctx, const mtc_camera* cam, const mtc_time* time, const mtc_sun* sun,
                    const mtc_cloud_params* params, uint32_t seed, float keep_fraction, int64_t* n_total) {
  int rc = use(m, "mtc_mlp_collect");
  if (rc) return rc;
  if (!ctx || !cam || !time || !sun || !params) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_collect: null pointer");
  if (ctx->desc.device != m->device) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_collect: context on another device");
  if (ctx->desc.kernel == MTC_KERNEL_PLAIN || !ctx->q_lf)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_collect: needs a FAST, PRECISION or LEARNED context");
  if (!ctx->volumes_ready) return mtc_fail(MTC_ERR_STATE, "mtc_mlp_collect: context has no volumes");
  if (!(keep_fraction > 0.0f && keep_fraction <= 1.0f))
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_collect: keep_fraction must be in (0, 1]");
  FrameConsts fc;
  rc = build_frame_consts(cam, nullptr, time, sun, params, ctx->desc.width, ctx->desc.height, &fc);
  if (rc) return rc;
  const double kf = (double)keep_fraction * 4294967296.0;
  const uint32_t keep = kf >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)kf;
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_collect: sync");
  MLP_OK(launch_light_samples(fc, ctx->q_lf, ctx->q_hf, ctx->q_curl, ctx->q_weather, volume_views(ctx), seed, keep,
                              m->xs, m->ys, m->keys, m->count, (unsigned long long)m->capacity, ctx->stream),
         "mtc_mlp_collect: launch");
  unsigned long long cnt = 0;
  MLP_OK(hipStreamSynchronize(ctx->stream), "mtc_mlp_collect: sync");
  MLP_OK(hipMemcpy(&cnt, m->count, sizeof(cnt), hipMemcpyDeviceToHost), "mtc_mlp_collect: count");
  m->n = (int64_t)std::min<unsigned long long>(cnt, (unsigned long long)m->capacity);
  if (n_total) *n_total = m->n;
  return MTC_OK;
}

int mtc_mlp_sample_count(mtc_mlp* m, int64_t* n) {
  if (!m || !n) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_sample_count: null pointer");
  *n = m->n;
  return MTC_OK;
}

int mtc_mlp_clear_samples(mtc_mlp* m) {
  int rc = use(m, "mtc_mlp_clear_samples");
  if (rc) return rc;
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_clear_samples: sync");
  const unsigned long long zero = 0;
  MLP_OK(hipMemcpy(m->count, &zero, sizeof(zero), hipMemcpyHostToDevice), "mtc_mlp_clear_samples");
  m->n = 0;
  return MTC_OK;
}

int mtc_mlp_set_samples(mtc_mlp* m, const float* xs, const float* ys, int64_t n) {
  int rc = use(m, "mtc_mlp_set_samples");
  if (rc) return rc;
  if (!xs || !ys || n < 0 || n > m->capacity)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_set_samples: null pointer or n outside 0..capacity");
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_set_samples: sync");
  MLP_OK(hipMemcpy(m->xs, xs, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice), "mtc_mlp_set_samples: xs");
  MLP_OK(hipMemcpy(m->ys, ys, (size_t)n * sizeof(float), hipMemcpyHostToDevice), "mtc_mlp_set_samples: ys");
  if (n) MLP_OK(hipMemset(m->keys, 0xFF, (size_t)n * sizeof(unsigned long long)), "mtc_mlp_set_samples: keys");
  const unsigned long long cnt = (unsigned long long)n;
  MLP_OK(hipMemcpy(m->count, &cnt, sizeof(cnt), hipMemcpyHostToDevice), "mtc_mlp_set_samples: count");
  m->n = n;
  return MTC_OK;
}

int mtc_mlp_read_samples(mtc_mlp* m, int64_t first, int64_t n, float* xs, float* ys, uint64_t* keys) {
  int rc = use(m, "mtc_mlp_read_samples");
  if (rc) return rc;
  if (first < 0 || n < 0 || first + n > m->n) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_read_samples: range outside the stored samples");
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_read_samples: sync");
  if (xs) MLP_OK(hipMemcpy(xs, m->xs + first * kIn, (size_t)n * kIn * sizeof(float), hipMemcpyDeviceToHost), "mtc_mlp_read_samples: xs");
  if (ys) MLP_OK(hipMemcpy(ys, m->ys + first, (size_t)n * sizeof(float), hipMemcpyDeviceToHost), "mtc_mlp_read_samples: ys");
  if (keys)
    MLP_OK(hipMemcpy(keys, m->keys + first, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost), "mtc_mlp_read_samples: keys");
  return MTC_OK;
}

int mtc_mlp_train(mtc_mlp* m, const mtc_mlp_train_desc* d, float* loss_curve) {
  int rc = use(m, "mtc_mlp_train");
  if (rc) return rc;
  if (!d) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_train: null descriptor");
  if (d->steps < 1 || d->steps > (1 << 24) || d->batch < 64 || d->batch > (1 << 20))
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_train: steps in 1..2^24, batch in 64..2^20");
  if (!(d->lr > 0.0f) || !(d->beta1 >= 0.0f && d->beta1 < 1.0f) || !(d->beta2 >= 0.0f && d->beta2 < 1.0f) ||
      !(d->eps > 0.0f) || d->seed == 0u)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_train: bad optimiser settings (lr > 0, betas in [0, 1), eps > 0, seed != 0)");
  const int64_t n_train = m->n - d->holdout;
  if (d->holdout < 0 || n_train < 1 || n_train > INT32_MAX)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_train: holdout leaves no training samples");
  if (m->loss_cap < d->steps) {
    (void)hipFree(m->loss);
    m->loss = nullptr;
    m->loss_cap = 0;
    MLP_OK(hipMalloc((void**)&m->loss, (size_t)d->steps * sizeof(float)), "mtc_mlp_train: loss buffer");
    m->loss_cap = d->steps;
  }
  MLP_OK(mlp_zero(m->loss, d->steps, m->stream), "mtc_mlp_train: zero");
  for (int s = 0; s < d->steps; ++s) {
    uint32_t seed = d->seed * 0x9E3779B1u + (uint32_t)s * 0x85EBCA77u;
    if (seed == 0u) seed = 1u;
    MLP_OK(mlp_train_step(m->xs, m->ys, (int)n_train, seed, d->batch, m->p, m->wb, m->grad, m->loss + s, m->stream),
           "mtc_mlp_train: step");
    MLP_OK(mlp_adam(m->p, m->grad, m->m1, m->m2, d->lr, d->beta1, d->beta2, d->eps, ++m->adam_step, m->stream),
           "mtc_mlp_train: adam");
    MLP_OK(repack(m), "mtc_mlp_train: pack");
  }
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_train: sync");
  if (loss_curve)
    MLP_OK(hipMemcpy(loss_curve, m->loss, (size_t)d->steps * sizeof(float), hipMemcpyDeviceToHost), "mtc_mlp_train: loss");
  // the old bins describe the old network
  m->calibrated = false;
  m->calib.mask = 0u;
  m->calib.lo = m->calib.hi = 0.0f;
  return MTC_OK;
}

int mtc_mlp_gradient(mtc_mlp* m, const float* xs, const float* ys, int32_t n, float* grad, float* loss) {
  int rc = use(m, "mtc_mlp_gradient");
  if (rc) return rc;
  if (!xs || !ys || !grad || n < 1 || n > kMaxDirect)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_gradient: null pointer or n outside 1..65536");
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_gradient: sync");
  MLP_OK(hipMemcpy(m->sx, xs, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice), "mtc_mlp_gradient: xs");
  MLP_OK(hipMemcpy(m->sy, ys, (size_t)n * sizeof(float), hipMemcpyHostToDevice), "mtc_mlp_gradient: ys");
  if (m->loss_cap < 1) {
    MLP_OK(hipMalloc((void**)&m->loss, sizeof(float)), "mtc_mlp_gradient: loss buffer");
    m->loss_cap = 1;
  }
  MLP_OK(mlp_zero(m->grad, kParams, m->stream), "mtc_mlp_gradient: zero");
  MLP_OK(mlp_zero(m->loss, 1, m->stream), "mtc_mlp_gradient: zero");
  MLP_OK(mlp_train_step(m->sx, m->sy, n, 0u, n, m->p, m->wb, m->grad, m->loss, m->stream), "mtc_mlp_gradient: step");
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_gradient: sync");
  MLP_OK(hipMemcpy(grad, m->grad, kParams * sizeof(float), hipMemcpyDeviceToHost), "mtc_mlp_gradient: grad");
  if (loss) MLP_OK(hipMemcpy(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost), "mtc_mlp_gradient: loss");
  MLP_OK(mlp_zero(m->grad, kParams, m->stream), "mtc_mlp_gradient: zero");
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_gradient: sync");
  return MTC_OK;
}

int mtc_mlp_predict(mtc_mlp* m, const float* xs, int32_t n, float* ys) {
  int rc = use(m, "mtc_mlp_predict");
  if (rc) return rc;
  if (!xs || !ys || n < 1 || n > kMaxDirect)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_predict: null pointer or n outside 1..65536");
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_predict: sync");
  MLP_OK(hipMemcpy(m->sx, xs, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice), "mtc_mlp_predict: xs");
  MLP_OK(mlp_forward(m->sx, n, m->frag, m->p, m->sy, m->stream), "mtc_mlp_predict: forward");
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_predict: sync");
  MLP_OK(hipMemcpy(ys, m->sy, (size_t)n * sizeof(float), hipMemcpyDeviceToHost), "mtc_mlp_predict: ys");
  return MTC_OK;
}

int mtc_mlp_calibrate(mtc_mlp* m, int64_t holdout, float max_rms, float fallback_density, mtc_mlp_calibration* out) {
  int rc = use(m, "mtc_mlp_calibrate");
  if (rc) return rc;
  if (holdout < 16 || holdout > m->n) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_calibrate: holdout outside 16..stored samples");
  if (!(max_rms > 0.0f) || !(fallback_density >= 0.0f))
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_mlp_calibrate: max_rms must be > 0 and fallback_density >= 0");
  const int64_t first = m->n - holdout;
  std::vector<float> y((size_t)holdout), pred((size_t)holdout);
  MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_calibrate: sync");
  MLP_OK(hipMemcpy(y.data(), m->ys + first, (size_t)holdout * sizeof(float), hipMemcpyDeviceToHost), "mtc_mlp_calibrate: ys");
  // predictions of the held-out inputs, kMaxDirect at a time through the staging buffer
  for (int64_t i = 0; i < holdout; i += kMaxDirect) {
    const int n = (int)std::min<int64_t>(kMaxDirect, holdout - i);
    MLP_OK(mlp_forward(m->xs + (first + i) * kIn, n, m->frag, m->p, m->sy, m->stream), "mtc_mlp_calibrate: forward");
    MLP_OK(hipMemcpyAsync(pred.data() + i, m->sy, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, m->stream),
           "mtc_mlp_calibrate: copy");
    MLP_OK(hipStreamSynchronize(m->stream), "mtc_mlp_calibrate: sync");
  }
  mtc_mlp_calibration c;
  default_calibration(&c);
  c.max_rms = max_rms;
  c.fallback_density = fallback_density;
  c.lo = *std::min_element(pred.begin(), pred.end());
  c.hi = *std::max_element(pred.begin(), pred.end());
  if (!(c.hi > c.lo)) c.hi = c.lo + 1.0f;
  const float inv = 1.0f / (c.hi - c.lo);
  double sum[16] = {0}, all = 0.0, mean = 0.0;
  for (int64_t i = 0; i < holdout; ++i) mean += y[i];
  mean /= (double)holdout;
  double var = 0.0;
  for (int64_t i = 0; i < holdout; ++i) {
    const double e = (double)pred[i] - (double)y[i];
    const int b = calib_bin(pred[i], c.hi, inv);
    sum[b] += e * e;
    c.count[b]++;
    all += e * e;
    var += ((double)y[i] - mean) * ((double)y[i] - mean);
  }
  for (int b = 0; b < 16; ++b) {
    c.rms[b] = c.count[b] ? (float)std::sqrt(sum[b] / c.count[b]) : 0.0f;
    if (c.count[b] && c.rms[b] > max_rms) c.mask |= 1u << b;
  }
  c.rms_all = (float)std::sqrt(all / (double)holdout);
  c.rms_mean_predictor = (float)std::sqrt(var / (double)holdout);
  m->calib = c;
  m->calibrated = true;
  if (out) *out = c;
  return MTC_OK;
}

int mtc_set_learned_light(mtc_ctx* ctx, mtc_mlp* m, int32_t audit_period) {
  if (!ctx || !m) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_set_learned_light: null pointer");
  if (ctx->desc.kernel != MTC_KERNEL_LEARNED)
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_set_learned_light: context is not MTC_KERNEL_LEARNED");
  if (ctx->desc.device != m->device) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_set_learned_light: network on another device");
  if (audit_period < 0 || audit_period > (1 << 20) || (audit_period & (audit_period - 1)))
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_set_learned_light: audit_period must be 0 or a power of two <= 2^20");
  // the frame kernels read the parameters on the context's stream
  if (hipSetDevice(m->device) != hipSuccess || hipStreamSynchronize(m->stream) != hipSuccess)
    return mtc_fail(MTC_ERR_HIP, "mtc_set_learned_light: sync");
  ctx->mlp = m;
  ctx->mlp_audit = audit_period;
  return MTC_OK;
}

}  // extern "C"

// march_plain.hip -- plain FP32 cloud ray-march kernel (kernel b, plain).
//
// Restates cloudRayMarch.comp's main (:690-826) and rayMarch (:565-688)
// for gfx950: one lane per marched ray, a wave64 covers an 8x8 tile of
// marched rays (32x32 pixels under the 4x4 selection), four waves per
// workgroup.  FP32 throughout in GLSL source order (-ffp-contract=off);
// the filtering is the hand-written sampler of device_common.h.  This is
// the baseline the optimised kernels are measured against.
#include "march_common.h"

namespace mtc {
namespace plain {

template <int FMT>
__device__ __forceinline__ c4 sample_lf(const VolumeViews& vv, f3 p) {
  if (FMT == MTC_FMT_FP16) return sample3d_f16((const uint2*)vv.lf, vv.lf_shift, p);
  return sample3d_u8((const uint32_t*)vv.lf, vv.lf_shift, p);
}
template <int FMT>
__device__ __forceinline__ c4 sample_hf(const VolumeViews& vv, f3 p) {
  if (FMT == MTC_FMT_FP16) return sample3d_f16((const uint2*)vv.hf, vv.hf_shift, p);
  return sample3d_u8((const uint32_t*)vv.hf, vv.hf_shift, p);
}

// sampleLowFrequency, :499-540
template <int FMT>
__device__ __forceinline__ float sample_low(const FrameConsts& fc, const VolumeViews& vv, f3 p, f3 unskewed, float h) {
  c4 n = sample_lf<FMT>(vv, p);
  float fbm = (n.g * 0.625f) + (n.b * 0.25f) + (n.a * 0.125f);
  fbm = clampf(fbm, 0.0f, 1.0f);
  float base = remapClamped(n.r, (fbm - 0.9f), 1.0f, 0.0f, 1.0f);
  float cov = fc.coverage_const;
  if (fc.coverage_mode == MTC_COVERAGE_WEATHER) {
    c4 w = sample2d_u8(vv.weather, vv.weather_shift, unskewed.x * fc.weather_scale, unskewed.z * fc.weather_scale);
    base = base * (height_gradient(h, w.g) * 0.5f);
    cov = w.r;
  }
  float bc = remapCBA(base, cov, 1.0f, 0.0f, 1.0f);
  return bc * cov;
}

// erodeCloudWithHighFrequency, :542-563
template <int FMT>
__device__ __forceinline__ float erode(const VolumeViews& vv, float base, f3 p, float h) {
  c4 cn = sample2d_u8(vv.curl, vv.curl_shift, p.x, p.y);
  float k = 1.0f - h;
  p.x = p.x + (cn.r * k) * 0.5f;
  p.y = p.y + (cn.g * k) * 0.5f;
  c4 hn = sample_hf<FMT>(vv, p);
  float fbm = (hn.r * 0.625f) + (hn.g * 0.25f) + (hn.b * 0.125f);
  float mod = clampf(mixf(fbm, 1.0f - fbm, clampf(h * 2.0f, 0.0f, 1.0f)), 0.0f, 1.0f);
  return remap(base, mod * 0.005f, 1.0f, 0.0f, 1.0f);
}

// GetLightEnergy, :331-388
__device__ __forceinline__ float light_energy(float h, float dl, float ds, float phase, float cosang, float brightness) {
  float prim = expf(-dl);
  float sec = expf(-dl);
  float att = fmaxf(remap(cosang, 0.7f, 1.0f, sec, sec * 0.25f), prim);
  float depth = 0.05f + powf(ds, clampf(remap(h * 0.125f, 0.3f, 0.85f, 0.5f, 2.0f), 0.5f, 2.0f));
  float vert = powf(clampf(remap(h * 1.5f, 0.07f, 0.34f, 0.1f, 1.0f), 0.1f, 1.0f), 0.8f);
  float insc = depth * vert;
  return att * prim * insc * phase * brightness;
}

struct Cnt {
  unsigned long long steps, cloud, cone, cone_er, early;
};

// rayMarch, :565-688
template <int FMT, bool COUNT>
__device__ f3 ray_march(const FrameConsts& fc, const VolumeViews& vv, const Ray& ray, f3 c, f3 startPos, float start_t,
                        float end_t, int pixelID, float& accum, Cnt& cnt) {
  float d = ray.d.y;
  float maxSteps;
  if (fc.step_mode == MTC_STEPS_FIXED) maxSteps = (float)fc.steps_max;
  else maxSteps = floorf(mixf((float)fc.steps_min, (float)fc.steps_max, 1.0f - d));
  float span = end_t - start_t;
  float stepSize = span / maxSteps;
  float transmittance = 1.0f;
  f3 ret = mk(0.0f, 0.0f, 0.0f);
  f3 lightDir = ld3(fc.light_dir);
  float cos_angle = dot(normalize(ray.d), lightDir);
  float HG_light = fmaxf(hg(cos_angle, fc.hg_g), fc.hg_si * hg(cos_angle, fc.hg_g2));
  const f3 wd = ld3(fc.wind_dir), woff = ld3(fc.wind_offset);
  const f3 inner = mk(c.x, fc.inner_off, c.z);

  // Iteration cap (absent in the GLSL): guarantees the loop ends even if
  // t + stepSize == t; far above any trip count of a valid scene.
  const int cap = 4 * max(fc.steps_min, fc.steps_max) + 64;
  int iter = 0;
  for (float t = start_t; t < end_t && iter < cap; t += stepSize, ++iter) {
    if (COUNT) cnt.steps++;
    int _index = (pixelID + (int)t) & 15;
    f3 jd = ld3(fc.step_jitter[_index]);
    f3 pos = ray.o + t * (ray.d + jd);
    f3 sp = (pos - inner) / fc.thickness;
    sp = sp / fc.sample_divisor;
    // getRelativeHeightInAtmosphere, :171-186
    float lr = length(pos - ray.o);
    float ls = length(startPos - ray.o);
    f3 nrm = normalize(pos - c);
    float cosT = dot(ray.d, nrm);
    float h = fabsf(cosT * (lr - ls)) / fc.thickness;
    // skewSamplePointWithWind, :489-497
    f3 sk = sp + ((h * wd) * fc.top_offset) * fc.wind_skew;
    sk = sk + woff;
    float base = sample_low<FMT>(fc, vv, sk, sp, h) * fc.base_density_factor;
    if (base > 0.0f) {
      if (COUNT) cnt.cloud++;
      float hfd = erode<FMT>(vv, base * fc.erosion_scale, sk, h);
      accum += hfd * fc.accum_scale;
      float dal = 0.0f;
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        if (i < fc.light_samples) {
          if (COUNT) cnt.cone++;
          f3 lp = pos + ((stepSize * ld3(fc.kern[i])) * (float)i);
          f3 slp = (lp - inner) / fc.thickness;
          float cb = sample_low<FMT>(fc, vv, slp, slp, h);
          if (cb > 0.0f) {
            if (COUNT) cnt.cone_er++;
            dal += erode<FMT>(vv, fc.cone_erosion_scale * cb, sk, h);
          }
        }
      }
      float E = light_energy(h, dal, base, HG_light, cos_angle, fc.brightness);
      transmittance = mixf(transmittance, E, 1.0f - accum);
      ret = ret + mk(transmittance, transmittance, transmittance);
    }
    if (accum >= 1.0f) {
      accum = 1.0f;
      if (COUNT) cnt.early++;
      break;
    }
  }
  return ret;
}

// main, :690-826.  full = 0: the 4x4 selection of frame fc.frame;
// full = 1: every pixel (jitter phase fc.frame).
template <int FMT, bool COUNT>
__global__ __launch_bounds__(256) void march_kernel(FrameConsts fc, VolumeViews vv, OutputViews ov, int full) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rx = blockIdx.x * 16 + (wave & 1) * 8 + (lane & 7);
  const int ry = blockIdx.y * 16 + (wave >> 1) * 8 + (lane >> 3);
  const int W = fc.width, H = fc.height;
  const int pixelID = fc.frame;
  int px, py;
  if (full) {
    px = rx;
    py = ry;
  } else {
    px = rx * 4 + pixelID / 4;
    py = ry * 4 + pixelID % 4;
  }
  const bool active = px < W && py < H;
  Cnt cnt = {0, 0, 0, 0, 0};
  unsigned long long n_ocean = 0, n_sky = 0, n_marched = 0;
  if (active) {
    float ux = (float)px / (float)W, uy = (float)py / (float)H;
    uy = 1.0f - uy;
    Ray ray = cast_ray(fc, fc.view, ux, uy, fc.jit_march[0], fc.jit_march[1]);
    float d = ray.d.y;
    float bgMul = fmaxf(fc.bg_min, d);
    float4 color, mask = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (d < 0.0f) {
      float a = -ray.d.y * 5.5f;
      color = make_float4(mixf(0.0f * 0.4f, 0.0f * 0.5f, a), mixf(0.16f * 0.4f, 0.73f * 0.5f, a),
                          mixf(0.51f * 0.4f, 0.95f * 0.5f, a), 1.0f);
      n_ocean = 1;
    } else {
      f3 bg = atmosphere(fc, ray.d);
      bg = bg * bgMul;
      if (d < fc.fade_start) {
        color = make_float4(bg.x, bg.y, bg.z, 1.0f);
        n_sky = 1;
      } else {
        n_marched = 1;
        f3 c = ray.o;
        c.y = -fc.earth_r;
        Isect in = ray_sphere(ray.o, ray.d, c, fc.r_in);
        Isect out = ray_sphere(ray.o, ray.d, c, fc.r_out);
        float accum = 0.0f;
        f3 col = ray_march<FMT, COUNT>(fc, vv, ray, c, in.point, in.t, out.t, pixelID, accum, cnt);
        float god = accum;
        accum = accum * smoothstepf(0.0f, 1.0f, fminf(1.0f, remap(ray.d.y, fc.fade_start, fc.fade_end, 0.0f, 1.0f)));
        color = make_float4(mixf(bg.x, col.x, accum), mixf(bg.y, col.y, accum), mixf(bg.z, col.z, accum), 1.0f);
        mask = encode_float_rgba(25.0f * fminf(0.05f, 1.0f - god));
        // :775-779, reachable only when cloud_fade_start < 0.05
        if (ray.d.y < 0.05f) {
          const float s = fmaxf(5.0f, ray.d.y);
          mask = make_float4(mask.x * s, mask.y * s, mask.z * s, mask.w * s);
        }
      }
    }
    const size_t o = (size_t)py * (size_t)W + (size_t)px;
    ov.color[o] = color;
    ov.mask[o] = mask;
  }
  if (COUNT) {
    wave_count(ov.counters + 0, n_marched);
    wave_count(ov.counters + 1, n_ocean);
    wave_count(ov.counters + 2, n_sky);
    wave_count(ov.counters + 3, cnt.steps);
    wave_count(ov.counters + 4, cnt.cloud);
    wave_count(ov.counters + 5, cnt.cone);
    wave_count(ov.counters + 6, cnt.cone_er);
    wave_count(ov.counters + 7, cnt.early);
  }
}

}  // namespace plain
}  // namespace mtc

// Host launcher: grid of 16x16-ray workgroups over the marched-ray lattice.
hipError_t launch_march_plain(const FrameConsts& fc, const VolumeViews& vv, const OutputViews& ov, int full,
                              hipStream_t stream, const KernelEvents& ke) {
  int rays_x = full ? fc.width : (fc.width + 3) / 4;
  int rays_y = full ? fc.height : (fc.height + 3) / 4;
  dim3 grid((rays_x + 15) / 16, (rays_y + 15) / 16);
  dim3 block(256);
  const bool count = ov.counters != nullptr;
  if (vv.format == MTC_FMT_FP16) {
    if (count) launch_timed(mtc::plain::march_kernel<MTC_FMT_FP16, true>, grid, block, stream, ke, fc, vv, ov, full);
    else launch_timed(mtc::plain::march_kernel<MTC_FMT_FP16, false>, grid, block, stream, ke, fc, vv, ov, full);
  } else {
    if (count) launch_timed(mtc::plain::march_kernel<MTC_FMT_UNORM8, true>, grid, block, stream, ke, fc, vv, ov, full);
    else launch_timed(mtc::plain::march_kernel<MTC_FMT_UNORM8, false>, grid, block, stream, ke, fc, vv, ov, full);
  }
  return hipGetLastError();
}

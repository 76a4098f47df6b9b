// reproject.hip -- 4x4 temporal reprojection update (kernel c).
//
// Restates reprojection.comp main (:193-244): per pixel, cast the current
// ray (no uv flip, the seq1/seq2 jitter table), hit the inner shell,
// project the hit into the old camera, and average `blur_samples` taps of
// the previous image along the motion vector.  One lane per pixel; a
// workgroup of 256 lanes covers a 64x4 pixel strip so each wave reads one
// contiguous 64-pixel row of the previous image per tap where the motion is
// small.  FP32 in GLSL order (-ffp-contract=off): the tap indices are
// bit-exact with the FP32 oracle.  The parts that depend only on uniforms
// (camera basis, eye + look, the ray origin's (rO - c) / R and C term, the
// tap fractions i / N) are folded once per frame on the host with the same
// FP32 operations (runtime_frame.cpp).
#include "device_common.h"

namespace mtc {
namespace reproj {

struct Ray {
  f3 o, d;
};

// castRay (:111-144) with the camera basis and eye + look folded per frame
__device__ __forceinline__ Ray cast_ray(const FrameConsts& fc, float ux, float uy) {
  float nx = ux * 2.0f - 1.0f, ny = uy * 2.0f - 1.0f;
  nx = nx + fc.jit_reproj[0];
  ny = ny + fc.jit_reproj[1];
  f3 cx = (nx * fc.tfx) * ld3(fc.rp_right);
  f3 cy = (ny * fc.tfy) * ld3(fc.rp_up);
  f3 eye = ld3(fc.eye);
  f3 p = (ld3(fc.rp_ref) + cx) + cy;
  Ray r;
  r.o = eye;
  r.d = normalize(p - eye);
  return r;
}

// raySphereIntersection (:147-191, point only) with the ray origin's
// (rO - c) / R and its C term folded per frame
__device__ __forceinline__ f3 inner_hit(const FrameConsts& fc, f3 rD, f3 c, float R) {
  f3 point = mk(0.0f, 0.0f, 0.0f);
  const f3 rO = ld3(fc.rp_ro);
  float A = dot(rD, rD);
  float B = 2.0f * dot(rD, rO);
  float C = fc.rp_C;
  float disc = B * B - 4.0f * A * C;
  if (disc < 0.0f) return point;
  float t = (-B - __builtin_sqrtf(disc)) / (2.0f * A);
  if (t < 0.0f) t = (-B + __builtin_sqrtf(disc)) / (2.0f * A);
  if (t >= 0.0f) {
    f3 p = rO + t * rD;
    p = p * R;
    point = p + c;
  }
  return point;
}

// SKIP_SELECTED: leave out the pixels this frame's march overwrites (the
// 4x4 phase (frame / 4, frame % 4)), so the two kernels write disjoint
// pixels and may run concurrently.
template <bool INDICES, bool SKIP_SELECTED>
__global__ __launch_bounds__(256) void reproject_kernel(FrameConsts fc, OutputViews ov) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int W = fc.width, H = fc.height;
  if (x >= W || y >= H) return;
  if (SKIP_SELECTED && (x & 3) == fc.frame / 4 && (y & 3) == fc.frame % 4) return;
  float ux = (float)x / (float)W, uy = (float)y / (float)H;
  Ray ray = cast_ray(fc, ux, uy);
  f3 c = ray.o;
  c.y = -fc.earth_r;
  f3 P = inner_hit(fc, ray.d, c, fc.r_in);
  const float* m = fc.old_view;
  // cameraOld.view * vec4(P, 1), columns summed left to right
  f3 q = mk(((m[0] * P.x + m[4] * P.y) + m[8] * P.z) + m[12] * 1.0f,
            ((m[1] * P.x + m[5] * P.y) + m[9] * P.z) + m[13] * 1.0f,
            ((m[2] * P.x + m[6] * P.y) + m[10] * P.z) + m[14] * 1.0f);
  f3 od = normalize(q);
  float nz = -od.z;
  od = od / nz;
  float ou = (od.x / fc.tfx) * 0.5f + 0.5f;
  float ovv = (od.y / fc.tfy) * 0.5f + 0.5f;
  float mx = ou - ux, my = ovv - uy;
  const int N = fc.blur_samples;
  const float fW = (float)W, fH = (float)H, maxx = (float)(W - 1), maxy = (float)(H - 1);
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const size_t o = (size_t)y * (size_t)W + (size_t)x;
  for (int i = 0; i < N; ++i) {
    const float fr = fc.rp_frac[i];  // float(i) / float(N)
    float bx = mx * fr, by = my * fr;
    // round() with ties to even, then clamp to the image (imageLoad of a
    // clamped ivec2)
    float iu = rintf((ou - bx) * fW);
    float iv = rintf((ovv - by) * fH);
    int sx = (int)fminf(fmaxf(iu, 0.0f), maxx);
    int sy = (int)fminf(fmaxf(iv, 0.0f), maxy);
    int li = sy * W + sx;
    if (INDICES) {
      ov.reproj_idx[o * (size_t)N + (size_t)i] = li;
    } else {
      float4 t = ov.prev[li];
      acc.x += t.x;
      acc.y += t.y;
      acc.z += t.z;
      acc.w += t.w;
    }
  }
  if (!INDICES) {
    const float fN = (float)N;
    ov.color[o] = make_float4(acc.x / fN, acc.y / fN, acc.z / fN, acc.w / fN);
  }
}

}  // namespace reproj
}  // namespace mtc

// mode 0: the update; 1: the tap-index map; 2: the update without the
// pixels the march of this frame writes
hipError_t launch_reproject(const FrameConsts& fc, const OutputViews& ov, int mode, hipStream_t stream) {
  dim3 grid((fc.width + 63) / 64, (fc.height + 3) / 4);
  dim3 block(256);
  if (mode == 1) hipLaunchKernelGGL((mtc::reproj::reproject_kernel<true, false>), grid, block, 0, stream, fc, ov);
  else if (mode == 2) hipLaunchKernelGGL((mtc::reproj::reproject_kernel<false, true>), grid, block, 0, stream, fc, ov);
  else hipLaunchKernelGGL((mtc::reproj::reproject_kernel<false, false>), grid, block, 0, stream, fc, ov);
  return hipGetLastError();
}

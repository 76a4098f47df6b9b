// reproject_pixel.h -- the 4x4 temporal reprojection update of one pixel
// (reprojection.comp main, :193-244).
//
// Per pixel: cast the current ray (no uv flip, the seq1/seq2 jitter table),
// hit the inner shell, project the hit into the old camera, and average
// `blur_samples` taps of the previous image along the motion vector.  FP32
// in GLSL order with one rounding per operation: every function body
// switches fp contraction off and spells its arithmetic out (no calls into
// helpers whose contraction depends on the including file), so the tap
// indices and the averages are bit-exact with the FP32 oracle whatever the
// including file's fp-contract setting.  The parts that depend only on
// uniforms (camera basis, eye + look, the ray origin's (rO - c) / R and C
// term, the tap fractions i / N) are folded once per frame on the host
// with the same FP32 operations (runtime_frame.cpp).
#pragma once
#include "device_common.h"

namespace mtc {
namespace reproj {

// castRay (:111-144) direction with the camera basis and eye + look folded
// per frame; the origin is the eye
__device__ __forceinline__ f3 ray_dir(const FrameConsts& fc, float ux, float uy) {
#pragma clang fp contract(off)
  float nx = ux * 2.0f - 1.0f, ny = uy * 2.0f - 1.0f;
  nx = nx + fc.jit_reproj[0];
  ny = ny + fc.jit_reproj[1];
  const float sx = nx * fc.tfx, sy = ny * fc.tfy;
  const float px = (fc.rp_ref[0] + sx * fc.rp_right[0]) + sy * fc.rp_up[0];
  const float py = (fc.rp_ref[1] + sx * fc.rp_right[1]) + sy * fc.rp_up[1];
  const float pz = (fc.rp_ref[2] + sx * fc.rp_right[2]) + sy * fc.rp_up[2];
  const float dx = px - fc.eye[0], dy = py - fc.eye[1], dz = pz - fc.eye[2];
  const float l = __builtin_sqrtf((dx * dx + dy * dy) + dz * dz);
  return mk(dx / l, dy / l, dz / l);
}

// raySphereIntersection (:147-191, point only) with the ray origin's
// (rO - c) / R and its C term folded per frame; c = (eye.x, -R_earth, eye.z)
__device__ __forceinline__ f3 inner_hit(const FrameConsts& fc, f3 rD) {
#pragma clang fp contract(off)
  f3 point = mk(0.0f, 0.0f, 0.0f);
  const float ox = fc.rp_ro[0], oy = fc.rp_ro[1], oz = fc.rp_ro[2];
  const float A = (rD.x * rD.x + rD.y * rD.y) + rD.z * rD.z;
  const float B = 2.0f * ((rD.x * ox + rD.y * oy) + rD.z * oz);
  const float C = fc.rp_C;
  const float disc = B * B - 4.0f * A * C;
  if (disc < 0.0f) return point;
  float t = (-B - __builtin_sqrtf(disc)) / (2.0f * A);
  if (t < 0.0f) t = (-B + __builtin_sqrtf(disc)) / (2.0f * A);
  if (t >= 0.0f) {
    const float R = fc.r_in;
    point = mk((ox + t * rD.x) * R + fc.eye[0], (oy + t * rD.y) * R + -fc.earth_r, (oz + t * rD.z) * R + fc.eye[2]);
  }
  return point;
}

// The update of pixel (x, y): the average of the taps into ov.color, or
// (INDICES) the tap indices into ov.reproj_idx.
template <bool INDICES>
__device__ __forceinline__ void reproject_pixel(const FrameConsts& fc, const OutputViews& ov, int x, int y) {
#pragma clang fp contract(off)
  const int W = fc.width, H = fc.height;
  const float ux = (float)x / (float)W, uy = (float)y / (float)H;
  const f3 d = ray_dir(fc, ux, uy);
  const f3 P = inner_hit(fc, d);
  const float* m = fc.old_view;
  // cameraOld.view * vec4(P, 1), columns summed left to right
  const float qx = ((m[0] * P.x + m[4] * P.y) + m[8] * P.z) + m[12] * 1.0f;
  const float qy = ((m[1] * P.x + m[5] * P.y) + m[9] * P.z) + m[13] * 1.0f;
  const float qz = ((m[2] * P.x + m[6] * P.y) + m[10] * P.z) + m[14] * 1.0f;
  const float ql = __builtin_sqrtf((qx * qx + qy * qy) + qz * qz);
  float odx = qx / ql, ody = qy / ql, odz = qz / ql;
  const float nz = -odz;
  odx = odx / nz;
  ody = ody / nz;
  const float ou = (odx / fc.tfx) * 0.5f + 0.5f;
  const float ovv = (ody / fc.tfy) * 0.5f + 0.5f;
  const float mx = ou - ux, my = ovv - uy;
  const int N = fc.blur_samples;
  const float fW = (float)W, fH = (float)H, maxx = (float)(W - 1), maxy = (float)(H - 1);
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const size_t o = (size_t)y * (size_t)W + (size_t)x;
  // taps in batches of kTapBatch: the batch's indices, then all its loads,
  // then the sum in tap order (the loads of a batch are in flight together
  // instead of one round trip per tap; the same operations in the same
  // order as one tap at a time)
  constexpr int kTapBatch = 8;
  for (int i0 = 0; i0 < N; i0 += kTapBatch) {
    int li[kTapBatch];
#pragma unroll
    for (int k = 0; k < kTapBatch; ++k) {
      const float fr = fc.rp_frac[min(i0 + k, N - 1)];  // float(i) / float(N)
      const float bx = mx * fr, by = my * fr;
      // round() with ties to even, then clamp to the image (imageLoad of a
      // clamped ivec2)
      const float iu = rintf((ou - bx) * fW);
      const float iv = rintf((ovv - by) * fH);
      const int sx = (int)fminf(fmaxf(iu, 0.0f), maxx);
      const int sy = (int)fminf(fmaxf(iv, 0.0f), maxy);
      li[k] = sy * W + sx;
    }
    if (INDICES) {
#pragma unroll
      for (int k = 0; k < kTapBatch; ++k)
        if (i0 + k < N) ov.reproj_idx[o * (size_t)N + (size_t)(i0 + k)] = li[k];
    } else {
      float4 t[kTapBatch];
#pragma unroll
      for (int k = 0; k < kTapBatch; ++k)
        if (i0 + k < N) t[k] = ov.prev[li[k]];
#pragma unroll
      for (int k = 0; k < kTapBatch; ++k) {
        if (i0 + k < N) {
          acc.x += t[k].x;
          acc.y += t[k].y;
          acc.z += t[k].z;
          acc.w += t[k].w;
        }
      }
    }
  }
  if (!INDICES) {
    const float fN = (float)N;
    ov.color[o] = make_float4(acc.x / fN, acc.y / fN, acc.z / fN, acc.w / fN);
  }
}

// True for the pixels the march of this frame writes (the 4x4 phase
// (frame / 4, frame % 4), main :777-786).
__device__ __forceinline__ bool marched_this_frame(const FrameConsts& fc, int x, int y) {
  return (x & 3) == fc.frame / 4 && (y & 3) == fc.frame % 4;
}

}  // namespace reproj
}  // namespace mtc

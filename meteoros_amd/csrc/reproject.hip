// reproject.hip -- 4x4 temporal reprojection update (kernel c).
//
// Restates reprojection.comp main (:193-244): per pixel, cast the current
// ray (no uv flip, the seq1/seq2 jitter table), hit the inner shell,
// project the hit into the old camera, and average `blur_samples` taps of
// the previous image along the motion vector.  One lane per pixel; a
// workgroup of 256 lanes covers a 64x4 pixel strip so each wave reads one
// contiguous 64-pixel row of the previous image per tap where the motion is
// small.  FP32 in GLSL order (-ffp-contract=off): the tap indices are
// bit-exact with the FP32 oracle.
#include "device_common.h"

namespace mtc {
namespace reproj {

struct Ray {
  f3 o, d;
};

__device__ __forceinline__ Ray cast_ray(const FrameConsts& fc, float ux, float uy) {
  const float* view = fc.view;
  f3 right = normalize(mk(view[0], view[4], view[8]));
  f3 up = normalize(mk(view[1], view[5], view[9]));
  f3 lk = normalize(mk(view[2], view[6], view[10]));
  f3 look = mk(-lk.x, -lk.y, -lk.z);
  float nx = ux * 2.0f - 1.0f, ny = uy * 2.0f - 1.0f;
  nx = nx + fc.jit_reproj[0];
  ny = ny + fc.jit_reproj[1];
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

// raySphereIntersection, reprojection.comp:147-191 (point only)
__device__ __forceinline__ f3 inner_hit(f3 rO, f3 rD, f3 c, float R) {
  f3 point = mk(0.0f, 0.0f, 0.0f);
  rO = rO - c;
  rO = rO / R;
  float A = dot(rD, rD);
  float B = 2.0f * dot(rD, rO);
  float C = dot(rO, rO) - 1.0f;
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

template <bool INDICES>
__global__ __launch_bounds__(256) void reproject_kernel(FrameConsts fc, OutputViews ov) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int W = fc.width, H = fc.height;
  if (x >= W || y >= H) return;
  float ux = (float)x / (float)W, uy = (float)y / (float)H;
  Ray ray = cast_ray(fc, ux, uy);
  f3 c = ray.o;
  c.y = -fc.earth_r;
  f3 P = inner_hit(ray.o, ray.d, c, fc.r_in);
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
    float fr = (float)i / (float)N;
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

// Tiled variant for the optimised frame path.  The taps of a pixel lie on
// the segment from old_uv (tap 0) towards uv (tap N-1), and their rounded,
// clamped indices are monotone along it, so the taps of a 64x4 tile fall in
// the bounding box of the tiles' first and last taps.  That box is staged
// in LDS with coalesced float4 loads and the taps gather from LDS; a box
// larger than the LDS tile falls back to global gathers.  Tap indices and
// summation order are those of reproject_kernel, so the output is
// bit-identical.  Pixels the march of this frame overwrites are skipped.
constexpr int kTileCap = 1536;  // float4 texels staged per workgroup (24 KiB)

template <bool SKIP_SELECTED>
__global__ __launch_bounds__(256) void reproject_tiled_kernel(FrameConsts fc, OutputViews ov) {
  __shared__ float4 tile[kTileCap];
  __shared__ int box[4];
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int W = fc.width, H = fc.height;
  const int N = fc.blur_samples;
  bool live = x < W && y < H;
  if (SKIP_SELECTED && live && (x & 3) == fc.frame / 4 && (y & 3) == fc.frame % 4) live = false;
  if (threadIdx.x == 0) {
    box[0] = 0x7fffffff;
    box[1] = 0x7fffffff;
    box[2] = -1;
    box[3] = -1;
  }
  __syncthreads();
  float ou = 0.0f, ovv = 0.0f, mx = 0.0f, my = 0.0f;
  const float fW = (float)W, fH = (float)H, maxx = (float)(W - 1), maxy = (float)(H - 1);
  if (live) {
    float ux = (float)x / (float)W, uy = (float)y / (float)H;
    Ray ray = cast_ray(fc, ux, uy);
    f3 c = ray.o;
    c.y = -fc.earth_r;
    f3 P = inner_hit(ray.o, ray.d, c, fc.r_in);
    const float* m = fc.old_view;
    f3 q = mk(((m[0] * P.x + m[4] * P.y) + m[8] * P.z) + m[12] * 1.0f,
              ((m[1] * P.x + m[5] * P.y) + m[9] * P.z) + m[13] * 1.0f,
              ((m[2] * P.x + m[6] * P.y) + m[10] * P.z) + m[14] * 1.0f);
    f3 od = normalize(q);
    float nz = -od.z;
    od = od / nz;
    ou = (od.x / fc.tfx) * 0.5f + 0.5f;
    ovv = (od.y / fc.tfy) * 0.5f + 0.5f;
    mx = ou - ux;
    my = ovv - uy;
    // first and last taps bound all of them
    int x0 = (int)fminf(fmaxf(rintf(ou * fW), 0.0f), maxx);
    int y0 = (int)fminf(fmaxf(rintf(ovv * fH), 0.0f), maxy);
    float fr = (float)(N - 1) / (float)N;
    int x1 = (int)fminf(fmaxf(rintf((ou - mx * fr) * fW), 0.0f), maxx);
    int y1 = (int)fminf(fmaxf(rintf((ovv - my * fr) * fH), 0.0f), maxy);
    atomicMin(&box[0], min(x0, x1));
    atomicMin(&box[1], min(y0, y1));
    atomicMax(&box[2], max(x0, x1));
    atomicMax(&box[3], max(y0, y1));
  }
  __syncthreads();
  const int bx0 = box[0], by0 = box[1];
  const int bw = box[2] - bx0 + 1, bh = box[3] - by0 + 1;
  const bool staged = box[2] >= 0 && bw * bh <= kTileCap;
  if (staged) {
    for (int e = threadIdx.x; e < bw * bh; e += 256) {
      int ry = e / bw;
      tile[e] = ov.prev[(size_t)(by0 + ry) * W + (bx0 + (e - ry * bw))];
    }
  }
  __syncthreads();
  if (!live) return;
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (int i = 0; i < N; ++i) {
    float fr = (float)i / (float)N;
    float bxo = mx * fr, byo = my * fr;
    float iu = rintf((ou - bxo) * fW);
    float iv = rintf((ovv - byo) * fH);
    int sx = (int)fminf(fmaxf(iu, 0.0f), maxx);
    int sy = (int)fminf(fmaxf(iv, 0.0f), maxy);
    float4 t = staged ? tile[(sy - by0) * bw + (sx - bx0)] : ov.prev[sy * W + sx];
    acc.x += t.x;
    acc.y += t.y;
    acc.z += t.z;
    acc.w += t.w;
  }
  const float fN = (float)N;
  ov.color[(size_t)y * (size_t)W + (size_t)x] = make_float4(acc.x / fN, acc.y / fN, acc.z / fN, acc.w / fN);
}

}  // namespace reproj
}  // namespace mtc

// mode 0: plain kernel; 1: tap-index map; 2: LDS-tiled kernel that skips
// the pixels the march overwrites
hipError_t launch_reproject(const FrameConsts& fc, const OutputViews& ov, int mode, hipStream_t stream) {
  dim3 grid((fc.width + 63) / 64, (fc.height + 3) / 4);
  dim3 block(256);
  if (mode == 1) hipLaunchKernelGGL((mtc::reproj::reproject_kernel<true>), grid, block, 0, stream, fc, ov);
  else if (mode == 2) hipLaunchKernelGGL((mtc::reproj::reproject_tiled_kernel<true>), grid, block, 0, stream, fc, ov);
  else hipLaunchKernelGGL((mtc::reproj::reproject_kernel<false>), grid, block, 0, stream, fc, ov);
  return hipGetLastError();
}

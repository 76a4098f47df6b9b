// post.hip -- the two post-process passes that consume the cloud stage's
// output: the god-ray radial blur (postProcess_GodRays.frag:66-150) and the
// Uncharted-2 tone map with Wang-hash dither (postProcess_ToneMap.frag:23-84).
//
// The reference runs both as full-screen fragment shaders; here each is a
// compute pass with one lane per pixel (64 x 4 pixel workgroups, rows
// coalesced).  The fragment's uv is its pixel centre, uv = (p + 0.5) / dim,
// and it reads and writes its own pixel.  (The reference stores to
// round(dim * uv), which at a pixel centre is a .5 tie: implementation-
// defined, and with ties-to-even two fragments share every even pixel while
// the odd ones are never written.)
//
// God rays.  100 bilinear samples of the decoded god-ray mask along the line
// from the pixel to the sun's screen position.  The mask is RGBA (the
// EncodeFloatRGBA of cloudRayMarch.comp:105-112); decoding is linear, so it
// is decoded once per frame into a footprint-packed map (each entry holds
// the 2x2 texels a bilinear tap reads, border texels included) and each of
// the 100 samples is one 16-byte load instead of four.  Sampler: linear,
// CLAMP_TO_BORDER with opaque black (Texture2D.cpp:75, Image.cpp:305-345).
//
// Tone map.  Bilinear read of the HDR result at the pixel centre (same
// sampler), Uncharted-2 curve with exposure 2.5 and white point 100, gamma
// 1/2.2, plus 0.01 x WangHashNoise(x, y, uint(total time)); stored as
// RGBA8 UNORM (the shader's rgba8 qualifier) or, for checking, as RGBA32F.
#include "device_common.h"

namespace mtc {
namespace post {

__device__ __forceinline__ float4 texel_or_border(const float4* __restrict__ img, int W, int H, int x, int y) {
  if (x < 0 || y < 0 || x >= W || y >= H) return make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  return img[(size_t)y * W + x];
}

// dot(x, 1 / (1, 255, 65025, 16581375)) (postProcess_GodRays.frag:36-43)
__device__ __forceinline__ float decode(float4 c) {
  const float d1 = 1.0f / 255.0f, d2 = 1.0f / 65025.0f, d3 = 1.0f / 16581375.0f;
  return ((c.x * 1.0f + c.y * d1) + c.z * d2) + c.w * d3;
}

// Entry (i, j), i in [-1, W-1], j in [-1, H-1], at ((j + 1) * (W + 1) + i + 1):
// decoded texels (i, j), (i+1, j), (i, j+1), (i+1, j+1).
__global__ __launch_bounds__(256) void decode_mask_kernel(const float4* __restrict__ mask, int W, int H,
                                                          float4* __restrict__ fp) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63) - 1;
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6) - 1;
  if (i >= W || j >= H) return;
  const float a = decode(texel_or_border(mask, W, H, i, j));
  const float b = decode(texel_or_border(mask, W, H, i + 1, j));
  const float c = decode(texel_or_border(mask, W, H, i, j + 1));
  const float d = decode(texel_or_border(mask, W, H, i + 1, j + 1));
  fp[(size_t)(j + 1) * (W + 1) + (i + 1)] = make_float4(a, b, c, d);
}

__global__ __launch_bounds__(256) void god_rays_kernel(PostConsts pc, const float4* __restrict__ fp,
                                                       float4* __restrict__ img) {
  const int W = pc.width, H = pc.height;
  const int px = blockIdx.x * 64 + (threadIdx.x & 63);
  const int py = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (px >= W || py >= H) return;
  float u = ((float)px + 0.5f) / (float)W, v = ((float)py + 0.5f) / (float)H;
  // delta_uv = ((uv - sunPos_ss) / numSamples) * ATMOSPHERE_DENSITY
  const float du = ((u - pc.sun_ss[0]) / 100.0f) * 1.0f, dv = ((v - pc.sun_ss[1]) / 100.0f) * 1.0f;
  const float Wf = (float)W, Hf = (float)H;
  float ar = 0.0f, ag = 0.0f, ab = 0.0f;
  for (int i = 0; i < 100; ++i) {
    const float x = u * Wf - 0.5f, y = v * Hf - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float wx = x - fx, wy = y - fy;
    // the walk stays inside the segment pixel -> sun (both in [0, 1]^2), so
    // floor lies in [-1, dim - 1]
    const int ix = min(max((int)fx, -1), W - 1), iy = min(max((int)fy, -1), H - 1);
    const float4 e = fp[(size_t)(iy + 1) * (W + 1) + (ix + 1)];
    const float a = lerpf(lerpf(e.x, e.y, wx), lerpf(e.z, e.w, wx), wy);
    // godRayLoop: (rgb * a) * (decay * WEIGHT_FOR_SAMPLE), decay = 1
    ar += (pc.light[0] * a) * 0.001f;
    ag += (pc.light[1] * a) * 0.001f;
    ab += (pc.light[2] * a) * 0.001f;
    u -= du;
    v -= dv;
  }
  const size_t o = (size_t)py * W + px;
  float4 c = img[o];
  // original + vec4(accum * EXPOSURE, 1) * blendFactor
  c.x += (ar * 1.0f) * pc.blend;
  c.y += (ag * 1.0f) * pc.blend;
  c.z += (ab * 1.0f) * pc.blend;
  c.w += 1.0f * pc.blend;
  img[o] = c;
}

// ((x (A x + C B) + D E) / (x (A x + B) + D F)) - E / F
__device__ __forceinline__ float uncharted2(float x) {
  const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
  return ((x * (A * x + C * B) + D * E) / (x * (A * x + B) + D * F)) - E / F;
}

__device__ __forceinline__ uint32_t wang_hash(uint32_t u, uint32_t v, uint32_t s) {
  uint32_t seed = (u * 1664525u + v) + s;
  seed = (seed ^ 61u) ^ (seed >> 16u);
  seed *= 9u;
  seed = seed ^ (seed >> 4u);
  seed *= 0x27d4eb2du;
  seed = seed ^ (seed >> 15u);
  return seed;
}

__device__ __forceinline__ uint32_t to_unorm8(float v) {
  const float c = fminf(fmaxf(v, 0.0f), 1.0f);
  return (uint32_t)__builtin_rintf(c * 255.0f);
}

template <bool F32OUT>
__global__ __launch_bounds__(256) void tone_map_kernel(const float4* __restrict__ img, int W, int H, uint32_t seed_s,
                                                       float whitemap, void* __restrict__ out) {
  const int px = blockIdx.x * 64 + (threadIdx.x & 63);
  const int py = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (px >= W || py >= H) return;
  const float u = ((float)px + 0.5f) / (float)W, v = ((float)py + 0.5f) / (float)H;
  const float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
  const float fx = floorf(x), fy = floorf(y);
  const float wx = x - fx, wy = y - fy;
  const int ix = (int)fx, iy = (int)fy;
  const float4 t00 = texel_or_border(img, W, H, ix, iy), t10 = texel_or_border(img, W, H, ix + 1, iy);
  const float4 t01 = texel_or_border(img, W, H, ix, iy + 1), t11 = texel_or_border(img, W, H, ix + 1, iy + 1);
  float rgb[3] = {lerpf(lerpf(t00.x, t10.x, wx), lerpf(t01.x, t11.x, wx), wy),
                  lerpf(lerpf(t00.y, t10.y, wx), lerpf(t01.y, t11.y, wx), wy),
                  lerpf(lerpf(t00.z, t10.z, wx), lerpf(t01.z, t11.z, wx), wy)};
  // noise = WangHashNoise(x, y, uint(time.y)) * 0.01
  const float noise = ((float)wang_hash((uint32_t)px, (uint32_t)py, seed_s) * (1.0f / 4294967296.0f)) * 0.01f;
  float o[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float c = uncharted2(2.5f * rgb[k]);
    c = c * whitemap;
    o[k] = powf(c, 1.0f / 2.2f) + noise;
  }
  const size_t i = (size_t)py * W + px;
  if (F32OUT) {
    ((float4*)out)[i] = make_float4(o[0], o[1], o[2], 1.0f);
  } else {
    ((uint32_t*)out)[i] = to_unorm8(o[0]) | (to_unorm8(o[1]) << 8) | (to_unorm8(o[2]) << 16) | (255u << 24);
  }
}

}  // namespace post
}  // namespace mtc

hipError_t launch_god_rays(const PostConsts& pc, const float4* mask, float4* footprint, float4* img, hipStream_t s) {
  const int W = pc.width, H = pc.height;
  dim3 gd((W + 1 + 63) / 64, (H + 1 + 3) / 4), gp((W + 63) / 64, (H + 3) / 4), block(256);
  hipLaunchKernelGGL(mtc::post::decode_mask_kernel, gd, block, 0, s, mask, W, H, footprint);
  hipLaunchKernelGGL(mtc::post::god_rays_kernel, gp, block, 0, s, pc, (const float4*)footprint, img);
  return hipGetLastError();
}

hipError_t launch_tone_map(const float4* img, int W, int H, uint32_t seed_s, float whitemap, int f32, void* out,
                           hipStream_t s) {
  dim3 grid((W + 63) / 64, (H + 3) / 4), block(256);
  if (f32) hipLaunchKernelGGL(mtc::post::tone_map_kernel<true>, grid, block, 0, s, img, W, H, seed_s, whitemap, out);
  else hipLaunchKernelGGL(mtc::post::tone_map_kernel<false>, grid, block, 0, s, img, W, H, seed_s, whitemap, out);
  return hipGetLastError();
}

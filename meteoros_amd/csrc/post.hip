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
// is decoded once per frame into a float map (border texels included), and
// each workgroup stages the part of that map its taps read at a sample into
// LDS, so the 100 taps per pixel read LDS rather than the texture path.
// Sampler: linear, CLAMP_TO_BORDER with opaque black (Texture2D.cpp:75,
// Image.cpp:305-345).
//
// TXAA.  Re-projects each pixel into the previous TXAA output through the
// inner atmosphere shell of the TXAA shader (R + 7.5 km), clips the
// history to the current frame's 3x3 / cross neighbourhood box and blends
// with a luminance-weighted feedback in [0, 0.5].  The reference reads and
// writes the current image in place (neighbours race); here the current
// frame is read-only and the result goes to the other history buffer.
//
// Tone map.  Bilinear read of the HDR result at the pixel centre (same
// sampler), Uncharted-2 curve with exposure 2.5 and white point 100, gamma
// 1/2.2, plus 0.01 x WangHashNoise(x, y, uint(total time)); stored as
// RGBA8 UNORM (the shader's rgba8 qualifier) or, for checking, as RGBA32F.
#include <mutex>

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

// The god-ray set-up launch: the blur's table, then the decoded mask.
//
// Table.  All samples of one pixel row share their vertical position: v
// starts at (py + 0.5) / H and steps by the row's own delta_uv.y, in the
// shader's FP32 order, so sample i of every pixel in the row reads the same
// texel rows with the same vertical weights.  One wave per row group of
// kGodRayGroupRows rows: lane = 16 * chunk + k walks row k's recurrence and
// writes samples 25 * chunk .. 25 * chunk + 24 (re-running the recurrence's
// subtractions up to its chunk, the same FP32 operations), with the group's
// lowest texel row (a 16-lane minimum) in the sample's header entry.
//
// Map.  One lane per map texel: dot(rgba, bitDec) of the mask texel, or of
// the sampler's opaque black border outside the image.  Decoding is linear,
// so decoding before filtering equals the shader's filtering before
// decoding up to FP32 rounding.  Each map row takes whole workgroups of 256
// texels, so the row, and the mask row it reads, are uniform: scalar row
// pointers and 32-bit column offsets, no per-lane division.
__global__ __launch_bounds__(256) void god_rays_setup_kernel(PostConsts pc, const float4* __restrict__ mask,
                                                             int table_blocks, float* __restrict__ map,
                                                             GodRayEntry* __restrict__ tab) {
  static_assert(kGodRayGroupRows * 4 == 64 && kGodRaySamples == 4 * 25,
                "one wave per row group: 16 rows x 4 chunks of 25 samples");
  const int W = pc.width, H = pc.height;
  if ((int)blockIdx.x < table_blocks) {
    const int g = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (g >= (H + kGodRayGroupRows - 1) / kGodRayGroupRows) return;
    const int k = lane & 15, chunk = lane >> 4;
    const int py = min(g * kGodRayGroupRows + k, H - 1);
    const float Hf = (float)H;
    float v = ((float)py + 0.5f) / Hf;
    // delta_uv.y = ((uv.y - sunPos_ss.y) / numSamples) * ATMOSPHERE_DENSITY
    const float dv = ((v - pc.sun_ss[1]) / (float)kGodRaySamples) * 1.0f;
    for (int s = 0; s < 25 * chunk; ++s) v -= dv;
    GodRayEntry* out = tab + (size_t)g * kGodRaySamples * kGodRayTableStride;
    for (int j = 0; j < 25; ++j) {
      const int i = 25 * chunk + j;
      const float y = v * Hf - 0.5f;
      const float fy = floorf(y);
      const float wy = y - fy;
      // the walk stays inside the segment pixel -> sun, so floor lies in [-1, H - 1]
      const int iy = min(max((int)fy, -1), H - 1);
      int lo = iy;
      for (int m = 1; m < 16; m <<= 1) lo = min(lo, __shfl_xor(lo, m, 16));
      out[(size_t)i * kGodRayTableStride + 1 + k] =
          GodRayEntry{(uint32_t)((iy - lo) * kGodRayStageCols * (int)sizeof(float)), 1.0f - wy, wy, 0u};
      if (k == 0) out[(size_t)i * kGodRayTableStride] = GodRayEntry{(uint32_t)(lo + 1), 0.0f, 0.0f, 0u};
      v -= dv;
    }
    return;
  }
  const int P = god_rays_map_pitch(W);
  const int bpr = (P + 255) / 256;  // workgroups per map row
  const int b = (int)blockIdx.x - table_blocks;
  const int row = b / bpr;
  const int col = (b - row * bpr) * 256 + (int)threadIdx.x;
  if (row >= god_rays_map_rows(H) || col >= P) return;
  const int y = row - 1, x = col - 4;
  float4 c = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  if (y >= 0 && y < H && x >= 0 && x < W)
    c = *(const float4*)((const char*)(mask + (size_t)y * W) + (uint32_t)x * 16u);
  *(float*)((char*)(map + (size_t)row * P) + (uint32_t)col * 4u) = decode(c);
}

// The radial blur.  Workgroup = 64 columns x kGodRayGroupRows rows, four
// waves of four rows per lane.  Pixels of one column start at the same u
// and step by the same delta_uv.x, so a lane computes each sample's
// horizontal position once for its four pixels, and at sample i the taps
// of the whole workgroup lie in one region of the decoded map: at most
// kGodRayStageRows rows (from the table's header) x kGodRayStageCols
// columns (from the first lane's column less one, aligned down to 4), and fewer
// further along the walk, where the samples of adjacent pixels close up.
// The workgroup stages the region of the sample's phase (kGodRayPhaseRows x
// kGodRayPhaseBlocks 16-byte loads: 324, 196, 100, 42; lane t loads slot t
// and, in the first phase, t + 256) into LDS, and every tap reads its four
// texels from LDS.  Staging runs two samples ahead in registers and one
// ahead in the second LDS buffer, the table entries one sample ahead in
// scalar registers; one barrier per sample.  Per tap: two ds_read2_b32 and
// six FMA-class operations, the same arithmetic as a per-pixel bilinear tap.
template <int ROWS, int BLOCKS>
struct GodRayShape {
  static constexpr int slots = ROWS * BLOCKS;
  static constexpr int blocks = BLOCKS;
  static_assert(ROWS <= kGodRayStageRows && 4 * BLOCKS <= kGodRayStageCols && slots <= 512,
                "a phase's region fits the LDS buffer and two loads per lane");
};
template <int P>
using GodRayPhase = GodRayShape<kGodRayPhaseRows[P], kGodRayPhaseBlocks[P]>;

__global__ __launch_bounds__(256) void god_rays_kernel(PostConsts pc, const float* __restrict__ map,
                                                       const GodRayEntry* __restrict__ tab, float4* __restrict__ img) {
  constexpr int REGION = kGodRayStageRows * kGodRayStageCols;
  static_assert(kGodRaySamples == 4 * kGodRayPhaseSamples && kGodRayPhaseSamples >= 3, "four phases");
  __shared__ __attribute__((aligned(16))) float stage[2 * REGION];
  typedef float f4v __attribute__((ext_vector_type(4)));
  typedef const __attribute__((address_space(1))) char gchar;
  const int W = pc.width, H = pc.height;
  const int P = god_rays_map_pitch(W);
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int px = blockIdx.x * 64 + lane;
  const int g = blockIdx.y;
  // lanes right of the image walk the last column's line and store nothing
  float u = ((float)min(px, W - 1) + 0.5f) / (float)W;
  // delta_uv.x = ((uv.x - sunPos_ss.x) / numSamples) * ATMOSPHERE_DENSITY
  const float du = ((u - pc.sun_ss[0]) / (float)kGodRaySamples) * 1.0f;
  const float Wf = (float)W;
  const GodRayEntry* __restrict__ gt = tab + (size_t)g * kGodRaySamples * kGodRayTableStride;
  // this lane's staging slots (row, 4-column block) of a phase's region, in
  // the map (goff*, for the samples being loaded) and in LDS (loff*, for the
  // samples being stored), set when the phase changes
  uint32_t goff0 = 0, goff1 = 0, loff0 = 0, loff1 = 0;
  auto map_slots = [&](auto S) {
    using Sh = decltype(S);
    goff0 = (uint32_t)((tid / Sh::blocks) * P + 4 * (tid % Sh::blocks)) * 4u;
    if (Sh::slots > 256) goff1 = (uint32_t)(((tid + 256) / Sh::blocks) * P + 4 * ((tid + 256) % Sh::blocks)) * 4u;
  };
  auto lds_slots = [&](auto S) {
    using Sh = decltype(S);
    loff0 = (uint32_t)((tid / Sh::blocks) * kGodRayStageCols + 4 * (tid % Sh::blocks)) * 4u;
    if (Sh::slots > 256)
      loff1 = (uint32_t)(((tid + 256) / Sh::blocks) * kGodRayStageCols + 4 * ((tid + 256) % Sh::blocks)) * 4u;
  };
  // sample position: x = u * W - 0.5 (sampler), weight and column; the
  // region's first stored column: one before the first lane's, aligned down
  // to 4.  The lanes' columns grow with the lane in exact arithmetic; in the
  // FP32 walk a lane can land one column below the first lane's where the
  // samples lie a hundredth of a texel apart (seen at widths of 8192 and
  // more), never further, so the region starts one column early.
  auto advance = [&](float& wx, int& ix, int& cs) {
    const float x = u * Wf - 0.5f;
    const float fx = floorf(x);
    wx = x - fx;
    // the walk stays between pixel and sun: floor lies in [-1, W - 1]
    ix = min(max((int)fx, -1), W - 1);
    cs = (__builtin_amdgcn_readfirstlane(ix) + 3) & ~3;
    u -= du;
  };
  f4v d0, d1;
  // the region's 16-byte blocks of the sample whose first map row is `row`
  auto load_region = [&](auto S, uint32_t row, int cs) {
    using Sh = decltype(S);
    const uint64_t rb = (uint64_t)map + ((uint64_t)row * (uint64_t)P + (uint64_t)cs) * sizeof(float);
    // the region's address is uniform: kept whole in scalar registers (each
    // half through uint32_t, since readfirstlane returns int)
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(rb >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)rb);
    gchar* b = (gchar*)(((uint64_t)hi << 32) | (uint64_t)lo);
    if (Sh::slots >= 256 || tid < Sh::slots) d0 = *(const __attribute__((address_space(1))) f4v*)(b + goff0);
    if (Sh::slots > 256 && tid + 256 < Sh::slots) d1 = *(const __attribute__((address_space(1))) f4v*)(b + goff1);
  };
  auto store_region = [&](auto S, char* wb) {
    using Sh = decltype(S);
    if (Sh::slots >= 256 || tid < Sh::slots) *(f4v*)(wb + loff0) = d0;
    if (Sh::slots > 256 && tid + 256 < Sh::slots) *(f4v*)(wb + loff1) = d1;
  };
  float wx_c, wx_n;
  int ix_c, ix_n, cs_c, cs_n;
  map_slots(GodRayPhase<0>{});
  lds_slots(GodRayPhase<0>{});
  advance(wx_c, ix_c, cs_c);
  load_region(GodRayPhase<0>{}, gt[0].a, cs_c);
  store_region(GodRayPhase<0>{}, (char*)stage);
  advance(wx_n, ix_n, cs_n);
  load_region(GodRayPhase<0>{}, gt[kGodRayTableStride].a, cs_n);
  GodRayEntry nx[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) nx[k] = gt[1 + 4 * w + k];
  uint32_t row2 = gt[2 * kGodRayTableStride].a;
  __syncthreads();
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  // sample i: its taps from stage[i & 1], sample i + 1's region (phase of
  // SW) into the other buffer, sample i + 2's region (phase of SL) loaded
  auto sample = [&](auto SW, auto SL, int i) {
    GodRayEntry cur[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) cur[k] = nx[k];
    const uint32_t row = row2;
    if (i + 1 < kGodRaySamples) {
#pragma unroll
      for (int k = 0; k < 4; ++k) nx[k] = gt[(size_t)(i + 1) * kGodRayTableStride + 1 + 4 * w + k];
    }
    if (i + 3 < kGodRaySamples) row2 = gt[(size_t)(i + 3) * kGodRayTableStride].a;
    const char* sb = (const char*)(stage + (i & 1) * REGION) + (uint32_t)(ix_c + 4 - cs_c) * 4u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float* q = (const float*)(sb + cur[k].a);
      const float a0 = q[0], a1 = q[1], b0 = q[kGodRayStageCols], b1 = q[kGodRayStageCols + 1];
      const float t = __fmaf_rn(wx_c, a1 - a0, a0), b = __fmaf_rn(wx_c, b1 - b0, b0);
      // godRayLoop with decay 1: the mask value summed here, the light
      // colour and WEIGHT_FOR_SAMPLE applied once below
      acc[k] = __fmaf_rn(cur[k].w1, b, __fmaf_rn(cur[k].w0, t, acc[k]));
    }
    if (i + 1 < kGodRaySamples) store_region(SW, (char*)(stage + ((i + 1) & 1) * REGION));
    wx_c = wx_n;
    ix_c = ix_n;
    cs_c = cs_n;
    if (i + 2 < kGodRaySamples) {
      advance(wx_n, ix_n, cs_n);
      load_region(SL, row, cs_n);
    }
    __syncthreads();
  };
  // Phase p covers samples [25 p, 25 p + 25).  Its loads start two samples
  // early (iteration 25 p - 2) and its stores one sample early, so each
  // phase change is one iteration of its own.  Every run of iterations,
  // including those single ones, is a loop unrolled by two: 110 us at 1080p
  // against 121.5 us not unrolled, in 62 VGPRs (eight waves per SIMD; with
  // the single iterations written as plain calls the kernel takes 68).
  constexpr int L = kGodRayPhaseSamples;
  int i = 0;
#pragma unroll 2
  for (; i < L - 2; ++i) sample(GodRayPhase<0>{}, GodRayPhase<0>{}, i);
  map_slots(GodRayPhase<1>{});
#pragma unroll 2
  for (; i < L - 1; ++i) sample(GodRayPhase<0>{}, GodRayPhase<1>{}, i);
  lds_slots(GodRayPhase<1>{});
#pragma unroll 2
  for (; i < 2 * L - 2; ++i) sample(GodRayPhase<1>{}, GodRayPhase<1>{}, i);
  map_slots(GodRayPhase<2>{});
#pragma unroll 2
  for (; i < 2 * L - 1; ++i) sample(GodRayPhase<1>{}, GodRayPhase<2>{}, i);
  lds_slots(GodRayPhase<2>{});
#pragma unroll 2
  for (; i < 3 * L - 2; ++i) sample(GodRayPhase<2>{}, GodRayPhase<2>{}, i);
  map_slots(GodRayPhase<3>{});
#pragma unroll 2
  for (; i < 3 * L - 1; ++i) sample(GodRayPhase<2>{}, GodRayPhase<3>{}, i);
  lds_slots(GodRayPhase<3>{});
#pragma unroll 2
  for (; i < kGodRaySamples; ++i) sample(GodRayPhase<3>{}, GodRayPhase<3>{}, i);
  if (px >= W) return;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int py = g * kGodRayGroupRows + 4 * w + k;
    if (py >= H) break;
    const size_t o = (size_t)py * W + px;
    float4 c = img[o];
    // original + vec4(accum * EXPOSURE, 1) * blendFactor, accum = light * sum * 0.001
    c.x += (((pc.light[0] * acc[k]) * 0.001f) * 1.0f) * pc.blend;
    c.y += (((pc.light[1] * acc[k]) * 0.001f) * 1.0f) * pc.blend;
    c.z += (((pc.light[2] * acc[k]) * 0.001f) * 1.0f) * pc.blend;
    c.w += 1.0f * pc.blend;
    img[o] = c;
  }
}

// a / b through the hardware reciprocal (about 1 ulp) and one FMA
// correction: q = a y, r = a - q b exactly, q + r y.  Four VALU operations
// instead of the IEEE division's dozen; the result is the correctly rounded
// quotient except in rare cases, where it is one ulp off.  For finite,
// normal, non-zero b (every denominator of the tone map and TXAA passes:
// image sizes, the Uncharted-2 curve's x (A x + B) + D F >= 0.06, lengths,
// 2A of the shell quadratic, the projected depth, the clip box's half
// extents (>= 1e-10) and the luminance floor (>= 0.2)).
struct Recip {
  float b, y;
};
__device__ __forceinline__ Recip recip(float b) { return Recip{b, __builtin_amdgcn_rcpf(b)}; }
__device__ __forceinline__ float divr(float a, const Recip& d) {
  const float q = a * d.y;
  const float r = __builtin_fmaf(-q, d.b, a);
  return __builtin_fmaf(r, d.y, q);
}
__device__ __forceinline__ float divr(float a, float b) { return divr(a, recip(b)); }
__device__ __forceinline__ f3 divr(f3 a, const Recip& d) { return mk(divr(a.x, d), divr(a.y, d), divr(a.z, d)); }
__device__ __forceinline__ f3 normalize_r(f3 a) { return divr(a, recip(length(a))); }
__device__ __forceinline__ float4 f4divr(float4 a, const Recip& d) {
  return make_float4(divr(a.x, d), divr(a.y, d), divr(a.z, d), divr(a.w, d));
}

// ((x (A x + C B) + D E) / (x (A x + B) + D F)) - E / F
__device__ __forceinline__ float uncharted2(float x) {
  const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
  return divr(x * (A * x + C * B) + D * E, x * (A * x + B) + D * F) - E / F;
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
  // one pixel row per wave: the row index is uniform (scalar arithmetic)
  const int py = blockIdx.y * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (px >= W || py >= H) return;
  const float u = divr((float)px + 0.5f, (float)W), v = divr((float)py + 0.5f, (float)H);
  const float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
  const float fx = floorf(x), fy = floorf(y);
  const float wx = x - fx, wy = y - fy;
  const int ix = (int)fx, iy = (int)fy;
  // At a pixel centre x and y lie within a rounding error of px and py, so
  // the taps are columns px - 1 .. px + 1 and rows py - 1 .. py + 1.  A wave
  // whose row and columns and their neighbours lie inside the image reads
  // them without bounds tests: its texel rows are uniform (v depends on py
  // only), so each tap is a row pointer in scalar registers plus a 32-bit
  // column offset.
  float4 t00, t10, t01, t11;
  const int x0 = blockIdx.x * 64;
  if (x0 >= 1 && x0 + 64 < W && py >= 1 && py + 1 < H) {
    const float4* r0 = img + (size_t)__builtin_amdgcn_readfirstlane(iy) * W;
    const char* c0 = (const char*)r0 + (uint32_t)ix * 16u;
    const char* c1 = (const char*)(r0 + W) + (uint32_t)ix * 16u;
    t00 = *(const float4*)c0;
    t10 = *(const float4*)(c0 + 16);
    t01 = *(const float4*)c1;
    t11 = *(const float4*)(c1 + 16);
  } else {
    t00 = texel_or_border(img, W, H, ix, iy);
    t10 = texel_or_border(img, W, H, ix + 1, iy);
    t01 = texel_or_border(img, W, H, ix, iy + 1);
    t11 = texel_or_border(img, W, H, ix + 1, iy + 1);
  }
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
    // GLSL defines pow(x, y) as exp2(y * log2(x)) (its precision is
    // inherited from those two): v_log_f32 / v_exp_f32, not the library
    // powf, which is ~100 instructions and made this pass VALU-bound
    o[k] = __builtin_amdgcn_exp2f((1.0f / 2.2f) * __builtin_amdgcn_logf(c)) + noise;
  }
  const size_t i = (size_t)py * W + px;
  if (F32OUT) {
    ((float4*)out)[i] = make_float4(o[0], o[1], o[2], 1.0f);
  } else {
    // W x H <= 16384^2 RGBA8 texels: the byte offset fits 32 bits
    *(uint32_t*)((char*)out + (uint32_t)i * 4u) =
        to_unorm8(o[0]) | (to_unorm8(o[1]) << 8) | (to_unorm8(o[2]) << 16) | (255u << 24);
  }
}

// ---------------------------------------------------------------------------
// TXAA (postProcess_TXAA.frag:171-270)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float4 unorm4(uint32_t t) {
  const float k = MTC_INV255;
  return make_float4((float)(t & 0xffu) * k, (float)((t >> 8) & 0xffu) * k, (float)((t >> 16) & 0xffu) * k,
                     (float)(t >> 24) * k);
}
__device__ __forceinline__ float4 f4min(float4 a, float4 b) {
  return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), fminf(a.w, b.w));
}
__device__ __forceinline__ float4 f4max(float4 a, float4 b) {
  return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w));
}
__device__ __forceinline__ float4 f4add(float4 a, float4 b) {
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ float4 f4scale(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }

// Texel (x, y) of a W x H RGBA8 image.  Images hold at most 16384^2 texels,
// so the byte offset fits 32 bits: one 32-bit multiply-add per texel and a
// load addressed from the image's base in scalar registers, instead of
// 64-bit address arithmetic per tap.
__device__ __forceinline__ uint32_t texel8(const uint32_t* __restrict__ img, int W, int x, int y) {
  const uint32_t off = ((uint32_t)y * (uint32_t)W + (uint32_t)x) * 4u;
  return *(const uint32_t*)((const char*)img + off);
}

// imageLoad of the current RGBA8 frame; outside the image it returns 0
__device__ __forceinline__ float4 load_cur(const uint32_t* __restrict__ img, int W, int H, int x, int y) {
  if (x < 0 || y < 0 || x >= W || y >= H) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  return unorm4(texel8(img, W, x, y));
}

// texture(prevFrameImage, uv): RGBA8 codes filtered, then scaled by 1/255;
// CLAMP_TO_BORDER with opaque black
__device__ __forceinline__ float4 sample_prev(const uint32_t* __restrict__ img, int W, int H, float u, float v) {
  const float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
  const float fx = floorf(x), fy = floorf(y);
  const float wx = x - fx, wy = y - fy;
  const int ix = (int)fminf(fmaxf(fx, -2.0f), (float)W), iy = (int)fminf(fmaxf(fy, -2.0f), (float)H);
  c4 t[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int xx = ix + (k & 1), yy = iy + (k >> 1);
    t[k] = (xx < 0 || yy < 0 || xx >= W || yy >= H) ? c4{0.0f, 0.0f, 0.0f, 255.0f} : unpack8(texel8(img, W, xx, yy));
  }
  c4 r = scale4(lerp4(lerp4(t[0], t[1], wx), lerp4(t[2], t[3], wx), wy), MTC_INV255);
  return make_float4(r.r, r.g, r.b, r.a);
}

__global__ __launch_bounds__(256) void txaa_kernel(TxaaConsts tc, const uint32_t* __restrict__ cur,
                                                   const uint32_t* __restrict__ prev, uint32_t* __restrict__ out) {
  const int W = tc.width, H = tc.height;
  const int px = blockIdx.x * 64 + (threadIdx.x & 63);
  // one pixel row per wave: the row index is uniform (scalar arithmetic)
  const int py = blockIdx.y * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (px >= W || py >= H) return;
  const float u = divr((float)px + 0.5f, (float)W), v = divr((float)py + 0.5f, (float)H);
  // castRay (:87-116) with the basis folded per frame
  float nx = u * 2.0f - 1.0f, ny = v * 2.0f - 1.0f;
  nx = nx + tc.jit[0];
  ny = ny + tc.jit[1];
  const f3 eye = ld3(tc.eye);
  const f3 cx = (nx * tc.tfx) * ld3(tc.right), cy = (ny * tc.tfy) * ld3(tc.up);
  const f3 pnt = (ld3(tc.ref) + cx) + cy;
  const f3 rD = normalize_r(pnt - eye);
  // raySphereIntersection (:119-165) with the inner atmosphere of the TXAA shader
  const f3 rO = ld3(tc.ro);
  f3 P = mk(0.0f, 0.0f, 0.0f);
  {
    const float A = dot(rD, rD), B = 2.0f * dot(rD, rO), C = tc.C;
    const float disc = B * B - 4.0f * A * C;
    if (!(disc < 0.0f)) {
      const Recip a2 = recip(2.0f * A);
      float t = divr(-B - __builtin_sqrtf(disc), a2);
      if (t < 0.0f) t = divr(-B + __builtin_sqrtf(disc), a2);
      if (t >= 0.0f) P = ((rO + t * rD) * tc.R) + ld3(tc.center);
    }
  }
  const float* m = tc.old_view;
  f3 q = mk(((m[0] * P.x + m[4] * P.y) + m[8] * P.z) + m[12] * 1.0f,
            ((m[1] * P.x + m[5] * P.y) + m[9] * P.z) + m[13] * 1.0f,
            ((m[2] * P.x + m[6] * P.y) + m[10] * P.z) + m[14] * 1.0f);
  f3 od = normalize_r(q);
  const float nz = -od.z;
  od = divr(od, recip(nz));
  const float ou = divr(od.x, tc.tfx) * 0.5f + 0.5f, ov = divr(od.y, tc.tfy) * 0.5f + 0.5f;
  // neighbourHoodClamping (:191-220): 3x3 and cross-5 min / max / mean.  A
  // wave is one pixel row of 64 columns; when its row and columns and their
  // neighbours all lie inside the image (every wave but those on the
  // border), the nine texels are loaded without the per-tap bounds tests.
  float4 ctl, ctc, ctr, cml, cmc, cmr, cbl, cbc, cbr;
  const int x0 = blockIdx.x * 64;
  if (x0 >= 1 && x0 + 64 < W && py >= 1 && py + 1 < H) {
    // byte offsets of the centre column in the three rows; the left and
    // right texels are the loads' immediate offsets -4 and +4
    const uint32_t rs = (uint32_t)W * 4u;
    const uint32_t om = ((uint32_t)py * (uint32_t)W + (uint32_t)px) * 4u, ot = om - rs, ob = om + rs;
    const char* c8 = (const char*)cur;
    ctl = unorm4(*(const uint32_t*)(c8 + ot - 4));
    ctc = unorm4(*(const uint32_t*)(c8 + ot));
    ctr = unorm4(*(const uint32_t*)(c8 + ot + 4));
    cml = unorm4(*(const uint32_t*)(c8 + om - 4));
    cmc = unorm4(*(const uint32_t*)(c8 + om));
    cmr = unorm4(*(const uint32_t*)(c8 + om + 4));
    cbl = unorm4(*(const uint32_t*)(c8 + ob - 4));
    cbc = unorm4(*(const uint32_t*)(c8 + ob));
    cbr = unorm4(*(const uint32_t*)(c8 + ob + 4));
  } else {
    ctl = load_cur(cur, W, H, px - 1, py - 1);
    ctc = load_cur(cur, W, H, px, py - 1);
    ctr = load_cur(cur, W, H, px + 1, py - 1);
    cml = load_cur(cur, W, H, px - 1, py);
    cmc = load_cur(cur, W, H, px, py);
    cmr = load_cur(cur, W, H, px + 1, py);
    cbl = load_cur(cur, W, H, px - 1, py + 1);
    cbc = load_cur(cur, W, H, px, py + 1);
    cbr = load_cur(cur, W, H, px + 1, py + 1);
  }
  float4 cmin = f4min(ctl, f4min(ctc, f4min(ctr, f4min(cml, f4min(cmc, f4min(cmr, f4min(cbl, f4min(cbc, cbr))))))));
  float4 cmax = f4max(ctl, f4max(ctc, f4max(ctr, f4max(cml, f4max(cmc, f4max(cmr, f4max(cbl, f4max(cbc, cbr))))))));
  float4 cavg = f4divr(f4add(f4add(f4add(f4add(f4add(f4add(f4add(f4add(ctl, ctc), ctr), cml), cmc), cmr), cbl), cbc),
                             cbr),
                       recip(9.0f));
  const float4 cmin5 = f4min(ctc, f4min(cml, f4min(cmc, f4min(cmr, cbc))));
  const float4 cmax5 = f4max(ctc, f4max(cml, f4max(cmc, f4max(cmr, cbc))));
  const float4 cavg5 = f4divr(f4add(f4add(f4add(f4add(ctc, cml), cmc), cmr), cbc), recip(5.0f));
  cmin = f4scale(f4add(cmin, cmin5), 0.5f);
  cmax = f4scale(f4add(cmax, cmax5), 0.5f);
  cavg = f4scale(f4add(cavg, cavg5), 0.5f);
  const float4 curr = cmc;
  float4 prv = sample_prev(prev, W, H, ou, ov);
  // clip_aabb(cmin.xyz, cmax.xyz, clamp(cavg, cmin, cmax), prev) (:170-189)
  {
    const float pw = fminf(fmaxf(cavg.w, cmin.w), cmax.w);
    const float pcx = 0.5f * (cmax.x + cmin.x), pcy = 0.5f * (cmax.y + cmin.y), pcz = 0.5f * (cmax.z + cmin.z);
    const float ex = 0.5f * (cmax.x - cmin.x) + 1e-10f, ey = 0.5f * (cmax.y - cmin.y) + 1e-10f,
                ez = 0.5f * (cmax.z - cmin.z) + 1e-10f;
    const float vx = prv.x - pcx, vy = prv.y - pcy, vz = prv.z - pcz, vw = prv.w - pw;
    const float ma = fmaxf(fabsf(divr(vx, ex)), fmaxf(fabsf(divr(vy, ey)), fabsf(divr(vz, ez))));
    if (ma > 1.0f) {
      const Recip rm = recip(ma);
      prv = make_float4(pcx + divr(vx, rm), pcy + divr(vy, rm), pcz + divr(vz, rm), pw + divr(vw, rm));
    }
  }
  // luminance-weighted feedback (:250-258)
  const float lum0 = (curr.x * 0.2125f + curr.y * 0.7154f) + curr.z * 0.0721f;
  const float lum1 = (prv.x * 0.2125f + prv.y * 0.7154f) + prv.z * 0.0721f;
  const float diff = divr(fabsf(lum0 - lum1), fmaxf(lum0, fmaxf(lum1, 0.2f)));
  const float w = 1.0f - diff;
  const float k = mixf(0.0f, 0.5f, w * w);
  const float4 o = make_float4(mixf(prv.x, curr.x, k), mixf(prv.y, curr.y, k), mixf(prv.z, curr.z, k),
                               mixf(prv.w, curr.w, k));
  out[(size_t)py * W + px] = to_unorm8(o.x) | (to_unorm8(o.y) << 8) | (to_unorm8(o.z) << 16) | (to_unorm8(o.w) << 24);
}

}  // namespace post
}  // namespace mtc

// ---------------------------------------------------------------------------
// Allocation-free god rays (mtc_god_rays_noalloc): the pass's scratch (the
// decoded mask and the blur's table) lives in this
// file-scope device array instead of a buffer the context allocates.  Each
// device holds its own copy, at its own address, from the first load of
// this library until the process ends.  In host code `scratch` names a
// host-side stand-in, never device memory: every launch takes the address
// hipGetSymbolAddress returns for the current device (god_rays_scratch
// below).  Every call on one device shares the array; the set-up launch
// rewrites every byte the blur reads, so nothing has to start at zero, and
// calls that overlap on one device race.
// ---------------------------------------------------------------------------
namespace mtc {
namespace post {
static __device__ __align__(256) float4 scratch[MTC_GOD_RAYS_NOALLOC_SCRATCH_BYTES / sizeof(float4)];
}  // namespace post
}  // namespace mtc

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

// Device address of `scratch` on `device`, which must be the current
// device; looked up once per device and kept.
hipError_t god_rays_scratch(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(mtc::post::scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

// `scratch` holds god_rays_scratch_bytes(W, H) bytes: the decoded map,
// then the blur's table.
hipError_t launch_god_rays(const PostConsts& pc, const float4* mask, void* scratch, float4* img, hipStream_t s) {
  const int W = pc.width, H = pc.height;
  float* map = (float*)scratch;
  GodRayEntry* tab = (GodRayEntry*)((char*)scratch + god_rays_map_span(W, H));
  const int groups = (H + kGodRayGroupRows - 1) / kGodRayGroupRows;
  const int table_blocks = (groups + 3) / 4;
  const long long map_blocks = (long long)god_rays_map_rows(H) * ((god_rays_map_pitch(W) + 255) / 256);
  hipLaunchKernelGGL(mtc::post::god_rays_setup_kernel, dim3((unsigned)(table_blocks + map_blocks)), dim3(256), 0, s, pc,
                     mask, table_blocks, map, tab);
  hipLaunchKernelGGL(mtc::post::god_rays_kernel, dim3((W + 63) / 64, groups), dim3(256), 0, s, pc, (const float*)map,
                     (const GodRayEntry*)tab, img);
  return hipGetLastError();
}

hipError_t launch_tone_map(const float4* img, int W, int H, uint32_t seed_s, float whitemap, int f32, void* out,
                           hipStream_t s) {
  dim3 grid((W + 63) / 64, (H + 3) / 4), block(256);
  if (f32) hipLaunchKernelGGL(mtc::post::tone_map_kernel<true>, grid, block, 0, s, img, W, H, seed_s, whitemap, out);
  else hipLaunchKernelGGL(mtc::post::tone_map_kernel<false>, grid, block, 0, s, img, W, H, seed_s, whitemap, out);
  return hipGetLastError();
}

hipError_t launch_txaa(const TxaaConsts& tc, const uint32_t* cur, const uint32_t* prev, uint32_t* out, hipStream_t s) {
  dim3 grid((tc.width + 63) / 64, (tc.height + 3) / 4), block(256);
  hipLaunchKernelGGL(mtc::post::txaa_kernel, grid, block, 0, s, tc, cur, prev, out);
  return hipGetLastError();
}

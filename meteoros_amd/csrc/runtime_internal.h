// runtime_internal.h -- shared declarations of the host runtime.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "meteoros_cloud.h"

// Record `msg` as this thread's last error and return `code`.
int mtc_fail(int code, const std::string& msg);

// Per-frame constants computed once on the host (in FP32 with the same
// formulas the per-ray code would use) and passed to the kernels by value.
// Everything here is the same for every ray of a frame, because every ray
// starts at the eye (cloudRayMarch.comp:222, :707).
struct FrameConsts {
  // camera (cloudRayMarch.comp:15-21, :199-207)
  float view[16];
  float old_view[16];
  float eye[3];  // ray origin = -camera.eye.xyz
  float tfx, tfy;
  float jit_march[2];   // getJitterOffset(frame, dim) numerators, march table (:114-132)
  float jit_reproj[2];  // same, reprojection table (reprojection.comp:72-90)
  int32_t frame;        // frameCountMod16
  int32_t width, height;
  // per-step direction jitter (:631-633): (jx, (jx+jy)*jitter_factor, jy)
  // for _index = 0..15, jx, jy = getJitterOffset(_index, ivec2(75))
  float step_jitter[16][3];
  // geometry
  float earth_r, r_in, r_out, thickness, inner_off;
  double earth_r_d, r_in_d, r_out_d;
  // march controls
  int32_t step_mode, steps_min, steps_max, light_samples, coverage_mode, blur_samples;
  float base_density_factor, sample_divisor, erosion_scale, accum_scale, cone_erosion_scale, brightness;
  float hg_g, hg_si, hg_g2;  // hg_g2 = 0.99 - silver_spread
  float wind_dir[3], top_offset, wind_skew;
  float wind_offset[3];  // ((wind_dir + (0, bias, 0)) * cloud_speed) * total_time
  float coverage_const, weather_scale;
  float bg_min, fade_start, fade_end;
  // lighting (:585-624): lightDir = normalize(SUN_LOCATION - eye), cone kernel
  float light_dir[3];
  float kern[6][3];
  // Preetham background (:401-467): per-frame parts
  float sky_sun_dir[3];  // normalize(BACKGROUND_SKY_SUN_LOCATION - eye)
  float sunE;            // sunIntensity * calcSunIntensity()
  float beta_r[3], beta_m[3], beta_sum[3];
  float ydot_clamped;    // clamp((1 - sunDir.y)^5, 0, 1)
  float sun_disk_lo, sun_disk_hi;
  // reprojection per-frame parts (reprojection.comp:111-144, :147-191, :233)
  float rp_right[3], rp_up[3];  // normalize(view row 0), normalize(view row 1)
  float rp_ref[3];              // eye + look, look = -normalize(view row 2)
  float rp_ro[3];               // (eye - (eye.x, -R_earth, eye.z)) / R_in
  float rp_C;                   // dot(rp_ro, rp_ro) - 1
  float rp_frac[64];            // i / motion_blur_samples
  // optimised marches: per-frame values they would otherwise derive per
  // ray (kept out of vector registers); same FP32 operations as the shader
  float mk_inv_thick;   // 1 / thickness
  float mk_inv_div;     // 1 / sample_divisor
  float mk_rel_y0;      // eye.y - inner_off: relative height offset of the eye
  float mk_c_y0;        // eye.y + earth_r: eye height above the planet centre
  float mk_wk[3];       // wind_dir * (top_offset * wind_skew)
  int32_t mk_cap;       // 4 * max(steps_min, steps_max) + 64 (loop guard)
};

// Timing events attached to one kernel dispatch (MTC_FLAG_TIMING): the
// runtime records the kernel's start and end with the dispatch itself
// (hipExtLaunchKernelGGL), not with separate markers that would stall the
// queue between kernels.  Both null: an ordinary launch.
struct KernelEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};

template <class K, class... Args>
inline void launch_timed(K kernel, dim3 grid, dim3 block, hipStream_t s, const KernelEvents& ke, Args... args) {
  if (ke.start || ke.stop) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ke.start, ke.stop, 0u, args...);
  else hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

// Longest-first dispatch of the learned-light march.  A march's tiles
// differ in cost by ~100x (sky rows a few us, horizon rows 100+ us per
// wave); dealt in raster order the long ones start late and the end of the
// launch runs on a draining chip.  Each workgroup of a learned march
// therefore files its tile under the level of its measured duration
// (kOrderLevels half-octave levels of the 100 MHz real-time clock) in slot
// `wr`; the next learned march deals workgroup r the r-th tile of slot `rd`
// taken longest level first, and clears slot `clr`'s counts for the march
// after it.  The three slots rotate per march.  A slot is used only when
// its counts add up to exactly the launch's tile count, which holds only
// after one complete march of the same tiling filed every tile once, so the
// mapping is always a permutation; otherwise (first frame, full frames)
// workgroup r marches tile r.  The order only changes which workgroup
// marches which tile when, never a result.  (The fast and precision
// marches deal their tiles in the strip order below instead: for them,
// resident waves scattered over the image lose more texel-cache locality
// than the shorter tail gains.)
constexpr int kOrderLevels = 16;
struct TileOrder {
  const uint32_t* rd = nullptr;  // [kOrderLevels counts][kOrderLevels x cap tiles]
  uint32_t* wr = nullptr;
  uint32_t* clr = nullptr;
  uint32_t cap = 0;              // tiles per level (= the largest ordered tiling)
};
constexpr size_t order_slot_words(uint32_t cap) { return kOrderLevels + (size_t)kOrderLevels * cap; }

// Fills FrameConsts from the boundary structs (runtime_frame.cpp).
int build_frame_consts(const mtc_camera* cur, const mtc_camera* old, const mtc_time* time, const mtc_sun* sun,
                       const mtc_cloud_params* p, int width, int height, FrameConsts* fc);

// Device views of the resident resources.
struct VolumeViews {
  const void* lf;        // uint32 RGBA8 texels or 4 x half texels
  const void* hf;
  const uint32_t* curl;  // RGBA8
  const uint32_t* weather;
  int32_t lf_dim, hf_dim, curl_dim, weather_dim;
  int32_t lf_shift, hf_shift, curl_shift, weather_shift;  // log2 of the dims
  int32_t format;
};

// Strip order of the fast, precision and batched marches.  Workgroups go
// to the 8 XCDs in turn (workgroup b to XCD (b + c) % 8, c fixed per launch)
// and each XCD has an L2 of its own.  In raster order every XCD's resident
// waves are spread over the whole width of the rows being marched, and each
// L2 holds the texels of all of them.  Here the tiles of each block of k
// rows are numbered column by column and cut into nstrip equal strips of
// run = tiles_x * k / nstrip tiles (k is a multiple of nstrip /
// gcd(tiles_x, nstrip), so run is whole); the tiles b = x, x + 8, x + 16,
// ... (one XCD) are strips x, x + 8, ..., x + nstrip - 8, block by block
// from the top, so that L2 serves an eighth of the area, while each XCD's
// strips, spread over the width, hold sky, horizon and ocean in the same
// shares as the others'.  Rows past the last whole block keep raster order.
// A bijection on [0, tiles_x * tiles_y) for any nstrip that is a multiple
// of 8 and such a k; nstrip = 0 keeps raster order.  The order never
// changes a tile's result.
__host__ __device__ inline void strip_tile(int b, int tiles_x, int tiles_y, int nstrip, int k, int& tx, int& ty) {
  tx = b % tiles_x;
  ty = b / tiles_x;
  if (nstrip <= 0 || b >= (tiles_y / k) * tiles_x * k) return;
  const int run = tiles_x * k / nstrip;
  const int per = nstrip / 8 * run;  // tiles of one block in one XCD's strips
  const int j = b >> 3;
  const int m = j % per;
  const int idx = ((b & 7) + 8 * (m / run)) * run + m % run;
  tx = idx / k;
  ty = (j / per) * k + idx % k;
}

// Strip order for a launch tiles_x tiles wide: nstrip, a multiple of 8,
// gives strips about 2.5 tiles (20 rays of a 4x4 frame) wide; k = nstrip /
// gcd(tiles_x, nstrip) rows per block, the fewest that split evenly.
// Measured on MI355X (precision march): 1080p 4x4 frames (60 tiles wide)
// 24 strips -8% against raster order, 32 -7%, 40 -7%, 48 -6%; 4K 4x4
// frames (120 tiles wide) 48 strips within 0.3% of raster order, 32 +2%,
// 64 +0.4%, 128 +1.3%.  Three times the rows per block: 1080p +0.7%, 4K
// -0.7%.
struct StripOrder {
  int nstrip, k;
};
__host__ __device__ inline StripOrder strip_order(int tiles_x) {
  StripOrder o;
  const int s = (tiles_x + 10) / 20;
  o.nstrip = 8 * (s > 1 ? s : 1);
  int a = tiles_x, b = o.nstrip;
  while (b) {
    const int t = a % b;
    a = b;
    b = t;
  }
  o.k = o.nstrip / a;  // nstrip / gcd(tiles_x, nstrip)
  return o;
}

// Batched rendering: one work item = one tile of one view.
struct BatchItem {
  int view;
  int tile;
};
// The batch's items are kept in 8 lists, list x holding each view's tiles
// b = x, x + 8, ... of the strip order with 8 strips (one strip of each view
// per list); a workgroup takes items from the list of the XCD it runs on,
// then from the others.  List x is
// items[off[x], off[x] + len[x]); its counter is word kBatchQueueStride * x
// of the queue buffer (one 128-byte line each).
constexpr int kBatchQueueStride = 32;
struct BatchQueues {
  int off[8], len[8];
};
struct BatchView {
  FrameConsts fc;
  unsigned long long offset;  // first texel of the view in the packed outputs
  int tiles_x;                // tiles per row of marched rays
  int pad;
};

struct OutputViews {
  float4* color;      // current result image, W*H RGBA32F
  const float4* prev; // previous result image
  float4* mask;       // god-ray mask, W*H RGBA32F
  unsigned long long* counters;  // 12 x u64 or nullptr
  int32_t* reproj_idx;           // optional tap-index map
};

// Per-frame constants of the god-ray pass (postProcess_GodRays.frag:72-98),
// folded on the host.
struct PostConsts {
  float sun_ss[2];  // clamp((proj * view * (0, 1, 0, 1)).xy + 1) / 2, 0, 1)
  float blend;      // dot(normalize((0, 1, 0) - eye), camera forward)
  float light[3];   // sunAndSky.lightColor.rgb
  int32_t width, height;
};

// God-ray blur geometry.  A workgroup (four waves) blurs 64 columns x
// kGodRayGroupRows pixel rows; each wave holds four of the rows per lane.
// At every sample the texels the workgroup's taps read lie in at most
// kGodRayStageRows rows x kGodRayStageCols columns of the decoded mask,
// which the workgroup stages into LDS.
constexpr int kGodRaySamples = 100;                    // MAX_SAMPLES, postProcess_GodRays.frag:25
constexpr int kGodRayGroupRows = 16;
constexpr int kGodRayStageRows = kGodRayGroupRows + 2;  // 16 rows at spacing <= 1 span <= 17 texel rows + 1
constexpr int kGodRayStageCols = 72;                   // 64 lanes at spacing <= 1 + 4-aligned start + 1
constexpr int kGodRayTableStride = kGodRayGroupRows + 1;  // entries per (row group, sample)

// The staged region shrinks along the walk.  Sample i of adjacent pixels
// lies (1 - i / 100) texels apart, so from sample 25 p on (phase p) the 16
// rows and 64 columns of a workgroup span at most ceil(15 s) + 2 texel rows
// and ceil(63 s) + 6 columns (the region starts one column before the first
// lane's, aligned down to 4), s = 1 - p / 4, with at least half a texel to
// spare for the walk's FP32 rounding.  Rows x 4-column blocks staged per
// phase (the LDS row stride stays kGodRayStageCols):
constexpr int kGodRayPhaseSamples = kGodRaySamples / 4;
constexpr int kGodRayPhaseRows[4] = {18, 14, 10, 7};
constexpr int kGodRayPhaseBlocks[4] = {18, 14, 10, 6};

// One entry of the blur's table, per (row group, sample): entry 0 holds in
// `a` the first decoded-map row staged for the sample; entry 1 + k, for
// pixel row k of the group, the byte offset in the staged region of the
// texel row its tap reads, with the vertical weights w0 = 1 - wy, w1 = wy.
struct GodRayEntry {
  uint32_t a;
  float w0, w1;
  uint32_t pad;
};

// The decoded mask: map row r + 1, column c + 4 hold texel (c, r) for
// r in [-1, H], c in [-4, W + 71] (border value outside the image); rows
// and columns beyond the image let every staged region be read whole.
__host__ __device__ inline int god_rays_map_pitch(int W) { return (W + 76 + 3) / 4 * 4; }
__host__ __device__ inline int god_rays_map_rows(int H) { return H + 19; }
inline size_t god_rays_map_bytes(int W, int H) {
  return (size_t)god_rays_map_rows(H) * (size_t)god_rays_map_pitch(W) * sizeof(float);
}
inline size_t god_rays_table_bytes(int H) {
  return (size_t)((H + kGodRayGroupRows - 1) / kGodRayGroupRows) * kGodRaySamples * kGodRayTableStride *
         sizeof(GodRayEntry);
}
// Bytes of the pass's scratch at W x H: the map (256-byte aligned size),
// then the table.
inline size_t god_rays_map_span(int W, int H) { return (god_rays_map_bytes(W, H) + 255) / 256 * 256; }
inline size_t god_rays_scratch_bytes(int W, int H) { return god_rays_map_span(W, H) + god_rays_table_bytes(H); }

// Per-frame constants of the TXAA pass (postProcess_TXAA.frag:228-246),
// folded on the host in FP32 with the shader's operation order.
struct TxaaConsts {
  float eye[3];           // eyePos = -camera.eye
  float right[3], up[3];  // normalize(view row 0), normalize(view row 1)
  float ref[3];           // eye + look
  float jit[2];           // getJitterOffset(frameCountMod16, dim)
  float tfx, tfy;         // camera.tanFovBy2
  float old_view[16];     // cameraOld.view
  float ro[3];            // (eye - earthCenter) / ATMOSPHERE_RADIUS_INNER
  float C;                // dot(ro, ro) - 1
  float center[3];        // earthCenter = (eye.x, -EARTH_RADIUS, eye.z)
  float R;                // ATMOSPHERE_RADIUS_INNER
  int32_t width, height;
};

// Calibration bin of a network prediction y: 16 bins over [lo, hi],
// square-root spaced from the top (bin b holds 16 sqrt((hi - y) / (hi - lo))
// in [b, b + 1)), so the dense, small-magnitude predictions of thin cloud
// get narrow bins of their own.  inv_span = 1 / (hi - lo).  The same FP32
// operations on the host (calibration) and in the frame kernel.
__host__ __device__ inline int calib_bin(float y, float hi, float inv_span) {
  float t = (hi - y) * inv_span;
  t = t > 0.0f ? t : 0.0f;
  const int b = (int)(16.0f * sqrtf(t));
  return b < 15 ? b : 15;
}

// Learned-light network as a frame kernel sees it.
struct LearnedLight {
  const void* frag;       // packed BF16 MFMA weight fragments
  const float* params;    // FP32 parameter vector (biases, output layer)
  float calib_hi, calib_inv;
  unsigned int calib_mask;
  float fallback_density;
  int audit;
};

// The context behind the opaque mtc_ctx handle.
struct mtc_ctx {
  mtc_desc desc;
  hipStream_t stream = nullptr;
  void* lf = nullptr;
  void* hf = nullptr;
  uint32_t* curl = nullptr;
  uint32_t* weather = nullptr;
  float4* img[2] = {nullptr, nullptr};
  float4* mask = nullptr;
  unsigned long long* counters = nullptr;
  // footprint-packed copies of the resources read by the fast kernels
  uint4* q_lf = nullptr;
  uint4* q_hf = nullptr;
  uint4* q_curl = nullptr;
  uint4* q_weather = nullptr;
  // MTC_FLAG_LDS_BRICKS: the LF's half2 (r, fbm) texels in 4^3 bricks
  uint32_t* q_lf_brick = nullptr;
  // MTC_KERNEL_LEARNED: longest-first tile order, 3 slots of
  // order_slot_words(order_cap) words rotated by order_seq
  uint32_t* order = nullptr;
  uint32_t order_cap = 0;
  uint64_t order_seq = 0;
  int latest = 0;
  int lanes_per_ray = 8;  // fast kernels: lanes cooperating on one ray
  bool volumes_ready = false;
  // MTC_FLAG_TIMING: event triples (before reprojection, between, after
  // march) of frames not yet folded into the totals
  std::vector<std::array<hipEvent_t, 4>> pending;
  std::vector<hipEvent_t> event_pool;
  mtc_timings totals = {0, 0.0, 0.0};
  int32_t timing_stride = 1;   // events on one frame in timing_stride
  uint64_t frames_seen = 0;    // frames since creation (for the stride)
  // post-process: the god-ray pass's scratch (decoded mask and blur table,
  // god_rays_scratch_bytes) and the display image (RGBA8 or RGBA32F),
  // allocated on first use
  void* ray_scratch = nullptr;
  void* display = nullptr;
  int display_format = -1;
  // TXAA: RGBA8 history / output ping-pong, txaa_latest = the last output
  uint32_t* txaa[2] = {nullptr, nullptr};
  int txaa_latest = 0;
  // MTC_KERNEL_LEARNED: attached network (not owned) and its audit period
  mtc_mlp* mlp = nullptr;
  int mlp_audit = 0;
  size_t lf_bytes = 0, hf_bytes = 0, curl_bytes = 0, weather_bytes = 0;
};

// Device views of a context's base resources.
VolumeViews volume_views(const mtc_ctx* c);

// The attached network's device views (mlp_api.cpp); MTC_ERR_STATE when
// none is attached.
int learned_light_of(const mtc_ctx* c, LearnedLight* out);

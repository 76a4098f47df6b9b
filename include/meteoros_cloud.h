/*
 * meteoros_cloud.h -- C ABI of the MI355X cloud compute stage.
 *
 * This header is the drop-in boundary that replaces the Vulkan compute
 * pipeline interface of Meteoros' cloud pass:
 *
 *   set 0  ping-pong result images       Renderer.cpp:924-927, :1034-1086
 *   set 1  LF/HF sampler3D, curl/weather sampler2D, god-ray image
 *                                         Renderer.cpp:930-938, :1087-1165
 *   set 2  CameraUBO (cloud pass)         camera.h:12-18, cloudRayMarch.comp:15-21
 *   set 3  TimeUBO                        Scene.h:12-21, cloudRayMarch.comp:23-31
 *   set 4  SunAndSkyUBO                   Sky.h:9-15, cloudRayMarch.comp:33-39
 *   reprojection set 2 = old CameraUBO    reprojection.comp:21-27
 *   dispatch order reprojection -> march  Renderer.cpp:653-722
 *
 * Plain pointers and sizes only.  Every entry point returns an mtc_status;
 * no C++ exception crosses the boundary.  A context is not internally
 * locked: use one context per host thread.
 */
#ifndef METEOROS_CLOUD_H
#define METEOROS_CLOUD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* status codes                                                        */
/* ------------------------------------------------------------------ */
typedef enum mtc_status {
  MTC_OK = 0,
  MTC_ERR_INVALID_ARG = 1,   /* null pointer, bad size, bad enum           */
  MTC_ERR_OUT_OF_MEMORY = 2, /* device or host allocation failed           */
  MTC_ERR_HIP = 3,           /* a HIP runtime call failed                  */
  MTC_ERR_UNSUPPORTED = 4,   /* valid request this build does not handle   */
  MTC_ERR_STATE = 5          /* call out of order (e.g. no volumes yet)    */
} mtc_status;

/* ------------------------------------------------------------------ */
/* uniform blocks, byte-identical to the reference's std140 UBOs       */
/* ------------------------------------------------------------------ */

/* CameraUBO (camera.h:12-18).  Matrices are column-major, view[c*4+r].
 * The shaders use eye = -eye.xyz as the ray origin
 * (cloudRayMarch.comp:707, reprojection.comp:205). */
typedef struct mtc_camera {
  float view[16];        /* offset   0 */
  float proj[16];        /* offset  64 */
  float eye[4];          /* offset 128 */
  float tan_fov_by2[2];  /* offset 144: (aspect*tan(fovy/2), tan(fovy/2)) */
} mtc_camera;            /* 152 bytes */

/* TimeUBO (Scene.h:12-21, cloudRayMarch.comp:23-31). */
typedef struct mtc_time {
  float halton1[4];      /* offset  0: H3(1..4)   */
  float halton2[4];      /* offset 16: H3(5..8)   */
  float halton3[4];      /* offset 32: H3(9..12)  */
  float halton4[4];      /* offset 48: H3(13..16) */
  float time[2];         /* offset 64: (delta, total) seconds */
  int32_t frame_count_mod16; /* offset 72 */
} mtc_time;              /* 76 bytes */

/* SunAndSkyUBO (Sky.h:9-15).  The shipped shader never reads it and uses
 * #define SUN_LOCATION / SUN_INTENSITY instead (cloudRayMarch.comp:94-99);
 * here sun_location.xyz is the lighting sun (SUN_LOCATION) and
 * sun_intensity scales the Preetham background (SUN_INTENSITY). */
typedef struct mtc_sun {
  float sun_location[4];  /* offset  0 */
  float sun_direction[4]; /* offset 16 (carried, unused by the stage) */
  float light_color[4];   /* offset 32 (carried, unused by the stage) */
  float sun_intensity;    /* offset 48 */
} mtc_sun;                /* 52 bytes */

/* ------------------------------------------------------------------ */
/* runtime configuration: every #define of cloudRayMarch.comp and       */
/* reprojection.comp that the configurations vary                       */
/* ------------------------------------------------------------------ */
enum {
  MTC_STEPS_REFERENCE = 0, /* floor(mix(steps_min, steps_max, 1-dir.y)), cloudRayMarch.comp:572 */
  MTC_STEPS_FIXED = 1      /* exactly steps_max as the divisor of the shell span */
};
enum {
  MTC_COVERAGE_CONSTANT = 0, /* cloud_coverage = coverage_constant (cloudRayMarch.comp:529) */
  MTC_COVERAGE_WEATHER = 1   /* coverage = weather.r, height gradient from weather.g
                                (the path commented out at cloudRayMarch.comp:515-525) */
};

typedef struct mtc_cloud_params {
  /* geometry, metres (cloudRayMarch.comp:75-78) */
  double earth_radius;        /* EARTH_RADIUS */
  double shell_inner_radius;  /* ATMOSPHERE_RADIUS_INNER */
  double shell_outer_radius;  /* ATMOSPHERE_RADIUS_OUTER */
  /* background sky sun (BACKGROUND_SKY_SUN_LOCATION, :95) */
  float sky_sun_location[4];
  /* primary march (:565-688) */
  int32_t step_mode;          /* MTC_STEPS_* */
  int32_t steps_min;          /* 35 */
  int32_t steps_max;          /* 60 */
  int32_t light_samples;      /* 6 (cone samples, :652), at most 6 */
  float base_density_factor;  /* 0.38 (:571) */
  float jitter_factor;        /* 1.18 (:568) */
  float jitter_dim;           /* 75, the ivec2(75) of :632 */
  float sample_divisor;       /* 8, :635 */
  float erosion_scale;        /* 1.4, :645 */
  float accum_scale;          /* 0.5, :648 */
  float cone_erosion_scale;   /* 1.5, :665 */
  float brightness;           /* 5.0, :672 */
  /* Henyey-Greenstein (:587-590) */
  float hg_eccentricity;      /* 0.6 */
  float hg_silver_intensity;  /* 0.7 */
  float hg_silver_spread;     /* 0.1 */
  /* wind (:89-91, :489-497) */
  float wind_direction[4];    /* (1,0,0) */
  float wind_upward_bias;     /* 0.1 */
  float cloud_speed;          /* 0.080 sample-space units per second */
  float cloud_top_offset;     /* 1.0 */
  float wind_skew;            /* 0.009 */
  /* coverage (:529) */
  int32_t coverage_mode;      /* MTC_COVERAGE_* */
  float coverage_constant;    /* 0.6 */
  float weather_scale;        /* weather uv = unskewed point.xz * weather_scale */
  /* composite (:713-716, :762) */
  float background_min;       /* 0.620 */
  float cloud_fade_start;     /* 0.06 */
  float cloud_fade_end;       /* 0.2 */
  /* reprojection (reprojection.comp:66) */
  int32_t motion_blur_samples;/* 10 */
  int32_t reserved[3];
} mtc_cloud_params;

/* ------------------------------------------------------------------ */
/* host-side uniform producers (camera.cpp, Scene.cpp, Sky.cpp)        */
/* ------------------------------------------------------------------ */

/* State of camera.h's Camera class (eye, ref, derived basis, lens). */
typedef struct mtc_camera_state {
  float eye[3];
  float ref[3];
  float forward[3];
  float right[3];
  float up[3];
  float world_up[3];
  int32_t width, height;
  float fovy_deg;
  float aspect;
  float near_clip, far_clip;
} mtc_camera_state;

/* Camera::Camera + RecomputeAttributes (camera.cpp:3-17, :68-77). */
int mtc_camera_init(mtc_camera_state* cam, const float eye[3], const float ref[3],
                    int32_t width, int32_t height, float fovy_deg,
                    float near_clip, float far_clip);
/* Camera::RotateAboutUp / RotateAboutRight (camera.cpp:79-96). */
int mtc_camera_rotate_about_up(mtc_camera_state* cam, float deg);
int mtc_camera_rotate_about_right(mtc_camera_state* cam, float deg);
/* Camera::TranslateAlong{Look,Right,Up} (camera.cpp:98-117). */
int mtc_camera_translate(mtc_camera_state* cam, float look, float right, float up);
/* Camera::UpdateBuffer (camera.cpp:31-42): glm lookAt / perspective (ZO),
 * y-flipped projection, tanFovBy2 with PI = 3.14159. */
int mtc_camera_ubo(const mtc_camera_state* cam, mtc_camera* out);

/* Scene::HaltonSequenceAt (Scene.cpp:125-138). */
float mtc_halton(int32_t index, int32_t base);
/* Scene::InitializeTime (Scene.cpp:86-119) with delta = total = 0. */
int mtc_time_init(mtc_time* t);
/* Scene::UpdateTime (Scene.cpp:65-85) with the wall clock replaced by dt. */
int mtc_time_advance(mtc_time* t, float dt);

/* Shipped constants of cloudRayMarch.comp / reprojection.comp. */
int mtc_params_reference(mtc_cloud_params* p, mtc_sun* sun);
/* The workload presets (1..5) of this stage; see mtc_preset_info. */
int mtc_params_preset(int32_t preset, mtc_cloud_params* p, mtc_sun* sun);
/* Sun at (elevation, azimuth) degrees, azimuth 0 = -z, at distance_m. */
int mtc_sun_from_angles(float elevation_deg, float azimuth_deg, float distance_m,
                        mtc_sun* sun, mtc_cloud_params* p);
/* Convert a wind speed in m/s into cloud_speed sample units per second:
 * one sample unit spans sample_divisor * shell thickness metres. */
float mtc_wind_speed_units(const mtc_cloud_params* p, float metres_per_second);

/* ------------------------------------------------------------------ */
/* volumes                                                             */
/* ------------------------------------------------------------------ */
enum {
  MTC_FMT_UNORM8 = 0, /* 4 x uint8 per texel, R8G8B8A8_UNORM (Sky.cpp:31-57) */
  MTC_FMT_FP16 = 1    /* 4 x IEEE half per texel */
};

/* Noise and map resources.  Layout of every array: x fastest, then y,
 * then z, tightly packed, 4 channels per texel
 * (ImageLoadingUtility.cpp:75-139, Image.cpp:181-212).  Dimensions must
 * be powers of two (REPEAT wrap is a mask): volumes up to 1024^3 with
 * MTC_KERNEL_PLAIN, up to 512^3 with the other kernels; maps up to
 * 8192^2. */
typedef struct mtc_volume_desc {
  int32_t lf_dim;       /* low-frequency Perlin-Worley volume, lf_dim^3   */
  int32_t hf_dim;       /* high-frequency Worley volume, hf_dim^3         */
  int32_t curl_dim;     /* curl-noise map, curl_dim^2                     */
  int32_t weather_dim;  /* weather map, weather_dim^2                     */
  int32_t format;       /* MTC_FMT_* of lf and hf (maps are always UNORM8) */
} mtc_volume_desc;

/* CPU-identical noise generator settings (kernel a). */
typedef struct mtc_noise_desc {
  uint32_t seed_volumes;   /* LF / HF / curl seed */
  uint32_t seed_weather;   /* weather map seed     */
} mtc_noise_desc;

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
enum {
  MTC_KERNEL_PLAIN = 0,     /* one lane per ray, FP32, GLSL op order          */
  MTC_KERNEL_FAST = 1,      /* FP16 texel cache + hoisted cone, FP32 accum    */
  MTC_KERNEL_PRECISION = 2, /* FP16 volumes, packed-FP16 filtering, FP64 shell */
  MTC_KERNEL_LEARNED = 3    /* FAST with the cone march answered by the MLP   */
};

enum {
  MTC_FLAG_COUNTERS = 1,     /* accumulate mtc_counters with device atomics  */
  MTC_FLAG_FULL_MARCH = 2,   /* march every pixel (no 4x4 selection)          */
  MTC_FLAG_TIMING = 4,       /* hipEvents around each kernel on its stream
                                (mtc_get_timings); frames keep their structure */
  MTC_FLAG_LDS_BRICKS = 8    /* fast / precision frames, 8 lanes per ray: the
                                primary steps' LF samples go through an 8^3
                                texel box per wave staged in LDS from a 4^3-
                                bricked copy of the LF (+4 bytes per LF texel);
                                results bit-identical to the default path */
};

typedef struct mtc_desc {
  int32_t width;
  int32_t height;
  mtc_volume_desc volumes;
  int32_t kernel;           /* MTC_KERNEL_* */
  int32_t flags;            /* MTC_FLAG_*   */
  int32_t device;           /* HIP device ordinal */
  int32_t lanes_per_ray;    /* fast kernels: lanes cooperating on one ray,
                               1, 2, 4 or 8 (0 = default 8) */
} mtc_desc;

/* Per-frame work counters (counting build / MTC_FLAG_COUNTERS). */
typedef struct mtc_counters {
  uint64_t rays_marched;     /* rays that entered rayMarch            */
  uint64_t rays_ocean;       /* dir.y < 0 early-outs                  */
  uint64_t rays_sky;         /* 0 <= dir.y < fade start early-outs    */
  uint64_t primary_steps;    /* executed primary iterations (ray-steps) */
  uint64_t cloud_steps;      /* primary iterations with base density > 0 */
  uint64_t cone_samples;     /* cone samples (learned frames: 1 per answered step) */
  uint64_t cone_erosions;    /* cone samples with density > 0          */
  uint64_t early_exits;      /* accumDensity >= 1 breaks              */
  uint64_t mlp_queries;      /* learned-light queries                 */
  uint64_t mlp_fallbacks;    /* learned-light queries sent to the cone */
  uint64_t mlp_audits;       /* network answers also cone-marched     */
  uint64_t mlp_audit_sq_err; /* sum over audits of (network - cone)^2
                                optical depth, in units of 1e-6       */
  uint64_t lf_brick_stages;  /* MTC_FLAG_LDS_BRICKS: LDS box stagings  */
  uint64_t lf_brick_samples; /* primary LF samples read from LDS       */
} mtc_counters;

/* Kernel time accumulated with hipEvents on the stream each kernel runs
 * on (MTC_FLAG_TIMING), milliseconds summed over `frames` frames.  The
 * fast, precision and learned-light contexts run a frame's reprojection
 * inside the march launch: march_ms covers both and reproject_ms stays 0. */
typedef struct mtc_timings {
  uint64_t frames;
  double reproject_ms;
  double march_ms;
} mtc_timings;

typedef struct mtc_ctx mtc_ctx;

/* Last error message of this thread (never NULL). */
const char* mtc_last_error(void);
/* Library version string. */
const char* mtc_version(void);
/* Number of HIP devices visible (0 when none; never an error). */
int mtc_device_count(int32_t* count);

int mtc_create(const mtc_desc* desc, mtc_ctx** out);
int mtc_destroy(mtc_ctx* ctx);

/* Upload volumes (Sky::CreateCloudResources, Sky.cpp:25-58).  lf/hf use
 * desc.volumes.format; curl and weather are UNORM8 RGBA.  Host memory is
 * only read during the call. */
int mtc_set_volumes(mtc_ctx* ctx, const void* lf, const void* hf,
                    const void* curl, const void* weather);
/* Generate the volumes on the device (kernel a). */
int mtc_generate_volumes(mtc_ctx* ctx, const mtc_noise_desc* noise);
/* Copy the resident volumes back to host memory (any pointer may be NULL). */
int mtc_get_volumes(mtc_ctx* ctx, void* lf, void* hf, void* curl, void* weather);

/* One frame of the cloud compute stage (Renderer.cpp:653-722):
 * reprojection of the previous result into the current image and the ray
 * march of the 1/16 pixels selected by time->frame_count_mod16, then the
 * ping-pong swap.  No pixel is written twice: plain contexts run the two
 * kernels in that order; the optimised contexts run one launch whose
 * reprojection leaves the marched pixels out.  Enqueued on `stream`
 * (hipStream_t, NULL = the context's stream); uniforms are captured by
 * value at enqueue time. */
int mtc_frame(mtc_ctx* ctx, const mtc_camera* cur, const mtc_camera* old,
              const mtc_time* time, const mtc_sun* sun,
              const mtc_cloud_params* params, void* stream);
/* The ray march alone over every pixel of the frame (no reprojection, no
 * swap); result lands in the current image. */
int mtc_march_full(mtc_ctx* ctx, const mtc_camera* cur, const mtc_time* time,
                   const mtc_sun* sun, const mtc_cloud_params* params, void* stream);

/* Wait for all work of the context. */
int mtc_synchronize(mtc_ctx* ctx);
/* Copy the latest result (RGBA32F, W*H*4 floats) and the god-ray mask
 * (RGBA32F EncodeFloatRGBA, W*H*4 floats).  Either pointer may be NULL. */
int mtc_read_output(mtc_ctx* ctx, float* rgba, float* mask, size_t n_floats);
/* Device pointers of the latest result and mask (valid until the next
 * mtc_frame / mtc_destroy). */
int mtc_output_device_ptrs(mtc_ctx* ctx, void** rgba, void** mask);
/* Previous-frame image used by the next reprojection. */
int mtc_set_history(mtc_ctx* ctx, const float* rgba, size_t n_floats);
int mtc_reset_history(mtc_ctx* ctx);

/* Counters accumulated since the last reset (needs MTC_FLAG_COUNTERS). */
int mtc_get_counters(mtc_ctx* ctx, mtc_counters* out);
int mtc_reset_counters(mtc_ctx* ctx);

/* Kernel timings since the last reset (needs MTC_FLAG_TIMING; waits for
 * the recorded events). */
int mtc_get_timings(mtc_ctx* ctx, mtc_timings* out);
int mtc_reset_timings(mtc_ctx* ctx);
/* With MTC_FLAG_TIMING, attach the timing events to one frame in every
 * `stride` (>= 1; default 1): the events cost each timed frame a few
 * percent, so a long sequence can be timed on a sample of its frames.  No
 * counterpart in the reference (a profiling control of this library). */
int mtc_set_timing_stride(mtc_ctx* ctx, int32_t stride);

/* Reprojection source-texel map for testing: for every pixel and tap the
 * linear index y*W+x of the previous-frame texel the tap reads
 * (reprojection.comp:233-240).  idx holds W*H*motion_blur_samples int32. */
int mtc_reprojection_indices(mtc_ctx* ctx, const mtc_camera* cur, const mtc_camera* old,
                             const mtc_time* time, const mtc_cloud_params* params,
                             int32_t* idx, size_t n_idx);

/* ------------------------------------------------------------------ */
/* loaders for the textures Meteoros ships (host only)                 */
/* ------------------------------------------------------------------ */
/* One 2D image (TGA or PNG) decoded to RGBA8, top row first, as
 * stbi_load(path, ..., 4) returns it (Texture2D::createTextureFromFile,
 * Sky.cpp:44-57).  width/height may be NULL; out may be NULL to query the
 * size. */
int mtc_load_image_rgba8(const char* path, int32_t* width, int32_t* height, uint8_t* out, size_t n_bytes);
/* An n^3 volume from n slices "<dir>/<base>(z+1).tga" into RGBA8 texels,
 * x fastest, then y, then z (ImageLoadingUtility.cpp:75-139); out holds
 * n^3 * 4 bytes, ready for mtc_set_volumes. */
int mtc_load_tga_slices(const char* dir, const char* base, int32_t n, uint8_t* out, size_t n_bytes);

/* ------------------------------------------------------------------ */
/* post-process passes that consume the stage's output                 */
/* ------------------------------------------------------------------ */
enum {
  MTC_DISPLAY_RGBA8 = 0,   /* RGBA8 UNORM, the tone-map shader's rgba8 store */
  MTC_DISPLAY_RGBA32F = 1  /* the same values before 8-bit quantisation       */
};

/* God-ray radial blur (postProcess_GodRays.frag:66-150; the reference
 * records it but leaves it commented out, Renderer.cpp:826-832): adds
 * 100 samples of the decoded god-ray mask along the pixel-to-sun line to
 * the latest result image, in place.  Uses camera view/proj/eye and
 * sun->light_color.  A sun behind the camera leaves the image unchanged. */
int mtc_god_rays(mtc_ctx* ctx, const mtc_camera* cam, const mtc_sun* sun, void* stream);
/* Sizes that mtc_god_rays_noalloc accepts: (width + 1) * (height + 1) up
 * to MTC_GOD_RAYS_NOALLOC_MAX_ENTRIES (e.g. 1920x1200 or 2048x1080),
 * height up to MTC_GOD_RAYS_NOALLOC_MAX_HEIGHT and width up to
 * MTC_GOD_RAYS_NOALLOC_MAX_WIDTH.  Within them its scratch (the decoded
 * mask, padded, and the blur's table of 1 700 bytes per pixel row) fits in
 * MTC_GOD_RAYS_NOALLOC_SCRATCH_BYTES (18 MiB). */
#define MTC_GOD_RAYS_NOALLOC_MAX_ENTRIES 2359296
#define MTC_GOD_RAYS_NOALLOC_MAX_HEIGHT 4320
#define MTC_GOD_RAYS_NOALLOC_MAX_WIDTH 16384
#define MTC_GOD_RAYS_NOALLOC_SCRATCH_BYTES 18874368
/* No counterpart in the reference: an entry point that allocates nothing,
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * god-ray pass of mtc_god_rays, without a context, on caller-owned device
 * memory of the current device: `mask` is the width*height RGBA32F god-ray
 * mask and `image` the width*height RGBA32F result, updated in place.  The
 * pass's scratch (the decoded mask and the blur's table) is a static
 * device array of MTC_GOD_RAYS_NOALLOC_SCRATCH_BYTES that each device holds
 * from the library's first load until the process ends; sizes beyond the
 * limits above are rejected (MTC_ERR_INVALID_ARG) before anything is
 * enqueued.  Enqueued on
 * `stream` (NULL = the null stream).  All calls on one device share that
 * array, so calls that overlap on one device race.  The first call on a
 * device looks up the array's address; make it outside graph capture. */
int mtc_god_rays_noalloc(int32_t width, int32_t height, const mtc_camera* cam, const mtc_sun* sun,
                         const float* mask, float* image, void* stream);
/* Uncharted-2 tone map + Wang-hash dither (postProcess_ToneMap.frag:68-84)
 * of the latest result image into the context's display image. */
int mtc_tone_map(mtc_ctx* ctx, const mtc_time* time, int32_t format, void* stream);
/* Copy the display image (W*H*4 bytes for RGBA8, W*H*16 for RGBA32F). */
int mtc_read_display(mtc_ctx* ctx, void* out, size_t n_bytes);
/* TXAA (postProcess_TXAA.frag:171-270; Renderer.cpp:840-846) of the RGBA8
 * display image against the previous TXAA output: reprojection through the
 * shader's own inner shell (R + 7.5 km), 3x3 / cross neighbourhood clip,
 * luminance-weighted feedback in [0, 0.5].  The result (RGBA8) becomes the
 * next call's history; the first history is black.  Needs an RGBA8
 * mtc_tone_map first.  The reference blends in place, reading neighbours
 * other fragments may have overwritten; here the current image is read-only. */
int mtc_txaa(mtc_ctx* ctx, const mtc_camera* cur, const mtc_camera* old, const mtc_time* time, void* stream);
/* Copy the latest TXAA output (W*H*4 bytes). */
int mtc_read_txaa(mtc_ctx* ctx, uint8_t* rgba8, size_t n_bytes);
/* Replace the TXAA history (W*H*4 bytes). */
int mtc_set_txaa_history(mtc_ctx* ctx, const uint8_t* rgba8, size_t n_bytes);

/* ------------------------------------------------------------------ */
/* batched views: many views of uneven size in one persistent launch  */
/* (no reference counterpart; the reference renders one view).        */
/* ------------------------------------------------------------------ */
typedef struct mtc_batch mtc_batch;

/* Launch shape of a batch; chosen by mtc_batch_autotune or by the caller. */
typedef struct mtc_batch_tuning {
  int32_t lanes_per_ray;   /* 4 or 8 lanes cooperate on one ray          */
  int32_t tile_rays_x;     /* tile width in rays; height = 256 / lanes / width */
  int32_t blocks_per_cu;   /* persistent workgroups per compute unit     */
  int32_t order;           /* 0: work items by descending view cost (steps x
                              rays), 1: in view order                    */
} mtc_batch_tuning;

/* A batch of n_views views (sizes fixed at creation) rendered with ctx's
 * volumes; ctx must be a MTC_KERNEL_FAST or MTC_KERNEL_PRECISION context.
 * Every pixel of every view is marched (no reprojection). */
int mtc_batch_create(mtc_ctx* ctx, int32_t n_views, const int32_t* widths, const int32_t* heights,
                     mtc_batch** out);
int mtc_batch_destroy(mtc_batch* b);
int mtc_batch_set_tuning(mtc_batch* b, const mtc_batch_tuning* t);
int mtc_batch_get_tuning(mtc_batch* b, mtc_batch_tuning* t);
/* Render all views: per-view uniform arrays of length n_views.  All views
 * must share coverage_mode.  The launch sequence (queue reset + persistent
 * kernel) is captured once per tuning into a hipGraph and replayed. */
int mtc_batch_render(mtc_batch* b, const mtc_camera* cams, const mtc_time* times, const mtc_sun* suns,
                     const mtc_cloud_params* params, void* stream);
/* Time candidate tunings on these inputs and keep the fastest; *best_ms is
 * its render time (median of the timed repeats). */
int mtc_batch_autotune(mtc_batch* b, const mtc_camera* cams, const mtc_time* times, const mtc_sun* suns,
                       const mtc_cloud_params* params, mtc_batch_tuning* best, double* best_ms);
/* Wall time of the last render measured with hipEvents (waits for it). */
int mtc_batch_last_ms(mtc_batch* b, double* ms);
/* Copy one view's RGBA32F colour and mask (W*H*4 floats each; either may be NULL). */
int mtc_batch_read(mtc_batch* b, int32_t view, float* rgba, float* mask, size_t n_floats);

/* ------------------------------------------------------------------ */
/* learned light: an MLP for the cone light march (no reference        */
/* counterpart; it stands in for GetLightEnergy's optical depth,       */
/* cloudRayMarch.comp:338-373)                                         */
/* ------------------------------------------------------------------ */
/* Network 9 -> 64 -> 64 -> 64 -> 1 with SiLU.  Parameters are one FP32
 * vector: W1[9][64], b1[64], W2[64][64], b2, W3[64][64], b3, w4[64], b4
 * (W[k][n]: input k, output n).  Inputs of an in-cloud step: position in
 * the shell (x, z modulo the weather tile, y modulo the LF tile), height
 * fraction, local density, sun direction (3), cos of the view ray's
 * zenith angle.  Output: log transmittance of the light march, -sum of
 * the cone's density terms. */
enum { MTC_MLP_INPUTS = 9, MTC_MLP_PARAMS = 9025 };

typedef struct mtc_mlp mtc_mlp;

typedef struct mtc_mlp_train_desc {
  int32_t steps;         /* Adam steps                                   */
  int32_t batch;         /* minibatch size                              */
  float lr;              /* Adam learning rate                          */
  float beta1, beta2, eps;
  uint32_t seed;         /* minibatch sampling seed (nonzero)           */
  int64_t holdout;       /* trailing samples kept out of training        */
} mtc_mlp_train_desc;

/* Held-out error of the network by predicted-value bin, and the fallback
 * rule a learned frame applies: a step goes to the cone when its local
 * density exceeds fallback_density (it then does not query the network)
 * or its prediction falls in a bin whose RMS error exceeds max_rms.  The
 * 16 bins cover [lo, hi] square-root spaced from the top: bin b holds
 * predictions y with 16 sqrt((hi - y) / (hi - lo)) in [b, b + 1). */
typedef struct mtc_mlp_calibration {
  float lo, hi;             /* bin range of the predicted value          */
  float rms[16];            /* held-out RMS error per bin                */
  uint32_t count[16];       /* held-out samples per bin                  */
  uint32_t mask;            /* bins sent to the cone                     */
  float max_rms;
  float fallback_density;
  float rms_all;            /* held-out RMS error over all bins          */
  float rms_mean_predictor; /* RMS of predicting the training mean       */
} mtc_mlp_calibration;

/* A network with room for `capacity` training samples on `device`.  The
 * parameters start from a seeded fan-in-scaled uniform draw (seed 1). */
int mtc_mlp_create(int32_t device, int64_t capacity, mtc_mlp** out);
int mtc_mlp_destroy(mtc_mlp* m);
/* Re-draw the parameters from `seed` and clear the optimiser state. */
int mtc_mlp_init(mtc_mlp* m, uint32_t seed);
int mtc_mlp_set_params(mtc_mlp* m, const float* params, size_t n);
int mtc_mlp_get_params(mtc_mlp* m, float* params, size_t n);
/* Append training samples: every pixel of ctx's view is marched with
 * these uniforms (ctx's volumes; ctx must not be MTC_KERNEL_PLAIN) and
 * each in-cloud step is kept with probability keep_fraction.  *n_total
 * (may be NULL) receives the stored sample count; samples beyond the
 * capacity are dropped. */
int mtc_mlp_collect(mtc_mlp* m, mtc_ctx* ctx, const mtc_camera* cam, const mtc_time* time,
                    const mtc_sun* sun, const mtc_cloud_params* params, uint32_t seed,
                    float keep_fraction, int64_t* n_total);
int mtc_mlp_sample_count(mtc_mlp* m, int64_t* n);
int mtc_mlp_clear_samples(mtc_mlp* m);
/* Replace the stored samples (xs: n * 9 floats, ys: n floats); their
 * keys read back as UINT64_MAX. */
int mtc_mlp_set_samples(mtc_mlp* m, const float* xs, const float* ys, int64_t n);
/* Copy samples [first, first + n): inputs, targets and the key of each
 * collected sample, py << 40 | px << 20 | step index.  Any output may be
 * NULL. */
int mtc_mlp_read_samples(mtc_mlp* m, int64_t first, int64_t n, float* xs, float* ys, uint64_t* keys);
/* Minibatch Adam on the stored samples (all but the trailing holdout),
 * loss = mean squared error.  loss_curve (may be NULL) receives the
 * minibatch loss of every step. */
int mtc_mlp_train(mtc_mlp* m, const mtc_mlp_train_desc* d, float* loss_curve);
/* Gradient of mean((f(x) - y)^2) over n host samples (n <= 65536) at the
 * current parameters, as the training step computes it; grad holds
 * MTC_MLP_PARAMS floats. */
int mtc_mlp_gradient(mtc_mlp* m, const float* xs, const float* ys, int32_t n, float* grad, float* loss);
/* Network outputs for n host inputs, through the frame kernel's
 * inference routine. */
int mtc_mlp_predict(mtc_mlp* m, const float* xs, int32_t n, float* ys);
/* Measure the held-out error and set the fallback rule. */
int mtc_mlp_calibrate(mtc_mlp* m, int64_t holdout, float max_rms, float fallback_density,
                      mtc_mlp_calibration* out);
/* Attach a network to a MTC_KERNEL_LEARNED context (it must outlive its
 * use there).  One in audit_period (power of two, 0 = none) of the steps
 * the network answers is also cone-marched and its error counted
 * (mtc_counters.mlp_audits / mlp_audit_sq_err). */
int mtc_set_learned_light(mtc_ctx* ctx, mtc_mlp* m, int32_t audit_period);

#ifdef __cplusplus
}
#endif
#endif /* METEOROS_CLOUD_H */

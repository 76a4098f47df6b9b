// batch.cpp -- many views of uneven size in one persistent launch.
//
// No reference counterpart: the reference renders one view per frame
// (Renderer.cpp:653-722).  A batch owns packed RGBA32F colour and mask
// outputs for all its views, a device array of per-view FrameConsts, and a
// list of (view, tile) work items.  A render uploads the views' uniforms,
// then replays a hipGraph holding [queue reset kernel, persistent march
// kernel]; the graph is captured once per (tuning, coverage mode, variant,
// item count).  The persistent grid pulls items from 8 atomic lists (each
// view's tiles dealt to the XCDs in column strips, one list per XCD; an
// XCD's workgroups take from the other lists once theirs is empty), so
// uneven views balance themselves; ordering items by descending view cost
// (steps x rays) keeps the longest work from landing last.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "runtime_internal.h"

hipError_t launch_march_batch(const BatchView* views, const BatchItem* items, const BatchQueues& bq, unsigned int* queue,
                              const uint4* lf, const uint4* hf, const uint4* curl, const uint4* weather,
                              const VolumeViews& vv, float4* color, float4* mask, int coverage_mode, int variant,
                              int lanes_per_ray, int tile_x, int blocks, hipStream_t s);

struct mtc_batch {
  mtc_ctx* ctx = nullptr;
  int n = 0;
  std::vector<int> w, h;
  std::vector<unsigned long long> offset;
  size_t total_px = 0;
  float4* color = nullptr;
  float4* mask = nullptr;
  BatchView* d_views = nullptr;
  BatchView* h_views = nullptr;  // pinned staging
  BatchItem* d_items = nullptr;
  size_t items_cap = 0;
  int n_items = 0;
  BatchQueues queues = {};
  unsigned int* d_queue = nullptr;  // 8 list counters, kBatchQueueStride words apart
  hipStream_t stream = nullptr;
  mtc_batch_tuning tuning = {8, 8, 4, 0};
  std::vector<int> item_steps;  // per-view steps the item order was built for
  mtc_batch_tuning item_tuning = {0, 0, 0, -1};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  mtc_batch_tuning graph_tuning = {0, 0, 0, -1};
  int graph_cov = -1, graph_items = -1;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int cus = 256;
  bool rendered = false;
};

namespace {

#define HIP_OK(expr, what)                                                                        \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) {                                                                       \
      return mtc_fail(_e == hipErrorOutOfMemory ? MTC_ERR_OUT_OF_MEMORY : MTC_ERR_HIP,            \
                      std::string(what) + ": " + hipGetErrorName(_e));                            \
    }                                                                                             \
  } while (0)

bool valid_tuning(const mtc_batch_tuning& t) {
  if (t.lanes_per_ray != 4 && t.lanes_per_ray != 8) return false;
  const int rays = 256 / t.lanes_per_ray;
  if (t.tile_rays_x < 1 || t.tile_rays_x > rays || rays % t.tile_rays_x != 0) return false;
  if (t.blocks_per_cu < 1 || t.blocks_per_cu > 16) return false;
  return t.order == 0 || t.order == 1;
}

bool same_tuning(const mtc_batch_tuning& a, const mtc_batch_tuning& b) {
  return a.lanes_per_ray == b.lanes_per_ray && a.tile_rays_x == b.tile_rays_x && a.blocks_per_cu == b.blocks_per_cu &&
         a.order == b.order;
}

void drop_graph(mtc_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// (Re)build the work-item list for the tuning's tile shape and order.
int build_items(mtc_batch* b, const mtc_cloud_params* params) {
  std::vector<int> steps(b->n);
  for (int v = 0; v < b->n; ++v) steps[v] = std::max(params[v].steps_min, params[v].steps_max);
  if (same_tuning(b->tuning, b->item_tuning) && steps == b->item_steps) return MTC_OK;
  const int tx = b->tuning.tile_rays_x, ty = 256 / b->tuning.lanes_per_ray / tx;
  std::vector<int> order(b->n);
  std::iota(order.begin(), order.end(), 0);
  if (b->tuning.order == 0) {
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return steps[a] > steps[c]; });
  }
  // each view's tiles in strip order, tile b to list b % 8 (the XCD that
  // claims list x runs column strip x of every view)
  std::vector<BatchItem> lists[8];
  for (int v : order) {
    const int tiles_x = (b->w[v] + tx - 1) / tx, tiles_y = (b->h[v] + ty - 1) / ty;
    // one strip per XCD: the lists balance the XCDs at the end (a workgroup
    // whose list is empty takes from the others), so the widest strips,
    // the fewest texels per L2, win.  Measured on MI355X (96-view set,
    // precision; strips of a 1080p view, 120 tiles wide): 8 strips 16.90 ms,
    // 16 16.97, 24 16.96, 48 (the frame march's rule) 17.31, 96 17.59.
    StripOrder so;
    so.nstrip = 8;
    so.k = 8 / std::gcd(tiles_x, 8);
    for (int t = 0; t < tiles_x * tiles_y; ++t) {
      int sx, sy;
      strip_tile(t, tiles_x, tiles_y, so.nstrip, so.k, sx, sy);
      lists[t & 7].push_back(BatchItem{v, sy * tiles_x + sx});
    }
  }
  std::vector<BatchItem> items;
  for (int x = 0; x < 8; ++x) {
    b->queues.off[x] = (int)items.size();
    b->queues.len[x] = (int)lists[x].size();
    items.insert(items.end(), lists[x].begin(), lists[x].end());
  }
  if (items.size() > b->items_cap) {
    (void)hipFree(b->d_items);
    b->d_items = nullptr;
    HIP_OK(hipMalloc((void**)&b->d_items, items.size() * sizeof(BatchItem)), "batch: items");
    b->items_cap = items.size();
  }
  HIP_OK(hipStreamSynchronize(b->stream), "batch: sync");
  HIP_OK(hipMemcpy(b->d_items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice), "batch: items");
  b->n_items = (int)items.size();
  b->item_steps = steps;
  b->item_tuning = b->tuning;
  return MTC_OK;
}

int capture(mtc_batch* b, int cov) {
  if (b->exec && same_tuning(b->graph_tuning, b->tuning) && b->graph_cov == cov && b->graph_items == b->n_items)
    return MTC_OK;
  drop_graph(b);
  mtc_ctx* c = b->ctx;
  const int variant = c->desc.kernel == MTC_KERNEL_PRECISION ? 1 : 0;
  HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "batch: begin capture");
  hipError_t e = launch_march_batch(b->d_views, b->d_items, b->queues, b->d_queue, c->q_lf, c->q_hf, c->q_curl,
                                    c->q_weather, volume_views(c), b->color, b->mask, cov, variant,
                                    b->tuning.lanes_per_ray, b->tuning.tile_rays_x,
                                    b->tuning.blocks_per_cu * b->cus, b->stream);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return mtc_fail(MTC_ERR_HIP, std::string("batch: capture failed: ") + hipGetErrorName(e != hipSuccess ? e : e2));
  }
  b->graph = g;
  HIP_OK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "batch: instantiate");
  b->graph_tuning = b->tuning;
  b->graph_cov = cov;
  b->graph_items = b->n_items;
  return MTC_OK;
}

int check_batch(mtc_batch* b, const char* fn) {
  if (!b) return mtc_fail(MTC_ERR_INVALID_ARG, std::string(fn) + ": null batch");
  if (hipSetDevice(b->ctx->desc.device) != hipSuccess) return mtc_fail(MTC_ERR_HIP, std::string(fn) + ": hipSetDevice");
  return MTC_OK;
}

}  // namespace

extern "C" {

int mtc_batch_create(mtc_ctx* ctx, int32_t n_views, const int32_t* widths, const int32_t* heights, mtc_batch** out) {
  if (!ctx || !widths || !heights || !out) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_batch_create: null pointer");
  *out = nullptr;
  if (ctx->desc.kernel != MTC_KERNEL_FAST && ctx->desc.kernel != MTC_KERNEL_PRECISION)
    return mtc_fail(MTC_ERR_UNSUPPORTED, "mtc_batch_create: needs a MTC_KERNEL_FAST or MTC_KERNEL_PRECISION context");
  if (n_views < 1 || n_views > 4096) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_batch_create: n_views must be in 1..4096");
  for (int v = 0; v < n_views; ++v)
    if (widths[v] < 1 || heights[v] < 1 || widths[v] > 16384 || heights[v] > 16384)
      return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_batch_create: view sizes must be in 1..16384");
  if (hipSetDevice(ctx->desc.device) != hipSuccess) return mtc_fail(MTC_ERR_HIP, "mtc_batch_create: hipSetDevice");
  mtc_batch* b = new mtc_batch;
  b->ctx = ctx;
  b->n = n_views;
  b->w.assign(widths, widths + n_views);
  b->h.assign(heights, heights + n_views);
  b->offset.resize(n_views);
  size_t off = 0;
  for (int v = 0; v < n_views; ++v) {
    b->offset[v] = off;
    off += (size_t)widths[v] * heights[v];
  }
  b->total_px = off;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->desc.device) == hipSuccess && prop.multiProcessorCount > 0)
    b->cus = prop.multiProcessorCount;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&b->color, off * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc((void**)&b->mask, off * sizeof(float4));
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_views, (size_t)n_views * sizeof(BatchView));
  if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_views, (size_t)n_views * sizeof(BatchView), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_queue, 8 * kBatchQueueStride * sizeof(unsigned int));
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e != hipSuccess) {
    std::string msg = std::string("mtc_batch_create: ") + hipGetErrorName(e);
    int code = e == hipErrorOutOfMemory ? MTC_ERR_OUT_OF_MEMORY : MTC_ERR_HIP;
    mtc_batch_destroy(b);
    return mtc_fail(code, msg);
  }
  *out = b;
  return MTC_OK;
}

int mtc_batch_destroy(mtc_batch* b) {
  if (!b) return MTC_OK;
  (void)hipSetDevice(b->ctx->desc.device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  (void)hipFree(b->color);
  (void)hipFree(b->mask);
  (void)hipFree(b->d_views);
  (void)hipFree(b->d_items);
  (void)hipFree(b->d_queue);
  if (b->h_views) (void)hipHostFree(b->h_views);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MTC_OK;
}

int mtc_batch_set_tuning(mtc_batch* b, const mtc_batch_tuning* t) {
  int rc = check_batch(b, "mtc_batch_set_tuning");
  if (rc) return rc;
  if (!t || !valid_tuning(*t))
    return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_batch_set_tuning: lanes 4|8, tile width dividing 256/lanes, 1..16 blocks/CU, order 0|1");
  b->tuning = *t;
  return MTC_OK;
}

int mtc_batch_get_tuning(mtc_batch* b, mtc_batch_tuning* t) {
  if (!b || !t) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_batch_get_tuning: null pointer");
  *t = b->tuning;
  return MTC_OK;
}

int mtc_batch_render(mtc_batch* b, const mtc_camera* cams, const mtc_time* times, const mtc_sun* suns,
                     const mtc_cloud_params* params, void* stream) {
  int rc = check_batch(b, "mtc_batch_render");
  if (rc) return rc;
  if (!cams || !times || !suns || !params) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_batch_render: null uniform array");
  if (!b->ctx->volumes_ready) return mtc_fail(MTC_ERR_STATE, "mtc_batch_render: context has no volumes");
  const int cov = params[0].coverage_mode;
  for (int v = 1; v < b->n; ++v)
    if (params[v].coverage_mode != cov) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_batch_render: views differ in coverage_mode");
  hipStream_t s = stream ? (hipStream_t)stream : b->stream;
  // the staging buffer may still feed the previous render's copy
  HIP_OK(hipStreamSynchronize(b->stream), "mtc_batch_render: sync");
  if (s != b->stream) HIP_OK(hipStreamSynchronize(s), "mtc_batch_render: sync");
  rc = build_items(b, params);
  if (rc) return rc;
  const int tx = b->tuning.tile_rays_x;
  for (int v = 0; v < b->n; ++v) {
    rc = build_frame_consts(&cams[v], nullptr, &times[v], &suns[v], &params[v], b->w[v], b->h[v], &b->h_views[v].fc);
    if (rc) return rc;
    b->h_views[v].offset = b->offset[v];
    b->h_views[v].tiles_x = (b->w[v] + tx - 1) / tx;
    b->h_views[v].pad = 0;
  }
  rc = capture(b, cov);
  if (rc) return rc;
  HIP_OK(hipMemcpyAsync(b->d_views, b->h_views, (size_t)b->n * sizeof(BatchView), hipMemcpyHostToDevice, s),
         "mtc_batch_render: uniforms");
  HIP_OK(hipEventRecord(b->ev0, s), "mtc_batch_render: event");
  HIP_OK(hipGraphLaunch(b->exec, s), "mtc_batch_render: graph launch");
  HIP_OK(hipEventRecord(b->ev1, s), "mtc_batch_render: event");
  b->rendered = true;
  return MTC_OK;
}

int mtc_batch_last_ms(mtc_batch* b, double* ms) {
  int rc = check_batch(b, "mtc_batch_last_ms");
  if (rc) return rc;
  if (!ms) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_batch_last_ms: null pointer");
  if (!b->rendered) return mtc_fail(MTC_ERR_STATE, "mtc_batch_last_ms: nothing rendered yet");
  HIP_OK(hipEventSynchronize(b->ev1), "mtc_batch_last_ms: sync");
  float f = 0.0f;
  HIP_OK(hipEventElapsedTime(&f, b->ev0, b->ev1), "mtc_batch_last_ms: elapsed");
  *ms = f;
  return MTC_OK;
}

int mtc_batch_autotune(mtc_batch* b, const mtc_camera* cams, const mtc_time* times, const mtc_sun* suns,
                       const mtc_cloud_params* params, mtc_batch_tuning* best, double* best_ms) {
  int rc = check_batch(b, "mtc_batch_autotune");
  if (rc) return rc;
  std::vector<mtc_batch_tuning> cands;
  for (int lanes : {8, 4})
    for (int tx : {8, 16, 32})
      for (int bpc : {2, 4, 8})
        for (int order : {0, 1}) {
          mtc_batch_tuning t = {lanes, tx, bpc, order};
          if (valid_tuning(t)) cands.push_back(t);
        }
  const mtc_batch_tuning keep = b->tuning;
  double best_t = 1e30;
  mtc_batch_tuning best_c = keep;
  for (const auto& t : cands) {
    b->tuning = t;
    rc = mtc_batch_render(b, cams, times, suns, params, nullptr);  // warm-up (and capture)
    if (rc) {
      b->tuning = keep;
      return rc;
    }
    double ts[3];
    for (double& x : ts) {
      rc = mtc_batch_render(b, cams, times, suns, params, nullptr);
      if (!rc) rc = mtc_batch_last_ms(b, &x);
      if (rc) {
        b->tuning = keep;
        return rc;
      }
    }
    std::sort(ts, ts + 3);
    if (ts[1] < best_t) {
      best_t = ts[1];
      best_c = t;
    }
  }
  b->tuning = best_c;
  if (best) *best = best_c;
  if (best_ms) *best_ms = best_t;
  return MTC_OK;
}

int mtc_batch_read(mtc_batch* b, int32_t view, float* rgba, float* mask, size_t n_floats) {
  int rc = check_batch(b, "mtc_batch_read");
  if (rc) return rc;
  if (view < 0 || view >= b->n) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_batch_read: view out of range");
  const size_t px = (size_t)b->w[view] * b->h[view];
  if (n_floats < px * 4) return mtc_fail(MTC_ERR_INVALID_ARG, "mtc_batch_read: buffer smaller than W*H*4 floats");
  HIP_OK(hipStreamSynchronize(b->stream), "mtc_batch_read: sync");
  HIP_OK(hipDeviceSynchronize(), "mtc_batch_read: sync");
  if (rgba) HIP_OK(hipMemcpy(rgba, b->color + b->offset[view], px * sizeof(float4), hipMemcpyDeviceToHost), "mtc_batch_read");
  if (mask) HIP_OK(hipMemcpy(mask, b->mask + b->offset[view], px * sizeof(float4), hipMemcpyDeviceToHost), "mtc_batch_read");
  return MTC_OK;
}

}  // extern "C"

// Batched rendering: many scenes through one set of wavefront launches
// (declared in include/rt_api.h).
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <new>
#include <unordered_set>
#include <vector>

#include "rt_internal.h"

#define RT_HIP(ctx, call)                                                                       \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return rt_set_error((ctx), RT_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

namespace {
constexpr int kBuildStreams = 8;
}

struct rt_batch {
  rt_ctx* ctx = nullptr;
  std::vector<rt_batch_item> items;
  std::vector<int> builds;  // the first item of each distinct BVH: rebuilds build each BVH once
  void* ws = nullptr;
  void* shade = nullptr;  // the scenes' shading tables, back to back
  size_t ws_bytes = 0, shade_bytes = 0;
  size_t npix = 0, ntiles = 0, nscenes = 0;
  int max_spp = 0, max_depth = 0;
  std::vector<uint32_t> live_tiles;  // per sample k: tiles of the scenes with more than k samples (a prefix)
  rt_tuning tuning;
  hipStream_t streams[kBuildStreams] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t joins[kBuildStreams] = {};
  rt_seg_build* seg = nullptr;  // one launch per build stage for all scenes (nullptr: per-scene builds)
};

extern "C" {

void rt_batch_destroy(rt_batch* b) {
  if (!b) return;
  for (int k = 0; k < kBuildStreams; ++k) {
    if (b->joins[k]) (void)hipEventDestroy(b->joins[k]);
    if (b->streams[k]) (void)hipStreamDestroy(b->streams[k]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  rt_seg_build_destroy(b->ctx, b->seg);
  rt_dev_free(b->ctx, b->ws, b->ws_bytes);
  rt_dev_free(b->ctx, b->shade, b->shade_bytes);
  delete b;
}

rt_status rt_batch_create(rt_ctx* ctx, const rt_batch_item* items, int32_t count, rt_batch** out) {
  if (!ctx || !out) return RT_ERR_INVALID;
  *out = nullptr;
  if (!items || count <= 0 || count > 65535) return rt_set_error(ctx, RT_ERR_INVALID, "need 1..65535 batch items");
  for (int i = 0; i < count; ++i) {
    const rt_batch_item& it = items[i];
    if (!it.bvh || !it.mats_dev || it.nmat <= 0 || !it.rgb_dev || !it.mesh.nx_dev || !it.mesh.ny_dev ||
        !it.mesh.nz_dev || !it.mesh.idx_dev || it.mesh.n != rt_bvh_num_triangles(it.bvh) || it.cam.width <= 0 ||
        it.cam.height <= 0 || it.opts.spp <= 0 || it.opts.depth < 0)
      return rt_set_error(ctx, RT_ERR_INVALID, "invalid batch item %d", i);
  }
  rt_batch* b = new (std::nothrow) rt_batch();
  if (!b) return rt_set_error(ctx, RT_ERR_OOM, "host allocation failed");
  b->ctx = ctx;
  b->items.assign(items, items + count);
  b->nscenes = (size_t)count;
  b->tuning = rt_default_tuning();
  std::vector<size_t> pix_off(count), tile_off(count), shade_off(count), tri_off(count);
  size_t shade_bytes = 0, ntris = 0;
  for (int i = 0; i < count; ++i) {
    const rt_camera& c = items[i].cam;
    tri_off[i] = ntris;
    ntris += (size_t)items[i].mesh.n;
    shade_off[i] = shade_bytes;
    shade_bytes += rt_align256(rt_shade_table_bytes((size_t)items[i].mesh.n));
    pix_off[i] = b->npix;
    b->npix += (size_t)c.width * c.height;
    b->max_spp = items[i].opts.spp > b->max_spp ? items[i].opts.spp : b->max_spp;
    b->max_depth = items[i].opts.depth > b->max_depth ? items[i].opts.depth : b->max_depth;
  }
  // Tiles go to the scenes in order of decreasing sample count (stable), so
  // the tiles that take sample k are a prefix: later samples launch their
  // camera and first shading passes over that prefix only.
  std::vector<int> by_spp(count);
  for (int i = 0; i < count; ++i) by_spp[i] = i;
  std::stable_sort(by_spp.begin(), by_spp.end(),
                   [&](int x, int y) { return items[x].opts.spp > items[y].opts.spp; });
  for (int i : by_spp) {
    const rt_camera& c = items[i].cam;
    tile_off[i] = b->ntiles;
    b->ntiles += (size_t)((c.width + 15) / 16) * ((c.height + 15) / 16);
  }
  b->live_tiles.assign(b->max_spp, 0u);
  for (int i = 0; i < count; ++i) {
    const rt_camera& c = items[i].cam;
    const uint32_t nt = (uint32_t)(((c.width + 15) / 16) * ((c.height + 15) / 16));
    for (int k = 0; k < items[i].opts.spp; ++k) b->live_tiles[k] += nt;
  }
  if (b->npix >= (1ull << 32) || b->ntiles * 256 >= (1ull << 32) || ntris >= (1ull << 32)) {
    rt_batch_destroy(b);
    return rt_set_error(ctx, RT_ERR_INVALID, "batch has too many pixels or triangles");
  }
  size_t bytes = rt_wavefront_bytes_batch(b->npix, b->ntiles, b->nscenes);
  hipError_t ae = rt_dev_alloc(ctx, &b->ws, bytes);
  if (ae == hipSuccess) b->ws_bytes = bytes;
  if (ae == hipSuccess) ae = rt_dev_alloc(ctx, &b->shade, shade_bytes);
  if (ae == hipSuccess) b->shade_bytes = shade_bytes;
  if (ae != hipSuccess) {
    rt_batch_destroy(b);
    return rt_set_error(ctx, RT_ERR_OOM, "batch workspace of %zu + %zu bytes", bytes, shade_bytes);
  }
  // Scene table, BVH views and tile -> scene map: the first three regions.
  const size_t sd = rt_scene_desc_bytes(), vb = rt_bvh_view_bytes();
  std::vector<char> table(sd * count), views(vb * count);
  std::vector<uint16_t> tile_scene(b->ntiles);
  for (int i = 0; i < count; ++i) {
    const rt_batch_item& it = items[i];
    rt_fill_scene(table.data() + sd * i, it.bvh, &it.mesh, it.mats_dev, it.nmat, &it.light, &it.cam, &it.opts,
                  it.rgb_dev, (uint32_t)pix_off[i], (uint32_t)tile_off[i], (uint32_t)tri_off[i],
                  (char*)b->shade + shade_off[i]);
    BvhView v = it.bvh->view();
    ::memcpy(views.data() + vb * i, &v, sizeof(v));
    size_t nt = (size_t)((it.cam.width + 15) / 16) * ((it.cam.height + 15) / 16);
    for (size_t t = 0; t < nt; ++t) tile_scene[tile_off[i] + t] = (uint16_t)i;
  }
  char* w = (char*)b->ws;
  size_t o0 = 0, o1 = rt_align256(sd * count), o2 = o1 + rt_align256(vb * count);
  hipError_t e = hipMemcpy(w + o0, table.data(), table.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(w + o1, views.data(), views.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(w + o2, tile_scene.data(), tile_scene.size() * 2, hipMemcpyHostToDevice);
  // The overflow counter accumulates over the batch's lifetime.
  if (e == hipSuccess)
    e = hipMemset(rt_wavefront_counts(b->ws, b->npix, b->ntiles, b->nscenes) + 8, 0, sizeof(uint32_t));
  for (int k = 0; e == hipSuccess && k < kBuildStreams; ++k) {
    e = hipStreamCreateWithFlags(&b->streams[k], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->joins[k], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  std::unordered_set<const rt_bvh*> seen;
  for (int i = 0; i < count; ++i)
    if (seen.insert(items[i].bvh).second) b->builds.push_back(i);
  std::vector<rt_batch_item> uniq;
  for (int k : b->builds) uniq.push_back(items[k]);
  if (e == hipSuccess) e = rt_seg_build_create(ctx, uniq.data(), (int)uniq.size(), &b->seg);
  (void)rt_ctx_overlap(ctx);  // created here, outside any graph capture of rt_batch_render
  if (e != hipSuccess) {
    rt_batch_destroy(b);
    return rt_set_error(ctx, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_DEVICE, "batch setup failed: %s",
                        hipGetErrorString(e));
  }
  *out = b;
  return RT_OK;
}

rt_status rt_batch_rebuild(rt_batch* b, void* stream) {
  if (!b) return RT_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  for (int k : b->builds) RT_HIP(b->ctx, rt_note_capture(b->items[k].bvh, s));
  // Every stage in one launch over all scenes, unless a scene also keeps
  // the nodes of layout 2, 4 or 8 (the segmented build fills layout 5's).
  bool seg = b->seg != nullptr;
  for (const rt_batch_item& it : b->items)
    seg = seg && !it.bvh->want_tnode && !it.bvh->want_wide4 && !it.bvh->want_wide8;
  for (int k : b->builds) const_cast<rt_bvh*>(b->items[k].bvh)->stage = kStageBuilt;
  if (seg) {
    RT_HIP(b->ctx, rt_seg_build_launch(b->seg, s));
    return RT_OK;
  }
  RT_HIP(b->ctx, hipEventRecord(b->fork, s));
  const size_t nb = b->builds.size();
  int used = (int)(nb < (size_t)kBuildStreams ? nb : (size_t)kBuildStreams);
  for (int k = 0; k < used; ++k) RT_HIP(b->ctx, hipStreamWaitEvent(b->streams[k], b->fork, 0));
  for (size_t j = 0; j < nb; ++j) {
    rt_batch_item& it = b->items[b->builds[j]];
    RT_HIP(b->ctx, rt_launch_build(const_cast<rt_bvh*>(it.bvh), &it.mesh, b->streams[j % kBuildStreams]));
  }
  for (int k = 0; k < used; ++k) {
    RT_HIP(b->ctx, hipEventRecord(b->joins[k], b->streams[k]));
    RT_HIP(b->ctx, hipStreamWaitEvent(s, b->joins[k], 0));
  }
  return RT_OK;
}

rt_status rt_batch_render(rt_batch* b, uint64_t* counters_dev, void* stream) {
  if (!b) return RT_ERR_INVALID;
  for (const rt_batch_item& it : b->items) {
    rt_status st = rt_bvh_require_built(it.bvh, "rt_batch_render");
    if (st != RT_OK) return st;
  }
  for (const rt_batch_item& it : b->items)
    RT_HIP(b->ctx, rt_ensure_layout(it.bvh, b->tuning.node_layout, (hipStream_t)stream));
  RT_HIP(b->ctx, rt_launch_wavefront(b->ws, b->npix, b->ntiles, b->nscenes, b->max_spp, b->max_depth,
                                     b->ctx->num_cus, &b->tuning, counters_dev, nullptr, (hipStream_t)stream,
                                     nullptr, nullptr, nullptr, &b->ctx->timer, rt_ctx_overlap(b->ctx),
                                     b->live_tiles.data()));
  return RT_OK;
}

rt_status rt_batch_set_tuning(rt_batch* b, const rt_tuning* t) {
  if (!b) return RT_ERR_INVALID;
  if (!rt_tuning_valid(t)) return rt_set_error(b->ctx, RT_ERR_INVALID, "invalid tuning");
  b->tuning = *t;
  for (const rt_batch_item& it : b->items) rt_keep_only_layout(it.bvh, t->node_layout);
  return RT_OK;
}

rt_status rt_batch_get_tuning(const rt_batch* b, rt_tuning* t) {
  if (!b || !t) return RT_ERR_INVALID;
  *t = b->tuning;
  return RT_OK;
}

rt_status rt_batch_autotune(rt_batch* b, void* stream, rt_tuning* best, float* best_ms) {
  if (!b) return RT_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  RT_HIP(b->ctx, hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone)
    return rt_set_error(b->ctx, RT_ERR_INVALID, "rt_batch_autotune times renders: not under graph capture");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  RT_HIP(b->ctx, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return rt_set_error(b->ctx, RT_ERR_DEVICE, "rt_batch_autotune: event creation failed");
  }
  const rt_tuning keep = b->tuning;
  // one untimed render (fills the layout), then the faster of two timed ones
  auto time = [&](const rt_tuning& t, float* ms) -> rt_status {
    b->tuning = t;
    rt_status r = rt_batch_render(b, nullptr, stream);
    *ms = -1.f;
    for (int rep = 0; r == RT_OK && rep < 2; ++rep) {
      float m = 0.f;
      if (hipEventRecord(e0, s) != hipSuccess) return rt_set_error(b->ctx, RT_ERR_DEVICE, "autotune timing failed");
      r = rt_batch_render(b, nullptr, stream);
      if (r == RT_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(&m, e0, e1) != hipSuccess))
        return rt_set_error(b->ctx, RT_ERR_DEVICE, "autotune timing failed");
      if (r == RT_OK && (*ms < 0.f || m < *ms)) *ms = m;
    }
    return r;
  };
  rt_tuning win = keep;
  float win_ms = 0.f;
  rt_status st = rt_autotune_descent(keep, time, &win, &win_ms);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->tuning = st == RT_OK ? win : keep;
  // the trials filled every layout: rebuilds keep only the one in use
  for (const rt_batch_item& it : b->items) rt_keep_only_layout(it.bvh, b->tuning.node_layout);
  if (st == RT_OK) {
    if (best) *best = win;
    if (best_ms) *best_ms = win_ms;
  }
  return st;
}

rt_status rt_batch_diagnostics(const rt_batch* b, void* stream, uint32_t out[1]) {
  if (!b || !out) return RT_ERR_INVALID;
  uint32_t* c = rt_wavefront_counts(b->ws, b->npix, b->ntiles, b->nscenes);
  RT_HIP(b->ctx, hipMemcpyAsync(out, c + 8, sizeof(uint32_t), hipMemcpyDeviceToHost, (hipStream_t)stream));
  RT_HIP(b->ctx, hipStreamSynchronize((hipStream_t)stream));
  return RT_OK;
}

}  // extern "C"

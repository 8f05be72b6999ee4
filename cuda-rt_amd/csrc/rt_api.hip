// C-ABI entry points of librtamd (declared in include/rt_api.h).
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>

#include "rt_internal.h"
#include "rt_math.h"

rt_status rt_set_error(rt_ctx* ctx, rt_status s, const char* fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return s;
}

#define RT_HIP(ctx, call)                                                                           \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess)                                                                           \
      return rt_set_error((ctx), RT_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_));     \
  } while (0)

namespace {

size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

bool mesh_ok(const rt_mesh_view* m) {
  return m && m->vx_dev && m->vy_dev && m->vz_dev && m->idx_dev && m->n > 0 && m->nv > 0;
}

// Scratch of rt_render_noalloc: the wavefront workspace followed by the
// shading table. Each GPU the library runs on holds its own copy (48 MiB)
// until the process ends. Host code never passes the array's name to a
// kernel: the name is a host-side stand-in, and the device address comes
// from hipGetSymbolAddress, once per device (noalloc_scratch_address).
constexpr size_t kNoAllocScratchBytes = size_t(48) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices];

// The device address of `scratch` on `device`, looked up on the first call
// for that device with the device current (the caller's current device is
// restored).
hipError_t noalloc_scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    int prev = -1;
    hipError_t e = hipGetDevice(&prev);
    if (e == hipSuccess) e = hipSetDevice(device);
    if (e != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return e;
    g_scratch_addr[device] = p;
  }
  *out = g_scratch_addr[device];
  return hipSuccess;
}

}  // namespace

static void destroy_overlap(rt_overlap& o) {
  for (hipEvent_t e : {o.start, o.filled, o.shaded, o.done})
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : o.flushed)
    if (e) (void)hipEventDestroy(e);
  if (o.aux) (void)hipStreamDestroy(o.aux);
  o = rt_overlap{};
}

const rt_overlap* rt_ctx_overlap(rt_ctx* ctx) {
  if (ctx->overlap_ready) return ctx->overlap.aux ? &ctx->overlap : nullptr;
  ctx->overlap_ready = true;
  rt_overlap& o = ctx->overlap;
  bool ok = hipStreamCreateWithFlags(&o.aux, hipStreamNonBlocking) == hipSuccess;
  for (hipEvent_t* e : {&o.start, &o.filled, &o.shaded, &o.done})
    ok = ok && hipEventCreateWithFlags(e, hipEventDisableTiming) == hipSuccess;
  for (hipEvent_t& e : o.flushed) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    destroy_overlap(o);
    return nullptr;
  }
  return &o;
}

hipError_t rt_dev_alloc(rt_ctx* ctx, void** p, size_t bytes) {
  *p = nullptr;
  if (ctx && ctx->fail_alloc_in > 0 && --ctx->fail_alloc_in == 0) return hipErrorOutOfMemory;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess || !*p) {
    (void)hipGetLastError();
    *p = nullptr;
    return e != hipSuccess ? e : hipErrorOutOfMemory;
  }
  if (ctx) ctx->dev_bytes += bytes;
  return hipSuccess;
}

void rt_dev_free(rt_ctx* ctx, void* p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  if (ctx) ctx->dev_bytes -= bytes;
}

rt_status rt_bvh_require_built(const rt_bvh* bvh, const char* who) {
  if (bvh->stage == kStageBuilt) return RT_OK;
  return rt_set_error(bvh->ctx, RT_ERR_INVALID, "%s: the BVH is not built (run rt_bvh_build_tree or rt_bvh_rebuild)",
                      who);
}

extern "C" {

const char* rt_version(void) { return "rtamd 0.1 (gfx950)"; }

rt_status rt_ctx_create(int32_t device, rt_ctx** out) {
  if (!out) return RT_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return RT_ERR_DEVICE;
  rt_ctx* c = new (std::nothrow) rt_ctx();
  if (!c) return RT_ERR_OOM;
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->num_cus = prop.multiProcessorCount;
    c->arch.assign(prop.gcnArchName, strnlen(prop.gcnArchName, sizeof(prop.gcnArchName)));
    c->arch = c->arch.substr(0, c->arch.find(':'));  // gfx950, without the target features
  }
  if (noalloc_scratch_address(device, &c->noalloc_scratch) != hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return RT_ERR_DEVICE;
  }
  *out = c;
  return RT_OK;
}

void rt_ctx_destroy(rt_ctx* ctx) {
  if (!ctx) return;
  rt_dev_free(ctx, ctx->workspace, ctx->workspace_bytes);
  rt_dev_free(ctx, ctx->shade, ctx->shade_bytes);
  rt_dev_free(ctx, ctx->diff, ctx->diff_bytes);
  for (auto& r : ctx->timer.recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : ctx->timer.pool) (void)hipEventDestroy(e);
  rt_dev_free(ctx, ctx->timer.dev_work, 6 * sizeof(unsigned long long));
  destroy_overlap(ctx->overlap);
  delete ctx;
}

const char* rt_last_error(const rt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }


rt_status rt_ctx_fail_alloc(rt_ctx* ctx, int32_t k) {
  if (!ctx || k < 0) return RT_ERR_INVALID;
  ctx->fail_alloc_in = k;
  return RT_OK;
}

rt_status rt_ctx_device_bytes(const rt_ctx* ctx, uint64_t* bytes) {
  if (!ctx || !bytes) return RT_ERR_INVALID;
  *bytes = ctx->dev_bytes;
  return RT_OK;
}

rt_status rt_ctx_set_kernel_timing(rt_ctx* ctx, int32_t enable) {
  if (!ctx) return RT_ERR_INVALID;
  if (enable && !ctx->timer.dev_work) {
    void* p = nullptr;
    if (rt_dev_alloc(ctx, &p, 6 * sizeof(unsigned long long)) != hipSuccess)
      return rt_set_error(ctx, RT_ERR_OOM, "kernel timing counters");
    RT_HIP(ctx, hipMemset(p, 0, 6 * sizeof(unsigned long long)));
    ctx->timer.dev_work = (unsigned long long*)p;
  }
  ctx->timer.enabled = enable != 0;
  return RT_OK;
}

rt_status rt_ctx_kernel_stats(rt_ctx* ctx, rt_kernel_stat out[2]) {
  if (!ctx || !out) return RT_ERR_INVALID;
  RT_HIP(ctx, hipDeviceSynchronize());
  for (int k = 0; k < 2; ++k) out[k] = rt_kernel_stat{0, 0.0, 0, 0, 0};
  rt_kernel_timer& T = ctx->timer;
  for (auto& r : T.recs) {
    float ms = 0.f;
    RT_HIP(ctx, hipEventElapsedTime(&ms, r.a, r.b));
    out[r.kind].launches += 1;
    out[r.kind].ms += ms;
    T.pool.push_back(r.a);
    T.pool.push_back(r.b);
  }
  T.recs.clear();
  if (T.dev_work) {
    unsigned long long w[6];
    RT_HIP(ctx, hipMemcpy(w, T.dev_work, sizeof(w), hipMemcpyDeviceToHost));
    RT_HIP(ctx, hipMemset(T.dev_work, 0, sizeof(w)));
    out[0].box_tests = w[0];
    out[0].tri_tests = w[1];
    out[1].box_tests = w[2];
    out[1].tri_tests = w[3];
    out[0].rays = w[4];
    out[1].rays = w[5];
  }
  return RT_OK;
}

static rt_status alloc_bvh(rt_ctx* ctx, int32_t n, int32_t bits, int32_t flags, rt_bvh** out) {
  rt_bvh* b = new (std::nothrow) rt_bvh();
  if (!b) return rt_set_error(ctx, RT_ERR_OOM, "host allocation failed");
  b->ctx = ctx;
  b->n = n;
  b->bits = bits;
  b->build_flags = flags;
  b->passes = bits == 30 ? 4 : 8;
  b->ntiles = (n + kSortTile - 1) / kSortTile;
  const size_t ks = bits == 30 ? 4 : 8;
  const size_t ni = (size_t)(n > 1 ? n - 1 : 1);
  const size_t nn = 2 * (size_t)n - 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += align_up(bytes);
    return o;
  };
  size_t o_tri = take((size_t)n * 6 * 4);
  size_t o_misc = take(kMiscWords * 4);
  size_t o_hist = off;  // hist, status and arrivals are contiguous (cleared together)
  off += (size_t)b->passes * 256 * 4;
  size_t o_status = off;
  off += (size_t)b->passes * b->ntiles * 256 * 4;
  size_t o_arr = off;
  off = align_up(off + ni * 4);
  size_t o_dep = take(ni * 64), o_def = take((size_t)n * 4), o_fetch = take(1024 * 4);
  size_t o_k0 = take((size_t)n * ks), o_k1 = take((size_t)n * ks);
  size_t o_i0 = take((size_t)n * 4), o_i1 = take((size_t)n * 4);
  size_t o_par = take(nn * 4), o_l = take(ni * 4), o_r = take(ni * 4), o_f = take(ni * 4);
  size_t o_nb = take(nn * 6 * 4), o_tn = take(ni * 64), o_lt = take(((size_t)n + 1) * 16 * kLeafStride);
  size_t o_w = take(ni * 128), o_wq = take(ni * 64), o_w8 = take(ni * 128), o_wf = take(ni);
  size_t o_lo = take((size_t)n * 4), o_def2 = take((size_t)n * 4), o_tr = take((size_t)n * 48);
  size_t o_wc = take(ni * 4), o_ws = take(((ni + 1023) / 1024 + 1) * 4), o_wsrc = take(ni * 16);
  void* p = nullptr;
  hipError_t e = rt_dev_alloc(ctx, &p, off);
  if (e != hipSuccess) {
    delete b;
    return rt_set_error(ctx, RT_ERR_OOM, "device allocation of %zu bytes for the BVH failed", off);
  }
  char* a = (char*)p;
  b->arena = p;
  b->arena_bytes = off;
  b->tri_box = (float*)(a + o_tri);
  b->misc = (uint32_t*)(a + o_misc);
  b->hist = (uint32_t*)(a + o_hist);
  b->status = (uint32_t*)(a + o_status);
  b->arrivals = (uint32_t*)(a + o_arr);
  b->deposits = (float4*)(a + o_dep);
  b->deferred = (uint32_t*)(a + o_def);
  b->deferred2 = (uint32_t*)(a + o_def2);
  b->trirec = (float4*)(a + o_tr);
  b->wcomp = (int32_t*)(a + o_wc);
  b->wscan = (uint32_t*)(a + o_ws);
  b->wsrc = (int4*)(a + o_wsrc);
  b->fetch = (uint32_t*)(a + o_fetch);
  b->keys[0] = a + o_k0;
  b->keys[1] = a + o_k1;
  b->ids[0] = (uint32_t*)(a + o_i0);
  b->ids[1] = (uint32_t*)(a + o_i1);
  b->parent = (int32_t*)(a + o_par);
  b->left = (int32_t*)(a + o_l);
  b->right = (int32_t*)(a + o_r);
  b->flags = (uint32_t*)(a + o_f);
  b->node_box = (float*)(a + o_nb);
  b->tnode = (float4*)(a + o_tn);
  b->ltri = (float4*)(a + o_lt);
  b->wide = (float4*)(a + o_w);
  b->wideq = (float4*)(a + o_wq);
  b->wide8 = (float4*)(a + o_w8);
  b->wide_flag = (uint8_t*)(a + o_wf);
  b->leaf_of = (int32_t*)(a + o_lo);
  // The diagnostic counters in misc[] accumulate over the BVH's lifetime;
  // the null leaf (record n, the child of every unused wide-node slot) is a
  // zero triangle that no build or refit writes.
  e = hipMemset(b->misc, 0, kMiscWords * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemset(b->ltri + kLeafStride * (size_t)n, 0, 16 * kLeafStride);
  if (e != hipSuccess) {
    rt_dev_free(ctx, p, off);
    delete b;
    return rt_set_error(ctx, RT_ERR_DEVICE, "hipMemset failed: %s", hipGetErrorString(e));
  }
  *out = b;
  return RT_OK;
}

static rt_status check_bits(rt_ctx* ctx, const rt_build_opts* opts, int* bits, int* flags);

rt_status rt_bvh_build(rt_ctx* ctx, const rt_mesh_view* mesh, const rt_build_opts* opts, void* stream,
                       rt_bvh** out) {
  if (!ctx || !out) return rt_set_error(ctx, RT_ERR_INVALID, "null context or output");
  *out = nullptr;
  if (!mesh_ok(mesh)) return rt_set_error(ctx, RT_ERR_INVALID, "mesh must have n > 0 triangles and positions");
  int bits = 30, flags = 0;
  rt_status st = check_bits(ctx, opts, &bits, &flags);
  if (st != RT_OK) return st;
  if (mesh->n >= (1 << 30)) return rt_set_error(ctx, RT_ERR_INVALID, "n must be < 2^30");
  RT_HIP(ctx, hipSetDevice(ctx->device));
  rt_bvh* b = nullptr;
  st = alloc_bvh(ctx, mesh->n, bits, flags, &b);
  if (st != RT_OK) return st;
  hipError_t e = rt_launch_build(b, mesh, (hipStream_t)stream);
  if (e != hipSuccess) {
    rt_bvh_destroy(b);
    return rt_set_error(ctx, RT_ERR_DEVICE, "build launch failed: %s", hipGetErrorString(e));
  }
  b->stage = kStageBuilt;
  *out = b;
  return RT_OK;
}

static rt_status check_bits(rt_ctx* ctx, const rt_build_opts* opts, int* bits, int* flags) {
  *bits = opts ? opts->morton_bits : 30;
  *flags = opts ? opts->flags : 0;
  if (*bits != 30 && *bits != 63) return rt_set_error(ctx, RT_ERR_INVALID, "morton_bits must be 30 or 63");
  if (*flags & ~RT_BUILD_KARRAS_SEARCH) return rt_set_error(ctx, RT_ERR_INVALID, "unknown build flags");
  return RT_OK;
}

rt_status rt_bvh_create(rt_ctx* ctx, int32_t n, const rt_build_opts* opts, rt_bvh** out) {
  if (!ctx || !out) return rt_set_error(ctx, RT_ERR_INVALID, "null context or output");
  *out = nullptr;
  if (n <= 0 || n >= (1 << 30)) return rt_set_error(ctx, RT_ERR_INVALID, "n must be in [1, 2^30)");
  int bits = 30, flags = 0;
  rt_status st = check_bits(ctx, opts, &bits, &flags);
  if (st != RT_OK) return st;
  RT_HIP(ctx, hipSetDevice(ctx->device));
  return alloc_bvh(ctx, n, bits, flags, out);
}

static rt_status stage_check(rt_bvh* bvh, int32_t need_min, int32_t need_max, const char* who, const char* before) {
  if (bvh->stage < need_min || bvh->stage > need_max)
    return rt_set_error(bvh->ctx, RT_ERR_INVALID, "%s: call %s first", who, before);
  return RT_OK;
}

rt_status rt_bvh_compute_bounds(rt_bvh* bvh, const rt_mesh_view* mesh, float* tri_lo_dev, float* tri_hi_dev,
                                void* stream) {
  if (!bvh) return RT_ERR_INVALID;
  if (!mesh_ok(mesh) || mesh->n != bvh->n)
    return rt_set_error(bvh->ctx, RT_ERR_INVALID, "rt_bvh_compute_bounds needs a mesh of the BVH's triangle count");
  RT_HIP(bvh->ctx, rt_launch_stage_bounds(bvh, mesh, tri_lo_dev, tri_hi_dev, (hipStream_t)stream));
  bvh->stage = kStageBounds;
  return RT_OK;
}

rt_status rt_bvh_compute_morton(rt_bvh* bvh, const double mn[3], const double mx[3], void* stream) {
  if (!bvh) return RT_ERR_INVALID;
  if ((mn == nullptr) != (mx == nullptr))
    return rt_set_error(bvh->ctx, RT_ERR_INVALID, "rt_bvh_compute_morton: give both bounds or neither");
  rt_status st = stage_check(bvh, kStageBounds, kStageBuilt, "rt_bvh_compute_morton", "rt_bvh_compute_bounds");
  if (st != RT_OK) return st;
  RT_HIP(bvh->ctx, rt_launch_stage_morton(bvh, mn, mx, (hipStream_t)stream));
  bvh->stage = kStageMorton;
  return RT_OK;
}

rt_status rt_bvh_sort_morton(rt_bvh* bvh, void* stream) {
  if (!bvh) return RT_ERR_INVALID;
  rt_status st = stage_check(bvh, kStageMorton, kStageMorton, "rt_bvh_sort_morton", "rt_bvh_compute_morton");
  if (st != RT_OK) return st;
  RT_HIP(bvh->ctx, rt_launch_stage_sort(bvh, (hipStream_t)stream));
  bvh->stage = kStageSorted;
  return RT_OK;
}

rt_status rt_bvh_build_tree(rt_bvh* bvh, const rt_mesh_view* mesh, void* stream) {
  if (!bvh) return RT_ERR_INVALID;
  if (!mesh_ok(mesh) || mesh->n != bvh->n)
    return rt_set_error(bvh->ctx, RT_ERR_INVALID, "rt_bvh_build_tree needs a mesh of the BVH's triangle count");
  rt_status st = stage_check(bvh, kStageSorted, kStageSorted, "rt_bvh_build_tree", "rt_bvh_sort_morton");
  if (st != RT_OK) return st;
  RT_HIP(bvh->ctx, rt_launch_stage_tree(bvh, mesh, (hipStream_t)stream));
  bvh->stage = kStageBuilt;
  return RT_OK;
}


rt_status rt_bvh_rebuild(rt_bvh* bvh, const rt_mesh_view* mesh, void* stream) {
  if (!bvh) return RT_ERR_INVALID;
  if (!mesh_ok(mesh) || mesh->n != bvh->n)
    return rt_set_error(bvh->ctx, RT_ERR_INVALID, "rebuild needs a mesh with the same triangle count");
  RT_HIP(bvh->ctx, rt_note_capture(bvh, (hipStream_t)stream));
  RT_HIP(bvh->ctx, rt_launch_build(bvh, mesh, (hipStream_t)stream));
  bvh->stage = kStageBuilt;
  return RT_OK;
}

rt_status rt_bvh_refit(rt_bvh* bvh, const rt_mesh_view* mesh, void* stream) {
  if (!bvh) return RT_ERR_INVALID;
  if (!mesh_ok(mesh) || mesh->n != bvh->n)
    return rt_set_error(bvh->ctx, RT_ERR_INVALID, "refit needs a mesh with the same triangle count");
  rt_status st = rt_bvh_require_built(bvh, "rt_bvh_refit");
  if (st != RT_OK) return st;
  RT_HIP(bvh->ctx, rt_note_capture(bvh, (hipStream_t)stream));
  RT_HIP(bvh->ctx, rt_launch_refit(bvh, mesh, (hipStream_t)stream));
  return RT_OK;
}

rt_status rt_bvh_export(const rt_bvh* bvh, const rt_bvh_export_bufs* dst, void* stream) {
  if (!bvh || !dst) return RT_ERR_INVALID;
  rt_status st = rt_bvh_require_built(bvh, "rt_bvh_export");
  if (st != RT_OK) return st;
  RT_HIP(bvh->ctx, rt_launch_export(bvh, dst, (hipStream_t)stream));
  return RT_OK;
}

rt_status rt_bvh_scene_bounds(const rt_bvh* bvh, void* stream, double mn[3], double mx[3]) {
  if (!bvh || !mn || !mx) return RT_ERR_INVALID;
  if (bvh->stage < kStageBounds)
    return rt_set_error(bvh->ctx, RT_ERR_INVALID, "rt_bvh_scene_bounds: call rt_bvh_compute_bounds first");
  uint32_t h[6 * kBoundsReplicas];
  RT_HIP(bvh->ctx, hipMemcpyAsync(h, bvh->misc + kMiscBounds, sizeof(h), hipMemcpyDeviceToHost, (hipStream_t)stream));
  RT_HIP(bvh->ctx, hipStreamSynchronize((hipStream_t)stream));
  for (int k = 0; k < 3; ++k) {
    uint32_t a = h[k], b = h[3 + k];
    for (int r = 1; r < kBoundsReplicas; ++r) {
      a = a < h[6 * r + k] ? a : h[6 * r + k];
      b = b > h[6 * r + 3 + k] ? b : h[6 * r + 3 + k];
    }
    mn[k] = (double)rt::ord2f(a);
    mx[k] = (double)rt::ord2f(b);
  }
  return RT_OK;
}

int32_t rt_bvh_num_triangles(const rt_bvh* bvh) { return bvh ? bvh->n : 0; }

rt_status rt_bvh_diagnostics(const rt_bvh* bvh, void* stream, uint32_t out[2]) {
  if (!bvh || !out) return RT_ERR_INVALID;
  uint32_t h[kMiscWords];
  RT_HIP(bvh->ctx, hipMemcpyAsync(h, bvh->misc, sizeof(h), hipMemcpyDeviceToHost, (hipStream_t)stream));
  RT_HIP(bvh->ctx, hipStreamSynchronize((hipStream_t)stream));
  out[0] = h[kMiscOverflow];
  out[1] = h[kMiscBuildError];
  return RT_OK;
}

void rt_bvh_destroy(rt_bvh* bvh) {
  if (!bvh) return;
  rt_dev_free(bvh->ctx, bvh->arena, bvh->arena_bytes);
  delete bvh;
}

rt_status rt_trace(const rt_bvh* bvh, const rt_mesh_view* mesh, const rt_rays* rays, int32_t mode, double tmin,
                   const rt_hits* hits, void* stream) {
  if (!bvh || !rays || !hits || !hits->tri_dev) return RT_ERR_INVALID;
  if (rays->n > 0 && (!rays->orig_dev || !rays->dir_dev))
    return rt_set_error(bvh->ctx, RT_ERR_INVALID, "null ray arrays");
  if (mode != RT_TRACE_CLOSEST && mode != RT_TRACE_ANY)
    return rt_set_error(bvh->ctx, RT_ERR_INVALID, "mode must be RT_TRACE_CLOSEST or RT_TRACE_ANY");
  rt_status st = rt_bvh_require_built(bvh, "rt_trace");
  if (st != RT_OK) return st;
  RT_HIP(bvh->ctx, rt_launch_trace(bvh, mesh, rays, mode, tmin, hits, (hipStream_t)stream));
  return RT_OK;
}

static rt_status grow(rt_ctx* ctx, void** buf, size_t* have, size_t need, hipStream_t stream, const char* what) {
  if (*have >= need) return RT_OK;
  // Under hipGraph capture nothing may be (re)allocated: the graph would keep
  // pointers into memory freed here.
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  RT_HIP(ctx, hipStreamIsCapturing(stream, &cap));
  if (cap != hipStreamCaptureStatusNone)
    return rt_set_error(ctx, RT_ERR_INVALID,
                        "%s of %zu bytes needed under graph capture, but the context holds %zu: make the same call "
                        "once before capturing",
                        what, need, *have);
  if (*buf) {
    RT_HIP(ctx, hipStreamSynchronize(stream));
    rt_dev_free(ctx, *buf, *have);
    *buf = nullptr;
    *have = 0;
  }
  if (rt_dev_alloc(ctx, buf, need) != hipSuccess) return rt_set_error(ctx, RT_ERR_OOM, "%s of %zu bytes", what, need);
  *have = need;
  return RT_OK;
}

static rt_status render_check(rt_ctx* ctx, const rt_bvh* bvh, const rt_mesh_view* mesh, const rt_material* mats_dev,
                              int32_t nmat, const rt_light* light, const rt_camera* cam, const rt_render_opts* opts,
                              float* rgb_dev, const char* who) {
  if (!bvh || !mesh_ok(mesh) || !mats_dev || nmat <= 0 || !light || !cam || !opts || !rgb_dev)
    return rt_set_error(ctx, RT_ERR_INVALID, "null argument to %s", who);
  if (!mesh->nx_dev || !mesh->ny_dev || !mesh->nz_dev)
    return rt_set_error(ctx, RT_ERR_INVALID, "%s needs shading normals", who);
  if (cam->width <= 0 || cam->height <= 0 || opts->spp <= 0 || opts->depth < 0)
    return rt_set_error(ctx, RT_ERR_INVALID, "bad image size, spp or depth");
  if (mesh->n != bvh->n) return rt_set_error(ctx, RT_ERR_INVALID, "mesh and BVH triangle counts differ");
  if (opts->flags & ~(RT_RENDER_RANDVEC_RTL | RT_RENDER_CLAMP | RT_RENDER_WAVEFRONT | RT_RENDER_ONE_STREAM))
    return rt_set_error(ctx, RT_ERR_INVALID, "unknown render flags");
  return rt_bvh_require_built(bvh, who);
}

rt_status rt_render(rt_ctx* ctx, const rt_bvh* bvh, const rt_mesh_view* mesh, const rt_material* mats_dev,
                    int32_t nmat, const rt_light* light, const rt_camera* cam, const rt_render_opts* opts,
                    float* rgb_dev, uint64_t* counters_dev, void* stream) {
  if (!ctx) return RT_ERR_INVALID;
  rt_status chk = render_check(ctx, bvh, mesh, mats_dev, nmat, light, cam, opts, rgb_dev, "rt_render");
  if (chk != RT_OK) return chk;
  void* ws = nullptr;
  void* shade = nullptr;
  if (opts->flags & RT_RENDER_WAVEFRONT) {
    hipStream_t s = (hipStream_t)stream;
    rt_status st = grow(ctx, &ctx->workspace, &ctx->workspace_bytes, rt_wavefront_bytes(cam->width, cam->height),
                        s, "wavefront workspace");
    if (st == RT_OK) st = grow(ctx, &ctx->shade, &ctx->shade_bytes, rt_shade_table_bytes(mesh->n), s, "shading table");
    if (st != RT_OK) return st;
    ws = ctx->workspace;
    shade = ctx->shade;
  } else if (counters_dev) {
    // the megakernel counts into 64 replicas of the counters (8 KiB)
    rt_status st = grow(ctx, &ctx->workspace, &ctx->workspace_bytes, 64 * 16 * sizeof(uint64_t),
                        (hipStream_t)stream, "counter replicas");
    if (st != RT_OK) return st;
    ws = ctx->workspace;
  }
  RT_HIP(ctx, rt_launch_render(bvh, mesh, mats_dev, nmat, light, cam, opts, rgb_dev, counters_dev, ws, shade,
                               (hipStream_t)stream, nullptr, nullptr, &ctx->timer,
                               shade && !(opts->flags & RT_RENDER_ONE_STREAM) ? rt_ctx_overlap(ctx) : nullptr));
  return RT_OK;
}

rt_status rt_render_autotune(rt_ctx* ctx, const rt_bvh* bvh, const rt_mesh_view* mesh, const rt_material* mats_dev,
                             int32_t nmat, const rt_light* light, const rt_camera* cam, const rt_render_opts* opts,
                             float* rgb_dev, void* stream, rt_tuning* best, float* best_ms) {
  if (!ctx) return RT_ERR_INVALID;
  rt_status st = render_check(ctx, bvh, mesh, mats_dev, nmat, light, cam, opts, rgb_dev, "rt_render_autotune");
  if (st != RT_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  RT_HIP(ctx, hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone)
    return rt_set_error(ctx, RT_ERR_INVALID, "rt_render_autotune times renders: not under graph capture");
  rt_render_opts o = *opts;
  o.flags |= RT_RENDER_WAVEFRONT;
  const uint32_t key = rt_tuning_key(bvh->n, (int64_t)cam->width * cam->height);
  const auto had = ctx->tuning.find(key);
  const bool had_entry = had != ctx->tuning.end();
  const rt_tuning before = had_entry ? had->second : rt_default_tuning();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  RT_HIP(ctx, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return rt_set_error(ctx, RT_ERR_DEVICE, "rt_render_autotune: event creation failed");
  }
  // one untimed render (fills the layout, warms the caches), then the
  // faster of two timed ones
  auto time = [&](const rt_tuning& t, float* ms) -> rt_status {
    ctx->tuning[key] = t;
    rt_status r = rt_render(ctx, bvh, mesh, mats_dev, nmat, light, cam, &o, rgb_dev, nullptr, stream);
    *ms = -1.f;
    for (int rep = 0; r == RT_OK && rep < 2; ++rep) {
      float m = 0.f;
      if (hipEventRecord(e0, s) != hipSuccess) return rt_set_error(ctx, RT_ERR_DEVICE, "autotune timing failed");
      r = rt_render(ctx, bvh, mesh, mats_dev, nmat, light, cam, &o, rgb_dev, nullptr, stream);
      if (r == RT_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(&m, e0, e1) != hipSuccess))
        return rt_set_error(ctx, RT_ERR_DEVICE, "autotune timing failed");
      if (r == RT_OK && (*ms < 0.f || m < *ms)) *ms = m;
    }
    return r;
  };
  rt_tuning win{};
  float win_ms = 0.f;
  st = rt_autotune_descent(before, time, &win, &win_ms);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != RT_OK) {
    if (had_entry)
      ctx->tuning[key] = before;
    else
      ctx->tuning.erase(key);
    return st;
  }
  ctx->tuning[key] = win;
  // the trials filled every layout: builds and refits keep only the winner's
  rt_keep_only_layout(bvh, win.node_layout);
  if (best) *best = win;
  if (best_ms) *best_ms = win_ms;
  return RT_OK;
}

rt_status rt_ctx_get_tuning(const rt_ctx* ctx, int32_t ntris, int64_t nrays, rt_tuning* out, int32_t* found) {
  if (!ctx || !out) return RT_ERR_INVALID;
  auto it = ctx->tuning.find(rt_tuning_key(ntris, nrays));
  if (found) *found = it != ctx->tuning.end();
  *out = it != ctx->tuning.end() ? it->second : rt_default_tuning();
  return RT_OK;
}

rt_status rt_ctx_set_tuning(rt_ctx* ctx, int32_t ntris, int64_t nrays, const rt_tuning* t) {
  if (!ctx) return RT_ERR_INVALID;
  const uint32_t key = rt_tuning_key(ntris, nrays);
  if (!t) {
    ctx->tuning.erase(key);
    return RT_OK;
  }
  if (!rt_tuning_valid(t)) return rt_set_error(ctx, RT_ERR_INVALID, "invalid tuning");
  ctx->tuning[key] = *t;
  return RT_OK;
}

static const char kTuningHeader[] = "# rtamd tuning table v1: arch log2_tris log2_rays stack_ring fetch_chunk "
                                    "waves_per_simd node_layout wg_threads";

rt_status rt_ctx_save_tuning(const rt_ctx* ctx, const char* path) {
  if (!ctx || !path) return RT_ERR_INVALID;
  FILE* f = fopen(path, "w");
  if (!f) return rt_set_error(const_cast<rt_ctx*>(ctx), RT_ERR_INVALID, "cannot write %s", path);
  bool ok = fprintf(f, "%s\n", kTuningHeader) > 0;
  for (const auto& kv : ctx->tuning) {
    const rt_tuning& t = kv.second;
    ok = ok && fprintf(f, "%s %u %u %d %d %d %d %d\n", ctx->arch.empty() ? "unknown" : ctx->arch.c_str(),
                       kv.first >> 8, kv.first & 255u, t.stack_ring, t.fetch_chunk, t.waves_per_simd,
                       t.node_layout, t.wg_threads ? t.wg_threads : 64) > 0;
  }
  ok = (fclose(f) == 0) && ok;
  return ok ? RT_OK : rt_set_error(const_cast<rt_ctx*>(ctx), RT_ERR_INVALID, "cannot write %s", path);
}

rt_status rt_ctx_load_tuning(rt_ctx* ctx, const char* path) {
  if (!ctx || !path) return RT_ERR_INVALID;
  FILE* f = fopen(path, "r");
  if (!f) return rt_set_error(ctx, RT_ERR_INVALID, "cannot open %s", path);
  std::map<uint32_t, rt_tuning> got;
  char line[512];
  int lineno = 0;
  rt_status st = RT_OK;
  while (st == RT_OK && fgets(line, sizeof(line), f)) {
    ++lineno;
    const char* p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '#' || *p == '\n' || *p == '\r' || *p == '\0') continue;
    char arch[64];
    unsigned a = 0, r = 0;
    rt_tuning t{};
    if (sscanf(p, "%63s %u %u %d %d %d %d %d", arch, &a, &r, &t.stack_ring, &t.fetch_chunk, &t.waves_per_simd,
               &t.node_layout, &t.wg_threads) != 8 ||
        a > 255 || r > 255 || !rt_tuning_valid(&t)) {
      st = rt_set_error(ctx, RT_ERR_INVALID, "%s:%d: not a tuning entry", path, lineno);
      break;
    }
    if (ctx->arch == arch) got[a << 8 | r] = t;  // other architectures' entries are skipped
  }
  fclose(f);
  if (st != RT_OK) return st;
  for (const auto& kv : got) ctx->tuning[kv.first] = kv.second;
  return RT_OK;
}

size_t rt_render_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

rt_status rt_render_noalloc(rt_ctx* ctx, const rt_bvh* bvh, const rt_mesh_view* mesh, const rt_material* mats_dev,
                            int32_t nmat, const rt_light* light, const rt_camera* cam, const rt_render_opts* opts,
                            float* rgb_dev, uint64_t* counters_dev, void* stream) {
  if (!ctx) return RT_ERR_INVALID;
  rt_status chk = render_check(ctx, bvh, mesh, mats_dev, nmat, light, cam, opts, rgb_dev, "rt_render_noalloc");
  if (chk != RT_OK) return chk;
  if (!ctx->noalloc_scratch) return rt_set_error(ctx, RT_ERR_DEVICE, "rt_render_noalloc: no scratch on this device");
  const size_t ws_bytes = rt_align256(rt_wavefront_bytes(cam->width, cam->height));
  const size_t need = ws_bytes + rt_shade_table_bytes((size_t)mesh->n);
  if (need > kNoAllocScratchBytes)
    return rt_set_error(ctx, RT_ERR_INVALID,
                        "rt_render_noalloc: a %dx%d image of %d triangles needs %zu bytes of scratch, more than its "
                        "%zu (use rt_render)",
                        cam->width, cam->height, mesh->n, need, kNoAllocScratchBytes);
  // The wavefront passes on the caller's stream: no auxiliary stream, events
  // or kernel timers, which would be created on first use.
  rt_render_opts o = *opts;
  o.flags |= RT_RENDER_WAVEFRONT | RT_RENDER_ONE_STREAM;
  char* ws = (char*)ctx->noalloc_scratch;
  RT_HIP(ctx, rt_launch_render(bvh, mesh, mats_dev, nmat, light, cam, &o, rgb_dev, counters_dev, ws, ws + ws_bytes,
                               (hipStream_t)stream, nullptr, nullptr, nullptr, nullptr));
  return RT_OK;
}

static rt_status diff_check(rt_ctx* ctx, const rt_bvh* bvh, const rt_mesh_view* mesh, const double* albedo,
                            double light_scale, const rt_light* light, const rt_camera* cam,
                            const rt_render_opts* opts) {
  if (!bvh || !mesh_ok(mesh) || !albedo || !light || !cam || !opts)
    return rt_set_error(ctx, RT_ERR_INVALID, "null argument to the differentiable render");
  if (!mesh->nx_dev || !mesh->ny_dev || !mesh->nz_dev)
    return rt_set_error(ctx, RT_ERR_INVALID, "the differentiable render needs shading normals");
  if (cam->width <= 0 || cam->height <= 0 || opts->spp <= 0 || opts->depth < 0)
    return rt_set_error(ctx, RT_ERR_INVALID, "bad image size, spp or depth");
  if (opts->depth > RT_DIFF_MAX_DEPTH)
    return rt_set_error(ctx, RT_ERR_UNSUPPORTED, "differentiable render depth %d exceeds %d", opts->depth,
                        RT_DIFF_MAX_DEPTH);
  if (mesh->n != bvh->n) return rt_set_error(ctx, RT_ERR_INVALID, "mesh and BVH triangle counts differ");
  if (!(light_scale > 0)) return rt_set_error(ctx, RT_ERR_INVALID, "light_scale must be positive");
  if (opts->flags & ~(RT_RENDER_RANDVEC_RTL | RT_RENDER_CLAMP | RT_RENDER_WAVEFRONT | RT_RENDER_ONE_STREAM))
    return rt_set_error(ctx, RT_ERR_INVALID, "unknown render flags");
  return rt_bvh_require_built(bvh, "the differentiable render");
}

size_t rt_diff_record_bytes(const rt_camera* cam, const rt_render_opts* opts) {
  if (!cam || !opts || cam->width <= 0 || cam->height <= 0 || opts->spp <= 0 || opts->depth < 0) return 0;
  return (size_t)cam->width * cam->height * opts->spp * (opts->depth + 1) * 2 * sizeof(int32_t);
}

rt_status rt_diff_forward(rt_ctx* ctx, const rt_bvh* bvh, const rt_mesh_view* mesh, const double* albedo_dev,
                          double light_scale, const rt_light* light, const rt_camera* cam,
                          const rt_render_opts* opts, double* image_dev, int32_t* records_dev, void* stream) {
  if (!ctx) return RT_ERR_INVALID;
  rt_status st = diff_check(ctx, bvh, mesh, albedo_dev, light_scale, light, cam, opts);
  if (st != RT_OK) return st;
  if (!image_dev || !records_dev) return rt_set_error(ctx, RT_ERR_INVALID, "null image or record buffer");
  hipStream_t s = (hipStream_t)stream;
  const size_t n = (size_t)mesh->n;
  const size_t pix = (size_t)cam->width * cam->height;
  const size_t mats_bytes = (n * sizeof(rt_material) + 255) & ~(size_t)255;
  const size_t ids_bytes = (n * sizeof(int32_t) + 255) & ~(size_t)255;
  st = grow(ctx, &ctx->diff, &ctx->diff_bytes, mats_bytes + ids_bytes + pix * 3 * sizeof(float), s,
            "differentiable render buffers");
  if (st != RT_OK) return st;
  st = grow(ctx, &ctx->workspace, &ctx->workspace_bytes, rt_wavefront_bytes(cam->width, cam->height), s,
            "wavefront workspace");
  if (st == RT_OK) st = grow(ctx, &ctx->shade, &ctx->shade_bytes, rt_shade_table_bytes(mesh->n), s, "shading table");
  if (st != RT_OK) return st;
  rt_material* mats = (rt_material*)ctx->diff;
  int32_t* ids = (int32_t*)((char*)ctx->diff + mats_bytes);
  float* rgb = (float*)((char*)ctx->diff + mats_bytes + ids_bytes);
  RT_HIP(ctx, rt_launch_fill_albedo(mats, ids, albedo_dev, mesh->n, s));
  RT_HIP(ctx, rt_launch_fill_i32(records_dev, rt_diff_record_bytes(cam, opts) / sizeof(int32_t), -1, s));
  rt_mesh_view m = *mesh;
  m.mat_id_dev = ids;
  rt_light L = *light;
  for (int k = 0; k < 3; ++k) L.color[k] *= light_scale;
  rt_render_opts o = *opts;
  o.flags = (o.flags & RT_RENDER_RANDVEC_RTL) | RT_RENDER_WAVEFRONT;
  RT_HIP(ctx, rt_launch_render(bvh, &m, mats, mesh->n, &L, cam, &o, rgb, nullptr, ctx->workspace, ctx->shade, s,
                               records_dev, image_dev, nullptr,
                               (opts->flags & RT_RENDER_ONE_STREAM) ? nullptr : rt_ctx_overlap(ctx)));
  return RT_OK;
}

rt_status rt_diff_backward(rt_ctx* ctx, const rt_bvh* bvh, const rt_mesh_view* mesh, const double* albedo_dev,
                           double light_scale, const rt_light* light, const rt_camera* cam,
                           const rt_render_opts* opts, const int32_t* records_dev, const double* dimage_dev,
                           double* g_albedo_dev, double* g_light_dev, double* g_pos_dev, void* stream) {
  if (!ctx) return RT_ERR_INVALID;
  rt_status st = diff_check(ctx, bvh, mesh, albedo_dev, light_scale, light, cam, opts);
  if (st != RT_OK) return st;
  if (!records_dev || !dimage_dev || !g_albedo_dev || !g_light_dev || !g_pos_dev)
    return rt_set_error(ctx, RT_ERR_INVALID, "null record, image gradient or gradient buffer");
  hipStream_t s = (hipStream_t)stream;
  rt_render_opts o = *opts;
  o.flags &= RT_RENDER_RANDVEC_RTL;
  RT_HIP(ctx, rt_launch_zero_f64(g_albedo_dev, 3 * (size_t)mesh->n, s));
  RT_HIP(ctx, rt_launch_zero_f64(g_light_dev, 1, s));
  RT_HIP(ctx, rt_launch_zero_f64(g_pos_dev, 3 * (size_t)mesh->nv, s));
  RT_HIP(ctx, rt_launch_adjoint(bvh, mesh, albedo_dev, light_scale, light, cam, &o, records_dev, dimage_dev,
                                g_albedo_dev, g_light_dev, g_pos_dev, s));
  return RT_OK;
}

rt_status rt_resolve_supersample(rt_ctx* ctx, const float* big_dev, int32_t width, int32_t height, int32_t ss,
                                 int32_t flags, float* out_dev, void* stream) {
  if (!ctx) return RT_ERR_INVALID;
  if (!big_dev || !out_dev || width <= 0 || height <= 0 || ss <= 0)
    return rt_set_error(ctx, RT_ERR_INVALID, "bad argument to rt_resolve_supersample");
  if (flags & ~RT_RESOLVE_REFERENCE_DIVISOR) return rt_set_error(ctx, RT_ERR_INVALID, "unknown resolve flags");
  RT_HIP(ctx, rt_launch_resolve(big_dev, width, height, ss, (flags & RT_RESOLVE_REFERENCE_DIVISOR) != 0, out_dev,
                                (hipStream_t)stream));
  return RT_OK;
}

}  // extern "C"

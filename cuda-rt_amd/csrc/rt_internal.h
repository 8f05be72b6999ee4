// Internal types shared by the HIP translation units of librtamd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/rt_api.h"

// Per-kernel timing of the wavefront traversal launches, for benchmarks:
// while enabled, HIP events bracket each traversal launch on its stream and
// the work counts are kept per kernel (dev_work: box and triangle tests of
// the closest-hit traversal, then of the any-hit traversal).
struct rt_kernel_timer {
  bool enabled = false;
  struct Rec {
    int kind;  // 0 closest hit, 1 any hit
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  unsigned long long* dev_work = nullptr;
};

// Wavefront renders keep up to this many samples' shadow work in flight
// (sample slots; the renders use fewer, see rt_trace.hip).
constexpr int kMaxSampleSlots = 6;

// Second stream and events of a context's wavefront renders: a sample's
// last any-hit pass runs there while the next samples are traced.
struct rt_overlap {
  hipStream_t aux = nullptr;
  hipEvent_t start = nullptr, filled = nullptr, shaded = nullptr, done = nullptr;
  hipEvent_t flushed[kMaxSampleSlots] = {};
};

struct rt_ctx {
  int32_t device = 0;
  int32_t num_cus = 256;
  std::string err;
  void* workspace = nullptr;     // wavefront render state, grown on demand
  size_t workspace_bytes = 0;
  void* shade = nullptr;         // per-leaf shading table of the last render (64 B per triangle)
  size_t shade_bytes = 0;
  void* diff = nullptr;          // differentiable render: materials, identity ids, float image
  size_t diff_bytes = 0;
  rt_kernel_timer timer;
  rt_overlap overlap;
  bool overlap_ready = false;
  void* noalloc_scratch = nullptr;  // this device's copy of rt_render_noalloc's static scratch array
  std::string arch;                 // the device's architecture name (gfx950), the tuning table's key
  std::map<uint32_t, rt_tuning> tuning;  // single-scene tunings by size bucket (rt_tuning_key)
  uint64_t dev_bytes = 0;           // device memory held by this context's objects (rt_dev_alloc)
  int32_t fail_alloc_in = 0;        // > 0: the fail_alloc_in-th next rt_dev_alloc fails (rt_ctx_fail_alloc)
};

// Every device allocation of the library: hipMalloc, counted in
// ctx->dev_bytes, failing with hipErrorOutOfMemory when the context's
// fault-injection countdown reaches it. rt_dev_free takes the size that was
// allocated.
hipError_t rt_dev_alloc(rt_ctx* ctx, void** p, size_t bytes);
void rt_dev_free(rt_ctx* ctx, void* p, size_t bytes);

// The context's overlap stream and events, created on first use (nullptr if
// that fails: the passes then run on one stream).
const rt_overlap* rt_ctx_overlap(rt_ctx* ctx);

// Device-side view of a built BVH, passed by value to the trace kernels.
// Node numbering follows the reference (inc/bvh.cuh:25-28): internal
// nodes [0, n-2], leaf of sorted position p at n-1+p.
struct BvhView {
  int32_t n;
  const float* node_box;   // (2n-1)*6: lo xyz, hi xyz
  const float4* tnode;     // (n-1)*4: child boxes and child refs (64 B per internal node)
  const float4* ltri;      // (n+1)*kLeafStride: vertices a, b, c of the leaf's triangle + triangle id (48 B
                           // used); record n is the null leaf (all zero, a degenerate triangle no ray hits)
  const float4* wide;      // (n-1)*8: 4-wide nodes (128 B), valid at even-depth internal nodes
  const float4* wideq;     // 8-bit quantized 4-wide nodes (64 B), packed: node k is the k-th even-depth node
  const float4* wide8;     // (n-1)*8: 8-wide quantized nodes (128 B), valid at depths divisible by 3
  const int32_t* leaf_of;  // n: sorted leaf position of each triangle (inverse of the sorted ids)
};

// 4-wide traversal node (128 B, eight float4), stored at the index of the
// binary internal node it replaces (the internal nodes at even depth). Its
// children are the binary node's children, each internal one replaced by
// its own two children, in sorted-leaf order; their fp32 boxes are the
// exact refit boxes:
//   [0] lo.x of children 0..3   [1] lo.y   [2] lo.z
//   [3] hi.x of children 0..3   [4] hi.y   [5] hi.z
//   [6] child refs (int bits): >= 0 wide node, < 0 ~(sorted leaf position);
//       an unused slot names the null leaf ~n (kNullLeaf) and has an empty
//       box (lo = +inf, hi = -inf), so the traversal needs no test for it:
//       the box is never entered, and if a child test let it through, the
//       null leaf's degenerate triangle fails the triangle test
//   [7] unused
// The quantized form (64 B, four float4) stores the node's box corner and,
// per axis, a power-of-two step; child planes are lo = corner + qlo*step
// (rounded down) and hi = corner + qhi*step (rounded up), so the decoded
// child boxes contain the exact ones:
//   [0] corner.xyz, step.x (a float: a power of two, or 0 for a flat axis)
//   [1] qlo of children 0..3 (byte j = child j) for x, y, z; step.y
//   [2] qhi likewise; step.z
//   An unused slot has qlo = 255 and qhi = 0 on every axis: an inverted box,
//   which the conservative child test rejects whenever the widening E of an
//   axis is below 255 steps (always, except for rays with coordinates of
//   order 2^24 steps; the null leaf covers those)
//   [3] child refs as in the 128-B form, except that a wide node is named
//       by its packed slot: the quantized nodes are stored densely, the k-th
//       even-depth internal node (in index order) at slot k
// 8-wide quantized node (128 B), stored at the binary internal nodes whose
// depth is a multiple of 3; its up to eight children are the binary
// descendants three levels down (or the leaves met before), in sorted-leaf
// order:
//   [0] corner.xyz, step exponents in bytes 0-2
//   [1] qlo.x of children 0-3, 4-7, qlo.y of 0-3, 4-7
//   [2] qlo.z of 0-3, 4-7, qhi.x of 0-3, 4-7
//   [3] qhi.y of 0-3, 4-7, qhi.z of 0-3, 4-7
//   [4] refs of children 0-3   [5] refs of children 4-7   [6], [7] unused
// The child ref of an unused slot of a BVH with n leaves: the null leaf, at
// sorted position n (one record past the last leaf).
__host__ __device__ constexpr int32_t kNullLeaf(int32_t n) { return ~n; }

// Traversal record of internal node i (64 B, four float4):
//   [0] = left.lo.xyz,  left.hi.x
//   [1] = left.hi.yz,   right.lo.xy
//   [2] = right.lo.z,   right.hi.xyz
//   [3] = left ref, right ref (int bits), 0, 0
// A ref >= 0 is an internal node; a ref < 0 is ~(sorted leaf position).
// Leaf record p (48 B): [0] = a.xyz, b.x; [1] = b.yz, c.xy; [2] = c.z, tri id bits, 0, 0.

struct rt_bvh {
  rt_ctx* ctx = nullptr;
  int32_t n = 0;
  int32_t bits = 30;
  int32_t passes = 4;
  int32_t ntiles = 0;
  // one device arena; the pointers below point into it
  void* arena = nullptr;
  size_t arena_bytes = 0;
  float* tri_box = nullptr;      // n*6
  int32_t build_flags = 0;
  uint32_t* misc = nullptr;      // see kMisc* below
  uint32_t* hist = nullptr;      // passes*256
  uint32_t* status = nullptr;    // passes*ntiles*256 (contiguous after hist)
  uint32_t* arrivals = nullptr;  // n-1 arrival counters of the top-level climb (contiguous after status)
  float4* deposits = nullptr;    // (n-1)*2 child deposits, 32 B each
  uint32_t* deferred = nullptr;  // n climb hand-offs to the top-level passes
  uint32_t* deferred2 = nullptr; // n: second hand-off list (the top-level passes alternate)
  uint32_t* fetch = nullptr;     // 1024 words: ray fetch counters of rt_trace
  void* keys[2] = {nullptr, nullptr};
  uint32_t* ids[2] = {nullptr, nullptr};
  int32_t* parent = nullptr;     // 2n-1
  int32_t* left = nullptr;       // max(n-1,1)
  int32_t* right = nullptr;      // max(n-1,1)
  uint32_t* flags = nullptr;     // max(n-1,1)
  float* node_box = nullptr;     // (2n-1)*6
  float4* tnode = nullptr;       // (n-1)*4
  float4* ltri = nullptr;        // n*kLeafStride
  float4* trirec = nullptr;      // n*3: the leaf records in triangle order (written with the triangle boxes)
  float4* wide = nullptr;        // (n-1)*8
  float4* wideq = nullptr;       // (n-1)*4 reserved; the packed nodes fill the front
  int32_t* wcomp = nullptr;      // n-1: packed index of each even-depth internal node
  int4* wsrc = nullptr;          // per packed node: the binary nodes of its four child slots (-1: empty)
  uint32_t* wscan = nullptr;     // per-1024-node block sums of the packing scan
  float4* wide8 = nullptr;       // (n-1)*8
  uint8_t* wide_flag = nullptr;  // n-1: bit 0 = even depth (4-wide node), bit 1 = depth divisible by 3 (8-wide)
  int32_t* leaf_of = nullptr;    // n: triangle id -> sorted leaf position
  // The binary records, the unquantized 4-wide nodes and the 8-wide nodes
  // are only kept up to date once a traversal has asked for them
  // (rt_ensure_layout); from then on every build and refit fills them.
  mutable bool want_tnode = false;  // binary records (layout 2, the megakernel); the search build always writes them
  mutable bool want_wide4 = false;
  mutable bool want_wide8 = false;
  // A rebuild or refit recorded in a hipGraph writes the layouts wanted when
  // it was recorded (cap_*, the intersection over every recording); a layout
  // wanted only later is refilled before every traversal that reads it, since
  // replays of that graph leave it stale.
  mutable bool captured = false, cap_tnode = true, cap_wide4 = true, cap_wide8 = true;
  // How far the staged build has got (kStage*); rt_bvh_build and
  // rt_bvh_rebuild go straight to kStageBuilt. Traversal, refit and export
  // need kStageBuilt.
  int32_t stage = 0;

  BvhView view() const { return BvhView{n, node_box, tnode, ltri, wide, wideq, wide8, leaf_of}; }
};

enum : int32_t { kStageCreated = 0, kStageBounds = 1, kStageMorton = 2, kStageSorted = 3, kStageBuilt = 4 };

// misc[] slots
constexpr int kBoundsReplicas = 8;  // scene-bound accumulators, spread to cut atomic contention
constexpr int kMiscBounds = 0;      // kBoundsReplicas * 6 words (min xyz, max xyz)
constexpr int kMiscTileCtr = 48;    // up to 8 words: onesweep tile counters
constexpr int kMiscOverflow = 56;   // 1 word: traversal stack overflows
constexpr int kMiscDeferred = 57;   // 1 word: number of climb hand-offs
constexpr int kMiscBuildError = 58; // 1 word: climb consistency failures (must stay 0)
constexpr int kMiscFetch = 59;      // 1 word: ray fetch counter of rt_trace
constexpr int kMiscDeferredB = 60;  // 1 word each: hand-off counts of the top-level passes
constexpr int kMiscDeferredC = 61;  //   (kMiscDeferred, kMiscDeferredB, kMiscDeferredC in rotation)
constexpr int kMiscWords = 64;

// Leaf records are kLeafStride float4 apart: the 48-B record (three
// vertices + triangle id) padded to 64 B, so no record straddles a
// 128-B cache line.
#ifndef RT_LEAF_STRIDE
#define RT_LEAF_STRIDE 4
#endif
constexpr int kLeafStride = RT_LEAF_STRIDE;

// Fused hierarchy + refit climb: one workgroup per chunk of sorted leaves
// (its LDS holds about 76 B per leaf of the chunk).
#ifndef RT_CLIMB_CHUNK
#define RT_CLIMB_CHUNK 512
#endif
#ifndef RT_CLIMB_THREADS
#define RT_CLIMB_THREADS RT_CLIMB_CHUNK
#endif
constexpr int kClimbThreads = RT_CLIMB_THREADS;
constexpr int kClimbChunk = RT_CLIMB_CHUNK;
constexpr int kClimbLeavesPerThread = kClimbChunk / kClimbThreads;
static_assert(kClimbChunk % kClimbThreads == 0, "a climb chunk is a whole number of leaves per thread");
constexpr int kClimbTopThreads = 1024;

// Sort tiling.
constexpr int kSortThreads = 256;
constexpr int kSortKeysPerThread = 16;
constexpr int kSortTile = kSortThreads * kSortKeysPerThread;

rt_status rt_set_error(rt_ctx* ctx, rt_status s, const char* fmt, ...);
// RT_OK if the BVH is built, else RT_ERR_INVALID with a message naming `who`.
rt_status rt_bvh_require_built(const rt_bvh* bvh, const char* who);

// Launchers (rt_build.hip).
hipError_t rt_launch_build(rt_bvh* b, const rt_mesh_view* m, hipStream_t s);
hipError_t rt_launch_refit(rt_bvh* b, const rt_mesh_view* m, hipStream_t s);
// The build's stages one at a time (rt_launch_build runs all four):
// triangle boxes + scene bounds (optionally copied out as n*3 arrays),
// Morton keys (normalised by mn/mx when both are given, else by the scene
// bounds), the radix sort, and the hierarchy with the traversal layouts.
hipError_t rt_launch_stage_bounds(rt_bvh* b, const rt_mesh_view* m, float* lo, float* hi, hipStream_t s);
hipError_t rt_launch_stage_morton(rt_bvh* b, const double* mn, const double* mx, hipStream_t s);
hipError_t rt_launch_stage_sort(rt_bvh* b, hipStream_t s);
hipError_t rt_launch_stage_tree(rt_bvh* b, const rt_mesh_view* m, hipStream_t s);
// Record that a rebuild or refit of b is being launched on s; if s is
// capturing a hipGraph, the layouts it writes are noted (see rt_bvh::captured).
hipError_t rt_note_capture(const rt_bvh* b, hipStream_t s);
// Before the first traversal of a BVH with node layout 2, 4 or 8: fill those
// nodes now and keep them filled by every later build and refit.
hipError_t rt_ensure_layout(const rt_bvh* b, int layout, hipStream_t s);
// Builds and refits stop filling the binary records and the unquantized
// 4-wide and 8-wide nodes unless `layout` traverses them; rt_ensure_layout
// fills them again on demand.
void rt_keep_only_layout(const rt_bvh* b, int layout);

// Segmented build of a batch's BVHs: each build stage is one launch over
// all of them. create returns *out = nullptr (and hipSuccess) when the BVHs
// cannot share launches (mixed key widths, the search hierarchy, n < 1).
// The launch builds as rt_launch_build does for every BVH, except the
// binary records and the unquantized 4-wide and 8-wide nodes: use it only
// while no BVH wants them (want_tnode, want_wide4, want_wide8).
struct rt_seg_build;
hipError_t rt_seg_build_create(rt_ctx* ctx, const rt_batch_item* items, int count, rt_seg_build** out);
hipError_t rt_seg_build_launch(const rt_seg_build* p, hipStream_t s);
void rt_seg_build_destroy(rt_ctx* ctx, rt_seg_build* p);
hipError_t rt_launch_export(const rt_bvh* b, const rt_bvh_export_bufs* dst, hipStream_t s);

// Launchers (rt_trace.hip).
hipError_t rt_launch_trace(const rt_bvh* b, const rt_mesh_view* m, const rt_rays* rays, int32_t mode, double tmin,
                           const rt_hits* hits, hipStream_t s);
// workspace == nullptr selects the megakernel, else the wavefront passes
// (rt_wavefront_bytes(width, height) bytes of device memory).
// records/color64 (optional, wavefront only): path record and FP64 image
// of the differentiable render.
hipError_t rt_launch_render(const rt_bvh* b, const rt_mesh_view* m, const rt_material* mats, int32_t nmat,
                            const rt_light* light, const rt_camera* cam, const rt_render_opts* opts, float* rgb,
                            uint64_t* counters, void* workspace, void* shade_table, hipStream_t s,
                            int32_t* records = nullptr, double* color64 = nullptr,
                            rt_kernel_timer* timer = nullptr, const rt_overlap* ov = nullptr);
size_t rt_wavefront_bytes(int32_t width, int32_t height);
// Per-leaf shading table of an n-triangle scene: 64 B per sorted leaf
// (three vertex normals and the material id), refreshed by every render.
inline size_t rt_shade_table_bytes(size_t n) { return n * 64; }
size_t rt_wavefront_bytes_batch(size_t npix, size_t ntiles, size_t nscenes);
size_t rt_scene_desc_bytes();
size_t rt_bvh_view_bytes();
void rt_fill_scene(void* table_entry, const rt_bvh* b, const rt_mesh_view* m, const rt_material* mats, int32_t nmat,
                   const rt_light* light, const rt_camera* cam, const rt_render_opts* opts, float* rgb,
                   uint32_t pix_off, uint32_t tile_off, uint32_t tri_off, void* shade_table);
hipError_t rt_launch_wavefront(void* workspace, size_t npix, size_t ntiles, size_t nscenes, int max_spp,
                               int max_depth, int num_cus, const rt_tuning* tuning, uint64_t* counters,
                               uint32_t* overflow, hipStream_t s, int32_t* records = nullptr,
                               double* color64 = nullptr, const void* single_scene = nullptr,
                               rt_kernel_timer* timer = nullptr, const rt_overlap* ov = nullptr,
                               const uint32_t* live_tiles = nullptr);
// Differentiable render (rt_diff.hip).
hipError_t rt_launch_zero_f64(double* p, size_t n, hipStream_t s);
hipError_t rt_launch_resolve(const float* big, int32_t W, int32_t H, int32_t ss, bool reference_divisor, float* out,
                             hipStream_t s);
hipError_t rt_launch_fill_i32(int32_t* p, size_t n, int32_t v, hipStream_t s);
hipError_t rt_launch_fill_albedo(rt_material* mats, int32_t* ids, const double* albedo, int32_t n, hipStream_t s);
hipError_t rt_launch_adjoint(const rt_bvh* b, const rt_mesh_view* m, const double* albedo, double light_scale,
                             const rt_light* light, const rt_camera* cam, const rt_render_opts* opts,
                             const int32_t* records, const double* dcolor, double* g_albedo, double* g_light,
                             double* g_pos, hipStream_t s);
uint32_t* rt_wavefront_counts(void* workspace, size_t npix, size_t ntiles, size_t nscenes);
rt_tuning rt_default_tuning();
bool rt_tuning_valid(const rt_tuning* t);
// Size bucket of a traversal: floor(log2) of the triangle count and of the
// rays of one pass (a render's width * height), 8 bits each.
uint32_t rt_tuning_key(int64_t ntris, int64_t nrays);
// The context's tuning for that size, or the default.
rt_tuning rt_ctx_tuning_for(const rt_ctx* ctx, int64_t ntris, int64_t nrays);

// Autotuning by coordinate descent from `start`: each parameter in turn
// (node layout, workgroup size, waves per SIMD, fetch chunk, stack ring)
// takes every value with the others fixed, and the fastest is kept.
// time(t, &ms) runs the work under t and reports its time.
template <class TimeFn>
rt_status rt_autotune_descent(rt_tuning start, TimeFn time, rt_tuning* best, float* best_ms) {
  rt_tuning cur = start;
  if (cur.wg_threads == 0) cur.wg_threads = 64;
  float cur_ms = 0.f;
  rt_status st = time(cur, &cur_ms);
  if (st != RT_OK) return st;
  const int layouts[] = {2, 4, 5, 8}, wgs[] = {64, 256}, waves[] = {2, 4, 8}, chunks[] = {16, 32, 64, 128},
            rings[] = {16, 32};
  auto try_value = [&](rt_tuning t) -> rt_status {
    float ms = 0.f;
    rt_status s = time(t, &ms);
    if (s == RT_OK && ms < cur_ms) {
      cur = t;
      cur_ms = ms;
    }
    return s;
  };
  for (int v : layouts)
    if (v != cur.node_layout) {
      rt_tuning t = cur;
      t.node_layout = v;
      if ((st = try_value(t)) != RT_OK) return st;
    }
  for (int v : wgs)
    if (v != cur.wg_threads) {
      rt_tuning t = cur;
      t.wg_threads = v;
      if ((st = try_value(t)) != RT_OK) return st;
    }
  for (int v : waves)
    if (v != cur.waves_per_simd) {
      rt_tuning t = cur;
      t.waves_per_simd = v;
      if ((st = try_value(t)) != RT_OK) return st;
    }
  for (int v : chunks)
    if (v != cur.fetch_chunk) {
      rt_tuning t = cur;
      t.fetch_chunk = v;
      if ((st = try_value(t)) != RT_OK) return st;
    }
  for (int v : rings)
    if (v != cur.stack_ring) {
      rt_tuning t = cur;
      t.stack_ring = v;
      if ((st = try_value(t)) != RT_OK) return st;
    }
  *best = cur;
  *best_ms = cur_ms;
  return RT_OK;
}

// Offsets of the per-batch tables inside a wavefront workspace (first three
// regions: scene table, BVH views, tile -> scene map).
constexpr size_t rt_align256(size_t v) { return (v + 255) & ~(size_t)255; }

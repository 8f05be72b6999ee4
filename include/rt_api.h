/* MI355X (gfx950) ray-tracer core: C-ABI boundary.
 *
 * This is the drop-in boundary for the reference's build / refit / trace
 * path (mbartling/cuda-rt). Each entry point names the reference interface
 * it replaces. All pointers named *_dev are device pointers (HBM); the
 * caller owns them. The library owns every BVH buffer until
 * rt_bvh_destroy. Calls are ordered on the given HIP stream (NULL = the
 * null stream) and do not synchronise unless stated. Errors are returned
 * as rt_status codes; rt_last_error() gives the message. The library never
 * exits the process.
 */
#ifndef RT_API_H
#define RT_API_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  RT_OK = 0,
  RT_ERR_INVALID = 1,     /* bad argument (null pointer, size, option) */
  RT_ERR_OOM = 2,         /* device allocation failed */
  RT_ERR_DEVICE = 3,      /* HIP runtime error */
  RT_ERR_UNSUPPORTED = 4, /* option not implemented */
  RT_ERR_OVERFLOW = 5     /* traversal stack overflow was recorded */
} rt_status;

typedef struct rt_ctx rt_ctx;
typedef struct rt_bvh rt_bvh;

/* Triangle mesh, fp32 SoA positions + int32 vertex indices.
 * Replaces the Vec3d vertices / normals / TriangleIndices arrays that
 * Scene_d uploads (src/scene.cpp:82-138, inc/scene.h:60-117). */
typedef struct {
  const float *vx_dev, *vy_dev, *vz_dev; /* nv positions */
  const float *nx_dev, *ny_dev, *nz_dev; /* nv shading normals (render only; may be NULL for build/trace) */
  const int32_t *idx_dev;                /* 3*n vertex indices (a, b, c) */
  const int32_t *mat_id_dev;             /* n material ids, or NULL for material 0 */
  int32_t nv;
  int32_t n;
} rt_mesh_view;

enum {
  RT_BUILD_DEFAULT = 0,       /* fused bottom-up hierarchy + refit climb */
  RT_BUILD_KARRAS_SEARCH = 1  /* per-node span/split search (kernels/bvh.cu:100-207), then the atomic refit */
};

typedef struct {
  int32_t morton_bits; /* 30 (reference, kernels/bvh.cu:22-32) or 63 */
  int32_t flags;       /* RT_BUILD_* */
} rt_build_opts;

/* Destination buffers for rt_bvh_export, all device pointers, in the
 * reference's layout (inc/bvh.cuh:18-28, kernels/bvh.cu:164-258):
 * nodes [0, n-2] are internal, [n-1, 2n-2] are leaves with
 * left == right == triangle id; the root's parent is -1. */
typedef struct {
  void *codes_dev;      /* n sorted Morton keys: uint32 (30-bit) or uint64 (63-bit) */
  uint32_t *ids_dev;    /* n sorted triangle ids (object_ids) */
  int32_t *parent_dev;  /* 2n-1 */
  int32_t *left_dev;    /* 2n-1 */
  int32_t *right_dev;   /* 2n-1 */
  float *lo_dev;        /* (2n-1)*3 node AABB minimum, xyz per node (sortedBBoxs) */
  float *hi_dev;        /* (2n-1)*3 node AABB maximum */
  float *tri_lo_dev;    /* n*3 per-triangle AABB minimum (BBoxs), may be NULL */
  float *tri_hi_dev;    /* n*3, may be NULL */
} rt_bvh_export_bufs;

/* Rays for rt_trace: AoS xyz doubles (the reference's ray, inc/ray.h:20-50). */
typedef struct {
  const double *orig_dev; /* 3*n */
  const double *dir_dev;  /* 3*n */
  int64_t n;
} rt_rays;

/* Hit records (the reference's isect, inc/isect.h:4-15). tri = -1 on miss.
 * alpha/beta are the Moller-Trumbore barycentrics of vertices b and c. */
typedef struct {
  int32_t *tri_dev;
  double *t_dev;     /* may be NULL */
  double *alpha_dev; /* may be NULL */
  double *beta_dev;  /* may be NULL */
} rt_hits;

enum {
  RT_TRACE_CLOSEST = 0, /* BVH_d::intersect, kernels/bvh.cu:382-414 */
  RT_TRACE_ANY = 1      /* BVH_d::intersectAny, kernels/bvh.cu:416-449 */
};

/* Material (inc/material.h:9-34). Trans() is true when kt == 0, as in the reference. */
typedef struct {
  double ka[3], kd[3], ks[3], kt[3], ke[3];
  double shininess;
} rt_material;

/* Positional light (inc/Light.h:10-44, kernels/Light.cu). */
typedef struct {
  double position[3];
  double color[3];
  double radius;       /* jitter: offset = rand3*radius - 0.5 (kernels/Light.cu:73-75) */
  double c0, c1, c2;   /* falloff min(1, 1/(c0 + c1 d + c2 d^2)) */
} rt_light;

/* Camera getters used by the render kernel (inc/camera.h:33-46). */
typedef struct {
  int32_t width, height;
  double fov_deg;      /* getFOV */
  double fstop;        /* getAperature() = 1/fstop; INFINITY gives a pinhole */
  double focal_point;  /* getFocalPoint */
  double eye[3];       /* getEye (used by the specular term) */
} rt_camera;

enum {
  RT_RENDER_RANDVEC_RTL = 1, /* draw randVec3d's components z, y, x instead of x, y, z */
  RT_RENDER_CLAMP = 2,       /* clamp each channel to [0,1] (kernels/scene.cu:137) */
  RT_RENDER_WAVEFRONT = 4,   /* ray-queue passes instead of the megakernel (same image) */
  RT_RENDER_ONE_STREAM = 8   /* wavefront: every pass on the caller's stream (by default the
                                last any-hit pass of a sample overlaps the next sample on a
                                second, context-owned stream; same image either way) */
};

typedef struct {
  int32_t spp;    /* samples per pixel (reference: 20, kernels/raytracer.cu:138) */
  int32_t depth;  /* diffuse bounces after the first hit; 0 = reference shading */
  uint32_t seed;  /* XORed into the seed index; 0 = reference seeding */
  int32_t flags;  /* RT_RENDER_* */
} rt_render_opts;

/* Library / context. */
const char *rt_version(void);
rt_status rt_ctx_create(int32_t device, rt_ctx **out);
void rt_ctx_destroy(rt_ctx *ctx);
const char *rt_last_error(const rt_ctx *ctx);

/* Build: Scene_d::computeBoundingBoxes + findMinMax (kernels/scene.cu:48-85)
 * followed by BVH_d::setUp -> computeMortonCodes -> sortMortonCodes ->
 * buildTree (src/bvh.cpp:3-32, kernels/bvh.cu:260-292). Allocates the BVH. */
rt_status rt_bvh_build(rt_ctx *ctx, const rt_mesh_view *mesh, const rt_build_opts *opts, void *stream,
                       rt_bvh **out);
/* Staged build: the calls of BVH_d::setUp one at a time (src/bvh.cpp:3-32,
 * inc/bvh.cuh:53-59), for callers written against the reference's stage
 * methods. rt_bvh_create allocates a BVH for n triangles without building
 * it (BVH_d::setUp's cudaMallocs, src/bvh.cpp:12-18); the four stages then
 * run in this order, each stream-ordered and allocation-free. After
 * rt_bvh_build_tree the BVH is the one rt_bvh_build makes when the Morton
 * bounds are the scene bounds. Out of order, a stage returns RT_ERR_INVALID;
 * traversal, refit and export need a built BVH (rt_bvh_build_tree,
 * rt_bvh_build or rt_bvh_rebuild). */
rt_status rt_bvh_create(rt_ctx *ctx, int32_t n, const rt_build_opts *opts, rt_bvh **out);
/* Scene_d::computeBoundingBoxes (kernels/scene.cu:48-57,88-120) with the
 * reduction of Scene_d::findMinMax (:72-85): per-triangle boxes, and the
 * scene bounds (read them with rt_bvh_scene_bounds). If tri_lo_dev /
 * tri_hi_dev are not NULL they also receive the boxes, n*3 floats each in
 * triangle order (the reference's BBoxs). */
rt_status rt_bvh_compute_bounds(rt_bvh *bvh, const rt_mesh_view *mesh, float *tri_lo_dev, float *tri_hi_dev,
                                void *stream);
/* BVH_d::computeMortonCodes(mMin, mMax) (kernels/bvh.cu:22-61,260-265):
 * keys of the box centroids normalised by mn / mx, or by the scene bounds of
 * rt_bvh_compute_bounds if both are NULL; object ids 0..n-1. */
rt_status rt_bvh_compute_morton(rt_bvh *bvh, const double mn[3], const double mx[3], void *stream);
/* BVH_d::sortMortonCodes (kernels/bvh.cu:267-273): stable sort of the keys
 * with their object ids. */
rt_status rt_bvh_sort_morton(rt_bvh *bvh, void *stream);
/* BVH_d::buildTree (kernels/bvh.cu:278-292): hierarchy and node boxes from
 * the sorted keys, then the traversal node layouts. */
rt_status rt_bvh_build_tree(rt_bvh *bvh, const rt_mesh_view *mesh, void *stream);
/* Rebuild the whole hierarchy for a mesh with the same triangle count into
 * an existing BVH without allocating (stream-ordered, graph-capturable). */
rt_status rt_bvh_rebuild(rt_bvh *bvh, const rt_mesh_view *mesh, void *stream);
/* Refit only: recompute triangle and node AABBs for moved vertices keeping
 * the topology. The reference runs RefitBounds only inside buildTree
 * (kernels/bvh.cu:288); this entry point re-zeroes the flags itself. */
rt_status rt_bvh_refit(rt_bvh *bvh, const rt_mesh_view *mesh, void *stream);
/* Copy the BVH into the reference layout (for parity checks). */
rt_status rt_bvh_export(const rt_bvh *bvh, const rt_bvh_export_bufs *dst, void *stream);
/* Scene bounds mMin / mMax as Scene_d::findMinMax returns them
 * (kernels/scene.cu:72-85). Synchronises the stream. */
rt_status rt_bvh_scene_bounds(const rt_bvh *bvh, void *stream, double mn[3], double mx[3]);
int32_t rt_bvh_num_triangles(const rt_bvh *bvh);
/* Diagnostic counters since the BVH was created: out[0] = traversal stack
 * overflows (rays reported as misses), out[1] = build consistency
 * failures. Both stay 0 in correct operation. Synchronises the stream. */
rt_status rt_bvh_diagnostics(const rt_bvh *bvh, void *stream, uint32_t out[2]);
void rt_bvh_destroy(rt_bvh *bvh);

/* Trace: BVH_d::intersect / intersectAny (kernels/bvh.cu:297-449).
 * tmin = 0 keeps the reference rule (every hit with |t| >= 1e-8 on the whole
 * line counts); tmin > 0 only accepts t >= tmin. */
rt_status rt_trace(const rt_bvh *bvh, const rt_mesh_view *mesh, const rt_rays *rays, int32_t mode, double tmin,
                   const rt_hits *hits, void *stream);

/* Render: runRayTracerKernelRec + traceRay + Material::shade +
 * Light::{getDirection, distanceAttenuation, shadowAttenuation}
 * (kernels/raytracer.cu:94-249, kernels/material.cu:6-40, kernels/Light.cu).
 * rgb_dev receives width*height*3 floats, row-flipped like the reference
 * image (kernels/raytracer.cu:106). counters_dev, if not NULL, receives
 * 5 uint64 counts added atomically: primary, shadow and bounce rays,
 * child-box slab tests (two per binary node entered, one per filled child
 * slot of a wide node entered) and triangle tests. The ray counts are
 * exact and count every shadow ray the reference casts (one per hit); a
 * shadow ray is traced only when its answer can change the image, i.e.
 * when some channel of (max(0, N.L) kd + specular ks) % light colour is
 * non-zero (the image is identical either way). The two work counts can
 * differ slightly between identical calls, because the 4-wide traversal
 * batches leaf tests across a wave and closest-hit pruning then depends on
 * which rays share a wave. Images and hits never depend on it.
 * RT_RENDER_WAVEFRONT keeps per-pixel state in a context-owned workspace
 * that the first such call for a given image size allocates, and copies the
 * mesh's normals and material ids into a per-leaf shading table (64 B per
 * triangle, also context-owned) at the start of every call; renders that
 * share a context must not overlap on different streams. */
rt_status rt_render(rt_ctx *ctx, const rt_bvh *bvh, const rt_mesh_view *mesh, const rt_material *mats_dev,
                    int32_t nmat, const rt_light *light, const rt_camera *cam, const rt_render_opts *opts,
                    float *rgb_dev, uint64_t *counters_dev, void *stream);

/* rt_render without allocation. No reference counterpart: it is for callers
 * that cannot allocate device memory (code under hipGraph capture, tight
 * serving loops, harnesses that forbid hipMalloc). It renders exactly as
 * rt_render with RT_RENDER_WAVEFRONT | RT_RENDER_ONE_STREAM (same image,
 * same counters), but keeps the wavefront workspace and the shading table
 * in a library-owned static device array of rt_render_noalloc_max_bytes()
 * (48 MiB) per GPU, found when the context is created. It calls no
 * hipMalloc/hipFree and creates no stream or event. A call whose workspace
 * (about 560 B per pixel) plus shading table (64 B per triangle) exceeds
 * the array returns RT_ERR_INVALID; 256 x 256 fits with up to 210,000
 * triangles. Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other streams, other contexts, other threads) race. */
rt_status rt_render_noalloc(rt_ctx *ctx, const rt_bvh *bvh, const rt_mesh_view *mesh, const rt_material *mats_dev,
                            int32_t nmat, const rt_light *light, const rt_camera *cam, const rt_render_opts *opts,
                            float *rgb_dev, uint64_t *counters_dev, void *stream);
size_t rt_render_noalloc_max_bytes(void);


/* Per-kernel timing for benchmarks (no reference counterpart). While
 * enabled, every closest-hit and any-hit traversal launch of rt_render and
 * rt_batch_render on this context is bracketed by HIP events on its stream,
 * and, for renders given counters_dev, the traversals' work counts are also
 * kept per kernel (a render without counters counts no work).
 * rt_ctx_kernel_stats synchronises the device, returns the totals since the
 * last call (out[0] closest hit, out[1] any hit) and resets them. */
typedef struct {
  uint64_t launches;
  double ms;            /* summed launch durations */
  uint64_t box_tests;   /* child-box tests */
  uint64_t tri_tests;   /* triangle tests */
  uint64_t rays;        /* rays traced (closest hit: primary + bounce; any hit: shadow rays queried) */
} rt_kernel_stat;
rt_status rt_ctx_set_kernel_timing(rt_ctx *ctx, int32_t enable);
rt_status rt_ctx_kernel_stats(rt_ctx *ctx, rt_kernel_stat out[2]);

/* Device memory accounting and allocation fault injection (no reference
 * counterpart: the reference exits the process when a cudaMalloc fails,
 * src/scene.cpp:71-81). Every device allocation the library makes for a
 * context's objects (BVH arenas, render and differentiable-render
 * workspaces, shading tables, batches, timers) goes through one allocator.
 * rt_ctx_device_bytes gives the bytes those objects hold now.
 * rt_ctx_fail_alloc(ctx, k) makes the k-th allocation from now on fail as a
 * failed hipMalloc would (k = 0 disarms it): the call that makes it returns
 * RT_ERR_OOM, frees what it had allocated, and leaves the context and its
 * other objects usable. Under hipGraph capture no call allocates: one that
 * would have to (a render with a larger image than the context's workspace
 * holds) returns RT_ERR_INVALID and leaves the capture intact. */
rt_status rt_ctx_fail_alloc(rt_ctx *ctx, int32_t k);
rt_status rt_ctx_device_bytes(const rt_ctx *ctx, uint64_t *bytes);

/* Supersampling resolve: AverageSuperSamplingKernel (kernels/scene.cu:127-155).
 * big_dev is a (width*ss) x (height*ss) image as rt_render writes it (3
 * floats per pixel, row-flipped); out_dev receives width x height. Each
 * sample is clamped to [0, 1]; the sum of the ss*ss samples is divided by
 * ss*ss, or by ss as the reference does when RT_RESOLVE_REFERENCE_DIVISOR
 * is set (the reference's /ss makes images ss times too bright for ss > 1;
 * both agree at ss = 1, the only value the reference runs). */
enum { RT_RESOLVE_REFERENCE_DIVISOR = 1 };
rt_status rt_resolve_supersample(rt_ctx *ctx, const float *big_dev, int32_t width, int32_t height, int32_t ss,
                                 int32_t flags, float *out_dev, void *stream);

/* ---- batched rendering (no reference counterpart: the reference renders
 * one scene per process, src/main.cpp:316-330). Many scenes, each with its
 * own BVH, camera, light, materials, spp and depth, go through one set of
 * wavefront launches whose persistent traversal kernels take rays of every
 * scene from one queue. Each scene's image is identical to rt_render's. */
typedef struct {
  int32_t stack_ring;      /* LDS traversal stack entries per lane: 16 or 32 */
  int32_t fetch_chunk;     /* rays a wave reserves per atomic: 16, 32, 64, 128, 256 or 512 */
  int32_t waves_per_simd;  /* persistent traversal workgroups per CU, 1..8 */
  int32_t node_layout;     /* 2: binary nodes, FP64 child tests, 64 B per node;
                              4: 4-wide nodes, fp32 conservative child tests, 128 B per node;
                              5: 4-wide nodes with 8-bit quantized child boxes, 64 B per node (default);
                              8: 8-wide nodes with 8-bit quantized child boxes, 128 B per node.
                              Every layout gives the same hits and images. */
  int32_t wg_threads;      /* persistent traversal workgroup: 64 (one wave, the default) or 256
                              (four waves); 0 means 64 */
} rt_tuning;

typedef struct {
  const rt_bvh *bvh;
  rt_mesh_view mesh;       /* device pointers; normals required */
  const rt_material *mats_dev;
  int32_t nmat;
  rt_light light;
  rt_camera cam;
  rt_render_opts opts;     /* RT_RENDER_WAVEFRONT is implied */
  float *rgb_dev;          /* width*height*3 floats, as rt_render */
} rt_batch_item;

typedef struct rt_batch rt_batch;

/* Allocate the batch workspace and upload the scene table (synchronous).
 * The items' device buffers must stay valid while the batch is used. */
rt_status rt_batch_create(rt_ctx *ctx, const rt_batch_item *items, int32_t count, rt_batch **out);
/* Rebuild every scene's BVH; builds run on parallel internal streams that
 * fork from and join back into `stream` (graph-capturable). */
rt_status rt_batch_rebuild(rt_batch *batch, void *stream);
/* Render every scene. Allocates nothing and never synchronises, so it can
 * be captured in a hipGraph. counters_dev as in rt_render (summed). */
rt_status rt_batch_render(rt_batch *batch, uint64_t *counters_dev, void *stream);
rt_status rt_batch_set_tuning(rt_batch *batch, const rt_tuning *tuning);
rt_status rt_batch_get_tuning(const rt_batch *batch, rt_tuning *tuning);
/* Time rt_batch_render under candidate tunings (synchronises; not under
 * graph capture), keep the fastest and return it and its time (the faster of
 * two timed renders after an untimed one). The search is a coordinate
 * descent from the batch's current tuning: node layout {2, 4, 5, 8}, then
 * workgroup size {64, 256}, waves per SIMD {2, 4, 8}, fetch chunk {16, 32,
 * 64, 128} and stack ring {16, 32}, each over its values with the others
 * fixed. Outputs do not depend on tuning. */
rt_status rt_batch_autotune(rt_batch *batch, void *stream, rt_tuning *best, float *best_ms);
/* out[0] = traversal stack overflows during batch renders (0 when healthy). */
rt_status rt_batch_diagnostics(const rt_batch *batch, void *stream, uint32_t out[1]);
void rt_batch_destroy(rt_batch *batch);

/* Single-scene tuning (no reference counterpart; the reference's launch
 * shape is fixed, kernels/raytracer.cu:55-70). A context keeps a table of
 * rt_tuning entries for its device architecture, keyed by the size of the
 * work: floor(log2(triangles)) and floor(log2(rays per pass)), the rays per
 * pass of a render being width * height. Wavefront rt_render and
 * rt_render_noalloc, rt_trace and the differentiable render use the entry
 * for their size, else the default tuning (layout 5, 64-thread workgroups).
 * Images and hits do not depend on the tuning.
 * rt_render_autotune times wavefront renders of this scene under candidate
 * tunings (the coordinate descent of rt_batch_autotune, from the table's
 * entry for this size or the default), stores the fastest in the table and
 * returns it and its time; rgb_dev receives the last trial's image. It
 * synchronises and is refused under graph capture. */
rt_status rt_render_autotune(rt_ctx *ctx, const rt_bvh *bvh, const rt_mesh_view *mesh, const rt_material *mats_dev,
                             int32_t nmat, const rt_light *light, const rt_camera *cam, const rt_render_opts *opts,
                             float *rgb_dev, void *stream, rt_tuning *best, float *best_ms);
/* The entry for ntris triangles and nrays rays per pass (*found = 0 and the
 * default tuning if there is none); set one (NULL removes it). */
rt_status rt_ctx_get_tuning(const rt_ctx *ctx, int32_t ntris, int64_t nrays, rt_tuning *out, int32_t *found);
rt_status rt_ctx_set_tuning(rt_ctx *ctx, int32_t ntris, int64_t nrays, const rt_tuning *tuning);
/* Write the table as text, one line per entry (architecture, the two size
 * buckets, the five parameters), replacing the file; read such a file and
 * add its entries for this context's architecture (entries of other
 * architectures are skipped; a malformed line or invalid tuning rejects the
 * whole file with RT_ERR_INVALID, leaving the table as it was). */
rt_status rt_ctx_save_tuning(const rt_ctx *ctx, const char *path);
rt_status rt_ctx_load_tuning(rt_ctx *ctx, const char *path);

/* ---- differentiable render (no reference counterpart: the reference
 * renders forward only, kernels/raytracer.cu:94-249). The scene is the
 * rt_render scene with one material per triangle, kd = albedo[t] (RGB,
 * doubles) and ka = ks = kt = ke = 0, and the light colour scaled by
 * light_scale. Derivatives of the image are taken with respect to the
 * albedos, light_scale and the vertex positions through the shading
 * (hit point, barycentrics, interpolated normal, light direction,
 * falloff, next bounce origin and direction). Which triangle each ray hits
 * and whether each shadow ray is blocked are held fixed: the visibility
 * (silhouette edge) terms are not included. Vertex normals are inputs. */
#define RT_DIFF_MAX_DEPTH 7

/* Bytes of the path record for one image: per pixel (py*width+px),
 * sample and bounce (depth+1 slots) a pair of int32: the sorted leaf
 * position hit (-1: miss) and the shadow ray's occluder code (-1: none,
 * -2: opaque, else the occluding leaf position). Slots after a path ends
 * hold -1. */
size_t rt_diff_record_bytes(const rt_camera *cam, const rt_render_opts *opts);
/* Forward render into image_dev (width*height*3 doubles, row-flipped like
 * rt_render, the unclamped sample mean) and record every path in
 * records_dev. RT_RENDER_CLAMP is ignored; the wavefront path is used. The
 * BVH must be built or refit on the mesh's current positions. */
rt_status rt_diff_forward(rt_ctx *ctx, const rt_bvh *bvh, const rt_mesh_view *mesh, const double *albedo_dev,
                          double light_scale, const rt_light *light, const rt_camera *cam,
                          const rt_render_opts *opts, double *image_dev, int32_t *records_dev, void *stream);
/* Adjoint render: given dLoss/dimage (layout of image_dev), overwrite
 * g_albedo_dev (n*3), g_light_dev (1) and g_pos_dev (nv*3, xyz per vertex)
 * with dLoss/dalbedo, dLoss/dlight_scale and dLoss/dposition. Arguments
 * other than the gradients must be those of the matching rt_diff_forward. */
rt_status rt_diff_backward(rt_ctx *ctx, const rt_bvh *bvh, const rt_mesh_view *mesh, const double *albedo_dev,
                           double light_scale, const rt_light *light, const rt_camera *cam,
                           const rt_render_opts *opts, const int32_t *records_dev, const double *dimage_dev,
                           double *g_albedo_dev, double *g_light_dev, double *g_pos_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif

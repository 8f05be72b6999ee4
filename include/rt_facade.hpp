// C++ facade over the C-ABI with the reference's class names, so code
// written against mbartling/cuda-rt's Scene_h / Scene_d / BVH_d / Camera /
// Light_h / RayTracer (inc/scene.h, inc/bvh.cuh, inc/camera.h,
// inc/Light.h, inc/raytracer.h) compiles against this library with a
// namespace change. Header-only; link with librtamd.so and the HIP runtime.
//
// What differs from the reference classes:
//  - errors throw rtamd::Error instead of printing and exit(1);
//  - the BVH_d::intersect / intersectAny of the reference are __device__
//    functions called per ray; here they trace a batch of rays from the
//    host (rt_trace);
//  - RayTracer(width, height, ss) renders width x height (the reference's
//    constructor copies width into height, inc/raytracer.h:36, so only
//    square images work there);
//  - samples per pixel (20 in the reference, kernels/raytracer.cu:137),
//    the seed, bounce depth and the RNG draw order are settable.
#pragma once
#include <hip/hip_runtime_api.h>

#include <stdint.h>

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rt_api.h"
#include "rt_io.h"

namespace rtamd {

struct Vec3d {
  double x = 0, y = 0, z = 0;
  Vec3d() = default;
  Vec3d(double a, double b, double c) : x(a), y(b), z(c) {}
};

class Error : public std::runtime_error {
 public:
  Error(rt_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
  rt_status status;
};

namespace detail {

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(RT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

inline void check(rt_status s, const rt_ctx* ctx) {
  if (s != RT_OK) throw Error(s, ctx ? rt_last_error(ctx) : "rt error");
}

// Owning device array.
template <class T>
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept { swap(o); }
  DevBuf& operator=(DevBuf&& o) noexcept {
    swap(o);
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  void resize(size_t n) {
    if (n == n_) return;
    reset();
    if (n) hip_check(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "hipMalloc");
    n_ = n;
  }
  void upload(const T* host, size_t n) {
    resize(n);
    if (n) hip_check(hipMemcpy(p_, host, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
  }
  void upload(const std::vector<T>& v) { upload(v.data(), v.size()); }
  void download(std::vector<T>& v) const {
    v.resize(n_);
    if (n_) hip_check(hipMemcpy(v.data(), p_, n_ * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H");
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  void swap(DevBuf& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
  }
  T* p_ = nullptr;
  size_t n_ = 0;
};

}  // namespace detail

// Camera (inc/camera.h:6-60; defaults kernels/camera.cu:9-22: fov 30,
// fstop 8, focal point 10, eye at the origin looking down -z).
class Camera {
 public:
  void setFstop(double fs) { fstop_ = fs; }
  void setfov(double f) { fov_ = f; }
  void setFOV(double f) { fov_ = f; }
  void setFocalPoint(double p) { focal_ = p; }
  void setEye(const Vec3d& e) { eye_ = e; }
  double getFocalPoint() const { return focal_; }
  const Vec3d& getEye() const { return eye_; }
  double getAperature() const { return 1 / fstop_; }
  double getFOV() const { return fov_; }
  rt_camera desc(int32_t width, int32_t height) const {
    rt_camera c{};
    c.width = width;
    c.height = height;
    c.fov_deg = fov_;
    c.fstop = fstop_;
    c.focal_point = focal_;
    c.eye[0] = eye_.x;
    c.eye[1] = eye_.y;
    c.eye[2] = eye_.z;
    return c;
  }

 private:
  double fstop_ = 8.0, fov_ = 30.0, focal_ = 10.0;
  Vec3d eye_;
};

// Light_h (inc/Light.h:7-29): positional light with radius jitter and
// falloff min(1, 1 / (a + b d + c d^2)).
struct Light_h {
  Vec3d color{1.0, 1.0, 1.0};
  Vec3d position{1.0, 1.0, 1.0};
  Vec3d orientation{0.5, -1.0, -1.0};
  double width = 1, height = 1, radius = 0.5;
  double constantTerm = 0, linearTerm = 0.01, quadraticTerm = 0.001;
  rt_light desc() const {
    rt_light l{};
    l.position[0] = position.x;
    l.position[1] = position.y;
    l.position[2] = position.z;
    l.color[0] = color.x;
    l.color[1] = color.y;
    l.color[2] = color.z;
    l.radius = radius;
    l.c0 = constantTerm;
    l.c1 = linearTerm;
    l.c2 = quadraticTerm;
    return l;
  }
};

// Host scene (inc/scene.h:25-58): triangles, normals, materials.
class Scene_h {
 public:
  Scene_h(int imageWidth = 512, int imageHeight = 512, int superSampling = 1)
      : imageWidth(imageWidth), imageHeight(imageHeight), superSampling(superSampling) {}

  // Scene_h::LoadObj (src/scene.cpp:6-58) through rt_obj_load.
  void LoadObj(const std::string& filename, const std::string& mtl_basepath) {
    rt_obj* o = nullptr;
    char err[512] = {0};
    rt_status s = rt_obj_load(filename.c_str(), mtl_basepath.c_str(), &o, err, sizeof(err));
    if (s != RT_OK) throw Error(s, err);
    int32_t nv = 0, n = 0, nmat = 0;
    rt_obj_counts(o, &nv, &n, &nmat, nullptr, nullptr, nullptr);
    vx.resize(nv), vy.resize(nv), vz.resize(nv), nx.resize(nv), ny.resize(nv), nz.resize(nv);
    idx.resize(3 * (size_t)n);
    material_ids.resize(n);
    materials.resize(nmat);
    rt_obj_mesh(o, vx.data(), vy.data(), vz.data(), nx.data(), ny.data(), nz.data(), idx.data(),
                material_ids.data());
    rt_obj_materials(o, materials.data());
    rt_obj_free(o);
  }
  void setCamera(Camera* c) { camera = c; }
  std::vector<float> getImage() const { return image; }
  int32_t numTriangles() const { return (int32_t)(idx.size() / 3); }

  std::vector<float> vx, vy, vz, nx, ny, nz;
  std::vector<int32_t> idx, material_ids;
  std::vector<rt_material> materials;
  int imageWidth, imageHeight, superSampling;
  Camera* camera = nullptr;
  std::vector<float> image;  // width*height*3, row-flipped, filled by RayTracer::pullRaytracedImage
};

// Per-triangle boxes on the device (the reference's BoundingBox* BBoxs,
// kernels/scene.cu:48-57): lo and hi, n*3 floats each, in triangle order.
struct BBoxes {
  float* lo_dev = nullptr;
  float* hi_dev = nullptr;
};

// Device BVH (inc/bvh.cuh:31-71): owns the LBVH. Built either the
// reference's way, setUp(..., mMin, mMax, ...) running computeMortonCodes,
// sortMortonCodes and buildTree (src/bvh.cpp:3-32), each a stage call of the
// C-ABI, or in one call (setUp(ctx, mesh, bits), rt_bvh_build).
class BVH_d {
 public:
  BVH_d() = default;
  BVH_d(const BVH_d&) = delete;
  BVH_d& operator=(const BVH_d&) = delete;
  ~BVH_d() { destroy(); }

  // One stream-ordered build (rt_bvh_build).
  void setUp(rt_ctx* ctx, const rt_mesh_view& mesh, int morton_bits = 30, void* stream = nullptr) {
    destroy();
    ctx_ = ctx;
    rt_build_opts o{morton_bits, RT_BUILD_DEFAULT};
    detail::check(rt_bvh_build(ctx, &mesh, &o, stream, &bvh_), ctx);
    bounds_ready_ = false;
  }

  // BVH_d::setUp(vertices, normals, BBoxs, t_indices, numTriangles,
  // materials, mMin, mMax, material_ids) (inc/bvh.cuh:53, src/bvh.cpp:3-32).
  // The mesh view carries the vertices, normals, triangle indices and
  // material ids. The BVH keeps its own copy of the triangle boxes: it reuses
  // the ones computeBoundingBoxes just made for this mesh, else recomputes
  // them (the same values as BBoxs). The Morton keys use mMin / mMax as
  // given, as the reference's computeMortonCodes does.
  void setUp(rt_ctx* ctx, const rt_mesh_view& mesh, const BBoxes& BBoxs, int numTriangles,
             const rt_material* materials, Vec3d mMin, Vec3d mMax, int morton_bits = 30, void* stream = nullptr) {
    (void)BBoxs;
    if (numTriangles != mesh.n) throw Error(RT_ERR_INVALID, "setUp: numTriangles differs from the mesh");
    materials_ = materials;
    stream_ = stream;
    ensure(ctx, mesh, morton_bits);
    if (!bounds_ready_ || !same_mesh(mesh)) {
      detail::check(rt_bvh_compute_bounds(bvh_, &mesh, nullptr, nullptr, stream), ctx_);
      mesh_ = mesh;
    }
    computeMortonCodes(mMin, mMax);
    sortMortonCodes();
    setupLeafNodes();
    buildTree();
  }
  // The triangle boxes of `mesh` into this BVH and into `out` (if its
  // pointers are set), with the scene bounds for Scene_d::findMinMax.
  void computeBoundingBoxes(rt_ctx* ctx, const rt_mesh_view& mesh, const BBoxes& out, int morton_bits = 30,
                            void* stream = nullptr) {
    stream_ = stream;
    ensure(ctx, mesh, morton_bits);
    detail::check(rt_bvh_compute_bounds(bvh_, &mesh, out.lo_dev, out.hi_dev, stream), ctx_);
    mesh_ = mesh;
    bounds_ready_ = true;
  }
  // kernels/bvh.cu:260-265.
  void computeMortonCodes(Vec3d& mMin, Vec3d& mMax) {
    const double mn[3] = {mMin.x, mMin.y, mMin.z}, mx[3] = {mMax.x, mMax.y, mMax.z};
    detail::check(rt_bvh_compute_morton(bvh_, mn, mx, stream_), ctx_);
  }
  // kernels/bvh.cu:267-273.
  void sortMortonCodes() { detail::check(rt_bvh_sort_morton(bvh_, stream_), ctx_); }
  // kernels/bvh.cu:275-277: does nothing in the reference either.
  void setupLeafNodes() {}
  // kernels/bvh.cu:278-292.
  void buildTree() {
    detail::check(rt_bvh_build_tree(bvh_, &mesh_, stream_), ctx_);
    bounds_ready_ = false;  // a later setUp recomputes the boxes (the vertices may have moved)
  }

  void rebuild(const rt_mesh_view& mesh, void* stream = nullptr) {
    detail::check(rt_bvh_rebuild(bvh_, &mesh, stream), ctx_);
  }
  void refit(const rt_mesh_view& mesh, void* stream = nullptr) {
    detail::check(rt_bvh_refit(bvh_, &mesh, stream), ctx_);
  }
  // Scene bounds of the last build or computeBoundingBoxes (synchronises).
  void sceneBounds(Vec3d& mn, Vec3d& mx, void* stream = nullptr) const {
    double a[3], b[3];
    detail::check(rt_bvh_scene_bounds(bvh_, stream, a, b), ctx_);
    mn = Vec3d(a[0], a[1], a[2]);
    mx = Vec3d(b[0], b[1], b[2]);
  }
  // Batched forms of the per-ray __device__ intersect / intersectAny
  // (kernels/bvh.cu:382-449).
  void intersect(const rt_mesh_view& mesh, const rt_rays& rays, const rt_hits& hits, void* stream = nullptr) const {
    detail::check(rt_trace(bvh_, &mesh, &rays, RT_TRACE_CLOSEST, 0.0, &hits, stream), ctx_);
  }
  void intersectAny(const rt_mesh_view& mesh, const rt_rays& rays, const rt_hits& hits, void* stream = nullptr) const {
    detail::check(rt_trace(bvh_, &mesh, &rays, RT_TRACE_ANY, 0.0, &hits, stream), ctx_);
  }
  void exportArrays(const rt_bvh_export_bufs& dst, void* stream = nullptr) const {
    detail::check(rt_bvh_export(bvh_, &dst, stream), ctx_);
  }
  rt_bvh* handle() const { return bvh_; }
  const rt_material* materials() const { return materials_; }

 private:
  void ensure(rt_ctx* ctx, const rt_mesh_view& mesh, int morton_bits) {
    if (bvh_ && ctx == ctx_ && rt_bvh_num_triangles(bvh_) == mesh.n && bits_ == morton_bits) return;
    destroy();
    ctx_ = ctx;
    bits_ = morton_bits;
    rt_build_opts o{morton_bits, RT_BUILD_DEFAULT};
    detail::check(rt_bvh_create(ctx, mesh.n, &o, &bvh_), ctx);
  }
  bool same_mesh(const rt_mesh_view& m) const {
    return m.vx_dev == mesh_.vx_dev && m.vy_dev == mesh_.vy_dev && m.vz_dev == mesh_.vz_dev &&
           m.idx_dev == mesh_.idx_dev && m.n == mesh_.n;
  }
  void destroy() {
    if (bvh_) rt_bvh_destroy(bvh_);
    bvh_ = nullptr;
    bounds_ready_ = false;
  }
  rt_ctx* ctx_ = nullptr;
  rt_bvh* bvh_ = nullptr;
  int bits_ = 30;
  rt_mesh_view mesh_{};
  const rt_material* materials_ = nullptr;
  void* stream_ = nullptr;
  bool bounds_ready_ = false;
};

// Device scene (inc/scene.h:60-117).
class Scene_d {
 public:
  explicit Scene_d(rt_ctx* ctx) : ctx_(ctx) {}

  // Scene_d::operator=(const Scene_h&) (src/scene.cpp:82-161): upload the
  // mesh, then computeBoundingBoxes, findMinMax and bvh.setUp, in the
  // reference's order.
  Scene_d& operator=(const Scene_h& h) {
    numTriangles = h.numTriangles();
    numVertices = (int)h.vx.size();
    numMaterials = (int)h.materials.size();
    imageWidth = h.imageWidth * h.superSampling;
    imageHeight = h.imageHeight * h.superSampling;
    if (numTriangles <= 0) throw Error(RT_ERR_INVALID, "scene has no triangles");
    vx_.upload(h.vx), vy_.upload(h.vy), vz_.upload(h.vz);
    nx_.upload(h.nx), ny_.upload(h.ny), nz_.upload(h.nz);
    idx_.upload(h.idx);
    mat_id_.upload(h.material_ids);
    if (h.materials.empty()) {
      rt_material m{};
      for (int k = 0; k < 3; ++k) m.kd[k] = 0.8;
      mats_.upload(&m, 1);
      numMaterials = 1;
    } else {
      mats_.upload(h.materials);
    }
    computeBoundingBoxes();
    Vec3d mMin, mMax;
    findMinMax(mMin, mMax);
    bvh.setUp(ctx_, mesh(), BBoxs(), numTriangles, materials(), mMin, mMax, morton_bits);
    return *this;
  }
  // kernels/scene.cu:48-57: per-triangle boxes into BBoxs().
  void computeBoundingBoxes() {
    bbox_lo_.resize(3 * (size_t)numTriangles);
    bbox_hi_.resize(3 * (size_t)numTriangles);
    bvh.computeBoundingBoxes(ctx_, mesh(), BBoxs(), morton_bits);
  }
  // kernels/scene.cu:72-85: the bounds of the boxes (+-1e9 start values).
  void findMinMax(Vec3d& mn, Vec3d& mx) const { bvh.sceneBounds(mn, mx); }
  rt_mesh_view mesh() const {
    rt_mesh_view m{};
    m.vx_dev = vx_.get(), m.vy_dev = vy_.get(), m.vz_dev = vz_.get();
    m.nx_dev = nx_.get(), m.ny_dev = ny_.get(), m.nz_dev = nz_.get();
    m.idx_dev = idx_.get();
    m.mat_id_dev = mat_id_.size() ? mat_id_.get() : nullptr;
    m.nv = numVertices;
    m.n = numTriangles;
    return m;
  }
  BBoxes BBoxs() const { return BBoxes{bbox_lo_.get(), bbox_hi_.get()}; }
  const rt_material* materials() const { return mats_.get(); }

  int numVertices = 0, numTriangles = 0, numMaterials = 0, imageWidth = 0, imageHeight = 0;
  int morton_bits = 30;
  BVH_d bvh;

 private:
  rt_ctx* ctx_;
  detail::DevBuf<float> vx_, vy_, vz_, nx_, ny_, nz_, bbox_lo_, bbox_hi_;
  detail::DevBuf<int32_t> idx_, mat_id_;
  detail::DevBuf<rt_material> mats_;
};

// Owns an rt_ctx (declared first in RayTracer, so it outlives the scene).
class Context {
 public:
  explicit Context(int device = 0) {
    rt_status s = rt_ctx_create(device, &ctx_);
    if (s != RT_OK) throw Error(s, "rt_ctx_create failed");
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  ~Context() {
    if (ctx_) rt_ctx_destroy(ctx_);
  }
  rt_ctx* get() const { return ctx_; }

 private:
  rt_ctx* ctx_ = nullptr;
};

// RayTracer (inc/raytracer.h:20-64, kernels/raytracer.cu:55-166).
class RayTracer {
 public:
  RayTracer(int imageWidth = 512, int imageHeight = 512, int superSampling = 1, int device = 0)
      : ctx_(device), hostScene(imageWidth, imageHeight, superSampling), deviceScene(ctx_.get()) {
    opts_.spp = 20;
    opts_.depth = 0;
    opts_.seed = 0;
    opts_.flags = RT_RENDER_WAVEFRONT;
  }
  RayTracer(const RayTracer&) = delete;
  RayTracer& operator=(const RayTracer&) = delete;

  void LoadObj(const std::string& filename, const std::string& mtl_basepath) {
    hostScene.LoadObj(filename, mtl_basepath);
  }
  void setCamera(Camera* c) { hostScene.setCamera(c); }
  void setHostLight(const Light_h& l) { hostLight = l; }
  void setUpDevice() { deviceScene = hostScene; }

  // Render every pixel of the (supersampled) image, then resolve.
  void run(void* stream = nullptr) {
    if (!hostScene.camera) throw Error(RT_ERR_INVALID, "no camera set");
    const int W = hostScene.imageWidth, H = hostScene.imageHeight, ss = hostScene.superSampling;
    rt_camera cam = hostScene.camera->desc(W * ss, H * ss);
    rt_light light = hostLight.desc();
    big_.resize((size_t)W * ss * H * ss * 3);
    small_.resize((size_t)W * H * 3);
    rt_render_opts o = opts_;
    o.flags &= ~RT_RENDER_CLAMP;  // the resolve clamps each sample
    rt_mesh_view m = deviceScene.mesh();
    rt_ctx* c = ctx_.get();
    detail::check(rt_render(c, deviceScene.bvh.handle(), &m, deviceScene.materials(), deviceScene.numMaterials,
                            &light, &cam, &o, big_.get(), nullptr, stream),
                  c);
    detail::check(rt_resolve_supersample(c, big_.get(), W, H, ss, resolve_flags_, small_.get(), stream), c);
  }
  void pullRaytracedImage() { small_.download(hostScene.image); }
  void writeImage(const std::string& filename) const {
    rt_status s = rt_write_bmp(filename.c_str(), hostScene.imageWidth, hostScene.imageHeight, hostScene.image.data());
    if (s != RT_OK) throw Error(s, "cannot write " + filename);
  }

  void setSamples(int spp) { opts_.spp = spp; }
  void setSeed(uint32_t seed) { opts_.seed = seed; }
  void setDepth(int depth) { opts_.depth = depth; }
  void setRenderFlags(int flags) { opts_.flags = flags; }
  void setResolveFlags(int flags) { resolve_flags_ = flags; }
  void setMortonBits(int bits) { deviceScene.morton_bits = bits; }
  const std::vector<float>& image() const { return hostScene.image; }
  rt_ctx* context() const { return ctx_.get(); }

 private:
  Context ctx_;

 public:
  Scene_h hostScene;
  Light_h hostLight;
  Scene_d deviceScene;

 private:
  rt_render_opts opts_{};
  int resolve_flags_ = 0;
  detail::DevBuf<float> big_, small_;
};

}  // namespace rtamd

// A caller written the way the reference drives its build, compiled against
// include/rt_facade.hpp: the Scene_h -> Scene_d assignment
// (src/scene.cpp:82-161), the same calls spelled out (computeBoundingBoxes,
// findMinMax, BVH_d::setUp with mMin / mMax, src/bvh.cpp:3-32), the stage
// methods called one by one, and the one-call build. It checks that these
// routes give the same tree, then writes, for the test to compare with the
// oracle, the tree built with caller-chosen Morton bounds and the per-triangle
// boxes of computeBoundingBoxes.
//
// Usage: facade_caller N SEED OUTDIR   (a random soup of N triangles, the
// generator rtamd.gen_soup uses, with its default shape parameters)
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "rt_facade.hpp"
#include "rt_scenes.h"

using namespace rtamd;

namespace {

struct Tree {
  std::vector<uint32_t> codes, ids;
  std::vector<int32_t> parent, left, right;
  std::vector<float> lo, hi;
};

Tree pull(const BVH_d& b, int n) {
  const size_t nn = 2 * (size_t)n - 1;
  detail::DevBuf<uint32_t> codes, ids;
  detail::DevBuf<int32_t> parent, left, right;
  detail::DevBuf<float> lo, hi;
  codes.resize(n), ids.resize(n), parent.resize(nn), left.resize(nn), right.resize(nn);
  lo.resize(3 * nn), hi.resize(3 * nn);
  rt_bvh_export_bufs e{};
  e.codes_dev = codes.get(), e.ids_dev = ids.get(), e.parent_dev = parent.get();
  e.left_dev = left.get(), e.right_dev = right.get(), e.lo_dev = lo.get(), e.hi_dev = hi.get();
  b.exportArrays(e);
  detail::hip_check(hipDeviceSynchronize(), "export");
  Tree t;
  codes.download(t.codes), ids.download(t.ids), parent.download(t.parent), left.download(t.left);
  right.download(t.right), lo.download(t.lo), hi.download(t.hi);
  return t;
}

template <class T>
bool same_bits(const std::vector<T>& a, const std::vector<T>& b) {
  return a.size() == b.size() && memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0;
}

bool same(const Tree& a, const Tree& b) {
  return same_bits(a.codes, b.codes) && same_bits(a.ids, b.ids) && same_bits(a.parent, b.parent) &&
         same_bits(a.left, b.left) && same_bits(a.right, b.right) && same_bits(a.lo, b.lo) &&
         same_bits(a.hi, b.hi);
}

template <class T>
void write(const std::string& dir, const char* name, const T* p, size_t count) {
  std::string path = dir + "/" + name;
  FILE* f = fopen(path.c_str(), "wb");
  if (!f || fwrite(p, sizeof(T), count, f) != count) {
    fprintf(stderr, "cannot write %s\n", path.c_str());
    exit(2);
  }
  fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s N SEED OUTDIR\n", argv[0]);
    return 2;
  }
  const int n = atoi(argv[1]);
  const uint64_t seed = strtoull(argv[2], nullptr, 10);
  const std::string out = argv[3];
  if (n < 1) return 2;
  try {
    Scene_h h(64, 64, 1);
    const size_t nv = 3 * (size_t)n;
    h.vx.resize(nv), h.vy.resize(nv), h.vz.resize(nv), h.nx.resize(nv), h.ny.resize(nv), h.nz.resize(nv);
    h.idx.resize(3 * (size_t)n);
    const double offset[3] = {-0.5, -0.5, -2.0};
    rt_gen_soup(n, seed, 0.002, 0.5, offset, 1.0, h.vx.data(), h.vy.data(), h.vz.data(), h.nx.data(), h.ny.data(),
                h.nz.data(), h.idx.data());
    h.material_ids.assign(n, 0);
    rt_material m{};
    for (int k = 0; k < 3; ++k) m.kd[k] = 0.8;
    h.materials.assign(1, m);

    Context ctx(0);
    // 1: the reference's assignment (upload, computeBoundingBoxes,
    //    findMinMax, bvh.setUp)
    Scene_d d(ctx.get());
    d = h;
    const Tree t1 = pull(d.bvh, n);
    // 2: the same calls made by the caller
    d.computeBoundingBoxes();
    Vec3d mMin, mMax;
    d.findMinMax(mMin, mMax);
    BVH_d b2;
    b2.setUp(ctx.get(), d.mesh(), d.BBoxs(), d.numTriangles, d.materials(), mMin, mMax);
    const Tree t2 = pull(b2, n);
    // 3: one call
    BVH_d b3;
    b3.setUp(ctx.get(), d.mesh(), 30);
    const Tree t3 = pull(b3, n);
    if (!same(t1, t2) || !same(t1, t3)) {
      fprintf(stderr, "routes differ: assignment vs setUp %d, assignment vs one call %d\n", (int)same(t1, t2),
              (int)same(t1, t3));
      return 1;
    }
    // 4: Morton bounds the caller chose (wider than the scene, not fp32
    //    values), through setUp and through the stage methods one by one
    Vec3d cmin(mMin.x - 0.25, mMin.y - 1.0 / 3.0, mMin.z - 0.1);
    Vec3d cmax(mMax.x + 0.5, mMax.y + 0.2, mMax.z + 1.0 / 7.0);
    BVH_d b4;
    b4.setUp(ctx.get(), d.mesh(), d.BBoxs(), d.numTriangles, d.materials(), cmin, cmax);
    const Tree t4 = pull(b4, n);
    b4.computeBoundingBoxes(ctx.get(), d.mesh(), BBoxes{});
    b4.computeMortonCodes(cmin, cmax);
    b4.sortMortonCodes();
    b4.setupLeafNodes();
    b4.buildTree();
    if (!same(t4, pull(b4, n))) {
      fprintf(stderr, "setUp and the stage methods differ\n");
      return 1;
    }
    if (same(t4, t1)) {
      fprintf(stderr, "the caller's bounds did not change the keys\n");
      return 1;
    }
    write(out, "codes.bin", t4.codes.data(), t4.codes.size());
    write(out, "ids.bin", t4.ids.data(), t4.ids.size());
    write(out, "parent.bin", t4.parent.data(), t4.parent.size());
    write(out, "left.bin", t4.left.data(), t4.left.size());
    write(out, "right.bin", t4.right.data(), t4.right.size());
    write(out, "lo.bin", t4.lo.data(), t4.lo.size());
    write(out, "hi.bin", t4.hi.data(), t4.hi.size());
    const double bounds[6] = {cmin.x, cmin.y, cmin.z, cmax.x, cmax.y, cmax.z};
    write(out, "bounds.bin", bounds, 6);
    const double sb[6] = {mMin.x, mMin.y, mMin.z, mMax.x, mMax.y, mMax.z};
    write(out, "scene_bounds.bin", sb, 6);
    std::vector<float> blo(3 * (size_t)n), bhi(3 * (size_t)n);
    detail::hip_check(hipMemcpy(blo.data(), d.BBoxs().lo_dev, blo.size() * 4, hipMemcpyDeviceToHost), "BBoxs");
    detail::hip_check(hipMemcpy(bhi.data(), d.BBoxs().hi_dev, bhi.size() * 4, hipMemcpyDeviceToHost), "BBoxs");
    write(out, "bbox_lo.bin", blo.data(), blo.size());
    write(out, "bbox_hi.bin", bhi.data(), bhi.size());
  } catch (const Error& e) {
    fprintf(stderr, "rtamd error %d: %s\n", (int)e.status, e.what());
    return 1;
  }
  printf("routes agree: assignment, setUp with mMin/mMax, stage methods, one call (n=%d)\n", n);
  return 0;
}

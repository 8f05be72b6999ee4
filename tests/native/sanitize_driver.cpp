// Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the CPU
// oracle (build stages, refit in shuffled orders, traversal, render with
// bounces and records), the OBJ/MTL loader, the BMP writer and the scene
// generators, on small inputs and the edge cases the build has (one and two
// triangles, duplicate keys, a flat axis, NaN and infinite vertices, an
// empty pixel list), plus malformed OBJ text. Exits non-zero on the first
// inconsistency; the sanitizers abort on the first memory or UB error.
//
// Usage: sanitize_driver OBJ_PATH MTL_DIR SCRATCH_DIR
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/rt_io.h"
#include "../../include/rt_scenes.h"
#include "../../oracle/rt_oracle.h"

namespace {

int g_fail = 0;
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      fprintf(stderr, "check failed at %s:%d: %s\n", __FILE__, __LINE__, #c); \
      g_fail = 1;                                                       \
    }                                                                   \
  } while (0)

struct Mesh {
  std::vector<float> vx, vy, vz, nx, ny, nz;
  std::vector<int32_t> idx, mat;
  int n() const { return (int)(idx.size() / 3); }
  orc_mesh view() const {
    orc_mesh m{};
    m.n = n();
    m.nv = (int32_t)vx.size();
    m.vx = vx.data(), m.vy = vy.data(), m.vz = vz.data();
    m.nx = nx.data(), m.ny = ny.data(), m.nz = nz.data();
    m.idx = idx.data();
    m.mat_id = mat.empty() ? nullptr : mat.data();
    return m;
  }
};

Mesh soup(int n, uint64_t seed) {
  Mesh m;
  const size_t nv = 3 * (size_t)n;
  m.vx.resize(nv), m.vy.resize(nv), m.vz.resize(nv), m.nx.resize(nv), m.ny.resize(nv), m.nz.resize(nv);
  m.idx.resize(3 * (size_t)n);
  const double off[3] = {-0.5, -0.5, -2.0};
  rt_gen_soup(n, seed, 0.002, 0.5, off, 1.0, m.vx.data(), m.vy.data(), m.vz.data(), m.nx.data(), m.ny.data(),
              m.nz.data(), m.idx.data());
  return m;
}

// Refit and brute-force boxes agree bit for bit, except where every
// triangle below a node is NaN on an axis: the refit's fminf/fmaxf unions
// keep the NaN, the brute force keeps its +-infinity start value.
bool same_boxes(const std::vector<float>& refit, const std::vector<float>& brute) {
  for (size_t k = 0; k < refit.size(); ++k) {
    if (isnan(refit[k]) && isinf(brute[k])) continue;
    if (memcmp(&refit[k], &brute[k], 4) != 0) return false;
  }
  return true;
}

void exercise(const Mesh& mesh, const char* what) {
  const orc_mesh m = mesh.view();
  const int n = m.n;
  const size_t nn = 2 * (size_t)n - 1;
  for (int bits : {30, 63}) {
    for (int threads : {1, 3}) {
      std::vector<uint64_t> codes(n);
      std::vector<uint32_t> ids(n);
      std::vector<int32_t> parent(nn), left(nn), right(nn);
      std::vector<float> lo(3 * nn), hi(3 * nn);
      double mn[3], mx[3];
      orc_build(&m, bits, codes.data(), ids.data(), parent.data(), left.data(), right.data(), lo.data(), hi.data(),
                mn, mx, threads);
      // the stages one by one, and the top-down definition of the tree
      std::vector<float> tlo(3 * (size_t)n), thi(3 * (size_t)n);
      orc_tri_aabb(&m, tlo.data(), thi.data());
      double smn[3], smx[3];
      orc_scene_bounds(n, tlo.data(), thi.data(), smn, smx);
      std::vector<int32_t> p2(nn), l2(nn), r2(nn), p3(nn), l3(nn), r3(nn);
      std::vector<uint32_t> sid(n);
      if (bits == 30) {
        std::vector<uint32_t> c(n);
        orc_morton30(n, tlo.data(), thi.data(), smn, smx, c.data());
        orc_sort30(n, c.data(), sid.data());
        orc_hierarchy30(n, c.data(), sid.data(), p2.data(), l2.data(), r2.data());
        orc_topdown30(n, c.data(), sid.data(), p3.data(), l3.data(), r3.data());
      } else {
        std::vector<uint64_t> c(n);
        orc_morton63(n, tlo.data(), thi.data(), smn, smx, c.data());
        orc_sort63(n, c.data(), sid.data());
        orc_hierarchy63(n, c.data(), sid.data(), p2.data(), l2.data(), r2.data());
        orc_topdown63(n, c.data(), sid.data(), p3.data(), l3.data(), r3.data());
      }
      CHECK(left == l2 && right == r2 && l2 == l3 && r2 == r3);
      // refit with the leaves in a shuffled order, and by brute force
      std::vector<int32_t> perm(n);
      std::iota(perm.begin(), perm.end(), 0);
      std::shuffle(perm.begin(), perm.end(), std::mt19937(17 + n));
      std::vector<float> lo2(3 * nn), hi2(3 * nn), lo3(3 * nn), hi3(3 * nn);
      orc_refit(n, p2.data(), l2.data(), r2.data(), tlo.data(), thi.data(), perm.data(), lo2.data(), hi2.data());
      orc_refit_brute(n, l2.data(), r2.data(), tlo.data(), thi.data(), lo3.data(), hi3.data());
      CHECK(same_boxes(lo2, lo3) && same_boxes(hi2, hi3));
      // traversal
      orc_bvh b{n, parent.data(), left.data(), right.data(), lo.data(), hi.data()};
      const int64_t nr = 300;
      std::vector<double> o(3 * nr), d(3 * nr), t(nr), al(nr), be(nr), nrm(3 * nr), tb(nr);
      std::vector<int32_t> tri(nr), trib(nr), leafpos(n);
      std::vector<int64_t> visits(2 * nr);
      std::mt19937 rng(5);
      std::uniform_real_distribution<double> u(-0.6, 0.6);
      for (int64_t r = 0; r < nr; ++r) {
        o[3 * r] = u(rng) * 0.01, o[3 * r + 1] = u(rng) * 0.01, o[3 * r + 2] = 0;
        d[3 * r] = u(rng), d[3 * r + 1] = u(rng), d[3 * r + 2] = -1;
      }
      for (int any = 0; any < 2; ++any)
        orc_trace(&m, &b, nr, o.data(), d.data(), any, tri.data(), t.data(), al.data(), be.data(), nrm.data(),
                  visits.data(), threads);
      for (int p = 0; p < n; ++p) leafpos[ids[p]] = p;
      orc_trace(&m, &b, nr, o.data(), d.data(), 0, tri.data(), t.data(), al.data(), be.data(), nrm.data(),
                visits.data(), threads);
      orc_trace_brute(&m, leafpos.data(), nr, o.data(), d.data(), trib.data(), tb.data());
      CHECK(tri == trib);
      // render: every pixel, a pixel list, an empty one; bounces with records
      orc_material mat{};
      for (int k = 0; k < 3; ++k) mat.kd[k] = 0.8, mat.ks[k] = 0.1;
      mat.shininess = 8;
      orc_light light{{1.0, 2.0, 0.5}, {1, 1, 1}, 0.5, 0.0, 0.01, 0.001};
      orc_camera cam{12, 8, 40.0, 20.0, 1.5, {0, 0, 0}};
      orc_render_opts opts{2, 2, 3, ORC_CLAMP};
      std::vector<float> rgb(12 * 8 * 3);
      std::vector<double> rgb64(12 * 8 * 3);
      std::vector<int32_t> rec(12 * 8 * 2 * 3 * 2), br(12 * 8 * 2 * 3);
      int64_t counts[3] = {0, 0, 0};
      orc_render(&m, &b, &mat, 1, &light, &cam, &opts, rgb.data(), nullptr, 0, counts, threads);
      CHECK(counts[0] == 12 * 8 * 2);
      const int64_t pix[3] = {0, 17, 95};
      orc_render64(&m, &b, &mat, 1, &light, &cam, &opts, rgb.data(), rgb64.data(), rec.data(), br.data(), pix, 3,
                   counts, threads);
      orc_render64(&m, &b, &mat, 1, &light, &cam, &opts, rgb.data(), rgb64.data(), rec.data(), br.data(), pix, 0,
                   counts, threads);
    }
  }
  fprintf(stderr, "%s: n=%d ok\n", what, n);
}

void write_text(const std::string& path, const char* text) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) {
    fprintf(stderr, "cannot write %s\n", path.c_str());
    exit(2);
  }
  fputs(text, f);
  fclose(f);
}

void load_obj(const std::string& path, const std::string& dir, bool must_load) {
  rt_obj* o = nullptr;
  char err[256] = {0};
  rt_status s = rt_obj_load(path.c_str(), dir.c_str(), &o, err, sizeof(err));
  if (s != RT_OK) {
    CHECK(!must_load);
    return;
  }
  int32_t nv = 0, n = 0, nmat = 0, nsh = 0, rv = 0, rvn = 0;
  rt_obj_counts(o, &nv, &n, &nmat, &nsh, &rv, &rvn);
  Mesh m;
  m.vx.resize(nv), m.vy.resize(nv), m.vz.resize(nv), m.nx.resize(nv), m.ny.resize(nv), m.nz.resize(nv);
  m.idx.resize(3 * (size_t)n), m.mat.resize(n);
  std::vector<double> p64(3 * (size_t)nv);
  std::vector<rt_material> mats(nmat);
  rt_obj_mesh(o, m.vx.data(), m.vy.data(), m.vz.data(), m.nx.data(), m.ny.data(), m.nz.data(), m.idx.data(),
              m.mat.data());
  rt_obj_positions64(o, p64.data());
  rt_obj_materials(o, mats.data());
  rt_obj_free(o);
  for (int32_t k = 0; k < 3 * n; ++k) CHECK(m.idx[k] >= 0 && m.idx[k] < nv);
  if (n > 0 && n <= 50000) {
    const orc_mesh v = m.view();
    std::vector<float> lo(3 * (size_t)n), hi(3 * (size_t)n);
    orc_tri_aabb(&v, lo.data(), hi.data());
  }
  fprintf(stderr, "%s: %d triangles, %d materials\n", path.c_str(), n, nmat);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    fprintf(stderr, "usage: %s OBJ_PATH MTL_DIR SCRATCH_DIR\n", argv[0]);
    return 2;
  }
  const std::string scratch = argv[3];
  for (int n : {1, 2, 3, 64, 700}) exercise(soup(n, 1000 + n), "soup");
  {
    int nv = 0, n = 0;
    rt_gen_icosphere_counts(3, &nv, &n);
    Mesh m;
    m.vx.resize(nv), m.vy.resize(nv), m.vz.resize(nv), m.nx.resize(nv), m.ny.resize(nv), m.nz.resize(nv);
    m.idx.resize(3 * (size_t)n);
    const double c[3] = {0, 0, -3};
    rt_gen_icosphere(3, 1.0, c, 0.05, 9, m.vx.data(), m.vy.data(), m.vz.data(), m.nx.data(), m.ny.data(),
                     m.nz.data(), m.idx.data());
    rt_gen_jitter(nv, m.vx.data(), m.vy.data(), m.vz.data(), 1e-3, 4, m.vx.data(), m.vy.data(), m.vz.data());
    exercise(m, "icosphere");
  }
  {
    Mesh m = soup(300, 7);  // duplicate keys: 100 copies of one triangle
    for (int t = 1; t < 100; ++t)
      for (int k = 0; k < 3; ++k) {
        m.vx[3 * t + k] = m.vx[k], m.vy[3 * t + k] = m.vy[k], m.vz[3 * t + k] = m.vz[k];
      }
    exercise(m, "duplicates");
    for (float& z : m.vz) z = -1.5f;  // flat along z: (c - min) / 0
    exercise(m, "flat axis");
    m.vx[5] = NAN;
    m.vy[30] = m.vy[31] = m.vy[32] = NAN;
    m.vz[60] = INFINITY;
    m.vx[90] = -INFINITY;
    exercise(m, "non-finite vertices");
  }
  load_obj(argv[1], argv[2], true);
  const std::string bad = scratch + "/bad.obj";
  write_text(bad,
             "v 1 2\nv 1e999 -1e-999 .5\nv +3 4e2 -.25e-1\nvn 0 0 1\nvt 0.5\n"
             "f 1 2 3\nf -1 -2 -3 -4\nf 1//1 2//1 3//1\nf 1/1/1 2/1/1 3/1/1\nf 7 8 9\nf 1 2\nf\n"
             "usemtl nothing\nmtllib missing.mtl\ng\ns off\no x\nf 1 2 3 1 2 3 1\n# end");
  load_obj(bad, scratch, false);
  write_text(bad, "");
  load_obj(bad, scratch, false);
  load_obj(scratch + "/does_not_exist.obj", scratch, false);
  std::vector<float> img(5 * 3 * 3);
  for (size_t k = 0; k < img.size(); ++k) img[k] = (float)k / (float)img.size() * 1.5f - 0.2f;
  CHECK(rt_write_bmp((scratch + "/img.bmp").c_str(), 5, 3, img.data()) == RT_OK);
  CHECK(rt_write_bmp((scratch + "/no/such/dir/img.bmp").c_str(), 5, 3, img.data()) != RT_OK);
  if (g_fail) return 1;
  printf("sanitized host code: all checks passed\n");
  return 0;
}

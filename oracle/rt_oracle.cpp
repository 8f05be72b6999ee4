// CPU oracle: double-precision restatement of the reference LBVH build,
// refit, traversal and thin-lens render (mbartling/cuda-rt).
//
// TEST INFRASTRUCTURE ONLY (checker and CPU baseline). See rt_oracle.h.
// Built with -ffp-contract=off so that every floating-point expression is
// evaluated exactly in the order written, like the reference's FP64 code.
#include "rt_oracle.h"
#include "rtmath_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V3 {
  double x, y, z;
};
static inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
static inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 scale(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }   // inc/vec.h:72-89
static inline V3 mulv(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }  // inc/vec.h:106 (%)
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }  // inc/vec.h:56
static inline V3 cross(V3 a, V3 b) {                                                // inc/vec.h:111
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// inc/vec.h:118-121: the length goes through sqrtf (float).
static inline V3 normalize(V3 a) {
  double sqinv = 1.0 / (double)sqrtf((float)(a.x * a.x + a.y * a.y + a.z * a.z));
  return scale(a, sqinv);
}
// inc/vec.h:137-140
static inline double norm(V3 a) { return (double)sqrtf((float)(a.x * a.x + a.y * a.y + a.z * a.z)); }
// inc/vec.h:158-165: minimum/maximum round both operands to float.
static inline double fmin_f(double a, double b) { return (double)fminf((float)a, (float)b); }
static inline double fmax_f(double a, double b) { return (double)fmaxf((float)a, (float)b); }

static const double kRayEps = 0.00000001;  // inc/common.h:2

static inline V3 vertex(const orc_mesh *m, int32_t i) {
  if (m->pos64) return v3(m->pos64[3 * (size_t)i], m->pos64[3 * (size_t)i + 1], m->pos64[3 * (size_t)i + 2]);
  return v3(m->vx[i], m->vy[i], m->vz[i]);
}
static inline V3 vnormal(const orc_mesh *m, int32_t i) { return v3(m->nx[i], m->ny[i], m->nz[i]); }

static inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }
static inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

// kernels/bvh.cu:72-88 (30-bit) and its 64-bit extension (fallback 64+clz).
template <typename K>
struct Delta;
template <>
struct Delta<uint32_t> {
  static int f(const uint32_t *k, int n, int i1, int i2) {
    int l = std::min(i1, i2), r = std::max(i1, i2);
    if (l < 0 || r >= n) return -1;
    uint32_t a = k[l], b = k[r];
    return a != b ? clz32(a ^ b) : 32 + clz32((uint32_t)(l ^ r));
  }
};
template <>
struct Delta<uint64_t> {
  static int f(const uint64_t *k, int n, int i1, int i2) {
    int l = std::min(i1, i2), r = std::max(i1, i2);
    if (l < 0 || r >= n) return -1;
    uint64_t a = k[l], b = k[r];
    return a != b ? clz64(a ^ b) : 64 + clz32((uint32_t)(l ^ r));
  }
};

// kernels/bvh.cu:92-99
static int sign_of(float x) {
  if (x > 0) return 1;
  if (x == 0.0f) return 0;
  if (x < 0) return -1;
  return 0;
}

template <typename K>
static void span_of(const K *k, int n, int i, int *first, int *last) {  // kernels/bvh.cu:100-133
  int d = sign_of((float)(Delta<K>::f(k, n, i, i + 1) - Delta<K>::f(k, n, i, i - 1)));
  int dmin = Delta<K>::f(k, n, i, i - d);
  int lmax = 2;
  while (Delta<K>::f(k, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0, t = lmax;
  do {
    t /= 2;
    if (Delta<K>::f(k, n, i, i + (l + t) * d) > dmin) l += t;
  } while (t > 1);
  *first = std::min(i, i + l * d);
  *last = std::max(i, i + l * d);
}

template <typename K>
static int split_of(const K *k, int n, int first, int last) {  // kernels/bvh.cu:134-162
  int common = Delta<K>::f(k, n, first, last);
  int lo = first, hi = last;
  do {
    int mid = (hi + lo) / 2;
    if (Delta<K>::f(k, n, lo, mid) > common)
      lo = mid;
    else
      hi = mid;
  } while (hi > lo + 1);
  return lo;
}

template <typename K>
static void hierarchy(int32_t n, const K *codes, const uint32_t *ids, int32_t *parent, int32_t *left,
                      int32_t *right) {  // kernels/bvh.cu:164-207
  if (n <= 0) return;
  for (int i = 0; i < 2 * n - 1; ++i) parent[i] = -1;
  for (int i = 0; i < n; ++i) {
    left[n - 1 + i] = right[n - 1 + i] = (int32_t)ids[i];
  }
  for (int i = 0; i < n - 1; ++i) {
    int first, last;
    span_of<K>(codes, n, i, &first, &last);
    int s = split_of<K>(codes, n, first, last);
    int c1 = (s == first) ? (n - 1 + s) : s;
    int c2 = (s + 1 == last) ? (n - 1 + s + 1) : s + 1;
    left[i] = c1;
    right[i] = c2;
    parent[c1] = i;
    parent[c2] = i;
  }
}

// Top-down radix tree straight from the definition: a node over sorted
// range [f, l] splits after the last position whose augmented key shares
// more than delta(f, l) leading bits with key f. Children that cover a
// range starting at position i (right child) or ending at i (left child)
// take Karras's index i.
template <typename K>
static void topdown(int32_t n, const K *codes, const uint32_t *ids, int32_t *parent, int32_t *left,
                    int32_t *right) {
  if (n <= 0) return;
  for (int i = 0; i < 2 * n - 1; ++i) parent[i] = -1;
  for (int i = 0; i < n; ++i) left[n - 1 + i] = right[n - 1 + i] = (int32_t)ids[i];
  if (n == 1) return;
  struct Item {
    int node, f, l;
  };
  std::vector<Item> stack;
  stack.push_back({0, 0, n - 1});
  while (!stack.empty()) {
    Item it = stack.back();
    stack.pop_back();
    int common = Delta<K>::f(codes, n, it.f, it.l);
    int s = it.f;  // linear scan: the last j with delta(f, j) > common
    for (int j = it.f + 1; j < it.l; ++j)
      if (Delta<K>::f(codes, n, it.f, j) > common) s = j;
    int c1 = (s == it.f) ? n - 1 + s : s;
    int c2 = (s + 1 == it.l) ? n - 1 + s + 1 : s + 1;
    left[it.node] = c1;
    right[it.node] = c2;
    parent[c1] = it.node;
    parent[c2] = it.node;
    if (s != it.f) stack.push_back({c1, it.f, s});
    if (s + 1 != it.l) stack.push_back({c2, s + 1, it.l});
  }
}

// inc/bbox.h:82-102 (Kay-Kajiya slab test, no t interval).
static inline bool box_hit(const float *lo, const float *hi, const double *e, const double *a) {
  double tmin[3], tmax[3];
  for (int i = 0; i < 3; ++i) {
    if (a[i] >= 0) {
      tmin[i] = ((double)lo[i] - e[i]) * a[i];
      tmax[i] = ((double)hi[i] - e[i]) * a[i];
    } else {
      tmax[i] = ((double)lo[i] - e[i]) * a[i];
      tmin[i] = ((double)hi[i] - e[i]) * a[i];
    }
  }
  double t0 = fmax(fmax(tmin[0], tmin[1]), tmin[2]);
  double t1 = fmin(fmin(tmax[0], tmax[1]), tmax[2]);
  return t0 <= t1;
}

struct Hit {
  double t, alpha, beta;
  V3 N;
  int32_t tri;
};

// kernels/bvh.cu:297-371 (Moller-Trumbore, no culling, |t| >= eps).
static bool tri_hit(const orc_mesh *m, int32_t tri, V3 p, V3 d, Hit *h) {
  const int32_t *ix = m->idx + 3 * (int64_t)tri;
  V3 a = vertex(m, ix[0]), b = vertex(m, ix[1]), c = vertex(m, ix[2]);
  V3 AB = sub(b, a), AC = sub(c, a);
  V3 P = cross(d, AC);
  double det = dot(AB, P);
  if ((double)fabsf((float)det) < kRayEps) return false;
  double inv = 1 / det;
  V3 T = sub(p, a);
  double alpha = dot(T, P) * inv;
  if (alpha < 0 || alpha > 1) return false;
  V3 Q = cross(T, AB);
  double beta = dot(d, Q) * inv;
  if (beta < 0 || alpha + beta > 1) return false;
  double t = dot(AC, Q) * inv;
  if ((double)fabsf((float)t) < kRayEps) return false;
  // Deviation: a NaN or infinite t, alpha or beta (only non-finite vertex or
  // ray coordinates give one) is not a hit. The reference accepts it
  // (kernels/bvh.cu:330-345 compare against NaN), and keeps a NaN t only if
  // it is the first hit its DFS finds (:394), so its answer depends on visit
  // order.
  if (!(fabs(t) < INFINITY && alpha == alpha && beta == beta)) return false;
  h->t = t;
  h->alpha = alpha;
  h->beta = beta;
  V3 N = add(add(scale(vnormal(m, ix[0]), 1.0 - (alpha + beta)), scale(vnormal(m, ix[1]), alpha)),
             scale(vnormal(m, ix[2]), beta));
  h->N = normalize(N);
  h->tri = tri;
  return true;
}

// kernels/bvh.cu:382-449. any: return the first hit in left-first DFS
// order. tmin_phys > 0 selects the bounce-ray rule (only t >= tmin_phys
// counts); 0 keeps the reference rule (any t with |t| >= 1e-8).
static bool traverse(const orc_mesh *m, const orc_bvh *b, V3 p, V3 d, bool any, double tmin_phys, Hit *out,
                     int64_t *visits) {
  const int STACK = 64;
  int stack[STACK];
  int top = STACK;
  double e[3] = {p.x, p.y, p.z};
  double a[3] = {1.0 / d.x, 1.0 / d.y, 1.0 / d.z};
  int n = b->n;
  if (n <= 0) return false;
  stack[--top] = 0;
  bool have = false;
  int64_t nv = 0, nt = 0;
  while (top != STACK) {
    int node = stack[top++];
    ++nv;
    if (!box_hit(b->lo + 3 * (int64_t)node, b->hi + 3 * (int64_t)node, e, a)) continue;
    if (b->left[node] == b->right[node]) {
      Hit cur;
      ++nt;
      if (tri_hit(m, b->left[node], p, d, &cur)) {
        if (tmin_phys > 0 && !(cur.t >= tmin_phys)) continue;
        if (!have || cur.t < out->t) {
          *out = cur;
          have = true;
          if (any) break;
        }
      }
    } else {
      // The reference writes both children before checking for overflow
      // (kernels/bvh.cu:402-407); an overflow reports a miss.
      if (top < 2) {
        have = false;
        break;
      }
      stack[--top] = b->right[node];
      stack[--top] = b->left[node];
    }
  }
  if (visits) {
    visits[0] += nv;
    visits[1] += nt;
  }
  return have;
}

// ---------------------------------------------------------------- mrand
static inline uint32_t xorshift(uint32_t &s) {  // kernels/mrand.cu:3-10
  s ^= (s << 13);
  s ^= (s >> 17);
  s ^= (s << 5);
  return s;
}
static inline uint32_t wang(uint32_t s) {  // kernels/mrand.cu:13-21
  s = (s ^ 61u) ^ (s >> 16);
  s *= 9u;
  s = s ^ (s >> 4);
  s *= 0x27d4eb2du;
  s = s ^ (s >> 15);
  return s;
}
static inline double rand_double(uint32_t &s) {  // kernels/mrand.cu:25-29
  uint32_t v = xorshift(s);
  return (double)((double)(float)v * 1.0 / 4294967296.0);
}
static inline V3 rand_vec(uint32_t &s, bool rtl) {  // kernels/mrand.cu:33-35
  if (!rtl) {
    double x = rand_double(s);
    double y = rand_double(s);
    double z = rand_double(s);
    return v3(x, y, z);
  }
  double z = rand_double(s);
  double y = rand_double(s);
  double x = rand_double(s);
  return v3(x, y, z);
}

static inline uint32_t seed_index(int px, int py, int width) {  // inc/indexing.h:4-9, 16x16 blocks
  int gx = width / 16;
  int bid = (px / 16) + (py / 16) * gx;
  return (uint32_t)(bid * 256 + (py % 16) * 16 + (px % 16));
}

struct RenderCtx {
  const orc_mesh *m;
  const orc_bvh *b;
  const orc_material *mats;
  int nmat;
  const orc_light *L;
  const orc_camera *cam;
  const orc_render_opts *o;
};

static inline V3 arr3(const double *p) { return v3(p[0], p[1], p[2]); }

static const orc_material &material_of(const RenderCtx &c, int32_t tri) {
  int id = c.m->mat_id ? c.m->mat_id[tri] : 0;
  if (id < 0 || id >= c.nmat) id = 0;
  return c.mats[id];
}

// kernels/Light.cu:5-53 (positional light): shadow factor.
// *occ (if given): -1 unshadowed, -2 opaque occluder, else the occluder's triangle.
static V3 shadow_atten(const RenderCtx &c, V3 q, V3 Ldir, int64_t *cnt, int32_t *occ) {
  Hit h;
  ++cnt[1];
  if (occ) *occ = -1;
  if (traverse(c.m, c.b, q, Ldir, true, 0.0, &h, nullptr)) {
    if (h.t < norm(sub(arr3(c.L->position), q))) {
      if (h.t < kRayEps) return v3(1, 1, 1);
      const orc_material &mt = material_of(c, h.tri);
      bool trans = (mt.kt[0] == 0) && (mt.kt[1] == 0) && (mt.kt[2] == 0);  // inc/material.h:31
      if (occ) *occ = trans ? h.tri : -2;
      if (trans) return mulv(arr3(mt.kd), sub(v3(1, 1, 1), arr3(mt.kt)));
      return v3(0, 0, 0);
    }
    return v3(1, 1, 1);
  }
  return v3(1, 1, 1);
}

// kernels/Light.cu:56-68
static double dist_atten(const RenderCtx &c, V3 q, bool *clamped) {
  double d = norm(sub(arr3(c.L->position), q));
  double f = c.L->c0 + (c.L->c1 + c.L->c2 * d) * d;
  f = (fabs(f) < kRayEps) ? 1.0 : 1.0 / f;
  *clamped = !(f < 1.0);
  return fmin(1.0, f);
}

// kernels/material.cu:6-40 with Light::getDirection (kernels/Light.cu:71-81).
// *br (if given) gets bit 0: N.L > 0, bit 1: falloff clamped to 1.
static V3 shade(const RenderCtx &c, V3 p, V3 d, const Hit &h, uint32_t &rng, int64_t *cnt, int32_t *occ,
                int32_t *br) {
  const orc_material &mt = material_of(c, h.tri);
  V3 q = add(p, scale(d, h.t));
  V3 I = add(arr3(mt.ke), mulv(arr3(mt.ka), v3(0, 0, 0)));
  V3 V = sub(arr3(c.cam->eye), q);
  V = v3(-V.x, -V.y, -V.z);
  V = normalize(V);
  V3 off = scale(rand_vec(rng, (c.o->flags & ORC_RANDVEC_RTL) != 0), c.L->radius);
  off = v3(off.x - 0.5, off.y - 0.5, off.z - 0.5);
  if (c.o->flags & ORC_PLANT_LIGHT_OFFSET) {  // planted mistake: (rand - 0.5) * radius
    const double h = 0.5 * c.L->radius;
    off = v3(off.x + 0.5 - h, off.y + 0.5 - h, off.z + 0.5 - h);
  }
  V3 L = normalize(sub(add(arr3(c.L->position), off), q));
  bool clamped;
  V3 atten = scale(shadow_atten(c, q, L, cnt, occ), dist_atten(c, q, &clamped));
  double nl = dot(h.N, L);
  if (br) *br = (nl > 0 ? 1 : 0) | (clamped ? 2 : 0);
  if (nl < 0) nl = 0;
  V3 diffuse = scale(arr3(mt.kd), nl);
  V3 R = add(scale(h.N, -2.0 * dot(L, h.N)), L);
  R = normalize(R);
  double rv = dot(R, V);
  double spec_f = (double)powf((float)std::max(0.0, rv), (float)mt.shininess);
  V3 spec = scale(arr3(mt.ks), spec_f);
  I = add(I, mulv(mulv(atten, add(diffuse, spec)), arr3(c.L->color)));
  return I;
}

// Cosine-weighted hemisphere direction about n, following the frame
// construction of src/ray.cpp:5-32 with theta = acos(sqrt(1-u1)) written
// as sin(theta) = sqrt(u1), cos(theta) = sqrt(1-u1).
static V3 cosine_dir(V3 n, double u1, double u2, int *axis) {
  double sphi, cphi;
  orc_math::sincos(2.0 * 3.141592653589793 * u2, &sphi, &cphi);
  double st = sqrt(u1), ct = sqrt(1.0 - u1);
  double xs = st * cphi, ys = ct, zs = st * sphi;
  V3 h = n;
  float ax = fabsf((float)h.x), ay = fabsf((float)h.y), az = fabsf((float)h.z);
  if (ax <= ay && ax <= az) {
    h.x = 1.0;
    *axis = 0;
  } else if (ay <= ax && ay <= az) {
    h.y = 1.0;
    *axis = 1;
  } else {
    h.z = 1.0;
    *axis = 2;
  }
  V3 x = normalize(cross(h, n));
  V3 z = normalize(cross(x, n));
  V3 dir = add(add(scale(x, xs), scale(n, ys)), scale(z, zs));
  return normalize(dir);
}

static const double kBounceTmin = 1e-6;

// kernels/raytracer.cu:168-249 plus the diffuse-bounce extension. rec (if
// given) receives per bounce (hit triangle or -1, occluder code of shade()),
// br (if given) per bounce the branch bits of shade() plus bit 2: normal
// flipped and bits 3-4: the bounce frame's replaced axis.
static V3 trace_path(const RenderCtx &c, V3 p, V3 d, uint32_t &rng, int64_t *cnt, int32_t *rec, int32_t *br) {
  V3 color = v3(0, 0, 0), thr = v3(1, 1, 1);
  for (int bounce = 0;; ++bounce) {
    Hit h;
    bool hit = traverse(c.m, c.b, p, d, false, bounce == 0 ? 0.0 : kBounceTmin, &h, nullptr);
    if (rec) {
      rec[2 * bounce] = hit ? h.tri : -1;
      rec[2 * bounce + 1] = -1;
    }
    if (!hit) {
      color = add(color, mulv(thr, v3(0.9, 0.9, 0.9)));
      break;
    }
    V3 s = shade(c, p, d, h, rng, cnt, rec ? rec + 2 * bounce + 1 : nullptr, br ? br + bounce : nullptr);
    color = add(color, mulv(thr, s));
    if (bounce >= c.o->depth) break;
    double u1 = rand_double(rng);
    double u2 = rand_double(rng);
    V3 n = h.N;
    const bool flip = dot(n, d) > 0;
    if (flip) n = v3(-n.x, -n.y, -n.z);
    V3 q = add(p, scale(d, h.t));
    const orc_material &mt = material_of(c, h.tri);
    thr = mulv(thr, arr3(mt.kd));
    p = q;
    int axis;
    d = cosine_dir(n, u1, u2, &axis);
    if (br) br[bounce] |= (flip ? 4 : 0) | (axis << 3);
    ++cnt[2];
  }
  return color;
}

// kernels/raytracer.cu:142-154: two draws, thin-lens origin, direction.
static void lens_ray(const orc_camera *cam, int32_t flags, double xx, double yy, uint32_t &rng, V3 *o_out,
                     V3 *d_out) {
  double aperture = 1 / cam->fstop;
  if (flags & ORC_PLANT_APERTURE) aperture *= 1.25;  // planted mistake
  double ang = 2 * M_PI * rand_double(rng);
  double u = rand_double(rng);
  double rad = (flags & ORC_PLANT_LENS_RADIUS) ? aperture * u / 2.0 : aperture * sqrt(u) / 2.0;
  double s, co;
  orc_math::sincos(ang, &s, &co);
  V3 o = v3(rad * co, rad * s, 0);
  V3 d = normalize(v3(xx, yy, -1.0));
  d = sub(scale(d, cam->focal_point), o);
  *o_out = o;
  *d_out = normalize(d);
}

static void pixel_xy(const orc_camera *cam, int px, int py, double *xx, double *yy) {  // kernels/raytracer.cu:126-132
  int W = cam->width, H = cam->height;
  double invW = 1.0 / double(W), invH = 1.0 / double(H);
  double aspect = double(W) / double(H);
  double angle = tan(M_PI * 0.5 * cam->fov_deg / 180.0);
  *xx = (2 * ((px + 0.5) * invW) - 1) * angle * aspect;
  *yy = (1 - 2 * ((py + 0.5) * invH)) * angle;
}

static void render_pixel(const RenderCtx &c, int px, int py, float *rgb, double *rgb64, int32_t *rec,
                         int32_t *branches, int64_t *cnt) {
  const orc_camera *cam = c.cam;
  int W = cam->width, H = cam->height;
  int64_t oidx = (int64_t)(H - py - 1) * W + px;  // kernels/raytracer.cu:106
  uint32_t rng = wang(seed_index(px, py, W) ^ c.o->seed);
  double xx, yy;
  pixel_xy(cam, px, py, &xx, &yy);
  V3 color = v3(0, 0, 0);
  int N = c.o->spp;
  for (int it = 0; it < N; ++it) {
    V3 o, d;
    lens_ray(cam, c.o->flags, xx, yy, rng, &o, &d);
    ++cnt[0];
    const int64_t slot = ((int64_t)(py * W + px) * N + it) * (c.o->depth + 1);
    int32_t *r = rec ? rec + 2 * slot : nullptr;
    int32_t *br = branches ? branches + slot : nullptr;
    for (int k = 0; k <= c.o->depth; ++k) {
      if (r) r[2 * k] = r[2 * k + 1] = -1;
      if (br) br[k] = -1;
    }
    color = add(color, trace_path(c, o, d, rng, cnt, r, br));
  }
  color = v3(color.x / double(N), color.y / double(N), color.z / double(N));
  if (c.o->flags & ORC_CLAMP) {  // kernels/scene.cu:137, inc/vec.h:123-134
    double *ch[3] = {&color.x, &color.y, &color.z};
    for (double *v : ch) {
      if (*v < 0) *v = 0.0;
      if (*v > 1) *v = 1.0;
    }
  }
  if (rgb) {
    rgb[3 * oidx + 0] = (float)color.x;
    rgb[3 * oidx + 1] = (float)color.y;
    rgb[3 * oidx + 2] = (float)color.z;
  }
  if (rgb64) {
    rgb64[3 * oidx + 0] = color.x;
    rgb64[3 * oidx + 1] = color.y;
    rgb64[3 * oidx + 2] = color.z;
  }
}

static void set_threads(int32_t nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

template <typename K>
static void stable_sort_keys(int32_t n, K *codes, uint32_t *ids) {
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return codes[a] < codes[b]; });
  std::vector<K> tmp(codes, codes + n);
  for (int i = 0; i < n; ++i) {
    codes[i] = tmp[perm[i]];
    ids[i] = perm[i];
  }
}

}  // namespace

extern "C" {

void orc_tri_aabb(const orc_mesh *m, float *lo, float *hi) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m->n; ++i) {
    const int32_t *ix = m->idx + 3 * i;
    V3 a = vertex(m, ix[0]), b = vertex(m, ix[1]), c = vertex(m, ix[2]);
    // kernels/scene.cu:110-120: max(a,b) then max(c, .); min likewise.
    V3 mx = v3(fmax_f(a.x, b.x), fmax_f(a.y, b.y), fmax_f(a.z, b.z));
    V3 mn = v3(fmin_f(a.x, b.x), fmin_f(a.y, b.y), fmin_f(a.z, b.z));
    mx = v3(fmax_f(c.x, mx.x), fmax_f(c.y, mx.y), fmax_f(c.z, mx.z));
    mn = v3(fmin_f(c.x, mn.x), fmin_f(c.y, mn.y), fmin_f(c.z, mn.z));
    lo[3 * i + 0] = (float)mn.x;
    lo[3 * i + 1] = (float)mn.y;
    lo[3 * i + 2] = (float)mn.z;
    hi[3 * i + 0] = (float)mx.x;
    hi[3 * i + 1] = (float)mx.y;
    hi[3 * i + 2] = (float)mx.z;
  }
}

void orc_scene_bounds(int32_t n, const float *lo, const float *hi, double *mn, double *mx) {
  double a[3] = {1e9, 1e9, 1e9}, b[3] = {-1e9, -1e9, -1e9};
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      a[k] = fmin_f(a[k], lo[3 * i + k]);
      b[k] = fmax_f(b[k], hi[3 * i + k]);
    }
  for (int k = 0; k < 3; ++k) {
    mn[k] = a[k];
    mx[k] = b[k];
  }
}

uint32_t orc_expand_bits10(uint32_t v) {  // kernels/bvh.cu:13-21 restated as a bit loop
  uint32_t r = 0;
  for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
  return r;
}

static inline uint64_t expand21(uint64_t v) {
  uint64_t r = 0;
  for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1ull) << (3 * b);
  return r;
}

// kernels/bvh.cu:22-32 and :52-58: centroid normalised in FP64, scaled,
// clamped with fmax/fmin (NaN -> 0), rounded to float, truncated.
static inline void unit_centroid(const float *lo, const float *hi, const double *mn, const double *mx,
                                 double *u) {
  for (int k = 0; k < 3; ++k) {
    double c = 0.5 * ((double)lo[k] + (double)hi[k]);
    u[k] = (c - mn[k]) / (mx[k] - mn[k]);
  }
}

void orc_morton30(int32_t n, const float *lo, const float *hi, const double *mn, const double *mx,
                  uint32_t *codes) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double u[3];
    unit_centroid(lo + 3 * i, hi + 3 * i, mn, mx, u);
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
      float f = (float)fmin(fmax(u[k] * 1024.0, 0.0), 1023.0);
      q[k] = (uint32_t)f;
    }
    codes[i] = orc_expand_bits10(q[0]) * 4 + orc_expand_bits10(q[1]) * 2 + orc_expand_bits10(q[2]);
  }
}

void orc_morton63(int32_t n, const float *lo, const float *hi, const double *mn, const double *mx,
                  uint64_t *codes) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double u[3];
    unit_centroid(lo + 3 * i, hi + 3 * i, mn, mx, u);
    uint64_t q[3];
    for (int k = 0; k < 3; ++k) {
      float f = (float)fmin(fmax(u[k] * 2097152.0, 0.0), 2097151.0);
      q[k] = (uint64_t)f;
    }
    codes[i] = (expand21(q[0]) << 2) | (expand21(q[1]) << 1) | expand21(q[2]);
  }
}

void orc_sort30(int32_t n, uint32_t *codes, uint32_t *ids) { stable_sort_keys<uint32_t>(n, codes, ids); }
void orc_sort63(int32_t n, uint64_t *codes, uint32_t *ids) { stable_sort_keys<uint64_t>(n, codes, ids); }

void orc_hierarchy30(int32_t n, const uint32_t *codes, const uint32_t *ids, int32_t *parent, int32_t *left,
                     int32_t *right) {
  hierarchy<uint32_t>(n, codes, ids, parent, left, right);
}
void orc_hierarchy63(int32_t n, const uint64_t *codes, const uint32_t *ids, int32_t *parent, int32_t *left,
                     int32_t *right) {
  hierarchy<uint64_t>(n, codes, ids, parent, left, right);
}
void orc_topdown30(int32_t n, const uint32_t *codes, const uint32_t *ids, int32_t *parent, int32_t *left,
                   int32_t *right) {
  topdown<uint32_t>(n, codes, ids, parent, left, right);
}
void orc_topdown63(int32_t n, const uint64_t *codes, const uint32_t *ids, int32_t *parent, int32_t *left,
                   int32_t *right) {
  topdown<uint64_t>(n, codes, ids, parent, left, right);
}

void orc_refit(int32_t n, const int32_t *parent, const int32_t *left, const int32_t *right, const float *tri_lo,
               const float *tri_hi, const int32_t *perm, float *lo, float *hi) {
  if (n <= 0) return;
  // Leaves: sortedBBoxs[LEAFIDX(i)] = BBoxs[ids[i]] (kernels/bvh.cu:169-173).
  for (int i = 0; i < n; ++i) {
    int leaf = n - 1 + i, tri = left[leaf];
    for (int k = 0; k < 3; ++k) {
      lo[3 * leaf + k] = tri_lo[3 * tri + k];
      hi[3 * leaf + k] = tri_hi[3 * tri + k];
    }
  }
  if (n == 1) return;
  std::vector<int> flags(n - 1, 0);
  for (int j = 0; j < n; ++j) {
    int i = perm ? perm[j] : j;
    int idx = n - 1 + i;
    do {  // kernels/bvh.cu:220-255
      idx = parent[idx];
      if (flags[idx]++ == 0) break;  // first arrival leaves the node to its sibling
      int l = left[idx], r = right[idx];
      for (int k = 0; k < 3; ++k) {
        lo[3 * idx + k] = (float)fmin_f(lo[3 * l + k], lo[3 * r + k]);
        hi[3 * idx + k] = (float)fmax_f(hi[3 * l + k], hi[3 * r + k]);
      }
    } while (idx != 0);
  }
}

void orc_refit_brute(int32_t n, const int32_t *left, const int32_t *right, const float *tri_lo,
                     const float *tri_hi, float *lo, float *hi) {
  if (n <= 0) return;
  // Post-order by explicit recursion over the topology: each node's box is
  // the min/max over the triangles of its subtree, folded in leaf order.
  std::vector<int> order;
  std::vector<int> st{0};
  if (n == 1) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = tri_lo[3 * left[0] + k];
      hi[k] = tri_hi[3 * left[0] + k];
    }
    return;
  }
  std::vector<std::vector<int>> tris(2 * n - 1);
  // collect leaves of each subtree
  std::vector<std::pair<int, int>> frames{{0, 0}};
  std::vector<int> post;
  while (!frames.empty()) {
    auto &f = frames.back();
    int node = f.first;
    if (left[node] == right[node]) {
      post.push_back(node);
      frames.pop_back();
      continue;
    }
    if (f.second == 0) {
      f.second = 1;
      frames.push_back({left[node], 0});
    } else if (f.second == 1) {
      f.second = 2;
      frames.push_back({right[node], 0});
    } else {
      post.push_back(node);
      frames.pop_back();
    }
  }
  for (int node : post) {
    if (left[node] == right[node]) {
      tris[node].push_back(left[node]);
    } else {
      tris[node] = tris[left[node]];
      tris[node].insert(tris[node].end(), tris[right[node]].begin(), tris[right[node]].end());
      tris[left[node]].clear();
      tris[left[node]].shrink_to_fit();
    }
    float a[3] = {INFINITY, INFINITY, INFINITY}, b[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int t : tris[node])
      for (int k = 0; k < 3; ++k) {
        a[k] = fminf(a[k], tri_lo[3 * t + k]);
        b[k] = fmaxf(b[k], tri_hi[3 * t + k]);
      }
    for (int k = 0; k < 3; ++k) {
      lo[3 * node + k] = a[k];
      hi[3 * node + k] = b[k];
    }
    if (left[node] == right[node]) continue;
    tris[right[node]].clear();
    tris[right[node]].shrink_to_fit();
  }
}

void orc_build(const orc_mesh *m, int32_t bits, uint64_t *codes, uint32_t *ids, int32_t *parent, int32_t *left,
               int32_t *right, float *lo, float *hi, double *mn, double *mx, int32_t nthreads) {
  set_threads(nthreads);
  int32_t n = m->n;
  if (n <= 0) return;
  std::vector<float> tlo(3 * (size_t)n), thi(3 * (size_t)n);
  orc_tri_aabb(m, tlo.data(), thi.data());
  orc_scene_bounds(n, tlo.data(), thi.data(), mn, mx);
  if (bits == 30) {
    std::vector<uint32_t> c32(n);
    orc_morton30(n, tlo.data(), thi.data(), mn, mx, c32.data());
    orc_sort30(n, c32.data(), ids);
    orc_hierarchy30(n, c32.data(), ids, parent, left, right);
    for (int i = 0; i < n; ++i) codes[i] = c32[i];
  } else {
    orc_morton63(n, tlo.data(), thi.data(), mn, mx, codes);
    orc_sort63(n, codes, ids);
    orc_hierarchy63(n, codes, ids, parent, left, right);
  }
  orc_refit(n, parent, left, right, tlo.data(), thi.data(), nullptr, lo, hi);
}

void orc_trace(const orc_mesh *m, const orc_bvh *b, int64_t nrays, const double *orig, const double *dir,
               int32_t any_hit, int32_t *tri, double *t, double *alpha, double *beta, double *normal,
               int64_t *visits, int32_t nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t r = 0; r < nrays; ++r) {
    Hit h;
    int64_t vc[2] = {0, 0};
    V3 p = v3(orig[3 * r], orig[3 * r + 1], orig[3 * r + 2]);
    V3 d = v3(dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]);
    bool ok = traverse(m, b, p, d, any_hit != 0, 0.0, &h, vc);
    tri[r] = ok ? h.tri : -1;
    if (t) t[r] = ok ? h.t : 0.0;
    if (alpha) alpha[r] = ok ? h.alpha : 0.0;
    if (beta) beta[r] = ok ? h.beta : 0.0;
    if (normal) {
      normal[3 * r] = ok ? h.N.x : 0.0;
      normal[3 * r + 1] = ok ? h.N.y : 0.0;
      normal[3 * r + 2] = ok ? h.N.z : 0.0;
    }
    if (visits) {
      visits[2 * r] = vc[0];
      visits[2 * r + 1] = vc[1];
    }
  }
}

void orc_trace_brute(const orc_mesh *m, const int32_t *leafpos, int64_t nrays, const double *orig,
                     const double *dir, int32_t *tri, double *t) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t r = 0; r < nrays; ++r) {
    V3 p = v3(orig[3 * r], orig[3 * r + 1], orig[3 * r + 2]);
    V3 d = v3(dir[3 * r], dir[3 * r + 1], dir[3 * r + 2]);
    int32_t best = -1;
    double bt = 0;
    for (int32_t k = 0; k < m->n; ++k) {
      Hit h;
      if (!tri_hit(m, k, p, d, &h)) continue;
      if (best < 0 || h.t < bt || (h.t == bt && leafpos[k] < leafpos[best])) {
        best = k;
        bt = h.t;
      }
    }
    tri[r] = best;
    t[r] = best >= 0 ? bt : 0.0;
  }
}

int32_t orc_intersect_triangle(const orc_mesh *m, int32_t tri, const double *orig, const double *dir, double *t,
                               double *alpha, double *beta, double *normal) {
  Hit h;
  bool ok = tri_hit(m, tri, arr3(orig), arr3(dir), &h);
  if (!ok) return 0;
  *t = h.t;
  *alpha = h.alpha;
  *beta = h.beta;
  normal[0] = h.N.x;
  normal[1] = h.N.y;
  normal[2] = h.N.z;
  return 1;
}

int32_t orc_box_hit(const float *lo, const float *hi, const double *orig, const double *dir) {
  double a[3] = {1.0 / dir[0], 1.0 / dir[1], 1.0 / dir[2]};
  return box_hit(lo, hi, orig, a) ? 1 : 0;
}

uint32_t orc_wang_hash(uint32_t s) { return wang(s); }
uint32_t orc_xorshift(uint32_t *state) { return xorshift(*state); }
double orc_rand_double(uint32_t *state) { return rand_double(*state); }
uint32_t orc_seed_index(int32_t px, int32_t py, int32_t width) { return seed_index(px, py, width); }
void orc_sincos(double x, double *s, double *c) { orc_math::sincos(x, s, c); }

void orc_primary_ray(const orc_camera *cam, const orc_render_opts *opts, int32_t px, int32_t py, int32_t sample,
                     double *orig, double *dir) {
  // Only valid for samples whose predecessors missed (a hit draws 3 more numbers).
  uint32_t rng = wang(seed_index(px, py, cam->width) ^ opts->seed);
  double xx, yy;
  pixel_xy(cam, px, py, &xx, &yy);
  V3 o, d;
  for (int k = 0; k <= sample; ++k) lens_ray(cam, opts->flags, xx, yy, rng, &o, &d);
  orig[0] = o.x; orig[1] = o.y; orig[2] = o.z;
  dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
}

void orc_cosine_dir(const double *n, double u1, double u2, double *dir) {
  int axis;
  V3 d = cosine_dir(v3(n[0], n[1], n[2]), u1, u2, &axis);
  dir[0] = d.x;
  dir[1] = d.y;
  dir[2] = d.z;
}

void orc_render(const orc_mesh *m, const orc_bvh *b, const orc_material *mats, int32_t nmat, const orc_light *light,
                const orc_camera *cam, const orc_render_opts *opts, float *rgb, const int64_t *pixels, int64_t npix,
                int64_t *counts, int32_t nthreads) {
  orc_render64(m, b, mats, nmat, light, cam, opts, rgb, nullptr, nullptr, nullptr, pixels, npix, counts, nthreads);
}

void orc_render64(const orc_mesh *m, const orc_bvh *b, const orc_material *mats, int32_t nmat,
                  const orc_light *light, const orc_camera *cam, const orc_render_opts *opts, float *rgb,
                  double *rgb64, int32_t *records, int32_t *branches, const int64_t *pixels, int64_t npix,
                  int64_t *counts, int32_t nthreads) {
  set_threads(nthreads);
  RenderCtx c{m, b, mats, nmat, light, cam, opts};
  int W = cam->width, H = cam->height;
  int64_t total = pixels ? npix : (int64_t)W * H;
  int64_t c0 = 0, c1 = 0, c2 = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : c0, c1, c2)
  for (int64_t k = 0; k < total; ++k) {
    int64_t pix = pixels ? pixels[k] : k;
    int px = (int)(pix % W), py = (int)(pix / W);
    int64_t cnt[3] = {0, 0, 0};
    render_pixel(c, px, py, rgb, rgb64, records, branches, cnt);
    c0 += cnt[0];
    c1 += cnt[1];
    c2 += cnt[2];
  }
  (void)H;
  if (counts) {
    counts[0] += c0;
    counts[1] += c1;
    counts[2] += c2;
  }
}

}  // extern "C"

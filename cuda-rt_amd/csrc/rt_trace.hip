// BVH traversal with ray-triangle intersection and the thin-lens render
// megakernel for gfx950.
//
// Reference path: inc/bbox.h:198-218 (slab test), kernels/bvh.cu:297-449
// (Moller-Trumbore, closest-hit and any-hit DFS), kernels/raytracer.cu:94-249
// (thin-lens ray generation, traceRay), kernels/material.cu:6-40 and
// kernels/Light.cu (shading, shadow ray), kernels/mrand.cu (RNG).
//
// Box and triangle tests are evaluated in FP64 exactly as the reference
// writes them (same operations, same order, no contraction), so the set of
// boxes a ray enters and the triangles it hits are the reference's.
// Closest hit is the minimum (t, sorted leaf position) over that set: the
// reference's strict "<" in left-first DFS order picks the same hit, since
// left-first DFS reaches leaves in sorted order. That makes the order free:
// the kernel descends nearer children first and skips boxes whose entry
// distance is beyond the best hit by more than a rounding margin. Any hit
// must return the first hit in DFS order, so it walks left child first.
#include <hip/hip_runtime.h>

#include "rt_internal.h"
#include "rt_math.h"

namespace {

using rt::d3;
using rt::kRayEps;
using rt::mk;

constexpr int kStack = 64;

struct RayD {
  d3 o, d;
  double ax, ay, az;
  bool sx, sy, sz;
};

__device__ __forceinline__ RayD make_ray(d3 o, d3 d) {
  RayD r;
  r.o = o;
  r.d = d;
  r.ax = 1.0 / d.x;
  r.ay = 1.0 / d.y;
  r.az = 1.0 / d.z;
  r.sx = r.ax >= 0;
  r.sy = r.ay >= 0;
  r.sz = r.az >= 0;
  return r;
}

// inc/bbox.h:198-218 on fp32 bounds widened exactly to double.
__device__ __forceinline__ bool slab(const RayD& r, float lx, float ly, float lz, float hx, float hy, float hz,
                                     double& t0, double& t1) {
  double nx = r.sx ? (double)lx : (double)hx, fx = r.sx ? (double)hx : (double)lx;
  double ny = r.sy ? (double)ly : (double)hy, fy = r.sy ? (double)hy : (double)ly;
  double nz = r.sz ? (double)lz : (double)hz, fz = r.sz ? (double)hz : (double)lz;
  double tminx = (nx - r.o.x) * r.ax, tmaxx = (fx - r.o.x) * r.ax;
  double tminy = (ny - r.o.y) * r.ay, tmaxy = (fy - r.o.y) * r.ay;
  double tminz = (nz - r.o.z) * r.az, tmaxz = (fz - r.o.z) * r.az;
  t0 = fmax(fmax(tminx, tminy), tminz);
  t1 = fmin(fmin(tmaxx, tmaxy), tmaxz);
  return t0 <= t1;
}

// kernels/bvh.cu:297-347 on the leaf record's vertices.
__device__ __forceinline__ bool tri_test(const RayD& r, const float4* __restrict__ L, double& t, double& al,
                                         double& be) {
  float4 q0 = L[0], q1 = L[1], q2 = L[2];
  d3 a = mk(q0.x, q0.y, q0.z), b = mk(q0.w, q1.x, q1.y), c = mk(q1.z, q1.w, q2.x);
  d3 AB = b - a, AC = c - a;
  d3 P = rt::cross(r.d, AC);
  double det = rt::dot(AB, P);
  if ((double)fabsf((float)det) < kRayEps) return false;
  double inv = 1 / det;
  d3 T = r.o - a;
  al = rt::dot(T, P) * inv;
  if (al < 0 || al > 1) return false;
  d3 Q = rt::cross(T, AB);
  be = rt::dot(r.d, Q) * inv;
  if (be < 0 || al + be > 1) return false;
  t = rt::dot(AC, Q) * inv;
  if ((double)fabsf((float)t) < kRayEps) return false;
  return true;
}

struct HitRec {
  double t, al, be;
  int32_t pos;  // sorted leaf position, -1 = none
};

__device__ __forceinline__ double prune_margin(double t) { return 1e-4 * fabs(t) + 1e-7; }

__device__ __forceinline__ void load_node(const float4* __restrict__ tn, int node, float4& a, float4& b, float4& c,
                                          float4& d) {
  const float4* p = tn + 4 * (size_t)node;
  a = p[0];
  b = p[1];
  c = p[2];
  d = p[3];
}

__device__ __forceinline__ void consider_leaf(const BvhView& B, const RayD& r, int pos, double tmin_phys,
                                              HitRec& best) {
  double t, al, be;
  if (!tri_test(r, B.ltri + 3 * (size_t)pos, t, al, be)) return;
  if (tmin_phys > 0 && !(t >= tmin_phys)) return;
  if (best.pos < 0 || t < best.t || (t == best.t && pos < best.pos)) {
    best.t = t;
    best.al = al;
    best.be = be;
    best.pos = pos;
  }
}

// Closest hit (kernels/bvh.cu:382-414): minimum (t, leaf position) over all
// triangles whose ancestors' boxes the line enters. tmin_phys > 0 restricts
// hits to t >= tmin_phys (bounce rays). Returns false on stack overflow
// (counted), with best.pos = -1.
__device__ bool closest_hit(const BvhView& B, const RayD& r, double tmin_phys, HitRec& best, uint32_t* overflow) {
  best.pos = -1;
  best.t = 0;
  const int n = B.n;
  if (n <= 0) return true;
  double t0, t1;
  const float* rb = B.node_box;
  if (!slab(r, rb[0], rb[1], rb[2], rb[3], rb[4], rb[5], t0, t1)) return true;
  if (n == 1) {
    consider_leaf(B, r, 0, tmin_phys, best);
    return true;
  }
  int stack[kStack];
  int sp = 0;
  int node = 0;
  // Each internal node is entered at most once per ray; the bound only
  // guards against a malformed tree.
  for (int guard = 0; guard < n; ++guard) {
    float4 q0, q1, q2, q3;
    load_node(B.tnode, node, q0, q1, q2, q3);
    double l0, l1, r0, r1;
    bool hl = slab(r, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, l0, l1);
    bool hr = slab(r, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, r0, r1);
    int cl = __float_as_int(q3.x), cr = __float_as_int(q3.y);
    if (tmin_phys > 0) {
      hl = hl && !(l1 < tmin_phys - prune_margin(tmin_phys));
      hr = hr && !(r1 < tmin_phys - prune_margin(tmin_phys));
    }
    if (best.pos >= 0) {
      double lim = best.t + prune_margin(best.t);
      hl = hl && !(l0 > lim);
      hr = hr && !(r0 > lim);
    }
    if (hl && cl < 0) {
      consider_leaf(B, r, ~cl, tmin_phys, best);
      hl = false;
    }
    if (hr && cr < 0) {
      consider_leaf(B, r, ~cr, tmin_phys, best);
      hr = false;
    }
    if (hl && hr) {
      int nearc = (l0 <= r0) ? cl : cr, farc = (l0 <= r0) ? cr : cl;
      if (sp == kStack) {
        atomicAdd(overflow, 1u);
        best.pos = -1;
        return false;
      }
      stack[sp++] = farc;
      node = nearc;
    } else if (hl) {
      node = cl;
    } else if (hr) {
      node = cr;
    } else {
      if (sp == 0) return true;
      node = stack[--sp];
    }
  }
  atomicAdd(overflow, 1u);
  best.pos = -1;
  return false;
}

// Any hit (kernels/bvh.cu:416-449): the first triangle in left-first DFS
// order whose test passes, whatever its t.
__device__ bool any_hit(const BvhView& B, const RayD& r, HitRec& hit, uint32_t* overflow) {
  hit.pos = -1;
  const int n = B.n;
  if (n <= 0) return false;
  double t0, t1, t, al, be;
  const float* rb = B.node_box;
  if (!slab(r, rb[0], rb[1], rb[2], rb[3], rb[4], rb[5], t0, t1)) return false;
  if (n == 1) {
    if (tri_test(r, B.ltri, t, al, be)) {
      hit.t = t; hit.al = al; hit.be = be; hit.pos = 0;
      return true;
    }
    return false;
  }
  int stack[kStack];
  int sp = 0;
  int node = 0;
  for (int guard = 0; guard < n; ++guard) {
    float4 q0, q1, q2, q3;
    load_node(B.tnode, node, q0, q1, q2, q3);
    double l0, l1, r0, r1;
    bool hl = slab(r, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, l0, l1);
    bool hr = slab(r, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, r0, r1);
    int cl = __float_as_int(q3.x), cr = __float_as_int(q3.y);
    int next = -1;
    bool have_next = false;
    if (hl) {
      if (cl < 0) {
        if (tri_test(r, B.ltri + 3 * (size_t)(~cl), t, al, be)) {
          hit.t = t; hit.al = al; hit.be = be; hit.pos = ~cl;
          return true;
        }
      } else {
        if (hr) {
          if (sp == kStack) {
            atomicAdd(overflow, 1u);
            return false;
          }
          stack[sp++] = cr;
        }
        next = cl;
        have_next = true;
      }
    }
    if (!have_next && hr) {
      if (cr < 0) {
        if (tri_test(r, B.ltri + 3 * (size_t)(~cr), t, al, be)) {
          hit.t = t; hit.al = al; hit.be = be; hit.pos = ~cr;
          return true;
        }
      } else {
        next = cr;
        have_next = true;
      }
    }
    while (!have_next) {
      if (sp == 0) return false;
      int ref = stack[--sp];
      if (ref < 0) {
        if (tri_test(r, B.ltri + 3 * (size_t)(~ref), t, al, be)) {
          hit.t = t; hit.al = al; hit.be = be; hit.pos = ~ref;
          return true;
        }
      } else {
        next = ref;
        have_next = true;
      }
    }
    node = next;
  }
  atomicAdd(overflow, 1u);
  return false;
}

__device__ __forceinline__ int32_t leaf_tri(const BvhView& B, int pos) {
  return __float_as_int(B.ltri[3 * (size_t)pos + 2].y);
}

// ------------------------------------------------------------ trace kernel
__global__ __launch_bounds__(256) void k_trace(BvhView B, const double* __restrict__ orig,
                                               const double* __restrict__ dir, long nrays, int mode, double tmin,
                                               int32_t* __restrict__ tri, double* __restrict__ tout,
                                               double* __restrict__ aout, double* __restrict__ bout,
                                               uint32_t* overflow) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nrays) return;
  RayD r = make_ray(mk(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]), mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
  HitRec h;
  if (mode == RT_TRACE_ANY)
    any_hit(B, r, h, overflow);
  else
    closest_hit(B, r, tmin, h, overflow);
  bool ok = h.pos >= 0;
  tri[i] = ok ? leaf_tri(B, h.pos) : -1;
  if (tout) tout[i] = ok ? h.t : 0.0;
  if (aout) aout[i] = ok ? h.al : 0.0;
  if (bout) bout[i] = ok ? h.be : 0.0;
}

// ------------------------------------------------------------ render
struct RenderParams {
  int W, H, spp, depth;
  uint32_t seed;
  int flags, nmat;
  double invW, invH, aspect, angle, aperture, focal;
  double eye[3];
  double lpos[3], lcol[3], lrad, c0, c1, c2;
};

struct MeshShade {
  const float *nx, *ny, *nz;
  const int32_t* idx;
  const int32_t* mat_id;
};

__device__ __forceinline__ d3 arr3(const double* p) { return mk(p[0], p[1], p[2]); }

__device__ __forceinline__ const rt_material& mat_of(const rt_material* mats, int nmat, const MeshShade& M, int tri) {
  int id = M.mat_id ? M.mat_id[tri] : 0;
  if (id < 0 || id >= nmat) id = 0;
  return mats[id];
}

// Smooth normal (kernels/bvh.cu:357-364).
__device__ __forceinline__ d3 shading_normal(const MeshShade& M, int tri, double al, double be) {
  int a = M.idx[3 * tri], b = M.idx[3 * tri + 1], c = M.idx[3 * tri + 2];
  d3 na = mk(M.nx[a], M.ny[a], M.nz[a]), nb = mk(M.nx[b], M.ny[b], M.nz[b]), nc = mk(M.nx[c], M.ny[c], M.nz[c]);
  d3 N = (na * (1.0 - (al + be)) + nb * al) + nc * be;
  return rt::normalize(N);
}

// Cosine-weighted direction about n: the frame of src/ray.cpp:17-30 with
// sin(theta) = sqrt(u1), cos(theta) = sqrt(1-u1), phi = 2 pi u2.
__device__ __forceinline__ d3 cosine_dir(d3 n, double u1, double u2) {
  double sphi, cphi;
  rt::sincos_cw(2.0 * 3.141592653589793 * u2, &sphi, &cphi);
  double st = sqrt(u1), ct = sqrt(1.0 - u1);
  double xs = st * cphi, ys = ct, zs = st * sphi;
  d3 h = n;
  float ax = fabsf((float)h.x), ay = fabsf((float)h.y), az = fabsf((float)h.z);
  if (ax <= ay && ax <= az)
    h.x = 1.0;
  else if (ay <= ax && ay <= az)
    h.y = 1.0;
  else
    h.z = 1.0;
  d3 x = rt::normalize(rt::cross(h, n));
  d3 z = rt::normalize(rt::cross(x, n));
  return rt::normalize((x * xs + n * ys) + z * zs);
}

constexpr double kBounceTmin = 1e-6;

__global__ __launch_bounds__(256) void k_render(BvhView B, MeshShade M, const rt_material* __restrict__ mats,
                                                RenderParams P, float* __restrict__ rgb,
                                                unsigned long long* counters, uint32_t* overflow) {
  const int tiles_x = (P.W + 15) / 16;
  const int tile = blockIdx.x;
  const int px = (tile % tiles_x) * 16 + (threadIdx.x & 15);
  const int py = (tile / tiles_x) * 16 + (threadIdx.x >> 4);
  const bool active = px < P.W && py < P.H;
  unsigned long long n_shadow = 0, n_bounce = 0;
  if (active) {
    uint32_t rng = rt::wang_hash(rt::seed_index(px, py, P.W) ^ P.seed);
    double xx = (2 * ((px + 0.5) * P.invW) - 1) * P.angle * P.aspect;
    double yy = (1 - 2 * ((py + 0.5) * P.invH)) * P.angle;
    const bool rtl = (P.flags & RT_RENDER_RANDVEC_RTL) != 0;
    const d3 lpos = arr3(P.lpos), lcol = arr3(P.lcol), eye = arr3(P.eye);
    d3 color = mk(0, 0, 0);
    for (int it = 0; it < P.spp; ++it) {
      double ang = 2 * M_PI * rt::rand_unit(rng);
      double rad = P.aperture * sqrt(rt::rand_unit(rng)) / 2.0;
      double sa, ca;
      rt::sincos_cw(ang, &sa, &ca);
      d3 o = mk(rad * ca, rad * sa, 0);
      d3 d = rt::normalize(mk(xx, yy, -1.0));
      d = d * P.focal - o;
      d = rt::normalize(d);
      d3 acc = mk(0, 0, 0), thr = mk(1, 1, 1);
      for (int bounce = 0;; ++bounce) {
        RayD ray = make_ray(o, d);
        HitRec h;
        closest_hit(B, ray, bounce == 0 ? 0.0 : kBounceTmin, h, overflow);
        if (h.pos < 0) {
          acc = acc + rt::hadamard(thr, mk(0.9, 0.9, 0.9));
          break;
        }
        const int tri = leaf_tri(B, h.pos);
        const rt_material& m = mat_of(mats, P.nmat, M, tri);
        d3 N = shading_normal(M, tri, h.al, h.be);
        // Material::shade (kernels/material.cu:6-40)
        d3 q = o + d * h.t;
        d3 I = arr3(m.ke) + rt::hadamard(arr3(m.ka), mk(0, 0, 0));
        d3 V = eye - q;
        V = mk(-V.x, -V.y, -V.z);
        V = rt::normalize(V);
        // Light::getDirection (kernels/Light.cu:71-81)
        double r1, r2, r3;
        if (!rtl) {
          r1 = rt::rand_unit(rng); r2 = rt::rand_unit(rng); r3 = rt::rand_unit(rng);
        } else {
          r3 = rt::rand_unit(rng); r2 = rt::rand_unit(rng); r1 = rt::rand_unit(rng);
        }
        d3 off = mk(r1, r2, r3) * P.lrad;
        off = mk(off.x - 0.5, off.y - 0.5, off.z - 0.5);
        d3 L = rt::normalize((lpos + off) - q);
        // Light::shadowAttenuation (kernels/Light.cu:5-53)
        d3 sh = mk(1, 1, 1);
        {
          RayD sray = make_ray(q, L);
          HitRec s;
          ++n_shadow;
          if (any_hit(B, sray, s, overflow)) {
            if (s.t < rt::length_f(lpos - q)) {
              if (!(s.t < kRayEps)) {
                const rt_material& sm = mat_of(mats, P.nmat, M, leaf_tri(B, s.pos));
                bool trans = (sm.kt[0] == 0) && (sm.kt[1] == 0) && (sm.kt[2] == 0);
                sh = trans ? rt::hadamard(arr3(sm.kd), mk(1, 1, 1) - arr3(sm.kt)) : mk(0, 0, 0);
              }
            }
          }
        }
        // Light::distanceAttenuation (kernels/Light.cu:56-68)
        double dl = rt::length_f(lpos - q);
        double fo = P.c0 + (P.c1 + P.c2 * dl) * dl;
        fo = (fabs(fo) < kRayEps) ? 1.0 : 1.0 / fo;
        fo = fmin(1.0, fo);
        d3 atten = sh * fo;
        double nl = rt::dot(N, L);
        if (nl < 0) nl = 0;
        d3 diffuse = arr3(m.kd) * nl;
        d3 spec = mk(0, 0, 0);
        if (!(m.ks[0] == 0 && m.ks[1] == 0 && m.ks[2] == 0)) {
          d3 R = N * (-2.0 * rt::dot(L, N)) + L;
          R = rt::normalize(R);
          double rv = rt::dot(R, V);
          double sf = (double)powf((float)fmax(0.0, rv), (float)m.shininess);
          spec = arr3(m.ks) * sf;
        }
        I = I + rt::hadamard(rt::hadamard(atten, diffuse + spec), lcol);
        acc = acc + rt::hadamard(thr, I);
        if (bounce >= P.depth) break;
        double u1 = rt::rand_unit(rng);
        double u2 = rt::rand_unit(rng);
        if (rt::dot(N, d) > 0) N = mk(-N.x, -N.y, -N.z);
        thr = rt::hadamard(thr, arr3(m.kd));
        o = q;
        d = cosine_dir(N, u1, u2);
        ++n_bounce;
      }
      color = color + acc;
    }
    color = mk(color.x / double(P.spp), color.y / double(P.spp), color.z / double(P.spp));
    if (P.flags & RT_RENDER_CLAMP) {
      color.x = color.x < 0 ? 0.0 : (color.x > 1 ? 1.0 : color.x);
      color.y = color.y < 0 ? 0.0 : (color.y > 1 ? 1.0 : color.y);
      color.z = color.z < 0 ? 0.0 : (color.z > 1 ? 1.0 : color.z);
    }
    size_t oi = 3 * ((size_t)(P.H - py - 1) * P.W + px);
    rgb[oi + 0] = (float)color.x;
    rgb[oi + 1] = (float)color.y;
    rgb[oi + 2] = (float)color.z;
  }
  if (counters) {
    unsigned long long prim = active ? (unsigned long long)P.spp : 0ull;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      prim += __shfl_xor(prim, off);
      n_shadow += __shfl_xor(n_shadow, off);
      n_bounce += __shfl_xor(n_bounce, off);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&counters[0], prim);
      atomicAdd(&counters[1], n_shadow);
      atomicAdd(&counters[2], n_bounce);
    }
  }
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace

hipError_t rt_launch_trace(const rt_bvh* b, const rt_mesh_view* m, const rt_rays* rays, int32_t mode, double tmin,
                           const rt_hits* hits, hipStream_t s) {
  (void)m;
  if (rays->n <= 0) return hipSuccess;
  k_trace<<<cdiv(rays->n, 256), 256, 0, s>>>(b->view(), rays->orig_dev, rays->dir_dev, rays->n, mode, tmin,
                                            hits->tri_dev, hits->t_dev, hits->alpha_dev, hits->beta_dev,
                                            b->misc + kMiscOverflow);
  return hipGetLastError();
}

hipError_t rt_launch_render(const rt_bvh* b, const rt_mesh_view* m, const rt_material* mats, int32_t nmat,
                            const rt_light* light, const rt_camera* cam, const rt_render_opts* opts, float* rgb,
                            uint64_t* counters, hipStream_t s) {
  RenderParams P;
  P.W = cam->width;
  P.H = cam->height;
  P.spp = opts->spp;
  P.depth = opts->depth;
  P.seed = opts->seed;
  P.flags = opts->flags;
  P.nmat = nmat;
  // kernels/raytracer.cu:126-133, evaluated once on the host.
  P.invW = 1.0 / double(P.W);
  P.invH = 1.0 / double(P.H);
  P.aspect = double(P.W) / double(P.H);
  P.angle = tan(M_PI * 0.5 * cam->fov_deg / 180.0);
  P.aperture = 1 / cam->fstop;
  P.focal = cam->focal_point;
  for (int k = 0; k < 3; ++k) {
    P.eye[k] = cam->eye[k];
    P.lpos[k] = light->position[k];
    P.lcol[k] = light->color[k];
  }
  P.lrad = light->radius;
  P.c0 = light->c0;
  P.c1 = light->c1;
  P.c2 = light->c2;
  MeshShade M{m->nx_dev, m->ny_dev, m->nz_dev, m->idx_dev, m->mat_id_dev};
  int tiles = cdiv(P.W, 16) * cdiv(P.H, 16);
  k_render<<<tiles, 256, 0, s>>>(b->view(), M, mats, P, rgb, (unsigned long long*)counters,
                                 b->misc + kMiscOverflow);
  return hipGetLastError();
}

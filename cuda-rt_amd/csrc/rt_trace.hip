// BVH traversal with ray-triangle intersection and the thin-lens render
// megakernel for gfx950.
//
// Reference path: inc/bbox.h:82-102 (slab test), kernels/bvh.cu:297-449
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

#include "../../include/rt_debug.h"
#include "rt_adjoint.h"
#include "rt_internal.h"
#include "rt_math.h"

namespace {

using rt::d3;
using rt::kRayEps;
using rt::mk;


// Loads through global-address-space pointers: the BVH arrays are reached
// through pointers stored in memory, which the compiler would otherwise
// load with flat instructions (whose completion is counted together with
// the LDS traversal stack's accesses).
typedef float f4v_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ldg4(const float4* p) {
  const f4v_t v = *(const __attribute__((address_space(1))) f4v_t*)(p);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float ldgf(const float* p) {
  return *(const __attribute__((address_space(1))) float*)(p);
}
__device__ __forceinline__ double ldgd(const double* p) {
  return *(const __attribute__((address_space(1))) double*)(p);
}

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

// inc/bbox.h:82-102 on fp32 bounds widened exactly to double.
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
  float4 q0 = ldg4(L), q1 = ldg4(L + 1), q2 = ldg4(L + 2);
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
  // only a non-finite vertex or ray coordinate gets here with a NaN or
  // infinite t, alpha or beta; such a triangle is not hit
  return fabs(t) < INFINITY && al == al && be == be;
}

struct HitRec {
  double t, al, be;
  int32_t pos;  // sorted leaf position, -1 = none
};

// Work counters: child-box slab tests (two per binary node entered, four
// per 4-wide node) and triangle tests.
struct TravStats {
  uint32_t visits = 0, tests = 0;
};

__device__ __forceinline__ double prune_margin(double t) { return 1e-4 * fabs(t) + 1e-7; }

__device__ __forceinline__ void load_node(const float4* __restrict__ tn, int node, float4& a, float4& b, float4& c,
                                          float4& d) {
  const float4* p = tn + 4 * (size_t)node;
  a = ldg4(p);
  b = ldg4(p + 1);
  c = ldg4(p + 2);
  d = ldg4(p + 3);
}

__device__ __forceinline__ void consider_leaf(const BvhView& B, const RayD& r, int pos, double tmin_phys,
                                              HitRec& best, TravStats& st) {
  double t, al, be;
  ++st.tests;
  if (!tri_test(r, B.ltri + kLeafStride * (size_t)pos, t, al, be)) return;
  if (tmin_phys > 0 && !(t >= tmin_phys)) return;
  if (best.pos < 0 || t < best.t || (t == best.t && pos < best.pos)) {
    best.t = t;
    best.al = al;
    best.be = be;
    best.pos = pos;
  }
}

// Traversal stack: per lane, the top RING entries in LDS (one column per
// thread: entry k of a lane at col[k * STRIDE], conflict-free) over a
// private spill area; depth beyond RING + kSpill is reported as an overflow
// and the ray as a miss. Two ways to keep the LDS part:
// - LIN (linear): when the LDS part is full, its lower half moves to the
//   spill area, and it comes back when the LDS part runs empty, so a push or
//   pop addresses col[d * STRIDE] with no wrap-around;
// - ring: entry k at col[(k % RING) * STRIDE]; past RING entries, each push
//   moves the oldest LDS entry to the spill area and each pop below it
//   brings one back.
// The persistent closest-hit traversal (stacks mostly shallower than RING)
// gains from the cheaper addressing; the any-hit traversal, whose stacks
// along whole lines often pass RING entries, from the single-entry spills.
// The per-pixel megakernel (binary nodes) keeps the ring.
constexpr int kThreads = 256;
constexpr int kRing = 32;
constexpr int kSpill = 64;

struct LaneStack {
  int* col;  // this lane's column of the workgroup's LDS stack
  int d = 0;   // LIN: entries in LDS; ring: entries in all
  int lo = 0;  // entries in the spill area (LIN: a multiple of RING / 2)
};

template <bool LIN = true>
__device__ __forceinline__ bool stack_empty(const LaneStack& S) {
  return LIN ? (S.d | S.lo) == 0 : S.d == 0;
}

// STRIDE = threads per workgroup (one LDS column per thread).
template <int RING = kRing, int STRIDE = kThreads, bool LIN = true>
__device__ __forceinline__ bool stack_push(LaneStack& S, int* spill, int v) {
  static_assert(kSpill % (RING / 2) == 0, "the spill area holds whole halves");
  if (!LIN) {
    if (S.d - S.lo == RING) {
      if (S.lo == kSpill) return false;
      spill[S.lo] = S.col[(S.lo & (RING - 1)) * STRIDE];
      ++S.lo;
    }
    S.col[(S.d & (RING - 1)) * STRIDE] = v;
    ++S.d;
    return true;
  }
  if (S.d == RING) {
    if (S.lo == kSpill) return false;
#pragma unroll
    for (int k = 0; k < RING / 2; ++k) spill[S.lo + k] = S.col[k * STRIDE];
#pragma unroll
    for (int k = 0; k < RING / 2; ++k) S.col[k * STRIDE] = S.col[(k + RING / 2) * STRIDE];
    S.lo += RING / 2;
    S.d = RING / 2;
  }
  S.col[S.d * STRIDE] = v;
  ++S.d;
  return true;
}

// Requires !stack_empty(S).
template <int RING = kRing, int STRIDE = kThreads, bool LIN = true>
__device__ __forceinline__ int stack_pop(LaneStack& S, int* spill) {
  if (!LIN) {
    --S.d;
    const int v = S.col[(S.d & (RING - 1)) * STRIDE];
    if (S.d == S.lo && S.lo > 0) {
      --S.lo;
      S.col[(S.lo & (RING - 1)) * STRIDE] = spill[S.lo];
    }
    return v;
  }
  if (S.d == 0) {
    S.lo -= RING / 2;
#pragma unroll
    for (int k = 0; k < RING / 2; ++k) S.col[k * STRIDE] = spill[S.lo + k];
    S.d = RING / 2;
  }
  --S.d;
  return S.col[S.d * STRIDE];
}

// Closest hit (kernels/bvh.cu:382-414): minimum (t, leaf position) over all
// triangles whose ancestors' boxes the line enters. tmin_phys > 0 restricts
// hits to t >= tmin_phys (bounce rays). Returns false on stack overflow
// (counted), with best.pos = -1.
__device__ bool closest_hit(const BvhView& B, const RayD& r, double tmin_phys, HitRec& best, uint32_t* overflow,
                            TravStats& st, int* lds_col) {
  best.pos = -1;
  best.t = 0;
  const int n = B.n;
  if (n <= 0) return true;
  double t0, t1;
  const float* rb = B.node_box;
  if (!slab(r, rb[0], rb[1], rb[2], rb[3], rb[4], rb[5], t0, t1)) return true;
  if (n == 1) {
    consider_leaf(B, r, 0, tmin_phys, best, st);
    return true;
  }
  LaneStack S;
  S.col = lds_col;
  int spill[kSpill];
  int node = 0;
  // Each internal node is entered at most once per ray; the bound only
  // guards against a malformed tree.
  for (int guard = 0; guard < n; ++guard) {
    float4 q0, q1, q2, q3;
    load_node(B.tnode, node, q0, q1, q2, q3);
    st.visits += 2;
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
      consider_leaf(B, r, ~cl, tmin_phys, best, st);
      hl = false;
    }
    if (hr && cr < 0) {
      consider_leaf(B, r, ~cr, tmin_phys, best, st);
      hr = false;
    }
    if (hl && hr) {
      int nearc = (l0 <= r0) ? cl : cr, farc = (l0 <= r0) ? cr : cl;
      if (!stack_push<kRing, kThreads, false>(S, spill, farc)) {
        atomicAdd(overflow, 1u);
        best.pos = -1;
        return false;
      }
      node = nearc;
    } else if (hl) {
      node = cl;
    } else if (hr) {
      node = cr;
    } else {
      if (stack_empty<false>(S)) return true;
      node = stack_pop<kRing, kThreads, false>(S, spill);
    }
  }
  atomicAdd(overflow, 1u);
  best.pos = -1;
  return false;
}

// Any hit (kernels/bvh.cu:416-449): the first triangle in left-first DFS
// order whose test passes, whatever its t.
__device__ bool any_hit(const BvhView& B, const RayD& r, HitRec& hit, uint32_t* overflow, TravStats& st,
                        int* lds_col) {
  hit.pos = -1;
  const int n = B.n;
  if (n <= 0) return false;
  double t0, t1, t, al, be;
  const float* rb = B.node_box;
  if (!slab(r, rb[0], rb[1], rb[2], rb[3], rb[4], rb[5], t0, t1)) return false;
  if (n == 1) {
    ++st.tests;
    if (tri_test(r, B.ltri, t, al, be)) {
      hit.t = t; hit.al = al; hit.be = be; hit.pos = 0;
      return true;
    }
    return false;
  }
  LaneStack S;
  S.col = lds_col;
  int spill[kSpill];
  int node = 0;
  for (int guard = 0; guard < n; ++guard) {
    float4 q0, q1, q2, q3;
    load_node(B.tnode, node, q0, q1, q2, q3);
    st.visits += 2;
    double l0, l1, r0, r1;
    bool hl = slab(r, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, l0, l1);
    bool hr = slab(r, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, r0, r1);
    int cl = __float_as_int(q3.x), cr = __float_as_int(q3.y);
    int next = -1;
    bool have_next = false;
    if (hl) {
      if (cl < 0) {
        ++st.tests;
        if (tri_test(r, B.ltri + kLeafStride * (size_t)(~cl), t, al, be)) {
          hit.t = t; hit.al = al; hit.be = be; hit.pos = ~cl;
          return true;
        }
      } else {
        if (hr && !stack_push<kRing, kThreads, false>(S, spill, cr)) {
          atomicAdd(overflow, 1u);
          return false;
        }
        next = cl;
        have_next = true;
      }
    }
    if (!have_next && hr) {
      if (cr < 0) {
        ++st.tests;
        if (tri_test(r, B.ltri + kLeafStride * (size_t)(~cr), t, al, be)) {
          hit.t = t; hit.al = al; hit.be = be; hit.pos = ~cr;
          return true;
        }
      } else {
        next = cr;
        have_next = true;
      }
    }
    while (!have_next) {
      if (stack_empty<false>(S)) return false;
      int ref = stack_pop<kRing, kThreads, false>(S, spill);
      if (ref < 0) {
        ++st.tests;
        if (tri_test(r, B.ltri + kLeafStride * (size_t)(~ref), t, al, be)) {
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
  return __float_as_int(B.ltri[kLeafStride * (size_t)pos + 2].y);
}

__device__ __forceinline__ int32_t leaf_tri_of(const BvhView& B, int pos) {
  return __float_as_int(ldgf((const float*)(B.ltri + kLeafStride * (size_t)pos + 2) + 1));
}

// Work counts of a traversal launch: child-box tests to stats[0], triangle
// tests to stats[1] of replica (workgroup & 63), replicas 16 words apart.
__device__ __forceinline__ void flush_stats(unsigned long long* stats, const TravStats& st) {
  if (!stats) return;
  unsigned long long vis = st.visits, tst = st.tests;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    vis += __shfl_xor(vis, off);
    tst += __shfl_xor(tst, off);
  }
  if ((threadIdx.x & 63) == 0) {
    unsigned long long* r = stats + (blockIdx.x & 63) * 16;
    if (vis) atomicAdd(&r[0], vis);
    if (tst) atomicAdd(&r[1], tst);
  }
}

// ------------------------------------------------------------ 4-wide nodes
// The reference tests a leaf's triangle when the FP64 slab test passes on
// the leaf's box and on every ancestor's box. Refit boxes nest exactly and
// the rounded FP64 test is monotone in the box (if a box passes, every box
// containing it passes), so that set is exactly the leaves whose own box
// passes. Internal children may therefore use any test that never rejects
// a box the FP64 test accepts: the 4-wide nodes use an fp32 slab test whose
// distances are widened by E = 6 eps (B + |o|) |1/d| (eps = 2^-24, B the
// largest coordinate magnitude of the scene box), which covers the rounding
// of 1/d, of o/d, of the decoded planes of quantized nodes and of the fused
// multiply-adds; a leaf child is decided by the FP64 test on its exact fp32
// box, as in the binary traversal.
struct RayF {
  float ix, iy, iz;  // sc / d (sc * +-1e30 for a zero component)
  float nx, ny, nz;  // sc * (o/d + E): near-plane offsets
  float fx, fy, fz;  // sc * (o/d - E): far-plane offsets
  float sx, sy, sz;  // 2^24 * sc / d: the plane scale of quantized nodes decoded as fp16 subnormals
  float sc;          // 1, or the power of two 2^-k that keeps every step * 2^24 * sc / d finite
};
// Every fp32 distance of a ray is its unscaled value times sc, so the child
// tests and the child order do not change: sc is a power of two, which
// scales exactly in the normal range, and the widening E, scaled with the
// rest, stays far above the subnormal range for any direction whose
// nonzero components are within a factor 2^180 of each other. The pruning bounds
// compared with these distances are scaled too (scaled_up / scaled_down).

// a = 1/d of the axis (RayD's ax, ay or az), clamped to +-1e30
__device__ __forceinline__ double clamped_inv(double a, double d) {
  return fabs(a) <= 1e30 ? a : copysign(1e30, d);
}

__device__ __forceinline__ void rayf_axis(double o, double inv, double bmax, float& iv, float& nn, float& ff) {
  const double oi = o * inv;
  const double E = 6.0 * 5.9604644775390625e-08 * (bmax + fabs(o)) * fabs(inv);
  iv = (float)inv;
  nn = (float)(oi + E);
  ff = (float)(oi - E);
}

__device__ __forceinline__ RayF make_rayf(const RayD& r, const float* root_box) {
  double bmax = 0;
  float b[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    b[k] = ldgf(root_box + k);
    bmax = fmax(bmax, fabs((double)b[k]));
  }
  const double ix = clamped_inv(r.ax, r.d.x), iy = clamped_inv(r.ay, r.d.y), iz = clamped_inv(r.az, r.d.z);
  RayF f;
  rayf_axis(r.o.x, ix, bmax, f.ix, f.nx, f.fx);
  rayf_axis(r.o.y, iy, bmax, f.iy, f.ny, f.fy);
  rayf_axis(r.o.z, iz, bmax, f.iz, f.nz, f.fz);
  f.sc = 1.0f;
  // A quantized node's step is the smallest power of two >= its extent /
  // 255 (2^-126 at least), so at most 2 * extent / 255 above that, and no
  // node is wider than the root: while max(extent * |1/d|) < 2^109, every
  // step * 2^24 / d is below 2^127. Beyond it (a zero direction component
  // in a scene a few thousand units wide) the ray's fp32 values are scaled
  // down by sc (exactly: a power of two); a non-finite root box takes the
  // smallest scale.
  const double sd = fmax(fmax(((double)b[3] - b[0]) * fabs(ix), ((double)b[4] - b[1]) * fabs(iy)),
                         ((double)b[5] - b[2]) * fabs(iz));
  if (!(sd < 0x1p109)) {
    // finite boxes give sd < 2^230 (extent < 2^129, |1/d| <= 1e30), so sc >= 2^-121
    const float sc = sd < 0x1p230 ? ldexpf(1.0f, 108 - ilogb(sd)) : 0x1p-126f;
    f.ix *= sc; f.iy *= sc; f.iz *= sc;
    f.nx *= sc; f.ny *= sc; f.nz *= sc;
    f.fx *= sc; f.fy *= sc; f.fz *= sc;
    f.sc = sc;
  }
  f.sx = f.ix * 16777216.0f;  // exact: |sc / d| <= 1e30
  f.sy = f.iy * 16777216.0f;
  f.sz = f.iz * 16777216.0f;
  return f;
}

// Float bounds of a double, rounded outward.
__device__ __forceinline__ float f_up(double x) {
  float f = (float)x;
  return (double)f < x ? __int_as_float(__float_as_int(f) + (f >= 0 ? 1 : -1)) : f;
}
__device__ __forceinline__ float f_down(double x) {
  float f = (float)x;
  return (double)f > x ? __int_as_float(__float_as_int(f) + (f > 0 ? -1 : 1)) : f;
}
// A distance bound in a ray's scaled fp32 units, rounded outward.
__device__ __forceinline__ float scaled_up(const RayF& f, double x) { return f_up(x * (double)f.sc); }
__device__ __forceinline__ float scaled_down(const RayF& f, double x) { return f_down(x * (double)f.sc); }

__device__ __forceinline__ float f4c(const float4& v, int j) { return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w; }

// The mirror image of the conservative child test: with the near and far
// offsets swapped, each fp32 entry distance is >= the reference's FP64 one
// and each exit distance <= it (the same bound E on the same roundings, now
// on the other side), so a box this test accepts is accepted by the FP64
// slab test (inc/bbox.h:82-102) too. Used for a leaf about to be taken: it
// is rejected only when the line passes within about E of the box's
// boundary (for an axis with a zero direction component, when the origin
// lies within 6 * 2^-24 * (B + |o|) of that slab's planes); the FP64 test
// decides those.
__device__ __forceinline__ bool inner_slab(const RayD& r, const RayF& f, float lx, float ly, float lz, float hx,
                                           float hy, float hz) {
  const float nx = r.sx ? lx : hx, fx = r.sx ? hx : lx;
  const float ny = r.sy ? ly : hy, fy = r.sy ? hy : ly;
  const float nz = r.sz ? lz : hz, fz = r.sz ? hz : lz;
  const float t0 = fmaxf(fmaxf(fmaf(nx, f.ix, -f.fx), fmaf(ny, f.iy, -f.fy)), fmaf(nz, f.iz, -f.fz));
  const float t1 = fminf(fminf(fmaf(fx, f.ix, -f.nx), fmaf(fy, f.iy, -f.ny)), fmaf(fz, f.iz, -f.nz));
  return t0 <= t1;
}

__device__ __forceinline__ void load_wide(const float4* __restrict__ wide, int node, float4 (&w)[7]) {
  const float4* p = wide + 8 * (size_t)node;
#pragma unroll
  for (int k = 0; k < 7; ++k) w[k] = ldg4(p + k);
}

// Conservative fp32 entry/exit distances of the four children.
__device__ __forceinline__ void wide_slabs(const RayD& r, const RayF& f, const float4 (&w)[7], float (&t0)[4],
                                           float (&t1)[4]) {
  const float4 nx = r.sx ? w[0] : w[3], fx = r.sx ? w[3] : w[0];
  const float4 ny = r.sy ? w[1] : w[4], fy = r.sy ? w[4] : w[1];
  const float4 nz = r.sz ? w[2] : w[5], fz = r.sz ? w[5] : w[2];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const float a = fmaf(f4c(nx, j), f.ix, -f.nx), b = fmaf(f4c(ny, j), f.iy, -f.ny), c = fmaf(f4c(nz, j), f.iz, -f.nz);
    const float d = fmaf(f4c(fx, j), f.ix, -f.fx), e = fmaf(f4c(fy, j), f.iy, -f.fy), g = fmaf(f4c(fz, j), f.iz, -f.fz);
    t0[j] = fmaxf(fmaxf(a, b), c);
    t1[j] = fminf(fminf(d, e), g);
  }
}

// fma((float)h, a, c) for the fp16 value h in the low (hi = 0) or high half
// of h2: one v_fma_mix_f32.
__device__ __forceinline__ float fma_h16(uint32_t h2, int hi, float a, float c) {
  const _Float16 h = __builtin_bit_cast(_Float16, (uint16_t)(hi ? h2 >> 16 : h2 & 0xffffu));
  return fmaf((float)h, a, c);
}

// The same distances from a quantized node: plane = corner + q * step,
// folded into t = q * (step / d) + (corner / d - offset). A v_perm_b32
// spreads two plane bytes into two fp16 halves with a zero byte above each,
// i.e. the fp16 subnormals q * 2^-24 (exact), which v_fma_mix_f32 takes
// directly: fma(q * 2^-24, step * 2^24 / d, c) is fma(q, step / d, c) bit
// for bit, since step * 2^24 / d is exact (step is a power of two) and the
// product is the same real number -- one instruction per plane instead of a
// byte conversion and an fma. (With the ray's scale sc, step * 2^24 * sc / d
// is below 2^127, so it is finite.)
__device__ __forceinline__ void wideq_slabs(const RayD& r, const RayF& f, const float4 (&q)[4], float (&t0)[4],
                                            float (&t1)[4]) {
  const float cnx = fmaf(q[0].x, f.ix, -f.nx), cfx = fmaf(q[0].x, f.ix, -f.fx);
  const float cny = fmaf(q[0].y, f.iy, -f.ny), cfy = fmaf(q[0].y, f.iy, -f.fy);
  const float cnz = fmaf(q[0].z, f.iz, -f.nz), cfz = fmaf(q[0].z, f.iz, -f.fz);
  const uint32_t lx = __float_as_uint(q[1].x), ly = __float_as_uint(q[1].y), lz = __float_as_uint(q[1].z);
  const uint32_t hx = __float_as_uint(q[2].x), hy = __float_as_uint(q[2].y), hz = __float_as_uint(q[2].z);
  const uint32_t nx = r.sx ? lx : hx, fx = r.sx ? hx : lx;
  const uint32_t ny = r.sy ? ly : hy, fy = r.sy ? hy : ly;
  const uint32_t nz = r.sz ? lz : hz, fz = r.sz ? hz : lz;
  const float ax = q[0].w * f.sx, ay = q[1].w * f.sy, az = q[2].w * f.sz;  // step * 2^24 / d
  const uint32_t h[6][2] = {
      {__builtin_amdgcn_perm(nx, nx, 0x0C010C00u), __builtin_amdgcn_perm(nx, nx, 0x0C030C02u)},
      {__builtin_amdgcn_perm(ny, ny, 0x0C010C00u), __builtin_amdgcn_perm(ny, ny, 0x0C030C02u)},
      {__builtin_amdgcn_perm(nz, nz, 0x0C010C00u), __builtin_amdgcn_perm(nz, nz, 0x0C030C02u)},
      {__builtin_amdgcn_perm(fx, fx, 0x0C010C00u), __builtin_amdgcn_perm(fx, fx, 0x0C030C02u)},
      {__builtin_amdgcn_perm(fy, fy, 0x0C010C00u), __builtin_amdgcn_perm(fy, fy, 0x0C030C02u)},
      {__builtin_amdgcn_perm(fz, fz, 0x0C010C00u), __builtin_amdgcn_perm(fz, fz, 0x0C030C02u)}};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int w = j >> 1, s = j & 1;
    const float a = fma_h16(h[0][w], s, ax, cnx), b = fma_h16(h[1][w], s, ay, cny), c = fma_h16(h[2][w], s, az, cnz);
    const float d = fma_h16(h[3][w], s, ax, cfx), e = fma_h16(h[4][w], s, ay, cfy), g = fma_h16(h[5][w], s, az, cfz);
    t0[j] = fmaxf(fmaxf(a, b), c);
    t1[j] = fminf(fminf(d, e), g);
  }
}

// Eight children of an 8-wide quantized node.
__device__ __forceinline__ void wide8_slabs(const RayF& f, const float4 (&q)[6], float (&t0)[8], float (&t1)[8]) {
  const uint32_t eb = __float_as_uint(q[0].w);
  const float sx = __uint_as_float((eb & 0xffu) << 23), sy = __uint_as_float(((eb >> 8) & 0xffu) << 23),
              sz = __uint_as_float(((eb >> 16) & 0xffu) << 23);
  const float ax = sx * f.ix, ay = sy * f.iy, az = sz * f.iz;
  const float cnx = fmaf(q[0].x, f.ix, -f.nx), cfx = fmaf(q[0].x, f.ix, -f.fx);
  const float cny = fmaf(q[0].y, f.iy, -f.ny), cfy = fmaf(q[0].y, f.iy, -f.fy);
  const float cnz = fmaf(q[0].z, f.iz, -f.nz), cfz = fmaf(q[0].z, f.iz, -f.fz);
  const bool px = f.ix >= 0, py = f.iy >= 0, pz = f.iz >= 0;
  const uint32_t lx[2] = {__float_as_uint(q[1].x), __float_as_uint(q[1].y)};
  const uint32_t ly[2] = {__float_as_uint(q[1].z), __float_as_uint(q[1].w)};
  const uint32_t lz[2] = {__float_as_uint(q[2].x), __float_as_uint(q[2].y)};
  const uint32_t hx[2] = {__float_as_uint(q[2].z), __float_as_uint(q[2].w)};
  const uint32_t hy[2] = {__float_as_uint(q[3].x), __float_as_uint(q[3].y)};
  const uint32_t hz[2] = {__float_as_uint(q[3].z), __float_as_uint(q[3].w)};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t nx = px ? lx[h] : hx[h], fx = px ? hx[h] : lx[h];
    const uint32_t ny = py ? ly[h] : hy[h], fy = py ? hy[h] : ly[h];
    const uint32_t nz = pz ? lz[h] : hz[h], fz = pz ? hz[h] : lz[h];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int sh = 8 * j;
      const float a = fmaf((float)((nx >> sh) & 0xffu), ax, cnx), b = fmaf((float)((ny >> sh) & 0xffu), ay, cny),
                  c = fmaf((float)((nz >> sh) & 0xffu), az, cnz);
      const float d = fmaf((float)((fx >> sh) & 0xffu), ax, cfx), e = fmaf((float)((fy >> sh) & 0xffu), ay, cfy),
                  g = fmaf((float)((fz >> sh) & 0xffu), az, cfz);
      t0[4 * h + j] = fmaxf(fmaxf(a, b), c);
      t1[4 * h + j] = fminf(fminf(d, e), g);
    }
  }
}

// The FP64 slab test of child j's exact box.
__device__ __forceinline__ bool wide_leaf_slab(const RayD& r, const float4 (&w)[7], int j, double& t0, double& t1) {
  return slab(r, f4c(w[0], j), f4c(w[1], j), f4c(w[2], j), f4c(w[3], j), f4c(w[4], j), f4c(w[5], j), t0, t1);
}

// ------------------------------------------------------------ persistent traversal
// Ray-queue traversal with dynamic ray fetch: each lane walks one ray a node
// per iteration and, as soon as its ray finishes, takes the next ray index
// from a chunk its wave reserved with one atomic (fetch_chunk rays), so the
// wave keeps its lanes busy until the queue drains instead of idling behind
// its longest ray. Closest-hit and any-hit use the same decisions as
// closest_hit()/any_hit() above, one node (or, for any-hit, one popped leaf)
// per step.
#ifndef RT_REFILL_MIN
#define RT_REFILL_MIN 16
#endif
constexpr int kRefillMin = RT_REFILL_MIN;  // idle lanes that trigger a refill (1..64)
#ifndef RT_LEAF_BATCH
#define RT_LEAF_BATCH 16
#endif
constexpr int kLeafBatch = RT_LEAF_BATCH;  // waiting leaves that trigger a batch of leaf tests
constexpr int kNoNode = 0x7fffffff;        // no node to visit (never a node index)
constexpr uint32_t kSegments = 8;       // one per XCD
constexpr uint32_t kFetchStride = 32;   // counters 128 bytes apart
constexpr uint32_t kFetchWords = kSegments * kFetchStride;

struct RayIn {
  const double* orig;  // component k of ray i's origin at orig[i*stride + k*comp]
  const double* dir;   // component k of ray i's direction at dir[i*stride + k*comp]
  size_t stride;       // between rays, in doubles
  size_t comp;         // between components, in doubles (1: xyz interleaved; the queue capacity: SoA)
  const uint32_t* count_dev;  // number of rays (device), or nullptr
  uint32_t count;             // used when count_dev is nullptr
  const uint32_t* perm;       // optional fetch order: the k-th ray fetched is ray perm[k]
  // Any hit only: a second queue in the same launch. Its *count_dev2 rays
  // follow the first queue's in fetch order, and each of its buffers (rays,
  // perm, scene, rec16) lies slot_bytes after the first queue's.
  const uint32_t* count_dev2;
  size_t slot_bytes;
};

// The wavefront passes keep each traversal result in one record, written
// with one or two 16-byte stores into a single 32-byte sector: the
// closest-hit record of a queue slot holds the sorted leaf position (-1: no
// hit) and t, alpha, beta; the any-hit record the leaf position and t. Split
// over separate arrays, the results of rays that finish at different times
// were written back as several partial sectors each.
struct alignas(32) HitRec32 {
  int32_t pos, pad;
  double t, al, be;
};
struct alignas(16) HitRec16 {
  int32_t pos, pad;
  double t;
};

struct HitOut {
  int32_t* pos;  // sorted leaf position or -1 (nullable)
  int32_t* tri;  // triangle id or -1 (nullable)
  double *t, *al, *be;  // nullable
  HitRec32* rec32;      // nullable: the record of ray i at rec32[i]
  HitRec16* rec16;      // nullable: the record of the k-th ray fetched at rec16[k]
};

// Results of ray i, the k-th ray fetched (k = i without a fetch order).
// WFO (the wavefront passes): only the record of the kernel's kind; the
// other outputs are compiled out, so their pointers take no registers.
template <bool ANY, bool WFO>
__device__ __forceinline__ void write_hit(const BvhView& B, const HitOut& O, uint32_t i, uint32_t k, int pos,
                                          double t, double al, double be) {
  if constexpr (WFO) {
    if constexpr (ANY) {
      HitRec16 h;
      h.pos = pos;
      h.pad = 0;
      h.t = pos >= 0 ? t : 0.0;
      O.rec16[k] = h;
    } else {
      HitRec32 h;
      h.pos = pos;
      h.pad = 0;
      h.t = pos >= 0 ? t : 0.0;
      h.al = pos >= 0 ? al : 0.0;
      h.be = pos >= 0 ? be : 0.0;
      O.rec32[i] = h;
    }
    return;
  }
  if (O.pos) O.pos[i] = pos;
  if (O.tri) O.tri[i] = pos >= 0 ? leaf_tri_of(B, pos) : -1;
  if (O.t) O.t[i] = pos >= 0 ? t : 0.0;
  if (O.al) O.al[i] = pos >= 0 ? al : 0.0;
  if (O.be) O.be[i] = pos >= 0 ? be : 0.0;
  if (O.rec32) {
    HitRec32 h;
    h.pos = pos;
    h.pad = 0;
    h.t = pos >= 0 ? t : 0.0;
    h.al = pos >= 0 ? al : 0.0;
    h.be = pos >= 0 ? be : 0.0;
    O.rec32[i] = h;
  }
  if (O.rec16) {
    HitRec16 h;
    h.pos = pos;
    h.pad = 0;
    h.t = pos >= 0 ? t : 0.0;
    O.rec16[k] = h;
  }
}

// views[scene[i]] is the BVH of ray i (scene == nullptr: views[0] for all).
// Closest-hit child keys: key = entry distance of a child the ray enters
// within [low, lim], +inf otherwise; returns the number of such children.
// LOW = false leaves out the test against low (for low = -inf). An unused
// slot needs no test of its own: its box is inverted, and it names the
// null leaf, which no ray hits.
template <bool LOW, int NW>
__device__ __forceinline__ int child_keys(const float (&t0)[NW], const float (&t1)[NW], const int (&cref)[NW],
                                          float lim, float low, float (&key)[NW], int (&ref)[NW]) {
  int m = 0;
#pragma unroll
  for (int j = 0; j < NW; ++j) {
    ref[j] = cref[j];
    bool h = (t0[j] <= t1[j]) & (t0[j] <= lim);
    if (LOW) h = h & (t1[j] >= low);
    key[j] = h ? t0[j] : __int_as_float(0x7f800000);
    m += h ? 1 : 0;
  }
  return m;
}

// LAYOUT 2: binary nodes, FP64 child tests; 4: 4-wide nodes (128 B);
// 5: quantized 4-wide nodes (64 B).
// Workgroups of the persistent traversal (rt_tuning::wg_threads): one wave
// by default, so a wave that has finished its rays frees its slot at once
// (for the kernel that runs concurrently on the other stream) instead of
// waiting for three others; four waves share one home ray segment.

// LOW = false: a closest-hit launch with tmin 0, whose child test can leave
// out the lower bound (t1 >= -inf is implied by t0 <= t1).
template <bool ANY, int RING, int LAYOUT, bool MULTI, bool LOW, int THREADS, bool WFO>
__global__ __launch_bounds__(THREADS) void k_trace_persistent(const BvhView* __restrict__ views,
                                                               const uint16_t* __restrict__ scene, RayIn R,
                                                               double tmin_phys, HitOut O, uint32_t* fetch_ctr,
                                                               uint32_t fetch_chunk, uint32_t* overflow,
                                                               unsigned long long* stats) {
  constexpr bool WIDE = LAYOUT != 2, QUANT = LAYOUT == 5 || LAYOUT == 8;
  __shared__ int s_stack[RING * THREADS];
  const uint32_t nrays0 = R.count_dev ? *R.count_dev : R.count;
  const uint32_t nrays = nrays0 + (ANY && R.count_dev2 ? *R.count_dev2 : 0u);
  const uint32_t kFetchChunk = fetch_chunk;
  BvhView B = views[0];
  int n = B.n;
  const uint32_t lane = __lane_id();
  const uint64_t mlt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  TravStats st;
  LaneStack S;
  S.col = s_stack + threadIdx.x;
  int spill[kSpill];
  RayD r;
  RayF rf;
  bool active = false;
  uint32_t rid = 0, kid = 0;  // the lane's ray and its fetch index
  int node = 0;
  int pending = kNoNode;  // 4-wide traversal: a leaf waiting for the wave's next batch of leaf tests
  double bt = 0, bal = 0, bbe = 0;
  int bpos = -1;
  float flim = __int_as_float(0x7f800000);  // 4-wide closest hit: fp32 pruning bound of the best hit so far
  uint32_t chunk_next = 0, chunk_end = 0;
  bool drained = false;
  // ray range split into kSegments contiguous segments of whole chunks
  const uint32_t seg_len = ((nrays + kSegments - 1) / kSegments + kFetchChunk - 1) / kFetchChunk * kFetchChunk;
  uint32_t seg = blockIdx.x % kSegments, tries = 0;
  while (true) {
    // ---- refill idle lanes from the wave's chunk
    // Refill only once enough lanes are idle: the ray setup (three FP64
    // divisions, the root-box test) runs for the whole wave when any lane
    // takes a ray, so taking rays in groups keeps it off most iterations.
    uint64_t need = __ballot(!active);
    if (need != 0 && !drained && __popcll(need) >= kRefillMin) {
      uint32_t k = (uint32_t)__popcll(need);
      uint32_t rank = (uint32_t)__popcll(need & mlt);
      uint32_t avail = chunk_end - chunk_next;
      uint32_t myid = 0xffffffffu;
      if (!active && rank < avail) myid = chunk_next + rank;
      uint32_t used = min(k, avail);
      chunk_next += used;
      if (k > avail) {
        // Next chunk: the workgroup's home segment first (one per XCD, so
        // an XCD's L2 serves one screen region), then the others in turn.
        uint32_t base = 0xffffffffu, end = 0;
        while (tries < kSegments) {
          uint32_t seg_lo = seg * seg_len, seg_hi = min(nrays, seg_lo + seg_len);
          uint32_t b = 0;
          if (lane == 0) b = atomicAdd(fetch_ctr + seg * kFetchStride, kFetchChunk);
          b = __shfl(b, 0) + seg_lo;
          if (b < seg_hi) {
            base = b;
            end = min(b + kFetchChunk, seg_hi);
            break;
          }
          seg = (seg + 1) % kSegments;
          ++tries;
        }
        if (base == 0xffffffffu) {
          drained = true;
        } else {
          chunk_next = base;
          chunk_end = end;
          uint32_t avail2 = chunk_end - chunk_next;
          uint32_t r2 = rank - avail;
          if (!active && rank >= avail && r2 < avail2) myid = chunk_next + r2;
          chunk_next += min(k - avail, avail2);
        }
      }
      if (myid != 0xffffffffu) {
        // second queue of a paired any-hit launch: its buffers are the
        // first queue's moved by slot_bytes (a multiple of 256)
        size_t so = 0;
        if (ANY && myid >= nrays0) {
          myid -= nrays0;
          so = R.slot_bytes;
        }
        kid = myid + (uint32_t)(so / sizeof(HitRec16));
        rid = R.perm ? ((const uint32_t*)((const char*)R.perm + so))[myid] : myid;
        if (MULTI) {
          B = views[((const uint16_t*)((const char*)scene + so))[rid]];
          n = B.n;
        }
        const double* po = (const double*)((const char*)R.orig + so) + (size_t)rid * R.stride;
        const double* pd = (const double*)((const char*)R.dir + so) + (size_t)rid * R.stride;
        r = make_ray(mk(ldgd(po), ldgd(po + R.comp), ldgd(po + 2 * R.comp)),
                     mk(ldgd(pd), ldgd(pd + R.comp), ldgd(pd + 2 * R.comp)));
        if (WIDE) rf = make_rayf(r, B.node_box);
        bpos = -1;
        bt = 0;
        flim = __int_as_float(0x7f800000);
        double t0, t1;
        const float* rb = B.node_box;
        if (n <= 0 || !slab(r, ldgf(rb), ldgf(rb + 1), ldgf(rb + 2), ldgf(rb + 3), ldgf(rb + 4), ldgf(rb + 5), t0, t1)) {
          write_hit<ANY, WFO>(B, O, rid, kid, -1, 0, 0, 0);
        } else if (n == 1) {
          double t, al, be;
          ++st.tests;
          bool ok = tri_test(r, B.ltri, t, al, be);
          if (!ANY && tmin_phys > 0 && !(t >= tmin_phys)) ok = false;
          write_hit<ANY, WFO>(B, O, rid, kid, ok ? 0 : -1, t, al, be);
        } else {
          active = true;
          node = 0;
          pending = kNoNode;
          S.d = 0;
          S.lo = 0;
        }
      }
    }
    if (__ballot(active) == 0) {
      if (drained) break;
      continue;
    }
    if (!active) continue;
    // ---- one traversal step
    bool done = false;
    if (WIDE) {
      // Leaf tests are batched across the wave: a lane that reaches a leaf
      // parks it in `pending` and keeps walking its stack; the wave tests
      // the parked leaves together once kLeafBatch lanes have one, or as
      // soon as a lane cannot go on (a second leaf, or nothing left to
      // walk). Leaves are still tested in the order the lane reaches them,
      // so any-hit keeps its left-first order.
      const bool has_pending = pending != kNoNode;
      const bool blocked = has_pending && (node == kNoNode || node < 0);
      const uint64_t wp = __ballot(has_pending);
      if (__ballot(blocked) != 0 || __popcll(wp) >= kLeafBatch) {
        if (has_pending) {
          const int pos = ~pending;
          pending = kNoNode;
          const float4* L = B.ltri + kLeafStride * (size_t)pos;
          // The reference tests a triangle only if the line enters its box
          // (the leaf's box test, kernels/bvh.cu:382-449). The quantized
          // nodes hold no exact leaf boxes, so that test is made here, after
          // the triangle test and only for a triangle that would be taken:
          // a hit needs both, and the box test is then rarely run at all.
          double t, al, be;
          ++st.tests;
          bool take = tri_test(r, L, t, al, be);
          if (!ANY) {
            if (take && tmin_phys > 0 && !(t >= tmin_phys)) take = false;
            // closest hit: the minimum (t, leaf position), as consider_leaf
            if (take && bpos >= 0 && !(t < bt || (t == bt && pos < bpos))) take = false;
          }
          if (take && (QUANT || !ANY)) {
            // leaf record: q0 = a.xyz, b.x; q1 = b.yz, c.xy; q2 = c.z, id. The
            // box is formed exactly as the build forms the triangle box.
            const float4 q0 = ldg4(L), q1 = ldg4(L + 1), q2 = ldg4(L + 2);
            const float ax = q0.x, ay = q0.y, az = q0.z, bx = q0.w, by = q1.x, bz = q1.y, cx = q1.z, cy = q1.w,
                        cz = q2.x;
            const float lx = fminf(cx, fminf(ax, bx)), ly = fminf(cy, fminf(ay, by)), lz = fminf(cz, fminf(az, bz));
            const float hx = fmaxf(cx, fmaxf(ax, bx)), hy = fmaxf(cy, fmaxf(ay, by)), hz = fmaxf(cz, fmaxf(az, bz));
            // the FP64 test only where the fp32 inner test cannot decide
            // (and for every ray whose fp32 distances are scaled)
            if (rf.sc != 1.0f || !inner_slab(r, rf, lx, ly, lz, hx, hy, hz)) {
              double l0, l1;
              take = slab(r, lx, ly, lz, hx, hy, hz, l0, l1);
            }
          }
          if (take) {
            bt = t;
            bal = al;
            bbe = be;
            bpos = pos;
            if (ANY) done = true;
            else flim = scaled_up(rf, bt + prune_margin(bt));
          }
        }
      }
      if (!done && node != kNoNode && node < 0 && pending == kNoNode) {
        pending = node;
        node = stack_empty<!ANY>(S) ? kNoNode : stack_pop<RING, THREADS, !ANY>(S, spill);
      }
      if (!done && node != kNoNode && node >= 0) {
        constexpr int NW = LAYOUT == 8 ? 8 : 4;
        float4 w[7];
        float t0[NW], t1[NW];
        int cref[NW];
        if constexpr (LAYOUT == 8) {
          float4 q[6];
          const float4* p = B.wide8 + 8 * (size_t)node;
#pragma unroll
          for (int k = 0; k < 6; ++k) q[k] = ldg4(p + k);
          wide8_slabs(rf, q, t0, t1);
#pragma unroll
          for (int j = 0; j < NW; ++j) cref[j] = __float_as_int(f4c(q[4 + j / 4], j & 3));
        } else if constexpr (QUANT) {
          float4 q[4];
          const float4* p = B.wideq + 4 * (size_t)node;
#pragma unroll
          for (int k = 0; k < 4; ++k) q[k] = ldg4(p + k);
          wideq_slabs(r, rf, q, t0, t1);
#pragma unroll
          for (int j = 0; j < NW; ++j) cref[j] = __float_as_int(f4c(q[3], j & 3));
        } else {
          load_wide(B.wide, node, w);
          wide_slabs(r, rf, w, t0, t1);
#pragma unroll
          for (int j = 0; j < NW; ++j) cref[j] = __float_as_int(f4c(w[6], j & 3));
        }
        // child-box tests of the filled slots only (an unused slot names
        // the null leaf); counted only when a caller asked
        if (stats) {
#pragma unroll
          for (int j = 0; j < NW; ++j) st.visits += cref[j] != kNullLeaf(n) ? 1u : 0u;
        }
        bool ovf = false;
        int next = kNoNode;
        if (!ANY) {
          const float lim = flim;
          const float low =
              tmin_phys > 0 ? scaled_down(rf, tmin_phys - prune_margin(tmin_phys)) : __int_as_float(0xff800000);
          float key[NW];
          int ref[NW];
          const int m = child_keys<LOW>(t0, t1, cref, lim, low, key, ref);
          // nearest child next, the others pushed far to near (sorting
          // networks of 5 and 19 comparators)
#define RT_CSWAP(a, b)                                  \
  if (key[b] < key[a]) {                                \
    float tk = key[a]; key[a] = key[b]; key[b] = tk;    \
    int tr = ref[a]; ref[a] = ref[b]; ref[b] = tr;      \
  }
          if constexpr (NW == 4) {
            RT_CSWAP(0, 1) RT_CSWAP(2, 3) RT_CSWAP(0, 2) RT_CSWAP(1, 3) RT_CSWAP(1, 2)
          } else {
            RT_CSWAP(0, 2) RT_CSWAP(1, 3) RT_CSWAP(4, 6) RT_CSWAP(5, 7) RT_CSWAP(0, 4) RT_CSWAP(1, 5) RT_CSWAP(2, 6)
            RT_CSWAP(3, 7) RT_CSWAP(0, 1) RT_CSWAP(2, 3) RT_CSWAP(4, 5) RT_CSWAP(6, 7) RT_CSWAP(2, 4) RT_CSWAP(3, 5)
            RT_CSWAP(1, 4) RT_CSWAP(3, 6) RT_CSWAP(1, 2) RT_CSWAP(3, 4) RT_CSWAP(5, 6)
          }
#undef RT_CSWAP
          if constexpr (NW == 4) {
            if (S.d <= RING - 3) {
              // room for three entries: push ref[m-1] .. ref[1] without a
              // branch per entry. Slots d .. d+2 are free, so the writes
              // past the m-1 kept entries are harmless.
              const int a0 = m >= 4 ? ref[3] : (m == 3 ? ref[2] : ref[1]);
              const int a1 = m >= 4 ? ref[2] : ref[1];
              int* top = S.col + S.d * THREADS;
              top[0] = a0;
              top[THREADS] = a1;
              top[2 * THREADS] = ref[1];
              S.d += m > 1 ? m - 1 : 0;
            } else {
#pragma unroll
              for (int k = NW - 1; k >= 1; --k)
                if (m > k) ovf = ovf || !stack_push<RING, THREADS, !ANY>(S, spill, ref[k]);
            }
          } else {
#pragma unroll
            for (int k = NW - 1; k >= 1; --k)
              if (m > k) ovf = ovf || !stack_push<RING, THREADS, !ANY>(S, spill, ref[k]);
          }
          if (m >= 1) next = ref[0];
        } else {
          // children in sorted-leaf order: the first passing one next, the
          // others pushed so that they pop in order (left-first DFS)
          if (LAYOUT != 4 && NW == 4 && S.d - S.lo <= RING - 3) {
            // room for three entries: plain ring writes, no per-entry check
#pragma unroll
            for (int j = NW - 1; j >= 0; --j) {
              const int ref = cref[j];
              if (t0[j] <= t1[j]) {
                if (next != kNoNode) {
                  S.col[(S.d & (RING - 1)) * THREADS] = next;  // ring stack
                  ++S.d;
                }
                next = ref;
              }
            }
          } else {
#pragma unroll
            for (int j = NW - 1; j >= 0; --j) {
              const int ref = cref[j];
              bool p = t0[j] <= t1[j];
              if (LAYOUT == 4 && p && ref < 0) {
                double l0, l1;
                p = wide_leaf_slab(r, w, j & 3, l0, l1);  // exact decision for a leaf child
              }
              if (p) {
                if (next != kNoNode) ovf = ovf || !stack_push<RING, THREADS, !ANY>(S, spill, next);
                next = ref;
              }
            }
          }
        }
        if (ovf) {
          atomicAdd(overflow, 1u);
          bpos = -1;
          done = true;
        } else if (next != kNoNode) {
          node = next;
        } else {
          node = stack_empty<!ANY>(S) ? kNoNode : stack_pop<RING, THREADS, !ANY>(S, spill);
        }
      }
      if (!done && node == kNoNode && pending == kNoNode) done = true;
    } else if (!ANY) {
      float4 q0, q1, q2, q3;
      load_node(B.tnode, node, q0, q1, q2, q3);
      st.visits += 2;
      double l0, l1, r0, r1;
      bool hl = slab(r, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, l0, l1);
      bool hr = slab(r, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, r0, r1);
      int cl = __float_as_int(q3.x), cr = __float_as_int(q3.y);
      if (tmin_phys > 0) {
        hl = hl && !(l1 < tmin_phys - prune_margin(tmin_phys));
        hr = hr && !(r1 < tmin_phys - prune_margin(tmin_phys));
      }
      if (bpos >= 0) {
        double lim = bt + prune_margin(bt);
        hl = hl && !(l0 > lim);
        hr = hr && !(r0 > lim);
      }
      HitRec best{bt, bal, bbe, bpos};
      if (hl && cl < 0) {
        consider_leaf(B, r, ~cl, tmin_phys, best, st);
        hl = false;
      }
      if (hr && cr < 0) {
        consider_leaf(B, r, ~cr, tmin_phys, best, st);
        hr = false;
      }
      bt = best.t;
      bal = best.al;
      bbe = best.be;
      bpos = best.pos;
      if (hl && hr) {
        int nearc = (l0 <= r0) ? cl : cr, farc = (l0 <= r0) ? cr : cl;
        if (!stack_push<RING, THREADS, !ANY>(S, spill, farc)) {
          atomicAdd(overflow, 1u);
          bpos = -1;
          done = true;
        }
        node = nearc;
      } else if (hl) {
        node = cl;
      } else if (hr) {
        node = cr;
      } else if (stack_empty<!ANY>(S)) {
        done = true;
      } else {
        node = stack_pop<RING, THREADS, !ANY>(S, spill);
      }
    } else {
      // node >= 0: internal node to enter; node < 0: leaf whose box passed.
      if (node < 0) {
        double t, al, be;
        ++st.tests;
        if (tri_test(r, B.ltri + kLeafStride * (size_t)(~node), t, al, be)) {
          bt = t; bal = al; bbe = be; bpos = ~node;
          done = true;
        } else if (stack_empty<!ANY>(S)) {
          done = true;
        } else {
          node = stack_pop<RING, THREADS, !ANY>(S, spill);
        }
      } else {
        float4 q0, q1, q2, q3;
        load_node(B.tnode, node, q0, q1, q2, q3);
        st.visits += 2;
        double l0, l1, r0, r1;
        bool hl = slab(r, q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, l0, l1);
        bool hr = slab(r, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w, r0, r1);
        int cl = __float_as_int(q3.x), cr = __float_as_int(q3.y);
        if (hl) {
          if (hr && !stack_push<RING, THREADS, !ANY>(S, spill, cr)) {
            atomicAdd(overflow, 1u);
            bpos = -1;
            done = true;
          }
          node = cl;
        } else if (hr) {
          node = cr;
        } else if (stack_empty<!ANY>(S)) {
          done = true;
        } else {
          node = stack_pop<RING, THREADS, !ANY>(S, spill);
        }
      }
    }
    if (done) {
      write_hit<ANY, WFO>(B, O, rid, kid, bpos, bt, bal, bbe);
      active = false;
    }
  }
  flush_stats(stats, st);
}

// ------------------------------------------------------------ shading
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
  // Optional per-leaf shading table, 64 B per sorted leaf position:
  // {na.xyz, material id bits}, {nb.xyz, 0}, {nc.xyz, 0}, {0}. The same
  // floats as the mesh's, laid out so a hit reads one record instead of
  // gathering index, normals and material id through the triangle id.
  const float4* lshade;
};

__device__ __forceinline__ d3 arr3(const double* p) { return mk(p[0], p[1], p[2]); }

__device__ __forceinline__ const rt_material& mat_of(const rt_material* mats, int nmat, const MeshShade& M, int tri) {
  int id = M.mat_id ? M.mat_id[tri] : 0;
  if (id < 0 || id >= nmat) id = 0;
  return mats[id];
}

// Smooth normal (kernels/bvh.cu:357-364).
__device__ __forceinline__ d3 blend_normal(d3 na, d3 nb, d3 nc, double al, double be) {
  d3 N = (na * (1.0 - (al + be)) + nb * al) + nc * be;
  return rt::normalize(N);
}
__device__ __forceinline__ d3 shading_normal(const MeshShade& M, int tri, double al, double be) {
  int a = M.idx[3 * tri], b = M.idx[3 * tri + 1], c = M.idx[3 * tri + 2];
  d3 na = mk(M.nx[a], M.ny[a], M.nz[a]), nb = mk(M.nx[b], M.ny[b], M.nz[b]), nc = mk(M.nx[c], M.ny[c], M.nz[c]);
  return blend_normal(na, nb, nc, al, be);
}

// Material and smooth normal of the triangle at sorted leaf `pos`, from the
// shading table when the render has one.
struct LeafShade {
  const rt_material* m;
  d3 na, nb, nc;
};
__device__ __forceinline__ const rt_material& clamp_material(const rt_material* mats, int nmat, int id) {
  if (id < 0 || id >= nmat) id = 0;
  return mats[id];
}
__device__ __forceinline__ LeafShade leaf_shade(const rt_material* mats, int nmat, const MeshShade& M,
                                                const BvhView& B, int pos) {
  LeafShade L;
  if (M.lshade) {
    const float4* r = M.lshade + 4 * (size_t)pos;
    float4 a = ldg4(r), b = ldg4(r + 1), c = ldg4(r + 2);
    L.m = &clamp_material(mats, nmat, __float_as_int(a.w));
    L.na = mk(a.x, a.y, a.z);
    L.nb = mk(b.x, b.y, b.z);
    L.nc = mk(c.x, c.y, c.z);
  } else {
    const int tri = leaf_tri_of(B, pos);
    L.m = &mat_of(mats, nmat, M, tri);
    int a = M.idx[3 * tri], b = M.idx[3 * tri + 1], c = M.idx[3 * tri + 2];
    L.na = mk(M.nx[a], M.ny[a], M.nz[a]);
    L.nb = mk(M.nx[b], M.ny[b], M.nz[b]);
    L.nc = mk(M.nx[c], M.ny[c], M.nz[c]);
  }
  return L;
}
__device__ __forceinline__ const rt_material& leaf_material(const rt_material* mats, int nmat, const MeshShade& M,
                                                            const BvhView& B, int pos) {
  if (M.lshade) return clamp_material(mats, nmat, __float_as_int(ldgf((const float*)(M.lshade + 4 * (size_t)pos) + 3)));
  return mat_of(mats, nmat, M, leaf_tri_of(B, pos));
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

// Thin-lens primary ray of the next sample (kernels/raytracer.cu:142-154).
__device__ __forceinline__ void lens_ray(const RenderParams& P, double xx, double yy, uint32_t& rng, d3& o, d3& d) {
  double ang = 2 * M_PI * rt::rand_unit(rng);
  double rad = P.aperture * sqrt(rt::rand_unit(rng)) / 2.0;
  double sa, ca;
  rt::sincos_cw(ang, &sa, &ca);
  o = mk(rad * ca, rad * sa, 0);
  d = rt::normalize(mk(xx, yy, -1.0));
  d = d * P.focal - o;
  d = rt::normalize(d);
}

// Everything Material::shade computes before the shadow query
// (kernels/material.cu:6-40, Light::getDirection / distanceAttenuation,
// kernels/Light.cu:56-81): the shadow ray (q, L), the emissive base
// I1 = ke + ka % 0, the falloff fo and the unshadowed (diffuse + specular) term.
struct ShadePrep {
  d3 q, L, I1, ds, N;
  double fo;
};

// Light sample point: three draws in the order of mrand's rand_vec
// (kernels/mrand.cu:33-35), offset by radius*r - 0.5 (kernels/Light.cu:71-81).
__device__ __forceinline__ d3 light_point(const RenderParams& P, uint32_t& rng) {
  double r1, r2, r3;
  if (!(P.flags & RT_RENDER_RANDVEC_RTL)) {
    r1 = rt::rand_unit(rng); r2 = rt::rand_unit(rng); r3 = rt::rand_unit(rng);
  } else {
    r3 = rt::rand_unit(rng); r2 = rt::rand_unit(rng); r1 = rt::rand_unit(rng);
  }
  d3 off = mk(r1, r2, r3) * P.lrad;
  off = mk(off.x - 0.5, off.y - 0.5, off.z - 0.5);
  return arr3(P.lpos) + off;
}

// Shadow-ray direction, distance falloff fo = min(1, 1/(c0 + c1 d + c2 d^2))
// to the light centre (kernels/Light.cu:56-68) and max(0, N . L).
struct LightTerms {
  d3 L;
  double fo, nl;
};
__device__ __forceinline__ LightTerms light_terms(const RenderParams& P, d3 N, d3 q, d3 lpoint) {
  LightTerms T;
  T.L = rt::normalize(lpoint - q);
  double dl = rt::length_f(arr3(P.lpos) - q);
  double fo = P.c0 + (P.c1 + P.c2 * dl) * dl;
  fo = (fabs(fo) < kRayEps) ? 1.0 : 1.0 / fo;
  T.fo = fmin(1.0, fo);
  double nl = rt::dot(N, T.L);
  T.nl = nl < 0 ? 0.0 : nl;
  return T;
}

__device__ __forceinline__ ShadePrep shade_prep(const RenderParams& P, const rt_material& m, d3 N, d3 o, d3 d,
                                                double t, uint32_t& rng) {
  ShadePrep S;
  const d3 eye = arr3(P.eye);
  d3 q = o + d * t;
  S.q = q;
  S.N = N;
  S.I1 = arr3(m.ke) + rt::hadamard(arr3(m.ka), mk(0, 0, 0));
  d3 V = eye - q;
  V = mk(-V.x, -V.y, -V.z);
  V = rt::normalize(V);
  LightTerms T = light_terms(P, N, q, light_point(P, rng));
  const d3 L = T.L;
  S.L = L;
  S.fo = T.fo;
  d3 diffuse = arr3(m.kd) * T.nl;
  d3 spec = mk(0, 0, 0);
  if (!(m.ks[0] == 0 && m.ks[1] == 0 && m.ks[2] == 0)) {
    d3 R = N * (-2.0 * rt::dot(L, N)) + L;
    R = rt::normalize(R);
    double rv = rt::dot(R, V);
    double sf = (double)powf((float)fmax(0.0, rv), (float)m.shininess);
    spec = arr3(m.ks) * sf;
  }
  S.ds = diffuse + spec;
  return S;
}

// Light::shadowAttenuation from the any-hit result (kernels/Light.cu:5-53).
__device__ __forceinline__ d3 shadow_factor(const RenderParams& P, const BvhView& B, const rt_material* mats,
                                            const MeshShade& M, d3 q, bool hit, const HitRec& s) {
  if (!hit) return mk(1, 1, 1);
  if (!(s.t < rt::length_f(arr3(P.lpos) - q))) return mk(1, 1, 1);
  if (s.t < kRayEps) return mk(1, 1, 1);
  const rt_material& sm = leaf_material(mats, P.nmat, M, B, s.pos);
  bool trans = (sm.kt[0] == 0) && (sm.kt[1] == 0) && (sm.kt[2] == 0);
  return trans ? rt::hadamard(arr3(sm.kd), mk(1, 1, 1) - arr3(sm.kt)) : mk(0, 0, 0);
}

// I = I1 + ((sh * fo) % (diffuse + specular)) % light colour.
__device__ __forceinline__ d3 shade_finish(const RenderParams& P, d3 I1, double fo, d3 ds, d3 sh) {
  return I1 + rt::hadamard(rt::hadamard(sh * fo, ds), arr3(P.lcol));
}

// True when the shadow query cannot change shade_finish's result: in every
// channel the unshadowed term ds (max(0, N.L) kd + specular ks) or the light
// colour is zero, so (sh * fo) % ds % colour is zero for every shadow factor
// sh (each is finite: 1, 0 or kd % (1 - kt); fo is at most 1). The image is
// then I1 whatever the query returns, and the query is not made.
__device__ __forceinline__ bool shadow_irrelevant(const RenderParams& P, d3 ds) {
  return (ds.x == 0 || P.lcol[0] == 0) && (ds.y == 0 || P.lcol[1] == 0) && (ds.z == 0 || P.lcol[2] == 0);
}

__device__ __forceinline__ void pixel_of_tile(const RenderParams& P, int tile, int lane, int& px, int& py) {
  const int tiles_x = (P.W + 15) / 16;
  px = (tile % tiles_x) * 16 + (lane & 15);
  py = (tile / tiles_x) * 16 + (lane >> 4);
}

__device__ __forceinline__ void pixel_xy(const RenderParams& P, int px, int py, double& xx, double& yy) {
  xx = (2 * ((px + 0.5) * P.invW) - 1) * P.angle * P.aspect;
  yy = (1 - 2 * ((py + 0.5) * P.invH)) * P.angle;
}

__device__ __forceinline__ void write_pixel(const RenderParams& P, float* rgb, int px, int py, d3 color) {
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

__device__ __forceinline__ void flush_counters(unsigned long long* counters, unsigned long long prim,
                                               unsigned long long shadow, unsigned long long bounce,
                                               const TravStats& st) {
  if (!counters) return;
  unsigned long long vis = st.visits, tst = st.tests;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    prim += __shfl_xor(prim, off);
    shadow += __shfl_xor(shadow, off);
    bounce += __shfl_xor(bounce, off);
    vis += __shfl_xor(vis, off);
    tst += __shfl_xor(tst, off);
  }
  if ((threadIdx.x & 63) == 0) {
    // one of 64 counter replicas, 16 words apart (summed after the launch)
    unsigned long long* r = counters + (blockIdx.x & 63) * 16;
    if (prim) atomicAdd(&r[0], prim);
    if (shadow) atomicAdd(&r[1], shadow);
    if (bounce) atomicAdd(&r[2], bounce);
    if (vis) atomicAdd(&r[3], vis);
    if (tst) atomicAdd(&r[4], tst);
  }
}

// ------------------------------------------------------------ megakernel
// One thread per pixel, all samples, bounces and shadow rays in one launch
// (kernels/raytracer.cu:94-166 + traceRay :168-249).
__global__ __launch_bounds__(kThreads) void k_render(BvhView B, MeshShade M, const rt_material* __restrict__ mats,
                                                     RenderParams P, float* __restrict__ rgb,
                                                     unsigned long long* counters, uint32_t* overflow) {
  __shared__ int s_stack[kRing * kThreads];
  int* col = s_stack + threadIdx.x;
  int px, py;
  pixel_of_tile(P, blockIdx.x, threadIdx.x, px, py);
  const bool active = px < P.W && py < P.H;
  unsigned long long n_shadow = 0, n_bounce = 0;
  TravStats st;
  if (active) {
    uint32_t rng = rt::wang_hash(rt::seed_index(px, py, P.W) ^ P.seed);
    double xx, yy;
    pixel_xy(P, px, py, xx, yy);
    d3 color = mk(0, 0, 0);
    for (int it = 0; it < P.spp; ++it) {
      d3 o, d;
      lens_ray(P, xx, yy, rng, o, d);
      d3 acc = mk(0, 0, 0), thr = mk(1, 1, 1);
      for (int bounce = 0;; ++bounce) {
        RayD ray = make_ray(o, d);
        HitRec h;
        closest_hit(B, ray, bounce == 0 ? 0.0 : kBounceTmin, h, overflow, st, col);
        if (h.pos < 0) {
          acc = acc + rt::hadamard(thr, mk(0.9, 0.9, 0.9));
          break;
        }
        const int tri = leaf_tri(B, h.pos);
        const rt_material& m = mat_of(mats, P.nmat, M, tri);
        ShadePrep S = shade_prep(P, m, shading_normal(M, tri, h.al, h.be), o, d, h.t, rng);
        ++n_shadow;  // the reference's shadow ray, traced only if it can change the image
        d3 sh = mk(1, 1, 1);
        if (!shadow_irrelevant(P, S.ds)) {
          HitRec s;
          const bool sh_hit = any_hit(B, make_ray(S.q, S.L), s, overflow, st, col);
          sh = shadow_factor(P, B, mats, M, S.q, sh_hit, s);
        }
        d3 I = shade_finish(P, S.I1, S.fo, S.ds, sh);
        acc = acc + rt::hadamard(thr, I);
        if (bounce >= P.depth) break;
        double u1 = rt::rand_unit(rng);
        double u2 = rt::rand_unit(rng);
        d3 N = S.N;
        if (rt::dot(N, d) > 0) N = mk(-N.x, -N.y, -N.z);
        thr = rt::hadamard(thr, arr3(m.kd));
        o = S.q;
        d = cosine_dir(N, u1, u2);
        ++n_bounce;
      }
      color = color + acc;
    }
    write_pixel(P, rgb, px, py, color);
  }
  flush_counters(counters, active ? (unsigned long long)P.spp : 0ull, n_shadow, n_bounce, st);
}

// ------------------------------------------------------------ wavefront
// The same per-pixel computation split into passes over a batch of scenes
// (one scene for rt_render): per sample a camera pass, then per bounce a
// persistent closest-hit traversal, a shading pass (which draws the light
// sample and, if the path continues, the bounce direction, and appends the
// shadow ray and the next ray to queues), a persistent any-hit traversal of
// the shadow queue and a pass that adds the shaded result. Ray slots carry a
// scene id; the scene table gives each scene's BVH, mesh, materials and
// camera. Per-pixel RNG state and partial sums stay in HBM between passes;
// each pixel's draws and additions happen in the megakernel's order, so
// both give the same image bit for bit.
struct SceneDesc {
  BvhView B;
  MeshShade M;
  const rt_material* mats;
  RenderParams P;
  float* rgb;
  uint32_t pix_off;   // first pixel of the scene in the per-pixel arrays
  uint32_t tile_off;  // first 16x16 tile of the scene
  uint32_t tri_off;   // triangles of the scenes before this one (batches)
};

struct WfState {
  const SceneDesc* scenes;
  const BvhView* views;    // views[s] = scenes[s].B (traversal reads only these)
  const uint16_t* tile_scene;  // scene of each tile
  uint32_t ntiles;
  uint32_t* rng;        // pixels
  double* color;        // pixels*3: sum over finished samples
  double* acc;          // pixels*3: a sample's path sum (one buffer per sample slot)
  double* thr;          // pixels*3: current path throughput
  double* ray[2];       // cap*6: ray origin, direction per queue slot
  uint32_t* pix[2];     // cap: global pixel of each queue slot (~0 = unused slot)
  uint16_t* rscene[2];  // cap: scene of each queue slot
  HitRec32* hrec;       // cap: closest-hit record of each queue slot
  // shadow queue of a sample slot: the last bounce's any-hit pass of
  // sample k may still run while the next samples are traced
  double* shq;          // cap*16: q, L, I1, ds, thr, fo
  uint32_t* spix;       // cap: global pixel of each shadow entry
  uint16_t* sscene;     // cap: scene of each shadow entry
  HitRec16* srec;       // cap: any-hit records in fetch order (rank in key order when sorted)
  uint32_t* srank;      // cap: rank of each shadow entry in key order (when sorted)
  uint32_t* scount;     // number of shadow entries of the sample slot (counts[2 + slot])
  uint16_t* skey;       // cap: sort key of each shadow ray: (scene, direction cell)
  uint32_t* sperm;      // cap: shadow rays in key order
  uint32_t* shist;      // sbins: key counts (zero between sorts)
  uint32_t* sbase;      // sbins: key offsets
  uint32_t sbins;       // number of keys: kShadowCells for one scene, nscenes << scell_bits for a batch
  int scell_bits;       // bits of the cell index in a batch key
  uint32_t* counts;     // [0],[1] ray queue sizes, [2 + slot] shadow entries per sample slot, [8] stack overflows
  uint32_t* fetch[2];   // kFetchWords each: closest-hit and any-hit fetch counters (the latter per slot)
  int2* rec;            // optional path record: (hit leaf, occluder code) per pixel, sample, bounce
  int rec_spp, rec_depth;
  double* color64;      // optional FP64 output (pixels*3, row-flipped like rgb, unclamped mean)
  size_t npix, cap;     // component strides: per-pixel arrays are SoA over npix, queues SoA over cap
  unsigned long long* rep;  // kCtrReplicas x kCtrStride ray counters, summed into the caller's at the end
};

// Ray and work counts go to one of kCtrReplicas copies of the counters (by
// workgroup), each on its own 128-byte line: all waves of a shading pass
// finish together, and one shared address would serialise their atomics.
// Words of a replica: 0 primary, 1 shadow, 2 bounce rays; 3, 4 box and
// triangle tests of the closest-hit traversal; 5, 6 those of the any-hit one.
constexpr int kCtrReplicas = 64;
constexpr int kCtrStride = 16;

// Shadow rays are traced in the order of their direction from the light
// (single-scene renders): lines through nearby cells share BVH nodes, so
// the any-hit pass finds more of them in cache. A cell is one of
// kShadowGrid x kShadowGrid on the square that the scene box projects onto,
// seen from the light, numbered in Morton order. 32 x 32 measured fastest on
// a 1M-triangle soup at 1024 x 1024: 16 x 16 sorts too coarsely, 64 x 64
// costs more in the sort than the any-hit pass gains.
#ifndef RT_SHADOW_SORT
#define RT_SHADOW_SORT 1
#endif
constexpr bool kShadowSort = RT_SHADOW_SORT != 0;
#ifndef RT_SHADOW_GRID
#define RT_SHADOW_GRID 32
#endif
constexpr int kShadowGrid = RT_SHADOW_GRID;  // cells per side, a power of two
constexpr int kShadowBits = __builtin_ctz(kShadowGrid);
constexpr int kShadowCells = kShadowGrid * kShadowGrid;
constexpr int kScanThreads = 1024;
static_assert((kShadowGrid & (kShadowGrid - 1)) == 0 && kShadowCells >= 64, "shadow grid");
// A batch sorts by (scene, cell) in a 16-bit key: the scene index in the
// high bits, the cell's Morton index below it. Past 64 scenes the cell
// index loses low bit pairs, i.e. the grid halves per side each time.
inline int batch_cell_bits(size_t nscenes) {
  int sb = 0;
  while (sb < 16 && (size_t(1) << sb) < nscenes) ++sb;
  const int cb = (16 - sb) & ~1;
  return cb < 2 * kShadowBits ? cb : 2 * kShadowBits;
}
inline uint32_t sort_bins(size_t nscenes) { return (uint32_t)(nscenes << batch_cell_bits(nscenes)); }
struct ShadowCells {
  float lx, ly, lz;           // light centre
  float wx, wy, wz;           // unit direction from the light to the box centre
  float ex, ey, ez, fx, fy, fz;  // two unit vectors across it
  float a0, as, b0, bs;       // cell = ((a - a0) * as, (b - b0) * bs) of the projection (a, b)
  bool ok;                    // false: a box corner is not in front of the light (every key 0)
};
__device__ __forceinline__ ShadowCells shadow_cells(const RenderParams& P, const float* root_box) {
  ShadowCells C;
  C.lx = (float)P.lpos[0];
  C.ly = (float)P.lpos[1];
  C.lz = (float)P.lpos[2];
  float cx = 0.5f * (root_box[0] + root_box[3]) - C.lx, cy = 0.5f * (root_box[1] + root_box[4]) - C.ly,
        cz = 0.5f * (root_box[2] + root_box[5]) - C.lz;
  float len = sqrtf(cx * cx + cy * cy + cz * cz);
  C.ok = len > 0.f;
  if (!C.ok || !(len < INFINITY)) {
    C.ok = false;
    return C;
  }
  C.wx = cx / len;
  C.wy = cy / len;
  C.wz = cz / len;
  // e = normalize(up x w), f = w x e, with up the axis least aligned with w
  float ux = 0.f, uy = 0.f, uz = 1.f;
  if (fabsf(C.wz) > 0.9f) {
    ux = 1.f;
    uz = 0.f;
  }
  float ex = uy * C.wz - uz * C.wy, ey = uz * C.wx - ux * C.wz, ez = ux * C.wy - uy * C.wx;
  float el = sqrtf(ex * ex + ey * ey + ez * ez);
  C.ex = ex / el;
  C.ey = ey / el;
  C.ez = ez / el;
  C.fx = C.wy * C.ez - C.wz * C.ey;
  C.fy = C.wz * C.ex - C.wx * C.ez;
  C.fz = C.wx * C.ey - C.wy * C.ex;
  float amin = INFINITY, amax = -INFINITY, bmin = INFINITY, bmax = -INFINITY;
  for (int k = 0; k < 8; ++k) {
    const float vx = root_box[(k & 1) ? 3 : 0] - C.lx, vy = root_box[(k & 2) ? 4 : 1] - C.ly,
                vz = root_box[(k & 4) ? 5 : 2] - C.lz;
    const float z = vx * C.wx + vy * C.wy + vz * C.wz;
    if (!(z > 1e-3f * len)) C.ok = false;
    const float a = (vx * C.ex + vy * C.ey + vz * C.ez) / z, b = (vx * C.fx + vy * C.fy + vz * C.fz) / z;
    amin = fminf(amin, a);
    amax = fmaxf(amax, a);
    bmin = fminf(bmin, b);
    bmax = fmaxf(bmax, b);
  }
  C.a0 = amin;
  C.b0 = bmin;
  C.as = amax > amin ? (float)kShadowGrid / (amax - amin) : 0.f;
  C.bs = bmax > bmin ? (float)kShadowGrid / (bmax - bmin) : 0.f;
  return C;
}
__device__ __forceinline__ uint16_t shadow_key(const ShadowCells& C, d3 q) {
  if (!C.ok) return 0;
  const float vx = (float)q.x - C.lx, vy = (float)q.y - C.ly, vz = (float)q.z - C.lz;
  const float z = vx * C.wx + vy * C.wy + vz * C.wz;
  if (!(z > 0.f)) return 0;
  const float a = (vx * C.ex + vy * C.ey + vz * C.ez) / z, b = (vx * C.fx + vy * C.fy + vz * C.fz) / z;
  const float top = (float)(kShadowGrid - 1);
  const uint32_t ia = (uint32_t)fminf(top, fmaxf(0.f, (a - C.a0) * C.as)),
                 ib = (uint32_t)fminf(top, fmaxf(0.f, (b - C.b0) * C.bs));
  uint32_t k = 0;
#pragma unroll
  for (int bit = 0; bit < kShadowBits; ++bit)
    k |= ((ia >> bit) & 1u) << (2 * bit) | ((ib >> bit) & 1u) << (2 * bit + 1);
  return (uint16_t)k;
}
__device__ __forceinline__ void rep_count(unsigned long long* rep, int k, bool pred) {
  const uint64_t m = __ballot(pred);
  if ((threadIdx.x & 63) == 0 && m)
    atomicAdd(&rep[(blockIdx.x & (kCtrReplicas - 1)) * kCtrStride + k], (unsigned long long)__popcll(m));
}

// Path record index of (global pixel, sample, bounce).
__device__ __forceinline__ size_t rec_index(const WfState& W, uint32_t p, int sample, int bounce) {
  return ((size_t)p * W.rec_spp + sample) * (W.rec_depth + 1) + bounce;
}

// Component-major (SoA) vectors: x at a[i], y at a[n + i], z at a[2n + i].
__device__ __forceinline__ d3 ldv(const double* a, size_t i, size_t n) { return mk(a[i], a[n + i], a[2 * n + i]); }
__device__ __forceinline__ void stv(double* a, size_t i, size_t n, d3 v) {
  a[i] = v.x;
  a[n + i] = v.y;
  a[2 * n + i] = v.z;
}
__device__ __forceinline__ void st3(double* p, d3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}

// Append to a queue with one atomic per workgroup; every thread must call it.
__device__ __forceinline__ uint32_t block_append(uint32_t* counter, bool want, uint32_t* s_tmp) {
  const uint32_t lane = __lane_id();
  const int w = threadIdx.x >> 6;
  uint64_t m = __ballot(want);
  if (lane == 0) s_tmp[w] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int j = 0; j < kThreads / 64; ++j) tot += s_tmp[j];
    s_tmp[kThreads / 64] = tot ? atomicAdd(counter, tot) : 0u;
  }
  __syncthreads();
  uint32_t base = s_tmp[kThreads / 64];
  for (int j = 0; j < w; ++j) base += s_tmp[j];
  uint64_t below = lane == 0 ? 0ull : (m & (~0ull >> (64 - lane)));
  uint32_t r = base + (uint32_t)__popcll(below);
  __syncthreads();
  return r;
}

// The scene table entry of scene s. A single-scene render passes its one
// entry as a kernel argument (MULTI = false), so every parameter is a
// uniform scalar load instead of a per-lane gather from the table.
template <bool MULTI>
__device__ __forceinline__ const SceneDesc& scene_of(const WfState& W, const SceneDesc& D1, int s) {
  if constexpr (MULTI)
    return W.scenes[s];
  else
    return D1;
}

// Global tile -> (scene, pixel); false outside the scene's image.
template <bool MULTI>
__device__ __forceinline__ bool tile_pixel(const WfState& W, const SceneDesc& D1, uint32_t gt, int lane, int& s,
                                           int& px, int& py, uint32_t& gp) {
  s = MULTI ? W.tile_scene[gt] : 0;
  const SceneDesc& D = scene_of<MULTI>(W, D1, s);
  pixel_of_tile(D.P, (int)(gt - D.tile_off), lane, px, py);
  if (px >= D.P.W || py >= D.P.H) return false;
  gp = D.pix_off + (uint32_t)(py * D.P.W + px);
  return true;
}

template <bool MULTI>
__global__ __launch_bounds__(kThreads) void k_wf_init(WfState W, SceneDesc D1) {
  if (blockIdx.x == 0) {
    // the queue counts; counts[8], a batch's stack-overflow count, keeps
    // accumulating over the batch's lifetime
    if (threadIdx.x < 16 && threadIdx.x != 8) W.counts[threadIdx.x] = 0u;
    for (int k = threadIdx.x; k < kCtrReplicas * kCtrStride; k += kThreads) W.rep[k] = 0ull;
    for (uint32_t k = threadIdx.x; k < W.sbins; k += kThreads) W.shist[k] = 0u;
  }
  int s, px, py;
  uint32_t gp;
  if (!tile_pixel<MULTI>(W, D1, blockIdx.x, threadIdx.x, s, px, py, gp)) return;
  const RenderParams& P = scene_of<MULTI>(W, D1, s).P;
  W.rng[gp] = rt::wang_hash(rt::seed_index(px, py, P.W) ^ P.seed);
}

// The image sum starts at zero. It is first added to by the shading pass of
// sample kSampleSlots (or the final flush), so this pass runs on the
// auxiliary stream, off the caller's path to the first closest-hit pass.
template <bool MULTI>
__global__ __launch_bounds__(kThreads) void k_wf_zero_color(WfState W, SceneDesc D1) {
  int s, px, py;
  uint32_t gp;
  if (!tile_pixel<MULTI>(W, D1, blockIdx.x, threadIdx.x, s, px, py, gp)) return;
  stv(W.color, gp, W.npix, mk(0, 0, 0));
}

// Start sample k: slot = tile*256 + lane holds the pixel's primary ray.
// Slots outside an image, or of a scene that has taken all its samples,
// get a NaN ray, which misses the root box. The sample's accumulator is
// set by its first shading pass.
template <bool MULTI>
__global__ __launch_bounds__(kThreads) void k_wf_camera(WfState W, SceneDesc D1, int k, uint32_t nslots) {
  const size_t slot = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) W.counts[0] = nslots;
  int s, px, py;
  uint32_t gp;
  bool ok = tile_pixel<MULTI>(W, D1, blockIdx.x, threadIdx.x, s, px, py, gp);
  const RenderParams& P = scene_of<MULTI>(W, D1, s).P;
  double* rr = W.ray[0] + slot;
  W.rscene[0][slot] = (uint16_t)s;
  if (!ok || k >= P.spp) {
    const double nan = __builtin_nan("");
    for (int j = 0; j < 6; ++j) rr[j * W.cap] = nan;
    W.pix[0][slot] = 0xffffffffu;
    return;
  }
  uint32_t rng = W.rng[gp];
  double xx, yy;
  pixel_xy(P, px, py, xx, yy);
  d3 o, d;
  lens_ray(P, xx, yy, rng, o, d);
  W.rng[gp] = rng;
  stv(rr, 0, W.cap, o);
  stv(rr + 3 * W.cap, 0, W.cap, d);
  W.pix[0][slot] = gp;
}

// Shading table of each scene (blockIdx.y), in triangle order so the
// index, normal and material reads are as coherent as the mesh; each
// triangle's record goes to its sorted leaf position.
__device__ __forceinline__ void shade_fill_one(const SceneDesc& D, int t) {
  const MeshShade& M = D.M;
  float4* out = const_cast<float4*>(M.lshade);
  if (!out) return;
  const int a = M.idx[3 * t], b = M.idx[3 * t + 1], c = M.idx[3 * t + 2];
  const int mat = M.mat_id ? M.mat_id[t] : 0;
  float4* r = out + 4 * (size_t)D.B.leaf_of[t];
  r[0] = make_float4(M.nx[a], M.ny[a], M.nz[a], __int_as_float(mat));
  r[1] = make_float4(M.nx[b], M.ny[b], M.nz[b], 0.f);
  r[2] = make_float4(M.nx[c], M.ny[c], M.nz[c], 0.f);
  r[3] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// A batch runs one thread per triangle of all its scenes in turn (the
// scene of triangle g is the last s with tri_off <= g), so every scene gets
// threads in proportion to its size.
template <bool MULTI>
__global__ __launch_bounds__(256) void k_shade_fill(const SceneDesc* scenes, int nscenes, SceneDesc D1) {
  if constexpr (MULTI) {
    const SceneDesc& last = scenes[nscenes - 1];
    const uint32_t total = last.tri_off + (uint32_t)last.B.n;
    for (uint32_t g = blockIdx.x * 256 + threadIdx.x; g < total; g += gridDim.x * 256) {
      int lo = 0, hi = nscenes - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (scenes[mid].tri_off <= g)
          lo = mid;
        else
          hi = mid - 1;
      }
      shade_fill_one(scenes[lo], (int)(g - scenes[lo].tri_off));
    }
  } else {
    for (int t = blockIdx.x * 256 + threadIdx.x; t < D1.B.n; t += gridDim.x * 256) shade_fill_one(D1, t);
  }
}

// One bounce's shading. The first bounce of sample k also takes over the
// slot's accumulator: with flush >= 0 the path sums still there (sample
// k - slots, whose shading has finished) go into the image first; then
// the accumulator starts from this pass's own terms.
template <bool MULTI>
__global__ __launch_bounds__(kThreads) void k_wf_shade(WfState W, SceneDesc D1, int bounce, int cur, int sample,
                                                       unsigned long long* counters, int flush) {
  __shared__ uint32_t s_tmp[8];
  const uint32_t n_in = W.counts[cur];
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  uint32_t p = i < n_in ? W.pix[cur][i] : 0xffffffffu;
  const bool live = p != 0xffffffffu;
  bool want_shadow = false, want_next = false, dark = false;
  ShadePrep S;
  d3 thr = mk(0, 0, 0), no = mk(0, 0, 0), nd = mk(0, 0, 0);
  uint16_t sc = 0, skey = 0;
  ShadowCells cells;
  cells.ok = false;
  if (!MULTI && kShadowSort && D1.B.n > 0) cells = shadow_cells(D1.P, D1.B.node_box);
  if (live) {
    sc = MULTI ? W.rscene[cur][i] : 0;
    const SceneDesc& D = scene_of<MULTI>(W, D1, sc);
    // a primary ray's throughput is 1 (the camera pass writes none)
    thr = bounce == 0 ? mk(1, 1, 1) : ldv(W.thr, p, W.npix);
    const HitRec32 hr = W.hrec[i];
    const int pos = hr.pos;
    if (W.rec) W.rec[rec_index(W, p, sample, bounce)] = make_int2(pos, -1);
    if (bounce == 0) {
      if (flush >= 0) stv(W.color, p, W.npix, ldv(W.color, p, W.npix) + ldv(W.acc, p, W.npix));
      stv(W.acc, p, W.npix, mk(0, 0, 0));
    }
    if (pos < 0) {
      stv(W.acc, p, W.npix, ldv(W.acc, p, W.npix) + rt::hadamard(thr, mk(0.9, 0.9, 0.9)));
    } else {
      d3 o = ldv(W.ray[cur], i, W.cap), d = ldv(W.ray[cur] + 3 * W.cap, i, W.cap);
      const double hh[3] = {hr.t, hr.al, hr.be};
      uint32_t rng = W.rng[p];
      const LeafShade LS = leaf_shade(D.mats, D.P.nmat, D.M, D.B, pos);
      const rt_material& m = *LS.m;
      S = shade_prep(D.P, m, blend_normal(LS.na, LS.nb, LS.nc, hh[1], hh[2]), o, d, hh[0], rng);
      // The shadow ray is queued only if its answer can change the image
      // (always for a path record, which names the occluder). Otherwise the
      // shading result is added here, exactly as the finish pass would.
      want_shadow = W.rec || !shadow_irrelevant(D.P, S.ds);
      dark = !want_shadow;
      if (dark)
        stv(W.acc, p, W.npix,
            ldv(W.acc, p, W.npix) + rt::hadamard(thr, shade_finish(D.P, S.I1, S.fo, S.ds, mk(1, 1, 1))));
      if (MULTI && kShadowSort)
        skey = (uint16_t)(((uint32_t)sc << W.scell_bits) |
                          (shadow_key(shadow_cells(D.P, D.B.node_box), S.q) >> (2 * kShadowBits - W.scell_bits)));
      if (bounce < D.P.depth) {
        double u1 = rt::rand_unit(rng);
        double u2 = rt::rand_unit(rng);
        d3 N = S.N;
        if (rt::dot(N, d) > 0) N = mk(-N.x, -N.y, -N.z);
        stv(W.thr, p, W.npix, rt::hadamard(thr, arr3(m.kd)));
        no = S.q;
        nd = cosine_dir(N, u1, u2);
        want_next = true;
      }
      W.rng[p] = rng;
    }
  }
  uint32_t ss = block_append(W.scount, want_shadow, s_tmp);
  if (want_shadow) {
    W.spix[ss] = p;
    W.sscene[ss] = sc;
    if (kShadowSort) W.skey[ss] = MULTI ? skey : shadow_key(cells, S.q);
    // shadow entry, component-major: q, L, I1, ds, thr, fo
    double* e = W.shq + ss;
    const size_t c = W.cap;
    stv(e, 0, c, S.q);
    stv(e + 3 * c, 0, c, S.L);
    stv(e + 6 * c, 0, c, S.I1);
    stv(e + 9 * c, 0, c, S.ds);
    if (bounce > 0) stv(e + 12 * c, 0, c, thr);  // first bounce: 1, not stored
    e[15 * c] = S.fo;
  }
  uint32_t ns = block_append(&W.counts[cur ^ 1], want_next, s_tmp);
  if (want_next) {
    W.pix[cur ^ 1][ns] = p;
    W.rscene[cur ^ 1][ns] = sc;
    stv(W.ray[cur ^ 1], ns, W.cap, no);
    stv(W.ray[cur ^ 1] + 3 * W.cap, ns, W.cap, nd);
  }
  // rays traced by the closest-hit pass of this bounce; the reference's
  // shadow rays not traced (the finish pass counts the traced ones), and the
  // traced ones for the any-hit kernel's statistics
  if (counters) {
    rep_count(W.rep, bounce == 0 ? 0 : 2, live);
    rep_count(W.rep, 1, dark);
    rep_count(W.rep, 7, want_shadow);
  }
}

// Counting sort of a sample's shadow queue by cell key, in three passes:
// per-cell counts, their exclusive scan, and the scatter of slot indices.
// The order within a cell depends on atomic timing; the any-hit pass writes
// every result to the ray's own slot, so the image does not depend on it.
// A workgroup sorts kSortPer * kThreads slots (kSortPer per thread, strided),
// so that few workgroups add their counts to the global ones.
constexpr int kSortPer = 16;
constexpr int kSortSpan = kSortPer * kThreads;

__global__ __launch_bounds__(kThreads) void k_sort_hist(WfState W) {
  __shared__ uint32_t h[kShadowCells];
  const uint32_t n = *W.scount, i0 = blockIdx.x * kSortSpan + threadIdx.x;
  if (blockIdx.x * kSortSpan >= n) return;
  for (int k = threadIdx.x; k < kShadowCells; k += kThreads) h[k] = 0u;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortPer; ++j) {
    const uint32_t i = i0 + j * kThreads;
    if (i < n) atomicAdd(&h[W.skey[i]], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kShadowCells; k += kThreads)
    if (h[k]) atomicAdd(&W.shist[k], h[k]);
}

// One workgroup; thread k scans cells [k*per, (k+1)*per).
__global__ __launch_bounds__(kScanThreads) void k_sort_scan(WfState W) {
  constexpr int per = kShadowCells >= kScanThreads ? kShadowCells / kScanThreads : 1;
  __shared__ uint32_t part[kScanThreads / 64];
  const int k = threadIdx.x, lane = k & 63, wv = k >> 6;
  const int c0 = k * per;
  uint32_t v[per], sum = 0;
#pragma unroll
  for (int j = 0; j < per; ++j) {
    v[j] = c0 + j < kShadowCells ? W.shist[c0 + j] : 0u;
    sum += v[j];
  }
  uint32_t x = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) part[wv] = x;
  __syncthreads();
  uint32_t run = x - sum;
  for (int j = 0; j < wv; ++j) run += part[j];
#pragma unroll
  for (int j = 0; j < per; ++j) {
    if (c0 + j < kShadowCells) {
      W.sbase[c0 + j] = run;
      W.shist[c0 + j] = 0u;  // ready for the next sort of this slot
    }
    run += v[j];
  }
}

__global__ __launch_bounds__(kThreads) void k_sort_scatter(WfState W) {
  __shared__ uint32_t h[kShadowCells];
  const uint32_t n = *W.scount, i0 = blockIdx.x * kSortSpan + threadIdx.x;
  if (blockIdx.x * kSortSpan >= n) return;
  for (int k = threadIdx.x; k < kShadowCells; k += kThreads) h[k] = 0u;
  __syncthreads();
  uint32_t key[kSortPer], rank[kSortPer];
#pragma unroll
  for (int j = 0; j < kSortPer; ++j) {
    const uint32_t i = i0 + j * kThreads;
    key[j] = i < n ? W.skey[i] : 0u;
    rank[j] = i < n ? atomicAdd(&h[key[j]], 1u) : 0u;
  }
  __syncthreads();
  // one global reservation per non-empty cell of the workgroup
  for (int k = threadIdx.x; k < kShadowCells; k += kThreads)
    if (h[k]) h[k] = atomicAdd(&W.sbase[k], h[k]);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortPer; ++j) {
    const uint32_t i = i0 + j * kThreads;
    if (i < n) {
      const uint32_t r = h[key[j]] + rank[j];
      W.sperm[r] = i;
      W.srank[i] = r;
    }
  }
}

// The batch form of the three passes, for up to 65536 keys: global counts,
// with one atomic per run of equal keys among a wave's lanes (neighbouring
// queue entries mostly share a key).
struct KeyRun {
  bool head;      // first lane of its run (and lane valid)
  uint32_t first; // first lane of this lane's run
  uint32_t len;   // run length (heads)
};
__device__ __forceinline__ KeyRun key_run(uint32_t key, bool valid) {
  const uint32_t lane = __lane_id();
  const uint32_t pk = __shfl_up(key, 1, 64);
  const bool pv = __shfl_up((int)valid, 1, 64) != 0;
  const bool start = lane == 0 || !valid || !pv || pk != key;
  const uint64_t heads = __ballot(start);
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  KeyRun r;
  r.first = 63u - (uint32_t)__clzll(heads & upto);
  const uint64_t above = heads & ~upto;
  r.len = (above ? (uint32_t)__ffsll((long long)above) - 1u : 64u) - lane;
  r.head = start && valid;
  return r;
}

__global__ __launch_bounds__(kThreads) void k_msort_hist(WfState W) {
  const uint32_t n = *W.scount;
  for (uint32_t b = blockIdx.x * kThreads; b < n; b += gridDim.x * kThreads) {
    const uint32_t i = b + threadIdx.x;
    const uint32_t key = i < n ? W.skey[i] : 0u;
    const KeyRun r = key_run(key, i < n);
    if (r.head) atomicAdd(&W.shist[key], r.len);
  }
}

// One workgroup; thread k scans keys [k*per, (k+1)*per).
__global__ __launch_bounds__(kScanThreads) void k_msort_scan(WfState W) {
  __shared__ uint32_t part[kScanThreads / 64];
  const uint32_t nb = W.sbins, per = (nb + kScanThreads - 1) / kScanThreads;
  const int k = threadIdx.x, lane = k & 63, wv = k >> 6;
  const uint32_t c0 = min(nb, k * per), c1 = min(nb, c0 + per);
  uint32_t sum = 0;
  for (uint32_t c = c0; c < c1; ++c) sum += W.shist[c];
  uint32_t x = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) part[wv] = x;
  __syncthreads();
  uint32_t run = x - sum;
  for (int j = 0; j < wv; ++j) run += part[j];
  for (uint32_t c = c0; c < c1; ++c) {
    const uint32_t v = W.shist[c];
    W.sbase[c] = run;
    W.shist[c] = 0u;  // ready for the next sort of this slot
    run += v;
  }
}

__global__ __launch_bounds__(kThreads) void k_msort_scatter(WfState W) {
  const uint32_t n = *W.scount;
  for (uint32_t b = blockIdx.x * kThreads; b < n; b += gridDim.x * kThreads) {
    const uint32_t i = b + threadIdx.x;
    const uint32_t key = i < n ? W.skey[i] : 0u;
    const KeyRun r = key_run(key, i < n);
    uint32_t base = 0;
    if (r.head) base = atomicAdd(&W.sbase[key], r.len);
    base = __shfl(base, (int)r.first, 64);
    if (i < n) {
      const uint32_t rk = base + (__lane_id() - r.first);
      W.sperm[rk] = i;
      W.srank[i] = rk;
    }
  }
}

// Occluder code of a shadow query: -1 = unshadowed (factor 1), -2 = opaque
// occluder (factor 0), else the leaf position of the transmitting occluder
// (factor kd % (1 - kt), kernels/Light.cu:17-26).
__device__ __forceinline__ int occluder_code(const SceneDesc& D, d3 q, const HitRec& s) {
  if (s.pos < 0 || !(s.t < rt::length_f(arr3(D.P.lpos) - q)) || s.t < kRayEps) return -1;
  const rt_material& sm = leaf_material(D.mats, D.P.nmat, D.M, D.B, s.pos);
  bool trans = (sm.kt[0] == 0) && (sm.kt[1] == 0) && (sm.kt[2] == 0);
  return trans ? s.pos : -2;
}

template <bool MULTI>
__global__ __launch_bounds__(kThreads) void k_wf_shadow_finish(WfState W, SceneDesc D1, int sample, int bounce,
                                                               unsigned long long* counters) {
  const uint32_t n_in = *W.scount;
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  const bool live = i < n_in;
  if (counters) rep_count(W.rep, 1, live);
  if (!live) return;
  const SceneDesc& D = scene_of<MULTI>(W, D1, MULTI ? W.sscene[i] : 0);
  const double* e = W.shq + i;
  const size_t c = W.cap;
  d3 q = ldv(e, 0, c);
  const HitRec16 r = W.srec[kShadowSort ? W.srank[i] : i];
  HitRec s;
  s.pos = r.pos;
  s.t = r.t;
  d3 I = shade_finish(D.P, ldv(e + 6 * c, 0, c), e[15 * c], ldv(e + 9 * c, 0, c),
                      shadow_factor(D.P, D.B, D.mats, D.M, q, s.pos >= 0, s));
  uint32_t p = W.spix[i];
  const d3 thr = bounce == 0 ? mk(1, 1, 1) : ldv(e + 12 * c, 0, c);
  stv(W.acc, p, W.npix, ldv(W.acc, p, W.npix) + rt::hadamard(thr, I));
  if (W.rec) W.rec[rec_index(W, p, sample, bounce)].y = occluder_code(D, q, s);
}

// Zero n 64-bit words (graph-safe replacement for a memset).
__global__ void k_zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Sum the replicas: counters[0..2] += the ray counts, counters[3..4] += the
// work counts of both traversals; per_kernel[0..3] (if given) += the work
// counts of the closest-hit and the any-hit traversal apart, per_kernel[4..5]
// += the rays each traced.
__global__ void k_wf_counters_out(const unsigned long long* rep, unsigned long long* counters,
                                  unsigned long long* per_kernel) {
  __shared__ unsigned long long s_sum[8];
  const int k = threadIdx.x;
  if (k < 8) {
    unsigned long long sum = 0;
    for (int r = 0; r < kCtrReplicas; ++r) sum += rep[r * kCtrStride + k];
    s_sum[k] = sum;
  }
  __syncthreads();
  if (k == 0 && counters) {
    for (int j = 0; j < 3; ++j)
      if (s_sum[j]) atomicAdd(&counters[j], s_sum[j]);
    if (s_sum[3] + s_sum[5]) atomicAdd(&counters[3], s_sum[3] + s_sum[5]);
    if (s_sum[4] + s_sum[6]) atomicAdd(&counters[4], s_sum[4] + s_sum[6]);
  }
  if (k == 1 && per_kernel) {
    for (int j = 0; j < 4; ++j) per_kernel[j] += s_sum[3 + j];
    per_kernel[4] += s_sum[0] + s_sum[2];
    per_kernel[5] += s_sum[7];
  }
}

// Reset the listed counters and, if fetch != nullptr, `words` fetch
// counters (graph-safe replacement for a memset).
__global__ void k_wf_reset(uint32_t* counts, uint32_t mask, uint32_t* fetch, uint32_t words) {
  if (counts && threadIdx.x < 32 && (mask >> threadIdx.x) & 1u) counts[threadIdx.x] = 0u;
  if (fetch)
    for (uint32_t k = threadIdx.x; k < words; k += blockDim.x) fetch[k] = 0u;
}

// The accumulators of the sample slots after the last sample.
struct AccSlots {
  double* acc[kMaxSampleSlots];
};

// End of the render: each slot's accumulator still holds the path sums of
// the pixel's last sample on that slot (samples spp - slots .. spp - 1 of
// its scene; earlier ones were added by the shading pass that reused the
// slot), added here in sample order.
template <bool MULTI>
__global__ __launch_bounds__(kThreads) void k_wf_flush_last(WfState W, SceneDesc D1, AccSlots A, int slots) {
  int s, px, py;
  uint32_t gp;
  if (!tile_pixel<MULTI>(W, D1, blockIdx.x, threadIdx.x, s, px, py, gp)) return;
  const int spp = scene_of<MULTI>(W, D1, s).P.spp;
  d3 c = ldv(W.color, gp, W.npix);
  for (int k = spp > slots ? spp - slots : 0; k < spp; ++k) c = c + ldv(A.acc[k % slots], gp, W.npix);
  stv(W.color, gp, W.npix, c);
}

template <bool MULTI>
__global__ __launch_bounds__(kThreads) void k_wf_resolve(WfState W, SceneDesc D1) {
  int s, px, py;
  uint32_t gp;
  if (!tile_pixel<MULTI>(W, D1, blockIdx.x, threadIdx.x, s, px, py, gp)) return;
  const SceneDesc& D = scene_of<MULTI>(W, D1, s);
  d3 c = ldv(W.color, gp, W.npix);
  if (D.rgb) write_pixel(D.P, D.rgb, px, py, c);
  if (W.color64) {
    size_t oi = 3 * ((size_t)(D.P.H - py - 1) * D.P.W + px);
    st3(W.color64 + oi, mk(c.x / double(D.P.spp), c.y / double(D.P.spp), c.z / double(D.P.spp)));
  }
}

// Copy a scene table that came in as a kernel argument to device memory
// (keeps single-scene renders free of host copies, so they can be captured).
__global__ void k_wf_put_scene(SceneDesc* dst, BvhView* views, uint16_t* tile_scene, SceneDesc d, uint32_t ntiles) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    dst[0] = d;
    views[0] = d.B;
  }
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x) tile_scene[t] = 0;
}

__global__ void k_put_view(BvhView* dst, BvhView v) {
  if (threadIdx.x == 0) *dst = v;
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

static RenderParams rt_make_params(const rt_camera* cam, const rt_light* light, const rt_render_opts* opts, int32_t nmat) {
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
  return P;
}

rt_tuning rt_default_tuning() {
  rt_tuning t;
#ifndef RT_DEF_RING
#define RT_DEF_RING 32
#endif
#ifndef RT_DEF_CHUNK
#define RT_DEF_CHUNK 16
#endif
#ifndef RT_DEF_WAVES
#define RT_DEF_WAVES 4
#endif
  t.stack_ring = RT_DEF_RING;
  t.fetch_chunk = RT_DEF_CHUNK;
  t.waves_per_simd = RT_DEF_WAVES;
#ifndef RT_DEFAULT_LAYOUT
#define RT_DEFAULT_LAYOUT 5
#endif
  t.node_layout = RT_DEFAULT_LAYOUT;
  t.wg_threads = 64;
  return t;
}

bool rt_tuning_valid(const rt_tuning* t) {
  return t && (t->stack_ring == 16 || t->stack_ring == 32) &&
         (t->fetch_chunk == 16 || t->fetch_chunk == 32 || t->fetch_chunk == 64 || t->fetch_chunk == 128 ||
          t->fetch_chunk == 256 || t->fetch_chunk == 512) &&
         t->waves_per_simd >= 1 && t->waves_per_simd <= 8 &&
         (t->node_layout == 2 || t->node_layout == 4 || t->node_layout == 5 || t->node_layout == 8) &&
         (t->wg_threads == 0 || t->wg_threads == 64 || t->wg_threads == 256);
}

static uint32_t floor_log2(int64_t v) {
  uint32_t k = 0;
  while (v > 1 && k < 255) {
    v >>= 1;
    ++k;
  }
  return k;
}

uint32_t rt_tuning_key(int64_t ntris, int64_t nrays) { return floor_log2(ntris) << 8 | floor_log2(nrays); }

rt_tuning rt_ctx_tuning_for(const rt_ctx* ctx, int64_t ntris, int64_t nrays) {
  if (ctx && !ctx->tuning.empty()) {
    auto it = ctx->tuning.find(rt_tuning_key(ntris, nrays));
    if (it != ctx->tuning.end()) return it->second;
  }
  return rt_default_tuning();
}

static int trace_threads(const rt_tuning& t) { return t.wg_threads == 256 ? 256 : 64; }

// waves_per_simd waves on each of a CU's 4 SIMDs, trace_threads(t) per workgroup
static int persistent_grid(int num_cus, const rt_tuning& t, uint32_t max_rays) {
  const int threads = trace_threads(t);
  int cap = num_cus * t.waves_per_simd * (4 * 64 / threads);
  int need = (int)((max_rays + threads - 1) / threads);
  return need < cap ? (need > 0 ? need : 1) : cap;
}

template <bool ANY, bool WFO>
static void launch_persistent(const rt_tuning& t, int grid, hipStream_t s, const BvhView* views,
                              const uint16_t* scene, RayIn R, double tmin, HitOut O, uint32_t* fetch,
                              uint32_t* overflow, unsigned long long* stats) {
  const uint32_t fc = t.fetch_chunk;
  // the default layout has closest-hit kernels without the lower bound
  // for launches with tmin 0
#define RT_LAUNCH_T(RINGV, LAYOUTV, TH)                                                                           \
  do {                                                                                                              \
    constexpr bool kNoLow = !ANY && LAYOUTV == 5;                                                                   \
    if (kNoLow && !(tmin > 0)) {                                                                                    \
      if (scene)                                                                                                    \
        k_trace_persistent<ANY, RINGV, LAYOUTV, true, !kNoLow, TH, WFO>                                                  \
            <<<grid, TH, 0, s>>>(views, scene, R, tmin, O, fetch, fc, overflow, stats);                            \
      else                                                                                                          \
        k_trace_persistent<ANY, RINGV, LAYOUTV, false, !kNoLow, TH, WFO>                                                 \
            <<<grid, TH, 0, s>>>(views, scene, R, tmin, O, fetch, fc, overflow, stats);                            \
    } else if (scene) {                                                                                             \
      k_trace_persistent<ANY, RINGV, LAYOUTV, true, true, TH, WFO>                                                       \
          <<<grid, TH, 0, s>>>(views, scene, R, tmin, O, fetch, fc, overflow, stats);                              \
    } else {                                                                                                        \
      k_trace_persistent<ANY, RINGV, LAYOUTV, false, true, TH, WFO>                                                      \
          <<<grid, TH, 0, s>>>(views, scene, R, tmin, O, fetch, fc, overflow, stats);                              \
    }                                                                                                               \
  } while (0)
#define RT_LAUNCH(RINGV, LAYOUTV)                   \
  do {                                              \
    if (trace_threads(t) == 256)                    \
      RT_LAUNCH_T(RINGV, LAYOUTV, 256);             \
    else                                            \
      RT_LAUNCH_T(RINGV, LAYOUTV, 64);              \
  } while (0)
  if (t.node_layout == 8) {
    if (t.stack_ring == 16) RT_LAUNCH(16, 8); else RT_LAUNCH(32, 8);
  } else if (t.node_layout == 5) {
    if (t.stack_ring == 16) RT_LAUNCH(16, 5); else RT_LAUNCH(32, 5);
  } else if (t.node_layout == 4) {
    if (t.stack_ring == 16) RT_LAUNCH(16, 4); else RT_LAUNCH(32, 4);
  } else {
    if (t.stack_ring == 16) RT_LAUNCH(16, 2); else RT_LAUNCH(32, 2);
  }
#undef RT_LAUNCH_T
#undef RT_LAUNCH
}

hipError_t rt_launch_trace(const rt_bvh* b, const rt_mesh_view* m, const rt_rays* rays, int32_t mode, double tmin,
                           const rt_hits* hits, hipStream_t s) {
  (void)m;
  if (rays->n <= 0) return hipSuccess;
  uint32_t* fetch = b->fetch;
  // fetch[kFetchWords] holds the BvhView the kernel reads
  BvhView* view = (BvhView*)(fetch + 512);
  k_wf_reset<<<1, 256, 0, s>>>(nullptr, 0u, fetch, kFetchWords);
  k_put_view<<<1, 1, 0, s>>>(view, b->view());
  RayIn R{rays->orig_dev, rays->dir_dev, 3, 1, nullptr, (uint32_t)rays->n, nullptr};
  HitOut O{nullptr, hits->tri_dev, hits->t_dev, hits->alpha_dev, hits->beta_dev, nullptr, nullptr};
  const rt_tuning t = rt_ctx_tuning_for(b->ctx, b->n, rays->n);
  hipError_t e = rt_ensure_layout(b, t.node_layout, s);
  if (e != hipSuccess) return e;
  int grid = persistent_grid(b->ctx ? b->ctx->num_cus : 256, t, (uint32_t)rays->n);
  if (mode == RT_TRACE_ANY)
    launch_persistent<true, false>(t, grid, s, view, nullptr, R, 0.0, O, fetch, b->misc + kMiscOverflow, nullptr);
  else
    launch_persistent<false, false>(t, grid, s, view, nullptr, R, tmin, O, fetch, b->misc + kMiscOverflow, nullptr);
  return hipGetLastError();
}

// Workspace layout of a wavefront render over `nscenes` scenes with `npix`
// pixels and `ntiles` tiles in total: regions every sample shares, then one
// block per sample slot. Sample k uses slot k % kSampleSlots; the last
// bounce's shadow work of sample k (on the auxiliary stream) reads and
// writes only its slot's block (and the atomic counter replicas), so sample
// k + 1 .. k + kSampleSlots - 1 may run on the caller's stream meanwhile.
// Each region is 256-byte aligned and disjoint from every other.
#ifndef RT_SAMPLE_SLOTS
#define RT_SAMPLE_SLOTS 2
#endif
constexpr int kSampleSlots = RT_SAMPLE_SLOTS;
// scount = counts + 2 + slot stays below counts[8], the batch overflow count
static_assert(kSampleSlots >= 2 && kSampleSlots <= kMaxSampleSlots && 2 + kMaxSampleSlots <= 8, "sample slots");
// The last bounce's any-hit passes of samples 2j and 2j + 1 (slots 0 and 1)
// run as one launch: a launch ends with a drain in which few lanes still
// have rays, and one launch over both samples' shadow rays pays it once.
#ifndef RT_PAIR_SHADOWS
#define RT_PAIR_SHADOWS 1
#endif
constexpr bool kPairShadows = RT_PAIR_SHADOWS != 0 && kSampleSlots == 2;

enum WfShared {
  kRScenes, kRViews, kRTileScene, kRRng, kRColor, kRThr, kRRay0, kRRay1, kRPix0, kRPix1, kRRscene0, kRRscene1,
  kRHrec, kRCounts, kRFetch0, kRRep, kRShared
};
enum WfSlot { kRAcc, kRShq, kRSpix, kRSscene, kRSrec, kRSrank, kRFetch1, kRSkey, kRSperm, kRShist, kRSbase, kRSlot };

struct WfLayout {
  size_t shared[kRShared], shared_bytes[kRShared];
  size_t slot[kMaxSampleSlots][kRSlot], slot_bytes[kRSlot];
  size_t total;
};

static WfLayout wf_layout(size_t npix, size_t ntiles, size_t nscenes, int nslots = kSampleSlots) {
  WfLayout L{};
  const size_t cap = ntiles * kThreads;
  const size_t bins = sort_bins(nscenes);
  const size_t sh[kRShared] = {nscenes * sizeof(SceneDesc), nscenes * sizeof(BvhView), ntiles * 2, npix * 4,
                               npix * 24, npix * 24, cap * 48, cap * 48, cap * 4, cap * 4, cap * 2, cap * 2,
                               cap * sizeof(HitRec32), 256, kFetchWords * 4, kCtrReplicas * kCtrStride * 8};
  const size_t sl[kRSlot] = {npix * 24, cap * 128, cap * 4, cap * 2, cap * sizeof(HitRec16), cap * 4,
                             kFetchWords * 4, cap * 2, cap * 4, bins * 4, bins * 4};
  size_t off = 0;
  for (int k = 0; k < kRShared; ++k) {
    L.shared[k] = off;
    L.shared_bytes[k] = sh[k];
    off += align256(sh[k]);
  }
  for (int k = 0; k < kRSlot; ++k) L.slot_bytes[k] = sl[k];
  for (int j = 0; j < nslots && j < kMaxSampleSlots; ++j)
    for (int k = 0; k < kRSlot; ++k) {
      L.slot[j][k] = off;
      off += align256(sl[k]);
    }
  L.total = off;
  return L;
}

// The state as the passes of a sample in `slot` see it: the slot's
// accumulator, shadow queue, shadow count, any-hit fetch counters and
// shadow-sort buffers; everything else is shared.
static WfState wf_state(void* ws, size_t npix, size_t ntiles, size_t nscenes, int slot = 0,
                        int nslots = kSampleSlots) {
  const WfLayout L = wf_layout(npix, ntiles, nscenes, nslots);
  char* w = (char*)ws;
  WfState W;
  W.scenes = (const SceneDesc*)(w + L.shared[kRScenes]);
  W.views = (const BvhView*)(w + L.shared[kRViews]);
  W.tile_scene = (const uint16_t*)(w + L.shared[kRTileScene]);
  W.ntiles = (uint32_t)ntiles;
  W.rng = (uint32_t*)(w + L.shared[kRRng]);
  W.color = (double*)(w + L.shared[kRColor]);
  W.thr = (double*)(w + L.shared[kRThr]);
  W.ray[0] = (double*)(w + L.shared[kRRay0]);
  W.ray[1] = (double*)(w + L.shared[kRRay1]);
  W.pix[0] = (uint32_t*)(w + L.shared[kRPix0]);
  W.pix[1] = (uint32_t*)(w + L.shared[kRPix1]);
  W.rscene[0] = (uint16_t*)(w + L.shared[kRRscene0]);
  W.rscene[1] = (uint16_t*)(w + L.shared[kRRscene1]);
  W.hrec = (HitRec32*)(w + L.shared[kRHrec]);
  W.counts = (uint32_t*)(w + L.shared[kRCounts]);
  W.fetch[0] = (uint32_t*)(w + L.shared[kRFetch0]);
  W.rep = (unsigned long long*)(w + L.shared[kRRep]);
  const size_t* S = L.slot[slot];
  W.acc = (double*)(w + S[kRAcc]);
  W.shq = (double*)(w + S[kRShq]);
  W.spix = (uint32_t*)(w + S[kRSpix]);
  W.sscene = (uint16_t*)(w + S[kRSscene]);
  W.srec = (HitRec16*)(w + S[kRSrec]);
  W.srank = (uint32_t*)(w + S[kRSrank]);
  W.fetch[1] = (uint32_t*)(w + S[kRFetch1]);
  W.skey = (uint16_t*)(w + S[kRSkey]);
  W.sperm = (uint32_t*)(w + S[kRSperm]);
  W.shist = (uint32_t*)(w + S[kRShist]);
  W.sbase = (uint32_t*)(w + S[kRSbase]);
  W.scount = W.counts + 2 + slot;
  W.sbins = sort_bins(nscenes);
  W.scell_bits = batch_cell_bits(nscenes);
  W.npix = npix;
  W.cap = ntiles * kThreads;
  return W;
}

// The regions a sample slot's state addresses, for the layout checks of the
// tests: each pointer of wf_state(slot) as an offset into the workspace,
// with the size of the buffer it points to.
extern "C" int32_t rt_wavefront_state_regions(uint64_t npix, uint64_t ntiles, uint64_t nscenes, int32_t nslots,
                                              int32_t slot, rt_wf_region* out, int32_t max, uint64_t* total) {
  if (nslots < 2 || nslots > kMaxSampleSlots || slot < 0 || slot >= nslots || nscenes < 1 || ntiles < 1) return -1;
  const WfLayout L = wf_layout(npix, ntiles, nscenes, nslots);
  if (total) *total = L.total;
  // a base address far from 0, so every offset is a plain difference
  char* base = (char*)(uintptr_t)(1ull << 40);
  const WfState W = wf_state(base, npix, ntiles, nscenes, slot, nslots);
  struct F {
    const char* name;
    const void* p;
    size_t bytes;
    int32_t shared;
  };
  const size_t* B = L.shared_bytes;
  const size_t* Q = L.slot_bytes;
  const F f[] = {
      {"scenes", W.scenes, B[kRScenes], 1},    {"views", W.views, B[kRViews], 1},
      {"tile_scene", W.tile_scene, B[kRTileScene], 1}, {"rng", W.rng, B[kRRng], 1},
      {"color", W.color, B[kRColor], 1},       {"thr", W.thr, B[kRThr], 1},
      {"ray0", W.ray[0], B[kRRay0], 1},        {"ray1", W.ray[1], B[kRRay1], 1},
      {"pix0", W.pix[0], B[kRPix0], 1},        {"pix1", W.pix[1], B[kRPix1], 1},
      {"rscene0", W.rscene[0], B[kRRscene0], 1}, {"rscene1", W.rscene[1], B[kRRscene1], 1},
      {"hrec", W.hrec, B[kRHrec], 1},
      {"counts", W.counts, B[kRCounts], 1},    {"fetch0", W.fetch[0], B[kRFetch0], 1},
      {"rep", W.rep, B[kRRep], 1},             {"acc", W.acc, Q[kRAcc], 0},
      {"shq", W.shq, Q[kRShq], 0},             {"spix", W.spix, Q[kRSpix], 0},
      {"sscene", W.sscene, Q[kRSscene], 0},    {"srec", W.srec, Q[kRSrec], 0},
      {"srank", W.srank, Q[kRSrank], 0},       {"fetch1", W.fetch[1], Q[kRFetch1], 0},
      {"skey", W.skey, Q[kRSkey], 0},          {"sperm", W.sperm, Q[kRSperm], 0},
      {"shist", W.shist, Q[kRShist], 0},       {"sbase", W.sbase, Q[kRSbase], 0},
      {"scount", W.scount, sizeof(uint32_t), 0},
  };
  const int32_t n = (int32_t)(sizeof(f) / sizeof(f[0]));
  for (int32_t k = 0; k < n && k < max && out; ++k)
    out[k] = rt_wf_region{f[k].name, (uint64_t)((const char*)f[k].p - base), (uint64_t)f[k].bytes, f[k].shared};
  return n;
}

uint32_t* rt_wavefront_counts(void* workspace, size_t npix, size_t ntiles, size_t nscenes) {
  return wf_state(workspace, npix, ntiles, nscenes).counts;
}

size_t rt_wavefront_bytes_batch(size_t npix, size_t ntiles, size_t nscenes) {
  return wf_layout(npix, ntiles, nscenes).total;
}

size_t rt_wavefront_bytes(int32_t width, int32_t height) {
  size_t ntiles = (size_t)cdiv(width, 16) * cdiv(height, 16);
  return rt_wavefront_bytes_batch((size_t)width * height, ntiles, 1);
}

// Fill the scene table of a batch (host side; the caller copies it to the device).
void rt_fill_scene(void* table_entry, const rt_bvh* b, const rt_mesh_view* m, const rt_material* mats, int32_t nmat,
                   const rt_light* light, const rt_camera* cam, const rt_render_opts* opts, float* rgb,
                   uint32_t pix_off, uint32_t tile_off, uint32_t tri_off, void* shade_table) {
  SceneDesc d;
  d.B = b->view();
  d.M = MeshShade{m->nx_dev, m->ny_dev, m->nz_dev, m->idx_dev, m->mat_id_dev, (const float4*)shade_table};
  d.mats = mats;
  d.P = rt_make_params(cam, light, opts, nmat);
  d.rgb = rgb;
  d.pix_off = pix_off;
  d.tile_off = tile_off;
  d.tri_off = tri_off;
  *(SceneDesc*)table_entry = d;
}
size_t rt_scene_desc_bytes() { return sizeof(SceneDesc); }
size_t rt_bvh_view_bytes() { return sizeof(BvhView); }

// Events of one timed traversal launch (recycled by rt_ctx_kernel_stats).
static hipError_t timer_mark(rt_kernel_timer* T, int kind, bool start, hipStream_t s) {
  if (start) {
    hipEvent_t ev[2];
    for (hipEvent_t& e : ev) {
      if (T->pool.empty()) {
        hipError_t err = hipEventCreate(&e);
        if (err != hipSuccess) return err;
      } else {
        e = T->pool.back();
        T->pool.pop_back();
      }
    }
    T->recs.push_back({kind, ev[0], ev[1]});
    return hipEventRecord(ev[0], s);
  }
  return hipEventRecord(T->recs.back().b, s);
}

// The wavefront passes over a batch whose scene table, views and
// tile->scene map are already in the workspace.
hipError_t rt_launch_wavefront(void* workspace, size_t npix, size_t ntiles, size_t nscenes, int max_spp,
                               int max_depth, int num_cus, const rt_tuning* tuning, uint64_t* counters,
                               uint32_t* overflow, hipStream_t s, int32_t* records, double* color64,
                               const void* single_scene, rt_kernel_timer* timer, const rt_overlap* ov,
                               const uint32_t* live_tiles) {
  // Wp[k % kSampleSlots] is the state sample k's passes see (its slot's
  // accumulator, shadow queue and shadow-sort buffers).
  WfState Wp[kSampleSlots];
  for (int j = 0; j < kSampleSlots; ++j) Wp[j] = wf_state(workspace, npix, ntiles, nscenes, j);
  for (WfState& W : Wp) {
    W.rec = (int2*)records;
    W.rec_spp = max_spp;
    W.rec_depth = max_depth;
    W.color64 = color64;
  }
  const WfState& W0 = Wp[0];
  // A paired any-hit launch addresses slot 1's shadow buffers as slot 0's
  // moved by one offset, and its records by fetch index + offset / 16.
  const size_t pair_bytes = (size_t)((const char*)Wp[1].shq - (const char*)Wp[0].shq);
  const auto moved = [&](const void* a, const void* b) { return (size_t)((const char*)b - (const char*)a); };
  const bool pair = kPairShadows && pair_bytes % 256 == 0 && moved(Wp[0].sperm, Wp[1].sperm) == pair_bytes &&
                    moved(Wp[0].srec, Wp[1].srec) == pair_bytes && moved(Wp[0].sscene, Wp[1].sscene) == pair_bytes &&
                    pair_bytes / sizeof(HitRec16) + ntiles * kThreads < (1ull << 31);
  rt_kernel_timer* T = timer && timer->enabled ? timer : nullptr;
  const bool multi = single_scene == nullptr;
  SceneDesc D1{};
  if (!multi) D1 = *(const SceneDesc*)single_scene;
#define RT_WFG(STREAM, GRID, STATE, KERNEL, ...)                                 \
  do {                                                                           \
    if (multi)                                                                   \
      KERNEL<true><<<(GRID), kThreads, 0, STREAM>>>(STATE, D1, __VA_ARGS__);     \
    else                                                                         \
      KERNEL<false><<<(GRID), kThreads, 0, STREAM>>>(STATE, D1, __VA_ARGS__);    \
  } while (0)
#define RT_WF(STREAM, STATE, KERNEL, ...) RT_WFG(STREAM, tiles, STATE, KERNEL, __VA_ARGS__)
#define RT_TRY(X)                          \
  do {                                     \
    hipError_t e_ = (X);                   \
    if (e_ != hipSuccess) return e_;       \
  } while (0)
  rt_tuning t = tuning ? *tuning : rt_default_tuning();
  unsigned long long* cnt = (unsigned long long*)counters;
  const size_t cap = ntiles * kThreads;
  const int tiles = (int)ntiles;
  const bool work = cnt != nullptr;  // count work only when the caller asked for counters
  // The last bounce's shadow sort, any-hit pass and shadow resolve of
  // sample k run on the auxiliary stream, overlapping the passes of the
  // next samples on s: they share no buffer but the replicated counters
  // (atomics). Sample k + kSampleSlots reuses sample k's slot, so its first
  // shading pass (which adds sample k's path sums to the colour sum) waits
  // for sample k's shadow resolve. With paired any-hit launches, samples
  // 2j and 2j + 1's shadow rays are traced together after sample 2j + 1's
  // shading, beside sample 2j + 2's camera and closest-hit passes.
  // The shading table is first read by sample 0's shading pass, so it is
  // filled on the auxiliary stream while sample 0's camera and closest-hit
  // passes run.
  hipStream_t sa = ov ? ov->aux : s;
  if (ov) {
    RT_TRY(hipEventRecord(ov->start, s));
    RT_TRY(hipStreamWaitEvent(sa, ov->start, 0));
  }
  if (multi) {
    k_shade_fill<true><<<4096, 256, 0, sa>>>(W0.scenes, (int)nscenes, D1);
  } else if (D1.M.lshade) {
    k_shade_fill<false><<<std::min(cdiv(D1.B.n, 256), 4096), 256, 0, sa>>>(W0.scenes, 1, D1);
  }
  if (ov) RT_TRY(hipEventRecord(ov->filled, sa));
  if (multi)
    k_wf_zero_color<true><<<tiles, kThreads, 0, sa>>>(W0, D1);
  else
    k_wf_zero_color<false><<<tiles, kThreads, 0, sa>>>(W0, D1);
  const int grid = persistent_grid(num_cus, t, (uint32_t)cap);
  uint32_t* ovf = overflow ? overflow : W0.counts + 8;
  if (multi)
    k_wf_init<true><<<tiles, kThreads, 0, s>>>(W0, D1);
  else
    k_wf_init<false><<<tiles, kThreads, 0, s>>>(W0, D1);
  // every slot starts with zero fetch counters and an empty shadow-sort
  // histogram (k_wf_init clears slot 0's; each sort clears its own after use)
  k_wf_reset<<<1, 256, 0, s>>>(nullptr, 0u, W0.fetch[0], kFetchWords);
  for (int j = 0; j < kSampleSlots; ++j) {
    k_wf_reset<<<1, 256, 0, s>>>(nullptr, 0u, Wp[j].fetch[1], kFetchWords);
    if (j > 0) k_wf_reset<<<1, 256, 0, s>>>(nullptr, 0u, Wp[j].shist, Wp[j].sbins);
  }
  for (int it = 0; it < max_spp; ++it) {
    const int par = it % kSampleSlots;
    const WfState& W = Wp[par];
    // tiles still sampling: with live_tiles (a batch whose scenes' tiles are
    // ordered by sample count, most first) a prefix of the tiles
    const int live = live_tiles ? (int)live_tiles[it] : tiles;
    RT_WFG(s, live, W, k_wf_camera, it, (uint32_t)live * kThreads);
    int cur = 0;
    for (int bounce = 0; bounce <= max_depth; ++bounce) {
      RayIn R{W.ray[cur], W.ray[cur] + 3 * cap, 1, cap, &W.counts[cur], 0, nullptr};
      HitOut O{nullptr, nullptr, nullptr, nullptr, nullptr, W.hrec, nullptr};
      // one scene: every ray uses views[0] (no per-ray scene lookup)
      if (T) RT_TRY(timer_mark(T, 0, true, s));
      launch_persistent<false, true>(t, grid, s, W.views, nscenes > 1 ? W.rscene[cur] : nullptr, R,
                               bounce == 0 ? 0.0 : kBounceTmin, O, W.fetch[0], ovf, work ? W.rep + 3 : nullptr);
      if (T) RT_TRY(timer_mark(T, 0, false, s));
      if (ov && it == 0 && bounce == 0) RT_TRY(hipStreamWaitEvent(s, ov->filled, 0));
      if (bounce == 0 && it >= kSampleSlots) {
        // the slot's last sample has finished its shading on the auxiliary
        // stream: its shadow count and any-hit fetch counters start over,
        // and this pass adds its path sums to the image
        if (ov) RT_TRY(hipStreamWaitEvent(s, ov->flushed[par], 0));
        k_wf_reset<<<1, 256, 0, s>>>(W.counts, 1u << (2 + par), W.fetch[1], kFetchWords);
      }
      RT_WFG(s, bounce == 0 ? live : tiles, W, k_wf_shade, bounce, cur, it, cnt,
             bounce == 0 && it >= kSampleSlots ? it - kSampleSlots : -1);
      // the queue traced above is consumed: clear its count and the fetch counters
      k_wf_reset<<<1, 256, 0, s>>>(W.counts, 1u << cur, W.fetch[0], kFetchWords);
      // earlier bounces share the accumulator with the next bounce's shading
      const bool overlap = ov && bounce == max_depth;
      hipStream_t sb = overlap ? sa : s;
      if (overlap) {
        RT_TRY(hipEventRecord(ov->shaded, s));
        RT_TRY(hipStreamWaitEvent(sa, ov->shaded, 0));
      }
      // Paired samples: the last bounce's any-hit pass of sample 2j waits
      // for sample 2j + 1's and runs in the same launch (one drain instead
      // of two); both samples' shadow finishes follow it.
      const bool last = bounce == max_depth;
      const bool defer = last && pair && par == 0 && it + 1 < max_spp;
      const bool paired = last && pair && par == 1;
      RayIn SR{W.shq, W.shq + 3 * cap, 1, cap, W.scount, 0, kShadowSort ? W.sperm : nullptr};
      if (kShadowSort && !multi) {
        const int sort_blocks = cdiv((long)cap, kSortSpan);
        k_sort_hist<<<sort_blocks, kThreads, 0, sb>>>(W);
        k_sort_scan<<<1, kScanThreads, 0, sb>>>(W);
        k_sort_scatter<<<sort_blocks, kThreads, 0, sb>>>(W);
      } else if (kShadowSort) {
        const int sort_blocks = std::min(cdiv((long)cap, kThreads), 4096);
        k_msort_hist<<<sort_blocks, kThreads, 0, sb>>>(W);
        k_msort_scan<<<1, kScanThreads, 0, sb>>>(W);
        k_msort_scatter<<<sort_blocks, kThreads, 0, sb>>>(W);
      }
      if (defer) {
        cur ^= 1;
        continue;
      }
      // a paired launch reads both slots' queues from slot 0's buffers
      const WfState& WA = paired ? Wp[0] : W;
      if (paired) {
        SR = RayIn{WA.shq, WA.shq + 3 * cap, 1, cap, WA.scount, 0, kShadowSort ? WA.sperm : nullptr,
                   Wp[1].scount, pair_bytes};
      }
      HitOut SO{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, WA.srec};
      if (T) RT_TRY(timer_mark(T, 1, true, sb));
      launch_persistent<true, true>(t, grid, sb, WA.views, nscenes > 1 ? WA.sscene : nullptr, SR, 0.0, SO,
                                    WA.fetch[1], ovf, work ? WA.rep + 5 : nullptr);
      if (T) RT_TRY(timer_mark(T, 1, false, sb));
      if (paired) RT_WF(sb, Wp[0], k_wf_shadow_finish, it - 1, bounce, cnt);
      RT_WF(sb, W, k_wf_shadow_finish, it, bounce, cnt);
      // the last bounce's shadow count and fetch counters are cleared before
      // the next sample on this slot is shaded
      if (bounce < max_depth) k_wf_reset<<<1, 256, 0, sb>>>(W.counts, 1u << (2 + par), W.fetch[1], kFetchWords);
      cur ^= 1;
    }
    // sample it's path sums are complete: the first shading pass of sample
    // it + kSampleSlots adds them to the image (a deferred sample's are
    // complete with its pair's)
    if (ov && pair && par == 0 && it + 1 < max_spp) continue;
    if (ov && pair && par == 1) RT_TRY(hipEventRecord(ov->flushed[0], sa));
    if (ov) RT_TRY(hipEventRecord(ov->flushed[par], sa));
  }
  if (ov) {
    RT_TRY(hipEventRecord(ov->done, sa));
    RT_TRY(hipStreamWaitEvent(s, ov->done, 0));
  }
  // every pixel's last samples (no later shading pass took them), in
  // sample order
  {
    AccSlots A;
    for (int j = 0; j < kSampleSlots; ++j) A.acc[j] = Wp[j].acc;
    RT_WF(s, W0, k_wf_flush_last, A, kSampleSlots);
  }
  if (multi)
    k_wf_resolve<true><<<tiles, kThreads, 0, s>>>(W0, D1);
  else
    k_wf_resolve<false><<<tiles, kThreads, 0, s>>>(W0, D1);
  if (work) k_wf_counters_out<<<1, 64, 0, s>>>(W0.rep, cnt, T ? T->dev_work : nullptr);
#undef RT_WF
#undef RT_WFG
#undef RT_TRY
  return hipGetLastError();
}

hipError_t rt_launch_render(const rt_bvh* b, const rt_mesh_view* m, const rt_material* mats, int32_t nmat,
                            const rt_light* light, const rt_camera* cam, const rt_render_opts* opts, float* rgb,
                            uint64_t* counters, void* workspace, void* shade_table, hipStream_t s,
                            int32_t* records, double* color64, rt_kernel_timer* timer, const rt_overlap* ov) {
  RenderParams P = rt_make_params(cam, light, opts, nmat);
  MeshShade M{m->nx_dev, m->ny_dev, m->nz_dev, m->idx_dev, m->mat_id_dev, nullptr};
  const int tiles = cdiv(P.W, 16) * cdiv(P.H, 16);
  unsigned long long* cnt = (unsigned long long*)counters;
  uint32_t* ovf = b->misc + kMiscOverflow;
  if (!workspace || !shade_table) {
    // megakernel (binary records); with counters, the workspace holds the counter replicas
    hipError_t e = rt_ensure_layout(b, 2, s);
    if (e != hipSuccess) return e;
    unsigned long long* rep = cnt ? (unsigned long long*)workspace : nullptr;
    if (rep) k_zero_u64<<<cdiv(kCtrReplicas * kCtrStride, 256), 256, 0, s>>>(rep, kCtrReplicas * kCtrStride);
    k_render<<<tiles, kThreads, 0, s>>>(b->view(), M, mats, P, rgb, rep, ovf);
    if (rep) k_wf_counters_out<<<1, 64, 0, s>>>(rep, cnt, nullptr);
    return hipGetLastError();
  }
  SceneDesc d;
  d.B = b->view();
  d.M = M;
  d.M.lshade = (const float4*)shade_table;
  d.mats = mats;
  d.P = P;
  d.rgb = rgb;
  d.pix_off = 0;
  d.tile_off = 0;
  d.tri_off = 0;
  const rt_tuning t = rt_ctx_tuning_for(b->ctx, b->n, (int64_t)P.W * P.H);
  hipError_t e = rt_ensure_layout(b, t.node_layout, s);
  if (e != hipSuccess) return e;
  WfState W = wf_state(workspace, (size_t)P.W * P.H, (size_t)tiles, 1);
  k_wf_put_scene<<<cdiv(tiles, 256), 256, 0, s>>>((SceneDesc*)W.scenes, (BvhView*)W.views,
                                                  (uint16_t*)W.tile_scene, d, (uint32_t)tiles);
  return rt_launch_wavefront(workspace, (size_t)P.W * P.H, (size_t)tiles, 1, P.spp, P.depth,
                             b->ctx ? b->ctx->num_cus : 256, &t, counters, ovf, s, records, color64, &d, timer, ov);
}

// ------------------------------------------------------------ differentiable render
// Per-triangle albedo as a material table: material t = (kd = albedo[t],
// ka = ks = kt = ke = 0), so Trans() is true and an occluder attenuates
// shadows by its albedo (kernels/Light.cu:22-24).
namespace {

__global__ void k_fill_albedo(rt_material* mats, int32_t* ids, const double* albedo, int32_t n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  ids[t] = t;
  rt_material m;
  for (int k = 0; k < 3; ++k) {
    m.ka[k] = 0.0;
    m.kd[k] = albedo[3 * (size_t)t + k];
    m.ks[k] = 0.0;
    m.kt[k] = 0.0;
    m.ke[k] = 0.0;
  }
  m.shininess = 0.0;
  mats[t] = m;
}

__global__ void k_zero_f64(double* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

__global__ void k_fill_i32(int32_t* p, size_t n, int32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

constexpr int kMaxDiffBounces = 8;

struct DiffBounce {
  d3 o, d, lpoint;
  double t, al, be;
  double g;           // fo * max(0, N . L) as the forward rounds it
  int tri, occ_tri;   // occ_tri: -1 unshadowed, -2 opaque, else triangle
  bool next, flip;
  double xs, ys, zs;
};

__device__ __forceinline__ rta::v3 to_a(d3 v) { return rta::V(v.x, v.y, v.z); }
__device__ __forceinline__ d3 from_a(rta::v3 v) { return mk(v.x, v.y, v.z); }

__device__ __forceinline__ d3 alb(const double* a, int tri) {
  return mk(a[3 * (size_t)tri], a[3 * (size_t)tri + 1], a[3 * (size_t)tri + 2]);
}
__device__ __forceinline__ void atomic_add3(double* dst, d3 v) {
  atomicAdd(dst + 0, v.x);
  atomicAdd(dst + 1, v.y);
  atomicAdd(dst + 2, v.z);
}

// Adjoint of the render: one thread per pixel replays each sample's path
// from the record (same draws, no traversal) and propagates dLoss/dcolor
// back to per-triangle albedo, light intensity and vertex positions.
__global__ __launch_bounds__(kThreads) void k_adjoint(BvhView B, MeshShade M, const float* __restrict__ vx,
                                                      const float* __restrict__ vy, const float* __restrict__ vz,
                                                      RenderParams P, const double* __restrict__ albedo,
                                                      double light_scale, const int2* __restrict__ rec,
                                                      const double* __restrict__ dcolor, double* g_alb,
                                                      double* g_light, double* g_pos) {
  int px, py;
  pixel_of_tile(P, blockIdx.x, threadIdx.x, px, py);
  double light_acc = 0.0;
  if (px < P.W && py < P.H) {
    const uint32_t p = (uint32_t)(py * P.W + px);
    uint32_t rng = rt::wang_hash(rt::seed_index(px, py, P.W) ^ P.seed);
    double xx, yy;
    pixel_xy(P, px, py, xx, yy);
    const size_t oi = 3 * ((size_t)(P.H - py - 1) * P.W + px);
    const d3 w = mk(dcolor[oi], dcolor[oi + 1], dcolor[oi + 2]);
    const double inv_spp = 1.0 / double(P.spp);
    const d3 color = mk(P.lcol[0] / light_scale, P.lcol[1] / light_scale, P.lcol[2] / light_scale);
    const d3 lcol = arr3(P.lcol), lcenter = arr3(P.lpos);
    for (int s = 0; s < P.spp; ++s) {
      d3 o, d;
      lens_ray(P, xx, yy, rng, o, d);
      DiffBounce br[kMaxDiffBounces];
      int nb = 0;
      bool miss = false;
      for (int k = 0; k <= P.depth; ++k) {
        int2 r = rec[((size_t)p * P.spp + s) * (P.depth + 1) + k];
        if (r.x < 0) {
          miss = true;
          break;
        }
        DiffBounce& b = br[nb++];
        b.o = o;
        b.d = d;
        b.tri = leaf_tri_of(B, r.x);
        RayD ray = make_ray(o, d);
        tri_test(ray, B.ltri + kLeafStride * (size_t)r.x, b.t, b.al, b.be);
        b.lpoint = light_point(P, rng);
        const d3 N0 = shading_normal(M, b.tri, b.al, b.be);
        const LightTerms T = light_terms(P, N0, o + d * b.t, b.lpoint);
        b.g = T.fo * T.nl;
        b.occ_tri = r.y >= 0 ? leaf_tri_of(B, r.y) : r.y;
        b.next = k < P.depth;
        if (b.next) {
          double u1 = rt::rand_unit(rng);
          double u2 = rt::rand_unit(rng);
          d3 N = N0;
          b.flip = rt::dot(N, d) > 0;
          if (b.flip) N = mk(-N.x, -N.y, -N.z);
          double sphi, cphi;
          rt::sincos_cw(2.0 * 3.141592653589793 * u2, &sphi, &cphi);
          b.xs = sqrt(u1) * cphi;
          b.ys = sqrt(1.0 - u1);
          b.zs = sqrt(u1) * sphi;
          o = o + d * b.t;
          d = cosine_dir(N, u1, u2);
        }
      }
      // Throughput before each bounce: thr_k = product of the albedos hit before k.
      d3 thr[kMaxDiffBounces + 1];
      thr[0] = mk(1, 1, 1);
      for (int k = 0; k < nb; ++k) thr[k + 1] = rt::hadamard(thr[k], alb(albedo, br[k].tri));
      // Reverse sweep. Each bounce term gives derivatives to its own albedo,
      // to every albedo in its throughput (products that leave that factor
      // out, so no division) and to its occluder's albedo; the position
      // adjoint flows back through the shading, the next ray and the hit.
      rta::v3 obar = rta::V(0, 0, 0), dbar = rta::V(0, 0, 0);
      for (int k = nb - 1; k >= 0; --k) {
        const DiffBounce& b = br[k];
        const d3 kd = alb(albedo, b.tri);
        d3 sh = mk(1, 1, 1);
        if (b.occ_tri == -2) sh = mk(0, 0, 0);
        if (b.occ_tri >= 0) sh = alb(albedo, b.occ_tri);
        const int ia = M.idx[3 * b.tri], ib = M.idx[3 * b.tri + 1], ic = M.idx[3 * b.tri + 2];
        const rta::v3 na = rta::V(M.nx[ia], M.ny[ia], M.nz[ia]), nbv = rta::V(M.nx[ib], M.ny[ib], M.nz[ib]),
                      nc = rta::V(M.nx[ic], M.ny[ic], M.nz[ic]);
        rta::Shade S = rta::shade_fwd(to_a(b.o), to_a(b.d), b.t, b.al, b.be, na, nbv, nc, to_a(b.lpoint),
                                      to_a(lcenter), P.c0, P.c1, P.c2);
        // term_k = inv_spp * w % thr_k % sh % kd % lcol * g
        // term_k = inv_spp * w % thr_k % sh % kd_k % lcol * g_k
        d3 base = rt::hadamard(rt::hadamard(w, thr[k]), rt::hadamard(sh, lcol)) * inv_spp;  // all but kd, g
        double gbar = rt::dot(rt::hadamard(base, kd), mk(1, 1, 1));
        light_acc += rt::dot(rt::hadamard(rt::hadamard(w, thr[k]), rt::hadamard(rt::hadamard(sh, kd), color)),
                             mk(1, 1, 1)) * inv_spp * b.g;
        // d/d kd_k
        atomic_add3(g_alb + 3 * (size_t)b.tri, base * b.g);
        // d/d albedo of earlier hits (through thr_k)
        for (int j = 0; j < k; ++j) {
          d3 ex = mk(1, 1, 1);
          for (int i = 0; i < k; ++i)
            if (i != j) ex = rt::hadamard(ex, alb(albedo, br[i].tri));
          d3 gj = rt::hadamard(rt::hadamard(rt::hadamard(w, ex), rt::hadamard(sh, kd)), lcol) * (inv_spp * b.g);
          atomic_add3(g_alb + 3 * (size_t)br[j].tri, gj);
        }
        // d/d occluder albedo
        if (b.occ_tri >= 0)
          atomic_add3(g_alb + 3 * (size_t)b.occ_tri,
                      rt::hadamard(rt::hadamard(rt::hadamard(w, thr[k]), kd), lcol) * (inv_spp * b.g));
        // positions: next-ray adjoint -> normal, then shading, then the hit
        rta::v3 Nextra = rta::V(0, 0, 0);
        rta::v3 qbar = obar;
        if (b.next) {
          rta::v3 Nf = b.flip ? -S.N : S.N;
          rta::Frame F = rta::frame_fwd(Nf, b.xs, b.ys, b.zs);
          rta::v3 nfbar = rta::frame_adj(F, Nf, dbar);
          Nextra = b.flip ? -nfbar : nfbar;
        }
        rta::ShadeAdj a = rta::shade_adj(S, na, nbv, nc, P.c1, P.c2, gbar, Nextra);
        qbar += a.q;
        // q = o + t d
        rta::v3 ob = qbar, db = qbar * b.t;
        double tbar = rta::dot(qbar, to_a(b.d));
        const rta::v3 A = rta::V(vx[ia], vy[ia], vz[ia]), Bv = rta::V(vx[ib], vy[ib], vz[ib]),
                      C = rta::V(vx[ic], vy[ic], vz[ic]);
        rta::TriAdj h = rta::hit_adj(A, Bv, C, to_a(b.o), to_a(b.d), b.t, b.al, b.be, tbar, a.al, a.be);
        ob += h.o;
        db += h.d;
        atomic_add3(g_pos + 3 * (size_t)ia, from_a(h.A));
        atomic_add3(g_pos + 3 * (size_t)ib, from_a(h.B));
        atomic_add3(g_pos + 3 * (size_t)ic, from_a(h.C));
        obar = ob;
        dbar = db;
      }
      if (miss) {
        // background term inv_spp * w % thr_nb * 0.9: albedos of all hits before the miss
        for (int j = 0; j < nb; ++j) {
          d3 ex = mk(1, 1, 1);
          for (int i = 0; i < nb; ++i)
            if (i != j) ex = rt::hadamard(ex, alb(albedo, br[i].tri));
          atomic_add3(g_alb + 3 * (size_t)br[j].tri, rt::hadamard(w, ex) * (0.9 * inv_spp));
        }
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) light_acc += __shfl_xor(light_acc, off);
  if ((threadIdx.x & 63) == 0 && light_acc != 0.0) atomicAdd(g_light, light_acc);
}

}  // namespace

hipError_t rt_launch_fill_albedo(rt_material* mats, int32_t* ids, const double* albedo, int32_t n, hipStream_t s) {
  k_fill_albedo<<<cdiv(n, 256), 256, 0, s>>>(mats, ids, albedo, n);
  return hipGetLastError();
}

hipError_t rt_launch_fill_i32(int32_t* p, size_t n, int32_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t blocks = (n + 255) / 256;
  k_fill_i32<<<(unsigned)(blocks < 65535 ? blocks : 65535), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t rt_launch_zero_f64(double* p, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t blocks = (n + 255) / 256;
  k_zero_f64<<<(unsigned)(blocks < 65535 ? blocks : 65535), 256, 0, s>>>(p, n);
  return hipGetLastError();
}

hipError_t rt_launch_adjoint(const rt_bvh* b, const rt_mesh_view* m, const double* albedo, double light_scale,
                             const rt_light* light, const rt_camera* cam, const rt_render_opts* opts,
                             const int32_t* records, const double* dcolor, double* g_albedo, double* g_light,
                             double* g_pos, hipStream_t s) {
  rt_light L = *light;
  for (int k = 0; k < 3; ++k) L.color[k] *= light_scale;
  RenderParams P = rt_make_params(cam, &L, opts, m->n);
  MeshShade M{m->nx_dev, m->ny_dev, m->nz_dev, m->idx_dev, m->mat_id_dev, nullptr};
  const int tiles = cdiv(P.W, 16) * cdiv(P.H, 16);
  k_adjoint<<<tiles, kThreads, 0, s>>>(b->view(), M, m->vx_dev, m->vy_dev, m->vz_dev, P, albedo, light_scale,
                                       (const int2*)records, dcolor, g_albedo, g_light, g_pos);
  return hipGetLastError();
}

// ------------------------------------------------------------ supersampling resolve
namespace {

// One thread per output pixel (kernels/scene.cu:127-155): clamp each of
// the ss x ss samples, sum, divide.
__global__ void k_resolve(const float* __restrict__ big, int W, int H, int ss, double div, float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= W || y >= H) return;
  double acc[3] = {0, 0, 0};
  for (int i = 0; i < ss; ++i)
    for (int j = 0; j < ss; ++j) {
      const float* p = big + 3 * ((size_t)(y * ss + i) * (size_t)(W * ss) + (x * ss + j));
      for (int c = 0; c < 3; ++c) {
        double v = p[c];
        v = v < 0 ? 0.0 : (v > 1 ? 1.0 : v);
        acc[c] += v;
      }
    }
  float* o = out + 3 * ((size_t)y * W + x);
  for (int c = 0; c < 3; ++c) o[c] = (float)(ss == 1 ? acc[c] : acc[c] / div);
}

}  // namespace

hipError_t rt_launch_resolve(const float* big, int32_t W, int32_t H, int32_t ss, bool reference_divisor, float* out,
                             hipStream_t s) {
  dim3 blk(16, 16), grd((unsigned)cdiv(W, 16), (unsigned)cdiv(H, 16));
  const double div = reference_divisor ? (double)ss : (double)ss * (double)ss;
  k_resolve<<<grd, blk, 0, s>>>(big, W, H, ss, div, out);
  return hipGetLastError();
}

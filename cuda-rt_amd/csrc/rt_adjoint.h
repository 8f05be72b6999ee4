// Reverse-mode derivatives of one path bounce for the differentiable
// render (host and device; no HIP dependency so it also builds with g++).
//
// Forward quantities follow the render kernel; every derivative treats the
// float-rounded square roots of the reference's normalize() as exact square
// roots. Visibility (which triangle a ray hits, whether a shadow ray is
// blocked) is held fixed: edge terms are not differentiated.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define RTA_HD __host__ __device__ inline
#else
#define RTA_HD inline
#endif

namespace rta {

struct v3 {
  double x, y, z;
};
RTA_HD v3 V(double x, double y, double z) { return v3{x, y, z}; }
RTA_HD v3 operator+(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
RTA_HD v3 operator-(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
RTA_HD v3 operator-(v3 a) { return V(-a.x, -a.y, -a.z); }
RTA_HD v3 operator*(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
RTA_HD v3 operator*(double s, v3 a) { return V(a.x * s, a.y * s, a.z * s); }
RTA_HD v3& operator+=(v3& a, v3 b) {
  a = a + b;
  return a;
}
RTA_HD double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RTA_HD v3 cross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
RTA_HD double len(v3 a) { return sqrt(dot(a, a)); }

// y = x / |x|  ->  xbar += (ybar - y (y . ybar)) / |x|
RTA_HD v3 normalize_adj(v3 x, v3 ybar) {
  double l = len(x);
  v3 y = x * (1.0 / l);
  return (ybar - y * dot(y, ybar)) * (1.0 / l);
}

// c = a x b  ->  abar += b x cbar, bbar += cbar x a
RTA_HD void cross_adj(v3 a, v3 b, v3 cbar, v3& abar, v3& bbar) {
  abar += cross(b, cbar);
  bbar += cross(cbar, a);
}

// Ray-triangle hit (t, alpha, beta) solves o + t d = A + alpha E1 + beta E2
// with E1 = B - A, E2 = C - A. Writing M = [-d | E1 | E2], the tangent is
// dx = M^-1 (do + t dd - w dA - alpha dB - beta dC), w = 1 - alpha - beta,
// so with r = M^-T xbar: obar += r, dbar += t r, Abar -= w r, Bbar -= alpha r,
// Cbar -= beta r. M^-T xbar = (xbar0 (c1 x c2) + xbar1 (c2 x c0) + xbar2 (c0 x c1)) / det M.
struct TriAdj {
  v3 o, d, A, B, C;
};
RTA_HD TriAdj hit_adj(v3 A, v3 B, v3 C, v3 o, v3 d, double t, double al, double be, double tbar, double abar,
                      double bbar) {
  (void)o;  // the tangent does not depend on the origin itself
  v3 c0 = -d, c1 = B - A, c2 = C - A;
  v3 k0 = cross(c1, c2), k1 = cross(c2, c0), k2 = cross(c0, c1);
  double det = dot(c0, k0);
  v3 r = (k0 * tbar + k1 * abar + k2 * bbar) * (1.0 / det);
  double w = 1.0 - al - be;
  TriAdj g;
  g.o = r;
  g.d = r * t;
  g.A = r * (-w);
  g.B = r * (-al);
  g.C = r * (-be);
  return g;
}

// Cosine-weighted direction about n (the frame of the render kernel):
// h = n with its smallest-magnitude component set to 1, x = normalize(h x n),
// z = normalize(x x n), dir = normalize(xs x + ys n + zs z).
struct Frame {
  v3 h, xx, x, zz, z, v, dir;
  int k;  // replaced component
  double xs, ys, zs;
};
RTA_HD Frame frame_fwd(v3 n, double xs, double ys, double zs) {
  Frame F;
  F.xs = xs;
  F.ys = ys;
  F.zs = zs;
  F.h = n;
  float ax = fabsf((float)n.x), ay = fabsf((float)n.y), az = fabsf((float)n.z);
  if (ax <= ay && ax <= az) {
    F.h.x = 1.0;
    F.k = 0;
  } else if (ay <= ax && ay <= az) {
    F.h.y = 1.0;
    F.k = 1;
  } else {
    F.h.z = 1.0;
    F.k = 2;
  }
  F.xx = cross(F.h, n);
  F.x = F.xx * (1.0 / len(F.xx));
  F.zz = cross(F.x, n);
  F.z = F.zz * (1.0 / len(F.zz));
  F.v = (F.x * xs + n * ys) + F.z * zs;
  F.dir = F.v * (1.0 / len(F.v));
  return F;
}
RTA_HD v3 frame_adj(const Frame& F, v3 n, v3 dirbar) {
  v3 vbar = normalize_adj(F.v, dirbar);
  v3 xbar = vbar * F.xs, nbar = vbar * F.ys, zbar = vbar * F.zs;
  v3 zzbar = normalize_adj(F.zz, zbar);
  cross_adj(F.x, n, zzbar, xbar, nbar);
  v3 xxbar = normalize_adj(F.xx, xbar);
  v3 hbar = V(0, 0, 0);
  cross_adj(F.h, n, xxbar, hbar, nbar);
  if (F.k != 0) nbar.x += hbar.x;
  if (F.k != 1) nbar.y += hbar.y;
  if (F.k != 2) nbar.z += hbar.z;
  return nbar;
}

// Shading geometry of one hit: smooth normal from fixed vertex normals,
// light direction to the jittered light point, distance falloff to the light
// centre, and the geometric factor g = fo * max(0, N . L).
struct Shade {
  v3 q, Nt, N, Lv, L, ld;
  double dist, den, fo, nl, g;
  bool fo_clamped, lit;
};
RTA_HD Shade shade_fwd(v3 o, v3 d, double t, double al, double be, v3 na, v3 nb, v3 nc, v3 lpoint, v3 lcenter,
                       double c0, double c1, double c2) {
  Shade S;
  S.q = o + d * t;
  S.Nt = (na * (1.0 - (al + be)) + nb * al) + nc * be;
  S.N = S.Nt * (1.0 / len(S.Nt));
  S.Lv = lpoint - S.q;
  S.L = S.Lv * (1.0 / len(S.Lv));
  S.ld = lcenter - S.q;
  S.dist = len(S.ld);
  S.den = c0 + (c1 + c2 * S.dist) * S.dist;
  double f = fabs(S.den) < 1e-8 ? 1.0 : 1.0 / S.den;
  S.fo_clamped = !(f < 1.0) || fabs(S.den) < 1e-8;
  S.fo = S.fo_clamped ? 1.0 : f;
  double nl = dot(S.N, S.L);
  S.lit = nl > 0;
  S.nl = S.lit ? nl : 0.0;
  S.g = S.fo * S.nl;
  return S;
}

struct ShadeAdj {
  v3 q, N;
  double al, be;
};
// Given gbar (and Nbar from later use of N), the adjoints of q, alpha, beta.
RTA_HD ShadeAdj shade_adj(const Shade& S, v3 na, v3 nb, v3 nc, double c1, double c2, double gbar, v3 Nbar_extra) {
  ShadeAdj a;
  a.q = V(0, 0, 0);
  double fobar = gbar * S.nl, nlbar = gbar * S.fo;
  v3 Nbar = Nbar_extra, Lbar = V(0, 0, 0);
  if (S.lit) {
    Nbar += S.L * nlbar;
    Lbar += S.N * nlbar;
  }
  if (!S.fo_clamped) {
    double distbar = fobar * (-S.fo * S.fo) * (c1 + 2.0 * c2 * S.dist);
    a.q += (S.ld * (1.0 / S.dist)) * (-distbar);
  }
  a.q += -normalize_adj(S.Lv, Lbar);
  v3 Ntbar = normalize_adj(S.Nt, Nbar);
  a.N = Nbar;
  a.al = dot(Ntbar, nb - na);
  a.be = dot(Ntbar, nc - na);
  return a;
}

}  // namespace rta

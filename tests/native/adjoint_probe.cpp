// Test-only C wrapper over the reverse-mode bounce derivatives
// (cuda-rt_amd/csrc/rt_adjoint.h) so the CPU tests can compare them with
// central finite differences of the matching double-precision forwards.
#include "../../cuda-rt_amd/csrc/rt_adjoint.h"

using rta::v3;
using rta::V;

static v3 ld(const double* p) { return V(p[0], p[1], p[2]); }
static void st(double* p, v3 v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}

extern "C" {

// in = A, B, C, o, d (15 doubles); out = t, alpha, beta (Moller-Trumbore in doubles).
void probe_hit_fwd(const double* in, double* out) {
  v3 A = ld(in), B = ld(in + 3), C = ld(in + 6), o = ld(in + 9), d = ld(in + 12);
  v3 AB = B - A, AC = C - A;
  v3 P = rta::cross(d, AC);
  double det = rta::dot(AB, P);
  v3 T = o - A;
  v3 Q = rta::cross(T, AB);
  out[0] = rta::dot(AC, Q) / det;
  out[1] = rta::dot(T, P) / det;
  out[2] = rta::dot(d, Q) / det;
}

// bar = (tbar, abar, bbar); out = adjoints of A, B, C, o, d.
void probe_hit_adj(const double* in, const double* bar, double* out) {
  double f[3];
  probe_hit_fwd(in, f);
  rta::TriAdj g = rta::hit_adj(ld(in), ld(in + 3), ld(in + 6), ld(in + 9), ld(in + 12), f[0], f[1], f[2], bar[0],
                               bar[1], bar[2]);
  st(out, g.A);
  st(out + 3, g.B);
  st(out + 6, g.C);
  st(out + 9, g.o);
  st(out + 12, g.d);
}

void probe_frame_fwd(const double* n, const double* s, double* dir) {
  rta::Frame F = rta::frame_fwd(ld(n), s[0], s[1], s[2]);
  st(dir, F.dir);
}

void probe_frame_adj(const double* n, const double* s, const double* dirbar, double* nbar) {
  rta::Frame F = rta::frame_fwd(ld(n), s[0], s[1], s[2]);
  st(nbar, rta::frame_adj(F, ld(n), ld(dirbar)));
}

// in = o, d, t, alpha, beta, na, nb, nc, lpoint, lcenter, c0, c1, c2 (30 doubles)
static rta::Shade shade_of(const double* in) {
  return rta::shade_fwd(ld(in), ld(in + 3), in[6], in[7], in[8], ld(in + 9), ld(in + 12), ld(in + 15), ld(in + 18),
                        ld(in + 21), in[24], in[25], in[26]);
}

// out = g, N
void probe_shade_fwd(const double* in, double* out) {
  rta::Shade S = shade_of(in);
  out[0] = S.g;
  st(out + 1, S.N);
}

// out = adjoints of q, alpha, beta given gbar and an extra adjoint of N.
void probe_shade_adj(const double* in, double gbar, const double* Nbar, double* out) {
  rta::Shade S = shade_of(in);
  rta::ShadeAdj a = rta::shade_adj(S, ld(in + 9), ld(in + 12), ld(in + 15), in[25], in[26], gbar, ld(Nbar));
  st(out, a.q);
  out[3] = a.al;
  out[4] = a.be;
}
}

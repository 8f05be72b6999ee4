// Host/device math shared by the build and render kernels.
//
// Everything here is compiled with -ffp-contract=off so each expression is
// evaluated exactly as written: the FP64 ray generation, box test and
// triangle test reproduce the reference's double-precision arithmetic
// operation by operation.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define RT_HD __host__ __device__ __forceinline__

namespace rt {

// inc/common.h:2
constexpr double kRayEps = 0.00000001;

struct d3 {
  double x, y, z;
};
RT_HD d3 mk(double x, double y, double z) { return d3{x, y, z}; }
RT_HD d3 operator+(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
RT_HD d3 operator-(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
RT_HD d3 operator*(d3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
RT_HD d3 hadamard(d3 a, d3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
RT_HD double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RT_HD d3 cross(d3 a, d3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
// The reference's normalize()/norm() take the length through sqrtf
// (inc/vec.h:118-140): the squared length is rounded to float first.
RT_HD d3 normalize(d3 a) {
  double inv = 1.0 / (double)sqrtf((float)(a.x * a.x + a.y * a.y + a.z * a.z));
  return a * inv;
}
RT_HD double length_f(d3 a) { return (double)sqrtf((float)(a.x * a.x + a.y * a.y + a.z * a.z)); }

// sin and cos of one double: Cody & Waite (1980) reduction by pi/2 split in
// three parts, Taylor series on [-pi/4, pi/4] to x^17 / x^18, Horner order.
// Valid for |x| < 2^20. The host and the device run the same operations.
RT_HD void sincos_cw(double x, double* s, double* c) {
  const double kPio2Hi = 1.5707963267341256;
  const double kPio2Mid = 6.077100506303966e-11;
  const double kPio2Lo = 2.0222662487959506e-21;
  const double kTwoOverPi = 0.6366197723675814;
  double kd = floor(x * kTwoOverPi + 0.5);
  double r = ((x - kd * kPio2Hi) - kd * kPio2Mid) - kd * kPio2Lo;
  double z = r * r;
  double ps = 1.0 / 355687428096000.0;
  ps = (-1.0 / 1307674368000.0) + z * ps;
  ps = (1.0 / 6227020800.0) + z * ps;
  ps = (-1.0 / 39916800.0) + z * ps;
  ps = (1.0 / 362880.0) + z * ps;
  ps = (-1.0 / 5040.0) + z * ps;
  ps = (1.0 / 120.0) + z * ps;
  ps = (-1.0 / 6.0) + z * ps;
  double sr = r + (r * z) * ps;
  double pc = 1.0 / 6402373705728000.0;
  pc = (-1.0 / 20922789888000.0) + z * pc;
  pc = (1.0 / 87178291200.0) + z * pc;
  pc = (-1.0 / 479001600.0) + z * pc;
  pc = (1.0 / 3628800.0) + z * pc;
  pc = (-1.0 / 40320.0) + z * pc;
  pc = (1.0 / 720.0) + z * pc;
  pc = (-1.0 / 24.0) + z * pc;
  pc = 0.5 + z * pc;
  double cr = 1.0 - z * pc;
  long k = (long)kd;
  switch (k & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

// mrand (kernels/mrand.cu:3-29): xorshift32 (13, 17, 5), Wang hash seeding,
// draws as double(float(u) / 2^32), so a draw can be exactly 1.0.
RT_HD uint32_t xorshift32(uint32_t& s) {
  s ^= (s << 13);
  s ^= (s >> 17);
  s ^= (s << 5);
  return s;
}
RT_HD uint32_t wang_hash(uint32_t s) {
  s = (s ^ 61u) ^ (s >> 16);
  s *= 9u;
  s = s ^ (s >> 4);
  s *= 0x27d4eb2du;
  s = s ^ (s >> 15);
  return s;
}
RT_HD double rand_unit(uint32_t& s) {
  uint32_t v = xorshift32(s);
  return (double)(float)v * 1.0 / 4294967296.0;
}
// Per-pixel seed index: the linear thread id of a 16x16-block launch with
// floor(width/16) blocks per row (inc/indexing.h:4-9, kernels/raytracer.cu:56-58).
RT_HD uint32_t seed_index(int px, int py, int width) {
  int gx = width / 16;
  int bid = (px / 16) + (py / 16) * gx;
  return (uint32_t)(bid * 256 + (py % 16) * 16 + (px % 16));
}

// Ordered-integer encoding of floats for atomicMin/atomicMax.
RT_HD uint32_t f2ord(float f) {
  uint32_t b = __builtin_bit_cast(uint32_t, f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
RT_HD float ord2f(uint32_t o) {
  uint32_t b = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  return __builtin_bit_cast(float, b);
}

}  // namespace rt

// Oracle copy of the double-precision sine/cosine used by the thin-lens ray generator.
//
// TEST INFRASTRUCTURE: this header belongs to the CPU oracle. It is checked
// against the host libm (tests/test_oracle.py) and the product's own device
// implementation is checked bit-for-bit against it on the GPU.
//
// Method: Cody & Waite (1980) argument reduction by pi/2 split into three
// parts (two 33-bit heads and a tail) followed by the Taylor series of sin
// and cos on [-pi/4, pi/4] truncated after the x^17 / x^18 terms, evaluated
// by Horner's rule in a fixed order. Valid for |x| < 2^20; the callers pass
// angles in [0, 2*pi].
#pragma once
#include <math.h>

namespace orc_math {

static const double kPio2Hi = 1.5707963267341256;      // pi/2 rounded to 33 bits
static const double kPio2Mid = 6.077100506303966e-11;  // next 33 bits
static const double kPio2Lo = 2.0222662487959506e-21;  // remainder
static const double kTwoOverPi = 0.6366197723675814;

static inline double sin_poly(double r) {
  double z = r * r;
  double p = 1.0 / 355687428096000.0;          // 1/17!
  p = (-1.0 / 1307674368000.0) + z * p;         // -1/15!
  p = (1.0 / 6227020800.0) + z * p;             // 1/13!
  p = (-1.0 / 39916800.0) + z * p;              // -1/11!
  p = (1.0 / 362880.0) + z * p;                 // 1/9!
  p = (-1.0 / 5040.0) + z * p;                  // -1/7!
  p = (1.0 / 120.0) + z * p;                    // 1/5!
  p = (-1.0 / 6.0) + z * p;                     // -1/3!
  return r + (r * z) * p;
}

static inline double cos_poly(double r) {
  double z = r * r;
  double p = 1.0 / 6402373705728000.0;         // 1/18!
  p = (-1.0 / 20922789888000.0) + z * p;        // -1/16!
  p = (1.0 / 87178291200.0) + z * p;            // 1/14!
  p = (-1.0 / 479001600.0) + z * p;             // -1/12!
  p = (1.0 / 3628800.0) + z * p;                // 1/10!
  p = (-1.0 / 40320.0) + z * p;                 // -1/8!
  p = (1.0 / 720.0) + z * p;                    // 1/6!
  p = (-1.0 / 24.0) + z * p;                    // -1/4!
  p = 0.5 + z * p;                              // 1/2! + z*(-1/4! + ...)
  return 1.0 - z * p;
}

static inline void sincos(double x, double* s, double* c) {
  double kd = floor(x * kTwoOverPi + 0.5);
  double r = ((x - kd * kPio2Hi) - kd * kPio2Mid) - kd * kPio2Lo;
  long k = (long)kd;
  double sr = sin_poly(r), cr = cos_poly(r);
  switch (k & 3) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

}  // namespace orc_math

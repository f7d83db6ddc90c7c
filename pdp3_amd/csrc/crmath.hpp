// Correctly rounded FP64 reciprocal square root for the FP64 plain form.  The
// reference evaluates pow(1 + u^2/c^2, -0.5) with glibc, whose result is the
// correctly rounded 1/sqrt in all but a few hundredths of a percent of cases
// (the device pow differs in ~2%).  The gamma factor is therefore taken as
// 1/sqrt(y) followed by a choice between the two neighbouring doubles made on
// the residual 1 - y c^2 evaluated with error-free products.
#pragma once

#include <hip/hip_runtime.h>

namespace pdp3 {

// two-sum / two-product error-free transformations
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// residual 1 - x*c^2 evaluated with error-free products (x*c^2 ~ 1)
__device__ __forceinline__ double rsqrt_residual(double x, double c) {
  double p = c * c;
  double pe = fma(c, c, -p);
  double q = x * p;
  double qe = fma(x, p, -q);
  double t = 1.0 - q;  // exact: q is within a few ulps of 1
  return (t - qe) - x * pe;
}

// 1/sqrt(x), correctly rounded except in exact-tie cases, for x > 0 finite
__device__ __forceinline__ double rsqrt_cr(double x) {
  double y = 1.0 / sqrt(x);
  double r0 = fabs(rsqrt_residual(x, y));
  double yl = nextafter(y, 0.0), yh = nextafter(y, 1e308);
  double rl = fabs(rsqrt_residual(x, yl)), rh = fabs(rsqrt_residual(x, yh));
  double best = y, rb = r0;
  if (rl < rb) {
    best = yl;
    rb = rl;
  }
  if (rh < rb) best = yh;
  return best;
}

// x / y, correctly rounded, from yr = 1/y correctly rounded: q = x yr lies
// within 1.5 ulp of x/y; one fma correction with the exact remainder x - q y
// makes it faithful, and a second one correctly rounded (Markstein's theorem:
// a faithful quotient corrected with the correctly rounded reciprocal is the
// correctly rounded quotient).  Five FP64 operations instead of the IEEE
// division sequence with its quarter-rate reciprocal, for divisors whose
// reciprocal is shared (uniform or tabulated); valid for finite operands
// without overflow or underflow, which is what the particle step divides.
// A zero x yields +0 (the IEEE division keeps the sign of a -0 dividend).
__device__ __forceinline__ double div_cr(double x, double y, double yr) {
  double q = x * yr;
  double r = fma(-q, y, x);
  q = fma(r, yr, q);
  r = fma(-q, y, x);
  return fma(r, yr, q);
}

}  // namespace pdp3

// gc_math.h -- the one transcendental on the Rodas3 path.
//
// Integrate_Cell_S computes Fac = FacSafe / Err**(ONE/ros_ELO) with
// ros_ELO = 3 (src/geos-chem/KPP/fullchem/gckpp_BatchIntegrator.F90:667); the
// exponent is the double y0 = RN(1/3) = 1/3 - 2^-54/3, and the reference CPU
// build takes x**y0 from libm pow.  This header gives a correctly rounded
// x**y0 for finite x > 0 (Err is floored at 1e-10 before the call, so that is
// the whole domain used), so the GPU result is the exact IEEE-rounded value:
//
//   x**y0 = cbrt(x) * x**(-d),  d = 1/3 - y0 = 2^-54/3
//         = c + r/(3c^2) - d*ln(x)*c + O(2^-100 c)
//
// where c is any double within a few ulps of cbrt(x) and r = x - c^3 is
// formed exactly with FMA-based two-products.  The sum is rounded once.
// The result is correctly rounded unless x**y0 lies within about 2^-95
// relative of a rounding midpoint.
//
// Usable from host (for testing against libm) and device code.
#pragma once
#include <math.h>

#ifndef GC_HD
#if defined(__HIPCC__)
#define GC_HD __host__ __device__
#else
#define GC_HD
#endif
#endif

GC_HD static inline double gc_pow_third(double x) {
  const double c = cbrt(x);
  // c^2 = c2h + c2l exactly
  const double c2h = c * c;
  const double c2l = fma(c, c, -c2h);
  // c^3 = c3h + c3l (c3l carries the two low-order parts)
  const double c3h = c2h * c;
  const double c3l = fma(c2h, c, -c3h) + c2l * c;
  // residual x - c^3; x - c3h is exact (Sterbenz: the two are within 2x)
  const double r = (x - c3h) - c3l;
  const double t1 = r / (3.0 * c2h);
  // d = 1/3 - RN(1/3) = 2^-54 / 3
  const double d = 1.8503717077085943e-17;
  const double t2 = -(d * log(x)) * c;
  return c + (t1 + t2);
}

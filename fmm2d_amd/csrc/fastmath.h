// FP64 natural logarithm and reciprocal for the pair kernels, written for the
// vector ALU: a few fused multiply-adds instead of the library log (~90
// instructions with its double-double steps) and the IEEE division sequence.
//
// log: x = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) with
//   s = (m - 1) / (m + 1), |s| <= 0.1716, evaluated as 2 s (1 + s^2/3 + ... +
//   s^20/21) by Horner's rule (the first omitted term is below 2.3e-17
//   relative), and e ln 2 added with ln 2 split into a head exact in its
//   product with e and a tail.  s comes from a reciprocal refined by Newton's
//   method and one residual correction.  Relative error a few units in the
//   last place over positive normal x (checked against long double on the
//   host); x must be positive and normal (the pair kernels pass |z|^2 of pairs
//   not excluded by the threshold).
// rcp: 1/x from the hardware approximation and two Newton steps (FMA), within
//   one unit in the last place of the correctly rounded value for normal x.
#pragma once
#include "common.h"

namespace fmm2d {

FMM2D_HD double fast_rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
#else
  double r = 1.0 / x;
#endif
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

FMM2D_HD double fast_log(double x) {
  int ex;
  double m = frexp(x, &ex);  // x = m 2^ex, m in [0.5, 1)
  if (m < 0.70710678118654752440) {
    m = 2.0 * m;
    ex -= 1;
  }
  const double t = m - 1.0;  // exact (m in [0.707, 1.414))
  const double d = m + 1.0;
  const double r = fast_rcp(d);
  double s = t * r;
  s = fma(fma(-s, d, t), r, s);  // residual correction
  const double z = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, z, 1.0 / 19.0);
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  // log m = 2 s + 2 s z p
  const double two_s = 2.0 * s;
  const double lm = fma(two_s * z, p, two_s);
  // ln 2 = kLn2Hi + kLn2Lo; kLn2Hi has 32 significant bits, so e * kLn2Hi is exact
  constexpr double kLn2Hi = 6.93147180369123816490e-01;
  constexpr double kLn2Lo = 1.90821492927058770002e-10;
  const double de = (double)ex;
  return fma(de, kLn2Hi, fma(de, kLn2Lo, lm));
}

}  // namespace fmm2d

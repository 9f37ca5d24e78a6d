// FP32 Hankel functions H0^(1)(x), H1^(1)(x) of a real argument x > 0, for
// the mixed-precision Helmholtz near field (real wavenumbers).
//
//   x < 2   ascending series in q = x^2/4 (Abramowitz & Stegun 9.1.10, 9.1.11,
//           9.1.13), nine terms, Horner's rule;
//   x >= 2  H_nu(x) = sqrt(2/(pi x)) e^{i (x - nu pi/2 - pi/4)} (P_nu + i Q_nu)(1/x),
//           P, Q degree-9 polynomials in t = 1/x fitted on [0, 1/2] (Hankel's
//           asymptotic form with its slowly varying amplitude fitted instead of
//           truncated); the phase x - pi/4 is reduced modulo 2 pi in FP64, so
//           its accuracy does not degrade with x.
// Coefficients: hank32_tables.h (gen_hank32.py).  Relative error ~1e-7 (FP32
// rounding), absolute near the zeros of J and Y; checked against scipy on the
// device (tests/test_direct_gpu.py).
#pragma once
#include "common.h"
#include "hank32_tables.h"

namespace fmm2d {

struct c32 {
  float re, im;
};
FMM2D_HD c32 operator*(c32 a, c32 b) { return c32{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
FMM2D_HD c32 operator+(c32 a, c32 b) { return c32{a.re + b.re, a.im + b.im}; }
FMM2D_HD c32 operator-(c32 a, c32 b) { return c32{a.re - b.re, a.im - b.im}; }
FMM2D_HD c32 operator*(float s, c32 a) { return c32{s * a.re, s * a.im}; }

template <int N>
FMM2D_HD float horner32(const float (&c)[N], float t) {
  float r = c[N - 1];
#pragma unroll
  for (int k = N - 2; k >= 0; --k) r = fmaf(r, t, c[k]);
  return r;
}

// xd: the argument in FP64 (its phase is reduced in FP64 for x >= 2)
FMM2D_HD void hank01_f32(double xd, c32* h0, c32* h1) {
  constexpr float kTwoOverPi = 0.63661977236758134f;
  const float x = (float)xd;
  if (x < 2.0f) {
    constexpr float A[FMM2D_H32_NSER] = {FMM2D_H32_A};
    constexpr float B[FMM2D_H32_NSER] = {FMM2D_H32_B};
    constexpr float C[FMM2D_H32_NSER] = {FMM2D_H32_C};
    constexpr float D[FMM2D_H32_NSER] = {FMM2D_H32_D};
    const float q = 0.25f * x * x;
    const float J0 = horner32(A, q);
    const float J1 = 0.5f * x * horner32(B, q);
    const float S0 = horner32(C, q);
    const float S1 = horner32(D, q);
    const float L = logf(0.5f * x) + 0.57721566490153286f;
    const float Y0 = kTwoOverPi * fmaf(L, J0, S0);
    const float Y1 = kTwoOverPi * (L * J1 - 1.0f / x) - (0.5f * kTwoOverPi * 0.5f) * x * S1;
    *h0 = c32{J0, Y0};
    *h1 = c32{J1, Y1};
  } else {
    constexpr float P0[FMM2D_H32_NFIT + 1] = {FMM2D_H32_P0};
    constexpr float Q0[FMM2D_H32_NFIT + 1] = {FMM2D_H32_Q0};
    constexpr float P1[FMM2D_H32_NFIT + 1] = {FMM2D_H32_P1};
    constexpr float Q1[FMM2D_H32_NFIT + 1] = {FMM2D_H32_Q1};
    const float t = 1.0f / x;
    const c32 m0{horner32(P0, t), horner32(Q0, t)};
    const c32 m1{horner32(P1, t), horner32(Q1, t)};
    const float amp = sqrtf(kTwoOverPi * t);
    // x - pi/4 reduced to [-pi, pi] in FP64
    constexpr double kTwoPi = 6.28318530717958647692;
    const double ph = xd - 0.78539816339744830962;
    const double red = ph - kTwoPi * rint(ph * (1.0 / kTwoPi));
    float s, c;
    sincosf((float)red, &s, &c);
    const c32 e{amp * c, amp * s};      // sqrt(2/(pi x)) e^{i (x - pi/4)}
    *h0 = m0 * e;
    *h1 = m1 * c32{e.im, -e.re};       // times e^{-i pi/2} = -i
  }
}

}  // namespace fmm2d

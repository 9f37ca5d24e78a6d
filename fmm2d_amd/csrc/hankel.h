// Hankel functions of the first kind, H0^(1)(z) and H1^(1)(z), for complex z.
//
// Replaces the reference's hank103 (common/hank103.f:116-243), which uses
// fitted coefficient tables; this evaluator is written from three textbook
// representations and is used by the P2P kernels and the Helmholtz operators:
//
//   |z| < 2.5          ascending series for J0, J1, Y0, Y1 (|z| < 1.75 when
//                      Im z > |Re z|)
//                      (Abramowitz & Stegun 9.1.10, 9.1.11, 9.1.13);
//   otherwise, |z|<17  the integral representation
//                        H_nu(z) = sqrt(2/(pi z)) e^{i(z - nu pi/2 - pi/4)} / Gamma(nu+1/2)
//                                  * int_0^inf e^{-t} t^{nu-1/2} (1 + i t/(2z))^{nu-1/2} dt
//                      (Hankel's integral, Watson, Treatise on Bessel Functions, sec. 7.2)
//                      evaluated by a 32-point generalized Gauss-Laguerre rule (gl_tables.h);
//   |z| >= 17          Hankel's asymptotic expansion (A&S 9.2.7), summed until
//                      the terms fall below 1e-17 relative.
//
// Real positive z = x (every real wavenumber) takes its own path, hank01_real:
//   x < 4              the ascending series as four polynomials of degree 16
//                      in q = x^2/4 (Horner's rule; coefficients exact
//                      rationals rounded once) and one logarithm;
//   x >= 4             H_nu(x) = sqrt(2/(pi x)) e^{i (x - nu pi/2 - pi/4)} M_nu(4/x),
//                      the amplitude M_nu as polynomials of degree 20 in s = 4/x
//                      (Chebyshev interpolants computed with mpmath; error about
//                      one unit in the last place of |M| ~ 1), so a pair costs a
//                      reciprocal, a square root, one sincos and 84 fused
//                      multiply-adds with no data-dependent loop.
//                      Both tables: hank64_tables.h (gen_hank64.py).
//
// Below the real axis (Im zk < 0: exponentially growing kernels) the rule
// above is used for arg z >= -0.3; deeper in the fourth quadrant the series
// (|z| < 5), Hankel's integral along the ray arg t = -pi/4, which keeps the
// integrand's branch point t = 2iz away from the path (|z| < 17), or the
// asymptotic expansion; the third quadrant by the continuation formula
// (A&S 9.1.36 with J = (H1 + H2)/2 and H2(w) = conj H1(conj w))
//   H_n(z) = (-1)^n [2 H_n(-z) + conj H_n(-conj z)],
// both of whose arguments lie in the first or the fourth quadrant.
//
// Accuracy (checked against mpmath): <= 5e-15 relative on the closed upper
// half-plane, which covers every Im(zk) >= 0, and <= 1e-14 below it.
#pragma once
#include "common.h"
#include "fastmath.h"
#include "gl_tables.h"
#include "hank64_tables.h"

namespace fmm2d {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEuler = 0.57721566490153286061;

constexpr int kSeriesMax = 26;

// nterm (<= kSeriesMax) ascending-series terms; the k-th term is bounded by
// (|z|^2/4)^k/(k!)^2.  The loop is unrolled so that the factors 1/k^2,
// 1/(k(k+1)), 1/k, 1/(k+1) are compile-time constants.
FMM2D_HD void hank01_series(cplx z, cplx* h0, cplx* h1, int nterm) {
  cplx q = -0.25 * (z * z);
  cplx t0 = mkc(1.0), t1 = mkc(1.0);
  cplx J0 = t0, S1sum = t1, S0 = mkc(0.0), S1 = mkc(1.0);  // k = 0 terms
  double Hk = 0.0;
#pragma unroll
  for (int k = 1; k <= kSeriesMax; ++k) {
    if (k > nterm) break;
    const double dk = (double)k;
    t0 = t0 * q * (1.0 / (dk * dk));
    t1 = t1 * q * (1.0 / (dk * (dk + 1.0)));
    Hk += 1.0 / dk;
    J0 += t0;
    S1sum += t1;
    S0 += Hk * t0;
    S1 += (2.0 * Hk + 1.0 / (dk + 1.0)) * t1;
  }
  cplx zh = 0.5 * z;
  cplx J1 = zh * S1sum;
  cplx L = logc(zh);
  L.re += kEuler;
  cplx Y0 = (2.0 / kPi) * (L * J0 - S0);
  cplx Y1 = (2.0 / kPi) * (L * J1) - (2.0 / kPi) * invc(z) - (1.0 / kPi) * (zh * S1);
  *h0 = cplx{J0.re - Y0.im, J0.im + Y0.re};
  *h1 = cplx{J1.re - Y1.im, J1.im + Y1.re};
}

// The same series for a real argument x > 0 (real wavenumbers): J and Y are
// real, H = J + i Y.
FMM2D_HD void hank01_series_real(double x, cplx* h0, cplx* h1, int nterm) {
  const double q = -0.25 * (x * x);
  double t0 = 1.0, t1 = 1.0, J0 = 1.0, S1sum = 1.0, S0 = 0.0, S1 = 1.0, Hk = 0.0;
#pragma unroll
  for (int k = 1; k <= kSeriesMax; ++k) {
    if (k > nterm) break;
    const double dk = (double)k;
    t0 = t0 * q * (1.0 / (dk * dk));
    t1 = t1 * q * (1.0 / (dk * (dk + 1.0)));
    Hk += 1.0 / dk;
    J0 += t0;
    S1sum += t1;
    S0 += Hk * t0;
    S1 += (2.0 * Hk + 1.0 / (dk + 1.0)) * t1;
  }
  const double xh = 0.5 * x;
  const double J1 = xh * S1sum;
  const double L = log(xh) + kEuler;
  const double Y0 = (2.0 / kPi) * (L * J0 - S0);
  const double Y1 = (2.0 / kPi) * (L * J1) - (2.0 / kPi) / x - (1.0 / kPi) * (xh * S1);
  *h0 = cplx{J0, Y0};
  *h1 = cplx{J1, Y1};
}

// sqrt(2/(pi z)) * exp(i (z - pi/4))
FMM2D_HD cplx hank_prefactor(cplx z) {
  cplx s = sqrtc(invc(z) * (2.0 / kPi));
  // exp(i z) first: subtracting pi/4 from a large Re z before the argument
  // reduction would cost |z|*1e-16 of phase accuracy
  cplx e = expc(cplx{-z.im, z.re}) * cplx{0.70710678118654752440, -0.70710678118654752440};
  return s * e;
}

FMM2D_HD void hank01_laguerre(cplx z, cplx* h0, cplx* h1) {
  static constexpr double gl_m_x[FMM2D_GL_N] = {FMM2D_GL_M_X};
  static constexpr double gl_m_w[FMM2D_GL_N] = {FMM2D_GL_M_W};
  static constexpr double gl_p_x[FMM2D_GL_N] = {FMM2D_GL_P_X};
  static constexpr double gl_p_w[FMM2D_GL_N] = {FMM2D_GL_P_W};
  cplx c = divc(cplx{0.0, 0.5}, z);  // i/(2z)
  cplx s0 = mkc(0.0), s1 = mkc(0.0);
#pragma unroll 4
  for (int k = 0; k < FMM2D_GL_N; ++k) {
    cplx u0 = cplx{1.0 + gl_m_x[k] * c.re, gl_m_x[k] * c.im};
    cplx u1 = cplx{1.0 + gl_p_x[k] * c.re, gl_p_x[k] * c.im};
    s0 += gl_m_w[k] * invc(sqrtc(u0));
    s1 += gl_p_w[k] * sqrtc(u1);
  }
  cplx p = hank_prefactor(z);
  *h0 = p * s0;
  cplx p1 = cplx{p.im, -p.re};  // p * e^{-i pi/2}
  *h1 = p1 * s1;
}

FMM2D_HD void hank01_asymptotic(cplx z, cplx* h0, cplx* h1) {
  cplx zi8 = invc(8.0 * z);
  cplx s0 = mkc(1.0), s1 = mkc(1.0), t0 = mkc(1.0), t1 = mkc(1.0);
  double best0 = 1.0, best1 = 1.0;
  bool go0 = true, go1 = true;
#pragma unroll 1
  for (int k = 1; k <= 40 && (go0 || go1); ++k) {
    double m = (double)(2 * k - 1);
    cplx f = zi8 * (1.0 / (double)k);
    f = cplx{-f.im, f.re};  // times i
    if (go0) {
      t0 = t0 * f * (0.0 - m * m);
      double a = absc(t0);
      if (a > best0) { go0 = false; } else { best0 = a; s0 += t0; if (a < 1e-17) go0 = false; }
    }
    if (go1) {
      t1 = t1 * f * (4.0 - m * m);
      double a = absc(t1);
      if (a > best1) { go1 = false; } else { best1 = a; s1 += t1; if (a < 1e-17) go1 = false; }
    }
  }
  cplx p = hank_prefactor(z);
  *h0 = p * s0;
  cplx p1 = cplx{p.im, -p.re};
  *h1 = p1 * s1;
}

// Hankel's integral along the ray t = s (1 - i), for the fourth quadrant below
// arg z = -0.3: the same nodes x_k with weights w_k e^{i x_k} (1 - i)^{a+1}.
FMM2D_HD void hank01_laguerre_rot(cplx z, cplx* h0, cplx* h1) {
  static constexpr double gl_m_x[FMM2D_GL_N] = {FMM2D_GL_M_X};
  static constexpr double gl_p_x[FMM2D_GL_N] = {FMM2D_GL_P_X};
  static constexpr double gl_m_wr[FMM2D_GL_N] = {FMM2D_GL_M_WRRE};
  static constexpr double gl_m_wi[FMM2D_GL_N] = {FMM2D_GL_M_WRIM};
  static constexpr double gl_p_wr[FMM2D_GL_N] = {FMM2D_GL_P_WRRE};
  static constexpr double gl_p_wi[FMM2D_GL_N] = {FMM2D_GL_P_WRIM};
  cplx c = divc(cplx{0.0, 0.5}, z) * cplx{1.0, -1.0};  // i (1 - i) / (2z)
  cplx s0 = mkc(0.0), s1 = mkc(0.0);
#pragma unroll 4
  for (int k = 0; k < FMM2D_GL_N; ++k) {
    cplx u0 = cplx{1.0 + gl_m_x[k] * c.re, gl_m_x[k] * c.im};
    cplx u1 = cplx{1.0 + gl_p_x[k] * c.re, gl_p_x[k] * c.im};
    s0 += cplx{gl_m_wr[k], gl_m_wi[k]} * invc(sqrtc(u0));
    s1 += cplx{gl_p_wr[k], gl_p_wi[k]} * sqrtc(u1);
  }
  cplx p = hank_prefactor(z);
  *h0 = p * s0;
  *h1 = cplx{p.im, -p.re} * s1;
}

// The ascending series, the quadrature and the asymptotic expansion for z in
// the closed upper half-plane or with arg z >= -0.3.
FMM2D_HD void hank01_upper(cplx z, double r, cplx* h0, cplx* h1) {
  // The series cancels like exp(|z| + Im z) relative to |H|: near the
  // positive imaginary axis it hands over to the quadrature already at 1.75.
  if (r < 1.75 || (r < 2.5 && fabs(z.re) >= z.im)) {
    // term bound below 1e-17 at k = 8 (r<0.5), 10 (r<1.2), 14 (r<2.5)
    const int nterm = r < 0.5 ? 8 : (r < 1.2 ? 10 : 14);
    if (z.im == 0.0 && z.re > 0.0)
      hank01_series_real(z.re, h0, h1, nterm);
    else
      hank01_series(z, h0, h1, nterm);
  } else if (r < 17.0) {
    hank01_laguerre(z, h0, h1);
  } else {
    hank01_asymptotic(z, h0, h1);
  }
}

// z in the fourth quadrant (Re z >= 0 > Im z).
FMM2D_HD void hank01_quad4(cplx z, double r, cplx* h0, cplx* h1) {
  if (atan2(z.im, z.re) >= -0.3)
    hank01_upper(z, r, h0, h1);
  else if (r < 5.0)
    hank01_series(z, h0, h1, 26);  // (6.25)^26/(26!)^2 < 1e-32
  else if (r < 17.0)
    hank01_laguerre_rot(z, h0, h1);
  else
    hank01_asymptotic(z, h0, h1);
}

// Below the real axis: out of line, so that the kernels whose arguments stay
// in the upper half-plane (every Im zk >= 0) carry only the call.
static __host__ __device__ __noinline__ void hank01_lower(cplx z, double r, cplx* h0, cplx* h1) {
  if (z.re >= 0.0) {
    hank01_quad4(z, r, h0, h1);
    return;
  }
  if (r < 1.75) {  // no cancellation below the axis at this size
    hank01_series(z, h0, h1, r < 0.5 ? 8 : (r < 1.2 ? 10 : 14));
    return;
  }
  cplx a0, a1, b0, b1;
  hank01_upper(cplx{-z.re, -z.im}, r, &a0, &a1);  // first quadrant
  hank01_quad4(cplx{-z.re, z.im}, r, &b0, &b1);   // fourth quadrant
  *h0 = cplx{2.0 * a0.re + b0.re, 2.0 * a0.im - b0.im};
  *h1 = cplx{-(2.0 * a1.re + b1.re), -(2.0 * a1.im - b1.im)};
}

// The coefficient rows of hank64_tables.h: NSER series rows (a, b, c, d) then
// NFIT + 1 amplitude rows (P0, Q0, P1, Q1).
constexpr int kH64Rows = FMM2D_H64_NSER + FMM2D_H64_NFIT + 1;
#if defined(__HIP_DEVICE_COMPILE__)
static __constant__ double kH64Tab[4 * kH64Rows] = {FMM2D_H64_SER, FMM2D_H64_AMP};
#endif

// H0^(1)(x), H1^(1)(x) for real x > 0.  On the device the coefficients are
// read with scalar loads from a constant table whose address passes through an
// empty asm statement: the compiler cannot hoist them out of the caller's loop
// (as literal constants, all 152 were hoisted into vector registers ahead of
// the loop, which then ran at one wave per SIMD or spilled).
FMM2D_HD void hank01_real(double x, cplx* h0, cplx* h1) {
#if defined(__HIP_DEVICE_COMPILE__)
  const __attribute__((address_space(4))) double* tab =
      (const __attribute__((address_space(4))) double*)kH64Tab;
  asm volatile("" : "+s"(tab));
#else
  static constexpr double tab[4 * kH64Rows] = {FMM2D_H64_SER, FMM2D_H64_AMP};
#endif
  if (x < FMM2D_H64_X1) {
    const double q = 0.25 * (x * x);
    constexpr int n = FMM2D_H64_NSER - 1;
    double a = tab[4 * n], b = tab[4 * n + 1], c = tab[4 * n + 2], d = tab[4 * n + 3];
#pragma unroll
    for (int k = n - 1; k >= 0; --k) {
      a = fma(a, q, tab[4 * k]);
      b = fma(b, q, tab[4 * k + 1]);
      c = fma(c, q, tab[4 * k + 2]);
      d = fma(d, q, tab[4 * k + 3]);
    }
    const double xh = 0.5 * x;
    const double J1 = xh * b;
    const double L = fast_log(xh) + kEuler;
    const double Y0 = (2.0 / kPi) * fma(L, a, c);
    const double Y1 = (2.0 / kPi) * (L * J1) - (2.0 / kPi) * fast_rcp(x) - (1.0 / kPi) * (xh * d);
    *h0 = cplx{a, Y0};
    *h1 = cplx{J1, Y1};
    return;
  }
  const double t = fast_rcp(x);
  const double s = FMM2D_H64_X1 * t;
  constexpr int o = 4 * FMM2D_H64_NSER, n = FMM2D_H64_NFIT;
  double p0 = tab[o + 4 * n], q0 = tab[o + 4 * n + 1], p1 = tab[o + 4 * n + 2],
         q1 = tab[o + 4 * n + 3];
#pragma unroll
  for (int k = n - 1; k >= 0; --k) {
    p0 = fma(p0, s, tab[o + 4 * k]);
    q0 = fma(q0, s, tab[o + 4 * k + 1]);
    p1 = fma(p1, s, tab[o + 4 * k + 2]);
    q1 = fma(q1, s, tab[o + 4 * k + 3]);
  }
  // sqrt(2/(pi x)) e^{i (x - pi/4)}: e^{i x} first, then the fixed rotation,
  // so the phase keeps the accuracy of sincos(x)
  double sn, cs;
  sincos(x, &sn, &cs);
  const double m = sqrt((2.0 / kPi) * t) * 0.70710678118654752440;
  const cplx e{m * (cs + sn), m * (sn - cs)};
  *h0 = e * cplx{p0, q0};
  const cplx m1 = e * cplx{p1, q1};
  *h1 = cplx{m1.im, -m1.re};  // times e^{-i pi/2}
}

// The general evaluator (any nonzero z; used for z off the positive real
// axis): out of line, so that kernels of real wavenumbers, whose arguments
// never come here, carry only the call and keep their registers for the real
// path.
static __host__ __device__ __noinline__ void hank01_complex(cplx z, cplx* h0, cplx* h1) {
  double r = absc(z);
  if (z.im >= 0.0)
    hank01_upper(z, r, h0, h1);
  else
    hank01_lower(z, r, h0, h1);
}

// H0^(1)(z), H1^(1)(z).  z must be nonzero.
FMM2D_HD void hank01(cplx z, cplx* h0, cplx* h1) {
  if (z.im == 0.0 && z.re > 0.0)
    hank01_real(z.re, h0, h1);
  else
    hank01_complex(z, h0, h1);
}

// hank01 as a call, for callers outside the pair loops (one evaluation per
// expansion): their kernels then keep no registers for the real path.
static __host__ __device__ __noinline__ void hank01_ool(cplx z, cplx* h0, cplx* h1) {
  hank01(z, h0, h1);
}

}  // namespace fmm2d

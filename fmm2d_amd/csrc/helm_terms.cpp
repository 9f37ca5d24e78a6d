// Helmholtz expansion-order rule (h2dterms, helmholtz/h2dterms.f:35-96).
//
// nterms = j + 1 for the first j >= 2 with
//   (|J_j(z2) H_j(z1)| + |J_{j-1}(z2) H_{j-1}(z1)|) |H_0(z1)|
//       < eps (|J_0(z2) H_0(z1)| + |J_1(z2) H_1(z1)|),
// z1 = 1.5 k bs, z2 = k bs sqrt(2)/2, searched up to j = 50000 (else ier 13).
// H_n by upward recurrence from H0, H1 (stable for H); J_n by Miller's
// backward recurrence normalised with the Jacobi-Anger identity at theta = pi,
//   exp(-i z) = J_0(z) + 2 sum_{n>=1} (-i)^n J_n(z),
// (exp(+i z) with (+i)^n below the real axis), which does not cancel for
// complex z.  Long double keeps the unnormalised recurrences in range.
#include <cmath>
#include <complex>
#include <vector>

#include "hankel.h"
#include "helm_ops.h"

namespace fmm2d {

using lc = std::complex<long double>;

// J_0..J_n(z) for n >= 1
std::vector<lc> bessel_j(lc z, int n) {
  std::vector<lc> out(n + 1, lc(0));
  if (std::abs(z) < 1e-200L) {
    out[0] = 1;
    return out;
  }
  const lc zi = 1.0L / z;
  // start index: forward recurrence from n until it has grown by 1e20
  int ntop = n + 1;
  {
    lc f0 = 0, f1 = 1;
    for (int i = n; i < n + 100000; ++i) {
      lc f2 = (2.0L * i) * zi * f1 - f0;
      f0 = f1;
      f1 = f2;
      if (std::norm(f2) > 1e40L) {
        ntop = i + 1;
        break;
      }
    }
  }
  std::vector<lc> f(ntop + 1, lc(0));
  f[ntop] = 0;
  f[ntop - 1] = 1;
  for (int i = ntop - 1; i >= 1; --i) {
    f[i - 1] = (2.0L * i) * zi * f[i] - f[i + 1];
    if (std::norm(f[i - 1]) > 1e300L) {  // rescale the tail
      for (int k = i - 1; k <= ntop; ++k) f[k] *= 1e-150L;
    }
  }
  const lc mul = (z.imag() < 0) ? lc(0, 1) : lc(0, -1);
  lc psi = 0, sn = 1;
  for (int i = 1; i <= ntop; ++i) {
    sn *= mul;
    psi += f[i] * sn;
  }
  psi = 2.0L * psi + f[0];
  const lc ez = (z.imag() < 0) ? std::exp(lc(0, 1) * z) : std::exp(lc(0, -1) * z);
  const lc sc = ez / psi;
  for (int i = 0; i <= n; ++i) out[i] = f[i] * sc;
  return out;
}

int h2dterms(double bsize, cplx zk, double eps, int* ier) {
  *ier = 0;
  const lc k(zk.re, zk.im);
  const lc z1 = k * (long double)bsize * 1.5L;
  const lc z2 = k * (long double)bsize * (std::sqrt(2.0L) / 2.0L);
  const int ntmax = 50000;
  cplx h0d, h1d;
  hank01(cplx{(double)z1.real(), (double)z1.imag()}, &h0d, &h1d);
  // search in growing windows so that small problems stay cheap
  int nwin = std::min(ntmax, std::max(64, (int)(4 * std::abs(z1)) + 64));
  for (;;) {
    std::vector<lc> J = bessel_j(z2, nwin);
    std::vector<long double> prod(nwin + 1);
    lc hm1 = lc(h0d.re, h0d.im), h = lc(h1d.re, h1d.im);
    const long double absh0 = std::abs(hm1);
    prod[0] = std::abs(J[0] * hm1);
    prod[1] = std::abs(J[1] * h);
    const lc zi = 1.0L / z1;
    for (int i = 1; i < nwin; ++i) {
      lc hn = (2.0L * i) * zi * h - hm1;
      hm1 = h;
      h = hn;
      prod[i + 1] = std::abs(J[i + 1]) * std::abs(h);
    }
    const long double x0 = prod[0] + prod[1];
    for (int j = 2; j <= nwin; ++j) {
      const long double x = (prod[j] + prod[j - 1]) * absh0;
      if (x < (long double)eps * x0) return j + 1;
    }
    if (nwin >= ntmax) break;
    nwin = std::min(ntmax, 2 * nwin);
  }
  *ier = 13;
  return 10001;
}

}  // namespace fmm2d

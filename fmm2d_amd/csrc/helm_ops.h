// Helmholtz FMM (hfmm2d) on the device.
#pragma once
#include "../../include/fmm2d_amd.h"
#include <complex>
#include <vector>

#include "common.h"

namespace fmm2d {

struct HelmArgs {
  int nd;
  double eps;
  cplx zk;
  int64_t ns, nt;
  const double* src;
  const double* targ;
  int ifcharge, ifdipole;
  const cplx* charge;
  const cplx* dipstr;
  const double* dipvec;
  int ifpgh, ifpghtarg;
  cplx *pot, *grad, *hess, *pottarg, *gradtarg, *hesstarg;
  const fmm2d_opts* opts;
};

// helmholtz/hfmm2d.f:19
int helmholtz_fmm(const HelmArgs& a, hipStream_t st);

// hfmm2d_mps (helmholtz/hfmm2d_mps.f:19): n multipole expansions (centres
// (2,n), scales, orders mterms, coefficients mpole(nd, -m:m) at the 1-based
// offsets impole) in, local expansions about the same centres out (same
// layout, written into local's blocks).  Device pointers.
struct HelmMpsArgs {
  int nd;
  double eps;
  cplx zk;
  int64_t n, ntot;
  const double* centers;
  const double* rscale;
  const int* mterms;
  const cplx* mpole;
  const int64_t* impole;
  cplx* local;
  const fmm2d_opts* opts;
};
int helmholtz_mps(const HelmMpsArgs& a, hipStream_t st);

// h2dterms (helmholtz/h2dterms.f:35-96); *ier = 13 when no order up to 50000
// meets the criterion.
int h2dterms(double bsize, cplx zk, double eps, int* ier);

// J_0(z) .. J_n(z) in long double (Miller's backward recurrence, normalised by
// the Jacobi-Anger identity), n >= 1.
std::vector<std::complex<long double>> bessel_j(std::complex<long double> z, int n);

}  // namespace fmm2d

// Laplace / Cauchy FMM (cfmm2d core, lfmm2d wrapper) on the device.
#pragma once
#include "../../include/fmm2d_amd.h"
#include "common.h"

namespace fmm2d {

struct CauchyArgs {
  int nd;
  double eps;
  int64_t ns, nt;
  const double* src;   // (2,ns) device
  const double* targ;  // (2,nt) device or null
  int ifcharge, ifdipole;
  const cplx* charge;   // (nd,ns)
  const cplx* dipstr;   // (nd,ns)
  const double* dipvec; // (nd,2,ns): lfmm2d, rfmm2d
  int ifpgh, ifpghtarg;
  cplx *pot, *grad, *hess, *pottarg, *gradtarg, *hesstarg;
  const fmm2d_opts* opts;  // may be null
};

// kind 0: cfmm2d (laplace/cfmm2d.f:23); kind 1: lfmm2d (laplace/lfmm2d.f:29);
// kind 2: rfmm2d (laplace/rfmm2d.f:27) -- charge, dipstr and the six outputs
// then point to real (double) arrays of the same shapes.
int laplace_fmm(int kind, const CauchyArgs& a, hipStream_t st);

// l2dterms (laplace/l2dterms.f:38-82) and lndiv2d (laplace/lndiv2d.f)
int l2dterms(double eps);
int lndiv2d(double eps, int64_t ns, int64_t nt);

}  // namespace fmm2d

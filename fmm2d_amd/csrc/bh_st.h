// Biharmonic and Stokes sums on the device (bh_st.hip): direct evaluators and
// the FMMs (biharmonic/bhfmm2d.f, stokes/stfmm2d.f).  All pointers are device
// pointers in the reference's column-major layouts.
#pragma once
#include "../../include/fmm2d_amd.h"
#include "common.h"

namespace fmm2d {

// bh2d_direct{c,d,cd}{p,g} (biharmonic/bhkernels2d.f:26-410): INCREMENTS
// vel(nd,nt) and, for pg = 2, grad(nd,3,nt)
struct BhDirectArgs {
  int nd;
  const double* src;  // (2,ns)
  int64_t ns;
  int ifcharge, ifdipole;
  const cplx* charge;  // (nd,2,ns)
  const cplx* dip;     // (nd,3,ns)
  const double* targ;  // (2,nt)
  int64_t nt;
  int pg;
  cplx* vel;
  cplx* grad;
  double thresh;
};
int bh_direct(const BhDirectArgs& a, hipStream_t st);

// st2ddirectstokg / st2ddirectstokstrsg (stokes/stokkernels2d.f:24, 169):
// INCREMENTS pot(nd,2,nt), pre(nd,nt), grad(nd,2,2,nt) (each may be null)
struct StDirectArgs {
  int nd;
  const double* src;
  int64_t ns;
  int ifstoklet;
  const double* stoklet;  // (nd,2,ns)
  int istress;            // 1: type-I stresslet
  const double* strslet;  // (nd,2,ns)
  const double* strsvec;  // (nd,2,ns)
  const double* targ;
  int64_t nt;
  double* pot;
  double* pre;
  double* grad;
  double thresh;
};
int st_direct(const StDirectArgs& a, hipStream_t st);

// bhfmm2d (biharmonic/bhfmm2d.f:1): OVERWRITES pot(nd,ns), grad(nd,3,ns),
// pottarg(nd,nt), gradtarg(nd,3,nt); Hessians are not computed (as in the
// reference)
struct BhArgs {
  int nd;
  double eps;
  int64_t ns;
  const double* src;
  int ifcharge;
  const cplx* charge;  // (nd,2,ns)
  int ifdipole;
  const cplx* dip;     // (nd,3,ns)
  int ifpgh;
  cplx* pot;
  cplx* grad;
  int64_t nt;
  const double* targ;
  int ifpghtarg;
  cplx* pottarg;
  cplx* gradtarg;
  const fmm2d_opts* opts;
};
int bh_fmm(const BhArgs& a, hipStream_t st);

// stfmm2d (stokes/stfmm2d.f:27): OVERWRITES pot(nd,2,ns), pre(nd,ns),
// grad(nd,2,2,ns) and the target outputs, as far as ifppreg / ifppregtarg ask
struct StArgs {
  int nd;
  double eps;
  int64_t ns;
  const double* src;
  int ifstoklet;
  const double* stoklet;
  int ifstrslet;
  const double* strslet;
  const double* strsvec;
  int ifppreg;
  double *pot, *pre, *grad;
  int64_t nt;
  const double* targ;
  int ifppregtarg;
  double *pottarg, *pretarg, *gradtarg;
  const fmm2d_opts* opts;
};
int st_fmm(const StArgs& a, hipStream_t st);

}  // namespace fmm2d

// Direct (all-pairs) evaluation entry used by the C-ABI.
#pragma once
#include <atomic>
#include "common.h"

namespace fmm2d {

struct DirectArgs {
  int family;        // 0 Cauchy (c2d), 1 Laplace (l2d), 2 Helmholtz (h2d)
  int nd;
  cplx zk;           // Helmholtz only
  const double* src; // (2,ns) column-major, device
  int64_t ns;
  int ifcharge, ifdipole;
  const cplx* charge;   // (nd,ns)
  const cplx* dipstr;   // (nd,ns)
  const double* dipvec; // (nd,2,ns), Laplace/Helmholtz only
  const double* targ;   // (2,nt)
  int64_t nt;
  int pg;               // 1 pot, 2 +grad, 3 +hess
  double thresh;
  cplx *pot, *grad, *hess;  // incremented
};

// noalloc: split partial results go to a static per-device array instead of
// stream-ordered allocations (fewer splits when they do not fit it), and no
// allocating call is made
int direct_eval(const DirectArgs& a, hipStream_t st, bool noalloc = false);

// r2d_direct{c,d,cd}{p,g,h} (laplace/rlapkernels2d.f): the Laplace kernels on
// real densities with real outputs.  The formulas are those of l2d_direct
// (laplace/lapkernels2d.f) with real arguments, so the real parts of the
// complex evaluation on the densities are exactly the real kernels' values.
int direct_eval_real(int nd, const double* src, int64_t ns, int ifcharge, const double* charge,
                     int ifdipole, const double* dipstr, const double* dipvec, const double* targ,
                     int64_t nt, int pg, double thresh, double* pot, double* grad, double* hess,
                     hipStream_t st);

// Bound on the blocks of one direct evaluation when the source range is split
// (fmm2d_set_direct_max_blocks; 0 = the default of about 2048).
extern std::atomic<int> g_direct_max_blocks;

}  // namespace fmm2d

// C-ABI entry points for direct evaluation and special functions.
// Declarations and reference citations: include/fmm2d_amd.h.
#include "../../include/fmm2d_amd.h"

#include "bh_st.h"
#include "common.h"
#include "p2p.h"

using namespace fmm2d;

static_assert(sizeof(fmm2d_cplx) == sizeof(cplx), "complex layout");

namespace {
int run_direct(int family, int nd, double zr, double zi, const double* sources, int64_t ns,
               int ifcharge, const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dipstr,
               const double* dipvec, const double* targ, int64_t nt, int ifpgh,
               fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, double thresh,
               void* stream, bool noalloc = false) {
  DirectArgs a;
  a.family = family;
  a.nd = nd;
  a.zk = cplx{zr, zi};
  a.src = sources;
  a.ns = ns;
  a.ifcharge = ifcharge;
  a.ifdipole = ifdipole;
  a.charge = reinterpret_cast<const cplx*>(charge);
  a.dipstr = reinterpret_cast<const cplx*>(dipstr);
  a.dipvec = dipvec;
  a.targ = targ;
  a.nt = nt;
  a.pg = ifpgh;
  a.thresh = thresh;
  a.pot = reinterpret_cast<cplx*>(pot);
  a.grad = reinterpret_cast<cplx*>(grad);
  a.hess = reinterpret_cast<cplx*>(hess);
  if ((ifcharge && !charge) || (ifdipole && !dipstr) || (ifdipole && family != 0 && !dipvec))
    return FMM2D_ERR_ARG;
  if (!pot || (ifpgh >= 2 && !grad) || (ifpgh >= 3 && !hess)) return FMM2D_ERR_ARG;
  if ((ns > 0 && !sources) || (nt > 0 && !targ)) return FMM2D_ERR_ARG;
  return direct_eval(a, reinterpret_cast<hipStream_t>(stream), noalloc);
}
}  // namespace

extern "C" {

int fmm2d_abi_version(void) { return FMM2D_ABI_VERSION; }

int fmm2d_c2d_direct(int nd, const double* sources, int64_t ns, int ifcharge,
                     const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dipstr,
                     const double* targ, int64_t nt, int ifpgh, fmm2d_cplx* pot,
                     fmm2d_cplx* grad, fmm2d_cplx* hess, double thresh, void* stream) {
  return run_direct(0, nd, 0.0, 0.0, sources, ns, ifcharge, charge, ifdipole, dipstr, nullptr,
                    targ, nt, ifpgh, pot, grad, hess, thresh, stream);
}

int fmm2d_l2d_direct(int nd, const double* sources, int64_t ns, int ifcharge,
                     const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dipstr,
                     const double* dipvec, const double* targ, int64_t nt, int ifpgh,
                     fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, double thresh,
                     void* stream) {
  return run_direct(1, nd, 0.0, 0.0, sources, ns, ifcharge, charge, ifdipole, dipstr, dipvec,
                    targ, nt, ifpgh, pot, grad, hess, thresh, stream);
}

int fmm2d_l2d_direct_noalloc(int nd, const double* sources, int64_t ns, int ifcharge,
                             const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dipstr,
                             const double* dipvec, const double* targ, int64_t nt, int ifpgh,
                             fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess,
                             double thresh, void* stream) {
  return run_direct(1, nd, 0.0, 0.0, sources, ns, ifcharge, charge, ifdipole, dipstr, dipvec,
                    targ, nt, ifpgh, pot, grad, hess, thresh, stream, true);
}

int fmm2d_h2d_direct(int nd, double zk_re, double zk_im, const double* sources, int64_t ns,
                     int ifcharge, const fmm2d_cplx* charge, int ifdipole,
                     const fmm2d_cplx* dipstr, const double* dipvec, const double* targ,
                     int64_t nt, int ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad,
                     fmm2d_cplx* hess, double thresh, void* stream) {
  return run_direct(2, nd, zk_re, zk_im, sources, ns, ifcharge, charge, ifdipole, dipstr, dipvec,
                    targ, nt, ifpgh, pot, grad, hess, thresh, stream);
}

int fmm2d_r2d_direct(int nd, const double* sources, int64_t ns, int ifcharge, const double* charge,
                     int ifdipole, const double* dipstr, const double* dipvec, const double* targ,
                     int64_t nt, int ifpgh, double* pot, double* grad, double* hess,
                     double thresh, void* stream) {
  return direct_eval_real(nd, sources, ns, ifcharge, charge, ifdipole, dipstr, dipvec, targ, nt,
                          ifpgh, thresh, pot, grad, hess, reinterpret_cast<hipStream_t>(stream));
}

int fmm2d_set_direct_max_blocks(int nblocks) {
  if (nblocks < 0) return FMM2D_ERR_ARG;
  return g_direct_max_blocks.exchange(nblocks);
}

int fmm2d_bh2d_direct(int nd, const double* sources, int64_t ns, int ifcharge,
                      const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dip,
                      const double* targ, int64_t nt, int ifpgh, fmm2d_cplx* vel,
                      fmm2d_cplx* grad, double thresh, void* stream) {
  if ((ns > 0 && !sources) || (nt > 0 && !targ)) return FMM2D_ERR_ARG;
  BhDirectArgs a{};
  a.nd = nd;
  a.src = sources;
  a.ns = ns;
  a.ifcharge = ifcharge;
  a.ifdipole = ifdipole;
  a.charge = reinterpret_cast<const cplx*>(charge);
  a.dip = reinterpret_cast<const cplx*>(dip);
  a.targ = targ;
  a.nt = nt;
  a.pg = ifpgh;
  a.vel = reinterpret_cast<cplx*>(vel);
  a.grad = reinterpret_cast<cplx*>(grad);
  a.thresh = thresh;
  return bh_direct(a, reinterpret_cast<hipStream_t>(stream));
}

int fmm2d_st2d_direct(int nd, const double* sources, int ifstoklet, const double* stoklet,
                      int istress, const double* strslet, const double* strsvec, int64_t ns,
                      const double* targ, int64_t nt, double* pot, double* pre, double* grad,
                      double thresh, void* stream) {
  if ((ns > 0 && !sources) || (nt > 0 && !targ)) return FMM2D_ERR_ARG;
  StDirectArgs a{};
  a.nd = nd;
  a.src = sources;
  a.ns = ns;
  a.ifstoklet = ifstoklet;
  a.stoklet = stoklet;
  a.istress = istress;
  a.strslet = strslet;
  a.strsvec = strsvec;
  a.targ = targ;
  a.nt = nt;
  a.pot = pot;
  a.pre = pre;
  a.grad = grad;
  a.thresh = thresh;
  return st_direct(a, reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"

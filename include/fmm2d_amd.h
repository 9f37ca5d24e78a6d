/*
 * fmm2d_amd -- C-ABI of the MI355X (gfx950) 2-D fast multipole library.
 *
 * Drop-in boundary for the hot path of flatironinstitute/fmm2d: the FMM
 * drivers (cfmm2d / lfmm2d / hfmm2d) and the direct evaluators
 * (c2d_direct* / l2d_direct* / h2d_direct*).  Every entry point mirrors the
 * argument meaning of the reference Fortran routine it replaces (cited per
 * function, paths relative to the reference's src/), with three changes:
 *   - arguments are passed by value and pointers are DEVICE pointers;
 *   - one entry point per family takes ifcharge/ifdipole/ifpgh flags instead
 *     of the reference's 27 fixed-flag wrappers;
 *   - an optional hipStream_t (as void*) orders the work; NULL = default stream.
 *
 * Array layouts are the reference's (Fortran, column-major, density index
 * fastest): sources(2,ns) = interleaved (x,y); charge(nd,ns); dipstr(nd,ns);
 * dipvec(nd,2,ns); pot(nd,n); Cauchy grad(nd,n), hess(nd,n); Laplace and
 * Helmholtz grad(nd,2,n), hess(nd,3,n).  Complex values are (re,im) double
 * pairs (fmm2d_cplx), layout-compatible with complex*16 / complex128.
 *
 * Return value: the reference's ier codes where they exist (0 ok, 4 workspace
 * allocation failure, 13 expansion-order search exhausted) and negative codes
 * for conditions the reference cannot report (-1 bad argument, -2 HIP error,
 * -3 unsupported option).
 */
#ifndef FMM2D_AMD_H
#define FMM2D_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FMM2D_ABI_VERSION 1

typedef struct {
  double re, im;
} fmm2d_cplx;

/* Library ABI version (FMM2D_ABI_VERSION).  No reference counterpart. */
int fmm2d_abi_version(void);

/*
 * Direct evaluation; INCREMENTS pot/grad/hess at the nt targets.
 *   replaces laplace/cauchykernels2d.f c2d_direct{c,d,cd}{p,g,h}
 *   (lines 3, 68, 145, 227, 292, 367, 447, 518, 599), e.g.
 *   c2d_directcg(nd,sources,ns,charge,targ,nt,pot,grad,thresh).
 * Pairs closer than thresh are skipped with each reference variant's own
 * rule (<= for the potential-only charge variants, < otherwise).
 */
int fmm2d_c2d_direct(int nd, const double* sources, int64_t ns, int ifcharge,
                     const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dipstr,
                     const double* targ, int64_t nt, int ifpgh, fmm2d_cplx* pot,
                     fmm2d_cplx* grad, fmm2d_cplx* hess, double thresh, void* stream);

/*   replaces laplace/lapkernels2d.f l2d_direct{c,d,cd}{p,g,h}
 *   (lines 8, 73, 151, 242, 314, 407, 513, 590, 695). */
int fmm2d_l2d_direct(int nd, const double* sources, int64_t ns, int ifcharge,
                     const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dipstr,
                     const double* dipvec, const double* targ, int64_t nt, int ifpgh,
                     fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, double thresh,
                     void* stream);

/*   replaces helmholtz/helmkernels2d.f h2d_direct{c,d,cd}{p,g,h}
 *   (lines 3, 71, 155, 250, 334, 435, 560, 651, 765); thresh applies to
 *   |zk|*|t-s| as in the reference. */
int fmm2d_h2d_direct(int nd, double zk_re, double zk_im, const double* sources, int64_t ns,
                     int ifcharge, const fmm2d_cplx* charge, int ifdipole,
                     const fmm2d_cplx* dipstr, const double* dipvec, const double* targ,
                     int64_t nt, int ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad,
                     fmm2d_cplx* hess, double thresh, void* stream);

/* H0^(1)(z[i]), H1^(1)(z[i]) for n device-resident arguments.
 *   replaces common/hank103.f:116 hank103(z,h0,h1,ifexpon) with ifexpon=1. */
int fmm2d_hank103(const fmm2d_cplx* z, int64_t n, fmm2d_cplx* h0, fmm2d_cplx* h1,
                  void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FMM2D_AMD_H */

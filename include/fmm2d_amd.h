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

/*
 * Optional controls and per-call diagnostics of the FMM drivers (NULL = the
 * reference's behaviour).  They generalise the reference's only tuning hook,
 * the *_ndiv entry points (laplace/lfmm2d_ndiv.f:29-32, laplace/cfmm2d_ndiv.f,
 * helmholtz/hfmm2d_ndiv.f:19-22: ndiv, idivflag, ifnear, timeinfo(8)).
 */
typedef struct {
  int32_t ndiv;        /* max points per leaf; <= 0: reference rule (lndiv2d / hndiv2d) */
  int32_t idivflag;    /* 0: refine on sources, 1: on targets, 2: on the max (pts_tree_build) */
  int32_t ifnear;      /* 1: add the list-1 near field (reference default), 0: omit it */
  int32_t precision;   /* 0: FP64 throughout; 1: mixed FP32/FP64 */
  int32_t n_nterms;    /* entries in nterms; 0: reference rule (l2dterms / h2dterms) */
  int32_t count_work;  /* 1: fill work[], near_pairs, m2l_pairs (one extra counting pass) */
  const int32_t* nterms; /* HOST array: expansion order for levels 0..n_nterms-1 */
  /* out: seconds per stage.  [0..7] are the reference's timeinfo(1..8): P2M,
   * list-4 P2L, M2M, M2L, L2L, list-3 M2P, L2P, near field (list-3 and L2P
   * run fused and are reported together in [6]); [8] tree + lists, [9] sort
   * and gather, [10] scatter, [11] whole call, [12] translation operators
   * (Helmholtz), [13] the M2L translate kernels alone ([3] then holds the
   * M2L planning: target classes and source slots), [15] the FP32 copies of
   * the strengths for the mixed-precision near field (Laplace; [7] then holds
   * the near-field kernels alone). */
  double timeinfo[16];
  int32_t nlevels, nboxes; /* out: tree depth and box count */
  int32_t pad1, pad2;
  /* out (count_work = 1): algorithmic work per stage, same slots as timeinfo: flops for
   * [0..7] (P2P 12+12*nd per Cauchy pair for pot+grad, M2L 8(p+1)^2 nd per
   * list-2 pair, 8(p+1) nd per point for P2M/L2P, 8(p+1)^2 nd per M2M/L2L
   * edge), bytes for [8] (40 per point + 80 per box); 0 when not counted. */
  double work[16];
  int64_t near_pairs;      /* out: source-target pairs in the near field */
  int64_t m2l_pairs;       /* out: list-2 translations performed */
  /* in: a handle from fmm2d_prepare for the same sources/targets arrays (NULL:
   * none); the tree build then starts from its sorted keys (no reference
   * counterpart) */
  void* prepared;
} fmm2d_opts;

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

/*
 * fmm2d_l2d_direct without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and increments as fmm2d_l2d_direct.  The partial results
 * of a source range split across workgroups go to a static 32 MiB array per
 * device (2^21 complex values) instead of a stream-ordered allocation; a call
 * whose split copies would not fit runs with fewer splits (none when one copy,
 * nd * nt * (1 + 2 [grad] + 3 [hess]) values, exceeds it), so no size is
 * refused and no allocating routine is called.  Every call on a device shares
 * that device's array: calls that overlap on one device (different streams or
 * host threads) race and must be serialised by the caller.
 */
int fmm2d_l2d_direct_noalloc(int nd, const double* sources, int64_t ns, int ifcharge,
                             const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dipstr,
                             const double* dipvec, const double* targ, int64_t nt, int ifpgh,
                             fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess,
                             double thresh, void* stream);

/*
 * Direct evaluation of the real Laplace kernels, INCREMENTS pot/grad/hess:
 *   replaces laplace/rlapkernels2d.f r2d_direct{c,d,cd}{p,g,h} (f2py: r2ddir,
 *   python/fmm2dpy/fmm2d.py:1370), real densities and outputs, the same
 *   per-variant exclusion rules as l2d_direct*.
 */
int fmm2d_r2d_direct(int nd, const double* sources, int64_t ns, int ifcharge, const double* charge,
                     int ifdipole, const double* dipstr, const double* dipvec, const double* targ,
                     int64_t nt, int ifpgh, double* pot, double* grad, double* hess,
                     double thresh, void* stream);

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

/*
 * Cauchy-kernel FMM; OVERWRITES the requested outputs.
 *   replaces laplace/cfmm2d.f:23 cfmm2d(nd,eps,ns,sources,ifcharge,charge,
 *   ifdipole,dipstr,iper,ifpgh,pot,grad,hess,nt,targ,ifpghtarg,pottarg,
 *   gradtarg,hesstarg,ier) -- ier is the return value; iper must be 0.
 * pot = sum c log|z-xi| + d/(z-xi) with the reference's analytic far-field
 * expansions; grad(nd,n) = d/dz, hess(nd,n) = d^2/dz^2.  Returns after the
 * outputs are complete on the device.
 */
int fmm2d_cfmm2d(int nd, double eps, int64_t ns, const double* sources, int ifcharge,
                 const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dipstr, int iper,
                 int ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, int64_t nt,
                 const double* targ, int ifpghtarg, fmm2d_cplx* pottarg, fmm2d_cplx* gradtarg,
                 fmm2d_cplx* hesstarg, fmm2d_opts* opts, void* stream);

/*
 * Laplace FMM with complex densities and dipole vectors; OVERWRITES outputs.
 *   replaces laplace/lfmm2d.f:29 lfmm2d(nd,eps,ns,sources,ifcharge,charge,
 *   ifdipole,dipstr,dipvec,iper,ifpgh,pot,grad,hess,nt,targ,ifpghtarg,
 *   pottarg,gradtarg,hesstarg,ier); grad(nd,2,n), hess(nd,3,n).
 */
int fmm2d_lfmm2d(int nd, double eps, int64_t ns, const double* sources, int ifcharge,
                 const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dipstr,
                 const double* dipvec, int iper, int ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad,
                 fmm2d_cplx* hess, int64_t nt, const double* targ, int ifpghtarg,
                 fmm2d_cplx* pottarg, fmm2d_cplx* gradtarg, fmm2d_cplx* hesstarg,
                 fmm2d_opts* opts, void* stream);

/*
 * rfmm2d (laplace/rfmm2d.f:27; f2py wrappers laplace/rfmm2dwrap.f): the Laplace
 * FMM on REAL charges (nd,ns), dipole strengths (nd,ns) and dipole vectors
 * (nd,2,ns) with REAL outputs pot (nd,n), grad (nd,2,n), hess (nd,3,n).  As the
 * reference, it runs the Cauchy FMM on the charges and on -(v1 + i v2) times
 * the dipole strengths and returns pot = Re, grad = (Re, -Im), hess = (Re,
 * -Im, -Re) (rfmm2d.f:65-139).
 */
int fmm2d_rfmm2d(int nd, double eps, int64_t ns, const double* sources, int ifcharge,
                 const double* charge, int ifdipole, const double* dipstr, const double* dipvec,
                 int iper, int ifpgh, double* pot, double* grad, double* hess, int64_t nt,
                 const double* targ, int ifpghtarg, double* pottarg, double* gradtarg,
                 double* hesstarg, fmm2d_opts* opts, void* stream);

/*
 * Helmholtz FMM; OVERWRITES outputs (ifpgh, ifpghtarg <= 2).
 *   replaces helmholtz/hfmm2d.f:19 hfmm2d(nd,eps,zk,ns,sources,ifcharge,
 *   charge,ifdipole,dipstr,dipvec,iper,ifpgh,pot,grad,hess,nt,targ,
 *   ifpghtarg,pottarg,gradtarg,hesstarg,ier).
 */
int fmm2d_hfmm2d(int nd, double eps, double zk_re, double zk_im, int64_t ns,
                 const double* sources, int ifcharge, const fmm2d_cplx* charge, int ifdipole,
                 const fmm2d_cplx* dipstr, const double* dipvec, int iper, int ifpgh,
                 fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, int64_t nt,
                 const double* targ, int ifpghtarg, fmm2d_cplx* pottarg,
                 fmm2d_cplx* gradtarg, fmm2d_cplx* hesstarg, fmm2d_opts* opts, void* stream);

/*
 * Helmholtz FMM whose sources are multipole expansions; OVERWRITES the local
 * expansions' blocks of local.
 *   replaces helmholtz/hfmm2d_mps.f:19 hfmm2d_mps(nd,eps,zk,nmpole,cmpole,
 *   rmpole,mterms,ntot,mpole,impole,local,ier) and, through opts->ndiv,
 *   helmholtz/hfmm2d_mps_ndiv.f:70 hfmm2d_mps_ndiv (ndiv = nmpole evaluates
 *   every pair directly, as hfmm2d_mps_direct).
 * Expansion i has centre cmpole(2,i), scale rmpole(i), order mterms(i) and
 * coefficients mpole(nd, -mterms(i):mterms(i)) starting at the 1-based offset
 * impole(i) of mpole (complex values, density fastest), meaning
 *   u(x) = sum_n mpole_n rmpole^|n| H_n(k r) e^{i n t}.
 * local receives, in the same layout and scaling (local_n rmpole^-|n|
 * J_n(k r) e^{i n t}), the local expansion at each centre of all the other
 * expansions (centres closer than |zk| boxsize 2^-51 are skipped).  ntot is
 * the length of mpole and local.  Default leaf capacity 1 (the reference's);
 * opts->ndiv, ifnear and nterms apply; precision must be 0.  Returns 13 when
 * h2dterms finds no order, FMM2D_ERR_UNSUPPORTED when an order exceeds the
 * translation kernels' limit (p_source + p_target <= 2047).
 */
int fmm2d_hfmm2d_mps(int nd, double eps, double zk_re, double zk_im, int64_t nmpole,
                     const double* cmpole, const double* rmpole, const int* mterms,
                     int64_t ntot, const fmm2d_cplx* mpole, const int64_t* impole,
                     fmm2d_cplx* local, fmm2d_opts* opts, void* stream);

/*
 * One FMM of a batch (fmm2d_batch): the arguments of fmm2d_cfmm2d (kind 0),
 * fmm2d_lfmm2d (kind 1), fmm2d_rfmm2d (kind 2: charge, dipstr and the outputs
 * point to doubles) or fmm2d_hfmm2d (kind 3: zk = zk_re + i zk_im), all device
 * pointers, iper = 0.  ier receives that call's return code.
 */
typedef struct fmm2d_job {
  int32_t kind, nd;
  double eps, zk_re, zk_im;
  int64_t ns;
  const double* sources;
  int32_t ifcharge, ifdipole;
  const void* charge;
  const void* dipstr;
  const double* dipvec;
  int32_t ifpgh, ifpghtarg;
  void* pot;
  void* grad;
  void* hess;
  int64_t nt;
  const double* targ;
  void* pottarg;
  void* gradtarg;
  void* hesstarg;
  fmm2d_opts* opts;
  int32_t ier, pad;
} fmm2d_job;

/*
 * Many independent FMMs in one call.  No counterpart in the reference (whose
 * drivers take one problem per call); for callers with many small problems of
 * uneven size, which one at a time leave the GPU idle between launches.
 * nstreams (1..16) host threads each take the next job, largest (ns + nt)
 * first, and run it on a stream of its own from a pool the library keeps per
 * device (default priority, never destroyed), so that several jobs' kernels
 * and host synchronisations overlap.  Returns when every job's device work is
 * done: 0, or the first nonzero ier in job order (FMM2D_ERR_ARG for a bad
 * njobs/jobs/nstreams or kind).
 */
int fmm2d_batch(int njobs, fmm2d_job* jobs, int nstreams);

/*
 * Adaptive quadtree (pts_tree_mem + pts_tree_build + pts_tree_sort,
 * common/pts_tree2d.f:34, 406, 1222) over device-resident points, numbered
 * box-for-box as the reference numbers it.  *handle receives an opaque tree;
 * copy it out with fmm2d_tree_get and release it with fmm2d_tree_free.
 */
int fmm2d_tree_build(const double* sources, int64_t ns, const double* targ, int64_t nt,
                     int idivflag, int ndiv, int nlmax, int* nlevels, int* nboxes,
                     void** handle, void* stream);

/*
 * Copy a tree to HOST arrays in the reference's conventions (1-based box and
 * point indices, -1 = none): laddr(2,0:nlevels), ilevel(nb), iparent(nb),
 * nchild(nb), ichild(4,nb), ncoll(nb), coll(9,nb), centers(2,nb),
 * boxsize(0:nlevels), isrc(ns), isrcse(2,nb), itarg(nt), itargse(2,nb).
 * When nlist1 is non-NULL the lists of computelists (common/tree_routs2d.f:486)
 * are also returned: nlistK(nb) and listK(mK,nb) with mK = 13, 27, 20, 5,
 * each row in ascending box order.  Any output pointer may be NULL.
 */
int fmm2d_tree_get(void* handle, int* laddr, int* ilevel, int* iparent, int* nchild,
                   int* ichild, int* ncoll, int* coll, double* centers, double* boxsize,
                   int* isrc, int* isrcse, int* itarg, int* itargse, int* nlist1, int* list1,
                   int* nlist2, int* list2, int* nlist3, int* list3, int* nlist4, int* list4);

void fmm2d_tree_free(void* handle);

/*
 * Leaf occupancy by level of the tree fmm2d_tree_build builds (the reference's
 * pts_tree_build, common/pts_tree2d.f:406, with leaf capacity ndiv), computed
 * on the device: hist[l] (HOST array of nbins <= 64 doubles) = the number of
 * points, sources plus targets, held by the leaves of level l, the last bin
 * taking every deeper level; *boxsize0 (HOST, may be NULL) = the root box
 * size.  No reference counterpart: these are the input features of the
 * parameter proposer (below), which the reference does not have.
 */
int fmm2d_tree_leaf_histogram(const double* sources, int64_t ns, const double* targ, int64_t nt,
                              int idivflag, int ndiv, int nbins, double* hist, double* boxsize0,
                              void* stream);

/*
 * The same histogram for the tree that refines every box holding more than
 * ndiv SOURCES (the refinement rule of pts_tree_build, common/pts_tree2d.f:
 * 551-631, with idivflag 0) without the level restriction of pts_tree_fix_lr,
 * computed from the sorted path keys of the points alone -- no tree is built,
 * which makes it several times cheaper than fmm2d_tree_leaf_histogram.  Levels
 * deeper than 20 count in the last bin (nbins <= 21 is exact).  No reference
 * counterpart: the parameter proposer's input features.
 */
int fmm2d_key_leaf_histogram(const double* sources, int64_t ns, const double* targ, int64_t nt,
                             int ndiv, int nbins, double* hist, double* boxsize0, void* stream);

/*
 * Prepare a tree build: the bounding square of the points and their sorted
 * path keys (the first steps of pts_tree_build, common/pts_tree2d.f:455-551,
 * which do not depend on the leaf capacity), kept on the device under
 * *handle.  A caller that chooses the leaf capacity and expansion orders from
 * the point distribution (fmm2d_prepared_leaf_histogram, then the proposer)
 * passes the handle in fmm2d_opts.prepared so the FMM does not sort the keys
 * again.  The arrays must keep their address and contents until the handle is
 * freed; the FMM uses the handle only when (sources, ns, targ, nt) are the
 * ones prepared.  Equal results with and without it.  No reference
 * counterpart: the reference has no parameter proposer.
 */
int fmm2d_prepare(const double* sources, int64_t ns, const double* targ, int64_t nt,
                  void** handle, void* stream);

void fmm2d_prepare_free(void* handle);

/*
 * fmm2d_key_leaf_histogram from a preparation's keys (no further bounding
 * box or sort).  No reference counterpart (the proposer's features).
 */
int fmm2d_prepared_leaf_histogram(void* handle, int ndiv, int nbins, double* hist,
                                  double* boxsize0, void* stream);

/* Expansion-order and leaf-size rules on the host:
 *   l2dterms  laplace/l2dterms.f:38     h2dterms helmholtz/h2dterms.f:35
 *   lndiv2d   laplace/lndiv2d.f         hndiv2d  helmholtz/hndiv2d.f */
int fmm2d_l2dterms(double eps, int* nterms);
int fmm2d_h2dterms(double bsize, double zk_re, double zk_im, double eps, int* nterms);
int fmm2d_ndiv(double eps, int64_t ns, int64_t nt, int helmholtz, int* ndiv, int* idivflag);

/*
 * Return the FMM drivers' cached device workspace on the current device to
 * the runtime (synchronises the device first).  No reference counterpart:
 * the reference allocates and frees its workspace inside every call
 * (laplace/cfmm2d.f:251-262, helmholtz/hfmm2d.f:354-371); here workspace
 * blocks are cached per (device, stream) across calls and kept until this is
 * called or the process ends.
 */
int fmm2d_release_workspace(void);

/*
 * Cap the device workspace the library holds per device (cached and in use),
 * in bytes; 0 = no cap (the default).  No reference counterpart: the reference
 * returns ier = 4 when a Fortran ALLOCATE of its workspace fails
 * (laplace/cfmm2d.f:251-262, helmholtz/hfmm2d.f:354-371); with a cap, a call
 * whose workspace would exceed it first returns the cached blocks and then
 * fails with that same code, FMM2D_ERR_ALLOC (4), for callers sharing the GPU.
 */
int fmm2d_set_workspace_limit(int64_t bytes);

/*
 * Bound the blocks of the Laplace / Cauchy / Helmholtz direct evaluations
 * (fmm2d_*2d_direct) whose targets are too few to fill the GPU: their source
 * range is split over blocks until about nblocks run (0 = the default, about
 * 2048, 8 per CU).  A bound of a few per XCD lets a sampled check run on a
 * second stream beside an FMM without taking every CU from it.  Returns the
 * previous bound.  No reference counterpart: for callers that overlap direct
 * sums with other work on one GPU.
 */
int fmm2d_set_direct_max_blocks(int nblocks);

/*
 * Parameter-proposer network (no reference counterpart: the reference picks
 * nterms and the leaf size by the fixed rules of laplace/l2dterms.f:38,
 * helmholtz/h2dterms.f:35 and laplace/lndiv2d.f; a proposal reaches the
 * drivers through fmm2d_opts.nterms / ndiv).  A perceptron nin -> 64 -> 64 ->
 * nout with tanh, tanh, linear layers in FP32.  params is a device array of
 * fmm2d_mlp_nparams(nin, nout) floats laid out W1[64][nin], b1[64],
 * W2[64][64], b2[64], W3[nout][64], b3[nout]; x[batch][nin], y[batch][nout].
 */
int64_t fmm2d_mlp_nparams(int nin, int nout);
int fmm2d_mlp_forward(int nin, int nout, const float* params, int batch, const float* x,
                      float* y, void* stream);
/* loss = sum_{b,o} weight[b][o] (y[b][o] - target[b][o])^2 / batch (weight NULL = 1) and
 * its gradient with respect to params (grad: nparams floats); loss: 1 device float. */
int fmm2d_mlp_loss_grad(int nin, int nout, const float* params, int batch, const float* x,
                        const float* target, const float* weight, float* grad, float* loss,
                        void* stream);
/* Adam step (Kingma & Ba 2015) on n parameters; step counts from 1. */
int fmm2d_mlp_adam(int64_t n, float* params, const float* grad, float* m, float* v, float lr,
                   float beta1, float beta2, float eps, int step, void* stream);

/*
 * Biharmonic sums (complex velocity of Stokes flow / elasticity).
 *   vel(z) = sum_j 2 c1_j log|z-z_j| + c2_j (z-z_j)/conj(z-z_j) + d1_j/(z-z_j)
 *            + d3_j/conj(z-z_j) + d2_j (z-z_j)/conj(z-z_j)^2
 * charge(nd,2,ns) holds (c1, c2), dip(nd,3,ns) holds (d1, d2, d3);
 * grad(nd,3,n) = (d/dz vel projected on z, projected on zbar, d/dzbar vel).
 *
 * fmm2d_bhfmm2d replaces biharmonic/bhfmm2d.f:1 bhfmm2d (same arguments, device
 * pointers, OVERWRITES pot/grad/pottarg/gradtarg; hess and hesstarg are not
 * written, as in the reference; iper must be 0).
 */
int fmm2d_bhfmm2d(int nd, double eps, int64_t ns, const double* sources, int ifcharge,
                  const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dip, int iper,
                  int ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, int64_t nt,
                  const double* targ, int ifpghtarg, fmm2d_cplx* pottarg,
                  fmm2d_cplx* gradtarg, fmm2d_cplx* hesstarg, fmm2d_opts* opts, void* stream);

/* bh2d_direct{c,d,cd}{p,g} (biharmonic/bhkernels2d.f:26, 82, 153, 210, 277, 339):
 * INCREMENTS vel(nd,nt) and, for ifpgh = 2, grad(nd,3,nt); pairs with
 * |t-s| <= thresh are skipped. */
int fmm2d_bh2d_direct(int nd, const double* sources, int64_t ns, int ifcharge,
                      const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dip,
                      const double* targ, int64_t nt, int ifpgh, fmm2d_cplx* vel,
                      fmm2d_cplx* grad, double thresh, void* stream);

/*
 * Stokes sums: Stokeslets stoklet(nd,2,ns) and type-I stresslets strslet,
 * strsvec (nd,2,ns); velocity pot(nd,2,n), pressure pre(nd,n), velocity
 * gradient grad(nd,2,2,n) (grad(l,i,j,k): ith component of the gradient of
 * velocity component j).
 *
 * fmm2d_stfmm2d replaces stokes/stfmm2d.f:27 stfmm2d (same arguments, device
 * pointers; ifppreg / ifppregtarg 1 = velocity, 2 = + pressure, 3 = + gradient;
 * OVERWRITES the requested outputs).
 */
int fmm2d_stfmm2d(int nd, double eps, int64_t ns, const double* sources, int ifstoklet,
                  const double* stoklet, int ifstrslet, const double* strslet,
                  const double* strsvec, int ifppreg, double* pot, double* pre, double* grad,
                  int64_t nt, const double* targ, int ifppregtarg, double* pottarg,
                  double* pretarg, double* gradtarg, fmm2d_opts* opts, void* stream);

/* st2ddirectstokg / st2ddirectstokstrsg (stokes/stokkernels2d.f:24, 169):
 * INCREMENTS pot, pre and grad at the targets (each may be NULL, then it is
 * not computed); istress = 1 adds the type-I stresslets; pairs with
 * |t-s| < thresh are skipped. */
int fmm2d_st2d_direct(int nd, const double* sources, int ifstoklet, const double* stoklet,
                      int istress, const double* strslet, const double* strsvec, int64_t ns,
                      const double* targ, int64_t nt, double* pot, double* pre, double* grad,
                      double thresh, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FMM2D_AMD_H */

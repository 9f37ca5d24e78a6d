/*
 * CPU oracle for fmm2d_amd -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's algorithms on the hot path
 * (flatironinstitute/fmm2d), used as the checker of the GPU library.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and never as the thing measured or shipped.
 *
 * Conventions: Fortran layouts (column-major, density index fastest),
 * complex values as interleaved (re, im) doubles, direct kernels INCREMENT
 * their outputs, tree indices returned 1-based like the reference.
 *
 * Pinning (tests/test_oracle.py): every function below is checked against
 * the reference library's outputs (the .npz fixtures in tests/golden, generated from the
 * reference Fortran built by oracle/ref.mk) and against closed forms:
 *   orc_direct    direct_ref.npz (all 27 variants x 3 families), single-source
 *                 closed forms, harmonicity / Helmholtz equation by finite
 *                 differences
 *   orc_bh_direct bh_ref.npz (reference bh2d_direct*, all six variants), the
 *                 bi-Laplacian and the z / zbar derivatives by finite differences
 *   orc_st_direct bh_ref.npz (reference st2ddirect*), divergence-free velocity,
 *                 gradient and Stokes momentum balance by finite differences
 *   orc_hank01    hankel.npz (mpmath values), Wronskian, recurrence
 *   orc_tree_*    tree_ref.npz, bit for bit
 *   orc_l2dterms, orc_lndiv2d   the reference routines (through ref_binding)
 */
#ifndef FMM2D_ORACLE_H
#define FMM2D_ORACLE_H
#include <stdint.h>

/* family: 0 Cauchy (laplace/cauchykernels2d.f), 1 Laplace
 * (laplace/lapkernels2d.f), 2 Helmholtz (helmholtz/helmkernels2d.f). */
void orc_direct(int family, int nd, double zk_re, double zk_im, const double* src, int ns,
                int ifc, const double* charge, int ifd, const double* dipstr,
                const double* dipvec, const double* targ, int nt, int pg, double thresh,
                double* pot, double* grad, double* hess);

/* H0^(1)(z), H1^(1)(z) (replaces common/hank103.f:116) for Im z >= 0. */
void orc_hank01(double zr, double zi, double* h0, double* h1);

/* pts_tree_build + pts_tree_sort (common/pts_tree2d.f:406, 1222) and
 * computelists (common/tree_routs2d.f:486). */
int orc_tree_build(const double* src, int ns, const double* targ, int nt, int idivflag,
                   int ndiv, int nlmax, int* nlevels, int* nboxes, void** handle);
int orc_tree_get(void* handle, int* laddr, int* ilevel, int* iparent, int* nchild, int* ichild,
                 int* ncoll, int* coll, double* centers, double* boxsize, int* isrc,
                 int* isrcse, int* itarg, int* itargse, int* nlist1, int* list1, int* nlist2,
                 int* list2, int* nlist3, int* list3, int* nlist4, int* list4);
void orc_tree_free(void* handle);

/* bh2d_direct{c,d,cd}{p,g} (biharmonic/bhkernels2d.f:26-410): complex
 * charge(nd,2,ns), dip(nd,3,ns), vel(nd,nt), grad(nd,3,nt), incremented. */
void orc_bh_direct(int nd, const double* src, int ns, int ifc, const double* charge, int ifd,
                   const double* dip, const double* targ, int nt, int pg, double thresh,
                   double* vel, double* grad);

/* st2ddirectstokg / st2ddirectstokstrsg (stokes/stokkernels2d.f:24, 169):
 * pot(nd,2,nt), pre(nd,nt), grad(nd,2,2,nt), incremented. */
void orc_st_direct(int nd, const double* src, int ns, int ifs, const double* stoklet, int ift,
                   const double* strslet, const double* strsvec, const double* targ, int nt,
                   double thresh, double* pot, double* pre, double* grad);

int orc_l2dterms(double eps);                          /* laplace/l2dterms.f:38 */
int orc_lndiv2d(double eps, int ns, int nt);           /* laplace/lndiv2d.f */

#endif

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

int orc_l2dterms(double eps);                          /* laplace/l2dterms.f:38 */
int orc_lndiv2d(double eps, int ns, int nt);           /* laplace/lndiv2d.f */

#endif

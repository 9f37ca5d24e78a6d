/*
 * CPU oracle for fmm2d_amd -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Plain-C restatement of the reference's hot-path algorithms.  Each function
 * cites the reference routine it follows (paths relative to the reference's
 * src/).  Serial and straightforward on purpose: it is the checker.
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zc;
typedef long double complex lzc;

/* ======================================================================
 * Hankel functions H0^(1), H1^(1) (replaces common/hank103.f:116-243).
 * Independent of the GPU evaluator: long double arithmetic with
 *   |z| < 2    ascending series for J0, J1, Y0, Y1 (A&S 9.1.10-13);
 *   |z| < 28   J_n by Miller's backward recurrence normalised with the
 *              Jacobi-Anger identity exp(-iz) = J0 + 2 sum (-i)^n J_n, and
 *              Y0, Y1 from the Neumann series
 *                Y0 = (2/pi)(ln(z/2)+g) J0 - (4/pi) sum_k (-1)^k J_2k / k
 *                Y1 = (2/pi)(ln(z/2)+g) J1 - (2/pi) J0/z
 *                     + (2/pi) sum_k (-1)^k (J_2k-1 - J_2k+1) / k
 *              (accurate for real z; loses about exp(2 Im z) otherwise);
 *   |z| >= 28  Hankel's asymptotic expansion (A&S 9.2.7).
 * Domain: Im z >= 0 with moderate Im z (the oracle is used with real zk).
 * ==================================================================== */
static const long double LPI = 3.141592653589793238462643383279503L;
static const long double LGAMMA_E = 0.577215664901532860606512090082402L;

static void hank_series(lzc z, lzc* h0, lzc* h1) {
  lzc q = -z * z / 4.0L, t0 = 1.0L, t1 = 1.0L, J0 = 1.0L, S1s = 1.0L, S0 = 0.0L, S1 = 1.0L;
  long double H = 0.0L;
  for (int k = 1; k <= 40; ++k) {
    t0 *= q / ((long double)k * k);
    t1 *= q / ((long double)k * (k + 1));
    H += 1.0L / k;
    J0 += t0;
    S1s += t1;
    S0 += H * t0;
    S1 += (2.0L * H + 1.0L / (k + 1)) * t1;
  }
  lzc J1 = z / 2.0L * S1s;
  lzc L = clogl(z / 2.0L) + LGAMMA_E;
  lzc Y0 = 2.0L / LPI * (L * J0 - S0);
  lzc Y1 = 2.0L / LPI * L * J1 - 2.0L / (LPI * z) - z / (2.0L * LPI) * S1;
  *h0 = J0 + I * Y0;
  *h1 = J1 + I * Y1;
}

static void hank_miller(lzc z, lzc* h0, lzc* h1) {
  int N = 2 * (int)(cabsl(z) / 2.0L + 30.0L);
  lzc* f = (lzc*)calloc((size_t)N + 2, sizeof(lzc));
  f[N + 1] = 0.0L;
  f[N] = 1e-300L;
  for (int n = N; n >= 1; --n) {
    f[n - 1] = (2.0L * n) / z * f[n] - f[n + 1];
    if (cabsl(f[n - 1]) > 1e300L)
      for (int k = n - 1; k <= N + 1; ++k) f[k] *= 1e-300L;
  }
  lzc psi = f[0], sn = 1.0L;
  for (int n = 1; n <= N; ++n) {
    sn *= -I;
    psi += 2.0L * sn * f[n];
  }
  lzc scale = cexpl(-I * z) / psi;
  for (int n = 0; n <= N; ++n) f[n] *= scale;
  lzc L = clogl(z / 2.0L) + LGAMMA_E;
  lzc s0 = 0.0L, s1 = 0.0L;
  for (int k = 1; 2 * k + 1 <= N; ++k) {
    long double sg = (k & 1) ? -1.0L : 1.0L;
    s0 += sg * f[2 * k] / k;
    s1 += sg * (f[2 * k - 1] - f[2 * k + 1]) / k;
  }
  lzc Y0 = 2.0L / LPI * L * f[0] - 4.0L / LPI * s0;
  lzc Y1 = 2.0L / LPI * L * f[1] - 2.0L / LPI * f[0] / z + 2.0L / LPI * s1;
  *h0 = f[0] + I * Y0;
  *h1 = f[1] + I * Y1;
  free(f);
}

static void hank_asym(lzc z, lzc* h0, lzc* h1) {
  lzc pre = csqrtl(2.0L / (LPI * z)) * cexpl(I * z) *
            cexpl(-I * LPI / 4.0L);
  lzc r[2];
  for (int nu = 0; nu < 2; ++nu) {
    long double mu = 4.0L * nu * nu;
    lzc s = 1.0L, t = 1.0L;
    long double best = 1.0L;
    for (int k = 1; k < 60; ++k) {
      long double m = 2.0L * k - 1.0L;
      t *= (mu - m * m) / (k * 8.0L * z) * I;
      if (cabsl(t) > best) break;
      best = cabsl(t);
      s += t;
      if (best < 1e-20L) break;
    }
    r[nu] = s;
  }
  *h0 = pre * r[0];
  *h1 = pre * cexpl(-I * LPI / 2.0L) * r[1];
}

void orc_hank01(double zr, double zi, double* h0, double* h1) {
  lzc z = (long double)zr + I * (long double)zi, a, b;
  long double r = cabsl(z);
  if (r < 2.0L) hank_series(z, &a, &b);
  else if (r < 28.0L) hank_miller(z, &a, &b);
  else hank_asym(z, &a, &b);
  h0[0] = (double)creall(a);
  h0[1] = (double)cimagl(a);
  h1[0] = (double)creall(b);
  h1[1] = (double)cimagl(b);
}

/* ======================================================================
 * Direct kernels.  Cauchy: laplace/cauchykernels2d.f:3-683; Laplace:
 * laplace/lapkernels2d.f:8-815; Helmholtz: helmholtz/helmkernels2d.f:3-904.
 * Exclusion rules as in each reference variant.
 * ==================================================================== */
static int skip(int family, int ifc, int ifd, int pg, double rr, double thresh, double az) {
  double t2 = thresh * thresh;
  if (family == 0) {
    if (ifc) return pg == 1 ? rr <= t2 : rr < t2;
    return pg == 1 ? az <= thresh : az < thresh;
  }
  if (family == 1) {
    int le = (pg == 1) || !ifc;
    return le ? rr <= t2 : rr < t2;
  }
  {
    int le = (pg == 1) && ifc && !ifd;
    return le ? az <= thresh : az < thresh;
  }
}

void orc_direct(int family, int nd, double zk_re, double zk_im, const double* src, int ns,
                int ifc, const double* charge, int ifd, const double* dipstr,
                const double* dipvec, const double* targ, int nt, int pg, double thresh,
                double* pot, double* grad, double* hess) {
  const zc* ch = (const zc*)charge;
  const zc* ds = (const zc*)dipstr;
  zc* P = (zc*)pot;
  zc* G = (zc*)grad;
  zc* H = (zc*)hess;
  const zc k = zk_re + I * zk_im, ik4 = I / 4.0;
  const int ng = family == 0 ? 1 : 2, nh = family == 0 ? 1 : 3;
  for (int j = 0; j < nt; ++j) {
    for (int i = 0; i < ns; ++i) {
      const double dx = targ[2 * j] - src[2 * i], dy = targ[2 * j + 1] - src[2 * i + 1];
      const double rr = dx * dx + dy * dy;
      const double r = sqrt(rr);
      const zc z = dx + I * dy;
      double az = (family == 2) ? cabs(k * r) : cabs(z);
      if (skip(family, ifc, ifd, pg, rr, thresh, az)) continue;
      for (int d = 0; d < nd; ++d) {
        zc* p = P + (size_t)nd * j + d;
        zc g[3] = {0, 0, 0}, h[3] = {0, 0, 0}, pp = 0;
        const zc c = ifc ? ch[(size_t)nd * i + d] : 0;
        const zc dd = ifd ? ds[(size_t)nd * i + d] : 0;
        double v1 = 0, v2 = 0;
        if (ifd && family != 0) {
          v1 = dipvec[(size_t)nd * (2 * i) + d];
          v2 = dipvec[(size_t)nd * (2 * i + 1) + d];
        }
        if (family == 0) {
          const zc zi = 1.0 / z;
          if (ifc) {
            pp += 0.5 * log(rr) * c;
            g[0] += zi * c;
            h[0] += -zi * zi * c;
          }
          if (ifd) {
            pp += zi * dd;
            g[0] += -zi * zi * dd;
            h[0] += 2.0 * zi * zi * zi * dd;
          }
        } else if (family == 1) {
          const double r2 = rr * rr, r3 = r2 * rr;
          if (ifc) {
            pp += 0.5 * log(rr) * c;
            g[0] += dx / rr * c;
            g[1] += dy / rr * c;
            h[0] += (rr - 2 * dx * dx) / r2 * c;
            h[1] += -2 * dx * dy / r2 * c;
            h[2] += (rr - 2 * dy * dy) / r2 * c;
          }
          if (ifd) {
            const zc d1 = dd * v1, d2 = dd * v2;
            pp += d1 * (-dx / rr) + d2 * (-dy / rr);
            const double dx1 = -(rr - 2 * dx * dx) / r2, dx2 = 2 * dx * dy / r2;
            const double dy2 = -(rr - 2 * dy * dy) / r2;
            g[0] += d1 * dx1 + d2 * dx2;
            g[1] += d1 * dx2 + d2 * dy2;
            const double a = -2 * dx * (dx * dx - 3 * dy * dy) / r3;
            const double b = 2 * dy * (dy * dy - 3 * dx * dx) / r3;
            h[0] += d1 * a + d2 * b;
            h[1] += d1 * b - d2 * a;
            h[2] += -d1 * a - d2 * b;
          }
        } else {
          double h0v[2], h1v[2];
          const zc kr = k * r;
          orc_hank01(creal(kr), cimag(kr), h0v, h1v);
          const zc h0 = h0v[0] + I * h0v[1], h1 = h1v[0] + I * h1v[1];
          if (ifc) {
            pp += h0 * c * ik4;
            const zc cd = -h1 * (k * ik4 / r) * c;
            g[0] += cd * dx;
            g[1] += cd * dy;
            const zc c2 = (k * ik4 / r) / rr * c, h2z = -kr * h0 + 2.0 * h1;
            h[0] += c2 * (h2z * dx * dx - rr * h1);
            h[1] += c2 * (h2z * dx * dy);
            h[2] += c2 * (h2z * dy * dy - rr * h1);
          }
          if (ifd) {
            const double ct = dx / r, st = dy / r;
            const zc h2 = 2.0 * h1 / kr - h0, h3 = 4.0 * h2 / kr - h1;
            pp += h1 / r * k * ik4 * dd * (dx * v1 + dy * v2);
            const zc cd = -k * k * ik4 * dd;
            const zc hxx = cd * (h2 * (ct * ct - 0.5) - h0 / 2.0), hxy = cd * (h2 * ct * st);
            const zc hyy = cd * (h2 * (st * st - 0.5) - h0 / 2.0);
            g[0] += hxx * v1 + hxy * v2;
            g[1] += hxy * v1 + hyy * v2;
            const zc q1 = (-h1 / 2.0 * (-1.5) - h3 / 2.0 * (ct * ct - 0.5 - st * st)) * ct;
            const zc q2 = (-h1 / 2.0 * (-0.5) - h3 / 2.0 * (1.5 * ct * ct - 0.5 * st * st)) * st;
            const zc q3 = (-h1 / 2.0 * (-0.5) - h3 / 2.0 * (1.5 * st * st - 0.5 * ct * ct)) * ct;
            const zc q4 = (-h1 / 2.0 * (-1.5) - h3 / 2.0 * (st * st - 0.5 - ct * ct)) * st;
            const zc c3 = -k * k * k * dd * ik4;
            h[0] += c3 * (q1 * v1 + q2 * v2);
            h[1] += c3 * (q2 * v1 + q3 * v2);
            h[2] += c3 * (q3 * v1 + q4 * v2);
          }
        }
        *p += pp;
        if (pg >= 2)
          for (int cc = 0; cc < ng; ++cc) G[(size_t)nd * (ng * (size_t)j + cc) + d] += g[cc];
        if (pg >= 3)
          for (int cc = 0; cc < nh; ++cc) H[(size_t)nd * (nh * (size_t)j + cc) + d] += h[cc];
      }
    }
  }
}

/* ======================================================================
 * Parameter rules
 * ==================================================================== */
int orc_l2dterms(double eps) { /* laplace/l2dterms.f:38-82 */
  double z2 = sqrt(2.0) / 2.0;
  for (int j = 2; j <= 100; ++j)
    if (cabs(cpow(z2, j) * (1.0 / cpow(1.5, j + 1))) < eps) return j;
  return 1;
}

int orc_lndiv2d(double eps, int ns, int nt) { /* laplace/lndiv2d.f */
  if (eps >= 0.5e-0) return 3;
  if (eps >= 0.5e-1) return 5;
  if (eps >= 0.5e-2) return 8;
  if (eps >= 0.5e-3) return 10;
  if (eps >= 0.5e-6) return 15;
  if (eps >= 0.5e-9) return 20;
  if (eps >= 0.5e-12) return 25;
  if (eps >= 0.5e-15) return 45;
  return ns + nt;
}

/* ======================================================================
 * Quadtree (common/pts_tree2d.f, common/tree_routs2d.f), 0-based inside.
 * ==================================================================== */
typedef struct {
  int nb, cap, nlevels, nlmax;
  double* centers;
  int *ilevel, *iparent, *nchild, *ichild, *nnbors, *nbors;
  int* laddr; /* 2 per level */
  double* boxsize;
  int ns, nt;
  int *isrc, *isrcse, *itarg, *itargse;
  int *nl[4], *ll[4];
} otree;

static const int kCap[4] = {13, 27, 20, 5};

static void grow(otree* t, int need) {
  if (need <= t->cap) return;
  int c = t->cap ? t->cap : 64;
  while (c < need) c *= 2;
  t->centers = realloc(t->centers, sizeof(double) * 2 * c);
  t->ilevel = realloc(t->ilevel, sizeof(int) * c);
  t->iparent = realloc(t->iparent, sizeof(int) * c);
  t->nchild = realloc(t->nchild, sizeof(int) * c);
  t->ichild = realloc(t->ichild, sizeof(int) * 4 * c);
  t->nnbors = realloc(t->nnbors, sizeof(int) * c);
  t->nbors = realloc(t->nbors, sizeof(int) * 9 * c);
  t->cap = c;
}

/* tree_refine_boxes / tree_refine_boxes_flag (tree_routs2d.f:9-56, 334-388):
 * children numbered in parent order, centres parent +- bs/2 */
static void make_children(otree* t, int b, double bs, int lev) {
  grow(t, t->nb + 4);
  t->nchild[b] = 4;
  for (int j = 1; j <= 4; ++j) {
    int c = t->nb++;
    double sx = (j % 2 == 1) ? -1.0 : 1.0, sy = (j <= 2) ? -1.0 : 1.0;
    t->centers[2 * c] = t->centers[2 * b] + sx * bs / 2;
    t->centers[2 * c + 1] = t->centers[2 * b + 1] + sy * bs / 2;
    t->iparent[c] = b;
    t->nchild[c] = 0;
    for (int l = 0; l < 4; ++l) t->ichild[4 * c + l] = -1;
    t->ichild[4 * b + j - 1] = c;
    t->ilevel[c] = lev;
  }
}

/* sort_pts_to_children (pts_tree2d.f:659-754): stable split on y then x */
static void sort_to_children(const otree* t, int b, const double* xy, int* ix, int* se) {
  int s = se[2 * b], e = se[2 * b + 1];
  int n = e - s + 1;
  int* tmp = malloc(sizeof(int) * (n > 0 ? n : 1));
  int i12 = s - 1, i34 = 0;
  for (int k = s; k <= e; ++k) {
    int p = ix[k];
    if (xy[2 * p + 1] - t->centers[2 * b + 1] < 0) ix[++i12] = p;
    else tmp[i34++] = p;
  }
  for (int k = 0; k < i34; ++k) ix[i12 + 1 + k] = tmp[k];
  int nsc[4] = {0, 0, 0, 0};
  for (int k = s; k <= i12; ++k) {
    int p = ix[k];
    if (xy[2 * p] - t->centers[2 * b] < 0) ix[s + nsc[0]++] = p;
    else tmp[nsc[1]++] = p;
  }
  for (int k = 0; k < nsc[1]; ++k) ix[s + nsc[0] + k] = tmp[k];
  for (int k = i12 + 1; k <= e; ++k) {
    int p = ix[k];
    if (xy[2 * p] - t->centers[2 * b] < 0) ix[i12 + 1 + nsc[2]++] = p;
    else tmp[nsc[3]++] = p;
  }
  for (int k = 0; k < nsc[3]; ++k) ix[i12 + 1 + nsc[2] + k] = tmp[k];
  int st = s;
  for (int q = 0; q < 4; ++q) {
    int c = t->ichild[4 * b + q];
    se[2 * c] = st;
    se[2 * c + 1] = st + nsc[q] - 1;
    st += nsc[q];
  }
  free(tmp);
}

/* computecoll (tree_routs2d.f:100-211); 1.05 is a single-precision literal */
static void computecoll(otree* t, const int* laddr, int nlevels) {
  const double c105 = (double)1.05f;
  t->nnbors[0] = 1;
  t->nbors[0] = 0;
  for (int k = 1; k < 9; ++k) t->nbors[k] = -1;
  for (int l = 1; l <= nlevels; ++l)
    for (int b = laddr[2 * l]; b <= laddr[2 * l + 1]; ++b) {
      int dad = t->iparent[b], cnt = 0;
      for (int k = 0; k < 9; ++k) t->nbors[9 * b + k] = -1;
      for (int i = 0; i < t->nnbors[dad]; ++i) {
        int j = t->nbors[9 * dad + i];
        for (int q = 0; q < 4; ++q) {
          int kb = t->ichild[4 * j + q];
          if (kb >= 0 && fabs(t->centers[2 * kb] - t->centers[2 * b]) <= c105 * t->boxsize[l] &&
              fabs(t->centers[2 * kb + 1] - t->centers[2 * b + 1]) <= c105 * t->boxsize[l])
            t->nbors[9 * b + cnt++] = kb;
        }
      }
      t->nnbors[b] = cnt;
    }
}

/* pts_tree_reorg (pts_tree2d.f:1066-1215) */
static void reorg(otree* t, int* laddr, const int* tail, int nlevels, int* iflag) {
  int nb = t->nb;
  int* map = malloc(sizeof(int) * nb);
  double* c0 = malloc(sizeof(double) * 2 * nb);
  int *lev0 = malloc(sizeof(int) * nb), *par0 = malloc(sizeof(int) * nb);
  int *nch0 = malloc(sizeof(int) * nb), *ich0 = malloc(sizeof(int) * 4 * nb);
  int* fl0 = malloc(sizeof(int) * nb);
  memcpy(c0, t->centers, sizeof(double) * 2 * nb);
  memcpy(lev0, t->ilevel, sizeof(int) * nb);
  memcpy(par0, t->iparent, sizeof(int) * nb);
  memcpy(nch0, t->nchild, sizeof(int) * nb);
  memcpy(ich0, t->ichild, sizeof(int) * 4 * nb);
  memcpy(fl0, iflag, sizeof(int) * nb);
  for (int l = 0; l <= 1 && l <= nlevels; ++l)
    for (int b = laddr[2 * l]; b <= laddr[2 * l + 1]; ++b) map[b] = b;
  int cur = laddr[4];
  for (int l = 2; l <= nlevels; ++l) {
    int lo = laddr[2 * l], hi = laddr[2 * l + 1];
    laddr[2 * l] = cur;
    for (int b = lo; b <= hi; ++b) map[b] = cur++;
    for (int b = tail[2 * l]; b <= tail[2 * l + 1]; ++b) map[b] = cur++;
    laddr[2 * l + 1] = cur - 1;
  }
  for (int b = 0; b < nb; ++b) {
    int m = map[b];
    t->centers[2 * m] = c0[2 * b];
    t->centers[2 * m + 1] = c0[2 * b + 1];
    t->ilevel[m] = lev0[b];
    t->nchild[m] = nch0[b];
    iflag[m] = fl0[b];
    t->iparent[m] = par0[b] < 0 ? -1 : map[par0[b]];
    for (int q = 0; q < 4; ++q) t->ichild[4 * m + q] = ich0[4 * b + q] < 0 ? -1 : map[ich0[4 * b + q]];
  }
  free(map); free(c0); free(lev0); free(par0); free(nch0); free(ich0); free(fl0);
}

static int touches(const otree* t, int a, int b, double dist) {
  return fabs(t->centers[2 * a] - t->centers[2 * b]) <= dist &&
         fabs(t->centers[2 * a + 1] - t->centers[2 * b + 1]) <= dist;
}

/* refine the flagged boxes of one range (tree_refine_boxes_flag) */
static void refine_flag(otree* t, int lo, int hi, int** iflag, int lev) {
  for (int b = lo; b <= hi; ++b) {
    if ((*iflag)[b] > 0) {
      int c0 = t->nb;
      make_children(t, b, t->boxsize[lev + 1], lev + 1);
      *iflag = realloc(*iflag, sizeof(int) * t->cap);
      for (int q = 0; q < 4; ++q) {
        (*iflag)[c0 + q] = 0;
        if ((*iflag)[b] == 1) (*iflag)[c0 + q] = 3;
      }
    }
  }
}

/* updateflags (tree_routs2d.f:217-326) */
static void updateflags(otree* t, int lo, int hi, int lev, int* iflag) {
  double distest = 1.05 * (t->boxsize[lev] + t->boxsize[lev + 1]) / 2.0;
  for (int b = lo; b <= hi; ++b) {
    if (iflag[b] != 3) continue;
    iflag[b] = 0;
    for (int i = 0; i < t->nnbors[b] && iflag[b] == 0; ++i) {
      int j = t->nbors[9 * b + i];
      for (int q = 0; q < 4; ++q) {
        int kb = t->ichild[4 * j + q];
        if (kb >= 0 && t->nchild[kb] > 0 && touches(t, kb, b, distest)) {
          iflag[b] = 1;
          break;
        }
      }
    }
  }
}

/* pts_tree_fix_lr (pts_tree2d.f:758-1061) */
static void fix_lr(otree* t, int* laddr) {
  int nl = t->nlevels;
  int* iflag = calloc((size_t)t->cap, sizeof(int));
  for (int l = nl; l >= 2; --l) {
    double distest = 1.05 * (t->boxsize[l - 1] + t->boxsize[l - 2]) / 2.0;
    for (int b = laddr[2 * l]; b <= laddr[2 * l + 1]; ++b) {
      int dad = t->iparent[b], gd = t->iparent[dad];
      for (int i = 0; i < t->nnbors[gd]; ++i) {
        int j = t->nbors[9 * gd + i];
        if (t->nchild[j] == 0 && iflag[j] == 0 && touches(t, j, dad, distest)) iflag[j] = 1;
      }
    }
  }
  for (int l = nl; l >= 1; --l) {
    double distest = 1.05 * (t->boxsize[l] + t->boxsize[l - 1]) / 2.0;
    for (int b = laddr[2 * l]; b <= laddr[2 * l + 1]; ++b) {
      if (iflag[b] != 1 && iflag[b] != 2) continue;
      int dad = t->iparent[b];
      for (int i = 0; i < t->nnbors[dad]; ++i) {
        int j = t->nbors[9 * dad + i];
        if (t->nchild[j] == 0 && iflag[j] == 0 && touches(t, j, b, distest)) iflag[j] = 2;
      }
    }
  }
  int* tail = malloc(sizeof(int) * 2 * (nl + 1));
  for (int l = 0; l <= nl; ++l) { tail[2 * l] = 0; tail[2 * l + 1] = -1; }
  for (int l = 1; l <= nl - 2; ++l) {
    tail[2 * (l + 1)] = t->nb;
    refine_flag(t, laddr[2 * l], laddr[2 * l + 1], &iflag, l);
    tail[2 * (l + 1) + 1] = t->nb - 1;
  }
  reorg(t, laddr, tail, nl, iflag);
  computecoll(t, laddr, nl);
  for (int b = 0; b < t->nb; ++b)
    if (iflag[b] != 3) iflag[b] = 0;
  for (int l = 0; l <= nl; ++l) { tail[2 * l] = 0; tail[2 * l + 1] = -1; }
  const double c105 = (double)1.05f;
  for (int l = 2; l <= nl - 2; ++l) {
    updateflags(t, laddr[2 * l], laddr[2 * l + 1], l, iflag);
    updateflags(t, tail[2 * l], tail[2 * l + 1], l, iflag);
    tail[2 * (l + 1)] = t->nb;
    int lo0 = laddr[2 * l], hi0 = laddr[2 * l + 1], lo1 = tail[2 * l], hi1 = tail[2 * l + 1];
    refine_flag(t, lo0, hi0, &iflag, l);
    refine_flag(t, lo1, hi1, &iflag, l);
    tail[2 * (l + 1) + 1] = t->nb - 1;
    for (int b = tail[2 * (l + 1)]; b <= tail[2 * (l + 1) + 1]; ++b) {
      int dad = t->iparent[b], cnt = 0;
      for (int i = 0; i < t->nnbors[dad]; ++i) {
        int j = t->nbors[9 * dad + i];
        for (int q = 0; q < 4; ++q) {
          int kb = t->ichild[4 * j + q];
          if (kb >= 0 &&
              fabs(t->centers[2 * kb] - t->centers[2 * b]) <= c105 * t->boxsize[l + 1] &&
              fabs(t->centers[2 * kb + 1] - t->centers[2 * b + 1]) <= c105 * t->boxsize[l + 1])
            t->nbors[9 * b + cnt++] = kb;
        }
      }
      t->nnbors[b] = cnt;
      for (int k = cnt; k < 9; ++k) t->nbors[9 * b + k] = -1;
    }
  }
  reorg(t, laddr, tail, nl, iflag);
  computecoll(t, laddr, nl);
  free(tail);
  free(iflag);
}

/* computelists (tree_routs2d.f:486-710), in the reference's order */
static void computelists(otree* t) {
  int nb = t->nb;
  for (int k = 0; k < 4; ++k) {
    t->nl[k] = calloc((size_t)nb, sizeof(int));
    t->ll[k] = calloc((size_t)nb * kCap[k], sizeof(int));
  }
  const double c105 = (double)1.05f;
#define PUSH(K, B, V)                                                     \
  do {                                                                    \
    if (t->nl[K][B] < kCap[K]) t->ll[K][(size_t)kCap[K] * (B) + t->nl[K][B]] = (V); \
    t->nl[K][B]++;                                                        \
  } while (0)
  if (t->nchild[0] == 0) PUSH(0, 0, 0);
  for (int l = 1; l <= t->nlevels; ++l)
    for (int b = t->laddr[2 * l]; b <= t->laddr[2 * l + 1]; ++b) {
      int dad = t->iparent[b];
      for (int i = 0; i < t->nnbors[dad]; ++i) {
        int j = t->nbors[9 * dad + i];
        for (int q = 0; q < t->nchild[j]; ++q) {
          int kb = t->ichild[4 * j + q];
          if (fabs(t->centers[2 * kb] - t->centers[2 * b]) >= c105 * t->boxsize[l] ||
              fabs(t->centers[2 * kb + 1] - t->centers[2 * b + 1]) >= c105 * t->boxsize[l])
            PUSH(1, b, kb);
        }
      }
      if (t->nchild[b] == 0) {
        for (int i = 0; i < t->nnbors[b]; ++i) {
          int j = t->nbors[9 * b + i];
          if (t->nchild[j] == 0) {
            PUSH(0, b, j);
          } else {
            double distest = 1.05 * (t->boxsize[l] + t->boxsize[l + 1]) / 2.0;
            for (int q = 0; q < t->nchild[j]; ++q) {
              int kb = t->ichild[4 * j + q];
              double xd = fabs(t->centers[2 * kb] - t->centers[2 * b]);
              double yd = fabs(t->centers[2 * kb + 1] - t->centers[2 * b + 1]);
              if (xd < distest && yd < distest) {
                PUSH(0, b, kb);
                PUSH(0, kb, b);
              } else {
                PUSH(2, b, kb);
                PUSH(3, kb, b);
              }
            }
          }
        }
      }
    }
#undef PUSH
}

int orc_tree_build(const double* src, int ns, const double* targ, int nt, int idivflag,
                   int ndiv, int nlmax, int* nlevels, int* nboxes, void** handle) {
  otree* t = calloc(1, sizeof(otree));
  t->ns = ns;
  t->nt = nt;
  t->nlmax = nlmax;
  t->laddr = calloc((size_t)2 * (nlmax + 2), sizeof(int));
  t->boxsize = calloc((size_t)nlmax + 2, sizeof(double));
  /* bounding square (pts_tree2d.f:455-517) */
  double xmin = src[0], xmax = src[0], ymin = src[1], ymax = src[1];
  for (int i = 0; i < ns; ++i) {
    if (src[2 * i] < xmin) xmin = src[2 * i];
    if (src[2 * i] > xmax) xmax = src[2 * i];
    if (src[2 * i + 1] < ymin) ymin = src[2 * i + 1];
    if (src[2 * i + 1] > ymax) ymax = src[2 * i + 1];
  }
  for (int i = 0; i < nt; ++i) {
    if (targ[2 * i] < xmin) xmin = targ[2 * i];
    if (targ[2 * i] > xmax) xmax = targ[2 * i];
    if (targ[2 * i + 1] < ymin) ymin = targ[2 * i + 1];
    if (targ[2 * i + 1] > ymax) ymax = targ[2 * i + 1];
  }
  t->boxsize[0] = xmax - xmin;
  if (ymax - ymin > t->boxsize[0]) t->boxsize[0] = ymax - ymin;
  grow(t, 1);
  t->nb = 1;
  t->centers[0] = (xmin + xmax) / 2;
  t->centers[1] = (ymin + ymax) / 2;
  t->ilevel[0] = 0;
  t->iparent[0] = -1;
  t->nchild[0] = 0;
  for (int q = 0; q < 4; ++q) t->ichild[q] = -1;
  /* point ranges as the refinement proceeds (0-based, inclusive ends) */
  int* ix = malloc(sizeof(int) * (ns > 0 ? ns : 1));
  int* it = malloc(sizeof(int) * (nt > 0 ? nt : 1));
  for (int i = 0; i < ns; ++i) ix[i] = i;
  for (int i = 0; i < nt; ++i) it[i] = i;
  int secap = 1024;
  int* ses = malloc(sizeof(int) * 2 * secap);
  int* set = malloc(sizeof(int) * 2 * secap);
  ses[0] = 0; ses[1] = ns - 1; set[0] = 0; set[1] = nt - 1;
  int nl = nlmax;
  for (int l = 0; l < nlmax; ++l) {
    int lo = t->laddr[2 * l], hi = t->laddr[2 * l + 1], any = 0;
    int nb0 = t->nb;
    for (int b = lo; b <= hi; ++b) {
      int nss = ses[2 * b + 1] - ses[2 * b] + 1, ntt = set[2 * b + 1] - set[2 * b] + 1;
      int nn = idivflag == 0 ? nss : (idivflag == 1 ? ntt : (nss > ntt ? nss : ntt));
      if (nn > ndiv) any = 1;
    }
    if (!any) { nl = l; break; }
    t->boxsize[l + 1] = t->boxsize[l] / 2;
    t->laddr[2 * (l + 1)] = t->nb;
    for (int b = lo; b <= hi; ++b) {
      int nss = ses[2 * b + 1] - ses[2 * b] + 1, ntt = set[2 * b + 1] - set[2 * b] + 1;
      int nn = idivflag == 0 ? nss : (idivflag == 1 ? ntt : (nss > ntt ? nss : ntt));
      if (nn <= ndiv) continue;
      make_children(t, b, t->boxsize[l + 1], l + 1);
      if (t->nb > secap) {
        while (secap < t->nb) secap *= 2;
        ses = realloc(ses, sizeof(int) * 2 * secap);
        set = realloc(set, sizeof(int) * 2 * secap);
      }
      sort_to_children(t, b, src, ix, ses);
      sort_to_children(t, b, targ, it, set);
    }
    t->laddr[2 * (l + 1) + 1] = t->nb - 1;
    (void)nb0;
  }
  t->nlevels = nl;
  free(ix); free(it); free(ses); free(set);
  for (int b = 0; b < t->nb; ++b) t->nnbors[b] = 0;
  computecoll(t, t->laddr, nl);
  if (nl >= 2) fix_lr(t, t->laddr);
  /* pts_tree_sort on the final tree (pts_tree2d.f:1222-1246) */
  int nb = t->nb;
  t->isrc = malloc(sizeof(int) * (ns > 0 ? ns : 1));
  t->itarg = malloc(sizeof(int) * (nt > 0 ? nt : 1));
  t->isrcse = malloc(sizeof(int) * 2 * nb);
  t->itargse = malloc(sizeof(int) * 2 * nb);
  for (int k = 0; k < 2; ++k) {
    int n = k ? nt : ns;
    int* ixx = k ? t->itarg : t->isrc;
    int* se = k ? t->itargse : t->isrcse;
    const double* xy = k ? targ : src;
    for (int i = 0; i < n; ++i) ixx[i] = i;
    se[0] = 0;
    se[1] = n - 1;
    for (int l = 0; l < nl; ++l)
      for (int b = t->laddr[2 * l]; b <= t->laddr[2 * l + 1]; ++b)
        if (t->nchild[b] > 0) sort_to_children(t, b, xy, ixx, se);
  }
  computelists(t);
  *nlevels = nl;
  *nboxes = nb;
  *handle = t;
  return 0;
}

int orc_tree_get(void* handle, int* laddr, int* ilevel, int* iparent, int* nchild, int* ichild,
                 int* ncoll, int* coll, double* centers, double* boxsize, int* isrc,
                 int* isrcse, int* itarg, int* itargse, int* nlist1, int* list1, int* nlist2,
                 int* list2, int* nlist3, int* list3, int* nlist4, int* list4) {
  otree* t = handle;
  int nb = t->nb, nl = t->nlevels;
#define P1(v) ((v) < 0 ? -1 : (v) + 1)
  for (int l = 0; l <= nl; ++l) {
    if (laddr) { laddr[2 * l] = t->laddr[2 * l] + 1; laddr[2 * l + 1] = t->laddr[2 * l + 1] + 1; }
    if (boxsize) boxsize[l] = t->boxsize[l];
  }
  for (int b = 0; b < nb; ++b) {
    if (ilevel) ilevel[b] = t->ilevel[b];
    if (iparent) iparent[b] = P1(t->iparent[b]);
    if (nchild) nchild[b] = t->nchild[b];
    if (ichild) for (int q = 0; q < 4; ++q) ichild[4 * b + q] = P1(t->ichild[4 * b + q]);
    if (ncoll) ncoll[b] = t->nnbors[b];
    if (coll) for (int q = 0; q < 9; ++q) coll[9 * b + q] = q < t->nnbors[b] ? P1(t->nbors[9 * b + q]) : -1;
    if (centers) { centers[2 * b] = t->centers[2 * b]; centers[2 * b + 1] = t->centers[2 * b + 1]; }
    if (isrcse) { isrcse[2 * b] = t->isrcse[2 * b] + 1; isrcse[2 * b + 1] = t->isrcse[2 * b + 1] + 1; }
    if (itargse) { itargse[2 * b] = t->itargse[2 * b] + 1; itargse[2 * b + 1] = t->itargse[2 * b + 1] + 1; }
  }
  if (isrc) for (int i = 0; i < t->ns; ++i) isrc[i] = t->isrc[i] + 1;
  if (itarg) for (int i = 0; i < t->nt; ++i) itarg[i] = t->itarg[i] + 1;
  int* ns_[4] = {nlist1, nlist2, nlist3, nlist4};
  int* ls_[4] = {list1, list2, list3, list4};
  for (int k = 0; k < 4; ++k) {
    if (!ns_[k] || !ls_[k]) continue;
    for (int b = 0; b < nb; ++b) {
      int n = t->nl[k][b] < kCap[k] ? t->nl[k][b] : kCap[k];
      ns_[k][b] = t->nl[k][b];
      for (int j = 0; j < kCap[k]; ++j)
        ls_[k][(size_t)kCap[k] * b + j] = j < n ? t->ll[k][(size_t)kCap[k] * b + j] + 1 : 0;
    }
  }
#undef P1
  return 0;
}

void orc_tree_free(void* handle) {
  otree* t = handle;
  if (!t) return;
  free(t->centers); free(t->ilevel); free(t->iparent); free(t->nchild); free(t->ichild);
  free(t->nnbors); free(t->nbors); free(t->laddr); free(t->boxsize);
  free(t->isrc); free(t->isrcse); free(t->itarg); free(t->itargse);
  for (int k = 0; k < 4; ++k) { free(t->nl[k]); free(t->ll[k]); }
  free(t);
}

/* ======================================================================
 * Biharmonic direct sums (biharmonic/bhkernels2d.f:26-410, bh2d_direct*):
 * complex velocity vel(nd,nt) and its three derivatives grad(nd,3,nt),
 * incremented; pairs with |t - s| <= thresh skipped.  charge(nd,2,ns) and
 * dip(nd,3,ns) are complex (interleaved doubles).
 * ==================================================================== */
void orc_bh_direct(int nd, const double* src, int ns, int ifc, const double* charge, int ifd,
                   const double* dip, const double* targ, int nt, int pg, double thresh,
                   double* vel, double* grad) {
  const zc* c = (const zc*)charge;
  const zc* d = (const zc*)dip;
  zc* v = (zc*)vel;
  zc* g = (zc*)grad;
  for (int j = 0; j < nt; ++j) {
    const zc zt = targ[2 * j] + I * targ[2 * j + 1];
    for (int i = 0; i < ns; ++i) {
      const zc z = zt - (src[2 * i] + I * src[2 * i + 1]);
      if (cabs(z) <= thresh) continue;
      const zc z1 = 1.0 / z, z2 = z1 * z1;
      for (int k = 0; k < nd; ++k) {
        zc dv = 0, g1 = 0, g2 = 0, g3 = 0;
        if (ifc) {
          const zc c1 = c[k + nd * (2 * i)], c2 = c[k + nd * (2 * i + 1)];
          dv += 2.0 * c1 * log(cabs(z)) + c2 * conj(z1) * z;
          g1 += c1 * z1;
          g2 += c2 * conj(z1);
          g3 += c1 * conj(z1) - c2 * conj(z2) * z;
        }
        if (ifd) {
          const zc d1 = d[k + nd * (3 * i)], d2 = d[k + nd * (3 * i + 1)],
                   d3 = d[k + nd * (3 * i + 2)];
          dv += d1 * z1 + d3 * conj(z1) + d2 * conj(z2) * z;
          g1 -= d1 * z2;
          g2 += d2 * conj(z2);
          g3 -= d3 * conj(z2) + 2.0 * d2 * conj(z2 * z1) * z;
        }
        v[k + nd * j] += dv;
        if (pg >= 2) {
          g[k + nd * (3 * j)] += g1;
          g[k + nd * (3 * j + 1)] += g2;
          g[k + nd * (3 * j + 2)] += g3;
        }
      }
    }
  }
}

/* ======================================================================
 * Stokes direct sums (stokes/stokkernels2d.f:24 st2ddirectstokg, :169
 * st2ddirectstokstrsg): velocity pot(nd,2,nt), pressure pre(nd,nt) and
 * velocity gradient grad(nd,2,2,nt) (grad(l,i,j,k) = d_i u_j), incremented;
 * pairs with |t - s|^2 < thresh^2 skipped.  With r = t - s:
 *   Stokeslet s:  u_i = -s_i log(r^2)/4 + r_i (r.s)/(2 r^2),  p = (r.s)/r^2
 *   stresslet (mu, nu): u_i = -2 r_i (r.mu)(r.nu)/r^4,
 *                       p = -4 (r.mu)(r.nu)/r^4 + 2 (mu.nu)/r^2
 * and the gradients as the reference writes them out.
 * ==================================================================== */
void orc_st_direct(int nd, const double* src, int ns, int ifs, const double* stoklet, int ift,
                   const double* strslet, const double* strsvec, const double* targ, int nt,
                   double thresh, double* pot, double* pre, double* grad) {
  const double th2 = thresh * thresh;
  for (int j = 0; j < nt; ++j) {
    for (int i = 0; i < ns; ++i) {
      const double x = targ[2 * j] - src[2 * i], y = targ[2 * j + 1] - src[2 * i + 1];
      const double r2 = x * x + y * y;
      if (r2 < th2) continue;
      const double r4 = r2 * r2, r6 = r4 * r2, rt = x * x - y * y;
      for (int k = 0; k < nd; ++k) {
        double* u = pot + k + nd * 2 * j;      /* u[0], u[nd] */
        double* gr = grad + k + nd * 4 * j;    /* gr[nd*(a + 2 b)] = d_a u_b */
        if (ifs) {
          const double s1 = stoklet[k + nd * (2 * i)], s2 = stoklet[k + nd * (2 * i + 1)];
          const double pl = x * s1 + y * s2, pl2 = x * s2 + y * s1;
          u[0] += -s1 * log(r2) / 4 + x * pl / r2 / 2;
          u[nd] += -s2 * log(r2) / 4 + y * pl / r2 / 2;
          gr[0] += -rt * pl / r4 / 2;
          gr[nd * 3] += rt * pl / r4 / 2;
          gr[nd * 1] += (pl2 * rt - s1 * 4 * y * x * x) / 2 / r4;
          gr[nd * 2] += -(pl2 * rt + s2 * 4 * x * y * y) / 2 / r4;
          pre[k + nd * j] += pl / r2;
        }
        if (ift) {
          const double m1 = strslet[k + nd * (2 * i)], m2 = strslet[k + nd * (2 * i + 1)];
          const double n1 = strsvec[k + nd * (2 * i)], n2 = strsvec[k + nd * (2 * i + 1)];
          const double pl = x * m1 + y * m2, pl2 = x * n1 + y * n2, pv = m1 * n1 + m2 * n2;
          u[0] += -2 * x * pl * pl2 / r4;
          u[nd] += -2 * y * pl * pl2 / r4;
          gr[0] += -2 * pl * pl2 / r4 - 2 * x * m1 * pl2 / r4 - 2 * x * n1 * pl / r4 +
                   8 * x * x * pl * pl2 / r6;
          gr[nd * 1] += -2 * x * m2 * pl2 / r4 - 2 * x * n2 * pl / r4 + 8 * x * y * pl * pl2 / r6;
          gr[nd * 2] += -2 * y * m1 * pl2 / r4 - 2 * y * n1 * pl / r4 + 8 * x * y * pl * pl2 / r6;
          gr[nd * 3] += -2 * pl * pl2 / r4 - 2 * y * m2 * pl2 / r4 - 2 * y * n2 * pl / r4 +
                        8 * y * y * pl * pl2 / r6;
          pre[k + nd * j] += -4 * pl * pl2 / r4 + 2 * pv / r2;
        }
      }
    }
  }
}

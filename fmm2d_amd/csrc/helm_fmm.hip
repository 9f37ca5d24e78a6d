// Helmholtz FMM on the device (FP64).
//
// Reference: helmholtz/hfmm2d.f (driver 19-420, hfmm2dmain 456-1620) and
// helmholtz/helmrouts2d.f (expansions and translations).  Expansions are the
// reference's scaled partial-wave series about each box centre c,
//   multipole  u(x) = (i/4) sum_n M_n s^|n| H_n(k r) e^{i n t}
//   local      u(x) = (i/4) sum_n L_n s^-|n| J_n(k r) e^{i n t},
// (r, t) the polar form of x - c, n = -p..p, s = min(|k boxsize / 2 pi|, 1)
// the level's rscale and p = nterms(level) from h2dterms.
//
// Storage: level l keeps its expansions in its own block of the multipole
// and local buffers, P_l = 2 p_l + 1 complex coefficients per (box, density),
// coefficient n at position p_l + n, so deep levels with short expansions
// do not pay for the long expansions of coarse levels.
//
// Translations (M2M, M2L, L2L) are the truncated Graf addition theorem,
//   T[i][j] = F_{i-j}(k d) e^{-i (i-j) a} (scaled), F = J (M2M, L2L) or H (M2L),
// (d, a) the polar form of c_source - c_target.  This is the operator the
// reference applies directly at low frequency (h2dmpmp, h2dmploc, h2dlocloc)
// and, through FFT-based diagonal forms, at high frequency (h2dmpmphf,
// h2d_diagtrans, h2dloclochf): with nsig >= 2 (p1 + p2) + 1 samples the
// diagonal forms are exact convolutions of the same truncated series, so
// dense operators reproduce both regimes.  They are built on the device from
// per-level Bessel/Hankel tables (host, long double) and applied with the
// shared MFMA translate kernel.
//
// Cylinder functions at points: J by Miller's algorithm written as a
// backward continued fraction for J_n/J_{n-1} (no overflow, no storage)
// with the sums evaluated by Horner's rule in the same backward sweep and
// normalised by exp(-i z) = J_0 + 2 sum (-i)^n J_n; H by the upward
// recurrence from H_0, H_1 (hankel.h), which is stable for H.
// Derivatives use C'_n relations: d/dx (C_n e^{int}) = (k/2)(C_{n-1}e^{i(n-1)t}
// - C_{n+1}e^{i(n+1)t}), d/dy = (ik/2)(... + ...), for C = J or H.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "devbuf.h"
#include "fmm_common.h"
#include "hankel32.h"
#include "helm_ops.h"
#include "lap_ops.h"
#include "p2p_math.h"
#include "scan.h"
#include "translate.h"
#include "tree.h"

namespace fmm2d {

namespace {

// Per-level expansion layout and parameters.
struct HLev {
  int64_t off;  // first coefficient of the level in the expansion buffers
  int p;        // nterms
  int box0;     // first box of the level
  double s;     // rscale
  int on;       // level takes part: Im(zk) boxsize < 40 (hfmm2d.f:711, 796)
  int pad;
};

constexpr double kZkiUpBound = 40.0;

__device__ __forceinline__ cplx* hexp(cplx* buf, const HLev& L, int b, int nd, int d) {
  return buf + L.off + ((int64_t)(b - L.box0) * nd + d) * (2 * L.p + 1) + L.p;
}
__device__ __forceinline__ const cplx* hexp(const cplx* buf, const HLev& L, int b, int nd, int d) {
  return buf + L.off + ((int64_t)(b - L.box0) * nd + d) * (2 * L.p + 1) + L.p;
}

// First index of Miller's backward recurrence for J_0..J_nmin(z): the
// three-term recurrence run forward from (0, 1) at nmin has grown by 1e20.
__device__ int miller_top(cplx z, int nmin) {
  if (abs2(z) == 0.0) return nmin + 1;
  const cplx zi = invc(z);
  cplx f0 = mkc(0.0), f1 = mkc(1.0);
  const int cap = nmin + 2 * (int)absc(z) + 1000;
  int i = nmin;
  for (; i < cap; ++i) {
    const cplx f2 = (2.0 * i) * (zi * f1) - f0;
    f0 = f1;
    f1 = f2;
    if (abs2(f2) > 1e40) break;
  }
  return i + 1;
}

// rho_n = F_n / F_{n-1} for the scaled F_n = J_n(z) / s^n from rho_{n+1}
// (F_{n-1} = (2 n s / z) F_n - s^2 F_{n+1}).  At a zero of J_{n-1} the
// denominator vanishes; it is moved off zero by 1e-30 of its leading term,
// which changes the computed F_{n-1} (zero) by that relative amount only.
__device__ __forceinline__ cplx miller_ratio(cplx zi, double s, int n, cplx rho1) {
  if (zi.re == 0.0 && zi.im == 0.0) return mkc(0.0);
  const cplx a = (2.0 * n * s) * zi;
  cplx den = a - (s * s) * rho1;
  const double a2 = abs2(a);
  if (abs2(den) < 1e-60 * a2) den.re += 1e-30 * sqrt(a2);
  return invc(den);
}

// Normalisation: exp(-i z) = F_0 + 2 sum_n (-i s)^n F_n for Im z >= 0 (the
// sum does not cancel there), exp(+i z) with (+i s)^n below the axis.
__device__ __forceinline__ cplx miller_unit(cplx z, double s) {
  return (z.im < 0.0) ? cplx{0.0, s} : cplx{0.0, -s};
}
__device__ __forceinline__ cplx miller_exp(cplx z) {
  return (z.im < 0.0) ? expc(cplx{-z.im, z.re}) : expc(cplx{z.im, -z.re});
}

// J_0..J_{m}(z) / s^n into out[0..m] (m >= 1), one thread.
__device__ void scaled_j(cplx z, double s, int m, cplx* out) {
  const int ntop = miller_top(z, m);
  const bool zero = (abs2(z) == 0.0);
  const cplx zi = zero ? mkc(0.0) : invc(z);
  const cplx us = miller_unit(z, s);
  cplx rho = mkc(0.0), U = mkc(2.0);
  for (int n = ntop; n >= 1; --n) {
    rho = miller_ratio(zi, s, n, rho);
    if (n <= m) out[n] = rho;
    U = mkc(n - 1 >= 1 ? 2.0 : 1.0) + us * (rho * U);
  }
  cplx F = miller_exp(z) / U;
  out[0] = F;
  for (int n = 1; n <= m; ++n) {
    F = F * out[n];
    out[n] = F;
  }
}

// H_0..H_m(z) s^n into out[0..m] (m >= 1), one thread.
__device__ void scaled_h(cplx z, double s, int m, cplx* out) {
  cplx h0, h1;
  hank01_ool(z, &h0, &h1);
  out[0] = h0;
  h1 = h1 * s;
  out[1] = h1;
  const cplx zi = invc(z);
  for (int n = 1; n < m; ++n) {
    const cplx h2 = ((2.0 * n * s) * zi) * h1 - (s * s) * h0;
    out[n + 1] = h2;
    h0 = h1;
    h1 = h2;
  }
}

// ---------------------------------------------------------- gathers
__global__ void k_gather_dipvec(int64_t n, int nd, const int* perm, const double* in,
                                double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t j = perm[i];
  for (int k = 0; k < 2 * nd; ++k) out[i * 2 * nd + k] = in[j * 2 * nd + k];
}

// ---------------------------------------------------------- P2M / P2L
// The cylinder functions of each source (J/s^n about its leaf's centre for
// P2M, H s^n about the target box's centre for P2L) are computed first, one
// thread per source or (box, source) pair, into rows G; then one 64-lane
// wavefront per box sums its rows, lane n adding to coefficients +-n:
//   charge  C_n += q G_n e^{-i n a},  C_{-n} += q (-1)^n G_n e^{i n a}
//   dipole  C_n += q (k/2) [(v1 - i v2) g_{n-1} f_lo - (v1 + i v2) g_{n+1} f_hi]
//           (g_j = G_|j| sign(j)^j e^{-i j a}; f_lo, f_hi powers of s from the
//            scaling, swapped for negative n)
// (h2dformmp{c,d,cd}, helmrouts2d.f:84-270 with ctompole/dtompole 1383-1490;
//  h2dformta{c,d,cd}, helmrouts2d.f:517-705).
template <int IFC, int IFD, int LOCAL>
__device__ void accumulate_batch(int nd, int p, cplx zk, double s, const cplx* G, const cplx* U,
                                 int nbat, int stride, int j0, const cplx* chs,
                                 const cplx* dps, const double* dvs, cplx* out_p0) {
  // LOCAL: (a_lo, a_hi) = (s, 1/s); multipole: (1/s, s)
  const double a_lo = LOCAL ? s : 1.0 / s, a_hi = LOCAL ? 1.0 / s : s;
  const cplx k2 = 0.5 * zk;
  for (int n = threadIdx.x; n <= p; n += kW) {
    const double sg = (n & 1) ? -1.0 : 1.0;
    for (int d = 0; d < nd; ++d) {
      cplx ap = mkc(0.0), am = mkc(0.0);
      for (int jj = 0; jj < nbat; ++jj) {
        const cplx* g = G + (int64_t)jj * stride;
        const cplx u = U[jj];  // e^{-i a}
        const cplx en = cpowi(u, n), ep = conjc(en);
        const int64_t j = j0 + jj;
        if (IFC) {
          const cplx q = chs[j * nd + d];
          const cplx t = q * g[n];
          ap += t * en;
          if (n > 0) am += sg * (t * ep);
        }
        if (IFD) {
          const cplx q = k2 * dps[j * nd + d];
          const double v1 = dvs[(2 * j) * nd + d], v2 = dvs[(2 * j + 1) * nd + d];
          const cplx al = cplx{v1, -v2}, be = cplx{v1, v2};
          const cplx ub = conjc(u);
          if (n == 0) {
            // g_{-1} = -G_1 e^{i a}, g_1 = G_1 e^{-i a}
            ap += (q * a_hi) * (-(al * (g[1] * ub)) - be * (g[1] * u));
          } else {
            // +n: g_{n-1} = G_{n-1} e^{-i(n-1)a}, g_{n+1} = G_{n+1} e^{-i(n+1)a}
            ap += q * ((a_lo * (al * (g[n - 1] * (en * ub)))) -
                       (a_hi * (be * (g[n + 1] * (en * u)))));
            // -n: g_{-n-1} = -sg G_{n+1} e^{i(n+1)a}, g_{-n+1} = -sg G_{n-1} e^{i(n-1)a}
            am += q * ((a_hi * (al * ((-sg) * (g[n + 1] * (ep * ub))))) -
                       (a_lo * (be * ((-sg) * (g[n - 1] * (ep * u))))));
          }
        }
      }
      cplx* c = out_p0 + (int64_t)d * (2 * p + 1);
      c[n] += ap;
      if (n > 0) c[-n] += am;
    }
  }
}

// P2M in two steps.  First, one thread per source: the row of scaled J_n(k r)
// / s^n, n = 0..p+1, about the centre of its leaf (its level's p and s) and
// e^{-i a}, into G and U; the rows depend only on the source, and a thread per
// source keeps every lane busy (a wavefront per leaf had one lane per source of
// a leaf of ~10).  Row i starts at G + roff[i] and holds p+2 values of its
// leaf's level (roff = exclusive scan of k_hp2m_rowlen), so G holds exactly
// the rows that are used, whatever the orders of the coarse levels.
__global__ void k_hp2m_rowlen(int64_t n, const HLev* lev, const int* leaf, const int* ilevel,
                              int64_t* len) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int l = ilevel[leaf[i]];
  len[i] = (l >= 2 && lev[l].on) ? lev[l].p + 2 : 0;
}

__global__ void k_hp2m_rows(int64_t n, cplx zk, const HLev* lev, const int* leaf,
                            const int* ilevel, const double* centers, const double* xs,
                            const int64_t* roff, cplx* G, cplx* U) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = leaf[i];
  const int l = ilevel[b];
  const HLev L = lev[l];
  if (l < 2 || !L.on) return;
  const double dx = xs[2 * i] - centers[2 * b], dy = xs[2 * i + 1] - centers[2 * b + 1];
  const double r = hypot(dx, dy);
  scaled_j(zk * r, L.s, L.p + 1, G + roff[i]);
  U[i] = (r > 0.0) ? cplx{dx / r, -dy / r} : mkc(1.0);
}

// Then one wavefront per leaf sums its sources' rows into the multipole
// (lane n: coefficients n and -n).
template <int IFC, int IFD>
__global__ __launch_bounds__(kW) void k_hp2m_g(int nd, cplx zk, const HLev* lev,
                                               const int* ilevel, const int* nchild,
                                               const int* sse, const int64_t* roff,
                                               const cplx* G, const cplx* U, const cplx* chs,
                                               const cplx* dps, const double* dvs, cplx* mp) {
  const int b = blockIdx.x;
  const int l = ilevel[b];
  const HLev L = lev[l];
  if (l < 2 || !L.on || nchild[b] != 0) return;
  const int s0 = sse[2 * b], s1 = sse[2 * b + 1];
  if (s1 <= s0) return;
  accumulate_batch<IFC, IFD, 0>(nd, L.p, zk, L.s, G + roff[s0], U + s0, s1 - s0, L.p + 2, s0,
                                chs, dps, dvs, hexp(mp, L, b, nd, 0));
}

// List 4 in three steps, like P2M.  cnt[b] = sources of the list-4 boxes of
// a box b that takes list-4 work, 0 otherwise.
__global__ void k_hp2l_count(int64_t nb, const HLev* lev, const int* ilevel, const int* n4,
                             const int* l4, const int* sse, const int* tse, int ifpgh,
                             int ifpght, int* cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int b = (int)i;
  const int l = ilevel[b];
  int s = 0;
  if (l >= 2 && lev[l].on && npts_eval(b, sse, tse, ifpgh, ifpght) > 0)
    for (int k = 0; k < n4[b]; ++k) {
      const int c = l4[(int64_t)kMaxList4 * b + k];
      s += sse[2 * c + 1] - sse[2 * c];
    }
  cnt[b] = s;
}

// Row storage of a box's pairs: cnt[b] rows of p+2 values of b's level.
__global__ void k_hp2l_rowlen(int64_t nb, const HLev* lev, const int* ilevel, const int* cnt,
                              int64_t* len) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  len[b] = (int64_t)cnt[b] * (lev[ilevel[b]].p + 2);
}

// One thread per (box b, list-4 source j) pair t, pairs numbered box after
// box (off = exclusive scan of cnt) and in list order within a box: the row
// of H_n(k r) s^n, n = 0..p+1, about b's centre and e^{-i a}, into G (b's rows
// from G + roff[b], p+2 values each) and U.
__global__ void k_hp2l_rows(int64_t np, int nb, cplx zk, const HLev* lev, const int* ilevel,
                            const double* centers, const int* l4, const int* sse,
                            const int* off, const double* xs, const int64_t* roff, cplx* G,
                            cplx* U) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np) return;
  int lo = 0, hi = nb - 1;  // last box with off[b] <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const int b = lo;
  int r = (int)(t - off[b]), j = 0;
  for (int k = 0; k < kMaxList4; ++k) {
    const int c = l4[(int64_t)kMaxList4 * b + k];
    const int sz = sse[2 * c + 1] - sse[2 * c];
    if (r < sz) {
      j = sse[2 * c] + r;
      break;
    }
    r -= sz;
  }
  const HLev L = lev[ilevel[b]];
  const double dx = xs[2 * j] - centers[2 * b], dy = xs[2 * j + 1] - centers[2 * b + 1];
  const double rr = hypot(dx, dy);
  scaled_h(zk * rr, L.s, L.p + 1, G + roff[b] + (t - off[b]) * (L.p + 2));
  U[t] = cplx{dx / rr, -dy / rr};
}

// One wavefront per box with list-4 work sums its pairs' rows into its local
// expansion, list-4 box after list-4 box (their sources are contiguous).
template <int IFC, int IFD>
__global__ __launch_bounds__(kW) void k_hp2l_g(int nd, cplx zk, const HLev* lev,
                                               const int* ilevel, const int* n4, const int* l4,
                                               const int* sse, const int* cnt, const int* off,
                                               const int64_t* roff, const cplx* G,
                                               const cplx* U, const cplx* chs, const cplx* dps,
                                               const double* dvs, cplx* loc) {
  const int b = blockIdx.x;
  if (cnt[b] == 0) return;
  const HLev L = lev[ilevel[b]];
  const int pstride = L.p + 2;
  cplx* out = hexp(loc, L, b, nd, 0);
  const int64_t pos0 = off[b];
  int64_t pos = pos0;
  const cplx* Gb = G + roff[b];  // row of pair pos at Gb + (pos - pos0) * pstride
  for (int k = 0; k < n4[b]; ++k) {
    const int c = l4[(int64_t)kMaxList4 * b + k];
    const int c0 = sse[2 * c], c1 = sse[2 * c + 1];
    if (c1 <= c0) continue;
    accumulate_batch<IFC, IFD, 1>(nd, L.p, zk, L.s, Gb + (pos - pos0) * pstride, U + pos, c1 - c0,
                                  pstride, c0, chs, dps, dvs, out);
    pos += c1 - c0;
  }
}

// ---------------------------------------------------------- evaluation
// Coefficients of the potential and its derivatives in the basis
// b_m = C_m e^{i m t} (scaled as the expansion), for m = -p-2 .. p+2.
// SIG = +1 for local expansions (J, coefficients scaled by s^|m|), -1 for
// multipoles (H, scaled by s^-|m|): a coefficient moved from index m' to m
// picks up s^{SIG (|m| - |m'|)}.
struct Deriv {
  const cplx* c;  // c[-p..p]
  int p;
  cplx k2;        // k/2
  double s, si;   // s, 1/s
  int sig;
  __device__ __forceinline__ cplx at(int m) const {
    return (m >= -p && m <= p) ? c[m] : mkc(0.0);
  }
  __device__ __forceinline__ double f(int m, int mp) const {
    const int e = sig * (abs(m) - abs(mp));
    return e == 0 ? 1.0 : e == 1 ? s : e == -1 ? si : e == 2 ? s * s : si * si;
  }
  // d/dx, d/dy coefficient at m
  __device__ __forceinline__ void grad(int m, cplx& gx, cplx& gy) const {
    const cplx up = f(m, m + 1) * at(m + 1), dn = f(m, m - 1) * at(m - 1);
    gx = k2 * (up - dn);
    gy = cplx{-k2.im, k2.re} * (up + dn);
  }
  __device__ __forceinline__ void hess(int m, cplx& hxx, cplx& hxy, cplx& hyy) const {
    const cplx kk = k2 * k2;
    const cplx up = f(m, m + 2) * at(m + 2), dn = f(m, m - 2) * at(m - 2), c0 = at(m);
    hxx = kk * (up - 2.0 * c0 + dn);
    hxy = cplx{-kk.im, kk.re} * (up - dn);
    hyy = -(kk * (up + 2.0 * c0 + dn));
  }
};

// Accumulated combinations at index n >= 0: X_n e^{int} + (-1)^n X_{-n} e^{-int}
template <int PG> struct Combo {
  cplx v[6];
  __device__ __forceinline__ void at(const Deriv& D, int n, cplx e) {
    const cplx eb = conjc(e);
    const double sg = (n & 1) ? -1.0 : 1.0;
    if (n == 0) {
      v[0] = D.at(0);
      if (PG >= 2) D.grad(0, v[1], v[2]);
      if (PG >= 3) D.hess(0, v[3], v[4], v[5]);
      return;
    }
    v[0] = D.at(n) * e + sg * (D.at(-n) * eb);
    if (PG >= 2) {
      cplx ax, ay, bx, by;
      D.grad(n, ax, ay);
      D.grad(-n, bx, by);
      v[1] = ax * e + sg * (bx * eb);
      v[2] = ay * e + sg * (by * eb);
    }
    if (PG >= 3) {
      cplx a0, a1, a2, b0, b1, b2;
      D.hess(n, a0, a1, a2);
      D.hess(-n, b0, b1, b2);
      v[3] = a0 * e + sg * (b0 * eb);
      v[4] = a1 * e + sg * (b1 * eb);
      v[5] = a2 * e + sg * (b2 * eb);
    }
  }
};

template <int PG> struct NComp {
  static constexpr int v = PG == 1 ? 1 : PG == 2 ? 3 : 6;
};

// Local expansion at a point (h2dtaeval{p,g,h}, helmrouts2d.f:711-960):
// sum_m X_m J_m(z)/s^|m| e^{imt} for the potential and derivative
// coefficients X, by the backward Miller/Horner sweep.
template <int PG>
__device__ void eval_local(const Deriv& D, cplx z, double s, cplx w, cplx* res) {
  const int nc = D.p + PG - 1;  // highest index with a coefficient
  const int ntop = miller_top(z, nc + 1);
  const bool zero = (abs2(z) == 0.0);
  const cplx zi = zero ? mkc(0.0) : invc(z);
  const cplx us = miller_unit(z, s);
  constexpr int NC = NComp<PG>::v;
  cplx T[NC];
#pragma unroll
  for (int a = 0; a < NC; ++a) T[a] = mkc(0.0);
  cplx rho = mkc(0.0), U = mkc(2.0);
  int n = ntop;
  for (; n > nc + 1; --n) {
    rho = miller_ratio(zi, s, n, rho);
    U = mkc(2.0) + us * (rho * U);
  }
  // n = nc + 1 .. 1: T <- X_{n-1} + rho_n T
  cplx e = cpowi(w, n - 1);
  const cplx wb = conjc(w);
  Combo<PG> cb;
  for (; n >= 1; --n) {
    rho = miller_ratio(zi, s, n, rho);
    U = mkc(n - 1 >= 1 ? 2.0 : 1.0) + us * (rho * U);
    cb.at(D, n - 1, e);
#pragma unroll
    for (int a = 0; a < NC; ++a) T[a] = cb.v[a] + rho * T[a];
    e = e * wb;
  }
  const cplx F0 = miller_exp(z) / U;
#pragma unroll
  for (int a = 0; a < NC; ++a) res[a] += F0 * T[a];
}

// Multipole expansion at a point (h2dmpeval{p,g,h}, helmrouts2d.f:274-510):
// forward sweep with the scaled upward recurrence for H.
template <int PG>
__device__ void eval_mpole(const Deriv& D, cplx z, double s, cplx w, cplx* res) {
  const int nc = D.p + PG - 1;
  constexpr int NC = NComp<PG>::v;
  cplx h0, h1;
  hank01_ool(z, &h0, &h1);
  h1 = h1 * s;
  const cplx zi = invc(z);
  cplx e = mkc(1.0);
  Combo<PG> cb;
  for (int n = 0; n <= nc; ++n) {
    cb.at(D, n, e);
#pragma unroll
    for (int a = 0; a < NC; ++a) res[a] += cb.v[a] * h0;
    const cplx h2 = ((2.0 * (n + 1) * s) * zi) * h1 - (s * s) * h0;
    h0 = h1;
    h1 = h2;
    e = e * w;
  }
}

// One thread per evaluation point: local expansion of its leaf and
// multipoles of the leaf's list-3 boxes (hfmm2d.f:1310-1520).
template <int PG>
__global__ __launch_bounds__(256) void k_heval(int64_t n, int nd, cplx zk, const HLev* lev, const double* xy,
                        const int* leaf, const int* ilevel, const double* centers, const int* n3,
                        const int* l3, const cplx* mp, const cplx* loc, cplx* pot, cplx* grad,
                        cplx* hess) {
  const int64_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = leaf[i];
  const double x = xy[2 * i], y = xy[2 * i + 1];
  const HLev Lb = lev[ilevel[b]];
  const cplx ik4 = cplx{0.0, 0.25};
  const int c3 = n3[b];
  for (int d = 0; d < nd; ++d) {
    constexpr int NC = NComp<PG>::v;
    cplx res[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) res[a] = mkc(0.0);
    if (Lb.on) {
      const double dx = x - centers[2 * b], dy = y - centers[2 * b + 1];
      const double r = hypot(dx, dy);
      const cplx w = (r > 0.0) ? cplx{dx / r, dy / r} : mkc(1.0);
      const Deriv D{hexp(loc, Lb, b, nd, d), Lb.p, 0.5 * zk, Lb.s, 1.0 / Lb.s, 1};
      eval_local<PG>(D, zk * r, Lb.s, w, res);
    }
    for (int k = 0; k < c3; ++k) {
      const int c = l3[(int64_t)kMaxList3 * b + k];
      const HLev Lc = lev[ilevel[c]];
      if (!Lc.on) continue;
      const double dx = x - centers[2 * c], dy = y - centers[2 * c + 1];
      const double r = hypot(dx, dy);
      const cplx w = cplx{dx / r, dy / r};
      const Deriv D{hexp(mp, Lc, c, nd, d), Lc.p, 0.5 * zk, Lc.s, 1.0 / Lc.s, -1};
      eval_mpole<PG>(D, zk * r, Lc.s, w, res);
    }
    pot[i * nd + d] += ik4 * res[0];
    if (PG >= 2) {
      grad[(2 * i) * nd + d] += ik4 * res[1];
      grad[(2 * i + 1) * nd + d] += ik4 * res[2];
    }
    if (PG >= 3) {
      hess[(3 * i) * nd + d] += ik4 * res[3];
      hess[(3 * i + 1) * nd + d] += ik4 * res[4];
      hess[(3 * i + 2) * nd + d] += ik4 * res[5];
    }
  }
}

// nd = 1: a block of kHEvB consecutive points shares out their (point,
// list-3 box) pairs evenly -- in rounds of kHEvB pairs, pair k to thread k mod
// kHEvB, its point found by binary search in the block's scan of list-3
// counts -- so leaves with different list-3 lengths in one wavefront leave no
// lane idle.  Each round's terms go to LDS, and each thread adds its own
// point's terms from there in pair order (a fixed order: the sums do not
// depend on scheduling); each thread then adds its own point's local
// expansion.
constexpr int kHEvB = 256;

template <int PG>
__global__ __launch_bounds__(kHEvB) void k_heval_b(int64_t n, cplx zk, const HLev* lev,
                                                   const double* xy, const int* leaf,
                                                   const int* ilevel, const double* centers,
                                                   const int* n3, const int* l3, const cplx* mp,
                                                   const cplx* loc, cplx* pot, cplx* grad,
                                                   cplx* hess) {
  constexpr int NC = NComp<PG>::v;
  static_assert(kHEvB == kScanThreads, "block_exclusive_scan spans the workgroup");
  __shared__ int scan_tmp[kScanWaves + 1];
  __shared__ int off[kHEvB];
  __shared__ int sb[kHEvB];
  __shared__ double sx[kHEvB], sy[kHEvB];
  __shared__ double part[2 * NC][kHEvB];  // re, im of each component: one round's terms
  const int tid = threadIdx.x;
  const int64_t i = xcd_block() * kHEvB + tid;
  const bool live = i < n;
  const int b = live ? leaf[i] : 0;
  const double x = live ? xy[2 * i] : 0.0, y = live ? xy[2 * i + 1] : 0.0;
  const int cnt = live ? n3[b] : 0;
  sb[tid] = b;
  sx[tid] = x;
  sy[tid] = y;
  double my[2 * NC];
#pragma unroll
  for (int a = 0; a < 2 * NC; ++a) my[a] = 0.0;
  int o, m;
  o = block_exclusive_scan(cnt, scan_tmp, &m);
  off[tid] = o;
  __syncthreads();
  for (int r0 = 0; r0 < m; r0 += kHEvB) {  // uniform
    const int k = r0 + tid;
    if (k < m) {
      int lo = 0, hi = kHEvB - 1;  // last slot t with off[t] <= k
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= k) lo = mid;
        else hi = mid - 1;
      }
      const int c = l3[(int64_t)kMaxList3 * sb[lo] + (k - off[lo])];
      const HLev Lc = lev[ilevel[c]];
      cplx res[NC];
#pragma unroll
      for (int a = 0; a < NC; ++a) res[a] = mkc(0.0);
      if (Lc.on) {
        const double dx = sx[lo] - centers[2 * c], dy = sy[lo] - centers[2 * c + 1];
        const double r = hypot(dx, dy);
        const cplx w = cplx{dx / r, dy / r};
        const Deriv D{hexp(mp, Lc, c, 1, 0), Lc.p, 0.5 * zk, Lc.s, 1.0 / Lc.s, -1};
        eval_mpole<PG>(D, zk * r, Lc.s, w, res);
      }
#pragma unroll
      for (int a = 0; a < NC; ++a) {
        part[2 * a][tid] = res[a].re;
        part[2 * a + 1][tid] = res[a].im;
      }
    }
    __syncthreads();
    // this point's pairs of the round, in pair order
    const int kb = max(o, r0), ke = min(o + cnt, r0 + kHEvB);
    for (int kk = kb; kk < ke; ++kk) {
#pragma unroll
      for (int a = 0; a < 2 * NC; ++a) my[a] += part[a][kk - r0];
    }
    __syncthreads();
  }
  if (!live) return;
  cplx res[NC];
#pragma unroll
  for (int a = 0; a < NC; ++a) res[a] = cplx{my[2 * a], my[2 * a + 1]};
  const HLev Lb = lev[ilevel[b]];
  if (Lb.on) {
    const double dx = x - centers[2 * b], dy = y - centers[2 * b + 1];
    const double r = hypot(dx, dy);
    const cplx w = (r > 0.0) ? cplx{dx / r, dy / r} : mkc(1.0);
    const Deriv D{hexp(loc, Lb, b, 1, 0), Lb.p, 0.5 * zk, Lb.s, 1.0 / Lb.s, 1};
    eval_local<PG>(D, zk * r, Lb.s, w, res);
  }
  const cplx ik4 = cplx{0.0, 0.25};
  pot[i] += ik4 * res[0];
  if (PG >= 2) {
    grad[2 * i] += ik4 * res[1];
    grad[2 * i + 1] += ik4 * res[2];
  }
  if (PG >= 3) {
    hess[3 * i] += ik4 * res[3];
    hess[3 * i + 1] += ik4 * res[4];
    hess[3 * i + 2] += ik4 * res[5];
  }
}

// ---------------------------------------------------------- near field
// hfmm2dpart_direct (hfmm2d.f:1560-1600 -> helmkernels2d.f) over list 1,
// one loop per point over the merged source ranges of its leaf's segment
// row (k_near_segs), so lanes of different leaves in one wavefront stay busy
// until their own totals run out.
template <int IFC, int IFD, int PG, int NDT>
__global__ __launch_bounds__(256) void k_hnear(int64_t n, int nd, int d0, cplx zk, const double* xy, const int* leaf,
                        const int2* seg, const double* xs, const cplx* chs, const cplx* dps,
                        const double* dvs, double thresh, cplx* pot, cplx* grad, cplx* hess) {
  const int64_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = leaf[i];
  const double x = xy[2 * i], y = xy[2 * i + 1];
  Acc<HELM, PG, NDT> a;
  a.zero();
  SegCursor sc;
  const int T = sc.init(seg + (int64_t)kSegRow * b);
  for (int t = 0; t < T; ++t) {
    const int j = sc.j;
    SrcStr<NDT> s;
#pragma unroll
    for (int dd = 0; dd < NDT; ++dd) {
      const bool ok = d0 + dd < nd;
      const int dc = ok ? d0 + dd : d0;
      if (IFC) s.charge[dd] = ok ? chs[(int64_t)j * nd + dc] : mkc(0.0);
      if (IFD) {
        s.dipstr[dd] = ok ? dps[(int64_t)j * nd + dc] : mkc(0.0);
        s.dv[dd][0] = dvs[(2 * (int64_t)j) * nd + dc];
        s.dv[dd][1] = dvs[(2 * (int64_t)j + 1) * nd + dc];
      }
    }
    pair_helm<IFC, IFD, PG, NDT>(x - xs[2 * (int64_t)j], y - xs[2 * (int64_t)j + 1], zk, s,
                                 thresh, a);
    sc.j = j + 1;
    sc.step();
  }
#pragma unroll
  for (int dd = 0; dd < NDT; ++dd) {
    if (d0 + dd >= nd) break;
    const int dc = d0 + dd;
    pot[i * nd + dc] += a.pot[dd];
    if (PG >= 2) {
      grad[(2 * i) * nd + dc] += a.grad[dd][0];
      grad[(2 * i + 1) * nd + dc] += a.grad[dd][1];
    }
    if (PG >= 3) {
      hess[(3 * i) * nd + dc] += a.hess[dd][0];
      hess[(3 * i + 1) * nd + dc] += a.hess[dd][1];
      hess[(3 * i + 2) * nd + dc] += a.hess[dd][2];
    }
  }
}

// Mixed precision (real wavenumber): the pair is formed in FP64 (displacement,
// |z|^2, r, and the phase of H inside hank01_f32), the Hankel pair and the
// kernel products in FP32, summed in FP32 over at most 32 sources and folded
// into FP64.  Pairs with k r < 1e-4 -- the point's own pair, duplicates and
// near-duplicates, where 1/z and the threshold rule matter -- take the exact
// FP64 formula and exclusion rule out of line.
template <int IFC, int IFD, int PG, int NDT> struct AccH32 {
  c32 pot[NDT];
  c32 grad[NDT][2];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int d = 0; d < NDT; ++d) {
      pot[d] = c32{0.f, 0.f};
      grad[d][0] = grad[d][1] = c32{0.f, 0.f};
    }
  }
};

template <int IFC, int IFD, int PG, int NDT>
__device__ __noinline__ void pair_helm_slow(double dx, double dy, cplx zk,
                                            const SrcStr<NDT>& s, double thresh,
                                            Acc<HELM, PG, NDT>& a) {
  pair_helm<IFC, IFD, PG, NDT, true>(dx, dy, zk, s, thresh, a);
}

template <int IFC, int IFD, int PG, int NDT>
__device__ __forceinline__ void pair_helm_f32(double dx, double dy, double rr, double k,
                                              const SrcStr<NDT>& s,
                                              AccH32<IFC, IFD, PG, NDT>& f) {
  const double r = sqrt(rr);
  c32 h0, h1;
  hank01_f32(k * r, &h0, &h1);
  const float rinv = 1.0f / (float)r, kf = (float)k, fx = (float)dx, fy = (float)dy;
  const float k4 = 0.25f * kf * rinv;
  if (IFC) {
    const c32 p0{-0.25f * h0.im, 0.25f * h0.re};  // (i/4) H0
    const c32 cdd{k4 * h1.im, -k4 * h1.re};       // -(ik/4) H1 / r
#pragma unroll
    for (int d = 0; d < NDT; ++d) {
      const c32 c{(float)s.charge[d].re, (float)s.charge[d].im};
      f.pot[d] = f.pot[d] + p0 * c;
      if (PG >= 2) {
        const c32 cd = cdd * c;
        f.grad[d][0] = f.grad[d][0] + fx * cd;
        f.grad[d][1] = f.grad[d][1] + fy * cd;
      }
    }
  }
  if (IFD) {
    const c32 cdd{-k4 * h1.im, k4 * h1.re};  // (ik/4) H1 / r
    c32 hf1{0.f, 0.f}, hf2{0.f, 0.f}, hf3{0.f, 0.f};
    const float kk4 = -0.25f * kf * kf;      // -(k^2) i/4 = (0, kk4)
    if (PG >= 2) {
      const float zi = rinv / kf;
      const c32 h2 = (2.0f * zi) * h1 - h0;
      const float ct = fx * rinv, st = fy * rinv;
      hf1 = (ct * ct - 0.5f) * h2 - 0.5f * h0;
      hf2 = (ct * st) * h2;
      hf3 = (st * st - 0.5f) * h2 - 0.5f * h0;
    }
#pragma unroll
    for (int d = 0; d < NDT; ++d) {
      const c32 ds{(float)s.dipstr[d].re, (float)s.dipstr[d].im};
      const float v1 = (float)s.dv[d][0], v2 = (float)s.dv[d][1];
      f.pot[d] = f.pot[d] + (fx * v1 + fy * v2) * (cdd * ds);
      if (PG >= 2) {
        const c32 cd{-kk4 * ds.im, kk4 * ds.re};
        const c32 hxx = cd * hf1, hxy = cd * hf2, hyy = cd * hf3;
        f.grad[d][0] = f.grad[d][0] + v1 * hxx + v2 * hxy;
        f.grad[d][1] = f.grad[d][1] + v1 * hxy + v2 * hyy;
      }
    }
  }
}

template <int IFC, int IFD, int PG, int NDT>
__device__ __forceinline__ void fold_h32(AccH32<IFC, IFD, PG, NDT>& f, Acc<HELM, PG, NDT>& a) {
#pragma unroll
  for (int d = 0; d < NDT; ++d) {
    a.pot[d] += cplx{(double)f.pot[d].re, (double)f.pot[d].im};
    if (PG >= 2) {
      a.grad[d][0] += cplx{(double)f.grad[d][0].re, (double)f.grad[d][0].im};
      a.grad[d][1] += cplx{(double)f.grad[d][1].re, (double)f.grad[d][1].im};
    }
  }
  f.zero();
}

// k_hnear with the mixed-precision pair (real zk, PG <= 2)
template <int IFC, int IFD, int PG, int NDT>
__global__ void k_hnear32(int64_t n, int nd, int d0, cplx zk, const double* xy, const int* leaf,
                          const int2* seg, const double* xs, const cplx* chs, const cplx* dps,
                          const double* dvs, double thresh, cplx* pot, cplx* grad, cplx* hess) {
  static_assert(PG <= 2, "hfmm2d has no Hessians");
  const int64_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = leaf[i];
  const double x = xy[2 * i], y = xy[2 * i + 1];
  const double k = zk.re;
  const double rr_slow = 1e-8 / (k * k);  // k r < 1e-4
  Acc<HELM, PG, NDT> a;
  a.zero();
  AccH32<IFC, IFD, PG, NDT> f;
  f.zero();
  SegCursor sc;
  const int T = sc.init(seg + (int64_t)kSegRow * b);
  for (int t = 0; t < T; ++t) {
    const int j = sc.j;
    SrcStr<NDT> s;
#pragma unroll
    for (int dd = 0; dd < NDT; ++dd) {
      const bool ok = d0 + dd < nd;
      const int dc = ok ? d0 + dd : d0;
      if (IFC) s.charge[dd] = ok ? chs[(int64_t)j * nd + dc] : mkc(0.0);
      if (IFD) {
        s.dipstr[dd] = ok ? dps[(int64_t)j * nd + dc] : mkc(0.0);
        s.dv[dd][0] = dvs[(2 * (int64_t)j) * nd + dc];
        s.dv[dd][1] = dvs[(2 * (int64_t)j + 1) * nd + dc];
      }
    }
    const double dx = x - xs[2 * (int64_t)j], dy = y - xs[2 * (int64_t)j + 1];
    const double rr = fma(dx, dx, dy * dy);
    if (rr < rr_slow) pair_helm_slow<IFC, IFD, PG, NDT>(dx, dy, zk, s, thresh, a);
    else pair_helm_f32<IFC, IFD, PG, NDT>(dx, dy, rr, k, s, f);
    if ((t & 31) == 31) fold_h32<IFC, IFD, PG, NDT>(f, a);
    sc.j = j + 1;
    sc.step();
  }
  fold_h32<IFC, IFD, PG, NDT>(f, a);
#pragma unroll
  for (int dd = 0; dd < NDT; ++dd) {
    if (d0 + dd >= nd) break;
    const int dc = d0 + dd;
    pot[i * nd + dc] += a.pot[dd];
    if (PG >= 2) {
      grad[(2 * i) * nd + dc] += a.grad[dd][0];
      grad[(2 * i + 1) * nd + dc] += a.grad[dd][1];
    }
  }
}

// ---------------------------------------------------------- operators
// Operator o: entry (i, j) (output i, input j, |i| <= pout, |j| <= pin) is
//   G[i - j] * w1[sa |i-j| + sb |j| + sc |i|] * w2[|i|]
// with G the scaled cylinder-function table of the offset and w1, w2 powers
// of the level scales, written straight into the swizzled MFMA A layout.
struct OpDesc {
  int64_t af;   // first double of the swizzled operator
  int pin, pout;
  int g;        // index of G[0] in the table
  int w1, w2;   // first power of each power table
  int sa, sb, sc;
  int pad;
};

__global__ void k_to_f32(int64_t n, const double* in, float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

__global__ void k_build_ops(const OpDesc* ops, const cplx* gt, const double* wt, double* af) {
  const OpDesc o = ops[blockIdx.y];
  const int Pin = 2 * o.pin + 1, Pout = 2 * o.pout + 1;
  const int RTT = (2 * Pout + 15) / 16, KS = (2 * Pin + 3) / 4;
  const int64_t total = (int64_t)RTT * KS * 64;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(e & 63);
    const int64_t rest = e >> 6;
    const int ks = (int)(rest % KS), rt = (int)(rest / KS);
    const int row = 16 * rt + (lane & 15), col = 4 * ks + (lane >> 4);
    double v = 0.0;
    if (row < 2 * Pout && col < 2 * Pin) {
      const int i = row / 2 - o.pout, j = col / 2 - o.pin;
      const int ex = o.sa * abs(i - j) + o.sb * abs(j) + o.sc * abs(i);
      const cplx t = gt[o.g + (i - j)] * (wt[o.w1 + ex] * wt[o.w2 + abs(i)]);
      const int ra = row & 1, cb = col & 1;
      v = (ra == cb) ? t.re : (ra == 0 ? -t.im : t.im);
    }
    af[o.af + e] = v;
  }
}

using lc = std::complex<long double>;

struct OpTables {
  std::vector<OpDesc> desc;
  std::vector<cplx> g;
  std::vector<double> w;
  int64_t af_total = 0;
};

cplx to_c(lc v) { return cplx{(double)v.real(), (double)v.imag()}; }

// G[m] = F_|m| sign^m e^{-i m a} for m = -mmax..mmax from f[0..mmax]
int push_table(OpTables& T, const std::vector<lc>& f, int mmax, long double ang) {
  const int g0 = (int)T.g.size() + mmax;
  T.g.resize(T.g.size() + 2 * mmax + 1);
  const lc u = std::polar(1.0L, -ang);
  lc e = 1.0L;
  for (int m = 0; m <= mmax; ++m) {
    const lc v = f[m] * e;                  // F_m e^{-i m a}
    T.g[g0 + m] = to_c(v);
    const lc vm = f[m] * std::conj(e);      // F_m e^{+i m a}
    T.g[g0 - m] = to_c((m & 1) ? -vm : vm);
    e *= u;
  }
  return g0;
}

int push_powers(OpTables& T, long double base, int emax) {
  const int w0 = (int)T.w.size();
  long double v = 1.0L;
  for (int e = 0; e <= emax; ++e) {
    T.w.push_back((double)v);
    v *= base;
  }
  return w0;
}

void add_op(OpTables& T, int pin, int pout, int g, int w1, int w2, int sa, int sb, int sc) {
  OpDesc o{};
  o.af = T.af_total;
  o.pin = pin;
  o.pout = pout;
  o.g = g;
  o.w1 = w1;
  o.w2 = w2;
  o.sa = sa;
  o.sb = sb;
  o.sc = sc;
  T.desc.push_back(o);
  T.af_total += afrag_stride(2 * pin + 1, 2 * pout + 1);
}

// scaled H_m(z) s^m, m = 0..mmax, long double upward recurrence from H0, H1
std::vector<lc> host_scaled_h(lc z, long double s, int mmax) {
  std::vector<lc> h(mmax + 2);
  cplx h0, h1;
  hank01(to_c(z), &h0, &h1);
  h[0] = lc(h0.re, h0.im);
  h[1] = lc(h1.re, h1.im) * s;
  for (int n = 1; n < mmax; ++n)
    h[n + 1] = (2.0L * n * s / z) * h[n] - s * s * h[n - 1];
  h.resize(mmax + 1);
  return h;
}

template <int PG>
void launch_heval(int64_t n, hipStream_t st, int nd, cplx zk, const HLev* lev, const double* xy,
                  const int* leaf, const Tree& t, const Lists& L, const cplx* mp, const cplx* loc,
                  cplx* p, cplx* g, cplx* h) {
  static_assert(kHEvB == 256, "launch1d_xcd launches blocks of 256 threads");
  if (nd == 1)
    launch1d_xcd(k_heval_b<PG>, n, st, zk, lev, xy, leaf, t.ilevel.p, t.centers.p, L.n3.p,
                 L.l3.p, mp, loc, p, g, h);
  else
    launch1d_xcd(k_heval<PG>, n, st, nd, zk, lev, xy, leaf, t.ilevel.p, t.centers.p, L.n3.p,
                 L.l3.p, mp, loc, p, g, h);
}

template <int IFC, int IFD, int NDT>
void launch_hnear_f(int pg, int64_t n, hipStream_t st, int nd, int d0, cplx zk, const double* xy,
                    const int* leaf, const int2* seg, const double* xs,
                    const cplx* chs, const cplx* dps, const double* dvs, double thresh, cplx* p,
                    cplx* g, cplx* h, bool f32) {
  if (f32 && pg == 1)
    launch1d_xcd(k_hnear32<IFC, IFD, 1, NDT>, n, st, nd, d0, zk, xy, leaf, seg, xs, chs, dps,
                 dvs, thresh, p, g, h);
  else if (f32)
    launch1d_xcd(k_hnear32<IFC, IFD, 2, NDT>, n, st, nd, d0, zk, xy, leaf, seg, xs, chs, dps,
                 dvs, thresh, p, g, h);
  else if (pg == 1)
    launch1d_xcd(k_hnear<IFC, IFD, 1, NDT>, n, st, nd, d0, zk, xy, leaf, seg, xs, chs, dps,
             dvs, thresh, p, g, h);
  else if (pg == 2)
    launch1d_xcd(k_hnear<IFC, IFD, 2, NDT>, n, st, nd, d0, zk, xy, leaf, seg, xs, chs, dps,
             dvs, thresh, p, g, h);
  else
    launch1d_xcd(k_hnear<IFC, IFD, 3, NDT>, n, st, nd, d0, zk, xy, leaf, seg, xs, chs, dps,
             dvs, thresh, p, g, h);
}

template <int NDT>
void launch_hnear_n(int pg, int ifc, int ifd, int64_t n, hipStream_t st, int nd, int d0, cplx zk,
                    const double* xy, const int* leaf, const int2* seg,
                    const double* xs, const cplx* chs, const cplx* dps, const double* dvs,
                    double thresh, cplx* p, cplx* g, cplx* h, bool f32) {
  if (ifc && !ifd)
    launch_hnear_f<1, 0, NDT>(pg, n, st, nd, d0, zk, xy, leaf, seg, xs, chs, dps, dvs, thresh,
                              p, g, h, f32);
  else if (!ifc && ifd)
    launch_hnear_f<0, 1, NDT>(pg, n, st, nd, d0, zk, xy, leaf, seg, xs, chs, dps, dvs, thresh,
                              p, g, h, f32);
  else
    launch_hnear_f<1, 1, NDT>(pg, n, st, nd, d0, zk, xy, leaf, seg, xs, chs, dps, dvs, thresh,
                              p, g, h, f32);
}

// f32: the mixed-precision pair (real zk and pg <= 2 only)
void launch_hnear(int pg, int ifc, int ifd, int64_t n, hipStream_t st, int nd, cplx zk,
                  const double* xy, const int* leaf, const int2* seg,
                  const double* xs, const cplx* chs, const cplx* dps, const double* dvs,
                  double thresh, cplx* p, cplx* g, cplx* h, bool f32 = false) {
  const int NDT = (nd == 1) ? 1 : 2;
  f32 = f32 && zk.im == 0.0 && pg <= 2;
  for (int d0 = 0; d0 < nd; d0 += NDT) {
    if (NDT == 1)
      launch_hnear_n<1>(pg, ifc, ifd, n, st, nd, d0, zk, xy, leaf, seg, xs, chs, dps, dvs,
                        thresh, p, g, h, f32);
    else
      launch_hnear_n<2>(pg, ifc, ifd, n, st, nd, d0, zk, xy, leaf, seg, xs, chs, dps, dvs,
                        thresh, p, g, h, f32);
  }
}

// ---------------------------------------------------- multipole sources (mps)
// hfmm2d_mps (helmholtz/hfmm2d_mps.f): the sources are multipole expansions,
// each with its own centre, scale and order, and the outputs are local
// expansions about the same centres.  A wavefront applies one translation
// between two expansions of arbitrary centres, scales and orders -- h2dmpmp
// (kind 0), h2dmploc (kind 1), h2dlocloc (kind 2), helmrouts2d.f:967, 1221,
// 1096.  In unscaled coefficients all three are the truncated Graf theorem
//   out_i += sum_j in_j F_{i-j}(k r) e^{-i (i-j) a},  F = J (kinds 0, 2), H (1),
// (r, a + pi) the polar form of c_out - c_in and F_{-n} = (-1)^n F_n.  Stored
// multipoles are true / s^|n| and stored locals true * s^|n|; the cylinder
// functions come scaled by the input's scale s1 (J_n / s1^n, H_n s1^n, as
// jbessel2d / h2dall), and each term takes the single power of s1 and s2 that
// gives the output's scaling (one exp2 per term, so no partial product
// overflows where the scaled result does not).
struct MpsExp {
  double x, y, s;  // centre, scale
  int p;           // order
  int pad;
  int64_t off;     // 0-based index of coefficient (density 0, n = 0) in mpole / local
};

constexpr int kMpsW = 64;
constexpr size_t kMpsLds = 64 * 1024;  // dynamic LDS of the mps translation kernels

// ft: 2 N + 2 complex of LDS, N = p1 + p2; in(n, d) = in[n sn + d sd], n = -p1..p1.
__device__ void mps_translate(int kind, cplx zk, int nd, const cplx* __restrict__ in, int64_t isn,
                              int64_t isd, double x1, double y1, double s1, int p1,
                              cplx* __restrict__ out, int64_t osn, int64_t osd, double x2,
                              double y2, double s2, int p2, cplx* ft) {
  const int N = p1 + p2;
  const double dx = x2 - x1, dy = y2 - y1;
  const double r = sqrt(dx * dx + dy * dy);
  const double th = atan2(dy, dx) - kPi;
  __syncthreads();  // the previous translation's reads of ft are done
  if (threadIdx.x == 0) {
    const cplx z = zk * r;
    if (kind == 1) scaled_h(z, s1, N > 1 ? N : 1, ft + N);
    else scaled_j(z, s1, N > 1 ? N : 1, ft + N);
  }
  __syncthreads();
  for (int n = threadIdx.x; n <= N; n += kMpsW) {
    const cplx f = ft[N + n];
    double sn, cs;
    sincos(n * th, &sn, &cs);
    ft[N + n] = f * cplx{cs, -sn};
    if (n > 0) {
      const cplx g = f * cplx{cs, sn};
      ft[N - n] = (n & 1) ? cplx{-g.re, -g.im} : g;
    }
  }
  __syncthreads();
  const double a = (kind == 1) ? -1.0 : 1.0, b = (kind == 2) ? -1.0 : 1.0;
  const double c = (kind == 0) ? -1.0 : 1.0;
  const double l1 = log2(s1), l2 = log2(s2);
  for (int i = (int)threadIdx.x - p2; i <= p2; i += kMpsW) {
    const double ei = c * (double)abs(i) * l2;
    for (int d = 0; d < nd; ++d) {
      cplx acc = mkc(0.0);
      for (int j = -p1; j <= p1; ++j) {
        const double e = (a * (double)abs(i - j) + b * (double)abs(j)) * l1 + ei;
        acc = acc + (in[j * isn + d * isd] * ft[N + i - j]) * exp2(e);
      }
      cplx* o = out + i * osn + d * osd;
      *o = *o + acc;
    }
  }
}

__global__ void k_mps_gather(int64_t n, const int* perm, const double* cen, const double* rs,
                             const int* mt, const int64_t* imp, MpsExp* e) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  MpsExp v;
  v.x = cen[2 * j];
  v.y = cen[2 * j + 1];
  v.s = rs[j];
  v.p = mt[j];
  v.pad = 0;
  v.off = imp[j] - 1;
  e[i] = v;
}

__global__ void k_mps_pmax(int64_t n, const int* mt, int* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  atomicMax(out, mt[i]);
  atomicMin(out + 1, mt[i]);
}

// Step 1 (hfmm2d_mps.f:505-530): each leaf's expansions shifted to the leaf's
// multipole (h2dmpmp), one wavefront per box.
__global__ __launch_bounds__(kMpsW) void k_mps_p2m(int nd, cplx zk, const HLev* lev,
                                                   const int* ilevel, const int* nchild,
                                                   const int* se, const double* centers,
                                                   const MpsExp* ex, const cplx* mpole,
                                                   cplx* mp) {
  extern __shared__ cplx ft[];
  const int b = blockIdx.x;
  const int l = ilevel[b];
  if (l < 2 || !lev[l].on || nchild[b] != 0) return;
  const HLev V = lev[l];
  cplx* o = hexp(mp, V, b, nd, 0);
  for (int e = se[2 * b]; e < se[2 * b + 1]; ++e) {
    const MpsExp E = ex[e];
    mps_translate(0, zk, nd, mpole + E.off + (int64_t)E.p * nd, nd, 1, E.x, E.y, E.s, E.p, o, 1,
                  2 * V.p + 1, centers[2 * b], centers[2 * b + 1], V.s, V.p, ft);
  }
}

// Step 2 (hfmm2d_mps.f:535-565): list-4 expansions into the box's local
// (h2dmploc), one wavefront per box.
__global__ __launch_bounds__(kMpsW) void k_mps_p2l(int nd, cplx zk, const HLev* lev,
                                                   const int* ilevel, const int* n4,
                                                   const int* l4, const int* se,
                                                   const double* centers, const MpsExp* ex,
                                                   const cplx* mpole, cplx* loc) {
  extern __shared__ cplx ft[];
  const int b = blockIdx.x;
  const int l = ilevel[b];
  if (l < 2 || !lev[l].on || n4[b] == 0) return;
  const HLev V = lev[l];
  cplx* o = hexp(loc, V, b, nd, 0);
  for (int k = 0; k < n4[b]; ++k) {
    const int jb = l4[(int64_t)b * kMaxList4 + k];
    for (int e = se[2 * jb]; e < se[2 * jb + 1]; ++e) {
      const MpsExp E = ex[e];
      mps_translate(1, zk, nd, mpole + E.off + (int64_t)E.p * nd, nd, 1, E.x, E.y, E.s, E.p, o,
                    1, 2 * V.p + 1, centers[2 * b], centers[2 * b + 1], V.s, V.p, ft);
    }
  }
}

// Steps 6-7 (hfmm2d_mps.f:880-945): one wavefront per expansion clears its
// local, adds the list-3 boxes' multipoles (h2dmploc; the reference gates
// them on the list-3 level, ilev + 1) and its leaf's local (h2dlocloc).
__global__ __launch_bounds__(kMpsW) void k_mps_far(int64_t n, int nd, cplx zk, const HLev* lev,
                                                   int nl, const int* leaf, const int* ilevel,
                                                   const int* n3, const int* l3,
                                                   const double* centers, const MpsExp* ex,
                                                   const cplx* mp, const cplx* loc,
                                                   cplx* local) {
  extern __shared__ cplx ft[];
  const int64_t e = blockIdx.x;
  const MpsExp E = ex[e];
  cplx* o = local + E.off + (int64_t)E.p * nd;
  for (int k = threadIdx.x; k < nd * (2 * E.p + 1); k += kMpsW) local[E.off + k] = mkc(0.0);
  const int b = leaf[e];
  const int l = ilevel[b];
  if (l >= 1 && l <= nl - 1 && lev[l + 1].on) {
    for (int k = 0; k < n3[b]; ++k) {
      const int jb = l3[(int64_t)b * kMaxList3 + k];
      const HLev W = lev[ilevel[jb]];
      mps_translate(1, zk, nd, hexp(mp, W, jb, nd, 0), 1, 2 * W.p + 1, centers[2 * jb],
                    centers[2 * jb + 1], W.s, W.p, o, nd, 1, E.x, E.y, E.s, E.p, ft);
    }
  }
  const HLev V = lev[l];
  if (V.on)
    mps_translate(2, zk, nd, hexp(loc, V, b, nd, 0), 1, 2 * V.p + 1, centers[2 * b],
                  centers[2 * b + 1], V.s, V.p, o, nd, 1, E.x, E.y, E.s, E.p, ft);
}

// Step 8 (hfmm2d_mps.f:960-1000): the expansions of the leaf's list-1 boxes
// whose centres are farther than thresh (so not the expansion itself),
// h2dmploc, one wavefront per expansion.
__global__ __launch_bounds__(kMpsW) void k_mps_near(int64_t n, int nd, cplx zk, double thresh,
                                                    const int* leaf, const int* n1,
                                                    const int* l1, const int* se,
                                                    const MpsExp* ex, const cplx* mpole,
                                                    cplx* local) {
  extern __shared__ cplx ft[];
  const int64_t e = blockIdx.x;
  const MpsExp E = ex[e];
  cplx* o = local + E.off + (int64_t)E.p * nd;
  const int b = leaf[e];
  for (int k = 0; k < n1[b]; ++k) {
    const int jb = l1[(int64_t)b * kMaxList1 + k];
    for (int j = se[2 * jb]; j < se[2 * jb + 1]; ++j) {
      const MpsExp S = ex[j];
      const double dx = S.x - E.x, dy = S.y - E.y;
      const double d = sqrt(dx * dx + dy * dy);
      if (!(d > thresh)) continue;
      mps_translate(1, zk, nd, mpole + S.off + (int64_t)S.p * nd, nd, 1, S.x, S.y, S.s, S.p, o,
                    nd, 1, E.x, E.y, E.s, E.p, ft);
    }
  }
}

// Box-to-box passes shared by hfmm2d and hfmm2d_mps (hfmm2d.f:1021-1300,
// hfmm2d_mps.f:520-760): translation operators for every level, then M2M
// (finest parents first), M2L level by level and L2L (coarsest first), on the
// per-level multipole and local blocks.  pg / pgt say which boxes hold
// evaluation points (sources / targets), as the planners need.  mixed: the M2L
// runs on v_mfma_f32_16x16x4f32 with FP32 copies of the operators (the
// swizzled A-fragment layout is the same for both element types) and of the
// multipoles, accumulating in FP32 and adding into the FP64 locals.
void helm_box_passes(const Tree& t, const Lists& L, const std::vector<HLev>& lev,
                     const std::vector<long double>& sl, const std::vector<int>& nterms,
                     int nd, cplx zk, int pg, int pgt, cplx* mpp, cplx* locp, Timer& tm,
                     hipStream_t st, bool mixed = false) {
  const int nl = t.nlevels;
  // ---- translation operators, level by level
  OpTables T;
  std::vector<int> op_m2l(nl + 1, -1), op_m2m(nl + 1, -1), op_l2l(nl + 1, -1);
  const lc k(zk.re, zk.im);
  for (int l = 1; l <= nl; ++l) {
    const long double bs = t.boxsize[l];
    const int p = nterms[l];
    if (l >= 2 && lev[l].on) {
      // M2L: source - target = (dx, dy) bs, max(|dx|,|dy|) >= 2 (h2dmploc,
      // helmrouts2d.f:1221-1380)
      op_m2l[l] = (int)T.desc.size();
      const int w1 = push_powers(T, sl[l], 2 * p);
      const int w2 = push_powers(T, 1.0L, p);
      std::vector<std::vector<lc>> hr(19);
      for (int dx = -3; dx <= 3; ++dx)
        for (int dy = -3; dy <= 3; ++dy) {
          const int r2 = dx * dx + dy * dy;
          if (std::abs(dx) <= 1 && std::abs(dy) <= 1) {
            add_op(T, p, p, 0, w1, w2, 0, 0, 0);  // never referenced: zero table slot
            T.desc.back().g = -1;
            continue;
          }
          if (hr[r2].empty())
            hr[r2] = host_scaled_h(k * (bs * std::sqrt((long double)r2)), sl[l], 2 * p);
          const int g = push_table(T, hr[r2], 2 * p, std::atan2((long double)dy, (long double)dx));
          add_op(T, p, p, g, w1, w2, -1, 1, 1);
        }
    }
    if (l <= nl - 1 && lev[l].on) {
      // M2M into level l from children at l + 1 (h2dmpmp, helmrouts2d.f:967-1090) and
      // L2L from level l to its children (h2dlocloc, helmrouts2d.f:1096-1215);
      // child - parent = (sx, sy) bs / 4
      const int pc = nterms[l + 1];
      const int mm = p + pc;
      const lc z = k * (bs * std::sqrt(2.0L) / 4.0L);
      std::vector<lc> J = bessel_j(z, std::max(mm, 1));
      std::vector<lc> Jc(J), Jp(J);  // scaled by the child's / parent's s^-m
      for (int m = 0; m <= mm; ++m) {
        Jc[m] = J[m] / std::pow(sl[l + 1], (long double)m);
        Jp[m] = J[m] / std::pow(sl[l], (long double)m);
      }
      op_m2m[l] = (int)T.desc.size();
      {
        const int w1 = push_powers(T, sl[l + 1], 2 * std::max(p, pc));
        const int w2 = push_powers(T, sl[l + 1] / sl[l], p);
        for (int q = 0; q < 4; ++q) {
          const long double sx = (q & 1) ? 1.0L : -1.0L, sy = (q & 2) ? 1.0L : -1.0L;
          const int g = push_table(T, Jc, mm, std::atan2(sy, sx));
          add_op(T, pc, p, g, w1, w2, 1, 1, -1);
        }
      }
      op_l2l[l] = (int)T.desc.size();
      {
        const int w1 = push_powers(T, sl[l], 2 * std::max(p, pc));
        const int w2 = push_powers(T, sl[l + 1] / sl[l], pc);
        for (int q = 0; q < 4; ++q) {
          const long double sx = (q & 1) ? 1.0L : -1.0L, sy = (q & 2) ? 1.0L : -1.0L;
          const int g = push_table(T, Jp, mm, std::atan2(-sy, -sx));
          add_op(T, p, pc, g, w1, w2, 1, -1, 1);
        }
      }
    }
  }
  DBuf<double> d_af(std::max<int64_t>(T.af_total, 1), st);
  {
    // drop the never-referenced slots from the build list
    std::vector<OpDesc> build;
    int64_t maxaf = 0;
    for (const OpDesc& od : T.desc)
      if (od.g >= 0) {
        build.push_back(od);
        maxaf = std::max(maxaf, afrag_stride(2 * od.pin + 1, 2 * od.pout + 1));
      }
    if (!build.empty()) {
      DBuf<OpDesc> d_desc = upload(build, st);
      DBuf<cplx> d_g = upload(T.g, st);
      DBuf<double> d_w = upload(T.w, st);
      const unsigned gx = (unsigned)std::min<int64_t>((maxaf + 255) / 256, 64);
      hipLaunchKernelGGL(k_build_ops, dim3(gx, (unsigned)build.size()), dim3(256), 0, st,
                         d_desc.p, d_g.p, d_w.p, d_af.p);
      hip_throw(hipGetLastError(), "build_ops");
      hip_throw(hipStreamSynchronize(st), "sync");
    }
  }
  auto opptr = [&](int first) { return d_af.p + T.desc[first].af; };
  DBuf<float> af32(mixed ? std::max<int64_t>(T.af_total, 1) : 0, st);
  if (mixed) launch1d(k_to_f32, T.af_total, st, (const double*)d_af.p, af32.p);
  tm.mark(12);

  ChildPlan cp;
  plan_children(t, pg, pgt, cp, st);
  DBuf<int> d_op4 = upload(std::vector<int>{0, 1, 2, 3}, st);
  // ---- step 3: M2M, finest parents first (hfmm2d.f:1021-1100)
  for (int l = nl - 1; l >= 1; --l) {
    if (op_m2m[l] < 0) continue;
    const HLev &P = lev[l], &C = lev[l + 1];
    launch_translate(0, 0, nd,
                     TrBufs{mpp + C.off, C.box0, 2 * C.p + 1, mpp + P.off, P.box0,
                            2 * P.p + 1},
                     4, {TrGroup{cp.lev_pos[l], cp.lev_pos[l + 1], 0, 0}}, cp.list.p, cp.up.p,
                     d_op4.p, opptr(op_m2m[l]), t.ilevel.p, nullptr, nullptr, st);
  }
  tm.mark(2);
  // ---- step 4: M2L level by level, targets grouped by child position
  {
    M2LPlan m2l;
    plan_m2l(t, L, pg, pgt, m2l, st);
    if (m2l.ntg > 0) {
      DBuf<int> d_soff = upload(m2l_slot_offsets(), st);
      // FP32 copy of every level's multipoles (mixed precision)
      const int64_t ncoef = lev[nl].off + (int64_t)(t.nboxes - t.lstart[nl]) * nd *
                                              (2 * lev[nl].p + 1);
      DBuf<float> mp32(mixed ? 2 * std::max<int64_t>(ncoef, 1) : 0, st);
      if (mixed) launch1d(k_to_f32, 2 * ncoef, st, (const double*)mpp, mp32.p);
      tm.mark(3);  // planning
      for (int l = 2; l <= nl; ++l) {
        if (op_m2l[l] < 0) continue;
        std::vector<TrGroup> groups;
        for (int q = 0; q < 4; ++q)
          groups.push_back(TrGroup{m2l.lev(q, l, nl), m2l.lev(q, l + 1, nl), 0, q});
        const HLev& V = lev[l];
        if (mixed)
          launch_translate(0, 0, nd,
                           TrBufs{mp32.p + 2 * V.off, V.box0, 2 * V.p + 1, locp + V.off,
                                  V.box0, 2 * V.p + 1},
                           kSlots, groups, m2l.tlist.p, L.s2.p, d_soff.p,
                           af32.p + T.desc[op_m2l[l]].af, t.ilevel.p, nullptr, nullptr, st,
                           true, true, true);
        else
          launch_translate(0, 0, nd,
                           TrBufs{mpp + V.off, V.box0, 2 * V.p + 1, locp + V.off, V.box0,
                                  2 * V.p + 1},
                           kSlots, groups, m2l.tlist.p, L.s2.p, d_soff.p, opptr(op_m2l[l]),
                           t.ilevel.p, nullptr, nullptr, st, false, true, true);
      }
      tm.mark(13);  // the M2L kernels alone
    }
  }
  tm.mark(3);
  // ---- step 5: L2L, coarsest first (hfmm2d.f:1225-1300)
  for (int l = 1; l <= nl - 1; ++l) {
    if (op_l2l[l] < 0) continue;
    const HLev &P = lev[l], &C = lev[l + 1];
    launch_translate(1, 0, nd,
                     TrBufs{locp + P.off, P.box0, 2 * P.p + 1, locp + C.off, C.box0,
                            2 * C.p + 1},
                     4, {TrGroup{cp.lev_pos[l], cp.lev_pos[l + 1], 0, 0}}, cp.list.p,
                     cp.down.p, d_op4.p, opptr(op_l2l[l]), t.ilevel.p, nullptr, nullptr, st);
  }
  tm.mark(4);
}

}  // namespace

int helmholtz_fmm(const HelmArgs& a, hipStream_t st) {
  if (a.nd < 1 || a.ns < 1 || a.nt < 0 || a.ifpgh < 0 || a.ifpgh > 3 || a.ifpghtarg < 0 ||
      a.ifpghtarg > 3)
    return FMM2D_ERR_ARG;
  if (!a.src || (a.nt > 0 && !a.targ) || (a.ifcharge && !a.charge) ||
      (a.ifdipole && (!a.dipstr || !a.dipvec)))
    return FMM2D_ERR_ARG;
  if ((a.ifpgh >= 1 && !a.pot) || (a.ifpgh >= 2 && !a.grad) || (a.ifpgh >= 3 && !a.hess) ||
      (a.ifpghtarg >= 1 && !a.pottarg) || (a.ifpghtarg >= 2 && !a.gradtarg) ||
      (a.ifpghtarg >= 3 && !a.hesstarg))
    return FMM2D_ERR_ARG;
  if (a.ns > INT32_MAX || a.nt > INT32_MAX) return FMM2D_ERR_ARG;
  if (a.zk.re == 0.0 && a.zk.im == 0.0) return FMM2D_ERR_ARG;
  if (!a.ifcharge && !a.ifdipole) return FMM2D_OK;
  const fmm2d_opts* o = a.opts;
  if (o && (o->precision < 0 || o->precision > 1)) return FMM2D_ERR_ARG;
  // precision 1: FP32 M2L on MFMA and FP32 near field with FP64 sums
  const bool mixed = o && o->precision == 1;
  if (o) {
    const int e = check_nterms_override(o->n_nterms, o->nterms, 1, kMaxNtermsHelmholtz);
    if (e != FMM2D_OK) return e;
  }
  try {
    double tinfo[16] = {0};
    Timer tm(o != nullptr, st);
    const int64_t ns = a.ns;
    const int64_t nt = (a.ifpghtarg > 0) ? a.nt : 0;
    const int nd = a.nd;
    const cplx zk = a.zk;
    int ndiv = lndiv2d(a.eps, ns, nt);  // hndiv2d (helmholtz/hndiv2d.f) is the same table
    int idivflag = 0;
    if (o && o->ndiv > 0) {
      ndiv = o->ndiv;
      idivflag = o->idivflag;
    }
    // ---- tree and lists
    Tree t;
    build_tree(a.src, ns, a.targ, nt, ndiv, idivflag, 51, t, st,
               o ? static_cast<const TreePrep*>(o->prepared) : nullptr);
    Lists L;
    build_lists(t, L, st);
    tm.mark(8);
    const int nl = t.nlevels, nb = t.nboxes;
    // ---- per-level orders, scales and layout (hfmm2d.f:317-326)
    std::vector<int> nterms(nl + 1);
    for (int l = 0; l <= nl; ++l) {
      int ier = 0;
      nterms[l] = h2dterms(t.boxsize[l], zk, a.eps, &ier);
      if (ier) return 13;
    }
    if (o && o->n_nterms > 0 && o->nterms)
      for (int l = 0; l <= nl; ++l) nterms[l] = o->nterms[std::min(l, o->n_nterms - 1)];
    const long double absk = std::hypot((long double)zk.re, (long double)zk.im);
    std::vector<HLev> lev(nl + 1);
    std::vector<long double> sl(nl + 1);
    int64_t ncoef = 0;
    for (int l = 0; l <= nl; ++l) {
      if (nterms[l] < 1 || nterms[l] > kMaxNtermsHelmholtz) return FMM2D_ERR_NTERMS;
      HLev& v = lev[l];
      sl[l] = std::min(absk * (long double)t.boxsize[l] / (2.0L * (long double)kPi), 1.0L);
      v.s = (double)sl[l];
      v.p = nterms[l];
      v.box0 = t.lstart[l];
      v.off = ncoef;
      v.on = (zk.im * t.boxsize[l] < kZkiUpBound) ? 1 : 0;
      const int nbl = ((l < nl) ? t.lstart[l + 1] : nb) - t.lstart[l];
      ncoef += (int64_t)nbl * nd * (2 * nterms[l] + 1);
    }
    DBuf<HLev> d_lev = upload(lev, st);

    // ---- sort and gather
    DBuf<double> xs(2 * ns, st), xt(2 * std::max<int64_t>(nt, 1), st);
    DBuf<cplx> chs(a.ifcharge ? ns * nd : 1, st), dps(a.ifdipole ? ns * nd : 1, st);
    DBuf<double> dvs(a.ifdipole ? 2 * ns * nd : 2, st);
    launch1d(k_gather_xy, ns, st, t.isrc.p, a.src, xs.p);
    if (nt > 0) launch1d(k_gather_xy, nt, st, t.itarg.p, a.targ, xt.p);
    if (a.ifcharge) launch1d(k_gather_dens, ns, st, nd, t.isrc.p, a.charge, chs.p);
    if (a.ifdipole) {
      launch1d(k_gather_dens, ns, st, nd, t.isrc.p, a.dipstr, dps.p);
      launch1d(k_gather_dipvec, ns, st, nd, t.isrc.p, a.dipvec, dvs.p);
    }
    DBuf<int> leaf_s(ns, st), leaf_t(std::max<int64_t>(nt, 1), st);
    launch1d(k_leaf_of, nb, st, t.nchild.p, t.isrcse.p, leaf_s.p);
    if (nt > 0) launch1d(k_leaf_of, nb, st, t.nchild.p, t.itargse.p, leaf_t.p);
    DBuf<cplx> mp(ncoef, st), loc(ncoef, st);
    mp.zero();
    loc.zero();
    const int pg = a.ifpgh, pgt = a.ifpghtarg;
    DBuf<cplx> ps(pg >= 1 ? ns * nd : 1, st), gs(pg >= 2 ? 2 * ns * nd : 1, st),
        hs(pg >= 3 ? 3 * ns * nd : 1, st);
    DBuf<cplx> pt(pgt >= 1 ? nt * nd : 1, st), gt(pgt >= 2 ? 2 * nt * nd : 1, st),
        ht(pgt >= 3 ? 3 * nt * nd : 1, st);
    ps.zero(); gs.zero(); hs.zero(); pt.zero(); gt.zero(); ht.zero();
    tm.mark(9);

    const int* sse = t.isrcse.p;
    const int* tse = t.itargse.p;
    // ---- step 1: P2M (rows of the sources, then their sums per leaf)
    {
      DBuf<int64_t> roff(ns, st), rtot(1, st);
      launch1d(k_hp2m_rowlen, ns, st, d_lev.p, leaf_s.p, t.ilevel.p, roff.p);
      exclusive_scan(roff.p, roff.p, ns, rtot.p, st);
      int64_t nrow = 0;
      readback(&nrow, rtot.p, sizeof(int64_t), st);
      DBuf<cplx> gJ(std::max<int64_t>(nrow, 1), st), uJ(ns, st);
      launch1d(k_hp2m_rows, ns, st, zk, d_lev.p, leaf_s.p, t.ilevel.p, t.centers.p, xs.p,
               (const int64_t*)roff.p, gJ.p, uJ.p);
#define FMM2D_P2M(C, D)                                                                       \
  hipLaunchKernelGGL((k_hp2m_g<C, D>), dim3(nb), dim3(kW), 0, st, nd, zk, d_lev.p, t.ilevel.p, \
                     t.nchild.p, sse, (const int64_t*)roff.p, gJ.p, uJ.p, chs.p, dps.p, dvs.p,  \
                     mp.p)
      if (a.ifcharge && !a.ifdipole) FMM2D_P2M(1, 0);
      else if (!a.ifcharge && a.ifdipole) FMM2D_P2M(0, 1);
      else FMM2D_P2M(1, 1);
#undef FMM2D_P2M
    }
    hip_throw(hipGetLastError(), "p2m");
    tm.mark(0);
    // ---- step 2: list 4 (pair counts, rows of the pairs, their sums per box)
    {
      DBuf<int> cnt4(nb, st), off4(nb, st), tot4(1, st);
      DBuf<int64_t> roff4(nb, st), rtot4(1, st);
      launch1d(k_hp2l_count, nb, st, d_lev.p, t.ilevel.p, L.n4.p, L.l4.p, sse, tse, pg, pgt,
               cnt4.p);
      exclusive_scan(cnt4.p, off4.p, nb, tot4.p, st);
      launch1d(k_hp2l_rowlen, nb, st, d_lev.p, t.ilevel.p, cnt4.p, roff4.p);
      exclusive_scan(roff4.p, roff4.p, nb, rtot4.p, st);
      int np4 = 0;
      int64_t nrow4 = 0;
      readback(&np4, tot4.p, sizeof(int), st);
      readback(&nrow4, rtot4.p, sizeof(int64_t), st);
      if (np4 > 0) {
        DBuf<cplx> gH(std::max<int64_t>(nrow4, 1), st), uH(np4, st);
        launch1d(k_hp2l_rows, np4, st, nb, zk, d_lev.p, t.ilevel.p, t.centers.p, L.l4.p, sse,
                 off4.p, xs.p, (const int64_t*)roff4.p, gH.p, uH.p);
#define FMM2D_P2L(C, D)                                                                       \
  hipLaunchKernelGGL((k_hp2l_g<C, D>), dim3(nb), dim3(kW), 0, st, nd, zk, d_lev.p, t.ilevel.p, \
                     L.n4.p, L.l4.p, sse, cnt4.p, off4.p, (const int64_t*)roff4.p, gH.p, uH.p, \
                     chs.p, dps.p, dvs.p, loc.p)
        if (a.ifcharge && !a.ifdipole) FMM2D_P2L(1, 0);
        else if (!a.ifcharge && a.ifdipole) FMM2D_P2L(0, 1);
        else FMM2D_P2L(1, 1);
#undef FMM2D_P2L
      }
    }
    hip_throw(hipGetLastError(), "p2l");
    tm.mark(1);

    helm_box_passes(t, L, lev, sl, nterms, nd, zk, pg, pgt, mp.p, loc.p, tm, st, mixed);
    // ---- steps 6+7: list 3 and L2P
    auto eval = [&](int p_, int64_t n, const double* xy, const int* lf, cplx* P_, cplx* G_,
                    cplx* H_) {
      if (p_ == 1) launch_heval<1>(n, st, nd, zk, d_lev.p, xy, lf, t, L, mp.p, loc.p, P_, G_, H_);
      else if (p_ == 2)
        launch_heval<2>(n, st, nd, zk, d_lev.p, xy, lf, t, L, mp.p, loc.p, P_, G_, H_);
      else launch_heval<3>(n, st, nd, zk, d_lev.p, xy, lf, t, L, mp.p, loc.p, P_, G_, H_);
    };
    if (pg > 0) eval(pg, ns, xs.p, leaf_s.p, ps.p, gs.p, hs.p);
    if (pgt > 0) eval(pgt, nt, xt.p, leaf_t.p, pt.p, gt.p, ht.p);
    tm.mark(6);
    // ---- step 8: near field, thresh = |zk| boxsize(0) 2^-51 (hfmm2d.f:1527)
    const int ifnear = o ? o->ifnear : 1;
    if (ifnear) {
      const double thresh = std::hypot(zk.re, zk.im) * t.boxsize[0] * std::ldexp(1.0, -51);
      DBuf<int2> seg((size_t)kSegRow * t.nboxes, st);
      launch1d(k_near_segs, t.nboxes, st, L.n1.p, L.l1.p, sse, seg.p);
      if (pg > 0)
        launch_hnear(pg, a.ifcharge, a.ifdipole, ns, st, nd, zk, xs.p, leaf_s.p, seg.p, xs.p,
                     chs.p, dps.p, dvs.p, thresh, ps.p, gs.p, hs.p, mixed);
      if (pgt > 0)
        launch_hnear(pgt, a.ifcharge, a.ifdipole, nt, st, nd, zk, xt.p, leaf_t.p, seg.p, xs.p,
                     chs.p, dps.p, dvs.p, thresh, pt.p, gt.p, ht.p, mixed);
    }
    tm.mark(7);
    // ---- unsort
    if (pg >= 1) launch1d(k_scatter, ns, st, nd, t.isrc.p, ps.p, a.pot);
    if (pg >= 2) launch1d(k_scatter, ns, st, 2 * nd, t.isrc.p, gs.p, a.grad);
    if (pg >= 3) launch1d(k_scatter, ns, st, 3 * nd, t.isrc.p, hs.p, a.hess);
    if (pgt >= 1) launch1d(k_scatter, nt, st, nd, t.itarg.p, pt.p, a.pottarg);
    if (pgt >= 2) launch1d(k_scatter, nt, st, 2 * nd, t.itarg.p, gt.p, a.gradtarg);
    if (pgt >= 3) launch1d(k_scatter, nt, st, 3 * nd, t.itarg.p, ht.p, a.hesstarg);
    tm.mark(10);
    std::vector<unsigned long long> hc(6 * (nl + 1), 0ull);
    const bool counting = o && o->count_work;
    if (counting) {
      DBuf<unsigned long long> dc(hc.size(), st);
      dc.zero();
      for (int l = 0; l <= nl; ++l) {
        const int b0 = t.lstart[l], b1 = (l < nl) ? t.lstart[l + 1] : nb;
        if (b1 > b0)
          launch1d(k_counts, b1 - b0, st, t.ilevel.p, t.iparent.p, t.nchild.p, t.ichild.p, sse,
                   tse, pg, pgt, L.n1.p, L.l1.p, L.n2.p, L.n3.p, L.n4.p, L.l4.p, b0,
                   dc.p + 6 * l);
      }
      hip_throw(hipMemcpyAsync(hc.data(), dc.p, sizeof(unsigned long long) * hc.size(),
                               hipMemcpyDeviceToHost, st), "copy");
    }
    hip_throw(hipStreamSynchronize(st), "sync");
    if (o) {
      fmm2d_opts* ow = const_cast<fmm2d_opts*>(o);
      tm.collect(tinfo);
      for (int kk = 0; kk < 16; ++kk) ow->timeinfo[kk] = tinfo[kk];
      ow->nlevels = nl;
      ow->nboxes = nb;
      if (counting) {
        // algorithmic flops: complex multiply-add = 8 flops; translations are
        // (2p+1)-square (M2L) or (2p_parent+1) x (2p_child+1) matrix-vector products
        const double ndd = (double)nd;
        double w[16] = {0};
        int64_t near = 0, m2l_pairs = 0;
        for (int l = 0; l <= nl; ++l) {
          const double Pl = 2.0 * nterms[l] + 1.0;
          const double Pc = (l < nl) ? 2.0 * nterms[l + 1] + 1.0 : Pl;
          const double Pp = (l > 0) ? 2.0 * nterms[l - 1] + 1.0 : Pl;
          const unsigned long long* c = &hc[6 * l];
          w[1] += 8.0 * Pl * ndd * (double)c[5];
          w[2] += 8.0 * Pl * Pc * ndd * (double)c[2];
          w[3] += 8.0 * Pl * Pl * ndd * (double)c[1];
          w[4] += 8.0 * Pl * Pp * ndd * (double)c[3];
          w[6] += 8.0 * Pl * ndd * (double)c[4];
          near += (int64_t)c[0];
          m2l_pairs += (int64_t)c[1];
        }
        // P2M and L2P at the finest order: a lower bound of their work
        w[0] = 8.0 * (2.0 * nterms[nl] + 1.0) * ndd * (double)ns;
        w[6] += 8.0 * (2.0 * nterms[nl] + 1.0) * ndd * (double)((pg > 0 ? ns : 0) + (pgt > 0 ? nt : 0));
        w[7] = 40.0 * ndd * (double)near;
        w[8] = 40.0 * (double)(ns + nt) + 80.0 * (double)nb;
        for (int kk = 0; kk < 16; ++kk) ow->work[kk] = w[kk];
        ow->near_pairs = near;
        ow->m2l_pairs = m2l_pairs;
      }
    }
    return FMM2D_OK;
  } catch (const HipError& e) {
    fprintf(stderr, "fmm2d_amd: %s failed: %s\n", e.what, hipGetErrorString(e.err));
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return FMM2D_ERR_ALLOC;
  } catch (...) {
    return FMM2D_ERR_HIP;
  }
}


int helmholtz_mps(const HelmMpsArgs& a, hipStream_t st) {
  if (a.nd < 1 || a.n < 1 || a.n > INT32_MAX || a.ntot < 1) return FMM2D_ERR_ARG;
  if (!a.centers || !a.rscale || !a.mterms || !a.mpole || !a.impole || !a.local)
    return FMM2D_ERR_ARG;
  if (a.zk.re == 0.0 && a.zk.im == 0.0) return FMM2D_ERR_ARG;
  const fmm2d_opts* o = a.opts;
  if (o && o->precision != 0) return FMM2D_ERR_UNSUPPORTED;
  if (o) {
    const int e = check_nterms_override(o->n_nterms, o->nterms, 1, kMaxNtermsHelmholtz);
    if (e != FMM2D_OK) return e;
  }
  try {
    double tinfo[16] = {0};
    Timer tm(o != nullptr, st);
    const int64_t n = a.n;
    const int nd = a.nd;
    const cplx zk = a.zk;
    // orders of the expansions (the LDS of the translation kernels holds
    // 2 (p1 + p2) + 2 coefficients)
    int pm[2] = {0, 0};
    {
      DBuf<int> d(2, st);
      const int init[2] = {INT32_MIN, INT32_MAX};
      upload_bytes(d.p, init, sizeof(init), st);
      launch1d(k_mps_pmax, n, st, a.mterms, d.p);
      readback(pm, d.p, sizeof(pm), st);
    }
    if (pm[1] < 0) return FMM2D_ERR_ARG;
    // ---- tree over the centres, one expansion per leaf by default
    // (hfmm2d_mps.f:100-150: idivflag 0, ndiv 1; hfmm2d_mps_ndiv.f: ndiv given)
    int ndiv = (o && o->ndiv > 0) ? o->ndiv : 1;
    Tree t;
    build_tree(a.centers, n, nullptr, 0, ndiv, 0, 51, t, st);
    Lists L;
    build_lists(t, L, st);
    tm.mark(8);
    const int nl = t.nlevels, nb = t.nboxes;
    // ---- per-level orders and scales (hfmm2d_mps.f:190-200)
    std::vector<int> nterms(nl + 1);
    for (int l = 0; l <= nl; ++l) {
      int ier = 0;
      nterms[l] = h2dterms(t.boxsize[l], zk, a.eps, &ier);
      if (ier) return 13;
    }
    if (o && o->n_nterms > 0 && o->nterms)
      for (int l = 0; l <= nl; ++l) nterms[l] = o->nterms[std::min(l, o->n_nterms - 1)];
    const long double absk = std::hypot((long double)zk.re, (long double)zk.im);
    std::vector<HLev> lev(nl + 1);
    std::vector<long double> sl(nl + 1);
    int64_t ncoef = 0;
    int pbox = 0;
    for (int l = 0; l <= nl; ++l) {
      if (nterms[l] < 1 || nterms[l] > kMaxNtermsHelmholtz) return FMM2D_ERR_NTERMS;
      HLev& v = lev[l];
      sl[l] = std::min(absk * (long double)t.boxsize[l] / (2.0L * (long double)kPi), 1.0L);
      v.s = (double)sl[l];
      v.p = nterms[l];
      v.box0 = t.lstart[l];
      v.off = ncoef;
      v.on = (zk.im * t.boxsize[l] < kZkiUpBound) ? 1 : 0;
      const int nbl = ((l < nl) ? t.lstart[l + 1] : nb) - t.lstart[l];
      ncoef += (int64_t)nbl * nd * (2 * nterms[l] + 1);
      pbox = std::max(pbox, nterms[l]);
    }
    const int nmax = 2 * std::max(pm[0], pbox);  // largest p1 + p2 of a translation
    const size_t lds = sizeof(cplx) * (2 * (size_t)nmax + 2);
    if (lds > kMpsLds) return FMM2D_ERR_UNSUPPORTED;
    DBuf<HLev> d_lev = upload(lev, st);
    DBuf<MpsExp> ex(n, st);
    launch1d(k_mps_gather, n, st, t.isrc.p, a.centers, a.rscale, a.mterms, a.impole, ex.p);
    DBuf<int> leaf(n, st);
    launch1d(k_leaf_of, nb, st, t.nchild.p, t.isrcse.p, leaf.p);
    DBuf<cplx> mp(ncoef, st), loc(ncoef, st);
    mp.zero();
    loc.zero();
    tm.mark(9);
    const int* se = t.isrcse.p;
    // ---- step 1: expansions into their leaves' multipoles
    hipLaunchKernelGGL(k_mps_p2m, dim3(nb), dim3(kMpsW), lds, st, nd, zk, d_lev.p, t.ilevel.p,
                       t.nchild.p, se, t.centers.p, ex.p, a.mpole, mp.p);
    hip_throw(hipGetLastError(), "mps p2m");
    tm.mark(0);
    // ---- step 2: list-4 expansions into box locals
    hipLaunchKernelGGL(k_mps_p2l, dim3(nb), dim3(kMpsW), lds, st, nd, zk, d_lev.p, t.ilevel.p,
                       L.n4.p, L.l4.p, se, t.centers.p, ex.p, a.mpole, loc.p);
    hip_throw(hipGetLastError(), "mps p2l");
    tm.mark(1);
    // ---- steps 3-5: M2M, M2L, L2L between boxes (every box holding an
    // expansion needs its local: pg = 1)
    helm_box_passes(t, L, lev, sl, nterms, nd, zk, 1, 0, mp.p, loc.p, tm, st);
    // ---- steps 6-7: list-3 multipoles and the leaf's local at every centre
    hipLaunchKernelGGL(k_mps_far, dim3((unsigned)n), dim3(kMpsW), lds, st, n, nd, zk, d_lev.p,
                       nl, leaf.p, t.ilevel.p, L.n3.p, L.l3.p, t.centers.p, ex.p, mp.p, loc.p,
                       a.local);
    hip_throw(hipGetLastError(), "mps far");
    tm.mark(6);
    // ---- step 8: list-1 expansions, thresh = |zk| boxsize(0) 2^-51 (hfmm2d_mps.f:962)
    const int ifnear = o ? o->ifnear : 1;
    if (ifnear) {
      const double thresh = std::hypot(zk.re, zk.im) * t.boxsize[0] * std::ldexp(1.0, -51);
      hipLaunchKernelGGL(k_mps_near, dim3((unsigned)n), dim3(kMpsW), lds, st, n, nd, zk, thresh,
                         leaf.p, L.n1.p, L.l1.p, se, ex.p, a.mpole, a.local);
      hip_throw(hipGetLastError(), "mps near");
    }
    tm.mark(7);
    hip_throw(hipStreamSynchronize(st), "sync");
    if (o) {
      fmm2d_opts* ow = const_cast<fmm2d_opts*>(o);
      tm.collect(tinfo);
      for (int kk = 0; kk < 16; ++kk) ow->timeinfo[kk] = tinfo[kk];
      ow->nlevels = nl;
      ow->nboxes = nb;
    }
    return FMM2D_OK;
  } catch (const HipError& e) {
    fprintf(stderr, "fmm2d_amd: %s failed: %s\n", e.what, hipGetErrorString(e.err));
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return FMM2D_ERR_ALLOC;
  } catch (...) {
    return FMM2D_ERR_HIP;
  }
}

}  // namespace fmm2d

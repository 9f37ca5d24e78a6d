// Pairwise source->target interaction math for the three kernel families.
//
// One call adds the contribution of ONE source to the accumulators of ONE
// target, for NDT densities at once.  The formulas are those the reference's
// direct kernels use, kept identical because they define the library's
// output convention (log|z| with no 1/(2 pi), Helmholtz scaled by i/4, Cauchy
// "grad" = d/dz):
//   Cauchy    laplace/cauchykernels2d.f:3-683   (c2d_direct{c,d,cd}{p,g,h})
//   Laplace   laplace/lapkernels2d.f:8-815      (l2d_direct{c,d,cd}{p,g,h})
//   Helmholtz helmholtz/helmkernels2d.f:3-904   (h2d_direct{c,d,cd}{p,g,h})
// including each variant's exclusion rule: a pair is skipped when
// |target - source| is below (or, for some variants, at) the threshold; see
// skip_pair().
#pragma once
#include "common.h"
#include "fastmath.h"
#include "hankel.h"

namespace fmm2d {

enum Family { CAUCHY = 0, LAPLACE = 1, HELM = 2 };

template <int FAM> struct FamDims {
  static constexpr int NG = (FAM == CAUCHY) ? 1 : 2;  // gradient components
  static constexpr int NH = (FAM == CAUCHY) ? 1 : 3;  // hessian components
};

// Accumulators for one target and NDT densities.
template <int FAM, int PG, int NDT> struct Acc {
  static constexpr int NG = FamDims<FAM>::NG, NH = FamDims<FAM>::NH;
  cplx pot[NDT];
  cplx grad[NDT][NG];
  cplx hess[NDT][NH];
  FMM2D_HD void zero() {
#pragma unroll
    for (int d = 0; d < NDT; ++d) {
      pot[d] = mkc(0.0);
#pragma unroll
      for (int c = 0; c < NG; ++c) grad[d][c] = mkc(0.0);
#pragma unroll
      for (int c = 0; c < NH; ++c) hess[d][c] = mkc(0.0);
    }
  }
};

// Source strengths for NDT densities.
template <int NDT> struct SrcStr {
  cplx charge[NDT];
  cplx dipstr[NDT];
  double dv[NDT][2];
};

// Exclusion rule of each reference variant.  Returns true if the pair must be
// skipped.  rr = |t-s|^2, absz = |zk|*|t-s| (Helmholtz) or |t-s| (Cauchy
// dipole-only).
template <int FAM, int IFC, int IFD, int PG>
FMM2D_HD bool skip_pair(double rr, double thresh, double thresh2, double absz) {
  if (FAM == CAUCHY) {
    const bool le = (PG == 1);
    if (IFC) return le ? (rr <= thresh2) : (rr < thresh2);
    return le ? (absz <= thresh) : (absz < thresh);
  } else if (FAM == LAPLACE) {
    const bool le = (PG == 1) || !IFC;
    return le ? (rr <= thresh2) : (rr < thresh2);
  } else {
    const bool le = (PG == 1) && IFC && !IFD;
    return le ? (absz <= thresh) : (absz < thresh);
  }
}

// ---------------------------------------------------------------- Cauchy
template <int IFC, int IFD, int PG, int NDT>
FMM2D_HD void pair_cauchy(double dx, double dy, const SrcStr<NDT>& s, double thresh,
                          double thresh2, Acc<CAUCHY, PG, NDT>& a) {
  double rr = dx * dx + dy * dy;
  double absz = 0.0;
  if (!IFC) absz = hypot(dx, dy);
  if (skip_pair<CAUCHY, IFC, IFD, PG>(rr, thresh, thresh2, absz)) return;
  const double rinv = fast_rcp(rr);
  cplx zinv = cplx{dx * rinv, -dy * rinv};  // 1/z
  cplx z2 = -(zinv * zinv);                  // -1/z^2
  cplx z3 = (-2.0) * (z2 * zinv);             // 2/z^3
  double lg = 0.0;
  if (IFC) lg = 0.5 * fast_log(rr);
#pragma unroll
  for (int d = 0; d < NDT; ++d) {
    if (IFC) {
      a.pot[d] += lg * s.charge[d];
      if (PG >= 2) a.grad[d][0] += zinv * s.charge[d];
      if (PG >= 3) a.hess[d][0] += z2 * s.charge[d];
    }
    if (IFD) {
      a.pot[d] += zinv * s.dipstr[d];
      if (PG >= 2) a.grad[d][0] += z2 * s.dipstr[d];
      if (PG >= 3) a.hess[d][0] += z3 * s.dipstr[d];
    }
  }
}

// ---------------------------------------------------------------- Laplace
template <int IFC, int IFD, int PG, int NDT>
FMM2D_HD void pair_laplace(double dx, double dy, const SrcStr<NDT>& s, double thresh,
                           double thresh2, Acc<LAPLACE, PG, NDT>& a) {
  double dx2 = dx * dx, dy2 = dy * dy;
  double rr = dx2 + dy2;
  if (skip_pair<LAPLACE, IFC, IFD, PG>(rr, thresh, thresh2, 0.0)) return;
  double rinv = fast_rcp(rr);
  double rinv2 = rinv * rinv;
  double px = dx * rinv, py = dy * rinv;  // charge gradient kernel
  double hxx = (rr - 2.0 * dx2) * rinv2, hxy = -2.0 * dx * dy * rinv2, hyy = (rr - 2.0 * dy2) * rinv2;
  double lg = 0.0;
  if (IFC) lg = 0.5 * fast_log(rr);
  // dipole third-derivative kernel
  double rinv3 = rinv2 * rinv;
  double t1 = -2.0 * dx * (dx2 - 3.0 * dy2) * rinv3;
  double t2 = 2.0 * dy * (dy2 - 3.0 * dx2) * rinv3;
#pragma unroll
  for (int d = 0; d < NDT; ++d) {
    if (IFC) {
      const cplx c = s.charge[d];
      a.pot[d] += lg * c;
      if (PG >= 2) { a.grad[d][0] += px * c; a.grad[d][1] += py * c; }
      if (PG >= 3) { a.hess[d][0] += hxx * c; a.hess[d][1] += hxy * c; a.hess[d][2] += hyy * c; }
    }
    if (IFD) {
      const cplx d1 = s.dipstr[d] * s.dv[d][0];
      const cplx d2 = s.dipstr[d] * s.dv[d][1];
      a.pot[d] += (-px) * d1 + (-py) * d2;
      if (PG >= 2) {
        // d/dx, d/dy of -(v . r)/|r|^2
        a.grad[d][0] += (-hxx) * d1 + (-hxy) * d2;
        a.grad[d][1] += (-hxy) * d1 + (-hyy) * d2;
      }
      if (PG >= 3) {
        a.hess[d][0] += t1 * d1 + t2 * d2;
        a.hess[d][1] += t2 * d1 + (-t1) * d2;
        a.hess[d][2] += (-t1) * d1 + (-t2) * d2;
      }
    }
  }
}

// ---------------------------------------------------------------- Helmholtz
// OOL: evaluate the Hankel pair by a call to the general evaluator (for rarely
// taken paths: their kernels then reserve neither registers nor scalar
// registers for the table-driven real path).
template <int IFC, int IFD, int PG, int NDT, bool OOL = false>
FMM2D_HD void pair_helm(double dx, double dy, cplx zk, const SrcStr<NDT>& s, double thresh,
                        Acc<HELM, PG, NDT>& a) {
  double rr = dx * dx + dy * dy;
  double r = sqrt(rr);
  cplx z = zk * r;
  // |z| of a real wavenumber is |Re z| exactly (what hypot returns for it)
  const double absz = (zk.im == 0.0) ? fabs(z.re) : absc(z);
  if (skip_pair<HELM, IFC, IFD, PG>(rr, thresh, 0.0, absz)) return;
  cplx h0, h1;
  if constexpr (OOL)
    hank01_complex(z, &h0, &h1);
  else
    hank01(z, &h0, &h1);
  const cplx ik4 = cplx{0.0, 0.25};
  const double rinv = fast_rcp(r);
  const cplx kq = zk * ik4;  // i k / 4
  if (IFC) {
    const cplx p0 = h0 * ik4;
    const cplx cdd = -(h1 * kq) * rinv;  // -(ik/4) H1 / r
    cplx hf1, hf2, hf3, cdd2;
    if (PG >= 3) {
      cdd2 = kq * (rinv / rr);
      const cplx h2z = 2.0 * h1 - z * h0;
      hf1 = h2z * (dx * dx) - rr * h1;
      hf2 = h2z * (dx * dy);
      hf3 = h2z * (dy * dy) - rr * h1;
      hf1 = cdd2 * hf1; hf2 = cdd2 * hf2; hf3 = cdd2 * hf3;
    }
#pragma unroll
    for (int d = 0; d < NDT; ++d) {
      const cplx c = s.charge[d];
      a.pot[d] += p0 * c;
      if (PG >= 2) {
        const cplx cd = cdd * c;
        a.grad[d][0] += cd * dx;
        a.grad[d][1] += cd * dy;
      }
      if (PG >= 3) {
        a.hess[d][0] += hf1 * c;
        a.hess[d][1] += hf2 * c;
        a.hess[d][2] += hf3 * c;
      }
    }
  }
  if (IFD) {
    const double ct = dx * rinv, st = dy * rinv;
    const cplx cdd = h1 * kq * rinv;  // (ik/4) H1 / r
    cplx h2, h3, hf1, hf2, hf3, cdd2, cdd3, q1, q2, q3, q4;
    if (PG >= 2) {
      const cplx zi = invc(z);
      h2 = 2.0 * (h1 * zi) - h0;
      cdd2 = -(zk * zk) * ik4;
      hf1 = h2 * (ct * ct - 0.5) - 0.5 * h0;
      hf2 = h2 * (ct * st);
      hf3 = h2 * (st * st - 0.5) - 0.5 * h0;
      if (PG >= 3) {
        h3 = 4.0 * (h2 * zi) - h1;
        cdd3 = -(zk * zk * zk) * ik4;
        const cplx mh1 = -0.5 * h1, mh3 = -0.5 * h3;
        q1 = (mh1 * (-1.5) + mh3 * (ct * ct - 0.5 - st * st)) * ct;
        q2 = (mh1 * (-0.5) + mh3 * (1.5 * ct * ct - 0.5 * st * st)) * st;
        q3 = (mh1 * (-0.5) + mh3 * (1.5 * st * st - 0.5 * ct * ct)) * ct;
        q4 = (mh1 * (-1.5) + mh3 * (st * st - 0.5 - ct * ct)) * st;
      }
    }
#pragma unroll
    for (int d = 0; d < NDT; ++d) {
      const double v1 = s.dv[d][0], v2 = s.dv[d][1];
      a.pot[d] += (cdd * s.dipstr[d]) * (dx * v1 + dy * v2);
      if (PG >= 2) {
        const cplx cd = cdd2 * s.dipstr[d];
        const cplx hxx = cd * hf1, hxy = cd * hf2, hyy = cd * hf3;
        a.grad[d][0] += hxx * v1 + hxy * v2;
        a.grad[d][1] += hxy * v1 + hyy * v2;
      }
      if (PG >= 3) {
        const cplx cd = cdd3 * s.dipstr[d];
        a.hess[d][0] += cd * (q1 * v1 + q2 * v2);
        a.hess[d][1] += cd * (q2 * v1 + q3 * v2);
        a.hess[d][2] += cd * (q3 * v1 + q4 * v2);
      }
    }
  }
}

}  // namespace fmm2d

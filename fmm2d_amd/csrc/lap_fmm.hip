// Laplace / Cauchy-kernel FMM on the device (FP64).
//
// Reference: laplace/cfmm2d.f (driver 23-436, cfmm2dmain 442-1341),
// laplace/lfmm2d.f (complex-density split 103-223), laplace/laprouts2d.f
// (expansions and translations).  Expansions are the reference's: multipole
// M_0 log(z) + sum M_n (r/z)^n and local sum L_n (z/r)^n about each box
// centre, scaled by the box size r (rscale = boxsize of the level).
//
// Work decomposition (one launch per step, or per level where the reference's
// loops carry a level dependency):
//   P2M        one 64-lane wavefront per leaf, lane = coefficient index
//   M2M / L2L  one wavefront per box, fixed 4 child-offset matrices
//   M2L        one wavefront per target box, lane = output coefficient; the
//              translation depends only on the offset (dx,dy) in [-3,3]^2 of
//              the source box (in box sizes) and is precomputed once
//              (49 (p+1)x(p+1) matrices; the level enters only through
//              log(boxsize) M_0 added to L_0)
//   list 4     one wavefront per box (sources of coarser leaves -> local)
//   list 3+L2P one thread per evaluation point
//   near field one thread per evaluation point over its leaf's list 1
// All kernels read and write the box expansions in place; there are no
// atomics in the FMM itself (pull form, as the reference's loops).
#include <algorithm>
#include <complex>
#include <map>
#include <mutex>
#include <vector>


#include "devbuf.h"
#include "fmm_common.h"
#include "lap_ops.h"
#include "p2p_math.h"
#include "scan.h"
#include "translate.h"
#include "tree.h"

namespace fmm2d {

// --------------------------------------------------------------- rules
int lndiv2d(double eps, int64_t ns, int64_t nt) {
  if (eps >= 0.5e-0) return 3;
  if (eps >= 0.5e-1) return 5;
  if (eps >= 0.5e-2) return 8;
  if (eps >= 0.5e-3) return 10;
  if (eps >= 0.5e-6) return 15;
  if (eps >= 0.5e-9) return 20;
  if (eps >= 0.5e-12) return 25;
  if (eps >= 0.5e-15) return 45;
  return (int)(ns + nt);
}

int l2dterms(double eps) {
  // first j >= 2 with (sqrt(2)/2)^j / 1.5^(j+1) < eps (j <= 100), else 1
  const double a = std::sqrt(2.0) / 2.0;
  for (int j = 2; j <= 100; ++j)
    if (std::pow(a, j) / std::pow(1.5, j + 1) < eps) return j;
  return 1;
}

namespace {

// lfmm2d split (laplace/lfmm2d.f:139-156): density d of the Laplace problem
// becomes Cauchy densities 2d (real part) and 2d+1 (imaginary part); a dipole
// (d, v) becomes -(v1 + i v2) times each part.  With half = 1 every density
// has zero imaginary part, the imaginary Cauchy densities vanish identically
// and only the real parts are kept (Cauchy density d).
// (with xy given, the point coordinates are gathered into oxy in the same pass)
// inv (if given) receives the inverse permutation: inv[perm[i]] = i.
__global__ void k_gather_split(int64_t n, int nd, int half, const int* perm, const cplx* ch,
                               const cplx* ds, const double* dv, cplx* och, cplx* ods,
                               const double* xy, double* oxy, int* inv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t j = perm[i];
  if (inv) inv[j] = (int)i;
  if (xy) {
    oxy[2 * i] = xy[2 * j];
    oxy[2 * i + 1] = xy[2 * j + 1];
  }
  const int np = half ? 1 : 2;
  for (int d = 0; d < nd; ++d) {
    const int64_t a = i * np * nd + np * d;
    if (ch) {
      const cplx c = ch[j * nd + d];
      och[a] = mkc(c.re);
      if (!half) och[a + 1] = mkc(c.im);
    }
    if (ds) {
      const cplx s = ds[j * nd + d];
      const cplx v = cplx{-dv[nd * (2 * j) + d], -dv[nd * (2 * j + 1) + d]};
      ods[a] = s.re * v;
      if (!half) ods[a + 1] = s.im * v;
    }
  }
}

// rfmm2d split (laplace/rfmm2d.f:65-84): real charge c -> Cauchy charge c,
// real dipole (d, v) -> Cauchy dipole -(v1 + i v2) d.
__global__ void k_gather_real(int64_t n, int nd, const int* perm, const double* ch,
                              const double* ds, const double* dv, cplx* och, cplx* ods) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t j = perm[i];
  for (int d = 0; d < nd; ++d) {
    if (ch) och[i * nd + d] = mkc(ch[j * nd + d]);
    if (ds) {
      const double s = ds[j * nd + d];
      ods[i * nd + d] = cplx{-dv[nd * (2 * j) + d] * s, -dv[nd * (2 * j + 1) + d] * s};
    }
  }
}

// rfmm2d unpack (laplace/rfmm2d.f:86-139): pot = Re, grad = (Re, -Im),
// hess = (Re, -Im, -Re) of the Cauchy outputs, into real arrays.
__global__ void k_scatter_real(int64_t n, int nd, int pg, const int* perm, const cplx* p1,
                               const cplx* g1, const cplx* h1, double* pot, double* grad,
                               double* hess) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t j = perm[i];
  for (int d = 0; d < nd; ++d) {
    const int64_t a = i * nd + d;
    pot[j * nd + d] = p1[a].re;
    if (pg >= 2) {
      grad[nd * (2 * j) + d] = g1[a].re;
      grad[nd * (2 * j + 1) + d] = -g1[a].im;
    }
    if (pg >= 3) {
      const double hxx = h1[a].re;
      hess[nd * (3 * j) + d] = hxx;
      hess[nd * (3 * j + 1) + d] = -h1[a].im;
      hess[nd * (3 * j + 2) + d] = -hxx;
    }
  }
}

// flag = 1 if some value has a non-zero (or NaN) imaginary part
__global__ void k_any_imag(int64_t n, const cplx* v, int* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double im = v[i].im;
  if (!(im == 0.0)) *flag = 1;
}

// lfmm2d unpack (laplace/lfmm2d.f:169-223).  in: Cauchy outputs of the split
// densities; pot(nd,n), grad(nd,2,n), hess(nd,3,n).  With half = 1 the
// imaginary densities' outputs are the exact zeros they would sum to.
__global__ void k_scatter_unpack(int64_t n, int nd, int half, int pg, const int* perm,
                                 const cplx* p1, const cplx* g1, const cplx* h1, cplx* pot,
                                 cplx* grad, cplx* hess) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t j = perm[i];
  const int np = half ? 1 : 2;
  const cplx z = mkc(0.0);
  for (int d = 0; d < nd; ++d) {
    const int64_t a = i * np * nd + np * d;
    const cplx pr = p1[a], pi = half ? z : p1[a + 1];
    pot[j * nd + d] = cplx{pr.re, pi.re};
    if (pg >= 2) {
      const cplx gr = g1[a], gi = half ? z : g1[a + 1];
      grad[nd * (2 * j) + d] = cplx{gr.re, gi.re};
      grad[nd * (2 * j + 1) + d] = cplx{-gr.im, -gi.im};
    }
    if (pg >= 3) {
      const cplx hr = h1[a], hi = half ? z : h1[a + 1];
      const cplx hxx = cplx{hr.re, hi.re};
      hess[nd * (3 * j) + d] = hxx;
      hess[nd * (3 * j + 1) + d] = cplx{-hr.im, -hi.im};
      hess[nd * (3 * j + 2) + d] = -hxx;
    }
  }
}

// k_scatter_unpack as a gather: thread k takes caller's point k from sorted
// position inv[k], so the outputs are written in the caller's order
// (consecutive threads, consecutive addresses) and the sorted outputs are read
// at scattered positions.  Same values as k_scatter_unpack.
__global__ void k_gather_unpack(int64_t n, int nd, int half, int pg, const int* inv,
                                const cplx* p1, const cplx* g1, const cplx* h1, cplx* pot,
                                cplx* grad, cplx* hess) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int64_t i = inv[j];
  const int np = half ? 1 : 2;
  const cplx z = mkc(0.0);
  for (int d = 0; d < nd; ++d) {
    const int64_t a = i * np * nd + np * d;
    const cplx pr = p1[a], pi = half ? z : p1[a + 1];
    pot[j * nd + d] = cplx{pr.re, pi.re};
    if (pg >= 2) {
      const cplx gr = g1[a], gi = half ? z : g1[a + 1];
      grad[nd * (2 * j) + d] = cplx{gr.re, gi.re};
      grad[nd * (2 * j + 1) + d] = cplx{-gr.im, -gi.im};
    }
    if (pg >= 3) {
      const cplx hr = h1[a], hi = half ? z : h1[a + 1];
      const cplx hxx = cplx{hr.re, hi.re};
      hess[nd * (3 * j) + d] = hxx;
      hess[nd * (3 * j + 1) + d] = cplx{-hr.im, -hi.im};
      hess[nd * (3 * j + 2) + d] = -hxx;
    }
  }
}

// ---------------------------------------------------------- P2M
// l2dformmp{c,d,cd} (laprouts2d.f:84-282) on leaves at levels >= 2.  One
// wavefront per leaf; sources go through LDS in batches of up to 64: lane j builds
// the power table w_j^n (w = (z_j - c)/r) by running products, then lane
// (n, d) sums  M_n += sum_j [-w_j^n/n c_j + w_j^(n-1)/r d_j]  (M_0 += c_j).
template <int IFC, int IFD>
__global__ __launch_bounds__(kW) void k_p2m(int nd, int P, const int* ilevel, const int* nchild,
                                            const int* nterms, const double* rsc,
                                            const double* centers, const int* sse,
                                            const double* xs, const cplx* chs, const cplx* dps,
                                            int bat, cplx* mp) {
  extern __shared__ cplx sh[];
  cplx* W = sh;                                    // [bat][P]
  int* J = reinterpret_cast<int*>(sh + bat * P);   // [bat]
  const int b = blockIdx.x;
  const int lev = ilevel[b];
  if (lev < 2 || nchild[b] != 0) return;
  const int s0 = sse[2 * b], s1 = sse[2 * b + 1];
  if (s1 <= s0) return;
  const int p = nterms[lev];
  const double r = rsc[lev], rinv = 1.0 / r;
  const double cx = centers[2 * b], cy = centers[2 * b + 1];
  const int lane = threadIdx.x;
  for (int j0 = s0; j0 < s1; j0 += bat) {
    const int cnt = min(bat, s1 - j0);
    if (lane < cnt) {
      const int j = j0 + lane;
      const cplx w = cplx{(xs[2 * j] - cx) * rinv, (xs[2 * j + 1] - cy) * rinv};
      cplx* row = W + lane * P;
      cplx v = mkc(1.0);
      row[0] = v;
      for (int n = 1; n <= p; ++n) {
        v = v * w;
        row[n] = v;
      }
      J[lane] = j;
    }
    __syncthreads();
    for (int idx = lane; idx < (p + 1) * nd; idx += kW) {
      const int n = idx / nd, d = idx - n * nd;
      const double cn = (n == 0) ? 0.0 : -1.0 / n;
      cplx acc = mkc(0.0);
      for (int jj = 0; jj < cnt; ++jj) {
        const int64_t j = J[jj];
        if (IFC) {
          const cplx c = chs[j * nd + d];
          acc += (n == 0) ? c : cn * (W[jj * P + n] * c);
        }
        if (IFD && n > 0) acc += rinv * (W[jj * P + n - 1] * dps[j * nd + d]);
      }
      mp[((int64_t)b * nd + d) * P + n] += acc;
    }
    __syncthreads();
  }
}

// l2dformta{c,d,cd} (laprouts2d.f:501-690): sources of the coarser leaves in
// list 4 of a box.  One wavefront per box; the list-4 sources are taken as one
// flat sequence in batches of up to 64: lane j stores u_j^n (u = r/z_j, z_j = source
// - centre), 1/z_j and log|z_j| in LDS, then lane (n, d) sums
//   L_0 += log|z_j| c_j,  L_n += -u_j^n/n c_j  (n >= 1),  L_n += -u_j^n/z_j d_j.
template <int IFC, int IFD>
__global__ __launch_bounds__(kW) void k_p2l(int nd, int P, const int* ilevel, const int* nterms,
                                            const double* rsc, const double* centers,
                                            const int* n4, const int* l4, const int* sse,
                                            const int* tse, int ifpgh, int ifpght,
                                            const double* xs, const cplx* chs, const cplx* dps,
                                            int bat, cplx* loc) {
  extern __shared__ cplx sh[];
  cplx* W = sh;                                        // [bat][P]
  cplx* ZI = sh + bat * P;                             // [bat]
  double* LG = reinterpret_cast<double*>(ZI + bat);    // [bat]
  int* J = reinterpret_cast<int*>(LG + bat);           // [bat]
  const int b = blockIdx.x;
  const int lev = ilevel[b];
  const int nbox = n4[b];
  if (lev < 2 || nbox == 0 || npts_eval(b, sse, tse, ifpgh, ifpght) == 0) return;
  const int p = nterms[lev];
  const double r = rsc[lev];
  const double cx = centers[2 * b], cy = centers[2 * b + 1];
  const int lane = threadIdx.x;
  // flat numbering of the list-4 sources: box k holds [pre[k], pre[k+1])
  int pre[kMaxList4 + 1], first[kMaxList4];
  pre[0] = 0;
  for (int k = 0; k < kMaxList4; ++k) {
    int c = -1;
    if (k < nbox) c = l4[(int64_t)kMaxList4 * b + k];
    first[k] = (c >= 0) ? sse[2 * c] : 0;
    pre[k + 1] = pre[k] + ((c >= 0) ? sse[2 * c + 1] - sse[2 * c] : 0);
  }
  const int total = pre[kMaxList4];
  for (int f0 = 0; f0 < total; f0 += bat) {
    const int cnt = min(bat, total - f0);
    if (lane < cnt) {
      const int f = f0 + lane;
      int k = 0;
      while (f >= pre[k + 1]) ++k;
      const int j = first[k] + (f - pre[k]);
      const cplx z0 = cplx{xs[2 * j] - cx, xs[2 * j + 1] - cy};
      const cplx zi = invc(z0);
      const cplx u = r * zi;
      cplx* row = W + lane * P;
      cplx v = mkc(1.0);
      row[0] = v;
      for (int n = 1; n <= p; ++n) {
        v = v * u;
        row[n] = v;
      }
      ZI[lane] = zi;
      LG[lane] = log(absc(z0));
      J[lane] = j;
    }
    __syncthreads();
    for (int idx = lane; idx < (p + 1) * nd; idx += kW) {
      const int n = idx / nd, d = idx - n * nd;
      const double cn = (n == 0) ? 0.0 : -1.0 / n;
      cplx acc = mkc(0.0);
      for (int jj = 0; jj < cnt; ++jj) {
        const int64_t j = J[jj];
        if (IFC) {
          const cplx c = chs[j * nd + d];
          acc += (n == 0) ? LG[jj] * c : cn * (W[jj * P + n] * c);
        }
        if (IFD) acc -= (ZI[jj] * W[jj * P + n]) * dps[j * nd + d];
      }
      loc[((int64_t)b * nd + d) * P + n] += acc;
    }
    __syncthreads();
  }
}

// Boxes with list-4 work (levels >= 2, a non-empty list 4, points to
// evaluate): flag, then (after an exclusive scan of the flags) the list in
// box order.
__global__ void k_p2l_flag(int64_t nb, const int* ilevel, const int* n4, const int* sse,
                           const int* tse, int ifpgh, int ifpght, int* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int b = (int)i;
  flag[b] = (ilevel[b] >= 2 && n4[b] > 0 && npts_eval(b, sse, tse, ifpgh, ifpght) > 0) ? 1 : 0;
}

__global__ void k_p2l_list(int64_t nb, const int* flag, const int* pos, int* list, int* cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  if (flag[i]) list[pos[i]] = (int)i;
  if (i == nb - 1) *cnt = pos[i] + flag[i];
}

// P2M (l2dformmp{c,d,cd}, laprouts2d.f:84-282) for expansions of at most PT
// coefficients: one thread per (box, density), for the leaves at levels >= 2,
// runs over the leaf's sources with the running products w^n, w = (z - c)/r, accumulating
// every coefficient in registers.  RE: real charges.
template <int PT, int IFC, int IFD, int RE>
__global__ __launch_bounds__(256) void k_p2m_reg(int64_t n, int nd, int P, const int* ilevel,
                                                 const int* nchild, const int* nterms,
                                                 const double* rsc, const double* centers,
                                                 const int* sse, const double* xs,
                                                 const cplx* chs, const cplx* dps, cplx* mp) {
  const int64_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = (int)(i / nd), d = (int)(i - (int64_t)b * nd);
  const int lev = ilevel[b];
  if (lev < 2 || nchild[b] != 0) return;
  const int s0 = sse[2 * b], s1 = sse[2 * b + 1];
  if (s1 <= s0) return;
  const int p = nterms[lev];
  const double rinv = 1.0 / rsc[lev];
  const double cx = centers[2 * b], cy = centers[2 * b + 1];
  cplx acc[PT];
#pragma unroll
  for (int k = 0; k < PT; ++k) acc[k] = mkc(0.0);
  // U sources at a time: U independent power chains in flight (4 while the
  // coefficients leave the registers for them, else 2)
  constexpr int U = PT <= 24 ? 4 : 2;
  for (int j = s0; j < s1; j += U) {
    cplx w[U], c[U], e[U], v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool ok = j + u < s1;
      const int jj = ok ? j + u : j;
      w[u] = cplx{(xs[2 * jj] - cx) * rinv, (xs[2 * jj + 1] - cy) * rinv};
      c[u] = (IFC && ok) ? chs[(int64_t)jj * nd + d] : mkc(0.0);
      e[u] = (IFD && ok) ? rinv * dps[(int64_t)jj * nd + d] : mkc(0.0);
      v[u] = mkc(1.0);  // w^(k-1)
    }
    if (IFC) {
      cplx s = c[0];
#pragma unroll
      for (int u = 1; u < U; ++u) s += c[u];
      acc[0] += s;
    }
#pragma unroll
    for (int k = 1; k < PT; ++k) {
      if (k <= p) {
        if (IFD) {
          cplx s = v[0] * e[0];
#pragma unroll
          for (int u = 1; u < U; ++u) s += v[u] * e[u];
          acc[k] += s;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = v[u] * w[u];
        if (IFC) {
          const double ck = -1.0 / k;
          if (RE) {
            double sr = v[0].re * c[0].re, si = v[0].im * c[0].re;
#pragma unroll
            for (int u = 1; u < U; ++u) {
              sr += v[u].re * c[u].re;
              si += v[u].im * c[u].re;
            }
            acc[k] += cplx{ck * sr, ck * si};
          } else {
            cplx s = v[0] * c[0];
#pragma unroll
            for (int u = 1; u < U; ++u) s += v[u] * c[u];
            acc[k] += ck * s;
          }
        }
      }
    }
  }
  cplx* out = mp + ((int64_t)b * nd + d) * P;
#pragma unroll
  for (int k = 0; k < PT; ++k)
    if (k <= p) out[k] += acc[k];
}

// Orders one wavefront's LDS writes before its later LDS reads (the
// wavefront's LDS accesses complete in order; this keeps the compiler from
// moving them across).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// FP32 complex of the mixed-precision P2L and evaluation kernels
struct fcx {
  float re, im;
};
__device__ __forceinline__ fcx fmulc(fcx a, fcx b) {
  return fcx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ fcx faddc(fcx a, fcx b) { return fcx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ fcx fscl(float s, fcx a) { return fcx{s * a.re, s * a.im}; }
__device__ __forceinline__ cplx todbl(fcx a) { return cplx{(double)a.re, (double)a.im}; }

// log|z| from |z|^2 in FP32 (hardware log2)
__device__ __forceinline__ float half_log_f32(float rr) {
  return 0.5f * __builtin_amdgcn_logf(rr) * 0.69314718055994531f;
}

// List-4 P2L (l2dformta{c,d,cd}, laprouts2d.f:501-690): the sources of a
// box's list-4 boxes into its local expansion.  Each wavefront takes boxes of
// `boxes` in a grid-stride loop (4 wavefronts per block, each with its own
// LDS rows).  A box's sources are one flat sequence taken in batches of up to
// `bat`: lane j writes row j, source j's terms S_j[n], n = 0..p, and lane
// (g, n) of the G = floor(64/(p+1)) lane groups adds rows g, g+G, ... of
// column n.  With u = r/(z - c) and e = -d/(z - c) (q: charge, d: Cauchy
// dipole strength)
//   T_0 = log|z - c| q + e,   T_n = -(1/n) (u^n q - n u^n e),
// the 1/n applied once to the column sum.  RE: real charges.  F32 (mixed
// precision): the terms in FP32 (displacement formed in FP64 and rounded
// once), rows of FP32 in LDS, each batch's column sums added to FP64 sums.
constexpr int kP2lWaves = 4;

template <int IFC, int IFD, int RE, int F32>
__global__ __launch_bounds__(kP2lWaves * kW) void k_p2l_w(
    const int* boxes, const int* nboxes, int nd, int P, const int* ilevel, const int* nterms,
    const double* rsc, const double* centers, const int* n4, const int* l4, const int* sse,
    const double* xs, const cplx* chs, const cplx* dps, int bat, cplx* loc) {
  extern __shared__ cplx shall[];  // [kP2lWaves][bat][P]
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  cplx* sh = shall + (size_t)wv * bat * P;
  float2* shf = reinterpret_cast<float2*>(sh);  // F32 rows, in the wave's own part
  const int nbx = *nboxes;
  for (int ib = blockIdx.x * kP2lWaves + wv; ib < nbx; ib += gridDim.x * kP2lWaves) {
    const int b = boxes[ib];
    const int lev = ilevel[b];
    // flat numbering of the sources: list-4 box k holds [pre[k], pre[k+1])
    int pre[kMaxList4 + 1], first[kMaxList4];
    pre[0] = 0;
    const int nbox = n4[b];
    for (int k = 0; k < kMaxList4; ++k) {
      const int c = (k < nbox) ? l4[(int64_t)kMaxList4 * b + k] : -1;
      first[k] = (c >= 0) ? sse[2 * c] : 0;
      pre[k + 1] = pre[k] + ((c >= 0) ? sse[2 * c + 1] - sse[2 * c] : 0);
    }
    const int total = pre[kMaxList4];
    const int p = nterms[lev], np1 = p + 1;
    const int G = min(kW / np1, bat);  // group sums use rows 0..G-1
    const int g = lane / np1, n = lane - g * np1;
    const bool act = g < G;
    const double r = rsc[lev];
    const double cx = centers[2 * b], cy = centers[2 * b + 1];
    for (int d = 0; d < nd; ++d) {
      cplx acc = mkc(0.0);
      for (int f0 = 0; f0 < total; f0 += bat) {
        const int cnt = min(bat, total - f0);
        if (lane < cnt) {
          const int f = f0 + lane;
          int k = 0;
          while (k + 1 < kMaxList4 && f >= pre[k + 1]) ++k;
          const int64_t j = first[k] + (f - pre[k]);
          cplx q = mkc(0.0), e = mkc(0.0);
          if (IFC) {
            q = chs[j * nd + d];
            if (RE) q.im = 0.0;
          }
          // z0 = source - centre of a well-separated box (no overflow guards)
          const double zx = xs[2 * j] - cx, zy = xs[2 * j + 1] - cy;
          if (F32) {
            float2* row = shf + lane * P;
            const float fx = (float)zx, fy = (float)zy;
            const float frr = fx * fx + fy * fy, fri = __builtin_amdgcn_rcpf(frr);
            const fcx zi = fcx{fx * fri, -fy * fri};
            const fcx u = fscl((float)r, zi);
            const fcx qf = fcx{(float)q.re, (float)q.im};
            fcx ef = fcx{0.f, 0.f};
            if (IFD) {
              const cplx dq = dps[j * nd + d];
              ef = fscl(-1.f, fmulc(zi, fcx{(float)dq.re, (float)dq.im}));
            }
            fcx t0 = ef;
            if (IFC) t0 = faddc(t0, fscl(half_log_f32(frr), qf));
            row[0] = float2{t0.re, t0.im};
            fcx v = fcx{1.f, 0.f};  // u^m
            for (int m = 1; m <= p; ++m) {
              v = fmulc(v, u);
              fcx s = fcx{0.f, 0.f};
              if (IFC) s = RE ? fcx{v.re * qf.re, v.im * qf.re} : fmulc(v, qf);
              if (IFD) s = faddc(s, fscl(-(float)m, fmulc(v, ef)));
              row[m] = float2{s.re, s.im};
            }
          } else {
            cplx* row = sh + lane * P;
            const double rr = zx * zx + zy * zy, ri = fast_rcp(rr);
            const cplx zi = cplx{zx * ri, -zy * ri};
            const cplx u = r * zi;
            if (IFD) e = -(zi * dps[j * nd + d]);
            cplx t0 = e;
            if (IFC) t0 += (0.5 * fast_log(rr)) * q;
            row[0] = t0;
            cplx v = mkc(1.0);  // u^m
            for (int m = 1; m <= p; ++m) {
              v = v * u;
              cplx s = mkc(0.0);
              if (IFC) s = RE ? cplx{v.re * q.re, v.im * q.re} : v * q;
              if (IFD) s -= (double)m * (v * e);
              row[m] = s;
            }
          }
        }
        wave_lds_sync();
        if (act) {
          if (F32) {
            fcx a32 = fcx{0.f, 0.f};
            for (int jj = g; jj < cnt; jj += G) {
              const float2 t = shf[jj * P + n];
              a32 = faddc(a32, fcx{t.x, t.y});
            }
            acc += todbl(a32);
          } else {
            for (int jj = g; jj < cnt; jj += G) acc += sh[jj * P + n];
          }
        }
        wave_lds_sync();
      }
      if (act) sh[g * P + n] = acc;
      wave_lds_sync();
      if (lane < np1) {
        cplx s = mkc(0.0);
        for (int gg = 0; gg < G; ++gg) s += sh[gg * P + lane];
        if (lane > 0) s = (-1.0 / lane) * s;
        loc[((int64_t)b * nd + d) * P + lane] += s;
      }
      wave_lds_sync();
    }
  }
}

// ---------------------------------------------------------- list 3 + L2P
// l2dtaeval{p,g,h} (laprouts2d.f:696-897) from the point's leaf and
// l2dmpeval{p,g,h} (laprouts2d.f:289-495) of the leaf's list-3 boxes.
constexpr int kEvU = 4;

// f = sum_{k=0}^{kmax} a_k t^k with f1 = f'(t) and f2 = f''(t)/2, by Horner's
// rule with synthetic division, coefficients a[k] read kEvU at a time from the
// highest down (loads issued together); a_0 is taken as zero when zero0.
template <int PG>
__device__ __forceinline__ void horner(const cplx* a, int kmax, bool zero0, cplx t, cplx& f,
                                       cplx& f1, cplx& f2) {
  f = f1 = f2 = mkc(0.0);
  for (int k0 = kmax; k0 >= 0; k0 -= kEvU) {
    cplx q[kEvU];
#pragma unroll
    for (int u = 0; u < kEvU; ++u) q[u] = a[max(k0 - u, 0)];
#pragma unroll
    for (int u = 0; u < kEvU; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        if (PG >= 3) f2 = f2 * t + f1;
        if (PG >= 2) f1 = f1 * t + f;
        f = f * t + ((zero0 && k == 0) ? mkc(0.0) : q[u]);
      }
    }
  }
}

// multipole of box c at (x, y), added to ap, ag, ah: M_0 log|z| + g(u),
// g(u) = sum_{n>=1} M_n u^n, u = r/z; d/dz = M_0/z - g'(u) u^2/r,
// d2/dz2 = -M_0/z^2 + (g''(u) u^4 + 2 g'(u) u^3)/r^2
template <int PG>
__device__ __forceinline__ void mp_eval(double x, double y, int c, int nd, int d, int P,
                                        const int* ilevel, const int* nterms,
                                        const double* rsc, const double* centers,
                                        const cplx* mp, cplx& ap, cplx& ag, cplx& ah) {
  const int lc = ilevel[c];
  const int pc = nterms[lc];
  const double rc = rsc[lc], rci = 1.0 / rc;
  // z = target - centre of a well-separated box: 1/z = conj(z) / |z|^2
  // and log|z| = log(|z|^2) / 2 need no overflow guards
  const double zx = x - centers[2 * c], zy = y - centers[2 * c + 1];
  const double rr = zx * zx + zy * zy, rcr = rc / rr;
  const cplx u = cplx{zx * rcr, -zy * rcr};  // r/z
  const cplx* M = mp + ((int64_t)c * nd + d) * P;
  const cplx m0 = M[0];
  cplx g, g1, g2;
  horner<PG>(M, pc, true, u, g, g1, g2);
  ap += (0.5 * fast_log(rr)) * m0 + g;
  if (PG >= 2) {
    const cplx u2 = u * u;
    ag += (u * rci) * m0 - (u2 * rci) * g1;
    if (PG >= 3) {
      const cplx u3 = u2 * u;
      ah += (rci * rci) * (-(u2 * m0) + 2.0 * ((u3 * u) * g2 + u3 * g1));
    }
  }
}

// local expansion of leaf b at (x, y): f(w) = sum L_n w^n, w = (z - c)/r;
// pot f, d/dz = f'(w)/r, d2/dz2 = f''(w)/r^2
template <int PG>
__device__ __forceinline__ void loc_eval(double x, double y, int b, int nd, int d, int P,
                                         const int* ilevel, const int* nterms,
                                         const double* rsc, const double* centers,
                                         const cplx* loc, cplx& ap, cplx& ag, cplx& ah) {
  const int lev = ilevel[b];
  const double r = rsc[lev], rinv = 1.0 / r;
  const cplx w = cplx{(x - centers[2 * b]) * rinv, (y - centers[2 * b + 1]) * rinv};
  cplx f1, f2;
  horner<PG>(loc + ((int64_t)b * nd + d) * P, nterms[lev], false, w, ap, f1, f2);
  ag = ah = mkc(0.0);
  if (PG >= 2) ag = f1 * rinv;
  if (PG >= 3) ah = (2.0 * rinv * rinv) * f2;
}

// Mixed precision (nd = 1): the same evaluations from FP32 copies of the
// expansions, in FP32 arithmetic.  The displacement from the box centre is
// formed in FP64 and rounded once; each evaluation's result is added to the
// point's FP64 sums.

// horner<PG> over interleaved FP32 coefficients a[2k], a[2k+1]
template <int PG>
__device__ __forceinline__ void horner32(const float* a, int kmax, bool zero0, fcx t, fcx& f,
                                         fcx& f1, fcx& f2) {
  f = f1 = f2 = fcx{0.f, 0.f};
  const float2* a2 = reinterpret_cast<const float2*>(a);
  for (int k0 = kmax; k0 >= 0; k0 -= kEvU) {
    float2 q[kEvU];
#pragma unroll
    for (int u = 0; u < kEvU; ++u) q[u] = a2[max(k0 - u, 0)];
#pragma unroll
    for (int u = 0; u < kEvU; ++u) {
      const int k = k0 - u;
      if (k >= 0) {
        if (PG >= 3) f2 = faddc(fmulc(f2, t), f1);
        if (PG >= 2) f1 = faddc(fmulc(f1, t), f);
        f = faddc(fmulc(f, t), (zero0 && k == 0) ? fcx{0.f, 0.f} : fcx{q[u].x, q[u].y});
      }
    }
  }
}

template <int PG>
__device__ __forceinline__ void mp_eval32(double x, double y, int c, int P, const int* ilevel,
                                          const int* nterms, const double* rsc,
                                          const double* centers, const float* mp, cplx& ap,
                                          cplx& ag, cplx& ah) {
  const int lc = ilevel[c];
  const int pc = nterms[lc];
  const double rc = rsc[lc];
  const double zx = x - centers[2 * c], zy = y - centers[2 * c + 1];
  const double rr = zx * zx + zy * zy, rcr = rc / rr;
  const fcx u = fcx{(float)(zx * rcr), (float)(-zy * rcr)};  // r/z
  const float rci = (float)(1.0 / rc);
  const float* M = mp + (int64_t)c * 2 * P;
  const fcx m0 = fcx{M[0], M[1]};
  fcx g, g1, g2;
  horner32<PG>(M, pc, true, u, g, g1, g2);
  const float lg = 0.5f * __builtin_amdgcn_logf((float)rr) * 0.69314718055994531f;
  ap += todbl(faddc(fscl(lg, m0), g));
  if (PG >= 2) {
    const fcx u2 = fmulc(u, u);
    const fcx t0 = fmulc(fscl(rci, u), m0), t1 = fmulc(fscl(rci, u2), g1);
    ag += todbl(fcx{t0.re - t1.re, t0.im - t1.im});
    if (PG >= 3) {
      const fcx u3 = fmulc(u2, u);
      const fcx a0 = fmulc(u2, m0);
      const fcx b0 = faddc(fmulc(fmulc(u3, u), g2), fmulc(u3, g1));
      ah += todbl(fscl(rci * rci, fcx{2.f * b0.re - a0.re, 2.f * b0.im - a0.im}));
    }
  }
}

template <int PG>
__device__ __forceinline__ void loc_eval32(double x, double y, int b, int P, const int* ilevel,
                                           const int* nterms, const double* rsc,
                                           const double* centers, const float* loc, cplx& ap,
                                           cplx& ag, cplx& ah) {
  const int lev = ilevel[b];
  const double r = rsc[lev], rinv = 1.0 / r;
  const fcx w = fcx{(float)((x - centers[2 * b]) * rinv), (float)((y - centers[2 * b + 1]) * rinv)};
  fcx f, f1, f2;
  horner32<PG>(loc + (int64_t)b * 2 * P, nterms[lev], false, w, f, f1, f2);
  ap = todbl(f);
  ag = ah = mkc(0.0);
  if (PG >= 2) ag = todbl(fscl((float)rinv, f1));
  if (PG >= 3) ah = todbl(fscl((float)(2.0 * rinv * rinv), f2));
}

// Thread per point (any nd): local expansion, then the point's list-3 boxes.
// The evaluation kernels are the first to write the sorted outputs: they
// assign, and the near field adds after them.
template <int PG>
__global__ void k_eval(int64_t n, int nd, int P, const double* xy, const int* leaf,
                       const int* ilevel, const int* nterms, const double* rsc,
                       const double* centers, const int* n3, const int* l3, const cplx* mp,
                       const cplx* loc, cplx* pot, cplx* grad, cplx* hess) {
  const int64_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = leaf[i];
  const double x = xy[2 * i], y = xy[2 * i + 1];
  const int c3 = n3[b];
  for (int d = 0; d < nd; ++d) {
    cplx ap, ag, ah;
    loc_eval<PG>(x, y, b, nd, d, P, ilevel, nterms, rsc, centers, loc, ap, ag, ah);
    for (int k3 = 0; k3 < c3; ++k3)
      mp_eval<PG>(x, y, l3[(int64_t)kMaxList3 * b + k3], nd, d, P, ilevel, nterms, rsc, centers,
                  mp, ap, ag, ah);
    pot[i * nd + d] = ap;
    if (PG >= 2) grad[i * nd + d] = ag;
    if (PG >= 3) hess[i * nd + d] = ah;
  }
}

// nd = 1: a block of kEvB consecutive points shares out their (point, list-3
// box) pairs evenly -- in rounds of kEvB pairs, pair k to thread k mod kEvB,
// its point found by binary search in the block's scan of list-3 counts --
// so leaves with different list-3 lengths in one wavefront leave no lane
// idle.  Each round's terms go to LDS, and each thread adds its own point's
// terms from there in pair order (a fixed order: the sums do not depend on
// scheduling).  Each thread then adds its own point's local expansion.
constexpr int kEvB = 256;

template <int PG, int F32>
__global__ __launch_bounds__(kEvB) void k_eval_b(int64_t n, int P, const double* xy,
                                                 const int* leaf, const int* ilevel,
                                                 const int* nterms, const double* rsc,
                                                 const double* centers, const int* n3,
                                                 const int* l3, const cplx* mp,
                                                 const cplx* loc, const float* mp32,
                                                 const float* loc32, cplx* pot, cplx* grad,
                                                 cplx* hess) {
  static_assert(kEvB == kScanThreads, "block_exclusive_scan spans the workgroup");
  __shared__ int scan_tmp[kScanWaves + 1];
  __shared__ int off[kEvB];
  __shared__ int sb[kEvB];
  __shared__ double sx[kEvB], sy[kEvB];
  __shared__ double part[2 * PG][kEvB];  // re, im of pot, grad, hess: one round's terms
  const int tid = threadIdx.x;
  const int64_t i = xcd_block() * kEvB + tid;
  const bool live = i < n;
  const int b = live ? leaf[i] : 0;
  const double x = live ? xy[2 * i] : 0.0, y = live ? xy[2 * i + 1] : 0.0;
  const int cnt = live ? n3[b] : 0;
  sb[tid] = b;
  sx[tid] = x;
  sy[tid] = y;
  double my[2 * PG];
#pragma unroll
  for (int a = 0; a < 2 * PG; ++a) my[a] = 0.0;
  int o, m;
  o = block_exclusive_scan(cnt, scan_tmp, &m);
  off[tid] = o;
  __syncthreads();
  for (int r0 = 0; r0 < m; r0 += kEvB) {  // uniform
    const int k = r0 + tid;
    if (k < m) {
      int lo = 0, hi = kEvB - 1;  // last slot t with off[t] <= k
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= k) lo = mid;
        else hi = mid - 1;
      }
      const int bt = sb[lo];
      cplx ap = mkc(0.0), ag = mkc(0.0), ah = mkc(0.0);
      const int c = l3[(int64_t)kMaxList3 * bt + (k - off[lo])];
      if (F32)
        mp_eval32<PG>(sx[lo], sy[lo], c, P, ilevel, nterms, rsc, centers, mp32, ap, ag, ah);
      else
        mp_eval<PG>(sx[lo], sy[lo], c, 1, 0, P, ilevel, nterms, rsc, centers, mp, ap, ag, ah);
      part[0][tid] = ap.re;
      part[1][tid] = ap.im;
      if (PG >= 2) {
        part[2][tid] = ag.re;
        part[3][tid] = ag.im;
      }
      if (PG >= 3) {
        part[4][tid] = ah.re;
        part[5][tid] = ah.im;
      }
    }
    __syncthreads();
    // this point's pairs of the round, in pair order
    const int kb = max(o, r0), ke = min(o + cnt, r0 + kEvB);
    for (int kk = kb; kk < ke; ++kk) {
#pragma unroll
      for (int a = 0; a < 2 * PG; ++a) my[a] += part[a][kk - r0];
    }
    __syncthreads();
  }
  if (!live) return;
  cplx ap, ag, ah;
  if (F32) loc_eval32<PG>(x, y, b, P, ilevel, nterms, rsc, centers, loc32, ap, ag, ah);
  else loc_eval<PG>(x, y, b, 1, 0, P, ilevel, nterms, rsc, centers, loc, ap, ag, ah);
  pot[i] = ap + cplx{my[0], my[1]};
  if (PG >= 2) grad[i] = ag + cplx{my[2], my[3]};
  if (PG >= 3) hess[i] = ah + cplx{my[4], my[5]};
}

// ---------------------------------------------------------- near field
// cfmm2dpart_direct (laplace/cfmm2d.f:1450-1616): the list-1 sources of the
// point's leaf through the c2d_direct* variant of the flags, one thread per
// evaluation point (loop structure at k_near).  Where U = 4 sources are taken
// together their loads are issued together, and the exclusion rule masks a
// pair's strengths instead of branching.
//
// RE = 1: the charges are real and only Re(pot) is consumed (lfmm2d and
// rfmm2d split to Cauchy densities with zero imaginary parts,
// laplace/lfmm2d.f:139-156, laplace/rfmm2d.f:135-148), so the charge terms
// take the real charge alone.
//
// F32 = 1 (mixed precision): the displacement is formed in FP64 and rounded
// to FP32; log|z|, 1/z, -1/z^2, 2/z^3 and their products with FP32 copies of
// the strengths are FP32 (hardware reciprocal and log2), summed in FP32 over
// at most kFoldT sources and added to FP64 accumulators after them.  The FP32 form
// takes the pairs with lo < |z|^2 < 1e20 (FP32), lo = max(1e-20, 4 thresh^2)
// rounded up: no such pair can be excluded, and the powers of 1/z stay in
// the normal FP32 range.  Every other pair (coincident points among them)
// goes through the exact FP64 exclusion rule and the FP64 form.

// c2d_direct* for one pair, FP64 (RE: real charges, Re(pot) only)
template <int IFC, int IFD, int PG, int NDT, int RE>
__device__ __forceinline__ void pair_c64(double dx, double dy, double rr, const SrcStr<NDT>& s,
                                         Acc<CAUCHY, PG, NDT>& a) {
  const double ri = fast_rcp(rr);
  const cplx zinv = cplx{dx * ri, -dy * ri};  // 1/z
  const cplx z2 = -(zinv * zinv);             // -1/z^2
  const cplx z3 = (-2.0) * (z2 * zinv);       // 2/z^3
  const double lg = IFC ? 0.5 * fast_log(rr) : 0.0;
#pragma unroll
  for (int d = 0; d < NDT; ++d) {
    if (IFC) {
      if (RE) {
        const double c = s.charge[d].re;
        a.pot[d].re += lg * c;
        if (PG >= 2) a.grad[d][0] += c * zinv;
        if (PG >= 3) a.hess[d][0] += c * z2;
      } else {
        a.pot[d] += lg * s.charge[d];
        if (PG >= 2) a.grad[d][0] += zinv * s.charge[d];
        if (PG >= 3) a.hess[d][0] += z2 * s.charge[d];
      }
    }
    if (IFD) {
      if (RE) {
        a.pot[d].re += zinv.re * s.dipstr[d].re - zinv.im * s.dipstr[d].im;
      } else {
        a.pot[d] += zinv * s.dipstr[d];
      }
      if (PG >= 2) a.grad[d][0] += z2 * s.dipstr[d];
      if (PG >= 3) a.hess[d][0] += z3 * s.dipstr[d];
    }
  }
}

// FP32 sums of one list-1 box
// (lr, li: charge times log2|z|^2, scaled by ln(2)/2 when folded)
template <int PG, int NDT> struct AccF {
  float pr[NDT], pi[NDT], lr[NDT], li[NDT], gr[NDT], gi[NDT], hr[NDT], hi[NDT];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int d = 0; d < NDT; ++d)
      pr[d] = pi[d] = lr[d] = li[d] = gr[d] = gi[d] = hr[d] = hi[d] = 0.0f;
  }
  __device__ __forceinline__ void fold(Acc<CAUCHY, PG, NDT>& a) const {
    constexpr double kHalfLn2 = 0.34657359027997264;
#pragma unroll
    for (int d = 0; d < NDT; ++d) {
      a.pot[d] += cplx{(double)pr[d] + kHalfLn2 * (double)lr[d],
                       (double)pi[d] + kHalfLn2 * (double)li[d]};
      if (PG >= 2) a.grad[d][0] += cplx{(double)gr[d], (double)gi[d]};
      if (PG >= 3) a.hess[d][0] += cplx{(double)hr[d], (double)hi[d]};
    }
  }
};

// FP32 strengths of one source: charge (cr, ci; ci unused with RE) and
// dipole strength (dr, di)
template <int NDT> struct StrF {
  float cr[NDT], ci[NDT], dr[NDT], di[NDT];
};

// FP32 strength layout: charges ns x nd floats (RE) or (re, im) pairs,
// dipole strengths (re, im) pairs
template <int IFC, int IFD, int NDT, int RE>
__device__ __forceinline__ void load_strf(int64_t j, int nd, int d0, const float* ch32,
                                          const float* dp32, StrF<NDT>& s) {
#pragma unroll
  for (int dd = 0; dd < NDT; ++dd) {
    const bool ok = dd == 0 || d0 + dd < nd;  // d0 < nd always
    const int64_t e = NDT == 1 ? j : j * nd + d0 + dd;  // NDT = 1 only with nd = 1
    s.cr[dd] = s.ci[dd] = s.dr[dd] = s.di[dd] = 0.0f;
    if (IFC && ok) {
      if (RE) {
        s.cr[dd] = ch32[e];
      } else {
        s.cr[dd] = ch32[2 * e];
        s.ci[dd] = ch32[2 * e + 1];
      }
    }
    if (IFD && ok) {
      s.dr[dd] = dp32[2 * e];
      s.di[dd] = dp32[2 * e + 1];
    }
  }
}

// the FP32 form; w = 0 drops the pair (its frr is replaced by 1)
template <int IFC, int IFD, int PG, int NDT, int RE>
__device__ __forceinline__ void pair_f32(float fx, float fy, float frr, bool w,
                                         const StrF<NDT>& s, AccF<PG, NDT>& f) {
  frr = w ? frr : 1.0f;
  const float ri = __builtin_amdgcn_rcpf(frr);
  const float z1r = fx * ri, z1i = -fy * ri;                          // 1/z
  const float z2r = z1i * z1i - z1r * z1r, z2i = -2.0f * z1r * z1i;   // -1/z^2
  const float z3r = -2.0f * (z2r * z1r - z2i * z1i);                 // 2/z^3
  const float z3i = -2.0f * (z2r * z1i + z2i * z1r);
  const float l2 = IFC ? __builtin_amdgcn_logf(frr) : 0.0f;           // log2 |z|^2
#pragma unroll
  for (int d = 0; d < NDT; ++d) {
    if (IFC) {
      const float cr = w ? s.cr[d] : 0.0f;
      f.lr[d] += l2 * cr;
      if (RE) {
        if (PG == 2) {
          // c/z = (c / |z|^2) conj(z): one product fewer than through 1/z
          const float cri = cr * ri;
          f.gr[d] += fx * cri;
          f.gi[d] -= fy * cri;
        }
        if (PG >= 3) {
          f.gr[d] += z1r * cr;
          f.gi[d] += z1i * cr;
          f.hr[d] += z2r * cr;
          f.hi[d] += z2i * cr;
        }
      } else {
        const float ci = w ? s.ci[d] : 0.0f;
        f.li[d] += l2 * ci;
        if (PG >= 2) { f.gr[d] += z1r * cr - z1i * ci; f.gi[d] += z1r * ci + z1i * cr; }
        if (PG >= 3) { f.hr[d] += z2r * cr - z2i * ci; f.hi[d] += z2r * ci + z2i * cr; }
      }
    }
    if (IFD) {
      const float cr = w ? s.dr[d] : 0.0f, ci = w ? s.di[d] : 0.0f;
      f.pr[d] += z1r * cr - z1i * ci;
      if (!RE) f.pi[d] += z1r * ci + z1i * cr;
      if (PG >= 2) { f.gr[d] += z2r * cr - z2i * ci; f.gi[d] += z2r * ci + z2i * cr; }
      if (PG >= 3) { f.hr[d] += z3r * cr - z3i * ci; f.hi[d] += z3r * ci + z3i * cr; }
    }
  }
}

constexpr int kNearU = 4;

// strengths of source j (zero beyond nd)
template <int IFC, int IFD, int NDT>
__device__ __forceinline__ void load_str(int64_t j, int nd, int d0, const cplx* chs,
                                         const cplx* dps, SrcStr<NDT>& s) {
#pragma unroll
  for (int dd = 0; dd < NDT; ++dd) {
    const bool ok = dd == 0 || d0 + dd < nd;  // d0 < nd always
    const int64_t e = NDT == 1 ? j : j * nd + d0 + dd;  // NDT = 1 only with nd = 1
    if (IFC) s.charge[dd] = ok ? chs[e] : mkc(0.0);
    if (IFD) s.dipstr[dd] = ok ? dps[e] : mkc(0.0);
  }
}

template <int NDT> __device__ __forceinline__ void zero_str(SrcStr<NDT>& s) {
#pragma unroll
  for (int dd = 0; dd < NDT; ++dd) s.charge[dd] = s.dipstr[dd] = mkc(0.0);
}

// one pair through the exact exclusion rule and the FP64 form
template <int IFC, int IFD, int PG, int NDT, int RE>
__device__ __forceinline__ void near_pair64(double dx, double dy, SrcStr<NDT>& s, double thresh,
                                            double thresh2, Acc<CAUCHY, PG, NDT>& a) {
  double rr = dx * dx + dy * dy;
  // exclusion rule of c2d_direct* (|z| computed only where it can decide)
  double absz = 2.0 * thresh;
  if (!IFC && rr <= 4.0 * thresh2) absz = hypot(dx, dy);
  if (skip_pair<CAUCHY, IFC, IFD, PG>(rr, thresh, thresh2, absz)) {
    zero_str(s);
    rr = 1.0;
    dx = 1.0;
    dy = 0.0;
  }
  pair_c64<IFC, IFD, PG, NDT, RE>(dx, dy, rr, s, a);
}

// The FP64 exclusion rule and form for one pair, added to the outputs of
// point i (k_near_slow: coincident or out-of-range pairs of the FP32 path).
template <int IFC, int IFD, int PG, int NDT, int RE>
__device__ __forceinline__ void slow_pair(double dx, double dy, int64_t j, int64_t i, int nd,
                                       int d0, const cplx* chs, const cplx* dps, double thresh,
                                       double thresh2, cplx* pot, cplx* grad, cplx* hess) {
  const double rr = dx * dx + dy * dy;
  double absz = 2.0 * thresh;
  if (!IFC && rr <= 4.0 * thresh2) absz = hypot(dx, dy);
  if (skip_pair<CAUCHY, IFC, IFD, PG>(rr, thresh, thresh2, absz)) return;
  SrcStr<NDT> s;
  load_str<IFC, IFD, NDT>(j, nd, d0, chs, dps, s);
  Acc<CAUCHY, PG, NDT> r;
  r.zero();
  pair_c64<IFC, IFD, PG, NDT, RE>(dx, dy, rr, s, r);
#pragma unroll
  for (int dd = 0; dd < NDT; ++dd) {
    if (d0 + dd >= nd) break;
    const int64_t o = i * nd + d0 + dd;
    if (RE) pot[o].re += r.pot[dd].re;
    else pot[o] += r.pot[dd];
    if (PG >= 2) grad[o] += r.grad[dd][0];
    if (PG >= 3) hess[o] += r.hess[dd][0];
  }
}

// |z|^2 of the FP32 form, with the contraction pinned so that k_near and
// k_near_slow classify every pair alike
__device__ __forceinline__ float near_frr(float fx, float fy) {
  return __builtin_fmaf(fx, fx, fy * fy);
}

// one pair of the mixed-precision path; ok = 0: the point's own pair,
// dropped.  Returns true for a pair the FP32 form does not take (coincident
// points or |z| out of the FP32 range), left to k_near_slow.
template <int IFC, int IFD, int PG, int NDT, int RE>
__device__ __forceinline__ bool near_pair32(double dx, double dy, bool ok, const StrF<NDT>& sf,
                                            float lo, AccF<PG, NDT>& f) {
  const float fx = (float)dx, fy = (float)dy;
  const float frr = near_frr(fx, fy);
  const bool in = frr > lo && frr < 1e20f;
  pair_f32<IFC, IFD, PG, NDT, RE>(fx, fy, frr, in && ok, sf, f);
  return !in && ok;
}

// Thread per evaluation point.  The sources of the point's leaf come from its
// segment row (k_near_segs) through one loop over all of them: kNearU
// consecutive sources per step while every active lane has that many left in
// its current range, one source per step otherwise, so lanes of different
// leaves in one wavefront stay busy until their own totals run out.  The step
// counter t advances alike on every active lane; FP32 sums are folded into
// the FP64 accumulators every kFoldT sources.  self: the points are the
// sources themselves (point i is source i), whose own pair the FP32 path drops
// by index.  A point with a pair the FP32 form does not take is appended to
// slow (at most once, so n entries suffice) for k_near_slow.
constexpr int kFoldT = 32;

template <int IFC, int IFD, int PG, int NDT, int F32, int RE>
__global__ __launch_bounds__(256) void k_near(int64_t n, int nd, int d0, const double* xy,
                                              const int* leaf, const int2* seg, int self,
                                              const double* __restrict__ xs, const cplx* chs,
                                              const cplx* dps, const float* __restrict__ ch32,
                                              const float* __restrict__ dp32, double thresh,
                                              float lo, int* slow, int* nslow, cplx* pot,
                                              cplx* grad, cplx* hess) {
  const int64_t i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = leaf[i];
  const double x = xy[2 * i], y = xy[2 * i + 1];
  const double thresh2 = thresh * thresh;
  const int iself = self ? (int)i : -1;
  bool any_slow = false;
  Acc<CAUCHY, PG, NDT> a;
  a.zero();
  AccF<PG, NDT> f;
  f.zero();
  SegCursor sc;
  const int T = sc.init(seg + (int64_t)kSegRow * b);
  int tf = kFoldT;
  for (int t = 0; t < T;) {
    const int j = sc.j;
    const double* px = xs + 2 * (int64_t)j;
    const int j0 = __builtin_amdgcn_readfirstlane(j);
    if (F32 && wave_all(j == j0 && sc.e - j >= kNearU)) {
      // every lane at the same source (a wavefront within one leaf): the
      // source loads are wave-uniform and go through the scalar unit instead
      // of every lane's vector load
      const double* p0 = xs + 2 * (int64_t)j0;
      double sx[kNearU], sy[kNearU];
      StrF<NDT> sf[kNearU];
#pragma unroll
      for (int u = 0; u < kNearU; ++u) {
        sx[u] = p0[2 * u];
        sy[u] = p0[2 * u + 1];
        load_strf<IFC, IFD, NDT, RE>(j0 + u, nd, d0, ch32, dp32, sf[u]);
      }
      const int ds = iself - j0;
#pragma unroll
      for (int u = 0; u < kNearU; ++u)
        any_slow |= near_pair32<IFC, IFD, PG, NDT, RE>(x - sx[u], y - sy[u], ds != u, sf[u], lo,
                                                       f);
      sc.j = j + kNearU;
      t += kNearU;
    } else if (wave_all(sc.e - j >= kNearU)) {
      if (F32) {
        double sx[kNearU], sy[kNearU];
        StrF<NDT> sf[kNearU];
#pragma unroll
        for (int u = 0; u < kNearU; ++u) {
          sx[u] = px[2 * u];
          sy[u] = px[2 * u + 1];
          load_strf<IFC, IFD, NDT, RE>(j + u, nd, d0, ch32, dp32, sf[u]);
        }
        const int ds = iself - j;
#pragma unroll
        for (int u = 0; u < kNearU; ++u)
          any_slow |= near_pair32<IFC, IFD, PG, NDT, RE>(x - sx[u], y - sy[u], ds != u, sf[u],
                                                         lo, f);
      } else {
        double sx[kNearU], sy[kNearU];
        SrcStr<NDT> s[kNearU];
#pragma unroll
        for (int u = 0; u < kNearU; ++u) {
          sx[u] = px[2 * u];
          sy[u] = px[2 * u + 1];
          load_str<IFC, IFD, NDT>(j + u, nd, d0, chs, dps, s[u]);
        }
#pragma unroll
        for (int u = 0; u < kNearU; ++u)
          near_pair64<IFC, IFD, PG, NDT, RE>(x - sx[u], y - sy[u], s[u], thresh, thresh2, a);
      }
      sc.j = j + kNearU;
      t += kNearU;
    } else {
      if (F32) {
        StrF<NDT> sf;
        load_strf<IFC, IFD, NDT, RE>(j, nd, d0, ch32, dp32, sf);
        any_slow |= near_pair32<IFC, IFD, PG, NDT, RE>(x - px[0], y - px[1], j != iself, sf, lo,
                                                       f);
      } else {
        SrcStr<NDT> s;
        load_str<IFC, IFD, NDT>(j, nd, d0, chs, dps, s);
        near_pair64<IFC, IFD, PG, NDT, RE>(x - px[0], y - px[1], s, thresh, thresh2, a);
      }
      sc.j = j + 1;
      t += 1;
    }
    sc.step();
    if (F32 && t >= tf) {
      f.fold(a);
      f.zero();
      tf = t + kFoldT;
    }
  }
  if (F32) {
    f.fold(a);
    if (any_slow) slow[atomicAdd(nslow, 1)] = (int)i;
  }
#pragma unroll
  for (int dd = 0; dd < NDT; ++dd) {
    if (d0 + dd >= nd) break;
    const int64_t o = i * nd + d0 + dd;
    if (RE) pot[o].re += a.pot[dd].re;
    else pot[o] += a.pot[dd];
    if (PG >= 2) grad[o] += a.grad[dd][0];
    if (PG >= 3) hess[o] += a.hess[dd][0];
  }
}

// The pairs of the points listed by k_near that its FP32 form left out,
// through the exact FP64 exclusion rule and the FP64 form, added to those
// points' outputs (grid-stride over the list).
template <int IFC, int IFD, int PG, int NDT, int RE>
__global__ __launch_bounds__(256) void k_near_slow(int nd, int d0, const double* xy,
                                                   const int* leaf, const int2* seg, int self,
                                                   const double* xs, const cplx* chs,
                                                   const cplx* dps, double thresh, float lo,
                                                   const int* slow, const int* nslow, cplx* pot,
                                                   cplx* grad, cplx* hess) {
  const int cnt = *nslow;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < cnt; q += gridDim.x * blockDim.x) {
    const int i = slow[q];
    const int iself = self ? i : -1;
    const double x = xy[2 * (int64_t)i], y = xy[2 * (int64_t)i + 1];
    SegCursor sc;
    const int T = sc.init(seg + (int64_t)kSegRow * leaf[i]);
    for (int t = 0; t < T; ++t) {
      const int j = sc.j;
      const double dx = x - xs[2 * (int64_t)j], dy = y - xs[2 * (int64_t)j + 1];
      const float frr = near_frr((float)dx, (float)dy);
      if (!(frr > lo && frr < 1e20f) && j != iself)
        slow_pair<IFC, IFD, PG, NDT, RE>(dx, dy, j, i, nd, d0, chs, dps, thresh, thresh * thresh,
                                         pot, grad, hess);
      sc.j = j + 1;
      sc.step();
    }
  }
}

// ---------------------------------------------------------- host side
using lcplx = std::complex<long double>;

long double binom(int n, int k) {
  if (k < 0 || k > n) return 0.0L;
  long double r = 1.0L;
  for (int i = 1; i <= k; ++i) r = r * (long double)(n - k + i) / (long double)i;
  return r;
}

cplx tocplx(lcplx z) { return cplx{(double)z.real(), (double)z.imag()}; }

// The fixed translation matrices (level independent for equal orders).
void make_matrices(int P, std::vector<cplx>& A, std::vector<cplx>& B, std::vector<cplx>& T) {
  A.assign(4 * (size_t)P * P, cplx{0, 0});
  B.assign(4 * (size_t)P * P, cplx{0, 0});
  T.assign(kOff * kOff * (size_t)P * P, cplx{0, 0});
  for (int q = 0; q < 4; ++q) {
    const long double sx = (q & 1) ? 1.0L : -1.0L, sy = (q & 2) ? 1.0L : -1.0L;
    // M2M: child - parent = (sx, sy) * r_child/2; r_parent = 2 r_child
    const lcplx w = lcplx(sx, sy) / 2.0L, wh = w / 2.0L, wi = 1.0L / w;
    for (int n = 0; n < P; ++n)
      for (int j = 0; j <= n; ++j) {
        lcplx v;
        if (n == 0) v = 1.0L;
        else if (j == 0) v = -std::pow(wh, n) / (long double)n;
        else v = binom(n - 1, j - 1) * std::pow(wi, j) * std::pow(wh, n);
        A[((size_t)q * P + j) * P + n] = tocplx(v);
      }
    // L2L: child - parent = (sx, sy) * r_parent/4; r_child = r_parent/2
    const lcplx v4 = lcplx(sx, sy) / 4.0L;
    for (int n = 0; n < P; ++n)
      for (int j = n; j < P; ++j) {
        const lcplx v = binom(j, n) * std::pow(v4, j - n) * std::pow(0.5L, n);
        B[((size_t)q * P + j) * P + n] = tocplx(v);
      }
  }
  // M2L: source - target = u * r with u = dx + i dy, max(|dx|,|dy|) >= 2
  for (int dx = -3; dx <= 3; ++dx)
    for (int dy = -3; dy <= 3; ++dy) {
      if (std::abs(dx) <= 1 && std::abs(dy) <= 1) continue;
      const int o = (dx + 3) * kOff + (dy + 3);
      const lcplx u((long double)dx, (long double)dy), ui = 1.0L / u, mui = -ui;
      for (int n = 0; n < P; ++n)
        for (int j = 0; j < P; ++j) {
          lcplx v;
          if (n == 0 && j == 0) v = std::log(std::abs(u));
          else if (n == 0) v = std::pow(mui, j);
          else if (j == 0) v = -std::pow(ui, n) / (long double)n;
          else v = binom(n + j - 1, j - 1) * std::pow(mui, j) * std::pow(ui, n);
          T[((size_t)o * P + j) * P + n] = tocplx(v);
        }
    }
}

// Translation operators of one expansion length P as MFMA fragments: the 49
// M2L offsets, 4 M2M and 4 L2L quadrants in FP64, and the M2L offsets in FP32
// for mixed precision.  They depend on P alone (the Laplace operators are
// level independent), so each (device, P) builds them once; they stay for
// the life of the process.
struct LapOps {
  double* af = nullptr;
  float* af32 = nullptr;
};

const LapOps& lap_ops(int P) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, LapOps> cache;
  int dev = 0;
  hip_throw(hipGetDevice(&dev), "device");
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find({dev, P});
  if (it != cache.end()) return it->second;
  std::vector<cplx> A, B, T;
  make_matrices(P, A, B, T);
  std::vector<cplx> ops(T);
  ops.insert(ops.end(), A.begin(), A.end());
  ops.insert(ops.end(), B.begin(), B.end());
  const std::vector<double> f64 = swizzle_ops(P, ops, kOff * kOff + 8);
  const std::vector<float> f32 = swizzle_ops<float>(P, T, kOff * kOff);
  LapOps o;
  hip_throw(hipMalloc(&o.af, f64.size() * sizeof(double)), "alloc");
  hip_throw(hipMalloc(&o.af32, f32.size() * sizeof(float)), "alloc");
  hip_throw(hipMemcpy(o.af, f64.data(), f64.size() * sizeof(double), hipMemcpyHostToDevice),
            "copy");
  hip_throw(hipMemcpy(o.af32, f32.data(), f32.size() * sizeof(float), hipMemcpyHostToDevice),
            "copy");
  return cache.emplace(std::make_pair(dev, P), o).first->second;
}

struct NearArgs {
  int nd;
  const double* xy;  // sorted evaluation points
  const int* leaf;   // leaf box of each
  const int2* seg;   // near-field segment rows (k_near_segs)
  int self;          // the points are the sorted sources
  int* slow;         // FP32 path: points with pairs left to k_near_slow (n entries)
  int* nslow;        // and their count
  const double* xs;
  const cplx* chs;
  const cplx* dps;
  const float* ch32;  // FP32 strengths (mixed precision), see load_strf
  const float* dp32;
  double thresh;
  cplx *p, *g, *h;
};

constexpr int kNearSlowBlocks = 256;

template <int PG, int NDT, int F32, int RE>
void launch_near_pg(int ifc, int ifd, int64_t n, hipStream_t st, int d0, const NearArgs& q) {
  // FP32 pairs need |z|^2 > max(1e-20, 4 thresh^2): no exclusion, normal range
  const float lo = (float)(std::max(1e-20, 4.0 * q.thresh * q.thresh) * 1.001);
#define FMM2D_NEAR_L(IFC, IFD)                                                                   \
  do {                                                                                           \
    if (F32) hip_throw(hipMemsetAsync(q.nslow, 0, sizeof(int), st), "hipMemsetAsync");          \
    launch1d_xcd(k_near<IFC, IFD, PG, NDT, F32, RE>, n, st, q.nd, d0, q.xy, q.leaf, q.seg,     \
                 q.self, q.xs, q.chs, q.dps, q.ch32, q.dp32, q.thresh, lo, q.slow, q.nslow,   \
                 q.p, q.g, q.h);                                                               \
    if (F32) {                                                                                   \
      hipLaunchKernelGGL((k_near_slow<IFC, IFD, PG, NDT, RE>), dim3(kNearSlowBlocks), dim3(256), \
                         0, st, q.nd, d0, q.xy, q.leaf, q.seg, q.self, q.xs, q.chs, q.dps,       \
                         q.thresh, lo, q.slow, q.nslow, q.p, q.g, q.h);                        \
      hip_throw(hipGetLastError(), "kernel launch");                                             \
    }                                                                                            \
  } while (0)
  if (ifc && !ifd) FMM2D_NEAR_L(1, 0);
  else if (!ifc && ifd) FMM2D_NEAR_L(0, 1);
  else FMM2D_NEAR_L(1, 1);
#undef FMM2D_NEAR_L
}

template <int F32, int RE>
void launch_near_f(int pg, int ifc, int ifd, int64_t n, hipStream_t st, const NearArgs& q) {
  const int NDT = (q.nd == 1) ? 1 : 2;
  for (int d0 = 0; d0 < q.nd; d0 += NDT) {
    if (NDT == 1) {
      if (pg == 1) launch_near_pg<1, 1, F32, RE>(ifc, ifd, n, st, d0, q);
      else if (pg == 2) launch_near_pg<2, 1, F32, RE>(ifc, ifd, n, st, d0, q);
      else launch_near_pg<3, 1, F32, RE>(ifc, ifd, n, st, d0, q);
    } else {
      if (pg == 1) launch_near_pg<1, 2, F32, RE>(ifc, ifd, n, st, d0, q);
      else if (pg == 2) launch_near_pg<2, 2, F32, RE>(ifc, ifd, n, st, d0, q);
      else launch_near_pg<3, 2, F32, RE>(ifc, ifd, n, st, d0, q);
    }
  }
}

// re: charges real and only Re(pot) consumed (lfmm2d, rfmm2d)
void launch_near(bool f32, bool re, int pg, int ifc, int ifd, int64_t n, hipStream_t st,
                 const NearArgs& q) {
  if (f32) {
    if (re) launch_near_f<1, 1>(pg, ifc, ifd, n, st, q);
    else launch_near_f<1, 0>(pg, ifc, ifd, n, st, q);
  } else {
    if (re) launch_near_f<0, 1>(pg, ifc, ifd, n, st, q);
    else launch_near_f<0, 0>(pg, ifc, ifd, n, st, q);
  }
}

// mixed precision: FP32 copy of the multipole expansions for the FP32 M2L
__global__ void k_to_f32(int64_t n, const double* in, float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

// FP32 real parts (the near field's real charges)
__global__ void k_re_to_f32(int64_t n, const cplx* in, float* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i].re;
}

// mp32, loc32: FP32 copies of the expansions (nd = 1, mixed precision), or null
void launch_eval(int pg, int64_t n, hipStream_t st, int nd, int P, const double* xy,
                 const int* leaf, const Tree& t, const int* nterms, const double* rsc,
                 const Lists& L, const cplx* mp, const cplx* loc, const float* mp32,
                 const float* loc32, cplx* p, cplx* g, cplx* h) {
  const bool f32 = nd == 1 && mp32 && loc32;
#define FMM2D_EVAL(PG)                                                                         \
  do {                                                                                         \
    if (f32)                                                                                   \
      launch1d_xcd(k_eval_b<PG, 1>, n, st, P, xy, leaf, t.ilevel.p, nterms, rsc, t.centers.p,  \
                   L.n3.p, L.l3.p, mp, loc, mp32, loc32, p, g, h);                             \
    else if (nd == 1)                                                                          \
      launch1d_xcd(k_eval_b<PG, 0>, n, st, P, xy, leaf, t.ilevel.p, nterms, rsc, t.centers.p,  \
                   L.n3.p, L.l3.p, mp, loc, (const float*)nullptr, (const float*)nullptr, p,   \
                   g, h);                                                                      \
    else                                                                                       \
      launch1d_xcd(k_eval<PG>, n, st, nd, P, xy, leaf, t.ilevel.p, nterms, rsc, t.centers.p,   \
                   L.n3.p, L.l3.p, mp, loc, p, g, h);                                          \
  } while (0)
  static_assert(kEvB == 256, "launch1d_xcd launches blocks of 256 threads");
  if (pg == 1) FMM2D_EVAL(1);
  else if (pg == 2) FMM2D_EVAL(2);
  else FMM2D_EVAL(3);
#undef FMM2D_EVAL
}

}  // namespace

int laplace_fmm(int kind, const CauchyArgs& a, hipStream_t st) {
  if (a.nd < 1 || a.ns < 1 || a.nt < 0 || a.ifpgh < 0 || a.ifpgh > 3 || a.ifpghtarg < 0 ||
      a.ifpghtarg > 3)
    return FMM2D_ERR_ARG;
  if (!a.src || (a.nt > 0 && !a.targ) || (a.ifcharge && !a.charge) ||
      (a.ifdipole && !a.dipstr) || (kind >= 1 && a.ifdipole && !a.dipvec))
    return FMM2D_ERR_ARG;
  if ((a.ifpgh >= 1 && !a.pot) || (a.ifpgh >= 2 && !a.grad) || (a.ifpgh >= 3 && !a.hess) ||
      (a.ifpghtarg >= 1 && !a.pottarg) || (a.ifpghtarg >= 2 && !a.gradtarg) ||
      (a.ifpghtarg >= 3 && !a.hesstarg))
    return FMM2D_ERR_ARG;
  if (a.ns > INT32_MAX || a.nt > INT32_MAX) return FMM2D_ERR_ARG;
  if (!a.ifcharge && !a.ifdipole) return FMM2D_OK;
  const fmm2d_opts* o = a.opts;
  if (o && (o->precision < 0 || o->precision > 1)) return FMM2D_ERR_ARG;
  if (o) {
    const int e = check_nterms_override(o->n_nterms, o->nterms, 2, kMaxNtermsLaplace);
    if (e != FMM2D_OK) return e;
  }
  // precision 1: FP32 M2L on MFMA and FP32 near-field kernels with FP64 sums;
  // tree, expansions, P2M/L2P, M2M/L2L stay FP64
  const bool mixed = o && o->precision == 1;
  try {
    double tinfo[16] = {0};
    Timer tm(o != nullptr, st);
    const int64_t ns = a.ns;
    const int64_t nt = (a.ifpghtarg > 0) ? a.nt : 0;
    // lfmm2d: 2 Cauchy densities per density, or 1 when all are real
    int half = 0;
    if (kind == 1) {
      DBuf<int> flag(1, st);
      flag.zero();
      if (a.ifcharge) launch1d(k_any_imag, a.ns * a.nd, st, a.charge, flag.p);
      if (a.ifdipole) launch1d(k_any_imag, a.ns * a.nd, st, a.dipstr, flag.p);
      int h = 0;
      readback(&h, flag.p, sizeof(int), st);
      half = (h == 0) ? 1 : 0;
    }
    const int nd = (kind == 1 && !half) ? 2 * a.nd : a.nd;  // Cauchy densities
    int ndiv = lndiv2d(a.eps, ns, nt);
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
    std::vector<int> nterms(nl + 1, l2dterms(a.eps));
    if (o && o->n_nterms > 0 && o->nterms)
      for (int l = 0; l <= nl; ++l) nterms[l] = o->nterms[std::min(l, o->n_nterms - 1)];
    int pmax = 0;
    for (int v : nterms) pmax = std::max(pmax, v);
    const int P = pmax + 1;
    std::vector<double> rsc(t.boxsize), logr(nl + 1);
    for (int l = 0; l <= nl; ++l) logr[l] = std::log(rsc[l]);
    const LapOps& ops = lap_ops(P);
    DBuf<int> d_nterms = upload(nterms, st);
    DBuf<double> d_rsc = upload(rsc, st), d_logr = upload(logr, st);

    // ---- sort and gather
    DBuf<double> xs(2 * ns, st), xt(2 * std::max<int64_t>(nt, 1), st);
    DBuf<cplx> chs(a.ifcharge ? ns * nd : 1, st), dps(a.ifdipole ? ns * nd : 1, st);
    DBuf<int> invs;  // lfmm2d: sorted position of each caller's source
    if (kind != 1) launch1d(k_gather_xy, ns, st, t.isrc.p, a.src, xs.p);
    if (nt > 0) launch1d(k_gather_xy, nt, st, t.itarg.p, a.targ, xt.p);
    if (kind == 0) {
      if (a.ifcharge) launch1d(k_gather_dens, ns, st, nd, t.isrc.p, a.charge, chs.p);
      if (a.ifdipole) launch1d(k_gather_dens, ns, st, nd, t.isrc.p, a.dipstr, dps.p);
    } else if (kind == 2) {
      launch1d(k_gather_real, ns, st, nd, t.isrc.p,
               a.ifcharge ? reinterpret_cast<const double*>(a.charge) : nullptr,
               a.ifdipole ? reinterpret_cast<const double*>(a.dipstr) : nullptr, a.dipvec,
               chs.p, dps.p);
    } else {
      if (a.ifpgh > 0) invs.alloc(ns, st);
      launch1d(k_gather_split, ns, st, a.nd, half, t.isrc.p, a.ifcharge ? a.charge : nullptr,
               a.ifdipole ? a.dipstr : nullptr, a.dipvec, chs.p, dps.p, a.src, xs.p,
               a.ifpgh > 0 ? invs.p : nullptr);
    }
    DBuf<int> leaf_s(ns, st), leaf_t(std::max<int64_t>(nt, 1), st);
    launch1d(k_leaf_of, nb, st, t.nchild.p, t.isrcse.p, leaf_s.p);
    if (nt > 0) launch1d(k_leaf_of, nb, st, t.nchild.p, t.itargse.p, leaf_t.p);
    const int64_t nexp = (int64_t)nb * nd * P;
    DBuf<cplx> mp(nexp, st), loc(nexp, st);
    mp.zero();
    loc.zero();
    const int pg = a.ifpgh, pgt = a.ifpghtarg;
    DBuf<cplx> ps(pg >= 1 ? ns * nd : 1, st), gs(pg >= 2 ? ns * nd : 1, st),
        hs(pg >= 3 ? ns * nd : 1, st);
    DBuf<cplx> pt(pgt >= 1 ? nt * nd : 1, st), gt(pgt >= 2 ? nt * nd : 1, st),
        ht(pgt >= 3 ? nt * nd : 1, st);
    // (the outputs need no clearing: list 3 + L2P writes every point first)
    tm.mark(9);

    const int* sse = t.isrcse.p;
    const int* tse = t.itargse.p;
    // sources per LDS batch of P2M / P2L: at most 64, and the tables within 64 KiB
    const size_t per_src = (size_t)(P + 1) * sizeof(cplx) + sizeof(double) + sizeof(int);
    const int bat = (int)std::max<size_t>(1, std::min<size_t>(kW, 65536 / per_src));
    const size_t lds_p2m = (size_t)bat * (P * sizeof(cplx) + sizeof(int));
    const size_t lds_p2l = (size_t)bat * per_src;
    // ---- step 1: P2M, step 2: list 4 (wavefronts over compact box lists up
    // to 64 coefficients; the LDS forms above beyond)
    const bool re_ch = kind != 0;  // real Cauchy charges
    DBuf<int> p2x_list(nb, st), p2x_cnt(1, st);
    // rows of P coefficients per wavefront: 32, or 16 beyond P = 32 (LDS
    // 4 x bat x P x 16 bytes <= 64 KiB)
    const int bat_x = P <= 32 ? 32 : 16;
    const size_t lds_p2x = (size_t)kP2lWaves * bat_x * P * sizeof(cplx);
    int p2x_grid = 0;
    if (P <= kW) {
      DBuf<int> flag(nb, st), pos(nb, st);
      launch1d(k_p2l_flag, nb, st, t.ilevel.p, L.n4.p, sse, tse, pg, pgt, flag.p);
      exclusive_scan(flag.p, pos.p, nb, (int*)nullptr, st);
      launch1d(k_p2l_list, nb, st, flag.p, pos.p, p2x_list.p, p2x_cnt.p);
      int dev = 0, ncu = 0, per_cu = 0;
      hip_throw(hipGetDevice(&dev), "device");
      hip_throw(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "cu");
      hip_throw(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_p2l_w<1, 1, 1, 0>,
                                                             kP2lWaves * kW, lds_p2x),
                "occupancy");
      p2x_grid = std::max(1, ncu * std::max(1, per_cu));
    }
#define FMM2D_P2L_W32(C, D, RE, F)                                                              \
  hipLaunchKernelGGL((k_p2l_w<C, D, RE, F>), dim3(p2x_grid), dim3(kP2lWaves * kW), lds_p2x,   \
                     st, p2x_list.p, p2x_cnt.p, nd, P, t.ilevel.p, d_nterms.p, d_rsc.p,        \
                     t.centers.p, L.n4.p, L.l4.p, sse, xs.p, chs.p, dps.p, bat_x, loc.p)
#define FMM2D_P2L_W(C, D, RE)                  \
  do {                                         \
    if (mixed) FMM2D_P2L_W32(C, D, RE, 1);     \
    else FMM2D_P2L_W32(C, D, RE, 0);           \
  } while (0)
#define FMM2D_P2M_REG(PT, C, D, RE)                                                           \
  launch1d_xcd(k_p2m_reg<PT, C, D, RE>, nbd, st, nd, P, t.ilevel.p, t.nchild.p, d_nterms.p,   \
               d_rsc.p, t.centers.p, sse, xs.p, chs.p, dps.p, mp.p)
#define FMM2D_P2M_PT(C, D, RE)                  \
  if (P <= 16) FMM2D_P2M_REG(16, C, D, RE);     \
  else if (P <= 20) FMM2D_P2M_REG(20, C, D, RE); \
  else if (P <= 24) FMM2D_P2M_REG(24, C, D, RE); \
  else if (P <= 32) FMM2D_P2M_REG(32, C, D, RE); \
  else if (P <= 48) FMM2D_P2M_REG(48, C, D, RE); \
  else FMM2D_P2M_REG(64, C, D, RE)
    if (P <= kW) {
      const int64_t nbd = (int64_t)nb * nd;
      if (a.ifcharge && !a.ifdipole) {
        if (re_ch) FMM2D_P2M_PT(1, 0, 1);
        else FMM2D_P2M_PT(1, 0, 0);
      } else if (!a.ifcharge && a.ifdipole) {
        FMM2D_P2M_PT(0, 1, 0);
      } else {
        if (re_ch) FMM2D_P2M_PT(1, 1, 1);
        else FMM2D_P2M_PT(1, 1, 0);
      }
      tm.mark(0);
      if (a.ifcharge && !a.ifdipole) {
        if (re_ch) FMM2D_P2L_W(1, 0, 1);
        else FMM2D_P2L_W(1, 0, 0);
      } else if (!a.ifcharge && a.ifdipole) {
        FMM2D_P2L_W(0, 1, 0);
      } else {
        if (re_ch) FMM2D_P2L_W(1, 1, 1);
        else FMM2D_P2L_W(1, 1, 0);
      }
      hip_throw(hipGetLastError(), "p2l");
    } else {
      if (a.ifcharge && !a.ifdipole)
        hipLaunchKernelGGL((k_p2m<1, 0>), dim3(nb), dim3(kW), lds_p2m, st, nd, P, t.ilevel.p, t.nchild.p,
                           d_nterms.p, d_rsc.p, t.centers.p, sse, xs.p, chs.p, dps.p, bat, mp.p);
      else if (!a.ifcharge && a.ifdipole)
        hipLaunchKernelGGL((k_p2m<0, 1>), dim3(nb), dim3(kW), lds_p2m, st, nd, P, t.ilevel.p, t.nchild.p,
                           d_nterms.p, d_rsc.p, t.centers.p, sse, xs.p, chs.p, dps.p, bat, mp.p);
      else
        hipLaunchKernelGGL((k_p2m<1, 1>), dim3(nb), dim3(kW), lds_p2m, st, nd, P, t.ilevel.p, t.nchild.p,
                           d_nterms.p, d_rsc.p, t.centers.p, sse, xs.p, chs.p, dps.p, bat, mp.p);
      hip_throw(hipGetLastError(), "p2m");
      tm.mark(0);
      // ---- step 2: list 4
      if (a.ifcharge && !a.ifdipole)
        hipLaunchKernelGGL((k_p2l<1, 0>), dim3(nb), dim3(kW), lds_p2l, st, nd, P, t.ilevel.p, d_nterms.p,
                           d_rsc.p, t.centers.p, L.n4.p, L.l4.p, sse, tse, pg, pgt, xs.p, chs.p,
                           dps.p, bat, loc.p);
      else if (!a.ifcharge && a.ifdipole)
        hipLaunchKernelGGL((k_p2l<0, 1>), dim3(nb), dim3(kW), lds_p2l, st, nd, P, t.ilevel.p, d_nterms.p,
                           d_rsc.p, t.centers.p, L.n4.p, L.l4.p, sse, tse, pg, pgt, xs.p, chs.p,
                           dps.p, bat, loc.p);
      else
        hipLaunchKernelGGL((k_p2l<1, 1>), dim3(nb), dim3(kW), lds_p2l, st, nd, P, t.ilevel.p, d_nterms.p,
                           d_rsc.p, t.centers.p, L.n4.p, L.l4.p, sse, tse, pg, pgt, xs.p, chs.p,
                           dps.p, bat, loc.p);
      hip_throw(hipGetLastError(), "p2l");
    }
#undef FMM2D_P2M_PT
#undef FMM2D_P2M_REG
#undef FMM2D_P2L_W
#undef FMM2D_P2L_W32
    tm.mark(1);
    // ---- translation operators as MFMA fragments: 49 M2L offsets, 4 M2M, 4 L2L
    const int opM2M = kOff * kOff, opL2L = opM2M + 4;
    const double* d_af = ops.af;
    std::vector<int2> band(nl + 1);
    for (int l = 0; l <= nl; ++l) band[l] = int2{0, 2 * (nterms[l] + 1)};
    DBuf<int2> d_band = upload(band, st);
    ChildPlan cp;
    plan_children(t, pg, pgt, cp, st);
    DBuf<int> d_op4 = upload(std::vector<int>{0, 1, 2, 3}, st);
    // ---- step 3: M2M, finest parents first
    for (int l = nl - 1; l >= 1; --l)
      launch_translate(0, 0, nd, TrBufs{mp.p, 0, P, mp.p, 0, P}, 4,
                       {TrGroup{cp.lev_pos[l], cp.lev_pos[l + 1], opM2M, 0}}, cp.list.p, cp.up.p,
                       d_op4.p, d_af, t.ilevel.p, d_band.p, d_logr.p, st);
    tm.mark(2);
    // ---- step 4: M2L, targets grouped by child position (box order within a group)
    // (mixed precision: FP32 copies of the multipoles, for the M2L and the
    // list-3 evaluations, and of the locals after L2L, for L2P)
    DBuf<float> mp32, loc32;
    if (mixed) {
      mp32.alloc(2 * nexp, st);
      launch1d(k_to_f32, 2 * nexp, st, reinterpret_cast<const double*>(mp.p), mp32.p);
    }
    {
      M2LPlan m2l;
      plan_m2l(t, L, pg, pgt, m2l, st);
      if (m2l.ntg > 0) {
        DBuf<int> d_soff = upload(m2l_slot_offsets(), st);
        std::vector<TrGroup> groups;
        for (int q = 0; q < 4; ++q)
          groups.push_back(TrGroup{m2l.cls_seg[q], m2l.cls_seg[q + 1], 0, q});
        if (mixed) {
          // FP32 operators and the FP32 copy of the multipoles
          tm.mark(3);  // planning
          launch_translate(0, 1, nd, TrBufs{mp32.p, 0, P, loc.p, 0, P}, kSlots, groups,
                           m2l.tlist.p, L.s2.p, d_soff.p, ops.af32, t.ilevel.p, d_band.p,
                           d_logr.p, st, true, true, true);
          tm.mark(13);  // the M2L kernel alone
        } else {
          tm.mark(3);  // planning
          launch_translate(0, 1, nd, TrBufs{mp.p, 0, P, loc.p, 0, P}, kSlots, groups,
                           m2l.tlist.p, L.s2.p, d_soff.p, d_af, t.ilevel.p, d_band.p,
                           d_logr.p, st, false, true, true);
          tm.mark(13);  // the M2L kernel alone
        }
      }
    }
    tm.mark(3);
    // ---- step 5: L2L, coarsest first
    for (int l = 1; l <= nl - 1; ++l)
      launch_translate(1, 0, nd, TrBufs{loc.p, 0, P, loc.p, 0, P}, 4,
                       {TrGroup{cp.lev_pos[l], cp.lev_pos[l + 1], opL2L, 0}}, cp.list.p,
                       cp.down.p, d_op4.p, d_af, t.ilevel.p, d_band.p, d_logr.p, st);
    tm.mark(4);
    // ---- steps 6+7: list 3 and L2P
    if (mixed && nd == 1) {
      loc32.alloc(2 * nexp, st);
      launch1d(k_to_f32, 2 * nexp, st, reinterpret_cast<const double*>(loc.p), loc32.p);
    }
    if (pg > 0)
      launch_eval(pg, ns, st, nd, P, xs.p, leaf_s.p, t, d_nterms.p, d_rsc.p, L, mp.p, loc.p,
                  mp32.p, loc32.p, ps.p, gs.p, hs.p);
    if (pgt > 0)
      launch_eval(pgt, nt, st, nd, P, xt.p, leaf_t.p, t, d_nterms.p, d_rsc.p, L, mp.p, loc.p,
                  mp32.p, loc32.p, pt.p, gt.p, ht.p);
    tm.mark(6);
    // ---- step 8: near field, thresh = boxsize(0) 2^-51 (cfmm2d.f:1276)
    const int ifnear = o ? o->ifnear : 1;
    if (ifnear) {
      const double thresh = t.boxsize[0] * std::ldexp(1.0, -51);
      const bool re = kind != 0;  // real Cauchy charges, Re(pot) consumed
      DBuf<float> ch32, dp32;
      if (mixed && a.ifcharge) {
        ch32 = DBuf<float>((re ? 1 : 2) * ns * nd, st);
        if (re) launch1d(k_re_to_f32, ns * nd, st, chs.p, ch32.p);
        else launch1d(k_to_f32, 2 * ns * nd, st, reinterpret_cast<const double*>(chs.p), ch32.p);
      }
      if (mixed && a.ifdipole) {
        dp32 = DBuf<float>(2 * ns * nd, st);
        launch1d(k_to_f32, 2 * ns * nd, st, reinterpret_cast<const double*>(dps.p), dp32.p);
      }
      DBuf<int2> seg((size_t)kSegRow * t.nboxes, st);
      launch1d(k_near_segs, t.nboxes, st, L.n1.p, L.l1.p, sse, seg.p);
      DBuf<int> slow, nslow;
      if (mixed) {
        slow = DBuf<int>((size_t)std::max(ns, nt), st);
        nslow = DBuf<int>(1, st);
      }
      tm.mark(15);  // FP32 strength copies, segment rows; [7] the near-field kernels alone
      if (pg > 0)
        launch_near(mixed, re, pg, a.ifcharge, a.ifdipole, ns, st,
                    NearArgs{nd, xs.p, leaf_s.p, seg.p, 1, slow.p, nslow.p, xs.p, chs.p, dps.p,
                             ch32.p, dp32.p, thresh, ps.p, gs.p, hs.p});
      if (pgt > 0)
        launch_near(mixed, re, pgt, a.ifcharge, a.ifdipole, nt, st,
                    NearArgs{nd, xt.p, leaf_t.p, seg.p, 0, slow.p, nslow.p, xs.p, chs.p, dps.p,
                             ch32.p, dp32.p, thresh, pt.p, gt.p, ht.p});
    }
    tm.mark(7);
    // ---- unsort
    if (kind == 0) {
      if (pg >= 1) launch1d(k_scatter, ns, st, nd, t.isrc.p, ps.p, a.pot);
      if (pg >= 2) launch1d(k_scatter, ns, st, nd, t.isrc.p, gs.p, a.grad);
      if (pg >= 3) launch1d(k_scatter, ns, st, nd, t.isrc.p, hs.p, a.hess);
      if (pgt >= 1) launch1d(k_scatter, nt, st, nd, t.itarg.p, pt.p, a.pottarg);
      if (pgt >= 2) launch1d(k_scatter, nt, st, nd, t.itarg.p, gt.p, a.gradtarg);
      if (pgt >= 3) launch1d(k_scatter, nt, st, nd, t.itarg.p, ht.p, a.hesstarg);
    } else if (kind == 2) {
      if (pg >= 1)
        launch1d(k_scatter_real, ns, st, nd, pg, t.isrc.p, ps.p, gs.p, hs.p,
                 reinterpret_cast<double*>(a.pot), reinterpret_cast<double*>(a.grad),
                 reinterpret_cast<double*>(a.hess));
      if (pgt >= 1)
        launch1d(k_scatter_real, nt, st, nd, pgt, t.itarg.p, pt.p, gt.p, ht.p,
                 reinterpret_cast<double*>(a.pottarg), reinterpret_cast<double*>(a.gradtarg),
                 reinterpret_cast<double*>(a.hesstarg));
    } else {
      if (pg >= 1)
        launch1d(k_gather_unpack, ns, st, a.nd, half, pg, invs.p, ps.p, gs.p, hs.p, a.pot, a.grad,
                 a.hess);
      if (pgt >= 1)
        launch1d(k_scatter_unpack, nt, st, a.nd, half, pgt, t.itarg.p, pt.p, gt.p, ht.p, a.pottarg,
                 a.gradtarg, a.hesstarg);
    }
    tm.mark(10);
    unsigned long long hc[6] = {0, 0, 0, 0, 0, 0};
    const bool counting = o && o->count_work;
    if (counting) {
      DBuf<unsigned long long> dc(6, st);
      dc.zero();
      launch1d(k_counts, nb, st, t.ilevel.p, t.iparent.p, t.nchild.p, t.ichild.p, sse, tse, pg,
               pgt, L.n1.p, L.l1.p, L.n2.p, L.n3.p, L.n4.p, L.l4.p, 0, dc.p);
      hip_throw(hipMemcpyAsync(hc, dc.p, sizeof(hc), hipMemcpyDeviceToHost, st), "copy");
    }
    hip_throw(hipStreamSynchronize(st), "sync");
    if (o) {
      fmm2d_opts* ow = const_cast<fmm2d_opts*>(o);
      tm.collect(tinfo);
      for (int k = 0; k < 16; ++k) ow->timeinfo[k] = tinfo[k];
      ow->nlevels = nl;
      ow->nboxes = nb;
      if (counting) {
        const double Pd = (double)P, ndd = (double)nd;
        const int64_t neval = (pg > 0 ? ns : 0) + (pgt > 0 ? nt : 0);
        const double gfac = (std::max(pg, pgt) >= 2) ? 2.0 : 1.0;
        double w[16] = {0};
        w[0] = 8.0 * Pd * ndd * (double)ns;
        w[1] = 8.0 * Pd * ndd * (double)hc[5];
        w[2] = 8.0 * Pd * Pd * ndd * (double)hc[2];
        w[3] = 8.0 * Pd * Pd * ndd * (double)hc[1];
        w[4] = 8.0 * Pd * Pd * ndd * (double)hc[3];
        w[5] = 0.0;  // list 3 is fused with L2P: counted in [6]
        w[6] = 8.0 * Pd * ndd * gfac * ((double)neval + (double)hc[4]);
        w[7] = (12.0 + 12.0 * ndd) * (double)hc[0];
        w[8] = 40.0 * (double)(ns + nt) + 80.0 * (double)nb;
        for (int k = 0; k < 16; ++k) ow->work[k] = w[k];
        ow->near_pairs = (int64_t)hc[0];
        ow->m2l_pairs = (int64_t)hc[1];
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

}  // namespace fmm2d

// Biharmonic and Stokes sums (biharmonic/bhfmm2d.f, biharmonic/bhkernels2d.f,
// stokes/stfmm2d.f, stokes/stokkernels2d.f): direct evaluators and FMMs.
//
// Direct evaluators: one thread per target and density, sources through LDS
// tiles of 256; with few target blocks the source range is split over blocks,
// each split writing a private copy that a second kernel adds in split order
// (deterministic), as in p2p.hip.
//
// bhfmm2d.  With zeta = z - z_j and a shifted origin o (z~ = z - o, the centre
// of the points' bounding box), the velocity of bhkernels2d.f
//   vel = sum 2 c1 log|zeta| + c2 zeta/conj(zeta) + d1/zeta + d3/conj(zeta)
//             + d2 zeta/conj(zeta)^2
// is  P0 + conj(P1 - G2) + z~ conj(P3 - G4)  where Pk, Gk, Hk are the
// potential, gradient and Hessian of five Cauchy sums (cauchykernels2d.f:
// pot = c log|zeta| + d/zeta, grad = c/zeta - d/zeta^2, hess = -c/zeta^2 +
// 2 d/zeta^3) with the charges c and dipoles d
//   k0: c1,        d1
//   k1: conj(c1),  conj(d3) - conj(c2) conj(z~_j)
//   k2: 0,        -conj(d2) conj(z~_j)
//   k3: 0,         conj(c2)
//   k4: 0,         conj(d2)
// and the three gradients of bhfmm2d.f:29-60 are
//   grad1 = G0,  grad2 = conj(P3 - G4),  grad3 = conj(G1 - H2) + z~ conj(G3 - H4)
// (each identity checked term by term against bh2d_directcdg,
// bhkernels2d.f:339-410).  The five sums run as ONE Cauchy FMM with 5 nd
// densities on one tree (laplace_fmm kind 0), one derivative order beyond the
// requested one.  The z~ and z~_j weighted terms cancel to O(|zeta|) in the
// sum, so the Cauchy FMM runs at eps / kBhEpsFactor.
//
// stfmm2d maps Stokeslets and type-I stresslets to biharmonic charges and
// dipoles and the biharmonic outputs back to velocity, pressure and velocity
// gradient exactly as stokes/stfmm2d.f:184-300 does (including its constant
// -sum(charge1) and the self term at the sources).
#include "../../include/fmm2d_amd.h"
#include "bh_st.h"
#include "devbuf.h"
#include "fastmath.h"
#include "lap_ops.h"
#include "split_sum.h"

namespace fmm2d {

namespace {

constexpr int kBT = 256;  // targets per block and sources per LDS tile
// Cauchy FMM accuracy relative to the requested eps (the O(|z~|/|zeta|)
// cancellation of the z~-weighted terms)
constexpr double kBhEpsFactor = 100.0;

FMM2D_HD cplx cj(cplx a) { return cplx{a.re, -a.im}; }

// ---------------------------------------------------------------- direct
struct BhTile {
  double x[kBT], y[kBT];
  cplx c1[kBT], c2[kBT], d1[kBT], d2[kBT], d3[kBT];
};

// bh2d_direct{c,d,cd}{p,g} (bhkernels2d.f:26-410); vel(nd,nt), grad(nd,3,nt)
template <int IFC, int IFD, int PG>
__global__ __launch_bounds__(kBT) void k_bh_direct(
    int nd, const double* __restrict__ src, int64_t ns, int64_t split_len,
    const cplx* __restrict__ charge, const cplx* __restrict__ dip,
    const double* __restrict__ targ, int64_t nt, double thresh, cplx* __restrict__ vel,
    cplx* __restrict__ grad, int64_t sv, int64_t sg) {
  __shared__ BhTile sm;
  const int64_t it = (int64_t)blockIdx.x * kBT + threadIdx.x;
  const int d = blockIdx.z;
  const int64_t s0 = (int64_t)blockIdx.y * split_len;
  const int64_t s1 = min(ns, s0 + split_len);
  const bool active = it < nt;
  double tx = 0.0, ty = 0.0;
  if (active) { tx = targ[2 * it]; ty = targ[2 * it + 1]; }
  cplx v = mkc(0.0), g1 = mkc(0.0), g2 = mkc(0.0), g3 = mkc(0.0);
  for (int64_t t0 = s0; t0 < s1; t0 += kBT) {
    const int nl = (int)min((int64_t)kBT, s1 - t0);
    __syncthreads();
    if ((int)threadIdx.x < nl) {
      const int64_t j = t0 + threadIdx.x;
      sm.x[threadIdx.x] = src[2 * j];
      sm.y[threadIdx.x] = src[2 * j + 1];
      if (IFC) {
        sm.c1[threadIdx.x] = charge[(int64_t)nd * (2 * j) + d];
        sm.c2[threadIdx.x] = charge[(int64_t)nd * (2 * j + 1) + d];
      }
      if (IFD) {
        sm.d1[threadIdx.x] = dip[(int64_t)nd * (3 * j) + d];
        sm.d2[threadIdx.x] = dip[(int64_t)nd * (3 * j + 1) + d];
        sm.d3[threadIdx.x] = dip[(int64_t)nd * (3 * j + 2) + d];
      }
    }
    __syncthreads();
    if (!active) continue;
    for (int k = 0; k < nl; ++k) {
      const double dx = tx - sm.x[k], dy = ty - sm.y[k];
      const double rr = dx * dx + dy * dy;
      if (sqrt(rr) <= thresh) continue;  // abs(zdis) .le. thresh
      const double ri = fast_rcp(rr);
      const cplx z = cplx{dx, dy};
      const cplx z1 = cplx{dx * ri, -dy * ri};   // 1/zeta
      const cplx z2 = z1 * z1;                   // 1/zeta^2
      const cplx z1c = cj(z1), z2c = cj(z2);
      if (IFC) {
        const cplx c1 = sm.c1[k], c2 = sm.c2[k];
        const cplx zr = z1c * z;  // zeta / conj(zeta)
        v += (fast_log(rr)) * c1 + c2 * zr;
        if (PG >= 2) {
          g1 += c1 * z1;
          g2 += c2 * z1c;
          g3 += c1 * z1c - c2 * (z2c * z);
        }
      }
      if (IFD) {
        const cplx d1 = sm.d1[k], d2 = sm.d2[k], d3 = sm.d3[k];
        v += d1 * z1 + d3 * z1c + d2 * (z2c * z);
        if (PG >= 2) {
          g1 -= d1 * z2;
          g2 += d2 * z2c;
          g3 -= d3 * z2c + 2.0 * (d2 * (cj(z2 * z1) * z));
        }
      }
    }
  }
  if (!active) return;
  const bool assign = sv != 0;  // split copies are written whole, once
  cplx* V = vel + blockIdx.y * sv;
  cplx* q = &V[(int64_t)nd * it + d];
  *q = assign ? v : *q + v;
  if (PG >= 2) {
    cplx* G = grad + blockIdx.y * sg;
    const cplx gg[3] = {g1, g2, g3};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      q = &G[(int64_t)nd * (3 * it + c) + d];
      *q = assign ? gg[c] : *q + gg[c];
    }
  }
}

struct StTile {
  double x[kBT], y[kBT];
  double s1[kBT], s2[kBT], m1[kBT], m2[kBT], n1[kBT], n2[kBT];
};

// st2ddirectstokg / st2ddirectstokstrsg (stokkernels2d.f:24-338), all three
// outputs; pot(nd,2,nt), pre(nd,nt), grad(nd,2,2,nt)
template <int IFS, int IFT>
__global__ __launch_bounds__(kBT) void k_st_direct(
    int nd, const double* __restrict__ src, int64_t ns, int64_t split_len,
    const double* __restrict__ stoklet, const double* __restrict__ strslet,
    const double* __restrict__ strsvec, const double* __restrict__ targ, int64_t nt,
    double thresh, double* __restrict__ pot, double* __restrict__ pre,
    double* __restrict__ grad, int assign, int64_t sp, int64_t sr, int64_t sg) {
  __shared__ StTile sm;
  const int64_t it = (int64_t)blockIdx.x * kBT + threadIdx.x;
  const int d = blockIdx.z;
  const int64_t s0 = (int64_t)blockIdx.y * split_len;
  const int64_t s1 = min(ns, s0 + split_len);
  const bool active = it < nt;
  double tx = 0.0, ty = 0.0;
  if (active) { tx = targ[2 * it]; ty = targ[2 * it + 1]; }
  const double th2 = thresh * thresh;
  double u1 = 0, u2 = 0, p = 0, g11 = 0, g21 = 0, g12 = 0, g22 = 0;
  for (int64_t t0 = s0; t0 < s1; t0 += kBT) {
    const int nl = (int)min((int64_t)kBT, s1 - t0);
    __syncthreads();
    if ((int)threadIdx.x < nl) {
      const int64_t j = t0 + threadIdx.x;
      sm.x[threadIdx.x] = src[2 * j];
      sm.y[threadIdx.x] = src[2 * j + 1];
      if (IFS) {
        sm.s1[threadIdx.x] = stoklet[(int64_t)nd * (2 * j) + d];
        sm.s2[threadIdx.x] = stoklet[(int64_t)nd * (2 * j + 1) + d];
      }
      if (IFT) {
        sm.m1[threadIdx.x] = strslet[(int64_t)nd * (2 * j) + d];
        sm.m2[threadIdx.x] = strslet[(int64_t)nd * (2 * j + 1) + d];
        sm.n1[threadIdx.x] = strsvec[(int64_t)nd * (2 * j) + d];
        sm.n2[threadIdx.x] = strsvec[(int64_t)nd * (2 * j + 1) + d];
      }
    }
    __syncthreads();
    if (!active) continue;
    for (int k = 0; k < nl; ++k) {
      const double dx = tx - sm.x[k], dy = ty - sm.y[k];
      const double r2 = dx * dx + dy * dy;
      if (r2 < th2) continue;
      const double ri2 = fast_rcp(r2);
      const double ri4 = ri2 * ri2;
      if (IFS) {
        const double a1 = sm.s1[k], a2 = sm.s2[k];
        const double lg = 0.25 * fast_log(r2);
        const double rt = dx * dx - dy * dy;
        const double pl = dx * a1 + dy * a2, pl2 = dx * a2 + dy * a1;
        u1 += -a1 * lg + 0.5 * dx * pl * ri2;
        u2 += -a2 * lg + 0.5 * dy * pl * ri2;
        g11 -= 0.5 * rt * pl * ri4;
        g22 += 0.5 * rt * pl * ri4;
        g21 += 0.5 * (pl2 * rt - 4.0 * a1 * dy * dx * dx) * ri4;
        g12 -= 0.5 * (pl2 * rt + 4.0 * a2 * dx * dy * dy) * ri4;
        p += pl * ri2;
      }
      if (IFT) {
        const double m1 = sm.m1[k], m2 = sm.m2[k], n1 = sm.n1[k], n2 = sm.n2[k];
        const double pl = dx * m1 + dy * m2, pl2 = dx * n1 + dy * n2;
        const double pv = m1 * n1 + m2 * n2;
        const double w = pl * pl2 * ri4, w6 = 8.0 * w * ri2;
        u1 -= 2.0 * dx * w;
        u2 -= 2.0 * dy * w;
        g11 += -2.0 * w - 2.0 * dx * (m1 * pl2 + n1 * pl) * ri4 + dx * dx * w6;
        g21 += -2.0 * dx * (m2 * pl2 + n2 * pl) * ri4 + dx * dy * w6;
        g12 += -2.0 * dy * (m1 * pl2 + n1 * pl) * ri4 + dx * dy * w6;
        g22 += -2.0 * w - 2.0 * dy * (m2 * pl2 + n2 * pl) * ri4 + dy * dy * w6;
        p += -4.0 * w + 2.0 * pv * ri2;
      }
    }
  }
  if (!active) return;
  // split copies (assign = 1) are written whole, once
  auto put = [&](double* base, int64_t stride, int64_t idx, double val) {
    double* q = base + blockIdx.y * stride + idx;
    *q = assign ? val : *q + val;
  };
  if (pot) {
    put(pot, sp, (int64_t)nd * (2 * it) + d, u1);
    put(pot, sp, (int64_t)nd * (2 * it + 1) + d, u2);
  }
  if (pre) put(pre, sr, (int64_t)nd * it + d, p);
  if (grad) {
    // grad(l, i, j, k): ith component of the gradient of velocity component j
    put(grad, sg, (int64_t)nd * (0 + 2 * (0 + 2 * it)) + d, g11);
    put(grad, sg, (int64_t)nd * (1 + 2 * (0 + 2 * it)) + d, g21);
    put(grad, sg, (int64_t)nd * (0 + 2 * (1 + 2 * it)) + d, g12);
    put(grad, sg, (int64_t)nd * (1 + 2 * (1 + 2 * it)) + d, g22);
  }
}

// source-range split so that a launch has about 2048 blocks
int64_t split_count(int64_t tblocks, int64_t ns) {
  int64_t nsplit = 1;
  if (tblocks < 2048) nsplit = (2048 + tblocks - 1) / tblocks;
  nsplit = std::min<int64_t>(nsplit, (ns + kBT - 1) / kBT);
  return std::max<int64_t>(1, std::min<int64_t>(nsplit, 65535));
}

// ---------------------------------------------------------------- bhfmm2d
// the five Cauchy densities per biharmonic density (layout (5 nd, ns))
__global__ void k_bh_densities(int64_t n, int nd, int ifc, int ifd, const double* src,
                               double ox, double oy, const cplx* charge, const cplx* dip,
                               cplx* cch, cplx* cdp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;  // n = points x densities
  const int64_t j = i / nd;
  const int d = (int)(i - j * nd);
  const cplx zj = cplx{src[2 * j] - ox, -(src[2 * j + 1] - oy)};  // conj(z~_j)
  const cplx c1 = ifc ? charge[(int64_t)nd * (2 * j) + d] : mkc(0.0);
  const cplx c2 = ifc ? charge[(int64_t)nd * (2 * j + 1) + d] : mkc(0.0);
  const cplx d1 = ifd ? dip[(int64_t)nd * (3 * j) + d] : mkc(0.0);
  const cplx d2 = ifd ? dip[(int64_t)nd * (3 * j + 1) + d] : mkc(0.0);
  const cplx d3 = ifd ? dip[(int64_t)nd * (3 * j + 2) + d] : mkc(0.0);
  const int64_t b = 5 * i;  // = (5 nd) j + 5 d
  cch[b + 0] = c1;
  cch[b + 1] = cj(c1);
  cch[b + 2] = mkc(0.0);
  cch[b + 3] = mkc(0.0);
  cch[b + 4] = mkc(0.0);
  cdp[b + 0] = d1;
  cdp[b + 1] = cj(d3) - cj(c2) * zj;
  cdp[b + 2] = -(cj(d2) * zj);
  cdp[b + 3] = cj(c2);
  cdp[b + 4] = cj(d2);
}

// the biharmonic outputs at n points from the Cauchy sums (5 nd densities)
__global__ void k_bh_combine(int64_t n, int nd, int pg, const double* xy, double ox, double oy,
                             const cplx* P, const cplx* G, const cplx* H, cplx* vel,
                             cplx* grad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;  // n = points x densities
  const int64_t j = i / nd;
  const int d = (int)(i - j * nd);
  const cplx z = cplx{xy[2 * j] - ox, xy[2 * j + 1] - oy};
  const int64_t b = 5 * i;
  const cplx C = P[b + 3] - G[b + 4];
  vel[i] = P[b + 0] + cj(P[b + 1] - G[b + 2]) + z * cj(C);
  if (pg >= 2) {
    grad[(int64_t)nd * (3 * j) + d] = G[b + 0];
    grad[(int64_t)nd * (3 * j + 1) + d] = cj(C);
    grad[(int64_t)nd * (3 * j + 2) + d] = cj(G[b + 1] - H[b + 2]) + z * cj(G[b + 3] - H[b + 4]);
  }
}

__global__ void k_bbox_centre(int64_t ns, const double* src, int64_t nt, const double* trg,
                              double* out) {
  // one block: min/max of x and y over sources and targets, then the centre
  __shared__ double r[4][256];
  double mn0 = 1e300, mx0 = -1e300, mn1 = 1e300, mx1 = -1e300;
  for (int64_t i = threadIdx.x; i < ns + nt; i += blockDim.x) {
    const double* p = i < ns ? src + 2 * i : trg + 2 * (i - ns);
    mn0 = fmin(mn0, p[0]);
    mx0 = fmax(mx0, p[0]);
    mn1 = fmin(mn1, p[1]);
    mx1 = fmax(mx1, p[1]);
  }
  r[0][threadIdx.x] = mn0;
  r[1][threadIdx.x] = mx0;
  r[2][threadIdx.x] = mn1;
  r[3][threadIdx.x] = mx1;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      r[0][threadIdx.x] = fmin(r[0][threadIdx.x], r[0][threadIdx.x + s]);
      r[1][threadIdx.x] = fmax(r[1][threadIdx.x], r[1][threadIdx.x + s]);
      r[2][threadIdx.x] = fmin(r[2][threadIdx.x], r[2][threadIdx.x + s]);
      r[3][threadIdx.x] = fmax(r[3][threadIdx.x], r[3][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = 0.5 * (r[0][0] + r[1][0]);
    out[1] = 0.5 * (r[2][0] + r[3][0]);
  }
}

// ---------------------------------------------------------------- stfmm2d
// stfmm2d.f:184-209: biharmonic charges and dipoles of the Stokes sources
__global__ void k_st_to_bh(int64_t n, int nd, int ifs, int ift, const double* stoklet,
                           const double* strslet, const double* strsvec, cplx* charge,
                           cplx* dip) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;  // n = points x densities
  const int64_t j = i / nd;
  const int d = (int)(i - j * nd);
  if (ifs) {
    const double a1 = stoklet[(int64_t)nd * (2 * j) + d], a2 = stoklet[(int64_t)nd * (2 * j + 1) + d];
    const cplx c1 = cplx{0.25 * a2, -0.25 * a1};  // (-i a1 + a2)/4
    charge[(int64_t)nd * (2 * j) + d] = c1;
    charge[(int64_t)nd * (2 * j + 1) + d] = cj(c1);
  }
  if (ift) {
    const double m1 = strslet[(int64_t)nd * (2 * j) + d], m2 = strslet[(int64_t)nd * (2 * j + 1) + d];
    const double n1 = strsvec[(int64_t)nd * (2 * j) + d], n2 = strsvec[(int64_t)nd * (2 * j + 1) + d];
    const cplx zd1 = cplx{0.5 * m2, -0.5 * m1};  // -(-m2 + i m1)/2
    const cplx zd2 = cplx{-n2, n1};              // -n2 + i n1
    const cplx p = zd1 * zd2;
    const cplx e1 = cplx{p.im, -p.re};           // -i zd1 zd2
    dip[(int64_t)nd * (3 * j) + d] = e1;
    dip[(int64_t)nd * (3 * j + 1) + d] = -cj(e1);
    const cplx s = zd1 * cj(zd2) + cj(zd1) * zd2;
    dip[(int64_t)nd * (3 * j + 2) + d] = cplx{-s.im, s.re};  // i s
  }
}

// zsum(d) = -sum_j charge1(d, j) (stfmm2d.f:186,198), one block per density
__global__ void k_st_zsum(int64_t ns, int nd, const cplx* charge, cplx* zsum) {
  __shared__ double r[2][256];
  const int d = blockIdx.x;
  double sr = 0.0, si = 0.0;
  for (int64_t j = threadIdx.x; j < ns; j += blockDim.x) {
    const cplx c = charge[(int64_t)nd * (2 * j) + d];
    sr += c.re;
    si += c.im;
  }
  r[0][threadIdx.x] = sr;
  r[1][threadIdx.x] = si;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      r[0][threadIdx.x] += r[0][threadIdx.x + s];
      r[1][threadIdx.x] += r[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) zsum[d] = cplx{-r[0][0], -r[1][0]};
}

// stfmm2d.f:226-300: velocity, pressure and gradient from the biharmonic
// outputs; self = 1 at the sources (adds the point's own charge1)
__global__ void k_st_from_bh(int64_t n, int nd, int ifpp, int self, const cplx* zsum,
                             const cplx* charge, const cplx* vel, const cplx* grad, double* pot,
                             double* pre, double* gst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;  // n = points x densities
  const int64_t j = i / nd;
  const int d = (int)(i - j * nd);
  if (ifpp >= 1) {
    cplx w = vel[i] + zsum[d];
    if (self && charge) w += charge[(int64_t)nd * (2 * j) + d];
    pot[(int64_t)nd * (2 * j) + d] = w.im;
    pot[(int64_t)nd * (2 * j + 1) + d] = -w.re;
  }
  if (ifpp >= 2) pre[i] = -4.0 * grad[(int64_t)nd * (3 * j) + d].im;
  if (ifpp >= 3) {
    const cplx a = grad[(int64_t)nd * (3 * j) + d], c = grad[(int64_t)nd * (3 * j + 2) + d];
    gst[(int64_t)nd * (0 + 2 * (0 + 2 * j)) + d] = c.im;
    gst[(int64_t)nd * (1 + 2 * (1 + 2 * j)) + d] = -c.im;
    gst[(int64_t)nd * (1 + 2 * (0 + 2 * j)) + d] = 2.0 * a.re - c.re;
    gst[(int64_t)nd * (0 + 2 * (1 + 2 * j)) + d] = -(2.0 * a.re + c.re);
  }
}

int bh_fmm_impl(const BhArgs& a, hipStream_t st) {
  const int nd = a.nd;
  const int64_t ns = a.ns, nt = (a.ifpghtarg > 0) ? a.nt : 0;
  // origin: centre of the bounding box
  DBuf<double> dorg(2, st);
  hipLaunchKernelGGL(k_bbox_centre, dim3(1), dim3(256), 0, st, ns, a.src, nt, a.targ, dorg.p);
  double org[2];
  readback(org, dorg.p, sizeof(org), st);
  const int ndc = 5 * nd;
  DBuf<cplx> cch((size_t)ndc * ns, st), cdp((size_t)ndc * ns, st);
  launch1d(k_bh_densities, ns * nd, st, nd, a.ifcharge, a.ifdipole, a.src, org[0], org[1],
           a.charge, a.dip, cch.p, cdp.p);
  const int pg = a.ifpgh > 0 ? a.ifpgh + 1 : 0, pgt = (nt > 0) ? a.ifpghtarg + 1 : 0;
  DBuf<cplx> P((size_t)ndc * ns * (pg >= 1), st), G((size_t)ndc * ns * (pg >= 2), st),
      H((size_t)ndc * ns * (pg >= 3), st);
  DBuf<cplx> PT((size_t)ndc * nt * (pgt >= 1), st), GT((size_t)ndc * nt * (pgt >= 2), st),
      HT((size_t)ndc * nt * (pgt >= 3), st);
  CauchyArgs c{};
  c.nd = ndc;
  c.eps = a.eps / kBhEpsFactor;
  c.ns = ns;
  c.nt = nt;
  c.src = a.src;
  c.targ = nt > 0 ? a.targ : nullptr;
  c.ifcharge = a.ifcharge;
  c.ifdipole = 1;
  c.charge = cch.p;
  c.dipstr = cdp.p;
  c.dipvec = nullptr;
  c.ifpgh = pg;
  c.ifpghtarg = pgt;
  c.pot = P.p;
  c.grad = G.p;
  c.hess = H.p;
  c.pottarg = PT.p;
  c.gradtarg = GT.p;
  c.hesstarg = HT.p;
  c.opts = a.opts;
  const int ier = laplace_fmm(0, c, st);
  if (ier != FMM2D_OK) return ier;
  if (a.ifpgh > 0)
    launch1d(k_bh_combine, ns * nd, st, nd, a.ifpgh, a.src, org[0], org[1], P.p, G.p, H.p,
             a.pot, a.grad);
  if (nt > 0)
    launch1d(k_bh_combine, nt * nd, st, nd, a.ifpghtarg, a.targ, org[0], org[1], PT.p, GT.p,
             HT.p, a.pottarg, a.gradtarg);
  hip_throw(hipGetLastError(), "kernel launch");
  return FMM2D_OK;
}

}  // namespace

int bh_direct(const BhDirectArgs& a, hipStream_t st) {
  if (a.nd < 1 || a.ns < 0 || a.nt < 0 || a.pg < 1 || a.pg > 2) return FMM2D_ERR_ARG;
  if (!a.ifcharge && !a.ifdipole) return FMM2D_ERR_ARG;
  if ((a.ifcharge && !a.charge) || (a.ifdipole && !a.dip) || !a.vel || (a.pg >= 2 && !a.grad))
    return FMM2D_ERR_ARG;
  if (a.nt == 0 || a.ns == 0) return FMM2D_OK;
  if ((int64_t)a.nd > 65535) return FMM2D_ERR_ARG;
  try {
    const int64_t tb = ceil_div(a.nt, kBT);
    int64_t nsplit = split_count(tb * a.nd, a.ns);
    const int64_t len = ((a.ns + nsplit - 1) / nsplit + kBT - 1) / kBT * kBT;
    nsplit = (a.ns + len - 1) / len;
    const int64_t nv = (int64_t)a.nd * a.nt, ng = a.pg >= 2 ? 3 * nv : 0;
    DBuf<cplx> work(nsplit > 1 ? (size_t)(nv + ng) * nsplit : 0, st);
    cplx* V = nsplit > 1 ? work.p : a.vel;
    cplx* G = nsplit > 1 ? work.p + nv * nsplit : a.grad;
    const int64_t sv = nsplit > 1 ? nv : 0, sg = nsplit > 1 ? ng : 0;
    dim3 grid((unsigned)tb, (unsigned)nsplit, (unsigned)a.nd);
#define FMM2D_BHD(C, D, P)                                                                    \
  hipLaunchKernelGGL((k_bh_direct<C, D, P>), grid, dim3(kBT), 0, st, a.nd, a.src, a.ns, len, \
                     a.charge, a.dip, a.targ, a.nt, a.thresh, V, G, sv, sg)
    const int cd = (a.ifcharge ? 1 : 0) + (a.ifdipole ? 2 : 0);
    if (a.pg == 1) {
      if (cd == 1) FMM2D_BHD(1, 0, 1);
      else if (cd == 2) FMM2D_BHD(0, 1, 1);
      else FMM2D_BHD(1, 1, 1);
    } else {
      if (cd == 1) FMM2D_BHD(1, 0, 2);
      else if (cd == 2) FMM2D_BHD(0, 1, 2);
      else FMM2D_BHD(1, 1, 2);
    }
#undef FMM2D_BHD
    hip_throw(hipGetLastError(), "kernel launch");
    if (nsplit > 1) {
      add_split_copies(a.vel, (const cplx*)V, nv, (int)nsplit, st);
      if (ng) add_split_copies(a.grad, (const cplx*)G, ng, (int)nsplit, st);
    }
    return FMM2D_OK;
  } catch (const HipError& e) {
    fprintf(stderr, "fmm2d_amd: %s failed: %s\n", e.what, hipGetErrorString(e.err));
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  }
}

int st_direct(const StDirectArgs& a, hipStream_t st) {
  if (a.nd < 1 || a.ns < 0 || a.nt < 0) return FMM2D_ERR_ARG;
  if ((a.ifstoklet && !a.stoklet) || (a.istress == 1 && (!a.strslet || !a.strsvec)))
    return FMM2D_ERR_ARG;
  if (a.nt == 0 || a.ns == 0 || (!a.ifstoklet && a.istress != 1)) return FMM2D_OK;
  if (!a.pot && !a.pre && !a.grad) return FMM2D_OK;
  if ((int64_t)a.nd > 65535) return FMM2D_ERR_ARG;
  try {
    const int64_t tb = ceil_div(a.nt, kBT);
    int64_t nsplit = split_count(tb * a.nd, a.ns);
    const int64_t len = ((a.ns + nsplit - 1) / nsplit + kBT - 1) / kBT * kBT;
    nsplit = (a.ns + len - 1) / len;
    const int64_t np = a.pot ? 2 * (int64_t)a.nd * a.nt : 0;
    const int64_t nr = a.pre ? (int64_t)a.nd * a.nt : 0;
    const int64_t ng = a.grad ? 4 * (int64_t)a.nd * a.nt : 0;
    DBuf<double> work(nsplit > 1 ? (size_t)(np + nr + ng) * nsplit : 0, st);
    double* Pp = a.pot ? (nsplit > 1 ? work.p : a.pot) : nullptr;
    double* Pr = a.pre ? (nsplit > 1 ? work.p + np * nsplit : a.pre) : nullptr;
    double* Pg = a.grad ? (nsplit > 1 ? work.p + (np + nr) * nsplit : a.grad) : nullptr;
    const int64_t sp = nsplit > 1 ? np : 0, sr = nsplit > 1 ? nr : 0, sg = nsplit > 1 ? ng : 0;
    dim3 grid((unsigned)tb, (unsigned)nsplit, (unsigned)a.nd);
    const int ifs = a.ifstoklet ? 1 : 0, ift = a.istress == 1 ? 1 : 0;
#define FMM2D_STD(S, T)                                                                       \
  hipLaunchKernelGGL((k_st_direct<S, T>), grid, dim3(kBT), 0, st, a.nd, a.src, a.ns, len,    \
                     a.stoklet, a.strslet, a.strsvec, a.targ, a.nt, a.thresh, Pp, Pr, Pg,    \
                     nsplit > 1 ? 1 : 0, sp, sr, sg)
    if (ifs && ift) FMM2D_STD(1, 1);
    else if (ifs) FMM2D_STD(1, 0);
    else FMM2D_STD(0, 1);
#undef FMM2D_STD
    hip_throw(hipGetLastError(), "kernel launch");
    if (nsplit > 1) {
      if (np) add_split_copies(a.pot, (const double*)Pp, np, (int)nsplit, st);
      if (nr) add_split_copies(a.pre, (const double*)Pr, nr, (int)nsplit, st);
      if (ng) add_split_copies(a.grad, (const double*)Pg, ng, (int)nsplit, st);
    }
    return FMM2D_OK;
  } catch (const HipError& e) {
    fprintf(stderr, "fmm2d_amd: %s failed: %s\n", e.what, hipGetErrorString(e.err));
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  }
}

int bh_fmm(const BhArgs& a, hipStream_t st) {
  if (a.nd < 1 || a.ns < 1 || a.nt < 0 || a.ifpgh < 0 || a.ifpgh > 3 || a.ifpghtarg < 0 ||
      a.ifpghtarg > 3)
    return FMM2D_ERR_ARG;
  if (!a.src || (a.nt > 0 && a.ifpghtarg > 0 && !a.targ) || (a.ifcharge && !a.charge) ||
      (a.ifdipole && !a.dip))
    return FMM2D_ERR_ARG;
  if ((a.ifpgh >= 1 && !a.pot) || (a.ifpgh >= 2 && !a.grad) ||
      (a.ifpghtarg >= 1 && a.nt > 0 && !a.pottarg) ||
      (a.ifpghtarg >= 2 && a.nt > 0 && !a.gradtarg))
    return FMM2D_ERR_ARG;
  if (!a.ifcharge && !a.ifdipole) return FMM2D_OK;
  // Hessians are not computed (bhfmm2d.f:85: "hessians not supported currently")
  BhArgs b = a;
  b.ifpgh = std::min(a.ifpgh, 2);
  b.ifpghtarg = std::min(a.ifpghtarg, 2);
  try {
    return bh_fmm_impl(b, st);
  } catch (const HipError& e) {
    fprintf(stderr, "fmm2d_amd: %s failed: %s\n", e.what, hipGetErrorString(e.err));
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return FMM2D_ERR_ALLOC;
  }
}

int st_fmm(const StArgs& a, hipStream_t st) {
  if (a.nd < 1 || a.ns < 1 || a.nt < 0 || a.ifppreg < 0 || a.ifppreg > 3 ||
      a.ifppregtarg < 0 || a.ifppregtarg > 3)
    return FMM2D_ERR_ARG;
  if (!a.src || (a.ifstoklet && !a.stoklet) || (a.ifstrslet && (!a.strslet || !a.strsvec)))
    return FMM2D_ERR_ARG;
  const int64_t nt = a.ifppregtarg > 0 ? a.nt : 0;
  if (nt > 0 && !a.targ) return FMM2D_ERR_ARG;
  if ((a.ifppreg >= 1 && !a.pot) || (a.ifppreg >= 2 && !a.pre) || (a.ifppreg >= 3 && !a.grad) ||
      (nt > 0 && ((a.ifppregtarg >= 1 && !a.pottarg) || (a.ifppregtarg >= 2 && !a.pretarg) ||
                  (a.ifppregtarg >= 3 && !a.gradtarg))))
    return FMM2D_ERR_ARG;
  try {
    const int nd = a.nd;
    const int64_t ns = a.ns;
    const int ifc = a.ifstoklet == 1 ? 1 : 0, ifd = a.ifstrslet == 1 ? 1 : 0;
    DBuf<cplx> charge((size_t)2 * nd * ns, st), dip((size_t)3 * nd * ns, st), zsum(nd, st);
    charge.zero();
    dip.zero();
    launch1d(k_st_to_bh, ns * nd, st, nd, ifc, ifd, a.stoklet, a.strslet, a.strsvec, charge.p,
             dip.p);
    hipLaunchKernelGGL(k_st_zsum, dim3(nd), dim3(256), 0, st, ns, nd, charge.p, zsum.p);
    const int pg = a.ifppreg == 0 ? 0 : (a.ifppreg == 1 ? 1 : 2);
    const int pgt = a.ifppregtarg == 0 ? 0 : (a.ifppregtarg == 1 ? 1 : 2);
    DBuf<cplx> vel((size_t)nd * ns * (pg >= 1), st), grad((size_t)3 * nd * ns * (pg >= 2), st);
    DBuf<cplx> velt((size_t)nd * nt * (pgt >= 1), st),
        gradt((size_t)3 * nd * nt * (pgt >= 2), st);
    int ier = FMM2D_OK;
    if (ifc || ifd) {
      BhArgs b{};
      b.nd = nd;
      b.eps = a.eps;
      b.ns = ns;
      b.src = a.src;
      b.ifcharge = ifc;
      b.charge = charge.p;
      b.ifdipole = ifd;
      b.dip = dip.p;
      b.ifpgh = pg;
      b.pot = vel.p;
      b.grad = grad.p;
      b.nt = nt;
      b.targ = a.targ;
      b.ifpghtarg = pgt;
      b.pottarg = velt.p;
      b.gradtarg = gradt.p;
      b.opts = a.opts;
      if (pg == 0) b.ifpgh = 0;
      ier = bh_fmm_impl(b, st);
      if (ier != FMM2D_OK) return ier;
    } else {
      vel.zero();
      grad.zero();
      velt.zero();
      gradt.zero();
    }
    if (a.ifppreg > 0)
      launch1d(k_st_from_bh, ns * nd, st, nd, a.ifppreg, 1, zsum.p, ifc ? charge.p : nullptr,
               vel.p, grad.p, a.pot, a.pre, a.grad);
    if (nt > 0)
      launch1d(k_st_from_bh, nt * nd, st, nd, a.ifppregtarg, 0, zsum.p, nullptr, velt.p,
               gradt.p, a.pottarg, a.pretarg, a.gradtarg);
    hip_throw(hipGetLastError(), "kernel launch");
    return ier;
  } catch (const HipError& e) {
    fprintf(stderr, "fmm2d_amd: %s failed: %s\n", e.what, hipGetErrorString(e.err));
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return FMM2D_ERR_ALLOC;
  }
}

}  // namespace fmm2d

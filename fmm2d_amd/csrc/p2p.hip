// All-pairs direct evaluation (the *2d_direct* evaluators and the l2ddir /
// c2ddir / h2ddir Python entry points).
//
// Reference: laplace/cauchykernels2d.f, laplace/lapkernels2d.f,
// helmholtz/helmkernels2d.f -- O(ns*nt) double loops that INCREMENT pot/grad/
// hess.  Here: one thread per target, sources streamed through LDS in tiles of
// 256, per-thread register accumulation over NDT densities.  When there are
// too few target blocks to fill the 256 CUs the four waves of a block share
// 64 targets and split each source tile, and the source range is split across
// blocks; each split writes a private copy of the outputs and a second kernel
// adds the copies into the caller's buffers in split order (deterministic).
#include <atomic>
#include <mutex>

#include "devbuf.h"
#include "p2p.h"
#include "p2p_math.h"
#include "split_sum.h"

namespace fmm2d {

constexpr int kTile = 256;

template <int FAM, int IFC, int IFD, int NDT> struct TileSmem {
  double x[kTile], y[kTile];
  cplx ch[IFC ? kTile * NDT : 1];
  cplx ds[IFD ? kTile * NDT : 1];
  double dv[(IFD && FAM != CAUCHY) ? kTile * NDT * 2 : 1];
};

// Output layout (reference, column-major): pot(nd,n); Cauchy grad(nd,n),
// hess(nd,n); Laplace/Helmholtz grad(nd,2,n), hess(nd,3,n).
//
// WS waves share one block's kTile/WS targets: wave w takes the w-th part of
// every source tile and the parts are added in LDS in wave order at the end
// (WS = 4 for launches with few targets, where a block of 256 targets would
// leave most lanes idle).
template <int FAM, int IFC, int IFD, int PG, int NDT, int WS>
__global__ __launch_bounds__(kTile) void direct_kernel(
    int nd, cplx zk, const double* __restrict__ src, int64_t ns, int64_t split_len,
    const cplx* __restrict__ charge, const cplx* __restrict__ dipstr,
    const double* __restrict__ dipvec, const double* __restrict__ targ, int64_t nt,
    double thresh, cplx* __restrict__ pot, cplx* __restrict__ grad, cplx* __restrict__ hess,
    int64_t split_stride_pot, int64_t split_stride_grad, int64_t split_stride_hess) {
  constexpr int NG = FamDims<FAM>::NG, NH = FamDims<FAM>::NH;
  constexpr int TT = kTile / WS;  // targets per block
  __shared__ TileSmem<FAM, IFC, IFD, NDT> sm;
  __shared__ cplx red[WS > 1 ? (WS - 1) * TT : 1];
  const int lane = (int)threadIdx.x % TT, w = (int)threadIdx.x / TT;
  const int64_t it = (int64_t)blockIdx.x * TT + lane;
  const int d0 = blockIdx.z * NDT;
  const int64_t s_begin = (int64_t)blockIdx.y * split_len;
  const int64_t s_end = min(ns, s_begin + split_len);
  const bool active = it < nt;
  double tx = 0.0, ty = 0.0;
  if (active) { tx = targ[2 * it]; ty = targ[2 * it + 1]; }
  const double thresh2 = thresh * thresh;
  Acc<FAM, PG, NDT> acc;
  acc.zero();
  for (int64_t t0 = s_begin; t0 < s_end; t0 += kTile) {
    const int nload = (int)min((int64_t)kTile, s_end - t0);
    __syncthreads();
    if ((int)threadIdx.x < nload) {
      const int64_t j = t0 + threadIdx.x;
      sm.x[threadIdx.x] = src[2 * j];
      sm.y[threadIdx.x] = src[2 * j + 1];
#pragma unroll
      for (int d = 0; d < NDT; ++d) {
        const bool ok = d0 + d < nd;
        if (IFC) sm.ch[threadIdx.x * NDT + d] = ok ? charge[(int64_t)nd * j + d0 + d] : mkc(0.0);
        if (IFD) sm.ds[threadIdx.x * NDT + d] = ok ? dipstr[(int64_t)nd * j + d0 + d] : mkc(0.0);
        if (IFD && FAM != CAUCHY) {
          sm.dv[(threadIdx.x * NDT + d) * 2 + 0] = ok ? dipvec[(int64_t)nd * (2 * j) + d0 + d] : 0.0;
          sm.dv[(threadIdx.x * NDT + d) * 2 + 1] = ok ? dipvec[(int64_t)nd * (2 * j + 1) + d0 + d] : 0.0;
        }
      }
    }
    __syncthreads();
    if (active) {
      const int kb = w * TT, ke = min(nload, kb + TT);
#pragma unroll(FAM == HELM ? 1 : 2)
      for (int k = kb; k < ke; ++k) {
        SrcStr<NDT> s;
#pragma unroll
        for (int d = 0; d < NDT; ++d) {
          if (IFC) s.charge[d] = sm.ch[k * NDT + d];
          if (IFD) s.dipstr[d] = sm.ds[k * NDT + d];
          if (IFD && FAM != CAUCHY) {
            s.dv[d][0] = sm.dv[(k * NDT + d) * 2];
            s.dv[d][1] = sm.dv[(k * NDT + d) * 2 + 1];
          }
        }
        const double dx = tx - sm.x[k], dy = ty - sm.y[k];
        if constexpr (FAM == CAUCHY)
          pair_cauchy<IFC, IFD, PG, NDT>(dx, dy, s, thresh, thresh2, acc);
        else if constexpr (FAM == LAPLACE)
          pair_laplace<IFC, IFD, PG, NDT>(dx, dy, s, thresh, thresh2, acc);
        else
          pair_helm<IFC, IFD, PG, NDT>(dx, dy, zk, s, thresh, acc);
      }
    }
  }
  if constexpr (WS > 1) {
    // wave 0 adds the other waves' parts in wave order
    auto fold = [&](cplx& v) {
      __syncthreads();
      if (w > 0) red[(w - 1) * TT + lane] = v;
      __syncthreads();
      if (w == 0) {
#pragma unroll
        for (int u = 0; u < WS - 1; ++u) v += red[u * TT + lane];
      }
    };
#pragma unroll
    for (int d = 0; d < NDT; ++d) {
      fold(acc.pot[d]);
      if (PG >= 2) {
#pragma unroll
        for (int c = 0; c < NG; ++c) fold(acc.grad[d][c]);
      }
      if (PG >= 3) {
#pragma unroll
        for (int c = 0; c < NH; ++c) fold(acc.hess[d][c]);
      }
    }
    if (w > 0) return;
  }
  if (!active) return;
  cplx* P = pot + blockIdx.y * split_stride_pot;
  cplx* G = grad + blockIdx.y * split_stride_grad;
  cplx* H = hess + blockIdx.y * split_stride_hess;
  // split copies (stride != 0) are written whole by exactly one block each
  // (target block x, split y, density chunk z), so they are assigned and need
  // no clearing; the caller's buffers are incremented
  const bool assign = split_stride_pot != 0;
#pragma unroll
  for (int d = 0; d < NDT; ++d) {
    if (d0 + d >= nd) break;
    const int dd = d0 + d;
    cplx* q = &P[(int64_t)nd * it + dd];
    *q = assign ? acc.pot[d] : *q + acc.pot[d];
    if (PG >= 2) {
#pragma unroll
      for (int c = 0; c < NG; ++c) {
        q = &G[(int64_t)nd * (NG * it + c) + dd];
        *q = assign ? acc.grad[d][c] : *q + acc.grad[d][c];
      }
    }
    if (PG >= 3) {
#pragma unroll
      for (int c = 0; c < NH; ++c) {
        q = &H[(int64_t)nd * (NH * it + c) + dd];
        *q = assign ? acc.hess[d][c] : *q + acc.hess[d][c];
      }
    }
  }
}

template <int FAM, int IFC, int IFD, int PG, int NDT>
static void launch_direct(int ws, dim3 grid, hipStream_t st, const DirectArgs& a,
                          int64_t split_len, cplx* pot, cplx* grad, cplx* hess, int64_t sp,
                          int64_t sg, int64_t sh) {
  if (NDT == 1 && ws == 4)
    hipLaunchKernelGGL((direct_kernel<FAM, IFC, IFD, PG, NDT, 4>), grid, dim3(kTile), 0, st,
                       a.nd, a.zk, a.src, a.ns, split_len, a.charge, a.dipstr, a.dipvec, a.targ,
                       a.nt, a.thresh, pot, grad, hess, sp, sg, sh);
  else
    hipLaunchKernelGGL((direct_kernel<FAM, IFC, IFD, PG, NDT, 1>), grid, dim3(kTile), 0, st,
                       a.nd, a.zk, a.src, a.ns, split_len, a.charge, a.dipstr, a.dipvec, a.targ,
                       a.nt, a.thresh, pot, grad, hess, sp, sg, sh);
}

template <int FAM, int IFC, int IFD, int NDT>
static void dispatch_pg(int ws, dim3 grid, hipStream_t st, const DirectArgs& a, int64_t sl,
                        cplx* p, cplx* g, cplx* h, int64_t sp, int64_t sg, int64_t sh) {
  if (a.pg == 1) launch_direct<FAM, IFC, IFD, 1, NDT>(ws, grid, st, a, sl, p, g, h, sp, sg, sh);
  else if (a.pg == 2) launch_direct<FAM, IFC, IFD, 2, NDT>(ws, grid, st, a, sl, p, g, h, sp, sg, sh);
  else launch_direct<FAM, IFC, IFD, 3, NDT>(ws, grid, st, a, sl, p, g, h, sp, sg, sh);
}

template <int FAM, int NDT>
static void dispatch_cd(int ws, const DirectArgs& a, dim3 grid, hipStream_t st, int64_t sl,
                        cplx* p, cplx* g, cplx* h, int64_t sp, int64_t sg, int64_t sh) {
  if (a.ifcharge && !a.ifdipole) dispatch_pg<FAM, 1, 0, NDT>(ws, grid, st, a, sl, p, g, h, sp, sg, sh);
  else if (!a.ifcharge && a.ifdipole) dispatch_pg<FAM, 0, 1, NDT>(ws, grid, st, a, sl, p, g, h, sp, sg, sh);
  else dispatch_pg<FAM, 1, 1, NDT>(ws, grid, st, a, sl, p, g, h, sp, sg, sh);
}

// Split copies of the no-allocation entry point (fmm2d_l2d_direct_noalloc):
// kScratchCplx complex values (32 MiB) per device, held until the process
// ends.  Host code never uses the name `scratch` as an address: every device's
// copy is reached through the pointer hipGetSymbolAddress returned for it.
constexpr int64_t kScratchCplx = int64_t(1) << 21;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) cplx scratch[kScratchCplx];

std::atomic<int> g_direct_max_blocks{0};

namespace {
std::mutex g_scratch_mu;
cplx* g_scratch_ptr[kMaxDevices] = {};

// this device's copy of `scratch` (looked up once per device, with it current)
int scratch_ptr(cplx** out) {
  int dev = 0;
  FMM2D_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return FMM2D_ERR_ARG;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    FMM2D_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<cplx*>(p);
  }
  *out = g_scratch_ptr[dev];
  return FMM2D_OK;
}
}  // namespace

int direct_eval(const DirectArgs& a, hipStream_t st, bool noalloc) {
  if (a.nd < 1 || a.ns < 0 || a.nt < 0 || a.pg < 1 || a.pg > 3) return FMM2D_ERR_ARG;
  if (!a.ifcharge && !a.ifdipole) return FMM2D_ERR_ARG;
  if (a.family < CAUCHY || a.family > HELM) return FMM2D_ERR_ARG;
  if (a.nt == 0 || a.ns == 0) return FMM2D_OK;
  const int NDT = (a.nd == 1) ? 1 : 2;
  const int nchunk = (a.nd + NDT - 1) / NDT;
  // few targets (one density): four waves share each block's 64 targets
  const int ws = (NDT == 1 && ceil_div(a.nt, kTile) < 2048) ? 4 : 1;
  const int64_t tblocks = ceil_div(a.nt, kTile / ws);
  const int NG = (a.family == CAUCHY) ? 1 : 2, NH = (a.family == CAUCHY) ? 1 : 3;
  const int64_t np = (int64_t)a.nd * a.nt;
  const int64_t ng = (a.pg >= 2) ? np * NG : 0;
  const int64_t nh = (a.pg >= 3) ? np * NH : 0;
  const int64_t per = np + ng + nh;  // complex values of one split copy
  // aim for >= 2048 blocks (8 per CU) before splitting the source range, or
  // for the caller's bound (fmm2d_set_direct_max_blocks)
  int64_t nsplit = 1;
  const int mb = g_direct_max_blocks.load(std::memory_order_relaxed);
  const int64_t want = mb > 0 ? mb : 2048;
  if (tblocks * nchunk < want) nsplit = (want + tblocks * nchunk - 1) / (tblocks * nchunk);
  const int64_t max_split = (a.ns + kTile - 1) / kTile;
  if (nsplit > max_split) nsplit = max_split;
  if (nsplit > 65535) nsplit = 65535;
  // without allocation the split copies must fit the static array: fewer
  // splits, and none when one copy alone does not fit
  if (noalloc && nsplit > 1 && per * nsplit > kScratchCplx) nsplit = kScratchCplx / per;
  if (nsplit < 1) nsplit = 1;
  const int64_t split_len = ((a.ns + nsplit - 1) / nsplit + kTile - 1) / kTile * kTile;
  nsplit = (a.ns + split_len - 1) / split_len;
  cplx *p = a.pot, *g = a.grad, *h = a.hess;
  cplx* work = nullptr;
  if (nsplit > 1) {
    if (noalloc) {
      const int e = scratch_ptr(&work);
      if (e != FMM2D_OK) return e;
    } else {
      FMM2D_HIP_TRY(hipMallocAsync((void**)&work, sizeof(cplx) * per * nsplit, st));
    }
    p = work;
    g = work + np * nsplit;
    h = work + (np + ng) * nsplit;
  }
  const int64_t sp = (nsplit > 1) ? np : 0, sg = (nsplit > 1) ? ng : 0, sh = (nsplit > 1) ? nh : 0;
  dim3 grid((unsigned)tblocks, (unsigned)nsplit, (unsigned)nchunk);
  if (NDT == 1) {
    if (a.family == CAUCHY) dispatch_cd<CAUCHY, 1>(ws, a, grid, st, split_len, p, g, h, sp, sg, sh);
    else if (a.family == LAPLACE) dispatch_cd<LAPLACE, 1>(ws, a, grid, st, split_len, p, g, h, sp, sg, sh);
    else dispatch_cd<HELM, 1>(ws, a, grid, st, split_len, p, g, h, sp, sg, sh);
  } else {
    if (a.family == CAUCHY) dispatch_cd<CAUCHY, 2>(ws, a, grid, st, split_len, p, g, h, sp, sg, sh);
    else if (a.family == LAPLACE) dispatch_cd<LAPLACE, 2>(ws, a, grid, st, split_len, p, g, h, sp, sg, sh);
    else dispatch_cd<HELM, 2>(ws, a, grid, st, split_len, p, g, h, sp, sg, sh);
  }
  FMM2D_HIP_TRY(hipGetLastError());
  if (nsplit > 1) {
    try {
      add_split_copies(a.pot, (const cplx*)p, np, (int)nsplit, st);
      if (ng) add_split_copies(a.grad, (const cplx*)g, ng, (int)nsplit, st);
      if (nh) add_split_copies(a.hess, (const cplx*)h, nh, (int)nsplit, st);
    } catch (const HipError& e) {
      if (!noalloc) (void)hipFreeAsync(work, st);
      return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
    }
    if (!noalloc) FMM2D_HIP_TRY(hipFreeAsync(work, st));
  }
  return FMM2D_OK;
}

namespace {
__global__ void k_real_to_cplx(int64_t n, const double* in, cplx* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = mkc(in[i]);
}
__global__ void k_add_real(int64_t n, const cplx* in, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += in[i].re;
}
}  // namespace

int direct_eval_real(int nd, const double* src, int64_t ns, int ifcharge, const double* charge,
                     int ifdipole, const double* dipstr, const double* dipvec, const double* targ,
                     int64_t nt, int pg, double thresh, double* pot, double* grad, double* hess,
                     hipStream_t st) {
  if (nd < 1 || ns < 0 || nt < 0 || pg < 1 || pg > 3 || (!ifcharge && !ifdipole))
    return FMM2D_ERR_ARG;
  if ((ifcharge && !charge) || (ifdipole && (!dipstr || !dipvec)) || !pot ||
      (pg >= 2 && !grad) || (pg >= 3 && !hess) || (ns > 0 && !src) || (nt > 0 && !targ))
    return FMM2D_ERR_ARG;
  if (nt == 0 || ns == 0) return FMM2D_OK;
  try {
    const int64_t nsd = ns * nd, ntd = nt * nd;
    DBuf<cplx> ch(ifcharge ? nsd : 0, st), ds(ifdipole ? nsd : 0, st);
    DBuf<cplx> p(ntd, st), g(pg >= 2 ? 2 * ntd : 0, st), h(pg >= 3 ? 3 * ntd : 0, st);
    if (ifcharge) launch1d(k_real_to_cplx, nsd, st, charge, ch.p);
    if (ifdipole) launch1d(k_real_to_cplx, nsd, st, dipstr, ds.p);
    p.zero();
    g.zero();
    h.zero();
    DirectArgs a;
    a.family = LAPLACE;
    a.nd = nd;
    a.zk = mkc(0.0);
    a.src = src;
    a.ns = ns;
    a.ifcharge = ifcharge;
    a.ifdipole = ifdipole;
    a.charge = ch.p;
    a.dipstr = ds.p;
    a.dipvec = dipvec;
    a.targ = targ;
    a.nt = nt;
    a.pg = pg;
    a.thresh = thresh;
    a.pot = p.p;
    a.grad = g.p;
    a.hess = h.p;
    const int ier = direct_eval(a, st);
    if (ier != FMM2D_OK) return ier;
    launch1d(k_add_real, ntd, st, p.p, pot);
    if (pg >= 2) launch1d(k_add_real, 2 * ntd, st, g.p, grad);
    if (pg >= 3) launch1d(k_add_real, 3 * ntd, st, h.p, hess);
  } catch (const HipError& e) {
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  }
  return FMM2D_OK;
}

}  // namespace fmm2d

// C-ABI entry points of the FMM drivers, the tree export and the parameter
// rules.  Declarations and reference citations: include/fmm2d_amd.h.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "../../include/fmm2d_amd.h"
#include "bh_st.h"
#include "helm_ops.h"
#include "lap_ops.h"
#include "tree.h"

using namespace fmm2d;

namespace fmm2d {

void keep_pool_memory() {
  static std::mutex mu;
  static std::vector<char> done;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return;
  std::lock_guard<std::mutex> lock(mu);
  if ((size_t)dev < done.size() && done[dev]) return;
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  if ((size_t)dev >= done.size()) done.resize(dev + 1, 0);
  done[dev] = 1;
}

}  // namespace fmm2d

namespace {

CauchyArgs cauchy_args(int nd, double eps, int64_t ns, const double* sources, int ifcharge,
                       const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dipstr,
                       const double* dipvec, int ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad,
                       fmm2d_cplx* hess, int64_t nt, const double* targ, int ifpghtarg,
                       fmm2d_cplx* pottarg, fmm2d_cplx* gradtarg, fmm2d_cplx* hesstarg,
                       const fmm2d_opts* opts) {
  CauchyArgs a;
  a.nd = nd;
  a.eps = eps;
  a.ns = ns;
  a.nt = nt;
  a.src = sources;
  a.targ = targ;
  a.ifcharge = ifcharge;
  a.ifdipole = ifdipole;
  a.charge = reinterpret_cast<const cplx*>(charge);
  a.dipstr = reinterpret_cast<const cplx*>(dipstr);
  a.dipvec = dipvec;
  a.ifpgh = ifpgh;
  a.ifpghtarg = ifpghtarg;
  a.pot = reinterpret_cast<cplx*>(pot);
  a.grad = reinterpret_cast<cplx*>(grad);
  a.hess = reinterpret_cast<cplx*>(hess);
  a.pottarg = reinterpret_cast<cplx*>(pottarg);
  a.gradtarg = reinterpret_cast<cplx*>(gradtarg);
  a.hesstarg = reinterpret_cast<cplx*>(hesstarg);
  a.opts = opts;
  return a;
}

// the batch streams of each device, created on first use and kept (blocks
// cached by the workspace are keyed by stream)
std::vector<hipStream_t> batch_streams(int dev, int n) {
  static std::mutex mu;
  static std::map<int, std::vector<hipStream_t>> pool;
  std::lock_guard<std::mutex> lock(mu);
  std::vector<hipStream_t>& v = pool[dev];
  while ((int)v.size() < n) {
    hipStream_t s = nullptr;
    hip_throw(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    v.push_back(s);
  }
  return std::vector<hipStream_t>(v.begin(), v.begin() + n);
}

int run_job(fmm2d_job& j, hipStream_t st);

struct TreeHandle {
  Tree t;
  Lists l;
  bool have_lists = false;
  hipStream_t st = nullptr;
};

}  // namespace

extern "C" {

int fmm2d_bhfmm2d(int nd, double eps, int64_t ns, const double* sources, int ifcharge,
                  const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dip, int iper,
                  int ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, int64_t nt,
                  const double* targ, int ifpghtarg, fmm2d_cplx* pottarg,
                  fmm2d_cplx* gradtarg, fmm2d_cplx* hesstarg, fmm2d_opts* opts, void* stream) {
  (void)hess;
  (void)hesstarg;
  if (iper != 0) return FMM2D_ERR_UNSUPPORTED;
  BhArgs a{};
  a.nd = nd;
  a.eps = eps;
  a.ns = ns;
  a.src = sources;
  a.ifcharge = ifcharge;
  a.charge = reinterpret_cast<const cplx*>(charge);
  a.ifdipole = ifdipole;
  a.dip = reinterpret_cast<const cplx*>(dip);
  a.ifpgh = ifpgh;
  a.pot = reinterpret_cast<cplx*>(pot);
  a.grad = reinterpret_cast<cplx*>(grad);
  a.nt = nt;
  a.targ = targ;
  a.ifpghtarg = ifpghtarg;
  a.pottarg = reinterpret_cast<cplx*>(pottarg);
  a.gradtarg = reinterpret_cast<cplx*>(gradtarg);
  a.opts = opts;
  keep_pool_memory();
  return bh_fmm(a, reinterpret_cast<hipStream_t>(stream));
}

int fmm2d_stfmm2d(int nd, double eps, int64_t ns, const double* sources, int ifstoklet,
                  const double* stoklet, int ifstrslet, const double* strslet,
                  const double* strsvec, int ifppreg, double* pot, double* pre, double* grad,
                  int64_t nt, const double* targ, int ifppregtarg, double* pottarg,
                  double* pretarg, double* gradtarg, fmm2d_opts* opts, void* stream) {
  StArgs a{};
  a.nd = nd;
  a.eps = eps;
  a.ns = ns;
  a.src = sources;
  a.ifstoklet = ifstoklet;
  a.stoklet = stoklet;
  a.ifstrslet = ifstrslet;
  a.strslet = strslet;
  a.strsvec = strsvec;
  a.ifppreg = ifppreg;
  a.pot = pot;
  a.pre = pre;
  a.grad = grad;
  a.nt = nt;
  a.targ = targ;
  a.ifppregtarg = ifppregtarg;
  a.pottarg = pottarg;
  a.pretarg = pretarg;
  a.gradtarg = gradtarg;
  a.opts = opts;
  keep_pool_memory();
  return st_fmm(a, reinterpret_cast<hipStream_t>(stream));
}

int fmm2d_cfmm2d(int nd, double eps, int64_t ns, const double* sources, int ifcharge,
                 const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dipstr, int iper,
                 int ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, int64_t nt,
                 const double* targ, int ifpghtarg, fmm2d_cplx* pottarg, fmm2d_cplx* gradtarg,
                 fmm2d_cplx* hesstarg, fmm2d_opts* opts, void* stream) {
  if (iper != 0) return FMM2D_ERR_UNSUPPORTED;
  CauchyArgs a = cauchy_args(nd, eps, ns, sources, ifcharge, charge, ifdipole, dipstr, nullptr,
                             ifpgh, pot, grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg,
                             hesstarg, opts);
  keep_pool_memory();
  return laplace_fmm(0, a, reinterpret_cast<hipStream_t>(stream));
}

int fmm2d_lfmm2d(int nd, double eps, int64_t ns, const double* sources, int ifcharge,
                 const fmm2d_cplx* charge, int ifdipole, const fmm2d_cplx* dipstr,
                 const double* dipvec, int iper, int ifpgh, fmm2d_cplx* pot, fmm2d_cplx* grad,
                 fmm2d_cplx* hess, int64_t nt, const double* targ, int ifpghtarg,
                 fmm2d_cplx* pottarg, fmm2d_cplx* gradtarg, fmm2d_cplx* hesstarg,
                 fmm2d_opts* opts, void* stream) {
  if (iper != 0) return FMM2D_ERR_UNSUPPORTED;
  CauchyArgs a = cauchy_args(nd, eps, ns, sources, ifcharge, charge, ifdipole, dipstr, dipvec,
                             ifpgh, pot, grad, hess, nt, targ, ifpghtarg, pottarg, gradtarg,
                             hesstarg, opts);
  keep_pool_memory();
  return laplace_fmm(1, a, reinterpret_cast<hipStream_t>(stream));
}

int fmm2d_rfmm2d(int nd, double eps, int64_t ns, const double* sources, int ifcharge,
                 const double* charge, int ifdipole, const double* dipstr, const double* dipvec,
                 int iper, int ifpgh, double* pot, double* grad, double* hess, int64_t nt,
                 const double* targ, int ifpghtarg, double* pottarg, double* gradtarg,
                 double* hesstarg, fmm2d_opts* opts, void* stream) {
  if (iper != 0) return FMM2D_ERR_UNSUPPORTED;
  CauchyArgs a = cauchy_args(nd, eps, ns, sources, ifcharge,
                             reinterpret_cast<const fmm2d_cplx*>(charge), ifdipole,
                             reinterpret_cast<const fmm2d_cplx*>(dipstr), dipvec, ifpgh,
                             reinterpret_cast<fmm2d_cplx*>(pot), reinterpret_cast<fmm2d_cplx*>(grad),
                             reinterpret_cast<fmm2d_cplx*>(hess), nt, targ, ifpghtarg,
                             reinterpret_cast<fmm2d_cplx*>(pottarg),
                             reinterpret_cast<fmm2d_cplx*>(gradtarg),
                             reinterpret_cast<fmm2d_cplx*>(hesstarg), opts);
  keep_pool_memory();
  return laplace_fmm(2, a, reinterpret_cast<hipStream_t>(stream));
}

int fmm2d_hfmm2d(int nd, double eps, double zk_re, double zk_im, int64_t ns,
                 const double* sources, int ifcharge, const fmm2d_cplx* charge, int ifdipole,
                 const fmm2d_cplx* dipstr, const double* dipvec, int iper, int ifpgh,
                 fmm2d_cplx* pot, fmm2d_cplx* grad, fmm2d_cplx* hess, int64_t nt,
                 const double* targ, int ifpghtarg, fmm2d_cplx* pottarg,
                 fmm2d_cplx* gradtarg, fmm2d_cplx* hesstarg, fmm2d_opts* opts, void* stream) {
  if (iper != 0) return FMM2D_ERR_UNSUPPORTED;
  HelmArgs a;
  a.nd = nd;
  a.eps = eps;
  a.zk = cplx{zk_re, zk_im};
  a.ns = ns;
  a.nt = nt;
  a.src = sources;
  a.targ = targ;
  a.ifcharge = ifcharge;
  a.ifdipole = ifdipole;
  a.charge = reinterpret_cast<const cplx*>(charge);
  a.dipstr = reinterpret_cast<const cplx*>(dipstr);
  a.dipvec = dipvec;
  a.ifpgh = ifpgh;
  a.ifpghtarg = ifpghtarg;
  a.pot = reinterpret_cast<cplx*>(pot);
  a.grad = reinterpret_cast<cplx*>(grad);
  a.hess = reinterpret_cast<cplx*>(hess);
  a.pottarg = reinterpret_cast<cplx*>(pottarg);
  a.gradtarg = reinterpret_cast<cplx*>(gradtarg);
  a.hesstarg = reinterpret_cast<cplx*>(hesstarg);
  a.opts = opts;
  keep_pool_memory();
  return helmholtz_fmm(a, reinterpret_cast<hipStream_t>(stream));
}

int fmm2d_hfmm2d_mps(int nd, double eps, double zk_re, double zk_im, int64_t nmpole,
                     const double* cmpole, const double* rmpole, const int* mterms,
                     int64_t ntot, const fmm2d_cplx* mpole, const int64_t* impole,
                     fmm2d_cplx* local, fmm2d_opts* opts, void* stream) {
  HelmMpsArgs a;
  a.nd = nd;
  a.eps = eps;
  a.zk = cplx{zk_re, zk_im};
  a.n = nmpole;
  a.ntot = ntot;
  a.centers = cmpole;
  a.rscale = rmpole;
  a.mterms = mterms;
  a.mpole = reinterpret_cast<const cplx*>(mpole);
  a.impole = impole;
  a.local = reinterpret_cast<cplx*>(local);
  a.opts = opts;
  keep_pool_memory();
  return helmholtz_mps(a, reinterpret_cast<hipStream_t>(stream));
}

int fmm2d_tree_build(const double* sources, int64_t ns, const double* targ, int64_t nt,
                     int idivflag, int ndiv, int nlmax, int* nlevels, int* nboxes,
                     void** handle, void* stream) {
  if (!handle || ns < 1 || nt < 0 || (nt > 0 && !targ) || !sources || ndiv < 1 ||
      idivflag < 0 || idivflag > 2 || nlmax < 1)
    return FMM2D_ERR_ARG;
  TreeHandle* h = new TreeHandle;
  h->st = reinterpret_cast<hipStream_t>(stream);
  try {
    keep_pool_memory();
    build_tree(sources, ns, targ, nt, ndiv, idivflag, nlmax, h->t, h->st);
    hip_throw(hipStreamSynchronize(h->st), "sync");
  } catch (const HipError& e) {
    delete h;
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  } catch (const std::bad_alloc&) {
    delete h;
    return FMM2D_ERR_ALLOC;
  } catch (...) {
    delete h;
    return FMM2D_ERR_HIP;
  }
  if (nlevels) *nlevels = h->t.nlevels;
  if (nboxes) *nboxes = h->t.nboxes;
  *handle = h;
  return FMM2D_OK;
}

int fmm2d_tree_leaf_histogram(const double* sources, int64_t ns, const double* targ, int64_t nt,
                              int idivflag, int ndiv, int nbins, double* hist, double* boxsize0,
                              void* stream) {
  if (!sources || ns < 1 || nt < 0 || (nt > 0 && !targ) || ndiv < 1 || idivflag < 0 ||
      idivflag > 2 || nbins < 1 || nbins > kLeafHistMax || !hist)
    return FMM2D_ERR_ARG;
  try {
    keep_pool_memory();
    const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    Tree t;
    build_tree(sources, ns, targ, nt, ndiv, idivflag, 51, t, st);
    leaf_level_histogram(t, nbins, hist, st);
    if (boxsize0) *boxsize0 = t.boxsize[0];
  } catch (const HipError& e) {
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return FMM2D_ERR_ALLOC;
  } catch (...) {
    return FMM2D_ERR_HIP;
  }
  return FMM2D_OK;
}

int fmm2d_key_leaf_histogram(const double* sources, int64_t ns, const double* targ, int64_t nt,
                             int ndiv, int nbins, double* hist, double* boxsize0, void* stream) {
  if (!sources || ns < 1 || nt < 0 || (nt > 0 && !targ) || ndiv < 1 || nbins < 1 ||
      nbins > kLeafHistMax || !hist || ns > INT32_MAX || nt > INT32_MAX)
    return FMM2D_ERR_ARG;
  try {
    keep_pool_memory();
    key_leaf_histogram(sources, ns, targ, nt, ndiv, nbins, hist, boxsize0,
                       reinterpret_cast<hipStream_t>(stream));
  } catch (const HipError& e) {
    fprintf(stderr, "fmm2d_amd: %s failed: %s\n", e.what, hipGetErrorString(e.err));
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return FMM2D_ERR_ALLOC;
  } catch (...) {
    return FMM2D_ERR_HIP;
  }
  return FMM2D_OK;
}

int fmm2d_prepare(const double* sources, int64_t ns, const double* targ, int64_t nt,
                  void** handle, void* stream) {
  if (!handle || !sources || ns < 1 || nt < 0 || (nt > 0 && !targ) || ns > INT32_MAX ||
      nt > INT32_MAX)
    return FMM2D_ERR_ARG;
  *handle = nullptr;
  try {
    keep_pool_memory();
    // nlmax 51: the depth limit of the FMM drivers' trees
    *handle = tree_prepare(sources, ns, nt > 0 ? targ : nullptr, nt, 51,
                           reinterpret_cast<hipStream_t>(stream));
  } catch (const HipError& e) {
    fprintf(stderr, "fmm2d_amd: %s failed: %s\n", e.what, hipGetErrorString(e.err));
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return FMM2D_ERR_ALLOC;
  } catch (...) {
    return FMM2D_ERR_HIP;
  }
  return FMM2D_OK;
}

void fmm2d_prepare_free(void* handle) { tree_prepare_free(static_cast<TreePrep*>(handle)); }

int fmm2d_prepared_leaf_histogram(void* handle, int ndiv, int nbins, double* hist,
                                  double* boxsize0, void* stream) {
  if (!handle || ndiv < 1 || nbins < 1 || nbins > kLeafHistMax || !hist) return FMM2D_ERR_ARG;
  try {
    keep_pool_memory();
    prepared_leaf_histogram(static_cast<const TreePrep*>(handle), ndiv, nbins, hist, boxsize0,
                            reinterpret_cast<hipStream_t>(stream));
  } catch (const HipError& e) {
    fprintf(stderr, "fmm2d_amd: %s failed: %s\n", e.what, hipGetErrorString(e.err));
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return FMM2D_ERR_ALLOC;
  } catch (...) {
    return FMM2D_ERR_HIP;
  }
  return FMM2D_OK;
}


int fmm2d_tree_get(void* handle, int* laddr, int* ilevel, int* iparent, int* nchild,
                   int* ichild, int* ncoll, int* coll, double* centers, double* boxsize,
                   int* isrc, int* isrcse, int* itarg, int* itargse, int* nlist1, int* list1,
                   int* nlist2, int* list2, int* nlist3, int* list3, int* nlist4, int* list4) {
  if (!handle) return FMM2D_ERR_ARG;
  TreeHandle* h = static_cast<TreeHandle*>(handle);
  Tree& t = h->t;
  const int nb = t.nboxes, nl = t.nlevels;
  auto get = [&](const DBuf<int>& d, size_t n, int* out, bool plus1) {
    if (!out || n == 0) return;
    hip_throw(hipMemcpyAsync(out, d.p, sizeof(int) * n, hipMemcpyDeviceToHost, h->st), "copy");
    hip_throw(hipStreamSynchronize(h->st), "sync");
    if (plus1)
      for (size_t i = 0; i < n; ++i) out[i] = out[i] < 0 ? -1 : out[i] + 1;
  };
  try {
    if (laddr)
      for (int l = 0; l <= nl; ++l) {
        laddr[2 * l] = t.lstart[l] + 1;
        laddr[2 * l + 1] = t.lend[l];
      }
    if (boxsize)
      for (int l = 0; l <= nl; ++l) boxsize[l] = t.boxsize[l];
    get(t.ilevel, nb, ilevel, false);
    get(t.iparent, nb, iparent, true);
    get(t.nchild, nb, nchild, false);
    get(t.ichild, 4 * (size_t)nb, ichild, true);
    get(t.ncoll, nb, ncoll, false);
    get(t.coll, 9 * (size_t)nb, coll, true);
    if (centers) {
      hip_throw(hipMemcpyAsync(centers, t.centers.p, sizeof(double) * 2 * nb,
                               hipMemcpyDeviceToHost, h->st), "copy");
      hip_throw(hipStreamSynchronize(h->st), "sync");
    }
    get(t.isrc, t.ns, isrc, true);
    get(t.itarg, t.nt, itarg, true);
    // [start, end) 0-based -> (first, last) 1-based
    for (int k = 0; k < 2; ++k) {
      int* out = k ? itargse : isrcse;
      if (!out) continue;
      get(k ? t.itargse : t.isrcse, 2 * (size_t)nb, out, false);
      for (int b = 0; b < nb; ++b) out[2 * b] += 1;
    }
    if (nlist1) {
      if (!h->have_lists) {
        build_lists(t, h->l, h->st);
        h->have_lists = true;
      }
      const int caps[4] = {kMaxList1, kMaxList2, kMaxList3, kMaxList4};
      int* ns[4] = {nlist1, nlist2, nlist3, nlist4};
      int* ls[4] = {list1, list2, list3, list4};
      const DBuf<int>* dn[4] = {&h->l.n1, &h->l.n2, &h->l.n3, &h->l.n4};
      const DBuf<int>* dl[4] = {&h->l.l1, &h->l.l2, &h->l.l3, &h->l.l4};
      for (int k = 0; k < 4; ++k) {
        if (!ns[k] || !ls[k]) continue;
        get(*dn[k], nb, ns[k], false);
        get(*dl[k], (size_t)caps[k] * nb, ls[k], true);
        for (int b = 0; b < nb; ++b) {
          int* row = ls[k] + (size_t)caps[k] * b;
          std::sort(row, row + ns[k][b]);
          for (int j = ns[k][b]; j < caps[k]; ++j) row[j] = 0;
        }
      }
    }
  } catch (const HipError&) {
    return FMM2D_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return FMM2D_ERR_ALLOC;
  } catch (...) {
    return FMM2D_ERR_HIP;
  }
  return FMM2D_OK;
}

void fmm2d_tree_free(void* handle) {
  TreeHandle* h = static_cast<TreeHandle*>(handle);
  if (!h) return;
  (void)hipStreamSynchronize(h->st);
  delete h;
}

int fmm2d_l2dterms(double eps, int* nterms) {
  if (!nterms) return FMM2D_ERR_ARG;
  *nterms = l2dterms(eps);
  return FMM2D_OK;
}

int fmm2d_h2dterms(double bsize, double zk_re, double zk_im, double eps, int* nterms) {
  if (!nterms) return FMM2D_ERR_ARG;
  int ier = 0;
  *nterms = h2dterms(bsize, cplx{zk_re, zk_im}, eps, &ier);
  return ier;
}

int fmm2d_ndiv(double eps, int64_t ns, int64_t nt, int helmholtz, int* ndiv, int* idivflag) {
  (void)helmholtz;  // lndiv2d and hndiv2d are the same table
  if (!ndiv || !idivflag) return FMM2D_ERR_ARG;
  *ndiv = lndiv2d(eps, ns, nt);
  *idivflag = 0;
  return FMM2D_OK;
}

int fmm2d_release_workspace(void) { return ws_release(); }

int fmm2d_set_workspace_limit(int64_t bytes) {
  if (bytes < 0) return FMM2D_ERR_ARG;
  ws_set_limit((size_t)bytes);
  return FMM2D_OK;
}

int fmm2d_batch(int njobs, fmm2d_job* jobs, int nstreams) {
  if (njobs < 0 || (njobs > 0 && !jobs) || nstreams < 1 || nstreams > 16) return FMM2D_ERR_ARG;
  if (njobs == 0) return FMM2D_OK;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return FMM2D_ERR_HIP;
  std::vector<hipStream_t> streams;
  try {
    streams = batch_streams(dev, std::min(nstreams, njobs));
  } catch (const HipError&) {
    return FMM2D_ERR_HIP;
  }
  std::vector<int> order(njobs);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return jobs[a].ns + jobs[a].nt > jobs[b].ns + jobs[b].nt;
  });
  std::atomic<int> next{0};
  std::vector<int> sync_err(streams.size(), 0);
  auto worker = [&](int w) {
    (void)hipSetDevice(dev);
    for (int k; (k = next.fetch_add(1)) < njobs;) {
      fmm2d_job& j = jobs[order[k]];
      try {
        j.ier = run_job(j, streams[w]);
      } catch (const std::bad_alloc&) {
        j.ier = FMM2D_ERR_ALLOC;
      } catch (...) {
        j.ier = FMM2D_ERR_HIP;
      }
    }
    if (hipStreamSynchronize(streams[w]) != hipSuccess) sync_err[w] = 1;
  };
  std::vector<std::thread> th;
  for (int w = 1; w < (int)streams.size(); ++w) th.emplace_back(worker, w);
  worker(0);
  for (auto& t : th) t.join();
  for (int e : sync_err)
    if (e) return FMM2D_ERR_HIP;
  for (int k = 0; k < njobs; ++k)
    if (jobs[k].ier != 0) return jobs[k].ier;
  return FMM2D_OK;
}

}  // extern "C"

namespace {

int run_job(fmm2d_job& j, hipStream_t st) {
  void* s = reinterpret_cast<void*>(st);
  auto* ch = static_cast<const fmm2d_cplx*>(j.charge);
  auto* ds = static_cast<const fmm2d_cplx*>(j.dipstr);
  auto* p = static_cast<fmm2d_cplx*>(j.pot);
  auto* g = static_cast<fmm2d_cplx*>(j.grad);
  auto* h = static_cast<fmm2d_cplx*>(j.hess);
  auto* pt = static_cast<fmm2d_cplx*>(j.pottarg);
  auto* gt = static_cast<fmm2d_cplx*>(j.gradtarg);
  auto* ht = static_cast<fmm2d_cplx*>(j.hesstarg);
  switch (j.kind) {
    case 0:
      return fmm2d_cfmm2d(j.nd, j.eps, j.ns, j.sources, j.ifcharge, ch, j.ifdipole, ds, 0,
                          j.ifpgh, p, g, h, j.nt, j.targ, j.ifpghtarg, pt, gt, ht, j.opts, s);
    case 1:
      return fmm2d_lfmm2d(j.nd, j.eps, j.ns, j.sources, j.ifcharge, ch, j.ifdipole, ds,
                          j.dipvec, 0, j.ifpgh, p, g, h, j.nt, j.targ, j.ifpghtarg, pt, gt, ht,
                          j.opts, s);
    case 2:
      return fmm2d_rfmm2d(j.nd, j.eps, j.ns, j.sources, j.ifcharge,
                          static_cast<const double*>(j.charge), j.ifdipole,
                          static_cast<const double*>(j.dipstr), j.dipvec, 0, j.ifpgh,
                          static_cast<double*>(j.pot), static_cast<double*>(j.grad),
                          static_cast<double*>(j.hess), j.nt, j.targ, j.ifpghtarg,
                          static_cast<double*>(j.pottarg), static_cast<double*>(j.gradtarg),
                          static_cast<double*>(j.hesstarg), j.opts, s);
    case 3:
      return fmm2d_hfmm2d(j.nd, j.eps, j.zk_re, j.zk_im, j.ns, j.sources, j.ifcharge, ch,
                          j.ifdipole, ds, j.dipvec, 0, j.ifpgh, p, g, h, j.nt, j.targ,
                          j.ifpghtarg, pt, gt, ht, j.opts, s);
    default:
      return FMM2D_ERR_ARG;
  }
}

}  // namespace

// Adaptive quadtree built on the GPU, box-for-box identical to the reference.
//
// Reference (common/pts_tree2d.f, common/tree_routs2d.f):
//   pts_tree_build 406-655   level-by-level refinement of boxes holding more
//                            than ndiv points, 4 children each, numbered in
//                            parent order (tree_refine_boxes 9-56)
//   computecoll 100-211      colleagues: same-level boxes with centres within
//                            1.05*bs (a single-precision 1.05, promoted)
//   pts_tree_fix_lr 758-1061 level restriction: flag / flag+ passes, refine,
//                            reorganise, then flag++ (updateflags 217-326)
//   pts_tree_reorg 1066-1215 per level: original boxes first, then new ones
//   pts_tree_sort 1222-1246  stable recursive partition of the points
//
// GPU formulation: one thread per point for classification into children
// (y - cy < 0 then x - cx < 0, as sort_pts_to_children 659-754), one thread
// per box for every box-level step, hipCUB scans where the reference uses
// cumsum, and the final point order as a stable radix sort of the points by
// the depth-first start offset of their leaf -- which is exactly the order the
// reference's recursive stable partition produces.  Box-level steps run level
// by level where the reference's loops carry a dependency between levels.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "tree.h"

namespace fmm2d {
namespace {

// The reference writes `1.05*boxsize(ilev)` with a default-real literal in
// computecoll / computelists: 1.05 rounded to single precision, then promoted.
constexpr double kColl = (double)1.05f;

struct Boxes {
  DBuf<double> centers;
  DBuf<int> ilevel, iparent, nchild, ichild, ncoll, coll, cs, ct, iflag;
  size_t cap = 0;
  void alloc(size_t c, hipStream_t st) {
    cap = c;
    centers.alloc(2 * c, st);
    ilevel.alloc(c, st);
    iparent.alloc(c, st);
    nchild.alloc(c, st);
    ichild.alloc(4 * c, st);
    ncoll.alloc(c, st);
    coll.alloc(9 * c, st);
    cs.alloc(c, st);
    ct.alloc(c, st);
    iflag.alloc(c, st);
  }
  void grow(size_t need, size_t keep) {
    if (need <= cap) return;
    size_t c = std::max(need, cap * 2);
    centers.grow(2 * c, 2 * keep);
    ilevel.grow(c, keep);
    iparent.grow(c, keep);
    nchild.grow(c, keep);
    ichild.grow(4 * c, 4 * keep);
    ncoll.grow(c, keep);
    coll.grow(9 * c, 9 * keep);
    cs.grow(c, keep);
    ct.grow(c, keep);
    iflag.grow(c, keep);
    cap = c;
  }
};

// ------------------------------------------------------------ bounding box
constexpr int kRedBlocks = 1024;

__global__ void k_bbox_partial(const double* __restrict__ src, int64_t ns,
                               const double* __restrict__ trg, int64_t nt, double* part) {
  __shared__ double sm[4][256];
  double v[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
  const int64_t n = ns + nt;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double* p = (i < ns) ? (src + 2 * i) : (trg + 2 * (i - ns));
    double x = p[0], y = p[1];
    v[0] = fmin(v[0], x);
    v[1] = fmax(v[1], x);
    v[2] = fmin(v[2], y);
    v[3] = fmax(v[3], y);
  }
  for (int k = 0; k < 4; ++k) sm[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sm[0][threadIdx.x] = fmin(sm[0][threadIdx.x], sm[0][threadIdx.x + s]);
      sm[1][threadIdx.x] = fmax(sm[1][threadIdx.x], sm[1][threadIdx.x + s]);
      sm[2][threadIdx.x] = fmin(sm[2][threadIdx.x], sm[2][threadIdx.x + s]);
      sm[3][threadIdx.x] = fmax(sm[3][threadIdx.x], sm[3][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = sm[threadIdx.x][0];
}

__global__ void k_bbox_final(const double* part, int nb, double* out) {
  __shared__ double sm[4][256];
  double v[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
  for (int i = threadIdx.x; i < nb; i += 256) {
    v[0] = fmin(v[0], part[4 * i]);
    v[1] = fmax(v[1], part[4 * i + 1]);
    v[2] = fmin(v[2], part[4 * i + 2]);
    v[3] = fmax(v[3], part[4 * i + 3]);
  }
  for (int k = 0; k < 4; ++k) sm[k][threadIdx.x] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 256; ++i) {
      v[0] = fmin(v[0], sm[0][i]);
      v[1] = fmax(v[1], sm[1][i]);
      v[2] = fmin(v[2], sm[2][i]);
      v[3] = fmax(v[3], sm[3][i]);
    }
    for (int k = 0; k < 4; ++k) out[k] = v[k];
  }
}

// ------------------------------------------------------------ refinement
__global__ void k_refine_flags(int64_t n, int lo, const int* cs, const int* ct, int idivflag,
                               int ndiv, int* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = lo + (int)i;
  int nn = cs[b];
  if (idivflag == 1) nn = ct[b];
  if (idivflag == 2) nn = max(cs[b], ct[b]);
  flag[i] = nn > ndiv ? 1 : 0;
}

// flag[i] = iflag[lo+i] > 0 (tree_refine_boxes_flag, tree_routs2d.f:334-388)
__global__ void k_iflag_pos(int64_t n, int lo, const int* iflag, int* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = iflag[lo + (int)i] > 0 ? 1 : 0;
}

__global__ void k_total(int n, const int* flag, const int* offs, int* total) {
  *total = (n > 0) ? offs[n - 1] + flag[n - 1] : 0;
}

// Children of the flagged boxes lo..lo+n-1 are numbered base + 4*offs + q, q =
// 0..3 for (-,-), (+,-), (-,+), (+,+) around the parent centre
// (tree_refine_boxes, tree_routs2d.f:25-52).  mode 0: first refinement (child
// counts zeroed); mode 1: level-restriction refinement (child iflag = 3 when
// the parent's flag is 1, else 0).
__global__ void k_make_children(int64_t n, int lo, const int* flag, const int* offs, int base,
                                double half, int lev, double* centers, int* ilevel, int* iparent,
                                int* nchild, int* ichild, int* cs, int* ct, int* iflag, int mode) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const int b = lo + (int)i;
  const int c0 = base + 4 * offs[i];
  const double cx = centers[2 * b], cy = centers[2 * b + 1];
  nchild[b] = 4;
  const int pf = (mode == 1) ? iflag[b] : 0;
  for (int q = 0; q < 4; ++q) {
    const int c = c0 + q;
    centers[2 * c] = cx + ((q & 1) ? half : -half);
    centers[2 * c + 1] = cy + ((q & 2) ? half : -half);
    iparent[c] = b;
    nchild[c] = 0;
    ichild[4 * c] = ichild[4 * c + 1] = ichild[4 * c + 2] = ichild[4 * c + 3] = -1;
    ichild[4 * b + q] = c;
    ilevel[c] = lev;
    cs[c] = 0;
    ct[c] = 0;
    iflag[c] = (mode == 1 && pf == 1) ? 3 : 0;
  }
}

// Move each point of a just-refined box into its child and count it.
__global__ void k_classify(int64_t n, const double* __restrict__ xy, int* pbox, const int* nchild,
                           const int* ichild, const double* centers, int* cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = pbox[i];
  if (nchild[b] == 0) return;
  const double x = xy[2 * i], y = xy[2 * i + 1];
  const int q = ((y - centers[2 * b + 1] < 0.0) ? 0 : 2) + ((x - centers[2 * b] < 0.0) ? 0 : 1);
  const int c = ichild[4 * b + q];
  pbox[i] = c;
  atomicAdd(&cnt[c], 1);
}

// ------------------------------------------------------------ colleagues
// computecoll (tree_routs2d.f:100-211) for boxes lo..lo+n-1 of one level.
__global__ void k_coll(int64_t n, int lo, double bs, const int* iparent, const int* ichild,
                       const double* centers, int* ncoll, int* coll) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = lo + (int)i;
  const int dad = iparent[b];
  const double cx = centers[2 * b], cy = centers[2 * b + 1];
  const double lim = kColl * bs;
  int cnt = 0;
  const int nd = ncoll[dad];
  for (int k = 0; k < nd; ++k) {
    const int j = coll[9 * dad + k];
    for (int q = 0; q < 4; ++q) {
      const int kb = ichild[4 * j + q];
      if (kb >= 0 && fabs(centers[2 * kb] - cx) <= lim && fabs(centers[2 * kb + 1] - cy) <= lim) {
        if (cnt < kMaxColl) coll[9 * b + cnt] = kb;
        ++cnt;
      }
    }
  }
  ncoll[b] = min(cnt, kMaxColl);
  for (int k = cnt; k < kMaxColl; ++k) coll[9 * b + k] = -1;
}

__global__ void k_root_coll(int* ncoll, int* coll) {
  ncoll[0] = 1;
  coll[0] = 0;
  for (int k = 1; k < 9; ++k) coll[k] = -1;
}

// ------------------------------------------------------------ level restriction
// flag pass (pts_tree2d.f:797-839): leaves among the grandparent's colleagues
// that touch the parent of a box at level lev.
__global__ void k_flag1(int64_t n, int lo, double distest, const int* iparent, const int* ncoll,
                        const int* coll, const int* nchild, const double* centers, int* iflag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = lo + (int)i;
  const int dad = iparent[b];
  const int gd = iparent[dad];
  const double dx0 = centers[2 * dad], dy0 = centers[2 * dad + 1];
  for (int k = 0; k < ncoll[gd]; ++k) {
    const int j = coll[9 * gd + k];
    if (nchild[j] == 0 && iflag[j] == 0) {
      if (fabs(centers[2 * j] - dx0) <= distest && fabs(centers[2 * j + 1] - dy0) <= distest)
        iflag[j] = 1;
    }
  }
}

// flag+ pass (pts_tree2d.f:845-889): leaves among the parent's colleagues
// touching a flagged box.
__global__ void k_flag2(int64_t n, int lo, double distest, const int* iparent, const int* ncoll,
                        const int* coll, const int* nchild, const double* centers, int* iflag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = lo + (int)i;
  const int f = iflag[b];
  if (f != 1 && f != 2) return;
  const int dad = iparent[b];
  const double bx = centers[2 * b], by = centers[2 * b + 1];
  for (int k = 0; k < ncoll[dad]; ++k) {
    const int j = coll[9 * dad + k];
    if (nchild[j] == 0 && iflag[j] == 0) {
      if (fabs(centers[2 * j] - bx) <= distest && fabs(centers[2 * j + 1] - by) <= distest)
        iflag[j] = 2;
    }
  }
}

// updateflags (tree_routs2d.f:217-326)
__global__ void k_updateflags(int64_t n, int lo, double distest, const int* ncoll, const int* coll,
                              const int* ichild, const int* nchild, const double* centers,
                              int* iflag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = lo + (int)i;
  if (iflag[b] != 3) return;
  int f = 0;
  const double bx = centers[2 * b], by = centers[2 * b + 1];
  for (int k = 0; k < ncoll[b] && !f; ++k) {
    const int j = coll[9 * b + k];
    for (int q = 0; q < 4; ++q) {
      const int kb = ichild[4 * j + q];
      if (kb >= 0 && nchild[kb] > 0 && fabs(centers[2 * kb] - bx) <= distest &&
          fabs(centers[2 * kb + 1] - by) <= distest) {
        f = 1;
        break;
      }
    }
  }
  iflag[b] = f;
}

__global__ void k_keep3(int64_t n, int* iflag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (iflag[i] != 3) iflag[i] = 0;
}

// ------------------------------------------------------------ reorganisation
__global__ void k_map_range(int64_t n, int from, int to, int* map) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  map[from + i] = to + (int)i;
}

__global__ void k_reorg_apply(int64_t n, const int* map, const double* c0, const int* lev0,
                              const int* par0, const int* nch0, const int* ich0, const int* fl0,
                              double* c1, int* lev1, int* par1, int* nch1, int* ich1, int* fl1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = map[i];
  c1[2 * m] = c0[2 * i];
  c1[2 * m + 1] = c0[2 * i + 1];
  lev1[m] = lev0[i];
  nch1[m] = nch0[i];
  fl1[m] = fl0[i];
  par1[m] = par0[i] < 0 ? -1 : map[par0[i]];
  for (int q = 0; q < 4; ++q) ich1[4 * m + q] = ich0[4 * i + q] < 0 ? -1 : map[ich0[4 * i + q]];
}

__global__ void k_remap(int64_t n, const int* map, int* v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  v[i] = map[v[i]];
}

// ------------------------------------------------------------ final point order
// Descend to the leaf of the final tree and count points per leaf.
__global__ void k_descend(int64_t n, const double* __restrict__ xy, int* pbox, const int* nchild,
                          const int* ichild, const double* centers, int* cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int b = pbox[i];
  const double x = xy[2 * i], y = xy[2 * i + 1];
  while (nchild[b] > 0) {
    const int q = ((y - centers[2 * b + 1] < 0.0) ? 0 : 2) + ((x - centers[2 * b] < 0.0) ? 0 : 1);
    b = ichild[4 * b + q];
  }
  pbox[i] = b;
  atomicAdd(&cnt[b], 1);
}

__global__ void k_sum_children(int64_t n, int lo, const int* nchild, const int* ichild, int* cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = lo + (int)i;
  if (nchild[b] == 0) return;
  int s = 0;
  for (int q = 0; q < 4; ++q) s += cnt[ichild[4 * b + q]];
  cnt[b] = s;
}

// se[2b] = first sorted position of box b, se[2b+1] = one past its last.
__global__ void k_starts(int64_t n, int lo, const int* nchild, const int* ichild, const int* cnt,
                         int* se) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = lo + (int)i;
  if (nchild[b] == 0) return;
  int s = se[2 * b];
  for (int q = 0; q < 4; ++q) {
    const int c = ichild[4 * b + q];
    se[2 * c] = s;
    se[2 * c + 1] = s + cnt[c];
    s += cnt[c];
  }
}

__global__ void k_root_se(int* se, int n) {
  se[0] = 0;
  se[1] = n;
}

__global__ void k_keys(int64_t n, const int* pbox, const int* se, int* key, int* val) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = se[2 * pbox[i]];
  val[i] = (int)i;
}

// ------------------------------------------------------------ host helpers
struct Builder {
  hipStream_t st;
  Boxes B;
  int nboxes = 0;
  DBuf<int> flag, offs, total;
  int* h_total = nullptr;

  explicit Builder(hipStream_t s) : st(s) {
    total.alloc(1, s);
    hip_throw(hipHostMalloc((void**)&h_total, sizeof(int)), "hipHostMalloc");
  }
  ~Builder() {
    if (h_total) (void)hipHostFree(h_total);
  }

  // exclusive scan of flag[0..n) into offs, returns the total
  int scan(int n) {
    if (n <= 0) return 0;
    size_t bytes = 0;
    hip_throw(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, flag.p, offs.p, n, st), "scan");
    DBuf<char> tmp(bytes, st);
    hip_throw(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, flag.p, offs.p, n, st), "scan");
    hipLaunchKernelGGL(k_total, dim3(1), dim3(1), 0, st, n, flag.p, offs.p, total.p);
    hip_throw(hipMemcpyAsync(h_total, total.p, sizeof(int), hipMemcpyDeviceToHost, st), "copy");
    hip_throw(hipStreamSynchronize(st), "sync");
    return *h_total;
  }
  void ensure_flag(int n) {
    if ((size_t)n > flag.n) {
      flag.alloc(std::max<size_t>(n, 2 * flag.n), st);
      offs.alloc(flag.n, st);
    }
  }
  // refine the boxes lo..lo+n-1 selected by flag[]; returns the count refined
  int refine(int lo, int n, double bs_child, int lev, int mode) {
    const int nref = scan(n);
    if (nref == 0) return 0;
    B.grow(nboxes + 4 * (size_t)nref, nboxes);
    launch1d(k_make_children, n, st, lo, flag.p, offs.p, nboxes, 0.5 * bs_child, lev,
             B.centers.p, B.ilevel.p, B.iparent.p, B.nchild.p, B.ichild.p, B.cs.p, B.ct.p,
             B.iflag.p, mode);
    nboxes += 4 * nref;
    return nref;
  }
  void computecoll(const std::vector<int>& ls, const std::vector<int>& le,
                   const std::vector<double>& bs, int nlevels) {
    hipLaunchKernelGGL(k_root_coll, dim3(1), dim3(1), 0, st, B.ncoll.p, B.coll.p);
    for (int l = 1; l <= nlevels; ++l)
      launch1d(k_coll, le[l] - ls[l], st, ls[l], bs[l], B.iparent.p, B.ichild.p, B.centers.p,
               B.ncoll.p, B.coll.p);
  }
  // pts_tree_reorg: per level >= 2, original boxes then the tail boxes.
  void reorg(std::vector<int>& ls, std::vector<int>& le, const std::vector<int>& ts,
             const std::vector<int>& te, int nlevels, DBuf<int>* pts[2], int64_t npts[2]) {
    DBuf<int> map(nboxes, st);
    for (int l = 0; l <= std::min(1, nlevels); ++l)
      launch1d(k_map_range, le[l] - ls[l], st, ls[l], ls[l], map.p);
    int cur = (nlevels >= 2) ? ls[2] : 0;
    for (int l = 2; l <= nlevels; ++l) {
      const int s = cur;
      launch1d(k_map_range, le[l] - ls[l], st, ls[l], cur, map.p);
      cur += le[l] - ls[l];
      launch1d(k_map_range, te[l] - ts[l], st, ts[l], cur, map.p);
      cur += te[l] - ts[l];
      ls[l] = s;
      le[l] = cur;
    }
    Boxes N;
    N.alloc(B.cap, st);
    launch1d(k_reorg_apply, nboxes, st, map.p, B.centers.p, B.ilevel.p, B.iparent.p, B.nchild.p,
             B.ichild.p, B.iflag.p, N.centers.p, N.ilevel.p, N.iparent.p, N.nchild.p,
             N.ichild.p, N.iflag.p);
    B.centers.swap(N.centers);
    B.ilevel.swap(N.ilevel);
    B.iparent.swap(N.iparent);
    B.nchild.swap(N.nchild);
    B.ichild.swap(N.ichild);
    B.iflag.swap(N.iflag);
    for (int k = 0; k < 2; ++k)
      if (npts[k] > 0) launch1d(k_remap, npts[k], st, map.p, pts[k]->p);
  }
};

}  // namespace

void build_tree(const double* src, int64_t ns, const double* trg, int64_t nt, int ndiv,
                int idivflag, int nlmax, Tree& t, hipStream_t st) {
  Builder bld(st);
  t.ns = ns;
  t.nt = nt;
  // ---- bounding square (pts_tree2d.f:455-517)
  double bb[4];
  {
    const int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, ceil_div(ns + nt, 256)));
    DBuf<double> part(4 * nb, st), out(4, st);
    hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(256), 0, st, src, ns, trg, nt, part.p);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, st, part.p, nb, out.p);
    hip_throw(hipMemcpyAsync(bb, out.p, sizeof(bb), hipMemcpyDeviceToHost, st), "copy");
    hip_throw(hipStreamSynchronize(st), "sync");
  }
  std::vector<double> bs(1);
  bs[0] = bb[1] - bb[0];
  const double sizey = bb[3] - bb[2];
  if (sizey > bs[0]) bs[0] = sizey;
  const double root[2] = {(bb[0] + bb[1]) / 2, (bb[2] + bb[3]) / 2};

  size_t cap0 = std::max<size_t>(1024, (size_t)(8 * (ns + nt) / std::max(ndiv, 1)));
  bld.B.alloc(cap0, st);
  bld.nboxes = 1;
  {
    const int none4[4] = {-1, -1, -1, -1};
    const int zero = 0, mone = -1;
    const int cnt[2] = {(int)ns, (int)nt};
    hip_throw(hipMemcpyAsync(bld.B.centers.p, root, sizeof(root), hipMemcpyHostToDevice, st), "c");
    hip_throw(hipMemcpyAsync(bld.B.ilevel.p, &zero, sizeof(int), hipMemcpyHostToDevice, st), "c");
    hip_throw(hipMemcpyAsync(bld.B.iparent.p, &mone, sizeof(int), hipMemcpyHostToDevice, st), "c");
    hip_throw(hipMemcpyAsync(bld.B.nchild.p, &zero, sizeof(int), hipMemcpyHostToDevice, st), "c");
    hip_throw(hipMemcpyAsync(bld.B.ichild.p, none4, sizeof(none4), hipMemcpyHostToDevice, st), "c");
    hip_throw(hipMemcpyAsync(bld.B.cs.p, &cnt[0], sizeof(int), hipMemcpyHostToDevice, st), "c");
    hip_throw(hipMemcpyAsync(bld.B.ct.p, &cnt[1], sizeof(int), hipMemcpyHostToDevice, st), "c");
    hip_throw(hipMemcpyAsync(bld.B.iflag.p, &zero, sizeof(int), hipMemcpyHostToDevice, st), "c");
    hip_throw(hipStreamSynchronize(st), "sync");
  }
  DBuf<int> pbs(std::max<int64_t>(ns, 1), st), pbt(std::max<int64_t>(nt, 1), st);
  pbs.zero();
  pbt.zero();

  // ---- level-by-level refinement (pts_tree2d.f:551-631)
  std::vector<int> ls{0}, le{1};
  int nlevels = nlmax;
  for (int l = 0; l < nlmax; ++l) {
    const int lo = ls[l], n = le[l] - ls[l];
    bld.ensure_flag(n);
    launch1d(k_refine_flags, n, st, lo, bld.B.cs.p, bld.B.ct.p, idivflag, ndiv, bld.flag.p);
    const int base = bld.nboxes;
    bs.push_back(bs[l] / 2);
    const int nref = bld.refine(lo, n, bs[l + 1], l + 1, 0);
    if (nref == 0) {
      bs.pop_back();
      nlevels = l;
      break;
    }
    ls.push_back(base);
    le.push_back(bld.nboxes);
    launch1d(k_classify, ns, st, src, pbs.p, bld.B.nchild.p, bld.B.ichild.p, bld.B.centers.p,
             bld.B.cs.p);
    launch1d(k_classify, nt, st, trg, pbt.p, bld.B.nchild.p, bld.B.ichild.p, bld.B.centers.p,
             bld.B.ct.p);
  }
  ls.resize(nlevels + 1);
  le.resize(nlevels + 1);
  bs.resize(nlevels + 1);
  bld.computecoll(ls, le, bs, nlevels);

  DBuf<int>* pts[2] = {&pbs, &pbt};
  int64_t npts[2] = {ns, nt};
  // ---- level restriction (pts_tree_fix_lr, pts_tree2d.f:758-1061)
  if (nlevels >= 2) {
    int* iflag = bld.B.iflag.p;
    hip_throw(hipMemsetAsync(iflag, 0, sizeof(int) * bld.nboxes, st), "memset");
    for (int l = nlevels; l >= 2; --l) {
      const double distest = 1.05 * (bs[l - 1] + bs[l - 2]) / 2.0;
      launch1d(k_flag1, le[l] - ls[l], st, ls[l], distest, bld.B.iparent.p, bld.B.ncoll.p,
               bld.B.coll.p, bld.B.nchild.p, bld.B.centers.p, iflag);
    }
    for (int l = nlevels; l >= 1; --l) {
      const double distest = 1.05 * (bs[l] + bs[l - 1]) / 2.0;
      launch1d(k_flag2, le[l] - ls[l], st, ls[l], distest, bld.B.iparent.p, bld.B.ncoll.p,
               bld.B.coll.p, bld.B.nchild.p, bld.B.centers.p, iflag);
    }
    std::vector<int> ts(nlevels + 1, 0), te(nlevels + 1, 0);
    for (int l = 1; l <= nlevels - 2; ++l) {
      ts[l + 1] = bld.nboxes;
      const int n = le[l] - ls[l];
      bld.ensure_flag(n);
      launch1d(k_iflag_pos, n, st, ls[l], bld.B.iflag.p, bld.flag.p);
      bld.refine(ls[l], n, bs[l + 1], l + 1, 1);
      te[l + 1] = bld.nboxes;
    }
    bld.reorg(ls, le, ts, te, nlevels, pts, npts);
    bld.computecoll(ls, le, bs, nlevels);
    launch1d(k_keep3, bld.nboxes, st, bld.B.iflag.p);
    std::fill(ts.begin(), ts.end(), 0);
    std::fill(te.begin(), te.end(), 0);
    for (int l = 2; l <= nlevels - 2; ++l) {
      const double distest = 1.05 * (bs[l] + bs[l + 1]) / 2.0;
      launch1d(k_updateflags, le[l] - ls[l], st, ls[l], distest, bld.B.ncoll.p, bld.B.coll.p,
               bld.B.ichild.p, bld.B.nchild.p, bld.B.centers.p, bld.B.iflag.p);
      launch1d(k_updateflags, te[l] - ts[l], st, ts[l], distest, bld.B.ncoll.p, bld.B.coll.p,
               bld.B.ichild.p, bld.B.nchild.p, bld.B.centers.p, bld.B.iflag.p);
      ts[l + 1] = bld.nboxes;
      {
        const int n = le[l] - ls[l];
        bld.ensure_flag(n);
        launch1d(k_iflag_pos, n, st, ls[l], bld.B.iflag.p, bld.flag.p);
        bld.refine(ls[l], n, bs[l + 1], l + 1, 1);
      }
      {
        const int n = te[l] - ts[l];
        bld.ensure_flag(n);
        launch1d(k_iflag_pos, n, st, ts[l], bld.B.iflag.p, bld.flag.p);
        bld.refine(ts[l], n, bs[l + 1], l + 1, 1);
      }
      te[l + 1] = bld.nboxes;
      launch1d(k_coll, te[l + 1] - ts[l + 1], st, ts[l + 1], bs[l + 1], bld.B.iparent.p,
               bld.B.ichild.p, bld.B.centers.p, bld.B.ncoll.p, bld.B.coll.p);
    }
    bld.reorg(ls, le, ts, te, nlevels, pts, npts);
    bld.computecoll(ls, le, bs, nlevels);
  }

  // ---- final point order (pts_tree_sort, pts_tree2d.f:1222-1246)
  const int nb = bld.nboxes;
  t.nlevels = nlevels;
  t.nboxes = nb;
  t.lstart = ls;
  t.lend = le;
  t.boxsize = bs;
  t.isrcse.alloc(2 * (size_t)nb, st);
  t.itargse.alloc(2 * (size_t)nb, st);
  t.isrc.alloc(std::max<int64_t>(ns, 1), st);
  t.itarg.alloc(std::max<int64_t>(nt, 1), st);
  DBuf<int>* se[2] = {&t.isrcse, &t.itargse};
  DBuf<int>* perm[2] = {&t.isrc, &t.itarg};
  const double* xy[2] = {src, trg};
  for (int k = 0; k < 2; ++k) {
    DBuf<int> cnt(nb, st);
    cnt.zero();
    se[k]->zero();
    launch1d(k_descend, npts[k], st, xy[k], pts[k]->p, bld.B.nchild.p, bld.B.ichild.p,
             bld.B.centers.p, cnt.p);
    for (int l = nlevels - 1; l >= 0; --l)
      launch1d(k_sum_children, le[l] - ls[l], st, ls[l], bld.B.nchild.p, bld.B.ichild.p, cnt.p);
    hipLaunchKernelGGL(k_root_se, dim3(1), dim3(1), 0, st, se[k]->p, (int)npts[k]);
    for (int l = 0; l < nlevels; ++l)
      launch1d(k_starts, le[l] - ls[l], st, ls[l], bld.B.nchild.p, bld.B.ichild.p, cnt.p, se[k]->p);
    if (npts[k] == 0) continue;
    DBuf<int> keys(npts[k], st), vals(npts[k], st), keys2(npts[k], st);
    launch1d(k_keys, npts[k], st, pts[k]->p, se[k]->p, keys.p, vals.p);
    int bits = 1;
    while ((int64_t(1) << bits) <= npts[k]) ++bits;
    size_t bytes = 0;
    hip_throw(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys.p, keys2.p, vals.p,
                                                 perm[k]->p, (int)npts[k], 0, bits, st), "sort");
    DBuf<char> tmp(bytes, st);
    hip_throw(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, keys.p, keys2.p, vals.p,
                                                 perm[k]->p, (int)npts[k], 0, bits, st), "sort");
  }
  t.centers = std::move(bld.B.centers);
  t.ilevel = std::move(bld.B.ilevel);
  t.iparent = std::move(bld.B.iparent);
  t.nchild = std::move(bld.B.nchild);
  t.ichild = std::move(bld.B.ichild);
  t.ncoll = std::move(bld.B.ncoll);
  t.coll = std::move(bld.B.coll);
}

}  // namespace fmm2d

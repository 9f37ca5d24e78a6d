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
// per box for every box-level step, device prefix sums (scan.h) where the
// reference uses cumsum, and the final point order as the points of each leaf (a contiguous
// range of the path order) in input order -- which is exactly the order the
// reference's recursive stable partition produces.  Box-level steps run level
// by level where the reference's loops carry a dependency between levels.

#include <algorithm>
#include <cmath>
#include <memory>

#include "radix_sort.h"
#include "scan.h"
#include "tree.h"

namespace fmm2d {
namespace {

// The reference writes `1.05*boxsize(ilev)` with a default-real literal in
// computecoll / computelists: 1.05 rounded to single precision, then promoted.
constexpr double kColl = (double)1.05f;
constexpr int kMaxLevels = 51;  // nlmax of the reference (pts_tree2d.f)
constexpr int kBatch = 8;             // refinement levels per host synchronisation
constexpr int64_t kBatchBoxes = 1 << 24;  // box capacity a batch may reserve

// copy the first n[k] 4-byte words of src[k] to dst[k] for k = blockIdx.y
struct WordCopies {
  const unsigned* src[10];
  unsigned* dst[10];
  int64_t n[10];
};
__global__ void k_copy_words(WordCopies c) {
  const int k = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c.n[k];
       i += (int64_t)gridDim.x * blockDim.x)
    c.dst[k][i] = c.src[k][i];
}

struct Boxes {
  DBuf<double> centers;
  DBuf<int> ilevel, iparent, nchild, ichild, ncoll, coll, ses, set, iflag;
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
    ses.alloc(2 * c, st);
    set.alloc(2 * c, st);
    iflag.alloc(c, st);
  }
  // capacity for `need` boxes, the first `keep` kept (one copy launch)
  void grow(size_t need, size_t keep) {
    if (need <= cap) return;
    const hipStream_t st = ilevel.st;
    Boxes b;
    b.alloc(std::max(need, cap * 2), st);
    if (keep > 0) {
      WordCopies c{};
      int k = 0;
      auto add = [&](const void* from, void* to, size_t words) {
        c.src[k] = static_cast<const unsigned*>(from);
        c.dst[k] = static_cast<unsigned*>(to);
        c.n[k++] = (int64_t)words;
      };
      add(centers.p, b.centers.p, 4 * keep);  // 2 doubles per box
      add(ilevel.p, b.ilevel.p, keep);
      add(iparent.p, b.iparent.p, keep);
      add(nchild.p, b.nchild.p, keep);
      add(ichild.p, b.ichild.p, 4 * keep);
      add(ncoll.p, b.ncoll.p, keep);
      add(coll.p, b.coll.p, 9 * keep);
      add(ses.p, b.ses.p, 2 * keep);
      add(set.p, b.set.p, 2 * keep);
      add(iflag.p, b.iflag.p, keep);
      const unsigned gx = (unsigned)std::min<int64_t>(ceil_div((int64_t)(9 * keep), 256), 1024);
      hipLaunchKernelGGL(k_copy_words, dim3(gx, 10), dim3(256), 0, st, c);
      hip_throw(hipGetLastError(), "box copy");
    }
    *this = std::move(b);
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

// ------------------------------------------------------------ point paths
// Each point's path through the tree is the sequence of its quadrants,
// computed with the reference's own centre arithmetic: child centre = parent
// centre +- bs_child/2 (one rounding), point below/left of the centre when
// y - cy < 0 / x - cx < 0.  The path does not depend on which boxes get
// refined, so it is computed once per point; digits are child indices 0..3 in
// the reference's order (1=(-,-), 2=(+,-), 3=(-,+), 4=(+,+)), most significant
// digit = level 1.  Levels 1..20 go to key1 (a 40-bit radix sort), levels
// 21..51 to key2 (only computed when the tree is that deep).  x - cx < 0 is
// tested as x < cx: with gradual underflow the difference of two distinct
// finite doubles is never rounded to zero and keeps its sign, and both forms
// are false for NaN.
constexpr int kD1 = 20;
constexpr int kD2Max = 31;

__global__ void k_path_keys(int64_t n, const double* __restrict__ xy, double cx0, double cy0,
                            const double* __restrict__ half, int d1, int d2,
                            unsigned long long* key1, unsigned long long* key2, int* iota) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = xy[2 * i], y = xy[2 * i + 1];
  double cx = cx0, cy = cy0;
  unsigned long long k = 0;
  for (int l = 1; l <= d1; ++l) {
    const int qx = (x < cx) ? 0 : 1, qy = (y < cy) ? 0 : 1;
    k = (k << 2) | (unsigned long long)(2 * qy + qx);
    const double h = half[l];
    cx += qx ? h : -h;
    cy += qy ? h : -h;
  }
  key1[i] = k;
  if (d2 > 0) {
    k = 0;
    for (int l = d1 + 1; l <= d1 + d2; ++l) {
      const int qx = (x < cx) ? 0 : 1, qy = (y < cy) ? 0 : 1;
      k = (k << 2) | (unsigned long long)(2 * qy + qx);
      const double h = half[l];
      cx += qx ? h : -h;
      cy += qy ? h : -h;
    }
    key2[i] = k;
  }
  if (iota) iota[i] = (int)i;
}

template <typename T>
__global__ void k_gather_by(int64_t n, const int* idx, const T* in, T* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = in[idx[i]];
}

struct Keys {
  const unsigned long long* k1;
  const unsigned long long* k2;  // null unless the tree is deeper than kD1
  int d1, d2;
};

__device__ __forceinline__ int digit_at(const Keys& K, int64_t i, int lev) {
  if (lev <= K.d1) return (int)((K.k1[i] >> (2 * (K.d1 - lev))) & 3ull);
  return (int)((K.k2[i] >> (2 * (K.d2 - (lev - K.d1)))) & 3ull);
}

// b[q] = first i in [lo0, hi0) whose digit at level lev is >= q + 1, q = 0..2
// (digits are sorted there because all keys in a box's range share the path
// above lev).  The three binary searches run interleaved over the whole
// range with their loads in flight together: at the first levels a box holds
// millions of points, and chained searches were ~60 dependent loads long.
__device__ void digit_bounds3(const Keys& K, int64_t lo0, int64_t hi0, int lev, int64_t* b) {
  int64_t lo[3] = {lo0, lo0, lo0}, hi[3] = {hi0, hi0, hi0};
  while (lo[0] < hi[0] || lo[1] < hi[1] || lo[2] < hi[2]) {
    int64_t mid[3];
    int dg[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) mid[q] = min(lo[q] + ((hi[q] - lo[q]) >> 1), hi0 - 1);
#pragma unroll
    for (int q = 0; q < 3; ++q) dg[q] = digit_at(K, mid[q], lev);  // lo0 <= mid < hi0
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (lo[q] < hi[q]) {
        if (dg[q] < q + 1) lo[q] = mid[q] + 1;
        else hi[q] = mid[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) b[q] = lo[q];
}

// ------------------------------------------------------------ refinement
__global__ void k_refine_flags(int64_t n, int lo, const int* ses, const int* set, int idivflag,
                               int ndiv, int* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = lo + (int)i;
  const int cs = ses[2 * b + 1] - ses[2 * b], ct = set[2 * b + 1] - set[2 * b];
  int nn = cs;
  if (idivflag == 1) nn = ct;
  if (idivflag == 2) nn = max(cs, ct);
  flag[i] = nn > ndiv ? 1 : 0;
}

// flag[i] = iflag[lo+i] > 0 (tree_refine_boxes_flag, tree_routs2d.f:334-388)
__global__ void k_total(int n, const int* flag, const int* offs, int* total) {
  *total = (n > 0) ? offs[n - 1] + flag[n - 1] : 0;
}

// Refinement of several levels between host synchronisations: dlev[2 l],
// dlev[2 l + 1] hold the first box and the box count of level l, written by
// k_level_next on the device; the launches cover an upper bound `ub` of the
// level's size and read the count from dlev.
__global__ void k_refine_flags_lv(int64_t ub, const int* dlev, int l, const int* ses,
                                  const int* set, int idivflag, int ndiv, int* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ub) return;
  const int lo = dlev[2 * l], n = dlev[2 * l + 1];
  int f = 0;
  if (i < n) {
    const int b = lo + (int)i;
    const int cs = ses[2 * b + 1] - ses[2 * b], ct = set[2 * b + 1] - set[2 * b];
    int nn = cs;
    if (idivflag == 1) nn = ct;
    if (idivflag == 2) nn = max(cs, ct);
    f = nn > ndiv ? 1 : 0;
  }
  flag[i] = f;
}

// level l+1 = the children of level l's flagged boxes, numbered from *dnb
__global__ void k_level_next(int64_t ub, const int* flag, const int* offs, int* dlev, int l,
                             int* dnb) {
  const int nref = offs[ub - 1] + flag[ub - 1];
  dlev[2 * (l + 1)] = *dnb;
  dlev[2 * (l + 1) + 1] = 4 * nref;
  *dnb += 4 * nref;
}

// box 1 (index 0): the root, holding every point
__global__ void k_init_root(double cx, double cy, int ns, int nt, double* centers, int* ilevel,
                            int* iparent, int* nchild, int* ichild, int* ses, int* set,
                            int* iflag) {
  centers[0] = cx;
  centers[1] = cy;
  ilevel[0] = 0;
  iparent[0] = -1;
  nchild[0] = 0;
  for (int q = 0; q < 4; ++q) ichild[q] = -1;
  ses[0] = 0;
  ses[1] = ns;
  set[0] = 0;
  set[1] = nt;
  iflag[0] = 0;
}

// cum[k] = exclusive-scan value at segment start cut[k]; cut[K] = n, the
// scanned length, gives the grand total
// Box ranges refined together with one scan (Builder::refine_segments),
// passed by value: segment j covers flag positions [pos[j], pos[j+1]) and
// boxes lo[j] ..; its children are at level lev[j] with half box size half[j].
constexpr int kRefSeg = 64;
struct RefSegs {
  int k;
  int pos[kRefSeg + 1];
  int lo[kRefSeg];
  int lev[kRefSeg];
  double half[kRefSeg];
};

// the segment of flag position i: the last j with pos[j] <= i (skips empty ones)
__device__ __forceinline__ int seg_of(const RefSegs& R, int64_t i) {
  int a = 0, b = R.k - 1;
  while (a < b) {
    const int m = (a + b + 1) >> 1;
    if (R.pos[m] <= i) a = m;
    else b = m - 1;
  }
  return a;
}

__global__ void k_iflag_segs(int64_t n, RefSegs R, const int* iflag, int* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = seg_of(R, i);
  flag[i] = iflag[R.lo[j] + (int)(i - R.pos[j])] > 0 ? 1 : 0;
}

// cum[k] = flagged boxes before segment k (k = 0..K)
__global__ void k_seg_totals(RefSegs R, const int* flag, const int* offs, int* cum) {
  const int n = R.pos[R.k];
  for (int k = threadIdx.x; k <= R.k; k += blockDim.x) {
    const int c = R.pos[k];
    cum[k] = (c < n) ? offs[c] : (n > 0 ? offs[n - 1] + flag[n - 1] : 0);
  }
}

// Children of the flagged boxes lo..lo+n-1 are numbered base + 4*offs + q, q =
// 0..3 for (-,-), (+,-), (-,+), (+,+) around the parent centre
// (tree_refine_boxes, tree_routs2d.f:25-52).  Their point ranges split the
// parent's by the path digit at level lev.  mode 1 (level-restriction
// refinement): child iflag = 3 when the parent's flag is 1, else 0.
// the four children of box b, numbered c0 .. c0+3, at level lev
__device__ void make_children_of(int b, int c0, double half, int lev, double* centers,
                                 int* ilevel, int* iparent, int* nchild, int* ichild, int* ses,
                                 int* set, int* iflag, int mode, const Keys& ks,
                                 const Keys& kt) {
  const double cx = centers[2 * b], cy = centers[2 * b + 1];
  nchild[b] = 4;
  const int pf = (mode == 1) ? iflag[b] : 0;
  int bs_[5], bt_[5];
  bs_[0] = ses[2 * b];
  bs_[4] = ses[2 * b + 1];
  bt_[0] = set[2 * b];
  bt_[4] = set[2 * b + 1];
  int64_t bq[3];
  digit_bounds3(ks, bs_[0], bs_[4], lev, bq);
  for (int q = 1; q < 4; ++q) bs_[q] = (int)bq[q - 1];
  digit_bounds3(kt, bt_[0], bt_[4], lev, bq);
  for (int q = 1; q < 4; ++q) bt_[q] = (int)bq[q - 1];
  for (int q = 0; q < 4; ++q) {
    const int c = c0 + q;
    centers[2 * c] = cx + ((q & 1) ? half : -half);
    centers[2 * c + 1] = cy + ((q & 2) ? half : -half);
    iparent[c] = b;
    nchild[c] = 0;
    ichild[4 * c] = ichild[4 * c + 1] = ichild[4 * c + 2] = ichild[4 * c + 3] = -1;
    ichild[4 * b + q] = c;
    ilevel[c] = lev;
    ses[2 * c] = bs_[q];
    ses[2 * c + 1] = bs_[q + 1];
    set[2 * c] = bt_[q];
    set[2 * c + 1] = bt_[q + 1];
    iflag[c] = (mode == 1 && pf == 1) ? 3 : 0;
  }
}

__global__ void k_make_children(int64_t n, int lo, const int* flag, const int* offs, int base,
                                double half, int lev, double* centers, int* ilevel, int* iparent,
                                int* nchild, int* ichild, int* ses, int* set, int* iflag, int mode,
                                Keys ks, Keys kt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  make_children_of(lo + (int)i, base + 4 * offs[i], half, lev, centers, ilevel, iparent, nchild,
                   ichild, ses, set, iflag, mode, ks, kt);
}

__global__ void k_make_children_segs(int64_t n, RefSegs R, const int* flag, const int* offs,
                                     int base, double* centers, int* ilevel, int* iparent,
                                     int* nchild, int* ichild, int* ses, int* set, int* iflag,
                                     Keys ks, Keys kt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const int j = seg_of(R, i);
  make_children_of(R.lo[j] + (int)(i - R.pos[j]), base + 4 * offs[i], R.half[j], R.lev[j],
                   centers, ilevel, iparent, nchild, ichild, ses, set, iflag, 1, ks, kt);
}

__global__ void k_make_children_lv(int64_t ub, const int* dlev, int l, const int* flag,
                                   const int* offs, double half, double* centers, int* ilevel,
                                   int* iparent, int* nchild, int* ichild, int* ses, int* set,
                                   int* iflag, Keys ks, Keys kt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ub || i >= dlev[2 * l + 1] || !flag[i]) return;
  make_children_of(dlev[2 * l] + (int)i, dlev[2 * (l + 1)] + 4 * offs[i], half, l + 1, centers,
                   ilevel, iparent, nchild, ichild, ses, set, iflag, 0, ks, kt);
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

// computecoll for every box at once (trees of at most kWalkLevels levels).
// The colleagues of a box are the boxes of its level whose integer
// coordinates (ix, iy) differ by at most one in each direction -- exactly the
// boxes the reference finds among the children of the parent's colleagues
// within 1.05 box sizes.  The reference lists them in the order of the
// parent's colleagues and then of the child slots (-,-), (+,-), (-,+), (+,+),
// which, by induction from the root, is increasing path order: the
// interleaved code of (iy, ix), iy bit above ix bit, coarsest level first.
constexpr int kWalkLevels = 29;

// (ix, iy) of every box, walking up to the root: the child centre is the
// parent centre +- half the child size, so the comparisons are exact.
__global__ void k_box_ij(int64_t nb, const int* iparent, const int* ilevel,
                         const double* centers, unsigned* ij) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  int cur = (int)i;
  const int lev = ilevel[cur];
  double cx = centers[2 * cur], cy = centers[2 * cur + 1];
  unsigned ix = 0, iy = 0;
  for (int t = 0; t < lev; ++t) {
    const int par = iparent[cur];
    const double px = centers[2 * par], py = centers[2 * par + 1];
    ix |= (cx > px ? 1u : 0u) << t;
    iy |= (cy > py ? 1u : 0u) << t;
    cur = par;
    cx = px;
    cy = py;
  }
  ij[2 * i] = ix;
  ij[2 * i + 1] = iy;
}

// spread the 29 bits of v to the even bit positions
__device__ __forceinline__ unsigned long long spread_bits(unsigned v) {
  unsigned long long x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

// interleaved path code of (ix, iy) at one level, iy bit above ix bit
__device__ __forceinline__ unsigned long long path_code(unsigned ix, unsigned iy) {
  return spread_bits(ix) | (spread_bits(iy) << 1);
}

// Colleagues of every box in one pass.  The neighbour (jx, jy) of box b
// (level lev, coordinates (ix, iy)) shares b's ancestors down to level
// m = lev - max(bitlength(ix ^ jx), bitlength(iy ^ jy)); it is found by going
// up lev - m parents and down lev - m child slots, slot 2 qy + qx taken from
// the bits of (jx, jy) (the reference's child order (-,-), (+,-), (-,+),
// (+,+)); a missing child ends the search.
__global__ void k_coll_walk(int64_t nb, const int* ilevel, const int* iparent, const int* ichild,
                            const unsigned* ij, int* ncoll, int* coll) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int b = (int)i;
  const int lev = ilevel[b];
  const long long side = 1ll << lev;
  const unsigned ix = ij[2 * i], iy = ij[2 * i + 1];
  int id[kMaxColl];
  unsigned long long code[kMaxColl];
  int cnt = 0;
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) {
      const long long jxl = (long long)ix + dx, jyl = (long long)iy + dy;
      if (jxl < 0 || jyl < 0 || jxl >= side || jyl >= side) continue;
      const unsigned jx = (unsigned)jxl, jy = (unsigned)jyl;
      int j = b;
      if (dx != 0 || dy != 0) {
        const unsigned diff = (ix ^ jx) | (iy ^ jy);
        const int up = 32 - __clz(diff);  // levels to the common ancestor
        int a = b;
        for (int t = 0; t < up; ++t) a = iparent[a];
        for (int t = up - 1; t >= 0 && a >= 0; --t) {
          const int q = (int)((((jy >> t) & 1u) << 1) | ((jx >> t) & 1u));
          a = ichild[4 * a + q];
        }
        j = a;
      }
      if (j < 0) continue;
      // insertion by path code
      const unsigned long long cj = path_code(jx, jy);
      int k = cnt++;
      while (k > 0 && code[k - 1] > cj) {
        code[k] = code[k - 1];
        id[k] = id[k - 1];
        --k;
      }
      code[k] = cj;
      id[k] = j;
    }
  ncoll[i] = cnt;
  for (int k = 0; k < kMaxColl; ++k) coll[9 * i + k] = k < cnt ? id[k] : -1;
}

__global__ void k_root_coll(int* ncoll, int* coll) {
  ncoll[0] = 1;
  coll[0] = 0;
  for (int k = 1; k < 9; ++k) coll[k] = -1;
}

// ------------------------------------------------------------ level restriction
// flag pass (pts_tree2d.f:797-839): leaves among the grandparent's colleagues
// that touch the parent of a box at level lev.
// It only turns flags 0 into 1, so the levels (the reference sweeps them from
// the finest up) run in one launch; distest = 1.05 (bs_{l-1} + bs_{l-2}) / 2 of
// the box's level l comes from `dist`.
struct LevelReals {
  double v[kMaxLevels + 1];
};

__global__ void k_flag1(int64_t n, int lo, LevelReals dist, const int* ilevel, const int* iparent,
                        const int* ncoll, const int* coll, const int* nchild,
                        const double* centers, int* iflag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = lo + (int)i;
  const double distest = dist.v[ilevel[b]];
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

// updateflags (tree_routs2d.f:217-326) over boxes lo1 .. lo1+n1-1, then
// lo2 .. lo2+(n-n1)-1
__global__ void k_updateflags(int64_t n, int n1, int lo1, int lo2, double distest,
                              const int* ncoll, const int* coll, const int* ichild,
                              const int* nchild, const double* centers, int* iflag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = i < n1 ? lo1 + (int)i : lo2 + (int)(i - n1);
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
// The new number of every box: the boxes are covered by segments (a level's
// original boxes, a level's tail boxes), sorted by first box; segment k moves
// [from[k], from[k+1]) to to[k], to[k] + 1, ...
constexpr int kMaxSeg = 2 * (kMaxLevels + 1);
struct SegTable {
  int n;
  int from[kMaxSeg], to[kMaxSeg];
};

__global__ void k_map_segments(int64_t nb, SegTable T, int* map) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  int lo = 0, hi = T.n - 1;  // last segment with from <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (T.from[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  map[i] = T.to[lo] + (int)(i - T.from[lo]);
}

__global__ void k_reorg_apply(int64_t n, const int* map, const double* c0, const int* lev0,
                              const int* par0, const int* nch0, const int* ich0, const int* fl0,
                              const int* s0, const int* t0, double* c1, int* lev1, int* par1,
                              int* nch1, int* ich1, int* fl1, int* s1, int* t1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = map[i];
  c1[2 * m] = c0[2 * i];
  c1[2 * m + 1] = c0[2 * i + 1];
  lev1[m] = lev0[i];
  nch1[m] = nch0[i];
  fl1[m] = fl0[i];
  s1[2 * m] = s0[2 * i];
  s1[2 * m + 1] = s0[2 * i + 1];
  t1[2 * m] = t0[2 * i];
  t1[2 * m + 1] = t0[2 * i + 1];
  par1[m] = par0[i] < 0 ? -1 : map[par0[i]];
  for (int q = 0; q < 4; ++q) ich1[4 * m + q] = ich0[4 * i + q] < 0 ? -1 : map[ich0[4 * i + q]];
}

// ------------------------------------------------------------ final point order
// Positions in path order are the reference's point positions up to the order
// inside each leaf, which the reference keeps in input order: label every
// point (in input order) with its leaf's first position, then stable-sort.
__global__ void k_leaf_label(int64_t n, const int* nchild, const int* se, const int* perm,
                             unsigned* label) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n || nchild[b] != 0) return;
  const int s = se[2 * b];
  for (int i = s; i < se[2 * b + 1]; ++i) label[perm[i]] = (unsigned)s;
}

// The same order without the sort: inside each leaf (a contiguous range of
// the path order) the input indices in increasing order, placed by rank.
// Leaves of up to kLeafSmall points: one thread each, ranks over a register
// copy; larger leaves are appended to `bigl` (in any order: each leaf is
// independent) for k_leaf_sort.
constexpr int kLeafSmall = 32;
constexpr int kLeafSortMax = 2048;
constexpr int kLeafSortWaves = 4;

__global__ void k_leaf_sort_small(int64_t nb, const int* nchild, const int* se, const int* perm,
                                  int* out, int* bigl, int* nbig) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb || nchild[i] != 0) return;
  const int s = se[2 * i], m = se[2 * i + 1] - s;
  if (m <= 0) return;
  if (m > kLeafSmall) {
    bigl[atomicAdd(nbig, 1)] = (int)i;
    return;
  }
  int v[kLeafSmall];
#pragma unroll
  for (int k = 0; k < kLeafSmall; ++k) v[k] = k < m ? perm[s + k] : 0x7fffffff;
#pragma unroll
  for (int k = 0; k < kLeafSmall; ++k) {
    int rank = 0;
#pragma unroll
    for (int q = 0; q < kLeafSmall; ++q) rank += (v[q] < v[k]) ? 1 : 0;
    if (k < m) out[s + rank] = v[k];
  }
}

// The leaves of `bigl`: one wavefront each in a grid-stride loop (4 per
// block, each with its own LDS); up to 64 points rank through lane shuffles,
// up to kLeafSortMax through LDS; a larger leaf sets *big and the caller falls
// back to the labelled sort.
__global__ __launch_bounds__(kLeafSortWaves * 64) void k_leaf_sort(const int* bigl,
                                                                   const int* nbig,
                                                                   const int* se, const int* perm,
                                                                   int* out, int* big) {
  __shared__ int shall[kLeafSortWaves][kLeafSortMax];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int* sh = shall[wv];
  const int nl = *nbig;
  for (int li = blockIdx.x * kLeafSortWaves + wv; li < nl; li += gridDim.x * kLeafSortWaves) {
    const int b = bigl[li];
    const int s = se[2 * b], m = se[2 * b + 1] - s;
    if (m > kLeafSortMax) {
      if (lane == 0) *big = 1;
      continue;
    }
    if (m <= 64) {
      const int v = lane < m ? perm[s + lane] : 0x7fffffff;
      int rank = 0;
      // j is uniform: read lane j through the scalar unit (v_readlane)
      for (int j = 0; j < m; ++j) rank += (__builtin_amdgcn_readlane(v, j) < v) ? 1 : 0;
      if (lane < m) out[s + rank] = v;
      continue;
    }
    for (int i = lane; i < m; i += 64) sh[i] = perm[s + i];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int i = lane; i < m; i += 64) {
      const int v = sh[i];
      int rank = 0;
      for (int j = 0; j < m; ++j) rank += (sh[j] < v) ? 1 : 0;
      out[s + rank] = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

__global__ void k_iota(int64_t n, int* v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (int)i;
}

// ------------------------------------------------------------ host helpers
// Points of one set (sources or targets) sorted by path key.
struct PathKeys {
  int64_t n = 0;
  int d1 = 0, d2 = 0;
  DBuf<unsigned long long> k1, k2;  // sorted
  DBuf<int> perm;                   // sorted position -> input index
  Keys view() const { return Keys{k1.p, k2.p, d1, d2}; }

  void build(const double* xy, int64_t np, double cx, double cy, const double* half, int dd1,
             int dd2, hipStream_t st) {
    n = np;
    d1 = dd1;
    d2 = dd2;
    perm.alloc(std::max<int64_t>(n, 1), st);
    k1.alloc(std::max<int64_t>(n, 1), st);
    k2.alloc(d2 > 0 ? std::max<int64_t>(n, 1) : 0, st);
    if (n == 0) return;
    DBuf<unsigned long long> u1(n, st), u2(d2 > 0 ? n : 0, st);
    DBuf<int> iota(n, st);
    launch1d(k_path_keys, n, st, xy, cx, cy, half, d1, d2, u1.p, u2.p, iota.p);
    if (d2 == 0) {
      radix_sort(u1.p, k1.p, iota.p, perm.p, n, 2 * d1, st);
    } else {
      // lexicographic (key1, key2): stable LSD passes, key2 first
      DBuf<unsigned long long> t2(n, st), t1(n, st);
      DBuf<int> p1(n, st);
      radix_sort(u2.p, t2.p, iota.p, p1.p, n, 2 * d2, st);
      launch1d(k_gather_by<unsigned long long>, n, st, p1.p, u1.p, t1.p);
      radix_sort(t1.p, k1.p, p1.p, perm.p, n, 2 * d1, st);
      launch1d(k_gather_by<unsigned long long>, n, st, perm.p, u2.p, k2.p);
    }
  }
};

struct Builder {
  hipStream_t st;
  Boxes B;
  int nboxes = 0;
  DBuf<int> flag, offs, total;
  Keys ks{}, kt{};
  // the centre comparisons of k_box_ij are exact when the domain is not far
  // from the origin relative to its size
  bool walk_ok = false;

  explicit Builder(hipStream_t s) : st(s) { total.alloc(64, s); }

  void scan_only(int n) { exclusive_scan(flag.p, offs.p, n, (int*)nullptr, st); }
  // exclusive scan of flag[0..n) into offs, returns the total
  int scan(int n) {
    if (n <= 0) return 0;
    scan_only(n);
    hipLaunchKernelGGL(k_total, dim3(1), dim3(1), 0, st, n, flag.p, offs.p, total.p);
    int h = 0;
    readback(&h, total.p, sizeof(int), st);
    return h;
  }
  // Refine several box ranges with one scan: segment k covers boxes
  // lo[k] .. lo[k]+n[k]-1 (flags from iflag > 0), children at level lev[k]
  // with box size bsc[k].  Children are numbered in segment order, exactly as
  // refining the segments one after the other.  first[k], last[k]: range of
  // the children of segment k.
  void refine_segments(const std::vector<int>& lo, const std::vector<int>& n,
                       const std::vector<int>& lev, const std::vector<double>& bsc,
                       std::vector<int>& first, std::vector<int>& last) {
    const size_t K = lo.size();
    if (K > (size_t)kRefSeg) throw HipError{hipErrorInvalidValue, "refinement segments"};
    RefSegs R{};
    R.k = (int)K;
    R.pos[0] = 0;
    for (size_t k = 0; k < K; ++k) {
      R.pos[k + 1] = R.pos[k] + n[k];
      R.lo[k] = lo[k];
      R.lev[k] = lev[k];
      R.half[k] = 0.5 * bsc[k];
    }
    first.assign(K, nboxes);
    last.assign(K, nboxes);
    const int ntot = R.pos[K];
    if (ntot == 0) return;
    ensure_flag(ntot);
    launch1d(k_iflag_segs, ntot, st, R, B.iflag.p, flag.p);
    scan_only(ntot);
    if (total.n < K + 1) total.alloc(K + 1, st);
    hipLaunchKernelGGL(k_seg_totals, dim3(1), dim3(64), 0, st, R, flag.p, offs.p, total.p);
    std::vector<int> cum(K + 1, 0);
    readback(cum.data(), total.p, sizeof(int) * (K + 1), st);
    const int nref = cum[K];
    if (nref == 0) return;
    B.grow(nboxes + 4 * (size_t)nref, nboxes);
    for (size_t k = 0; k < K; ++k) {
      first[k] = nboxes + 4 * cum[k];
      last[k] = nboxes + 4 * cum[k + 1];
    }
    launch1d(k_make_children_segs, ntot, st, R, flag.p, offs.p, nboxes, B.centers.p,
             B.ilevel.p, B.iparent.p, B.nchild.p, B.ichild.p, B.ses.p, B.set.p, B.iflag.p, ks,
             kt);
    nboxes += 4 * nref;
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
             B.centers.p, B.ilevel.p, B.iparent.p, B.nchild.p, B.ichild.p, B.ses.p, B.set.p,
             B.iflag.p, mode, ks, kt);
    nboxes += 4 * nref;
    return nref;
  }
  void computecoll(const std::vector<int>& ls, const std::vector<int>& le,
                   const std::vector<double>& bs, int nlevels) {
    // all levels at once by walking the tree from every box (integer
    // coordinates exact: at most kWalkLevels levels, domain near the origin)
    if (nlevels <= kWalkLevels && walk_ok) {
      DBuf<unsigned> ij(2 * (int64_t)nboxes, st);
      launch1d(k_box_ij, nboxes, st, B.iparent.p, B.ilevel.p, B.centers.p, ij.p);
      launch1d(k_coll_walk, nboxes, st, B.ilevel.p, B.iparent.p, B.ichild.p, ij.p, B.ncoll.p,
               B.coll.p);
      return;
    }
    hipLaunchKernelGGL(k_root_coll, dim3(1), dim3(1), 0, st, B.ncoll.p, B.coll.p);
    for (int l = 1; l <= nlevels; ++l)
      launch1d(k_coll, le[l] - ls[l], st, ls[l], bs[l], B.iparent.p, B.ichild.p, B.centers.p,
               B.ncoll.p, B.coll.p);
  }
  // pts_tree_reorg: per level >= 2, original boxes then the tail boxes.
  void reorg(std::vector<int>& ls, std::vector<int>& le, const std::vector<int>& ts,
             const std::vector<int>& te, int nlevels) {
    DBuf<int> map(nboxes, st);
    std::vector<std::pair<int, int>> seg;  // (from, to) of non-empty segments
    for (int l = 0; l <= std::min(1, nlevels); ++l)
      if (le[l] > ls[l]) seg.push_back({ls[l], ls[l]});
    int cur = (nlevels >= 2) ? ls[2] : 0;
    for (int l = 2; l <= nlevels; ++l) {
      const int s = cur;
      if (le[l] > ls[l]) seg.push_back({ls[l], cur});
      cur += le[l] - ls[l];
      if (te[l] > ts[l]) seg.push_back({ts[l], cur});
      cur += te[l] - ts[l];
      ls[l] = s;
      le[l] = cur;
    }
    std::sort(seg.begin(), seg.end());
    SegTable T{};
    T.n = (int)seg.size();
    if (T.n > kMaxSeg) throw HipError{hipErrorInvalidValue, "tree reorganisation segments"};
    for (int k = 0; k < T.n; ++k) {
      T.from[k] = seg[k].first;
      T.to[k] = seg[k].second;
    }
    launch1d(k_map_segments, nboxes, st, T, map.p);
    Boxes N;
    N.alloc(B.cap, st);
    launch1d(k_reorg_apply, nboxes, st, map.p, B.centers.p, B.ilevel.p, B.iparent.p, B.nchild.p,
             B.ichild.p, B.iflag.p, B.ses.p, B.set.p, N.centers.p, N.ilevel.p, N.iparent.p,
             N.nchild.p, N.ichild.p, N.iflag.p, N.ses.p, N.set.p);
    B.centers.swap(N.centers);
    B.ilevel.swap(N.ilevel);
    B.iparent.swap(N.iparent);
    B.nchild.swap(N.nchild);
    B.ichild.swap(N.ichild);
    B.iflag.swap(N.iflag);
    B.ses.swap(N.ses);
    B.set.swap(N.set);
  }
};

}  // namespace

// The steps of the tree build that do not depend on the leaf capacity: the
// bounding square, the box sizes, the points' path keys sorted (sources and
// targets separately).  Kept by a caller that needs the sorted keys before it
// chooses the leaf capacity (the parameter proposer) and handed to the build.
struct TreePrep {
  const double* src = nullptr;
  const double* trg = nullptr;
  int64_t ns = 0, nt = 0;
  int nlmax = 0;
  std::vector<double> bsall;  // box size of every possible level
  double root[2] = {0.0, 0.0};
  bool walk_ok = false;
  DBuf<double> dhalf;
  PathKeys pk[2];
};

namespace {
void prepare_into(TreePrep& P, const double* src, int64_t ns, const double* trg, int64_t nt,
                  int nlmax, hipStream_t st) {
  if (nlmax > kMaxLevels) throw HipError{hipErrorInvalidValue, "nlmax beyond 51"};
  P.src = src;
  P.trg = trg;
  P.ns = ns;
  P.nt = nt;
  P.nlmax = nlmax;
  // ---- bounding square (pts_tree2d.f:455-517)
  double bb[4];
  {
    const int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, ceil_div(ns + nt, 256)));
    DBuf<double> part(4 * nb, st), out(4, st);
    hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(256), 0, st, src, ns, trg, nt, part.p);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, st, part.p, nb, out.p);
    readback(bb, out.p, sizeof(bb), st);
  }
  // box sizes of every possible level (boxsize(l+1) = boxsize(l)/2)
  std::vector<double>& bsall = P.bsall;
  bsall.assign(nlmax + 2, 0.0);
  bsall[0] = bb[1] - bb[0];
  const double sizey = bb[3] - bb[2];
  if (sizey > bsall[0]) bsall[0] = sizey;
  for (int l = 1; l < (int)bsall.size(); ++l) bsall[l] = bsall[l - 1] / 2;
  P.root[0] = (bb[0] + bb[1]) / 2;
  P.root[1] = (bb[2] + bb[3]) / 2;
  P.walk_ok = std::max(std::fabs(P.root[0]), std::fabs(P.root[1])) <= std::ldexp(bsall[0], 20);

  // ---- point paths, sorted (sources and targets separately)
  std::vector<double> half(bsall.size());
  for (size_t l = 0; l < half.size(); ++l) half[l] = 0.5 * bsall[l];
  P.dhalf.alloc(half.size(), st);
  upload_bytes(P.dhalf.p, half.data(), sizeof(double) * half.size(), st);
  const int d1 = std::min(kD1, nlmax);
  P.pk[0].build(src, ns, P.root[0], P.root[1], P.dhalf.p, d1, 0, st);
  P.pk[1].build(trg, nt, P.root[0], P.root[1], P.dhalf.p, d1, 0, st);
}
}  // namespace

TreePrep* tree_prepare(const double* src, int64_t ns, const double* trg, int64_t nt, int nlmax,
                       hipStream_t st) {
  std::unique_ptr<TreePrep> P(new TreePrep());
  prepare_into(*P, src, ns, trg, nt, nlmax, st);
  return P.release();
}

void tree_prepare_free(TreePrep* p) { delete p; }

bool tree_prepare_matches(const TreePrep* p, const double* src, int64_t ns, const double* trg,
                          int64_t nt, int nlmax) {
  return p && p->src == src && p->ns == ns && p->nt == nt && (nt == 0 || p->trg == trg) &&
         p->nlmax == nlmax;
}

void build_tree(const double* src, int64_t ns, const double* trg, int64_t nt, int ndiv,
                int idivflag, int nlmax, Tree& t, hipStream_t st, const TreePrep* prep) {
  if (nlmax > kMaxLevels) throw HipError{hipErrorInvalidValue, "nlmax beyond 51"};
  Builder bld(st);
  t.ns = ns;
  t.nt = nt;
  // bounding square, box sizes and sorted path keys: the caller's, or made here
  TreePrep own;
  const TreePrep* P = prep;
  if (!tree_prepare_matches(prep, src, ns, trg, nt, nlmax)) {
    prepare_into(own, src, ns, trg, nt, nlmax, st);
    P = &own;
  }
  const std::vector<double>& bsall = P->bsall;
  const double* root = P->root;
  bld.walk_ok = P->walk_ok;
  const double* dhalf_p = P->dhalf.p;
  const int d1 = std::min(kD1, nlmax);
  // the keys in use: the prepared ones, or deeper ones made below (points
  // closer than 2^-21 of the domain)
  PathKeys pkx[2];
  const PathKeys* pk[2] = {&P->pk[0], &P->pk[1]};
  bld.ks = pk[0]->view();
  bld.kt = pk[1]->view();

  size_t cap0 = std::max<size_t>(1024, (size_t)(8 * (ns + nt) / std::max(ndiv, 1)));
  bld.B.alloc(cap0, st);
  bld.nboxes = 1;
  hipLaunchKernelGGL(k_init_root, dim3(1), dim3(1), 0, st, root[0], root[1], (int)ns, (int)nt,
                     bld.B.centers.p, bld.B.ilevel.p, bld.B.iparent.p, bld.B.nchild.p,
                     bld.B.ichild.p, bld.B.ses.p, bld.B.set.p, bld.B.iflag.p);

  // ---- level-by-level refinement (pts_tree2d.f:551-631).  While the
  // bound on a level's size is moderate, kBatch levels run between host
  // synchronisations (level ranges kept on the device); a refined box holds
  // more than ndiv points, so a level has at most 4 (ns + nt) / (ndiv + 1)
  // boxes.  Otherwise, and for the key-extension check at level d1, one level
  // at a time with its exact count.
  std::vector<int> ls{0}, le{1};
  std::vector<double> bs{bsall[0]};
  int nlevels = nlmax;
  const int64_t lev_cap = 4 * ((ns + nt) / (std::max(ndiv, 0) + 1) + 1);
  DBuf<int> dlev(2 * (nlmax + 2), st), dnb(1, st);
  int l = 0;
  bool done = false;
  while (!done && l < nlmax) {
    const int lo = ls[l], n = le[l] - ls[l];
    // the batch: levels l .. l+K-1 refined, sizes bounded by ub[k]
    int K = std::min(kBatch, nlmax - l);
    if (pk[0]->d2 == 0 && l < d1) K = std::min(K, d1 - l);
    // as many levels as fit the capacity a batch may reserve
    const int64_t limit = std::max<int64_t>(kBatchBoxes, 2 * (int64_t)bld.nboxes);
    std::vector<int64_t> ub{n};
    int64_t reserve = 0;
    for (int k = 1; k <= K; ++k) {
      const int64_t u = std::min<int64_t>(4 * ub[k - 1], lev_cap);
      if (reserve + u > limit) {
        K = k - 1;
        break;
      }
      ub.push_back(u);
      reserve += u;
    }
    const bool batch = (pk[0]->d2 != 0 || l < d1) && K > 1;
    if (!batch) {
      // one level, exact (pts_tree2d.f:551-631 as written)
      bld.ensure_flag(n);
      launch1d(k_refine_flags, n, st, lo, bld.B.ses.p, bld.B.set.p, idivflag, ndiv, bld.flag.p);
      if (l + 1 > d1 && pk[0]->d2 == 0) {
        // a box at level d1 may need splitting: extend the paths (rare: only
        // for points closer than 2^-21 of the domain)
        if (bld.scan(n) > 0) {
          const int d2 = std::min(kD2Max, std::min(62, nlmax) - d1);
          pkx[0].build(src, ns, root[0], root[1], dhalf_p, d1, d2, st);
          pkx[1].build(trg, nt, root[0], root[1], dhalf_p, d1, d2, st);
          pk[0] = &pkx[0];
          pk[1] = &pkx[1];
          bld.ks = pk[0]->view();
          bld.kt = pk[1]->view();
        }
      }
      const int base = bld.nboxes;
      bs.push_back(bsall[l + 1]);
      const int nref = bld.refine(lo, n, bs[l + 1], l + 1, 0);
      if (nref == 0) {
        bs.pop_back();
        nlevels = l;
        break;
      }
      ls.push_back(base);
      le.push_back(bld.nboxes);
      ++l;
      continue;
    }
    bld.B.grow(bld.nboxes + (size_t)reserve, bld.nboxes);
    int64_t ubmax = 1;
    for (int k = 0; k < K; ++k) ubmax = std::max(ubmax, ub[k]);
    bld.ensure_flag((int)ubmax);
    {
      const int h[3] = {lo, n, bld.nboxes};
      upload_bytes(dlev.p + 2 * l, h, 2 * sizeof(int), st);
      upload_bytes(dnb.p, h + 2, sizeof(int), st);
      for (int k = 0; k < K; ++k) {
        const int L = l + k;
        launch1d(k_refine_flags_lv, ub[k], st, dlev.p, L, bld.B.ses.p, bld.B.set.p, idivflag,
                 ndiv, bld.flag.p);
        bld.scan_only((int)ub[k]);
        hipLaunchKernelGGL(k_level_next, dim3(1), dim3(1), 0, st, ub[k], bld.flag.p, bld.offs.p,
                           dlev.p, L, dnb.p);
        launch1d(k_make_children_lv, ub[k], st, dlev.p, L, bld.flag.p, bld.offs.p,
                 0.5 * bsall[L + 1], bld.B.centers.p, bld.B.ilevel.p, bld.B.iparent.p,
                 bld.B.nchild.p, bld.B.ichild.p, bld.B.ses.p, bld.B.set.p, bld.B.iflag.p, bld.ks,
                 bld.kt);
      }
      std::vector<int> hl(2 * (K + 1));
      int hnb = 0;
      readback(hl.data(), dlev.p + 2 * l, sizeof(int) * 2 * (K + 1), st);
      readback(&hnb, dnb.p, sizeof(int), st);
      bld.nboxes = hnb;
      for (int k = 1; k <= K; ++k) {
        if (hl[2 * k + 1] == 0) {
          nlevels = l + k - 1;
          done = true;
          break;
        }
        bs.push_back(bsall[l + k]);
        ls.push_back(hl[2 * k]);
        le.push_back(hl[2 * k] + hl[2 * k + 1]);
      }
    }
    l += K;
  }
  ls.resize(nlevels + 1);
  le.resize(nlevels + 1);
  bs.resize(nlevels + 1);
  bld.computecoll(ls, le, bs, nlevels);

  // ---- level restriction (pts_tree_fix_lr, pts_tree2d.f:758-1061)
  std::vector<int> ts1(nlevels + 1, 0), te1(nlevels + 1, 0);
  int added1 = 0;
  if (nlevels >= 2) {
    int* iflag = bld.B.iflag.p;
    hip_throw(hipMemsetAsync(iflag, 0, sizeof(int) * bld.nboxes, st), "memset");
    {
      LevelReals dist{};
      for (int l = 2; l <= nlevels; ++l) dist.v[l] = 1.05 * (bs[l - 1] + bs[l - 2]) / 2.0;
      launch1d(k_flag1, le[nlevels] - ls[2], st, ls[2], dist, bld.B.ilevel.p, bld.B.iparent.p,
               bld.B.ncoll.p, bld.B.coll.p, bld.B.nchild.p, bld.B.centers.p, iflag);
    }
    for (int l = nlevels; l >= 1; --l) {
      const double distest = 1.05 * (bs[l] + bs[l - 1]) / 2.0;
      launch1d(k_flag2, le[l] - ls[l], st, ls[l], distest, bld.B.iparent.p, bld.B.ncoll.p,
               bld.B.coll.p, bld.B.nchild.p, bld.B.centers.p, iflag);
    }
    {
      // the flags are fixed before this pass, so all levels refine in one scan
      std::vector<int> lo, n, lev, first, last;
      std::vector<double> bsc;
      for (int l = 1; l <= nlevels - 2; ++l) {
        lo.push_back(ls[l]);
        n.push_back(le[l] - ls[l]);
        lev.push_back(l + 1);
        bsc.push_back(bs[l + 1]);
      }
      bld.refine_segments(lo, n, lev, bsc, first, last);
      for (int l = 1; l <= nlevels - 2; ++l) {
        ts1[l + 1] = first[l - 1];
        te1[l + 1] = last[l - 1];
        added1 += last[l - 1] - first[l - 1];
      }
    }
  }
  // With no box added by the flagged refinement no box carries flag 3, the
  // updateflags sweep below refines nothing and both reorganisations are the
  // identity: the tree is already level-restricted.
  if (nlevels >= 2 && added1 > 0) {
    bld.reorg(ls, le, ts1, te1, nlevels);
    bld.computecoll(ls, le, bs, nlevels);
    launch1d(k_keep3, bld.nboxes, st, bld.B.iflag.p);
    std::vector<int> ts(nlevels + 1, 0), te(nlevels + 1, 0);
    int added2 = 0;
    for (int l = 2; l <= nlevels - 2; ++l) {
      const double distest = 1.05 * (bs[l] + bs[l + 1]) / 2.0;
      // the original then the tail boxes of level l (their flags are independent)
      launch1d(k_updateflags, (le[l] - ls[l]) + (te[l] - ts[l]), st, le[l] - ls[l], ls[l], ts[l],
               distest, bld.B.ncoll.p, bld.B.coll.p, bld.B.ichild.p, bld.B.nchild.p,
               bld.B.centers.p, bld.B.iflag.p);
      {
        // the original then the tail boxes of level l, in one scan
        std::vector<int> first, last;
        bld.refine_segments({ls[l], ts[l]}, {le[l] - ls[l], te[l] - ts[l]}, {l + 1, l + 1},
                            {bs[l + 1], bs[l + 1]}, first, last);
        ts[l + 1] = first[0];
        te[l + 1] = last[1];
        added2 += te[l + 1] - ts[l + 1];
      }
      launch1d(k_coll, te[l + 1] - ts[l + 1], st, ts[l + 1], bs[l + 1], bld.B.iparent.p,
               bld.B.ichild.p, bld.B.centers.p, bld.B.ncoll.p, bld.B.coll.p);
    }
    // (with nothing added the second reorganisation is the identity and the
    // colleagues are those computed above)
    if (added2 > 0) {
      bld.reorg(ls, le, ts, te, nlevels);
      bld.computecoll(ls, le, bs, nlevels);
    }
  }

  // ---- final point order (pts_tree_sort, pts_tree2d.f:1222-1246)
  const int nb = bld.nboxes;
  t.nlevels = nlevels;
  t.nboxes = nb;
  t.lstart = ls;
  t.lend = le;
  t.boxsize = bs;
  t.isrc.alloc(std::max<int64_t>(ns, 1), st);
  t.itarg.alloc(std::max<int64_t>(nt, 1), st);
  DBuf<int>* perm[2] = {&t.isrc, &t.itarg};
  const int* se[2] = {bld.B.ses.p, bld.B.set.p};
  const int64_t npts[2] = {ns, nt};
  DBuf<int> big(4, st), bigl(nb, st);
  big.zero();
  for (int k = 0; k < 2; ++k) {
    if (npts[k] == 0) continue;
    // big[k]: a leaf beyond kLeafSortMax; big[2 + k]: leaves beyond kLeafSmall
    launch1d(k_leaf_sort_small, nb, st, bld.B.nchild.p, se[k], pk[k]->perm.p, perm[k]->p, bigl.p,
             big.p + 2 + k);
    const unsigned grid = (unsigned)std::min<int64_t>(4096, ceil_div(nb, kLeafSortWaves));
    hipLaunchKernelGGL(k_leaf_sort, dim3(grid), dim3(kLeafSortWaves * 64), 0, st, bigl.p,
                       big.p + 2 + k, se[k], pk[k]->perm.p, perm[k]->p, big.p + k);
    hip_throw(hipGetLastError(), "kernel launch");
  }
  int hbig[2] = {0, 0};
  readback(hbig, big.p, sizeof(hbig), st);
  for (int k = 0; k < 2; ++k) {
    if (npts[k] == 0 || !hbig[k]) continue;
    // a leaf too large for k_leaf_sort: stable sort of all points by leaf
    DBuf<unsigned> label(npts[k], st), lsorted(npts[k], st);
    DBuf<int> iota(npts[k], st);
    launch1d(k_leaf_label, nb, st, bld.B.nchild.p, se[k], pk[k]->perm.p, label.p);
    int bits = 1;
    while ((int64_t(1) << bits) <= npts[k]) ++bits;
    // values: input indices 0..n-1 in input order
    launch1d(k_iota, npts[k], st, iota.p);
    radix_sort(label.p, lsorted.p, iota.p, perm[k]->p, npts[k], bits, st);
  }
  t.isrcse = std::move(bld.B.ses);
  t.itargse = std::move(bld.B.set);
  t.centers = std::move(bld.B.centers);
  t.ilevel = std::move(bld.B.ilevel);
  t.iparent = std::move(bld.B.iparent);
  t.nchild = std::move(bld.B.nchild);
  t.ichild = std::move(bld.B.ichild);
  t.ncoll = std::move(bld.B.ncoll);
  t.coll = std::move(bld.B.coll);
}

// Points (sources + targets) held by the leaves of each level, levels beyond
// nbins - 1 counted in the last bin: per-block counts in LDS, then one global
// atomic per bin and block (integer counts, so the result does not depend on
// the order of the additions).
__global__ void k_leaf_hist(int64_t nb, const int* nchild, const int* ilevel, const int* ses,
                            const int* set, int nbins, unsigned long long* hist) {
  __shared__ unsigned long long h[kLeafHistMax];
  for (int k = threadIdx.x; k < nbins; k += blockDim.x) h[k] = 0;
  __syncthreads();
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb && nchild[b] == 0) {
    const int n = (ses[2 * b + 1] - ses[2 * b]) + (set[2 * b + 1] - set[2 * b]);
    if (n > 0) atomicAdd(&h[min(ilevel[b], nbins - 1)], (unsigned long long)n);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nbins; k += blockDim.x)
    if (h[k]) atomicAdd(&hist[k], h[k]);
}

// ---------------------------------------------------- key leaf histogram
// Leaf level of every point in the tree refined while a box holds more than
// ndiv sources (pts_tree_build's rule, common/pts_tree2d.f:551-631, without
// the level restriction of pts_tree_fix_lr), from the sorted source path
// keys alone: with w(j) = common levels of keys j and j + ndiv, the level-l box
// of a point holds more than ndiv sources iff some window j..j+ndiv of sorted
// sources lies in it -- for a source at sorted position i, iff w(j) >= l for a
// j in [i - ndiv, i]; for a target whose key sorts at position pos, iff
// min(w(j), common levels(key_j, target)) >= l for a j in [pos - ndiv - 1, pos].
// The leaf level is one more than the largest such l (0 if the root is a leaf).
__device__ __forceinline__ int key_common_levels(unsigned long long a, unsigned long long b,
                                                 int d1) {
  const unsigned long long x = a ^ b;
  if (x == 0) return d1;
  return (__clzll((long long)x) - (64 - 2 * d1)) >> 1;
}

__global__ void k_window_levels(int64_t m, int ndiv, int d1, const unsigned long long* keys,
                                unsigned char* w) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) w[j] = (unsigned char)key_common_levels(keys[j], keys[j + ndiv], d1);
}

// h[b] += the lanes of the wave that hold bin b (valid lanes only): one LDS
// add per distinct bin of the wave instead of one per lane
__device__ __forceinline__ void wave_bin_add(unsigned long long* h, bool valid, int b) {
  const int lane = threadIdx.x & 63;
  unsigned long long act = __ballot(valid);
  while (act) {  // uniform
    const int leader = __ffsll((long long)act) - 1;
    const int bl = __builtin_amdgcn_readlane(b, leader);
    const unsigned long long same = __ballot(valid && b == bl);
    if (lane == leader) atomicAdd(&h[bl], (unsigned long long)__popcll(same));
    act &= ~same;
  }
}

// Sources' leaf levels for ndiv <= kWinMax, kHistTile sorted keys per block of
// 256 threads.  Source i's leaf level is 1 + the largest w[j] =
// common levels(ks[j], ks[j + ndiv]) over the windows j in [i - ndiv, i]
// (0 outside [0, ns - ndiv - 1]): the maxima over ndiv + 1 consecutive w by
// doubling in LDS (after t passes m[x] = max w[x .. x + 2^t - 1]; the window
// is the union of two such ranges).
constexpr int kHistTile = 2048;
constexpr int kWinMax = 2048;
__global__ __launch_bounds__(256) void k_src_leaf_hist(int64_t ns, int ndiv, int d1,
                                                       const unsigned long long* __restrict__ ks,
                                                       int nbins, unsigned long long* hist) {
  __shared__ unsigned char m[kHistTile + kWinMax];
  __shared__ unsigned long long h[kLeafHistMax];
  for (int k = threadIdx.x; k < nbins; k += blockDim.x) h[k] = 0;
  const int64_t i0 = (int64_t)blockIdx.x * kHistTile;
  const int64_t last = ns - ndiv - 1;  // last window start
  const int len = kHistTile + ndiv;
  for (int x = threadIdx.x; x < len; x += 256) {
    const int64_t j = i0 - ndiv + x;
    m[x] = (j >= 0 && j <= last) ? (unsigned char)key_common_levels(ks[j], ks[j + ndiv], d1)
                                 : (unsigned char)0;
  }
  __syncthreads();
  constexpr int kPer = (kHistTile + kWinMax) / 256;
  int span = 1;
  while (2 * span <= ndiv + 1) {  // uniform
    unsigned char v[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int x = threadIdx.x + r * 256;
      v[r] = x < len ? m[x] : (unsigned char)0;
      if (x + span < len) v[r] = max(v[r], m[x + span]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
      const int x = threadIdx.x + r * 256;
      if (x < len) m[x] = v[r];
    }
    __syncthreads();
    span *= 2;
  }
  // span <= ndiv + 1 < 2 span: window [xi - ndiv, xi] = ranges from xi - ndiv and xi - span + 1
#pragma unroll
  for (int r = 0; r < kHistTile / 256; ++r) {
    const int64_t i = i0 + threadIdx.x + r * 256;
    const bool valid = i < ns;
    int b = 0;
    if (valid) {
      const int xi = (int)(i - i0) + ndiv;
      const int best = max((int)m[xi - ndiv], (int)m[xi - span + 1]);
      b = min(best + 1, nbins - 1);
    }
    wave_bin_add(h, valid, b);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nbins; k += blockDim.x)
    if (h[k]) atomicAdd(&hist[k], h[k]);
}

// Leaf levels of the cnt points from i0 on (sources below ns, then targets) by
// direct window scans: the general path (targets; ndiv > kWinMax; ns <= ndiv).
__global__ void k_key_leaf_hist(int64_t cnt, int64_t i0, int64_t ns, int ndiv, int d1,
                                const unsigned long long* __restrict__ skeys,
                                const unsigned char* __restrict__ w,
                                const unsigned long long* __restrict__ tkeys, int nbins,
                                unsigned long long* hist) {
  __shared__ unsigned long long h[kLeafHistMax];
  for (int k = threadIdx.x; k < nbins; k += blockDim.x) h[k] = 0;
  __syncthreads();
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = i0 + g;
  int bin = 0;
  if (g < cnt) {
    int lev = 0;
    if (ns > ndiv) {
      const int64_t last = ns - ndiv - 1;  // last window start
      int best = 0;
      if (i < ns) {
        for (int64_t j = max((int64_t)0, i - ndiv); j <= min(i, last); ++j)
          best = max(best, (int)w[j]);
      } else {
        const unsigned long long t = tkeys[i - ns];
        int64_t lo = 0, hi = ns;  // pos = first sorted source key >= t
        while (lo < hi) {
          const int64_t mid = (lo + hi) >> 1;
          if (skeys[mid] < t) lo = mid + 1;
          else hi = mid;
        }
        for (int64_t j = max((int64_t)0, lo - ndiv - 1); j <= min(lo, last); ++j)
          best = max(best, min((int)w[j], key_common_levels(skeys[j], t, d1)));
      }
      lev = best + 1;
    }
    bin = min(lev, nbins - 1);
  }
  wave_bin_add(h, g < cnt, bin);
  __syncthreads();
  for (int k = threadIdx.x; k < nbins; k += blockDim.x)
    if (h[k]) atomicAdd(&hist[k], h[k]);
}

namespace {
// hist[l] from the sorted source keys ks and the target keys kt (any order),
// both of d1 levels
void hist_from_keys(const unsigned long long* ks, int64_t ns, const unsigned long long* kt,
                    int64_t nt, int d1, int ndiv, int nbins, double* hist, hipStream_t st) {
  DBuf<unsigned long long> d(nbins, st);
  d.zero();
  // sources by LDS window maxima when the window fits, the rest by scans
  const bool tiled = ns > ndiv && ndiv <= kWinMax;
  if (tiled)
    hipLaunchKernelGGL(k_src_leaf_hist, dim3((unsigned)ceil_div(ns, (int64_t)kHistTile)),
                       dim3(256), 0, st, ns, ndiv, d1, ks, nbins, d.p);
  const int64_t i0 = tiled ? ns : 0;
  if (ns + nt > i0) {
    const int64_t m = ns > ndiv ? ns - ndiv : 0;
    DBuf<unsigned char> w(std::max<int64_t>(m, 1), st);
    launch1d(k_window_levels, m, st, ndiv, d1, ks, w.p);
    launch1d(k_key_leaf_hist, ns + nt - i0, st, i0, ns, ndiv, d1, ks,
             (const unsigned char*)w.p, kt, nbins, d.p);
  }
  std::vector<unsigned long long> h(nbins);
  readback(h.data(), d.p, sizeof(unsigned long long) * nbins, st);
  for (int k = 0; k < nbins; ++k) hist[k] = (double)h[k];
}
}  // namespace

void key_leaf_histogram(const double* src, int64_t ns, const double* trg, int64_t nt, int ndiv,
                        int nbins, double* hist, double* boxsize0, hipStream_t st) {
  if (nbins < 1 || nbins > kLeafHistMax) throw HipError{hipErrorInvalidValue, "histogram bins"};
  double bb[4];
  {
    const int nb = (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, ceil_div(ns + nt, 256)));
    DBuf<double> part(4 * nb, st), out(4, st);
    hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(256), 0, st, src, ns, trg, nt, part.p);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, st, part.p, nb, out.p);
    readback(bb, out.p, sizeof(bb), st);
  }
  const double bs = std::max(bb[1] - bb[0], bb[3] - bb[2]);
  if (boxsize0) *boxsize0 = bs;
  // levels up to nbins - 1 tell every bin apart (the last bin takes the leaf
  // levels from nbins - 1 on): shorter keys, fewer sort passes
  const int d1 = std::max(1, std::min(kD1, nbins - 1));
  // half[l] = boxsize(l) / 2 with boxsize(l + 1) = boxsize(l) / 2, as in build_tree
  std::vector<double> half(d1 + 2);
  double bsl = bs;
  for (int l = 0; l < (int)half.size(); ++l) {
    half[l] = 0.5 * bsl;
    bsl = bsl / 2;
  }
  DBuf<double> dhalf(half.size(), st);
  upload_bytes(dhalf.p, half.data(), sizeof(double) * half.size(), st);
  const double cx = (bb[0] + bb[1]) / 2, cy = (bb[2] + bb[3]) / 2;
  DBuf<unsigned long long> k0(ns, st), ks(ns, st), kt(std::max<int64_t>(nt, 1), st);
  launch1d(k_path_keys, ns, st, src, cx, cy, (const double*)dhalf.p, d1, 0, k0.p,
           (unsigned long long*)nullptr, (int*)nullptr);
  if (nt > 0)
    launch1d(k_path_keys, nt, st, trg, cx, cy, (const double*)dhalf.p, d1, 0, kt.p,
             (unsigned long long*)nullptr, (int*)nullptr);
  radix_sort(k0.p, ks.p, (const int*)nullptr, (int*)nullptr, ns, 2 * d1, st);
  hist_from_keys(ks.p, ns, kt.p, nt, d1, ndiv, nbins, hist, st);
}

void prepared_leaf_histogram(const TreePrep* P, int ndiv, int nbins, double* hist,
                             double* boxsize0, hipStream_t st) {
  if (nbins < 1 || nbins > kLeafHistMax) throw HipError{hipErrorInvalidValue, "histogram bins"};
  if (boxsize0) *boxsize0 = P->bsall[0];
  // the prepared keys hold min(kD1, nlmax) levels; every level from
  // nbins - 1 on falls in the last bin, so deeper keys give the same counts
  hist_from_keys(P->pk[0].k1.p, P->ns, P->pk[1].k1.p, P->nt, P->pk[0].d1, ndiv, nbins, hist, st);
}

void leaf_level_histogram(const Tree& t, int nbins, double* hist, hipStream_t st) {
  if (nbins < 1 || nbins > kLeafHistMax) throw HipError{hipErrorInvalidValue, "histogram bins"};
  DBuf<unsigned long long> d(nbins, st);
  d.zero();
  launch1d(k_leaf_hist, t.nboxes, st, t.nchild.p, t.ilevel.p, t.isrcse.p, t.itargse.p, nbins,
           d.p);
  std::vector<unsigned long long> h(nbins);
  readback(h.data(), d.p, sizeof(unsigned long long) * nbins, st);
  for (int k = 0; k < nbins; ++k) hist[k] = (double)h[k];
}

}  // namespace fmm2d

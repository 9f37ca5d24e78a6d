// Pieces shared by the FMM drivers (Laplace/Cauchy and Helmholtz): point
// gather/scatter, leaf lookup, the child and list-2 work plans that feed the
// MFMA translate kernel, work counters, upload and stage timing.
//
// Everything here has internal linkage (anonymous namespace), so each driver
// translation unit carries its own copy of the kernels.
#pragma once

#include <algorithm>
#include <vector>

#include "devbuf.h"
#include "scan.h"
#include "tree.h"

namespace fmm2d {
namespace {

constexpr int kW = 64;  // wavefront size
constexpr int kOff = 7;  // offsets -3..3 per axis

FMM2D_HD cplx cpowi(cplx w, int k) {
  cplx r = mkc(1.0), b = w;
  while (k) {
    if (k & 1) r = r * b;
    b = b * b;
    k >>= 1;
  }
  return r;
}

__device__ __forceinline__ int npts_eval(int b, const int* sse, const int* tse, int ifpgh,
                                         int ifpght) {
  int n = 0;
  if (ifpgh > 0) n += sse[2 * b + 1] - sse[2 * b];
  if (ifpght > 0) n += tse[2 * b + 1] - tse[2 * b];
  return n;
}


// ---------------------------------------------------------- gather / scatter
__global__ void k_gather_xy(int64_t n, const int* perm, const double* xy, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  out[2 * i] = xy[2 * (int64_t)j];
  out[2 * i + 1] = xy[2 * (int64_t)j + 1];
}

// cfmm2d: densities copied as they are.
__global__ void k_gather_dens(int64_t n, int nd, const int* perm, const cplx* in, cplx* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t j = perm[i];
  for (int d = 0; d < nd; ++d) out[i * nd + d] = in[j * nd + d];
}


// out[perm[i]] = in[i] for m values per point
__global__ void k_scatter(int64_t n, int m, const int* perm, const cplx* in, cplx* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t j = perm[i];
  for (int k = 0; k < m; ++k) out[j * m + k] = in[i * m + k];
}


// leaf[i] = the leaf holding sorted point i.  A wavefront takes 64 boxes and
// writes their point ranges one box after the other with all 64 lanes
// (consecutive lanes on consecutive points).
__global__ void k_leaf_of(int64_t n, const int* nchild, const int* se, int* leaf) {
  const int lane = threadIdx.x & 63;
  const int64_t b0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63);
  const int64_t b = b0 + lane;
  int s = 0, e = 0;
  if (b < n && nchild[b] == 0) {
    s = se[2 * b];
    e = se[2 * b + 1];
  }
  for (int k = 0; k < 64; ++k) {
    const int sk = __builtin_amdgcn_readlane(s, k), ek = __builtin_amdgcn_readlane(e, k);
    for (int i = sk + lane; i < ek; i += 64) leaf[i] = (int)(b0 + k);
  }
}


// ---------------------------------------------------------- near-field segments
// The sources a leaf interacts with directly (the boxes of its list 1,
// common/tree_routs2d.f:486-710) as a few contiguous ranges of the sorted
// sources: empty boxes dropped, ranges sorted by start and merged where one
// ends at the next one's start (sibling leaves are adjacent in the sorted
// order).  Row b of kSegRow int2: [0] = {sources in total, ranges}, then
// the ranges [s0, s1), then an end mark {0, 0} (every box has a row; that of
// a box with an empty list 1 holds no range).
constexpr int kSegRow = 16;
static_assert(kSegRow >= kMaxList1 + 2, "segment row holds the list-1 ranges and the end mark");

__global__ __launch_bounds__(256) void k_near_segs(int64_t nb, const int* n1, const int* l1,
                                                   const int* sse, int2* seg) {
  __shared__ int2 buf[256][kMaxList1];
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int c1 = n1[b];
  int2* r = buf[threadIdx.x];
  int m = 0;
  for (int k = 0; k < c1; ++k) {  // insertion sort by start
    const int c = l1[(int64_t)kMaxList1 * b + k];
    const int s0 = sse[2 * c], s1 = sse[2 * c + 1];
    if (s1 <= s0) continue;
    int q = m++;
    for (; q > 0 && r[q - 1].x > s0; --q) r[q] = r[q - 1];
    r[q] = make_int2(s0, s1);
  }
  int2* out = seg + (int64_t)kSegRow * b;
  int ns = 0, tot = 0;
  for (int k = 0; k < m;) {
    int2 s = r[k++];
    for (; k < m && r[k].x == s.y; ++k) s.y = r[k].y;
    out[1 + ns++] = s;
    tot += s.y - s.x;
  }
  out[1 + ns] = make_int2(0, 0);
  out[0] = make_int2(tot, ns);
}

// true when pred holds on every active lane of the wavefront
__device__ __forceinline__ bool wave_all(bool pred) { return __ballot(pred) == __ballot(1); }

// Walks the ranges of one segment row: current range [j, e) and the next
// one prefetched, advanced when the current one runs out.
struct SegCursor {
  const int2* row;
  int j, e, k, last;
  int2 nxt;
  __device__ __forceinline__ int init(const int2* r) {
    row = r;
    const int2 hd = r[0];
    last = hd.y + 1;  // slot of the end mark
    const int2 cur = r[1];
    nxt = r[min(2, last)];
    k = 3;
    j = cur.x;
    e = cur.y;
    return hd.x;
  }
  __device__ __forceinline__ void step() {
    if (j == e) {
      j = nxt.x;
      e = nxt.y;
      nxt = row[min(k, last)];
      ++k;
    }
  }
};


// ---------------------------------------------------------- list-2 plan
// The M2L of one target box is sum over its list-2 sources of T_o M_c, with o
// the source's offset.  For a target in child position q of its parent the
// possible offsets form a fixed set of 27 (the 6x6 children of the parent's
// 3x3 neighbourhood minus the 3x3 adjacent ones), so targets of the same q are
// processed together by the translate kernel, one offset slot after another.
__device__ __forceinline__ bool m2l_target(int b, const int* ilevel, const int* n2, const int* sse,
                                           const int* tse, int ifpgh, int ifpght) {
  return ilevel[b] >= 2 && n2[b] > 0 && npts_eval(b, sse, tse, ifpgh, ifpght) > 0;
}

// flag[p] = 1 if some child of box p is an M2L target
__global__ void k_m2l_parent_flag(int64_t n, const int* ilevel, const int* nchild,
                                  const int* ichild, const int* n2, const int* sse,
                                  const int* tse, int ifpgh, int ifpght, int* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int f = 0;
  if (nchild[i] > 0)
    for (int q = 0; q < 4; ++q) {
      const int c = ichild[4 * i + q];
      if (c >= 0 && m2l_target(c, ilevel, n2, sse, tse, ifpgh, ifpght)) f = 1;
    }
  flag[i] = f;
}

// *total = flags set in flag[0..n), from its exclusive scan pos
// out[pos[i]] = i for every flagged i (pos = exclusive scan of flag)
__global__ void k_compact_flagged(int64_t n, const int* flag, const int* pos, int* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) out[pos[i]] = (int)i;
}

// tlist[q * np + k] = child q of parent plist[k] if an M2L target, else -1
__global__ void k_m2l_aligned(int64_t np, const int* plist, const int* ilevel, const int* ichild,
                              const int* n2, const int* sse, const int* tse, int ifpgh,
                              int ifpght, int* tlist) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int p = plist[k];
  for (int q = 0; q < 4; ++q) {
    const int c = ichild[4 * p + q];
    tlist[q * np + k] =
        (c >= 0 && m2l_target(c, ilevel, n2, sse, tse, ifpgh, ifpght)) ? c : -1;
  }
}

// Non-leaf boxes (in box order, hence level order) and their 4-child slots for
// M2M (a child without sources is skipped, laprouts2d.f via cfmm2d.f:962-983)
// and L2L (all children when the parent has points to evaluate, cfmm2d.f:1042-1076).
__global__ void k_nonleaf_flag(int64_t n, const int* nchild, int* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = nchild[i] > 0 ? 1 : 0;
}

__global__ void k_child_slots(int64_t n, const int* list, const int* ichild, const int* sse,
                              const int* tse, int ifpgh, int ifpght, int* up, int* down) {
  const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= n) return;
  const int b = list[pos];
  const bool ev = npts_eval(b, sse, tse, ifpgh, ifpght) > 0;
  for (int qq = 0; qq < 4; ++qq) {
    const int c = ichild[4 * b + qq];
    up[pos * 4 + qq] = (sse[2 * c + 1] > sse[2 * c]) ? c : -1;
    down[pos * 4 + qq] = ev ? c : -1;
  }
}

// out[l] = first position of list (sorted) at or beyond lstart[l], one thread
// per l (launch >= nl + 2 threads)
__global__ void k_level_bounds(const int* list, int n, const int* lstart, int nl, int* out) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l > nl + 1) return;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (list[mid] < lstart[l]) lo = mid + 1;
    else hi = mid;
  }
  out[l] = lo;
}


// ---------------------------------------------------------- work counters
// cnt[0] near pairs, [1] list-2 pairs, [2] M2M edges, [3] L2L edges,
// [4] list-3 (point, box) pairs, [5] list-4 (source, box) pairs; M2M edges
// count at the parent, L2L edges at the child
__global__ void k_counts(int64_t n, const int* ilevel, const int* iparent, const int* nchild,
                         const int* ichild, const int* sse, const int* tse, int ifpgh,
                         int ifpght, const int* n1, const int* l1, const int* n2, const int* n3,
                         const int* n4, const int* l4, int b0, unsigned long long* cnt) {
  // boxes b0 .. b0+n-1; cnt is the same for every lane, so the compiler
  // combines each wavefront's atomics into one
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = b0 + (int)i;
  const int lev = ilevel[b];
  const int ne = npts_eval(b, sse, tse, ifpgh, ifpght);
  unsigned long long near = 0, l3 = 0, l4c = 0, m2m = 0, l2l = 0, m2l = 0;
  if (nchild[b] == 0) {
    unsigned long long s = 0;
    for (int k = 0; k < n1[b]; ++k) {
      const int c = l1[(int64_t)kMaxList1 * b + k];
      s += sse[2 * c + 1] - sse[2 * c];
    }
    near = s * (unsigned long long)ne;
    l3 = (unsigned long long)n3[b] * ne;
  } else if (lev >= 1) {
    for (int q = 0; q < 4; ++q) {
      const int c = ichild[4 * b + q];
      if (sse[2 * c + 1] > sse[2 * c]) ++m2m;
    }
  }
  if (lev >= 2 && ne > 0) {
    m2l = n2[b];
    for (int k = 0; k < n4[b]; ++k) {
      const int c = l4[(int64_t)kMaxList4 * b + k];
      l4c += sse[2 * c + 1] - sse[2 * c];
    }
    if (npts_eval(iparent[b], sse, tse, ifpgh, ifpght) > 0) l2l = 1;
  }
  if (near) atomicAdd(&cnt[0], near);
  if (m2l) atomicAdd(&cnt[1], m2l);
  if (m2m) atomicAdd(&cnt[2], m2m);
  if (l2l) atomicAdd(&cnt[3], l2l);
  if (l3) atomicAdd(&cnt[4], l3);
  if (l4c) atomicAdd(&cnt[5], l4c);
}


template <typename T> DBuf<T> upload(const std::vector<T>& v, hipStream_t st) {
  static_assert(sizeof(T) % 4 == 0, "upload_bytes moves 4-byte words");
  DBuf<T> d(v.size(), st);
  upload_bytes(d.p, v.data(), sizeof(T) * v.size(), st);  // v may die after return
  return d;
}


struct Timer {
  bool on;
  hipStream_t st;
  std::vector<hipEvent_t> ev;
  std::vector<int> slot;
  Timer(bool enable, hipStream_t s) : on(enable), st(s) { mark(-1); }
  void mark(int s) {
    if (!on) return;
    hipEvent_t e;
    (void)hipEventCreate(&e);
    (void)hipEventRecord(e, st);
    ev.push_back(e);
    slot.push_back(s);
  }
  // after a stream sync: accumulate interval k (ending at event k) into slot
  void collect(double* out) {
    if (!on) return;
    for (size_t k = 1; k < ev.size(); ++k) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev[k - 1], ev[k]);
      if (slot[k] >= 0) out[slot[k]] += 1e-3 * ms;
    }
    float tot = 0.f;
    (void)hipEventElapsedTime(&tot, ev.front(), ev.back());
    out[11] = 1e-3 * tot;
  }
  ~Timer() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};


// Non-leaf boxes level by level with their child slots: up[4 pos + q] is
// child q if it holds sources (M2M input), down[4 pos + q] is child q if the
// parent has points to evaluate (L2L output).  list[lev_pos[l] .. lev_pos[l+1])
// are the non-leaf boxes of level l.
struct ChildPlan {
  DBuf<int> list, up, down;
  std::vector<int> lev_pos;
};

void plan_children(const Tree& t, int pg, int pgt, ChildPlan& cp, hipStream_t st) {
  const int nb = t.nboxes, nl = t.nlevels;
  cp.list.alloc(nb, st);
  cp.lev_pos.assign(nl + 2, 0);
  DBuf<int> flag(nb, st), cnt(1, st), dlb(nl + 2, st);
  launch1d(k_nonleaf_flag, nb, st, t.nchild.p, flag.p);
  {
    // stream compaction of the flagged box ids (exclusive scan, then scatter)
    DBuf<int> pos(nb, st);
    exclusive_scan(flag.p, pos.p, nb, cnt.p, st);
    launch1d(k_compact_flagged, nb, st, (const int*)flag.p, (const int*)pos.p, cp.list.p);
  }
  int nnl = 0;
  readback(&nnl, cnt.p, sizeof(int), st);
  std::vector<int> lsx(t.lstart);
  lsx.push_back(nb);
  DBuf<int> d_lsx = upload(lsx, st);
  if (nl + 2 > 64) throw HipError{hipErrorInvalidValue, "tree levels beyond 62"};
  hipLaunchKernelGGL(k_level_bounds, dim3(1), dim3(64), 0, st, cp.list.p, nnl, d_lsx.p, nl,
                     dlb.p);
  cp.up.alloc(4 * (size_t)std::max(nnl, 1), st);
  cp.down.alloc(4 * (size_t)std::max(nnl, 1), st);
  launch1d(k_child_slots, nnl, st, cp.list.p, t.ichild.p, t.isrcse.p, t.itargse.p, pg, pgt,
           cp.up.p, cp.down.p);
  readback(cp.lev_pos.data(), dlb.p, sizeof(int) * (nl + 2), st);
}

// seg[q * (nl + 2) + l]: first position in tlist of class q at level >= l
// M2L targets (level >= 2, non-empty list 2, points to evaluate) by child
// position q (their sources per offset slot are Lists::s2).  The classes are
// aligned by parent: entry k of class q is the q-th child of the k-th parent
// with an M2L target among its children (parents in box order), or -1 when
// that child is no target -- so sibling targets, whose list-2 sources are
// mostly the same boxes, sit at the same position of the four classes (few
// entries are -1: a refined box's children nearly always all hold points).
// cls_seg[q] .. cls_seg[q+1] is class q, and lev_seg[q][l] .. lev_seg[q][l+1]
// its part at level l.
struct M2LPlan {
  DBuf<int> tlist;
  int cls_seg[5] = {0, 0, 0, 0, 0};
  std::vector<int> lev_seg;  // 4 x (nl + 2)
  int ntg = 0;
  int lev(int q, int l, int nl) const { return lev_seg[(size_t)q * (nl + 2) + l]; }
};

void plan_m2l(const Tree& t, const Lists& L, int pg, int pgt, M2LPlan& mp, hipStream_t st) {
  const int nb = t.nboxes, nl = t.nlevels;
  DBuf<int> flag(nb, st), pos(nb + 1, st), cnt(1, st);
  launch1d(k_m2l_parent_flag, nb, st, t.ilevel.p, t.nchild.p, t.ichild.p, L.n2.p, t.isrcse.p,
           t.itargse.p, pg, pgt, flag.p);
  exclusive_scan(flag.p, pos.p, nb, cnt.p, st);
  int np = 0;
  readback(&np, cnt.p, sizeof(int), st);
  DBuf<int> plist(std::max(np, 1), st);
  launch1d(k_compact_flagged, nb, st, (const int*)flag.p, (const int*)pos.p, plist.p);
  mp.tlist.alloc(4 * (size_t)std::max(np, 1), st);
  launch1d(k_m2l_aligned, np, st, plist.p, t.ilevel.p, t.ichild.p, L.n2.p, t.isrcse.p,
           t.itargse.p, pg, pgt, mp.tlist.p);
  // targets of level l are children of parents of level l - 1: their rows
  // start at the first parent at or beyond lstart[l - 1]
  std::vector<int> lsx(nl + 2, 0);
  for (int l = 1; l <= nl + 1; ++l) lsx[l] = (l - 1 <= nl) ? t.lstart[l - 1] : nb;
  lsx[0] = 0;
  DBuf<int> d_lsx = upload(lsx, st);
  DBuf<int> dlb(nl + 2, st);
  if (nl + 2 > 64) throw HipError{hipErrorInvalidValue, "tree levels beyond 62"};
  hipLaunchKernelGGL(k_level_bounds, dim3(1), dim3(64), 0, st, plist.p, np, d_lsx.p, nl,
                     dlb.p);
  std::vector<int> lb(nl + 2, 0);
  readback(lb.data(), dlb.p, sizeof(int) * lb.size(), st);
  for (int q = 0; q <= 4; ++q) mp.cls_seg[q] = q * np;
  mp.lev_seg.assign(4 * (size_t)(nl + 2), 0);
  for (int q = 0; q < 4; ++q)
    for (int l = 0; l <= nl + 1; ++l) mp.lev_seg[(size_t)q * (nl + 2) + l] = q * np + lb[l];
  mp.ntg = 4 * np;
}

// slot -> offset index (dx+3)*7 + (dy+3) for each child position q
std::vector<int> m2l_slot_offsets() {
  std::vector<int> soff(4 * kSlots, 0);
  for (int q = 0; q < 4; ++q)
    for (int dx = -3; dx <= 3; ++dx)
      for (int dy = -3; dy <= 3; ++dy) {
        const int sl = m2l_slot(q, dx, dy);
        if (sl >= 0) soff[q * kSlots + sl] = (dx + 3) * kOff + (dy + 3);
      }
  return soff;
}

}  // namespace
}  // namespace fmm2d

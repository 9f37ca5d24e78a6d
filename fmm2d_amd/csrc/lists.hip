// Interaction lists 1-4 (computelists, common/tree_routs2d.f:486-710).
//
// The reference builds them in one serial sweep that also appends to the
// lists of OTHER boxes (list1 and list4 of the finer neighbours of a leaf).
// Here every box gathers its own entries ("pull" form), so each box's
// lanes write only its own rows and no atomics are needed:
//   list2(b): children of the parent's colleagues not adjacent to b
//             (|dx| >= 1.05 bs or |dy| >= 1.05 bs, single-precision 1.05);
//   list1(b), leaf b: leaf colleagues (b included), children of non-leaf
//             colleagues that touch b, and coarser leaves (colleagues of b's
//             parent) that touch b;
//   list3(b), leaf b: children of non-leaf colleagues that do not touch b;
//   list4(b): coarser leaves among the parent's colleagues that do not
//             touch b.
// "Touch" uses distest = 1.05*(bs_l + bs_{l+1})/2 with strict <, as the
// reference.  Entries are the same sets as the reference's; their order
// within a row may differ (only summation order depends on it).
#include <algorithm>

#include "tree.h"

namespace fmm2d {
namespace {

constexpr double kColl = (double)1.05f;

// Three boxes per wavefront, boxes taken in a grid-stride loop (kListWaves
// wavefronts per block).  Box s of a wavefront has lanes 18 s .. 18 s + 17; its
// lane l owns the candidates 2 l and 2 l + 1 of the (colleague k, child slot q)
// sequence c = 4 k + q.  Entries are appended in candidate order through
// ballots restricted to the box's lanes, so every row is in the order of a
// serial sweep over (k, q); the M2L slot rows are assembled in LDS.
constexpr int kListWaves = 4;
constexpr int kBoxLanes = 18;  // 2 kBoxLanes >= 4 kMaxColl candidates
constexpr int kBoxesPerWave = 3;

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Appends the entries (e0 if p0, then e1 if p1, for each lane in order) of
// the lanes in `seg` after `cnt` entries of `row`; returns the new count.
__device__ __forceinline__ int append2(bool p0, int e0, bool p1, int e1, int lane,
                                       unsigned long long seg, int cnt, int cap, int* row) {
  const unsigned long long m0 = __ballot(p0) & seg, m1 = __ballot(p1) & seg;
  const unsigned long long lt = (1ull << lane) - 1ull;
  const int before = __popcll(m0 & lt) + __popcll(m1 & lt);
  if (p0 && cnt + before < cap) row[cnt + before] = e0;
  const int r1 = before + (p0 ? 1 : 0);
  if (p1 && cnt + r1 < cap) row[cnt + r1] = e1;
  return cnt + __popcll(m0) + __popcll(m1);
}

__global__ __launch_bounds__(kListWaves * 64) void k_lists(
    int nb, const int* ilevel, const int* iparent, const int* nchild, const int* ichild,
    const int* ncoll, const int* coll, const double* centers, const double* bsl, int* n1, int* l1,
    int* n2, int* l2, int* n3, int* l3, int* n4, int* l4, int* s2) {
  __shared__ int slots[kListWaves][kBoxesPerWave][kSlots];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sb = lane / kBoxLanes;  // box of this lane in the wavefront (3: none)
  const int ll = lane - sb * kBoxLanes;
  const bool inbox = sb < kBoxesPerWave;
  const unsigned long long seg =
      inbox ? (((1ull << kBoxLanes) - 1ull) << (sb * kBoxLanes)) : 0ull;
  int* ssh = slots[wv][inbox ? sb : 0];
  const int k = ll >> 1, q0 = 2 * (ll & 1);  // candidates (k, q0), (k, q0 + 1)
  const int64_t step = (int64_t)gridDim.x * kListWaves * kBoxesPerWave;
  for (int64_t base = ((int64_t)blockIdx.x * kListWaves + wv) * kBoxesPerWave; base < nb;
       base += step) {
    const int64_t bb = base + sb;
    const bool act = inbox && bb < nb;
    const int b = act ? (int)bb : 0;
    const int lev = ilevel[b];
    const double bx = centers[2 * b], by = centers[2 * b + 1];
    int* r1 = l1 + (int64_t)kMaxList1 * b;
    int* r2 = l2 + (int64_t)kMaxList2 * b;
    int* r3 = l3 + (int64_t)kMaxList3 * b;
    int* r4 = l4 + (int64_t)kMaxList4 * b;
    int c1 = 0, c2 = 0, c3 = 0, c4 = 0;
    if (act)
      for (int t = ll; t < kSlots; t += kBoxLanes) ssh[t] = -1;
    wave_sync_lds();
    const bool root = act && b == 0;
    const bool nonroot = act && b != 0;
    // the root: list 1 = itself when it is a leaf
    c1 = append2(root && ll == 0 && nchild[0] == 0, 0, false, 0, lane, seg, c1, kMaxList1, r1);
    // list 2 from the children of the parent's colleagues, and their slots
    const int dad = nonroot ? iparent[b] : 0;
    const unsigned long long mq =
        __ballot(nonroot && ll < 3 && ichild[4 * dad + ll] == b) & seg;
    const unsigned long long mqs = inbox ? (mq >> (sb * kBoxLanes)) : 0ull;
    const int qb = mqs ? __ffsll((long long)mqs) - 1 : 3;  // child position of b
    const int ncd = nonroot ? ncoll[dad] : 0;
    const double lim = kColl * bsl[lev];
    const double binv = 1.0 / bsl[lev];
    int kb0 = -1, kb1 = -1;
    if (k < ncd) {
      const int j = coll[9 * dad + k];
      const int nch = nchild[j];
      if (q0 < nch) kb0 = ichild[4 * j + q0];
      if (q0 + 1 < nch) kb1 = ichild[4 * j + q0 + 1];
    }
    bool far0 = false, far1 = false;
    if (kb0 >= 0) {
      const double dx = centers[2 * kb0] - bx, dy = centers[2 * kb0 + 1] - by;
      far0 = fabs(dx) >= lim || fabs(dy) >= lim;
      if (far0) {
        const int sl = m2l_slot(qb, (int)rint(dx * binv), (int)rint(dy * binv));
        if (sl >= 0) ssh[sl] = kb0;
      }
    }
    if (kb1 >= 0) {
      const double dx = centers[2 * kb1] - bx, dy = centers[2 * kb1 + 1] - by;
      far1 = fabs(dx) >= lim || fabs(dy) >= lim;
      if (far1) {
        const int sl = m2l_slot(qb, (int)rint(dx * binv), (int)rint(dy * binv));
        if (sl >= 0) ssh[sl] = kb1;
      }
    }
    c2 = append2(far0, kb0, far1, kb1, lane, seg, c2, kMaxList2, r2);
    // entries contributed by coarser leaves (colleagues of the parent), one
    // lane per colleague (the even lanes)
    {
      bool to1 = false, to4 = false;
      int j = -1;
      if (nonroot && lev >= 2 && (ll & 1) == 0 && k < ncd) {
        j = coll[9 * dad + k];
        if (nchild[j] == 0) {
          const double dist = 1.05 * (bsl[lev - 1] + bsl[lev]) / 2.0;
          const double xd = fabs(centers[2 * j] - bx), yd = fabs(centers[2 * j + 1] - by);
          to1 = xd < dist && yd < dist;
          to4 = !to1;
        }
      }
      c1 = append2(to1, j, false, 0, lane, seg, c1, kMaxList1, r1);
      c4 = append2(to4, j, false, 0, lane, seg, c4, kMaxList4, r4);
    }
    // own entries of a leaf: a leaf colleague itself (on the lane of its slot
    // 0), or the children of a non-leaf colleague
    {
      bool a0 = false, a1 = false, t0 = false, t1 = false;  // to list 1 / list 3
      int e0 = -1, e1 = -1;
      if (nonroot && nchild[b] == 0 && k < ncoll[b]) {
        const int j = coll[9 * b + k];
        const int nch = nchild[j];
        const double dist = 1.05 * (bsl[lev] + bsl[lev + 1]) / 2.0;
        if (nch == 0) {
          if (q0 == 0) {
            e0 = j;
            a0 = true;
          }
        } else {
          if (q0 < nch) {
            e0 = ichild[4 * j + q0];
            const double xd = fabs(centers[2 * e0] - bx), yd = fabs(centers[2 * e0 + 1] - by);
            a0 = xd < dist && yd < dist;
            t0 = !a0;
          }
          if (q0 + 1 < nch) {
            e1 = ichild[4 * j + q0 + 1];
            const double xd = fabs(centers[2 * e1] - bx), yd = fabs(centers[2 * e1 + 1] - by);
            a1 = xd < dist && yd < dist;
            t1 = !a1;
          }
        }
      }
      c1 = append2(a0, e0, a1, e1, lane, seg, c1, kMaxList1, r1);
      c3 = append2(t0, e0, t1, e1, lane, seg, c3, kMaxList3, r3);
    }
    wave_sync_lds();
    if (act) {
      for (int t = ll; t < kSlots; t += kBoxLanes) s2[(int64_t)kSlots * b + t] = ssh[t];
      if (ll == 0) {
        n1[b] = min(c1, kMaxList1);
        n2[b] = min(c2, kMaxList2);
        n3[b] = min(c3, kMaxList3);
        n4[b] = min(c4, kMaxList4);
      }
    }
    wave_sync_lds();
  }
}
}  // namespace

void build_lists(const Tree& t, Lists& L, hipStream_t st) {
  const int nb = t.nboxes;
  L.n1.alloc(nb, st);
  L.n2.alloc(nb, st);
  L.n3.alloc(nb, st);
  L.n4.alloc(nb, st);
  L.l1.alloc((size_t)kMaxList1 * nb, st);
  L.l2.alloc((size_t)kMaxList2 * nb, st);
  L.l3.alloc((size_t)kMaxList3 * nb, st);
  L.l4.alloc((size_t)kMaxList4 * nb, st);
  L.s2.alloc((size_t)kSlots * nb, st);
  // boxsize per level, with one extra (finer) entry for the leaf test at the
  // deepest level, where computelists never reaches it (no non-leaf colleague)
  std::vector<double> bsl(t.boxsize);
  bsl.push_back(bsl.back() / 2);
  DBuf<double> dbs(bsl.size(), st);
  upload_bytes(dbs.p, bsl.data(), sizeof(double) * bsl.size(), st);
  if (nb > 0) {
    const unsigned grid =
        (unsigned)std::min<int64_t>(8192, ceil_div(nb, kListWaves * kBoxesPerWave));
    hipLaunchKernelGGL(k_lists, dim3(grid), dim3(kListWaves * 64), 0, st, nb, t.ilevel.p,
                       t.iparent.p, t.nchild.p, t.ichild.p, t.ncoll.p, t.coll.p, t.centers.p,
                       dbs.p, L.n1.p, L.l1.p, L.n2.p, L.l2.p, L.n3.p, L.l3.p, L.n4.p, L.l4.p,
                       L.s2.p);
    hip_throw(hipGetLastError(), "kernel launch");
  }
}  // dbs is released stream-ordered, after k_lists

}  // namespace fmm2d

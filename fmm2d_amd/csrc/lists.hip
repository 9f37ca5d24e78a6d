// Interaction lists 1-4 (computelists, common/tree_routs2d.f:486-710).
//
// The reference builds them in one serial sweep that also appends to the
// lists of OTHER boxes (list1 and list4 of the finer neighbours of a leaf).
// Here every box gathers its own entries ("pull" form), so each box's
// wavefront writes only its own rows and no atomics are needed:
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

// One wavefront per box, boxes taken in a grid-stride loop (kListWaves per
// block).  Lane l < 36 owns candidate (k, q) = (l / 4, l % 4): child slot q of
// colleague k.  Entries are appended in lane order through ballots, so every
// row is in the order of a serial sweep over (k, q); the M2L slot row is
// assembled in the wavefront's LDS and written once.
constexpr int kListWaves = 4;

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// appends the lanes with `pred` (in lane order) after `cnt` entries of `row`
__device__ __forceinline__ int append(bool pred, int val, int lane, int cnt, int cap, int* row) {
  const unsigned long long m = __ballot(pred);
  const int r = __popcll(m & ((1ull << lane) - 1ull));
  if (pred && cnt + r < cap) row[cnt + r] = val;
  return cnt + __popcll(m);
}

__global__ __launch_bounds__(kListWaves * 64) void k_lists(
    int nb, const int* ilevel, const int* iparent, const int* nchild, const int* ichild,
    const int* ncoll, const int* coll, const double* centers, const double* bsl, int* n1, int* l1,
    int* n2, int* l2, int* n3, int* l3, int* n4, int* l4, int* s2) {
  __shared__ int slots[kListWaves][kSlots];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int* ssh = slots[wv];
  const int k = lane >> 2, q = lane & 3;
  for (int b = blockIdx.x * kListWaves + wv; b < nb; b += gridDim.x * kListWaves) {
    const int lev = ilevel[b];
    const double bx = centers[2 * b], by = centers[2 * b + 1];
    int* r1 = l1 + (int64_t)kMaxList1 * b;
    int* r2 = l2 + (int64_t)kMaxList2 * b;
    int* r3 = l3 + (int64_t)kMaxList3 * b;
    int* r4 = l4 + (int64_t)kMaxList4 * b;
    int c1 = 0, c2 = 0, c3 = 0, c4 = 0;
    if (lane < kSlots) ssh[lane] = -1;
    wave_sync_lds();
    if (b == 0) {
      c1 = append(lane == 0 && nchild[0] == 0, 0, lane, c1, kMaxList1, r1);
    } else {
      const int dad = iparent[b];
      const double lim = kColl * bsl[lev];
      const double binv = 1.0 / bsl[lev];
      // child position of b (the first slot holding b; 3 if none of 0..2 does)
      const unsigned long long mq = __ballot(lane < 3 && ichild[4 * dad + lane] == b);
      const int qb = mq ? __ffsll((long long)mq) - 1 : 3;
      const int ncd = ncoll[dad];
      // list 2 from the children of the parent's colleagues, and their slots
      int kb = -1;
      if (lane < 4 * kMaxColl && k < ncd) {
        const int j = coll[9 * dad + k];
        if (q < nchild[j]) kb = ichild[4 * j + q];
      }
      bool far = false;
      if (kb >= 0) {
        const double dx = centers[2 * kb] - bx, dy = centers[2 * kb + 1] - by;
        far = fabs(dx) >= lim || fabs(dy) >= lim;
        if (far) {
          const int sl = m2l_slot(qb, (int)rint(dx * binv), (int)rint(dy * binv));
          if (sl >= 0) ssh[sl] = kb;
        }
      }
      c2 = append(far, kb, lane, c2, kMaxList2, r2);
      // entries contributed by coarser leaves (colleagues of the parent)
      if (lev >= 2) {
        const double dist = 1.05 * (bsl[lev - 1] + bsl[lev]) / 2.0;
        bool to1 = false, to4 = false;
        int j = -1;
        if (lane < ncd) {
          j = coll[9 * dad + lane];
          if (nchild[j] == 0) {
            const double xd = fabs(centers[2 * j] - bx), yd = fabs(centers[2 * j + 1] - by);
            to1 = xd < dist && yd < dist;
            to4 = !to1;
          }
        }
        c1 = append(to1, j, lane, c1, kMaxList1, r1);
        c4 = append(to4, j, lane, c4, kMaxList4, r4);
      }
      // own entries of a leaf
      if (nchild[b] == 0) {
        const double dist = 1.05 * (bsl[lev] + bsl[lev + 1]) / 2.0;
        const int ncb = ncoll[b];
        bool to1 = false, to3 = false;
        int e = -1;
        if (lane < 4 * kMaxColl && k < ncb) {
          const int j = coll[9 * b + k];
          const int nch = nchild[j];
          if (nch == 0) {
            if (q == 0) {
              e = j;
              to1 = true;
            }
          } else if (q < nch) {
            e = ichild[4 * j + q];
            const double xd = fabs(centers[2 * e] - bx), yd = fabs(centers[2 * e + 1] - by);
            to1 = xd < dist && yd < dist;
            to3 = !to1;
          }
        }
        c1 = append(to1, e, lane, c1, kMaxList1, r1);
        c3 = append(to3, e, lane, c3, kMaxList3, r3);
      }
    }
    wave_sync_lds();
    if (lane < kSlots) s2[(int64_t)kSlots * b + lane] = ssh[lane];
    if (lane == 0) {
      n1[b] = min(c1, kMaxList1);
      n2[b] = min(c2, kMaxList2);
      n3[b] = min(c3, kMaxList3);
      n4[b] = min(c4, kMaxList4);
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
  hip_throw(hipMemcpyAsync(dbs.p, bsl.data(), sizeof(double) * bsl.size(), hipMemcpyHostToDevice,
                           st), "copy");
  hip_throw(hipStreamSynchronize(st), "sync");  // bsl dies on return
  if (nb > 0) {
    const unsigned grid = (unsigned)std::min<int64_t>(8192, ceil_div(nb, kListWaves));
    hipLaunchKernelGGL(k_lists, dim3(grid), dim3(kListWaves * 64), 0, st, nb, t.ilevel.p,
                       t.iparent.p, t.nchild.p, t.ichild.p, t.ncoll.p, t.coll.p, t.centers.p,
                       dbs.p, L.n1.p, L.l1.p, L.n2.p, L.l2.p, L.n3.p, L.l3.p, L.n4.p, L.l4.p,
                       L.s2.p);
    hip_throw(hipGetLastError(), "kernel launch");
  }
}  // dbs is released stream-ordered, after k_lists

}  // namespace fmm2d

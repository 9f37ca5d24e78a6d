// Interaction lists 1-4 (computelists, common/tree_routs2d.f:486-710).
//
// The reference builds them in one serial sweep that also appends to the
// lists of OTHER boxes (list1 and list4 of the finer neighbours of a leaf).
// Here every box gathers its own entries ("pull" form), so one thread per
// box writes only its own rows and no atomics are needed:
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
#include "tree.h"

namespace fmm2d {
namespace {

constexpr double kColl = (double)1.05f;

__global__ void k_lists(int64_t n, const int* ilevel, const int* iparent, const int* nchild,
                        const int* ichild, const int* ncoll, const int* coll,
                        const double* centers, const double* bsl, int* n1, int* l1, int* n2,
                        int* l2, int* n3, int* l3, int* n4, int* l4, int* s2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = (int)i;
  const int lev = ilevel[b];
  const double bx = centers[2 * b], by = centers[2 * b + 1];
  int c1 = 0, c2 = 0, c3 = 0, c4 = 0;
  int* srow = s2 + (int64_t)kSlots * b;
  for (int k = 0; k < kSlots; ++k) srow[k] = -1;
  if (b == 0) {
    if (nchild[0] == 0) l1[c1++] = 0;
  } else {
    const int dad = iparent[b];
    const double lim = kColl * bsl[lev];
    const double binv = 1.0 / bsl[lev];
    int qb = 0;  // child position of b
    while (qb < 3 && ichild[4 * dad + qb] != b) ++qb;
    // list 2, and each entry's M2L offset slot
    for (int k = 0; k < ncoll[dad]; ++k) {
      const int j = coll[9 * dad + k];
      for (int q = 0; q < nchild[j]; ++q) {
        const int kb = ichild[4 * j + q];
        const double dx = centers[2 * kb] - bx, dy = centers[2 * kb + 1] - by;
        if (fabs(dx) >= lim || fabs(dy) >= lim) {
          if (c2 < kMaxList2) l2[(int64_t)kMaxList2 * b + c2] = kb;
          ++c2;
          const int sl = m2l_slot(qb, (int)rint(dx * binv), (int)rint(dy * binv));
          if (sl >= 0) srow[sl] = kb;
        }
      }
    }
    // entries contributed by coarser leaves (colleagues of the parent)
    if (lev >= 2) {
      const double dist = 1.05 * (bsl[lev - 1] + bsl[lev]) / 2.0;
      for (int k = 0; k < ncoll[dad]; ++k) {
        const int j = coll[9 * dad + k];
        if (nchild[j] != 0) continue;
        const double xd = fabs(centers[2 * j] - bx), yd = fabs(centers[2 * j + 1] - by);
        if (xd < dist && yd < dist) {
          if (c1 < kMaxList1) l1[(int64_t)kMaxList1 * b + c1] = j;
          ++c1;
        } else {
          if (c4 < kMaxList4) l4[(int64_t)kMaxList4 * b + c4] = j;
          ++c4;
        }
      }
    }
    // own entries of a leaf
    if (nchild[b] == 0) {
      const double dist = 1.05 * (bsl[lev] + bsl[lev + 1]) / 2.0;
      for (int k = 0; k < ncoll[b]; ++k) {
        const int j = coll[9 * b + k];
        if (nchild[j] == 0) {
          if (c1 < kMaxList1) l1[(int64_t)kMaxList1 * b + c1] = j;
          ++c1;
        } else {
          for (int q = 0; q < nchild[j]; ++q) {
            const int kb = ichild[4 * j + q];
            const double xd = fabs(centers[2 * kb] - bx), yd = fabs(centers[2 * kb + 1] - by);
            if (xd < dist && yd < dist) {
              if (c1 < kMaxList1) l1[(int64_t)kMaxList1 * b + c1] = kb;
              ++c1;
            } else {
              if (c3 < kMaxList3) l3[(int64_t)kMaxList3 * b + c3] = kb;
              ++c3;
            }
          }
        }
      }
    }
  }
  n1[b] = min(c1, kMaxList1);
  n2[b] = min(c2, kMaxList2);
  n3[b] = min(c3, kMaxList3);
  n4[b] = min(c4, kMaxList4);
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
  launch1d(k_lists, nb, st, t.ilevel.p, t.iparent.p, t.nchild.p, t.ichild.p, t.ncoll.p, t.coll.p,
           t.centers.p, dbs.p, L.n1.p, L.l1.p, L.n2.p, L.l2.p, L.n3.p, L.l3.p, L.n4.p, L.l4.p,
           L.s2.p);
}  // dbs is released stream-ordered, after k_lists

}  // namespace fmm2d

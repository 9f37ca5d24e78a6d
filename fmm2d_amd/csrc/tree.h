// Adaptive quadtree on the device, numbered exactly as the reference numbers
// it, plus the interaction lists.
#pragma once
#include <vector>

#include "devbuf.h"

namespace fmm2d {

constexpr int kMaxColl = 9;    // colleagues incl. self
constexpr int kMaxList1 = 13;  // computemnlists, common/tree_routs2d.f:459-479
constexpr int kMaxList2 = 27;
constexpr int kMaxList3 = 20;
constexpr int kMaxList4 = 5;

struct Tree {
  int nlevels = 0;
  int nboxes = 0;
  int64_t ns = 0, nt = 0;
  std::vector<int> lstart, lend;  // level l boxes are [lstart[l], lend[l]) (0-based)
  std::vector<double> boxsize;    // boxsize[l], l = 0..nlevels
  // device, indexed by 0-based box id; "none" is -1
  DBuf<double> centers;  // 2 per box
  DBuf<int> ilevel, iparent, nchild, ichild /*4*/, ncoll, coll /*9*/;
  DBuf<int> isrcse, itargse;  // [start, end) per box, 2 per box
  DBuf<int> isrc, itarg;      // sorted position -> original point index
};

// M2L offset slots.  A box in child position q of its parent (0..3 for
// (-,-), (+,-), (-,+), (+,+)) can only have list-2 sources at the offsets
// (dx, dy) (in box sizes) of the 6x6 children of its parent's 3x3
// neighbourhood, minus the 3x3 adjacent ones: 27 slots per q.
constexpr int kSlots = 27;

// slot of offset (dx,dy) for child position q, or -1 (adjacent / impossible)
FMM2D_HD int m2l_slot(int q, int dx, int dy) {
  const int bx = (q & 1) ? -3 : -2, by = (q & 2) ? -3 : -2;
  const int ix = dx - bx, iy = dy - by;
  if (ix < 0 || ix > 5 || iy < 0 || iy > 5) return -1;
  if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1) return -1;
  int s = 0;
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      const int ex = bx + a, ey = by + b;
      if (ex >= -1 && ex <= 1 && ey >= -1 && ey <= 1) continue;
      if (a == ix && b == iy) return s;
      ++s;
    }
  return -1;
}

struct Lists {
  DBuf<int> n1, l1, n2, l2, n3, l3, n4, l4;  // counts and [box][cap] tables
  DBuf<int> s2;  // [box][27]: list-2 source of each M2L offset slot, or -1
};

// The part of the tree build that does not depend on the leaf capacity: the
// bounding square and the points' sorted path keys (opaque; tree.hip).
struct TreePrep;
TreePrep* tree_prepare(const double* src, int64_t ns, const double* trg, int64_t nt, int nlmax,
                       hipStream_t st);
void tree_prepare_free(TreePrep* p);
// true when p was prepared for exactly these arrays and nlmax
bool tree_prepare_matches(const TreePrep* p, const double* src, int64_t ns, const double* trg,
                          int64_t nt, int nlmax);

// Build the tree of the reference's pts_tree_build + pts_tree_sort
// (common/pts_tree2d.f:406-655, 1222-1246) over the given points.
// src/trg: (2,n) interleaved device arrays.  prep: a preparation of the same
// points (used when tree_prepare_matches, else ignored), or null.  Throws
// HipError on HIP failure.
void build_tree(const double* src, int64_t ns, const double* trg, int64_t nt, int ndiv,
                int idivflag, int nlmax, Tree& t, hipStream_t st,
                const TreePrep* prep = nullptr);

// Lists 1-4 of computelists (common/tree_routs2d.f:486-710), as sets.
void build_lists(const Tree& t, Lists& l, hipStream_t st);

// hist[l] (host, nbins <= kLeafHistMax values) = points, sources plus targets,
// held by the leaves of level l; the last bin also takes every deeper level.
constexpr int kLeafHistMax = 64;
void leaf_level_histogram(const Tree& t, int nbins, double* hist, hipStream_t st);

// hist[l] as above for the tree refined while a box holds more than ndiv
// sources, without level restriction, computed from the sorted path keys
// only (no tree is built); *boxsize0 = the root box size.
void key_leaf_histogram(const double* src, int64_t ns, const double* trg, int64_t nt, int ndiv,
                        int nbins, double* hist, double* boxsize0, hipStream_t st);
// The same from a preparation's keys (min(20, nlmax) levels: equal to
// key_leaf_histogram when nbins - 1 <= that depth).
void prepared_leaf_histogram(const TreePrep* p, int ndiv, int nbins, double* hist,
                             double* boxsize0, hipStream_t st);

}  // namespace fmm2d

// Expansion-to-expansion translations (M2M, M2L, L2L) on MFMA, shared by the
// Laplace/Cauchy and Helmholtz drivers.
//
// A translation is y = T x with T a complex Pout x Pin operator and x, y box
// expansions.  In real form T becomes a 2Pout x 2Pin matrix over interleaved
// (re, im) coefficients; operators are pre-swizzled into the A-operand layout
// of v_mfma_f64_16x16x4f64 (lane l holds row l%16, column l/16 of a 16x4
// block; the result tile holds row l/16 + 4 i, column l%16 in element i).
//
// Slot s of target t reads source slots[row * nslots + s] (-1: empty), row =
// t's position in tlist, or its box id (slots_by_box).
// Work is split into groups: group g owns targets tlist[seg0[g]..seg1[g])
// and tiles [tile0[g], tile0[g+1]) of 16 columns; within a group every target has the same
// slot -> operator map (opbase[g] + slot_op[cls[g]*nslots + s]), so the A
// operand is uniform across a wavefront, whose 16 columns are (target,
// density) pairs.  MODE 0 gathers: dst[t] += sum_s T_s src[slot(t,s)]; MODE 1
// scatters: dst[slot(t,s)] += T_s src[t].  Box b's expansion for density d
// starts at ((b - box0) nd + d) P of its buffer (box0 = first box of the
// buffer's level, 0 for a buffer holding all levels).  With a band table,
// band[level] = [lo, hi) are the interleaved coefficient rows that level uses
// (the rest stay zero); without one every row is used.  With LOGT the Laplace
// M2L term log(boxsize) M_0 is added to L_0.
#pragma once
#include <vector>

#include "common.h"
#include "devbuf.h"

namespace fmm2d {

typedef double d4v __attribute__((ext_vector_type(4)));

// Group table of one launch, passed by value: group g owns targets
// tlist[seg0[g]..seg1[g]) and tiles [tile0[g], tile0[g+1]).
constexpr int kTrMaxGroups = 8;
// Groups of one launch.  Wave tiles: the first ilv tiles of every group go
// round-robin over the groups (tile t of group g is wave t * ng + g), the
// rest group after group from ng * ilv on (group g's from rem0[g]); ilv = 0
// is group after group throughout.  Interleaving the M2L classes (the four
// child positions) puts siblings, whose list-2 sources are mostly the same
// boxes, into neighbouring waves of one workgroup.
struct TrTable {
  int ng, ilv;
  int tile0[kTrMaxGroups + 1];
  int rem0[kTrMaxGroups + 1];
  int seg0[kTrMaxGroups], seg1[kTrMaxGroups];
  int opbase[kTrMaxGroups], cls[kTrMaxGroups];
};

// RT row tiles are accumulated per pass over the k steps; with KSU > 0 the k
// loop is unrolled to KSU steps (operators of at most 16 RT rows and 4 KSU
// columns, the whole operator in one pass), with KSU == 0 it is a runtime
// loop and the row tiles are covered in chunks of RT.
typedef float f4v __attribute__((ext_vector_type(4)));

// One 16x16x4 MFMA step in the element type T (double: v_mfma_f64_16x16x4f64,
// float: v_mfma_f32_16x16x4f32), and the output row of accumulator element i
// for lane group kr: the two instructions lay their 16x16 results out
// differently (f64: row kr + 4 i, f32: row 4 kr + i).
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef d4v acc_t;
  static __device__ __forceinline__ acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int kr, int i) { return kr + 4 * i; }
};
template <> struct Mfma<float> {
  typedef f4v acc_t;
  static __device__ __forceinline__ acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int kr, int i) { return 4 * kr + i; }
};

// orders a wavefront's LDS writes before its later LDS reads and vice versa
// (the wavefront's LDS accesses complete in order; this keeps the compiler
// from moving them across)
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// T is the arithmetic type of the operator (Afrag) and the source
// expansions (src, 2 Pin values of T per box and density); the results are
// added to the FP64 destination expansions.
template <typename T, int RT, int KSU, int MODE, int LOGT>
__global__ __launch_bounds__(256) void k_translate(
    int nd, int Pin, int Pout, int nslots, int64_t nwaves, int per_xcd, TrTable tab,
    const int* tlist, const int* slots, int slots_by_box, const int* slot_op, const T* Afrag,
    int64_t afs,
    int sbox0, int dbox0, const int* ilevel, const int2* band, const double* logr, const T* src,
    cplx* dst, int ssplit, int ksplit, double* part, int64_t* dstoff) {
  typedef typename Mfma<T>::acc_t acc_t;
  // Workgroups are dispatched round-robin over the 8 XCDs; block b works on
  // logical block (b % 8) * per_xcd + b / 8 so that each XCD walks a
  // contiguous range of targets, whose neighbouring source boxes then stay
  // in that XCD's L2.
  const int64_t lblk = (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  const int64_t wv0 = lblk * 4 + (threadIdx.x >> 6);
  if (wv0 >= nwaves) return;
  const int lane = threadIdx.x & 63;
  // Split launches (few targets, long operators): wavefront wv0 also takes
  // one of ssplit groups of slots and one of ksplit ranges of k steps.  MODE
  // 0 (every slot adds to the column's own target) writes the partial sums of
  // group sk to part[sk][column][row] (all 16 RTT rows, 0 outside the band)
  // and the target's offset in dst to dstoff[column], for k_tr_split_sum to
  // add in group order.  MODE 1 (each slot's target is written by one slot of
  // one column): slot groups add directly; with k ranges split, the partial
  // sums go to part[kg][column][slot][row] and dstoff[column][slot], for
  // k_tr_split_sum to add in k-range order.
  const int nsk = ssplit * ksplit;
  const int sk = (int)(wv0 % nsk);
  const int64_t wv = wv0 / nsk;
  const bool split = nsk > 1;
  // With the k loop not unrolled (long operators), the chunks of RT row tiles
  // go to separate wavefronts: wavefront wv works on tile wv / nrc, chunk
  // wv % nrc (the rows are disjoint outputs).
  const int RTT0 = (2 * Pout + 15) / 16;
  const int nrc = KSU > 0 ? 1 : (RTT0 + RT - 1) / RT;
  const int64_t wave = wv / nrc;
  const int rc0 = (int)(wv - wave * nrc) * RT;
  int g = 0;
  int64_t tl;  // tile within group g
  if (wave < (int64_t)tab.ilv * tab.ng) {
    g = (int)(wave % tab.ng);
    tl = wave / tab.ng;
  } else {
    const int64_t w2 = wave - (int64_t)tab.ilv * tab.ng;
    while (g + 1 < tab.ng && w2 >= tab.rem0[g + 1]) ++g;
    tl = tab.ilv + (w2 - tab.rem0[g]);
  }
  const int q = tab.cls[g];
  const int64_t col = tl * 16 + (lane & 15);
  const int64_t tpos = tab.seg0[g] + col / nd;
  const int d = (int)(col % nd);
  // a target list entry may be -1 (no target at that position)
  const int tb = tpos < tab.seg1[g] ? tlist[tpos] : -1;
  const bool valid = tb >= 0;
  // slot row of this column's target: by position in tlist, or by box id
  const int64_t srow = slots_by_box ? (int64_t)tb : tpos;
  const int kr = lane >> 4;
  const int RTT = (2 * Pout + 15) / 16, KS = (2 * Pin + 3) / 4;
  const int2 full_in = int2{0, 2 * Pin}, full_out = int2{0, 2 * Pout};
  const int ob = tab.opbase[g];
  // FP32 gather mode with 8 k steps (rows of 32 coefficients): per slot the
  // wavefront first stages the 16 source rows of its columns in LDS
  // (consecutive lanes load consecutive coefficients of a row, two rows per
  // load instead of sixteen), then reads the B operands from there.  Measured
  // slower, and so not used, for 12 k steps (rows straddle the loads) and for
  // FP64 (twice the LDS, lower occupancy).
  constexpr bool STAGE = MODE == 0 && KSU == 8 && sizeof(T) == 4;
  constexpr int SW = STAGE ? 4 * KSU : 1;  // staged coefficients per row
  __shared__ T sbuf[STAGE ? 4 : 1][STAGE ? 16 * SW : 1];
  __shared__ int4 sinfo[STAGE ? 4 : 1][16];
  const int wib = STAGE ? (int)(threadIdx.x >> 6) : 0;
  const int sg = sk / ksplit, kg = sk - sg * ksplit;
  const int s_lo = (int)((int64_t)nslots * sg / ssplit);
  const int s_hi = (int)((int64_t)nslots * (sg + 1) / ssplit);
  const int k_lo = (int)((int64_t)KS * kg / ksplit);
  const int k_hi = (int)((int64_t)KS * (kg + 1) / ksplit);
  auto put = [&](double* a, double v) { *a += v; };
  for (int rc = rc0; rc < (KSU > 0 ? RTT : rc0 + 1); rc += RT) {
    if (MODE == 0) {
      acc_t acc[RT];
#pragma unroll
      for (int r = 0; r < RT; ++r) acc[r] = acc_t{0, 0, 0, 0};
      T m0 = 0;
      for (int s = s_lo; s < s_hi; ++s) {
        const int c = valid ? slots[srow * nslots + s] : -1;
        const T* A = Afrag + (int64_t)(ob + slot_op[q * nslots + s]) * afs;
        const T* M = nullptr;
        int klo = 0, khi = 0;
        if (c >= 0) {
          M = src + ((int64_t)(c - sbox0) * nd + d) * 2 * Pin;
          const int2 bd = band ? band[ilevel[c]] : full_in;
          klo = bd.x;
          khi = bd.y;
        }
        if (STAGE) {
          if (kr == 0) sinfo[wib][lane] = int4{c, klo, c >= 0 ? khi : 0, d};
          wave_lds_fence();
#pragma unroll
          for (int e0 = 0; e0 < 16 * SW; e0 += 64) {
            const int e = e0 + lane, row = e / SW, k = e - row * SW;
            const int4 in = sinfo[wib][row];
            T v = T(0);
            if (k >= in.y && k < in.z)
              v = src[((int64_t)(in.x - sbox0) * nd + in.w) * 2 * Pin + k];
            sbuf[wib][e] = v;
          }
          wave_lds_fence();
        }
        auto step = [&](int ks) {
          const int k = 4 * ks + kr;
          const T b = STAGE ? sbuf[wib][(lane & 15) * SW + k]
                            : ((k >= klo && k < khi) ? M[k] : T(0));
          if (LOGT && rc == 0 && ks == 0 && lane < 32) m0 += b;
#pragma unroll
          for (int r = 0; r < RT; ++r)
            if (RT == 1 || KSU > 0 || rc + r < RTT)
              acc[r] = Mfma<T>::step(A[((int64_t)(rc + r) * KS + ks) * 64 + lane], b, acc[r]);
        };
        if (KSU > 0) {
#pragma unroll
          for (int ks = 0; ks < KSU; ++ks)
            if (ks < KS) step(ks);
        } else {
          for (int ks = k_lo; ks < k_hi; ++ks) step(ks);
        }
        if (STAGE) wave_lds_fence();  // the next slot overwrites the rows
      }
      if (split) {
        const int64_t ncols = (int64_t)tab.tile0[tab.ng] * 16;
        const int64_t colg = (int64_t)(tab.tile0[g] + tl) * 16 + (lane & 15);
        double* Pp = part + ((int64_t)sk * ncols + colg) * (16 * RTT);
        if (sk == 0 && rc == 0 && kr == 0)
          dstoff[colg] = valid ? ((int64_t)(tb - dbox0) * nd + d) * 2 * Pout : -1;
        const int2 bd = valid ? (band ? band[ilevel[tb]] : full_out) : int2{0, 0};
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = 16 * (rc + r) + Mfma<T>::row(kr, i);
            if (rc + r < RTT) Pp[row] = (row >= bd.x && row < bd.y) ? (double)acc[r][i] : 0.0;
          }
      } else if (valid) {
        const int2 bd = band ? band[ilevel[tb]] : full_out;
        double* L = reinterpret_cast<double*>(dst + ((int64_t)(tb - dbox0) * nd + d) * Pout);
        // level-dependent part of L_0: lanes 0..15 hold Re M_0 (row 0), lanes
        // 16..31 Im M_0 (row 1) of their column
        if (LOGT && rc == 0 && lane < 32) L[kr] += logr[ilevel[tb]] * (double)m0;
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int row = 16 * (rc + r) + Mfma<T>::row(kr, i);
            if (rc + r < RTT && row >= bd.x && row < bd.y) put(&L[row], (double)acc[r][i]);
          }
      }
    } else {
      const T* M = valid ? src + ((int64_t)(tb - sbox0) * nd + d) * 2 * Pin : nullptr;
      int klo = 0, khi = 0;
      if (valid) {
        const int2 bd = band ? band[ilevel[tb]] : full_in;
        klo = bd.x;
        khi = bd.y;
      }
      // the source coefficients are reused by every slot: keep them in
      // registers when the k loop is unrolled
      T bv[KSU > 0 ? KSU : 1];
      if (KSU > 0) {
#pragma unroll
        for (int ks = 0; ks < (KSU > 0 ? KSU : 1); ++ks) {
          const int k = 4 * ks + kr;
          bv[ks] = (ks < KS && k >= klo && k < khi) ? M[k] : T(0);
        }
      }
      for (int s = s_lo; s < s_hi; ++s) {
        const int c = valid ? slots[srow * nslots + s] : -1;
        const T* A = Afrag + (int64_t)(ob + slot_op[q * nslots + s]) * afs;
        acc_t acc[RT];
#pragma unroll
        for (int r = 0; r < RT; ++r) acc[r] = acc_t{0, 0, 0, 0};
        auto step = [&](int ks, T b) {
#pragma unroll
          for (int r = 0; r < RT; ++r)
            if (RT == 1 || KSU > 0 || rc + r < RTT)
              acc[r] = Mfma<T>::step(A[((int64_t)(rc + r) * KS + ks) * 64 + lane], b, acc[r]);
        };
        if (KSU > 0) {
#pragma unroll
          for (int ks = 0; ks < (KSU > 0 ? KSU : 1); ++ks)
            if (ks < KS) step(ks, bv[ks]);
        } else {
          for (int ks = k_lo; ks < k_hi; ++ks) {
            const int k = 4 * ks + kr;
            step(ks, (k >= klo && k < khi) ? M[k] : T(0));
          }
        }
        if (ksplit > 1) {
          // partial over k range kg of (column, slot): part[kg][column][slot][row]
          const int64_t ncs = (int64_t)tab.tile0[tab.ng] * 16 * nslots;
          const int64_t cs = ((int64_t)(tab.tile0[g] + tl) * 16 + (lane & 15)) * nslots + s;
          double* Pp = part + ((int64_t)kg * ncs + cs) * (16 * RTT);
          if (kg == 0 && rc == 0 && kr == 0)
            dstoff[cs] = c >= 0 ? ((int64_t)(c - dbox0) * nd + d) * 2 * Pout : -1;
          const int2 bd = c >= 0 ? (band ? band[ilevel[c]] : full_out) : int2{0, 0};
#pragma unroll
          for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int row = 16 * (rc + r) + Mfma<T>::row(kr, i);
              if (rc + r < RTT) Pp[row] = (row >= bd.x && row < bd.y) ? (double)acc[r][i] : 0.0;
            }
        } else if (c >= 0) {
          const int2 bd = band ? band[ilevel[c]] : full_out;
          double* L = reinterpret_cast<double*>(dst + ((int64_t)(c - dbox0) * nd + d) * Pout);
#pragma unroll
          for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int row = 16 * (rc + r) + Mfma<T>::row(kr, i);
              if (rc + r < RTT && row >= bd.x && row < bd.y) put(&L[row], (double)acc[r][i]);
            }
        }
      }
    }
  }
}

// Real interleaved form of square complex operators ops[o][j][n] (input j,
// output n), swizzled as Afrag[o][row tile][k step][lane].
template <typename T = double>
inline std::vector<T> swizzle_ops(int P, const std::vector<cplx>& ops, int nops) {
  const int K2 = 2 * P, RTT = (K2 + 15) / 16, KS = (K2 + 3) / 4;
  std::vector<T> out((size_t)nops * RTT * KS * 64, T(0));
  for (int o = 0; o < nops; ++o)
    for (int r = 0; r < RTT; ++r)
      for (int ks = 0; ks < KS; ++ks)
        for (int l = 0; l < 64; ++l) {
          const int row = 16 * r + (l & 15), col = 4 * ks + (l >> 4);
          double v = 0.0;
          if (row < K2 && col < K2) {
            const cplx t = ops[((size_t)o * P + col / 2) * P + row / 2];
            const int a = row & 1, b = col & 1;
            v = (a == b) ? t.re : (a == 0 ? -t.im : t.im);
          }
          out[(((size_t)o * RTT + r) * KS + ks) * 64 + l] = (T)v;
        }
  return out;
}

// doubles per swizzled operator of Pout x Pin complex entries
inline int64_t afrag_stride(int Pin, int Pout) {
  return (int64_t)((2 * Pout + 15) / 16) * ((2 * Pin + 3) / 4) * 64;
}
inline int64_t afrag_stride(int P) { return afrag_stride(P, P); }

struct TrGroup {
  int seg0, seg1;  // targets tlist[seg0..seg1)
  int opbase, cls;
};

// dst[dstoff[c] + row] += sum over the nsk split groups, in group order, of
// part[k][c][row] (rows 0 .. rows-1 of the ncols columns; dstoff -1: none)
static __global__ void k_tr_split_sum(int64_t n, int64_t ncols, int rstride, int rows, int nsk,
                               const double* __restrict__ part, const int64_t* __restrict__ dstoff,
                               double* dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t c = t / rows;
  const int row = (int)(t - c * rows);
  const int64_t o = dstoff[c];
  if (o < 0) return;
  double s = 0.0;
  for (int k = 0; k < nsk; ++k) s += part[((int64_t)k * ncols + c) * rstride + row];
  dst[o + row] += s;
}

// Buffers of one translate launch.  src holds 2 Pin values per (box,
// density) of the launch's element type (double: the cplx expansions).
struct TrBufs {
  const void* src;
  int sbox0, Pin;
  cplx* dst;
  int dbox0, Pout;
};

// Launch translate passes over the given groups (at most kTrMaxGroups per
// launch; launches are stream-ordered, so groups must not depend on each other).
// fp32: operators (Afrag) and sources are float, accumulation FP32 on
// v_mfma_f32_16x16x4f32 (gather mode only); otherwise FP64.
// Launches with fewer wavefronts than this split their slots (and, for long
// operators whose k loop is not unrolled, their k steps) over more
// wavefronts, whose partial sums are added in a fixed order (no atomics: the
// results do not depend on scheduling): small problems and coarse levels
// would otherwise run a few dozen wavefronts through 27 slots each.
constexpr int64_t kTrSplitWaves = 4096;

inline void launch_translate(int mode, int logt, int nd, const TrBufs& io, int nslots,
                             const std::vector<TrGroup>& groups, const int* tlist,
                             const int* slots, const int* slot_op, const void* Afrag,
                             const int* ilevel, const int2* band, const double* logr,
                             hipStream_t st, bool fp32 = false, bool slots_by_box = false,
                             bool interleave = false) {
  const int Pin = io.Pin, Pout = io.Pout;
  const int64_t afs = afrag_stride(Pin, Pout);
  const int RTT = (2 * Pout + 15) / 16, KS = (2 * Pin + 3) / 4;
  if (fp32 && mode != 0) throw HipError{hipErrorInvalidValue, "translate fp32 scatter"};
  size_t gi = 0;
  while (gi < groups.size()) {
    TrTable tab{};
    int tiles = 0;
    for (; gi < groups.size() && tab.ng < kTrMaxGroups; ++gi) {
      const TrGroup& g = groups[gi];
      if (g.seg1 <= g.seg0) continue;
      tab.tile0[tab.ng] = tiles;
      tab.seg0[tab.ng] = g.seg0;
      tab.seg1[tab.ng] = g.seg1;
      tab.opbase[tab.ng] = g.opbase;
      tab.cls[tab.ng] = g.cls;
      ++tab.ng;
      tiles += (int)(((int64_t)(g.seg1 - g.seg0) * nd + 15) / 16);
    }
    if (tab.ng == 0) continue;
    tab.tile0[tab.ng] = tiles;
    tab.ilv = 0;
    if (interleave) {
      tab.ilv = tiles;
      for (int g = 0; g < tab.ng; ++g) tab.ilv = std::min(tab.ilv, tab.tile0[g + 1] - tab.tile0[g]);
    }
    tab.rem0[0] = 0;
    for (int g = 0; g < tab.ng; ++g)
      tab.rem0[g + 1] = tab.rem0[g] + (tab.tile0[g + 1] - tab.tile0[g]) - tab.ilv;
#define FMM2D_TR(T, RT, KSU, MODE, LOGT)                                                        \
  {                                                                                             \
    const int64_t nw0 = (int64_t)tiles * (KSU > 0 ? 1 : (RTT + RT - 1) / RT);                   \
    int ssplit = 1, ksplit = 1;                                                                 \
    if (!LOGT && nw0 < kTrSplitWaves) {                                                         \
      ssplit = (int)std::min<int64_t>(nslots, (kTrSplitWaves + nw0 - 1) / nw0);                 \
      if (KSU == 0)                                                                             \
        ksplit = (int)std::max<int64_t>(1, std::min<int64_t>(KS / 16, (kTrSplitWaves + nw0 *   \
                                          ssplit - 1) / (nw0 * ssplit)));                      \
    }                                                                                           \
    const int64_t nw = nw0 * ssplit * ksplit;                                                   \
    const int64_t nblk = (nw + 3) / 4;                                                          \
    const int per_xcd = (int)((nblk + 7) / 8);                                                  \
    const int nsk = ssplit * ksplit;                                                            \
    /* partial sums: MODE 0 per (split group, column), MODE 1 per (k range, column, slot) */    \
    const int64_t ncols = (int64_t)tiles * 16 * (MODE == 0 ? 1 : nslots);                       \
    const int npart = MODE == 0 ? nsk : ksplit;                                                 \
    const bool gsplit = npart > 1;                                                              \
    DBuf<double> part(gsplit ? (size_t)npart * ncols * 16 * RTT : 0, st);                       \
    DBuf<int64_t> dstoff(gsplit ? (size_t)ncols : 0, st);                                       \
    hipLaunchKernelGGL((k_translate<T, RT, KSU, MODE, LOGT>), dim3((unsigned)(8 * per_xcd)),    \
                       dim3(256), 0, st, nd, Pin, Pout, nslots, nw, per_xcd, tab, tlist, slots, \
                       (int)slots_by_box, slot_op, static_cast<const T*>(Afrag), afs, io.sbox0, \
                       io.dbox0, ilevel, band, logr, static_cast<const T*>(io.src), io.dst,     \
                       ssplit, ksplit, part.p, dstoff.p);                                       \
    if (gsplit)                                                                                 \
      launch1d(k_tr_split_sum, ncols * 2 * Pout, st, ncols, 16 * RTT, 2 * Pout, npart,          \
               (const double*)part.p, (const int64_t*)dstoff.p,                                 \
               reinterpret_cast<double*>(io.dst));                                              \
  }
#define FMM2D_TR_RT(T, MODE, LOGT)                 \
  switch (KS <= 4 * RTT ? RTT : 0) {               \
    case 1: FMM2D_TR(T, 1, 4, MODE, LOGT); break;  \
    case 2: FMM2D_TR(T, 2, 8, MODE, LOGT); break;  \
    case 3: FMM2D_TR(T, 3, 12, MODE, LOGT); break; \
    case 4: FMM2D_TR(T, 4, 16, MODE, LOGT); break; \
    default: FMM2D_TR(T, 4, 0, MODE, LOGT); break; \
  }
// FP64 also unrolls operators of 5..8 row tiles (expansions of up to 64
// complex coefficients: Helmholtz orders up to 31)
#define FMM2D_TR_RT64(MODE, LOGT)                            \
  switch (KS <= 4 * RTT ? RTT : 0) {                         \
    case 1: FMM2D_TR(double, 1, 4, MODE, LOGT); break;       \
    case 2: FMM2D_TR(double, 2, 8, MODE, LOGT); break;       \
    case 3: FMM2D_TR(double, 3, 12, MODE, LOGT); break;      \
    case 4: FMM2D_TR(double, 4, 16, MODE, LOGT); break;      \
    case 5: FMM2D_TR(double, 5, 20, MODE, LOGT); break;      \
    case 6: FMM2D_TR(double, 6, 24, MODE, LOGT); break;      \
    case 7: FMM2D_TR(double, 7, 28, MODE, LOGT); break;      \
    case 8: FMM2D_TR(double, 8, 32, MODE, LOGT); break;      \
    default: FMM2D_TR(double, 4, 0, MODE, LOGT); break;      \
  }
    if (fp32) {
      if (logt) {
        FMM2D_TR_RT(float, 0, 1)
      } else {
        FMM2D_TR_RT(float, 0, 0)
      }
    } else if (mode == 0 && logt) {
      FMM2D_TR_RT(double, 0, 1)
    } else if (mode == 0) {
      FMM2D_TR_RT64(0, 0)
    } else {
      FMM2D_TR_RT64(1, 0)
    }
#undef FMM2D_TR_RT64
#undef FMM2D_TR_RT
#undef FMM2D_TR
    hip_throw(hipGetLastError(), "translate");
  }
}

}  // namespace fmm2d

// Shared types and helpers for the fmm2d_amd HIP library (gfx950).
//
// Complex numbers are a plain {re, im} pair of doubles: layout-compatible with
// C99 `double complex`, numpy complex128 and Fortran complex*16, which is what
// the C-ABI hands us.  All arithmetic is written out so the compiler sees
// ordinary FP64 operations (no library complex type on the device).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <new>
#include <cmath>

#define FMM2D_HD __host__ __device__ __forceinline__

struct cplx {
  double re, im;
};

FMM2D_HD cplx mkc(double r, double i = 0.0) { return cplx{r, i}; }
FMM2D_HD cplx operator+(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
FMM2D_HD cplx operator-(cplx a, cplx b) { return cplx{a.re - b.re, a.im - b.im}; }
FMM2D_HD cplx operator-(cplx a) { return cplx{-a.re, -a.im}; }
FMM2D_HD cplx operator*(cplx a, cplx b) {
  return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
FMM2D_HD cplx operator*(double s, cplx a) { return cplx{s * a.re, s * a.im}; }
FMM2D_HD cplx operator*(cplx a, double s) { return cplx{s * a.re, s * a.im}; }
FMM2D_HD cplx& operator+=(cplx& a, cplx b) { a.re += b.re; a.im += b.im; return a; }
FMM2D_HD cplx& operator-=(cplx& a, cplx b) { a.re -= b.re; a.im -= b.im; return a; }
FMM2D_HD cplx conjc(cplx a) { return cplx{a.re, -a.im}; }
FMM2D_HD double abs2(cplx a) { return a.re * a.re + a.im * a.im; }
FMM2D_HD double absc(cplx a) { return hypot(a.re, a.im); }
// a / b, scaled to avoid overflow for large |b| (Smith's algorithm).
FMM2D_HD cplx divc(cplx a, cplx b) {
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return cplx{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  } else {
    double r = b.re / b.im, d = b.re * r + b.im;
    return cplx{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
}
FMM2D_HD cplx operator/(cplx a, cplx b) { return divc(a, b); }
FMM2D_HD cplx operator/(cplx a, double s) { return cplx{a.re / s, a.im / s}; }
FMM2D_HD cplx invc(cplx b) { return divc(cplx{1.0, 0.0}, b); }
FMM2D_HD cplx expc(cplx a) {
  double e = exp(a.re), s, c;
  sincos(a.im, &s, &c);
  return cplx{e * c, e * s};
}
// principal log
FMM2D_HD cplx logc(cplx a) { return cplx{log(absc(a)), atan2(a.im, a.re)}; }
// principal square root (branch cut on the negative real axis)
FMM2D_HD cplx sqrtc(cplx a) {
  double m = absc(a);
  if (m == 0.0) return cplx{0.0, 0.0};
  double t = sqrt(0.5 * (m + fabs(a.re)));
  if (a.re >= 0.0) return cplx{t, 0.5 * a.im / t};
  return cplx{0.5 * fabs(a.im) / t, copysign(t, a.im)};
}

// ---- error codes shared by the C-ABI -------------------------------------
// 0 = ok, 4 = workspace allocation failure (the reference's ier=4),
// 13 = expansion-order search exhausted (the reference's ier=13),
// negative = errors the reference cannot produce (bad argument, HIP failure).
enum {
  FMM2D_OK = 0,
  FMM2D_ERR_ALLOC = 4,
  FMM2D_ERR_NTERMS = 13,
  FMM2D_ERR_ARG = -1,
  FMM2D_ERR_HIP = -2,
  FMM2D_ERR_UNSUPPORTED = -3,
};

// Bounds of the per-level expansion orders a caller may pass in
// fmm2d_opts.nterms (the reference's rules give at most ~50 for Laplace at
// eps = 1e-16; Helmholtz orders grow with zk times the box size).
constexpr int kMaxNtermsLaplace = 200;
constexpr int kMaxNtermsHelmholtz = 2000;

// Checks fmm2d_opts.nterms before any device work: FMM2D_ERR_ARG for a
// negative count or a missing array, FMM2D_ERR_NTERMS for an order outside
// [lo, hi], else FMM2D_OK.
static inline int check_nterms_override(int n, const int32_t* v, int lo, int hi) {
  if (n < 0 || (n > 0 && !v)) return FMM2D_ERR_ARG;
  for (int i = 0; i < n; ++i)
    if (v[i] < lo || v[i] > hi) return FMM2D_ERR_NTERMS;
  return FMM2D_OK;
}

#define FMM2D_HIP_TRY(expr)                                                   \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      fprintf(stderr, "fmm2d_amd: HIP error %s at %s:%d: %s\n",               \
              hipGetErrorString(_e), __FILE__, __LINE__, #expr);              \
      return _e == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;     \
    }                                                                         \
  } while (0)

static inline unsigned ceil_div(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

#ifdef __HIPCC__
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  With gridDim.x a multiple of 8 this maps block b to logical block
// (b % 8) * (gridDim.x / 8) + b / 8, so each XCD walks one contiguous range
// of the work and neighbouring items share its L2.  Grids of fewer than 8
// blocks keep their order.
__device__ __forceinline__ int64_t xcd_block() {
  const unsigned g8 = gridDim.x >> 3;
  return g8 ? (int64_t)(blockIdx.x & 7) * g8 + (blockIdx.x >> 3) : (int64_t)blockIdx.x;
}
#endif

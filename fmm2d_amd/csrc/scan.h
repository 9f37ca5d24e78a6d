// Device-wide exclusive prefix sums (the role of common/cumsum.f:39 in the
// reference's tree build, and of the offset scans of the drivers).
//
// Reduce-then-scan in three launches over tiles of kScanTile elements: each
// workgroup sums its tile (coalesced, kScanItems strided loads per thread), one
// workgroup scans the tile sums, and each workgroup scans its tile again in
// kScanItems rounds of 256 consecutive elements (wave64 shuffle scans plus a
// 4-wave carry in LDS), starting from its tile's offset.  Deterministic (fixed
// summation order), in place allowed (out == in).  Traffic: 2 reads + 1 write
// of the array; the tile sums are n / 4096 elements.
#pragma once
#include "common.h"
#include "devbuf.h"

namespace fmm2d {

constexpr int kScanThreads = 256;
constexpr int kScanWaves = kScanThreads / 64;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

// Exclusive scan of v over a workgroup of NW waves; *total = the sum.  lds
// holds NW + 1 values.  Ends with a barrier, so lds may be reused.
template <int NW, typename T>
__device__ __forceinline__ T block_exclusive_scan_w(T v, T* lds, T* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const T y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    T s = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      const T t = lds[i];
      lds[i] = s;
      s += t;
    }
    lds[NW] = s;
  }
  __syncthreads();
  const T ex = x - v + lds[w];
  *total = lds[NW];
  __syncthreads();
  return ex;
}

// the same over a workgroup of kScanThreads (256) threads
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* lds, T* total) {
  return block_exclusive_scan_w<kScanWaves>(v, lds, total);
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const T* __restrict__ in,
                                                              int64_t n, T* part) {
  __shared__ T lds[kScanWaves + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x;
  T s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const int64_t j = base + (int64_t)i * kScanThreads;
    if (j < n) s += in[j];
  }
  T tot;
  block_exclusive_scan(s, lds, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// One workgroup: exclusive scan of the nblk tile sums in place; *total = sum.
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_top(T* part, int64_t nblk, T* total) {
  __shared__ T lds[kScanWaves + 1];
  T carry = 0;
  for (int64_t b0 = 0; b0 < nblk; b0 += kScanThreads) {
    const int64_t j = b0 + threadIdx.x;
    const T v = j < nblk ? part[j] : T(0);
    T tot;
    const T ex = block_exclusive_scan(v, lds, &tot);
    if (j < nblk) part[j] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_down(const T* in, T* out, int64_t n,
                                                            const T* part) {
  __shared__ T lds[kScanWaves + 1];
  T carry = part[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x;
#pragma unroll 1
  for (int i = 0; i < kScanItems; ++i) {
    const int64_t j = base + (int64_t)i * kScanThreads;
    if (j - threadIdx.x >= n) break;  // uniform: the round starts past the end
    const T v = j < n ? in[j] : T(0);
    T tot;
    const T ex = block_exclusive_scan(v, lds, &tot);
    if (j < n) out[j] = carry + ex;
    carry += tot;
  }
}

// out[i] = in[0] + ... + in[i-1] for i < n; *total (device, may be null) =
// the sum of all n.  Stream-ordered, no host synchronisation.
// Short arrays in one launch: a single workgroup walks the array in tiles of
// kScanTile, each thread scanning kScanItems consecutive elements in
// registers and the workgroup scanning the thread totals.  The tree build
// scans dozens of per-level arrays of a few thousand elements, where three
// launches (and a partials buffer) cost more than the work.
// Arrays up to kScanSingleMax elements: one workgroup of kScan1Threads, each
// thread summing a contiguous chunk, one block scan of the chunk sums, then
// each thread writing its chunk's prefix sums (reading each element before
// writing it, so in place is allowed).  One launch where the three-launch scan
// is bound by launch latency.
constexpr int kScan1Threads = 1024;
constexpr int64_t kScanSingleMax = 16 * kScanTile;

template <typename T>
__global__ __launch_bounds__(kScan1Threads) void k_scan_single(const T* in, T* out, int64_t n,
                                                               T* total) {
  __shared__ T lds[kScan1Threads / 64 + 1];
  const int64_t chunk = (n + kScan1Threads - 1) / kScan1Threads;
  const int64_t b = (int64_t)threadIdx.x * chunk, e = min(n, b + chunk);
  T s = 0;
#pragma unroll 16
  for (int64_t j = b; j < e; ++j) s += in[j];
  T tot;
  T run = block_exclusive_scan_w<kScan1Threads / 64>(s, lds, &tot);
#pragma unroll 16
  for (int64_t j = b; j < e; ++j) {
    const T v = in[j];
    out[j] = run;
    run += v;
  }
  if (threadIdx.x == 0 && total) *total = tot;
}

template <typename T>
void exclusive_scan(const T* in, T* out, int64_t n, T* total, hipStream_t st) {
  if (n <= 0) {
    if (total) hip_throw(hipMemsetAsync(total, 0, sizeof(T), st), "hipMemsetAsync");
    return;
  }
  if (n <= kScanSingleMax) {
    hipLaunchKernelGGL(k_scan_single<T>, dim3(1), dim3(kScan1Threads), 0, st, in, out, n, total);
    hip_throw(hipGetLastError(), "scan");
    return;
  }
  const int64_t nblk = ceil_div(n, (int64_t)kScanTile);
  DBuf<T> part(nblk, st);
  hipLaunchKernelGGL(k_scan_reduce<T>, dim3(nblk), dim3(kScanThreads), 0, st, in, n, part.p);
  hipLaunchKernelGGL(k_scan_top<T>, dim3(1), dim3(kScanThreads), 0, st, part.p, nblk, total);
  hipLaunchKernelGGL(k_scan_down<T>, dim3(nblk), dim3(kScanThreads), 0, st, in, out, n,
                     (const T*)part.p);
  hip_throw(hipGetLastError(), "scan");
}

}  // namespace fmm2d

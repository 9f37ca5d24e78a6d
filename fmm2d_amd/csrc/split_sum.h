// Sum of split copies: out[i] += sum_{k < nsplit} part[k n + i] in a fixed
// order (the direct evaluators split a long source range over blocks, each
// writing a private copy of the outputs).
//
// With many splits (the few-target case: up to 2048 copies of a few hundred
// values) a thread per element would issue its nsplit loads one after the
// other (~1 ms for 2048 splits); there one wavefront takes one element, lane l
// the copies l, l + 64, ..., and a butterfly over the 64 lanes adds the lane
// sums in the same order every run.
#pragma once
#include "common.h"
#include "devbuf.h"

namespace fmm2d {

template <typename T>
__global__ void k_split_sum_thread(int64_t n, T* __restrict__ out, const T* __restrict__ part,
                                   int nsplit) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T s = out[i];
  for (int k = 0; k < nsplit; ++k) s += part[(int64_t)k * n + i];
  out[i] = s;
}

__device__ __forceinline__ double wave_xor(double v, int m) { return __shfl_xor(v, m, 64); }
__device__ __forceinline__ cplx wave_xor(cplx v, int m) {
  return cplx{__shfl_xor(v.re, m, 64), __shfl_xor(v.im, m, 64)};
}

template <typename T>
__global__ __launch_bounds__(256) void k_split_sum_wave(int64_t n, T* __restrict__ out,
                                                        const T* __restrict__ part, int nsplit) {
  const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);  // uniform per wavefront
  const int lane = threadIdx.x & 63;
  if (e >= n) return;
  T s{};
  for (int k = lane; k < nsplit; k += 64) s += part[(int64_t)k * n + e];
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) s += wave_xor(s, m);
  if (lane == 0) out[e] += s;
}

template <typename T>
inline void add_split_copies(T* out, const T* part, int64_t n, int nsplit, hipStream_t st) {
  if (n <= 0) return;
  if (nsplit >= 32)
    hipLaunchKernelGGL(k_split_sum_wave<T>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, n,
                       out, part, nsplit);
  else
    hipLaunchKernelGGL(k_split_sum_thread<T>, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, out,
                       part, nsplit);
  hip_throw(hipGetLastError(), "kernel launch");
}

}  // namespace fmm2d

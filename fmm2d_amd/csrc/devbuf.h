// Stream-ordered device buffers (hipMallocAsync / hipFreeAsync on the call's
// stream; the default memory pool keeps freed blocks for reuse, so repeated
// FMM calls do not go back to the driver).
#pragma once
#include "common.h"
#include <utility>

namespace fmm2d {

struct HipError {
  hipError_t err;
  const char* what;
};

inline void hip_throw(hipError_t e, const char* what) {
  if (e != hipSuccess) throw HipError{e, what};
}

template <typename T> struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  hipStream_t st = nullptr;
  DBuf() = default;
  DBuf(size_t count, hipStream_t s) { alloc(count, s); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept { swap(o); }
  DBuf& operator=(DBuf&& o) noexcept { reset(); swap(o); return *this; }
  ~DBuf() { reset(); }
  void swap(DBuf& o) { std::swap(p, o.p); std::swap(n, o.n); std::swap(st, o.st); }
  void alloc(size_t count, hipStream_t s) {
    reset();
    st = s;
    n = count;
    if (count) hip_throw(hipMallocAsync((void**)&p, sizeof(T) * count, s), "hipMallocAsync");
  }
  void zero() {
    if (n) hip_throw(hipMemsetAsync(p, 0, sizeof(T) * n, st), "hipMemsetAsync");
  }
  void fill_bytes(int v) {
    if (n) hip_throw(hipMemsetAsync(p, v, sizeof(T) * n, st), "hipMemsetAsync");
  }
  // grow to at least count elements, keeping the first `keep` elements
  void grow(size_t count, size_t keep) {
    if (count <= n) return;
    T* q = nullptr;
    hip_throw(hipMallocAsync((void**)&q, sizeof(T) * count, st), "hipMallocAsync");
    if (keep) hip_throw(hipMemcpyAsync(q, p, sizeof(T) * keep, hipMemcpyDeviceToDevice, st), "copy");
    if (p) hip_throw(hipFreeAsync(p, st), "hipFreeAsync");
    p = q;
    n = count;
  }
  void reset() {
    if (p) (void)hipFreeAsync(p, st);
    p = nullptr;
    n = 0;
  }
  T* get() const { return p; }
};

template <typename F, typename... A>
inline void launch1d(F kernel, int64_t n, hipStream_t st, A... args) {
  if (n <= 0) return;
  hipLaunchKernelGGL(kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, args...);
  hip_throw(hipGetLastError(), "kernel launch");
}

// Grid size rounded up to a multiple of the 8 XCDs, for kernels that index
// their work with xcd_block().
inline unsigned xcd_grid(int64_t nblocks) {
  return nblocks < 8 ? (unsigned)nblocks : (unsigned)(((nblocks + 7) / 8) * 8);
}

// launch1d for kernels whose thread index is xcd_block() * 256 + threadIdx.x
template <typename F, typename... A>
inline void launch1d_xcd(F kernel, int64_t n, hipStream_t st, A... args) {
  if (n <= 0) return;
  hipLaunchKernelGGL(kernel, dim3(xcd_grid(ceil_div(n, 256))), dim3(256), 0, st, n, args...);
  hip_throw(hipGetLastError(), "kernel launch");
}

}  // namespace fmm2d

// Device workspace buffers.  Blocks come from a cache kept per (device,
// stream): a freed block goes back to its stream's cache and is handed out
// again to a later request on the same stream, so reuse is ordered after
// every earlier use by the stream itself.  Returning large blocks to the
// runtime costs up to milliseconds of host time each, which inside an FMM
// call (dozens of temporaries between host synchronisations) would dominate
// the small stages; fmm2d_release_workspace() hands the cache back.
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

// Small transfers in stream order (hostio.hip).  readback copies `bytes` of
// device memory to the host and waits for them (a kernel stores them into
// pinned memory that the host polls: a few microseconds where a copy and a
// stream synchronisation take about twenty); upload_bytes copies host memory
// to the device without waiting (in a kernel's arguments; `src` may be reused
// at once).  Sizes beyond the limits, or not a multiple of 4, copy and
// synchronise.
constexpr size_t kReadbackMax = 4096;
constexpr size_t kUploadInline = 1024;
void readback(void* dst, const void* src, size_t bytes, hipStream_t st);
void upload_bytes(void* dst, const void* src, size_t bytes, hipStream_t st);

// Workspace cache (ws.cpp).  ws_alloc returns at least `bytes` bytes usable
// in stream order on `st`; ws_free returns the block to st's cache.
void* ws_alloc(size_t bytes, hipStream_t st);
void ws_free(void* p, hipStream_t st);
// Free every cached block of the current device (after a device sync).
int ws_release();
// Cap the bytes the cache holds per device (cached + in use); 0 = no cap.  A
// request beyond it first frees the cached blocks, then fails as out of memory.
void ws_set_limit(size_t bytes);

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
    if (count) p = static_cast<T*>(ws_alloc(sizeof(T) * count, s));
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
    T* q = static_cast<T*>(ws_alloc(sizeof(T) * count, st));
    if (keep) hip_throw(hipMemcpyAsync(q, p, sizeof(T) * keep, hipMemcpyDeviceToDevice, st), "copy");
    if (p) ws_free(p, st);
    p = q;
    n = count;
  }
  void reset() {
    if (p) ws_free(p, st);
    p = nullptr;
    n = 0;
  }
  T* get() const { return p; }
};

// Keep freed stream-ordered workspace in the current device's default memory
// pool instead of returning it to the driver at every synchronisation (the
// pool's default release threshold is 0), so that repeated calls and the
// buffers a call allocates between its host synchronisations are served from
// the pool.  Done once per device; the memory stays with the process.
void keep_pool_memory();

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

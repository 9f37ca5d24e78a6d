// Small host <-> device transfers in stream order (devbuf.h: readback,
// upload_bytes).
//
// readback: a one-block kernel stores the words into pinned host memory and
// then, after a system-scope fence, a sequence number; the host polls for
// that number.  A copy into pageable memory followed by a stream
// synchronisation takes about 20 microseconds per round trip on MI355X, this
// about 6, and an FMM call makes dozens of such round trips (level counts,
// list sizes).  The pinned boxes are kept in a pool for the life of the
// process; one box serves one readback at a time.
//
// upload_bytes: the words travel in the arguments of a store kernel, which the
// runtime copies at launch, so no synchronisation keeps the host memory alive.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "devbuf.h"

namespace fmm2d {

namespace {

constexpr int kBoxHead = 16;  // words before the data (the sequence number is word 0)

__global__ void k_readback(const unsigned* __restrict__ src, int nw, unsigned* box,
                           unsigned seq) {
  for (int i = threadIdx.x; i < nw; i += blockDim.x) box[kBoxHead + i] = src[i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(box, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

struct InlineWords {
  unsigned w[kUploadInline / 4];
};

__global__ void k_store_inline(unsigned* __restrict__ dst, int nw, InlineWords a) {
  for (int i = threadIdx.x; i < nw; i += blockDim.x) dst[i] = a.w[i];
}

struct Box {
  unsigned* h = nullptr;  // pinned: [0] sequence number, [kBoxHead ..] data
  unsigned seq = 0;
};

std::mutex g_box_mu;
std::vector<Box*> g_boxes;  // free boxes (never returned to the runtime)

Box* box_acquire() {
  {
    std::lock_guard<std::mutex> g(g_box_mu);
    if (!g_boxes.empty()) {
      Box* b = g_boxes.back();
      g_boxes.pop_back();
      return b;
    }
  }
  void* p = nullptr;
  hip_throw(hipHostMalloc(&p, 4 * kBoxHead + kReadbackMax, hipHostMallocCoherent),
            "hipHostMalloc (readback box)");
  Box* b = new Box;
  b->h = static_cast<unsigned*>(p);
  __atomic_store_n(b->h, 0u, __ATOMIC_RELEASE);
  return b;
}

void box_release(Box* b) {
  std::lock_guard<std::mutex> g(g_box_mu);
  g_boxes.push_back(b);
}

}  // namespace

void readback(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes == 0) return;
  if (bytes > kReadbackMax || bytes % 4 != 0) {
    hip_throw(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st), "copy");
    hip_throw(hipStreamSynchronize(st), "sync");
    return;
  }
  Box* b = box_acquire();
  struct Release {
    Box* b;
    ~Release() { box_release(b); }
  } rel{b};
  if (++b->seq == 0) b->seq = 1;  // 0 is the initial word
  const unsigned seq = b->seq;
  hipLaunchKernelGGL(k_readback, dim3(1), dim3(256), 0, st, static_cast<const unsigned*>(src),
                     (int)(bytes / 4), b->h, seq);
  hip_throw(hipGetLastError(), "readback launch");
  // poll for up to 200 microseconds, then block on the stream (which also
  // reports a failure of the work queued before)
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned it = 1;; ++it) {
    if (__atomic_load_n(b->h, __ATOMIC_ACQUIRE) == seq) break;
    if ((it & 1023u) == 0 &&
        std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(200)) {
      hip_throw(hipStreamSynchronize(st), "sync");
      if (__atomic_load_n(b->h, __ATOMIC_ACQUIRE) != seq)
        throw HipError{hipErrorUnknown, "readback"};
      break;
    }
  }
  std::memcpy(dst, b->h + kBoxHead, bytes);
}

void upload_bytes(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes == 0) return;
  if (bytes > kUploadInline || bytes % 4 != 0) {
    hip_throw(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), "upload");
    hip_throw(hipStreamSynchronize(st), "sync");  // src may die after return
    return;
  }
  InlineWords a;
  std::memcpy(a.w, src, bytes);
  hipLaunchKernelGGL(k_store_inline, dim3(1), dim3(256), 0, st, static_cast<unsigned*>(dst),
                     (int)(bytes / 4), a);
  hip_throw(hipGetLastError(), "upload launch");
}

}  // namespace fmm2d

// Workspace cache behind DBuf (devbuf.h).
//
// Blocks are cached per (device, stream).  A block freed by a call on stream
// S is handed out again only to requests on S, after everything S already
// queued, which is the same ordering guarantee stream-ordered allocation
// gives.  Sizes are rounded (256 B below 1 MiB, 1 MiB above) and a cached
// block serves requests down to half its size.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

#include "devbuf.h"

namespace fmm2d {

namespace {

using Key = std::pair<int, uintptr_t>;  // (device, stream)

struct Owner {
  Key key;
  size_t bytes;
};

std::mutex g_mu;
std::map<Key, std::multimap<size_t, void*>> g_free;  // cached blocks by size
std::unordered_map<void*, Owner> g_owner;            // every block we hold
std::map<int, size_t> g_held;                        // bytes held per device
size_t g_limit = 0;                                  // 0: no limit

size_t round_size(size_t b) {
  const size_t mib = size_t(1) << 20;
  if (b < mib) return (b + 255) & ~size_t(255);
  return (b + mib - 1) & ~(mib - 1);
}

// free the cached (not in-use) blocks of one device; caller holds g_mu
void drop_cached(int dev) {
  for (auto it = g_free.begin(); it != g_free.end();) {
    if (it->first.first != dev) {
      ++it;
      continue;
    }
    for (auto& kv : it->second) {
      (void)hipFree(kv.second);
      g_owner.erase(kv.second);
      g_held[dev] -= kv.first;
    }
    it = g_free.erase(it);
  }
}

}  // namespace

void* ws_alloc(size_t bytes, hipStream_t st) {
  int dev = 0;
  hip_throw(hipGetDevice(&dev), "hipGetDevice");
  const size_t want = round_size(bytes == 0 ? 1 : bytes);
  const Key key{dev, reinterpret_cast<uintptr_t>(st)};
  std::lock_guard<std::mutex> lock(g_mu);
  auto& fl = g_free[key];
  auto it = fl.lower_bound(want);
  if (it != fl.end() && it->first <= 2 * want) {
    void* p = it->second;
    fl.erase(it);
    return p;
  }
  if (g_limit && g_held[dev] + want > g_limit) {
    // over the caller's cap: give the cached blocks back first
    (void)hipDeviceSynchronize();
    drop_cached(dev);
    if (g_held[dev] + want > g_limit) throw HipError{hipErrorOutOfMemory, "workspace limit"};
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e == hipErrorOutOfMemory) {
    // give the cached blocks of this device back and try once more
    (void)hipGetLastError();
    (void)hipDeviceSynchronize();
    drop_cached(dev);
    e = hipMalloc(&p, want);
  }
  hip_throw(e, "hipMalloc");
  g_owner[p] = Owner{key, want};
  g_held[dev] += want;
  return p;
}

void ws_free(void* p, hipStream_t st) {
  (void)st;
  if (!p) return;
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_owner.find(p);
  if (it == g_owner.end()) return;
  g_free[it->second.key].emplace(it->second.bytes, p);
}

int ws_release() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return FMM2D_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_mu);
  if (hipDeviceSynchronize() != hipSuccess) return FMM2D_ERR_HIP;
  drop_cached(dev);
  return FMM2D_OK;
}

void ws_set_limit(size_t bytes) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_limit = bytes;
}

}  // namespace fmm2d

// Stable LSD radix sort of 32- or 64-bit keys (optionally with 32-bit values) on the
// device: the point order of the tree build (common/pts_tree2d.f:1222-1246,
// pts_tree_sort, is a stable partition by the quadrant at every level; sorting
// the path keys is the same order).
//
// Per 8-bit digit, three steps:
//   1. k_rs_hist: every tile of kRsTile keys counts its digits (wave-level
//      peer groups from 8 ballots, one LDS add per group) into
//      hist[digit][tile] (digit-major);
//   2. exclusive_scan over hist (scan.h): offs[digit][tile] = where the tile's
//      keys of that digit start in the output;
//   3. k_rs_scatter: every tile ranks its keys stably (the four waves take
//      contiguous quarters of the tile in rounds of 64 consecutive keys: rank =
//      keys of the digit in earlier waves + earlier rounds + lower lanes),
//      places them in LDS in digit order and writes them out with consecutive
//      threads on consecutive positions of each digit's run.
// Traffic per pass: keys read twice, keys (and values) read and written once.
#pragma once
#include "common.h"
#include "devbuf.h"
#include "scan.h"

namespace fmm2d {

constexpr int kRsThreads = 256;
constexpr int kRsWaves = kRsThreads / 64;
constexpr int kRsItems = 8;                      // keys per thread
constexpr int kRsTile = kRsThreads * kRsItems;   // keys per tile
constexpr int kRsWaveKeys = kRsTile / kRsWaves;  // keys per wave (kRsItems rounds of 64)

__device__ __forceinline__ unsigned long long rs_lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// lanes of the wave whose 8-bit digit equals this lane's (among `active`)
__device__ __forceinline__ unsigned long long rs_peers(unsigned d, unsigned long long active) {
  unsigned long long m = active;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const unsigned long long bal = __ballot((d >> b) & 1u);
    m &= ((d >> b) & 1u) ? bal : ~bal;
  }
  return m;
}

template <typename K>
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const K* __restrict__ keys,
                                                        int64_t n, int shift, int ntiles,
                                                        unsigned* __restrict__ hist) {
  __shared__ unsigned h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t base = (int64_t)blockIdx.x * kRsTile + (int64_t)w * kRsWaveKeys + lane;
#pragma unroll
  for (int r = 0; r < kRsItems; ++r) {
    const int64_t j = base + 64 * r;
    const bool ok = j < n;
    const unsigned long long act = __ballot(ok);
    if (act == 0) break;  // uniform
    const unsigned d = ok ? (unsigned)((keys[j] >> shift) & K(255)) : 0u;
    const unsigned long long m = rs_peers(d, act);
    // the highest lane of each peer group adds the group's size
    if (ok && (m >> lane) == 1ull) atomicAdd(&h[d], (unsigned)__popcll(m));
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

template <typename K, bool HAS_V>
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    const K* __restrict__ kin, K* __restrict__ kout,
    const int* __restrict__ vin, int* __restrict__ vout, int64_t n, int shift, int ntiles,
    const unsigned* __restrict__ offs) {
  __shared__ K sk[kRsTile];
  __shared__ int sv[HAS_V ? kRsTile : 1];
  __shared__ unsigned wcnt[kRsWaves][256];
  __shared__ unsigned start[256];  // tile-local start of each digit's run
  __shared__ unsigned gbase[256];  // output position of the tile's first key of each digit
  __shared__ unsigned lds_scan[kScanWaves + 1];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t tile0 = (int64_t)blockIdx.x * kRsTile;
  const int64_t base = tile0 + (int64_t)w * kRsWaveKeys + lane;
#pragma unroll
  for (int q = 0; q < kRsWaves; ++q) wcnt[q][threadIdx.x] = 0;
  __syncthreads();
  K key[kRsItems];
  int val[kRsItems];
  unsigned long long act[kRsItems], peer[kRsItems];
  unsigned dig[kRsItems];
  // phase A: digit counts of each wave's quarter (peer groups per round)
#pragma unroll
  for (int r = 0; r < kRsItems; ++r) {
    const int64_t j = base + 64 * r;
    const bool ok = j < n;
    act[r] = __ballot(ok);
    key[r] = ok ? kin[j] : K(0);
    if (HAS_V) val[r] = ok ? vin[j] : 0;
    dig[r] = (unsigned)((key[r] >> shift) & K(255));
    peer[r] = act[r] ? rs_peers(dig[r], act[r]) : 0ull;
    if (ok && (peer[r] >> lane) == 1ull) wcnt[w][dig[r]] += (unsigned)__popcll(peer[r]);
  }
  __syncthreads();
  {
    // per digit (thread = digit): wave prefixes, tile-local start, output base
    const int d = threadIdx.x;
    unsigned s = 0;
#pragma unroll
    for (int q = 0; q < kRsWaves; ++q) {
      const unsigned c = wcnt[q][d];
      wcnt[q][d] = s;
      s += c;
    }
    unsigned tot;
    const unsigned st = block_exclusive_scan(s, lds_scan, &tot);
    start[d] = st;
    gbase[d] = offs[(int64_t)d * ntiles + blockIdx.x];
#pragma unroll
    for (int q = 0; q < kRsWaves; ++q) wcnt[q][d] += st;  // running local positions
  }
  __syncthreads();
  // phase B: stable ranks, keys placed in LDS in digit order
#pragma unroll
  for (int r = 0; r < kRsItems; ++r) {
    const bool ok = (act[r] >> lane) & 1ull;
    if (act[r] == 0) break;  // uniform
    unsigned pos = 0;
    if (ok) pos = wcnt[w][dig[r]] + (unsigned)__popcll(peer[r] & rs_lanemask_lt());
    __builtin_amdgcn_wave_barrier();  // every peer has read the running position
    if (ok && (peer[r] >> lane) == 1ull) wcnt[w][dig[r]] += (unsigned)__popcll(peer[r]);
    __builtin_amdgcn_wave_barrier();
    if (ok) {
      sk[pos] = key[r];
      if (HAS_V) sv[pos] = val[r];
    }
  }
  __syncthreads();
  // phase C: consecutive threads write consecutive positions of each run
  const int cnt = (int)min((int64_t)kRsTile, n - tile0);
#pragma unroll
  for (int r = 0; r < kRsItems; ++r) {
    const int p = r * kRsThreads + threadIdx.x;
    if (p >= cnt) break;
    const K k = sk[p];
    const unsigned d = (unsigned)((k >> shift) & K(255));
    const int64_t dst = (int64_t)gbase[d] + (p - (int)start[d]);
    kout[dst] = k;
    if (HAS_V) vout[dst] = sv[p];
  }
}

// Stable sort of n unsigned keys (32 or 64 bits) by their low `bits` bits
// (values, if given, move with their keys).  kin / vin are not modified; kout /
// vout receive the result.  n < 2^32.
template <typename K>
void radix_sort(const K* kin, K* kout, const int* vin, int* vout, int64_t n, int bits,
                hipStream_t st) {
  if (n <= 0) return;
  const bool hv = vin != nullptr;
  const int npass = bits <= 0 ? 0 : (bits + 7) / 8;
  if (npass == 0) {
    hip_throw(hipMemcpyAsync(kout, kin, sizeof(K) * n, hipMemcpyDeviceToDevice, st), "copy");
    if (hv)
      hip_throw(hipMemcpyAsync(vout, vin, sizeof(int) * n, hipMemcpyDeviceToDevice, st), "copy");
    return;
  }
  const int ntiles = (int)ceil_div(n, (int64_t)kRsTile);
  DBuf<unsigned> hist((size_t)256 * ntiles, st);
  DBuf<K> kt(npass > 1 ? n : 0, st);
  DBuf<int> vt((npass > 1 && hv) ? n : 0, st);
  const K* ks = kin;
  const int* vs = vin;
  for (int p = 0; p < npass; ++p) {
    // the last pass writes kout: odd passes from the end go to the scratch
    const bool to_out = ((npass - 1 - p) % 2) == 0;
    K* kd = to_out ? kout : kt.p;
    int* vd = to_out ? vout : vt.p;
    hipLaunchKernelGGL(k_rs_hist<K>, dim3(ntiles), dim3(kRsThreads), 0, st, ks, n, 8 * p,
                       ntiles, hist.p);
    exclusive_scan(hist.p, hist.p, (int64_t)256 * ntiles, (unsigned*)nullptr, st);
    if (hv)
      hipLaunchKernelGGL((k_rs_scatter<K, true>), dim3(ntiles), dim3(kRsThreads), 0, st, ks, kd, vs,
                         vd, n, 8 * p, ntiles, (const unsigned*)hist.p);
    else
      hipLaunchKernelGGL((k_rs_scatter<K, false>), dim3(ntiles), dim3(kRsThreads), 0, st, ks, kd,
                         (const int*)nullptr, (int*)nullptr, n, 8 * p, ntiles,
                         (const unsigned*)hist.p);
    hip_throw(hipGetLastError(), "radix sort");
    ks = kd;
    vs = vd;
  }
}

}  // namespace fmm2d

// Wave- and block-level scans shared by the device-wide scan (scan.hip) and the
// batched rebuild of the SPH step: 256-thread blocks, 1024-element tiles.
#pragma once

#include "gclsph_internal.hpp"

namespace gclsph {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kTile = kScanThreads * kScanItems;  // 1024

template <typename T>
__device__ inline T wave_inclusive_scan(T v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Exclusive block scan over kScanThreads values; returns the thread's exclusive
// prefix and writes the block total to *total.
template <typename T>
__device__ inline T block_exclusive_scan(T v, T* total) {
  __shared__ T wsum[kScanThreads / 64];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  T inc = wave_inclusive_scan(v, lane);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  T pre = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kScanThreads / 64; ++k) {
    if (k < w) pre += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  *total = tot;
  return pre + inc - v;
}

}  // namespace gclsph

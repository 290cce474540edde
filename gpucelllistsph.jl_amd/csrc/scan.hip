// Device-wide exclusive scan of int32 counts (cell histograms, per-particle pair
// counts).  Three passes over 1024-element tiles: tile sums, a single-block scan
// of the tile sums, then the tile scans with their prefix added.  Every kernel
// takes an optional device flag and returns at once while it is zero, so a
// rebuild that the device decides against costs only the launches.
#include "blockscan.hpp"

namespace gclsph {
namespace {

__global__ void k_tile_sums(const int32_t* __restrict__ in, int64_t n, int64_t* __restrict__ tsum,
                            const int* flag, int32_t* max_out) {
  if (flag && *flag == 0) return;
  int64_t base = int64_t(blockIdx.x) * kTile;
  int64_t s = 0;
  int32_t mx = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + int64_t(k) * kScanThreads + threadIdx.x;
    if (i < n) {
      int32_t v = in[i];
      s += v;
      mx = v > mx ? v : mx;
    }
  }
  int64_t tot;
  block_exclusive_scan<int64_t>(s, &tot);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
  if (max_out) {  // tile maximum, reduced by k_scan_tile_sums
    __shared__ int32_t wmx[kScanThreads / 64];
    for (int o = 32; o > 0; o >>= 1) {
      int32_t t = __shfl_xor(mx, o, 64);
      mx = t > mx ? t : mx;
    }
    if ((threadIdx.x & 63) == 0) wmx[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      int32_t m = wmx[0];
      for (int w = 1; w < kScanThreads / 64; ++w) m = wmx[w] > m ? wmx[w] : m;
      reinterpret_cast<int32_t*>(tsum + gridDim.x + 2)[blockIdx.x] = m;
    }
  }
}

// Single block: exclusive scan of tile sums in place; tsum[nt] = grand total;
// with max_out, also the maximum of the tile maxima stored after the sums.
__global__ void k_scan_tile_sums(int64_t* __restrict__ tsum, int64_t nt, const int* flag, int32_t* max_out) {
  if (flag && *flag == 0) return;
  if (max_out) {
    const int32_t* tm = reinterpret_cast<const int32_t*>(tsum + nt + 2);
    int32_t mx = 0;
    for (int64_t i = threadIdx.x; i < nt; i += kScanThreads) mx = tm[i] > mx ? tm[i] : mx;
    for (int o = 32; o > 0; o >>= 1) {
      int32_t t = __shfl_xor(mx, o, 64);
      mx = t > mx ? t : mx;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(max_out, mx);
  }
  int64_t per = (nt + kScanThreads - 1) / kScanThreads;
  int64_t b = int64_t(threadIdx.x) * per, e = b + per < nt ? b + per : nt;
  int64_t s = 0;
  for (int64_t i = b; i < e; ++i) s += tsum[i];
  int64_t tot;
  int64_t pre = block_exclusive_scan<int64_t>(s, &tot);
  for (int64_t i = b; i < e; ++i) {
    int64_t v = tsum[i];
    tsum[i] = pre;
    pre += v;
  }
  if (threadIdx.x == 0) tsum[nt] = tot;
}

template <typename O>
__global__ void k_tile_scan(const int32_t* __restrict__ in, O* __restrict__ out, int64_t n,
                            const int64_t* __restrict__ tsum, int64_t nt, const int* flag) {
  if (flag && *flag == 0) return;
  int64_t base = int64_t(blockIdx.x) * kTile + int64_t(threadIdx.x) * kScanItems;
  int32_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    v[k] = i < n ? in[i] : 0;
    s += v[k];
  }
  int64_t tot;
  int64_t pre = block_exclusive_scan<int64_t>(s, &tot) + tsum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    if (i < n) out[i] = O(pre);
    pre += v[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = O(tsum[nt]);
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v, const int* flag) {
  if (flag && *flag == 0) return;
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

template <typename O>
void scan_impl(const int32_t* in, O* out, int64_t n, int64_t* tmp, const int* flag, hipStream_t s,
               int32_t* max_out) {
  int64_t nt = (n + kTile - 1) / kTile;
  if (nt < 1) nt = 1;
  k_tile_sums<<<unsigned(nt), kScanThreads, 0, s>>>(in, n, tmp, flag, max_out);
  GCL_HIP(hipGetLastError());
  k_scan_tile_sums<<<1, kScanThreads, 0, s>>>(tmp, nt, flag, max_out);
  GCL_HIP(hipGetLastError());
  k_tile_scan<O><<<unsigned(nt), kScanThreads, 0, s>>>(in, out, n, tmp, nt, flag);
  GCL_HIP(hipGetLastError());
}

}  // namespace

// tile sums (nt + 2 int64) followed by the tile maxima (nt int32)
size_t scan_tmp_elems(int64_t n) {
  int64_t nt = (n + kTile - 1) / kTile;
  return size_t(nt + 2 + (nt + 1) / 2 + 1);
}

void exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, int64_t* tmp, const int* flag,
                        hipStream_t s, int32_t* max_out) {
  scan_impl<int32_t>(in, out, n, tmp, flag, s, max_out);
}

void exclusive_scan_i64(const int32_t* in, int64_t* out, int64_t n, int64_t* tmp, const int* flag,
                        hipStream_t s) {
  scan_impl<int64_t>(in, out, n, tmp, flag, s, nullptr);
}

void fill_i32(int32_t* p, int64_t n, int32_t v, const int* flag, hipStream_t s) {
  int64_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  k_fill_i32<<<unsigned(blocks), kBlock, 0, s>>>(p, n, v, flag);
  GCL_HIP(hipGetLastError());
}

}  // namespace gclsph

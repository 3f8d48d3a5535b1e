// Exclusive prefix sum of int32 on the device (three launches: tile scan, scan of tile sums
// in one workgroup, tile offsets), and order-preserving stream compaction built on it.
#pragma once
#include "common.h"

namespace dlp {

constexpr int kScanItems = 4;                       // items per thread
constexpr int kScanTile = kBlock * kScanItems;      // 1024 items per tile

constexpr inline size_t scan_workspace_bytes(int n) {
  const size_t tiles = (size_t)(n + kScanTile - 1) / kScanTile + 1;
  return ((tiles * 4 + 255) & ~(size_t)255);
}

// Inclusive scan across the 64 lanes of a wavefront.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive prefix and the
// block total in *total (all threads).
__device__ __forceinline__ int block_excl_scan(int v, int* total) {
  __shared__ int wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int inc = wave_incl_scan(v);
  if (lane == 63) wsum[wid] = inc;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kBlock / 64; ++k) {
    if (k < wid) base += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

// out[i] = sum_{k<i} in[k] (in and out may alias); *total = sum of all (device int32, may be NULL).
int exclusive_scan_i32(int n, const int32_t* in, int32_t* out, int32_t* total, void* ws,
                       size_t ws_bytes, hipStream_t st);

// Order-preserving compaction: indices i with mask[i] != 0 -> out, *count (device).
int compact_u8(int n, const uint8_t* mask, int32_t* out, int32_t* count, void* ws,
               size_t ws_bytes, hipStream_t st);

}  // namespace dlp

// Device prefix sums and order-preserving compaction.
#include "scan.h"

namespace dlp {

__global__ __launch_bounds__(256) void k_scan_tiles(int n, const int32_t* __restrict__ in,
                                                    int32_t* __restrict__ out,
                                                    int32_t* __restrict__ tile_sum) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int x[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    x[k] = (base + k < n) ? in[base + k] : 0;
    s += x[k];
  }
  int tot;
  int pre = block_excl_scan(s, &tot);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = pre;
    pre += x[k];
  }
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_sums(int ntiles, int32_t* __restrict__ tile_sum,
                                                   int32_t* __restrict__ total) {
  int carry = 0;
  for (int b = 0; b < ntiles; b += kBlock) {
    const int i = b + threadIdx.x;
    const int v = i < ntiles ? tile_sum[i] : 0;
    int tot;
    const int pre = block_excl_scan(v, &tot);
    if (i < ntiles) tile_sum[i] = carry + pre;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(int n, int32_t* __restrict__ out,
                                                  const int32_t* __restrict__ tile_sum) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += tile_sum[i / kScanTile];
}

// out[i] = sum_{k<i} in[k] (in and out may alias); *total = sum of all (device int32, may be NULL).
int exclusive_scan_i32(int n, const int32_t* in, int32_t* out, int32_t* total, void* ws,
                              size_t ws_bytes, hipStream_t st) {
  DLP_CHECK_ARG(ws_bytes >= scan_workspace_bytes(n), "scan: workspace too small");
  const int ntiles = (n + kScanTile - 1) / kScanTile;
  int32_t* tile_sum = (int32_t*)ws;
  if (n > 0) {
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(kBlock), 0, st, n, in, out, tile_sum);
    DLP_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, st, ntiles, tile_sum, total);
  DLP_CHECK_LAUNCH();
  if (n > 0) {
    hipLaunchKernelGGL(k_scan_add, dim3(grid_for(n)), dim3(kBlock), 0, st, n, out, tile_sum);
    DLP_CHECK_LAUNCH();
  }
  return DLP_OK;
}


// Flags from a byte mask, scanned; then each set index writes itself at its prefix.
__global__ void k_mask_to_i32(int n, const uint8_t* __restrict__ mask, int32_t* __restrict__ f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = mask[i] != 0;
}
__global__ void k_scatter_set(int n, const uint8_t* __restrict__ mask,
                              const int32_t* __restrict__ pos, int32_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && mask[i]) out[pos[i]] = (int32_t)i;
}

int compact_u8(int n, const uint8_t* mask, int32_t* out, int32_t* count, void* ws,
               size_t ws_bytes, hipStream_t st) {
  const size_t a = ((size_t)n * 4 + 255) & ~(size_t)255;
  DLP_CHECK_ARG(ws_bytes >= a + scan_workspace_bytes(n), "flatnonzero: workspace too small");
  int32_t* pos = (int32_t*)ws;
  if (n > 0) {
    hipLaunchKernelGGL(k_mask_to_i32, dim3(grid_for(n)), dim3(kBlock), 0, st, n, mask, pos);
    DLP_CHECK_LAUNCH();
  }
  int rc = exclusive_scan_i32(n, pos, pos, count, (char*)ws + a, ws_bytes - a, st);
  if (rc) return rc;
  if (n > 0) {
    hipLaunchKernelGGL(k_scatter_set, dim3(grid_for(n)), dim3(kBlock), 0, st, n, mask, pos, out);
    DLP_CHECK_LAUNCH();
  }
  return DLP_OK;
}

}  // namespace dlp

extern "C" size_t dlp_scan_workspace_bytes(int n) {
  if (n < 1) n = 1;
  return (((size_t)n * 4 + 255) & ~(size_t)255) + dlp::scan_workspace_bytes(n);
}

extern "C" int dlp_flatnonzero_u8(int n, const uint8_t* mask, int32_t* out, int32_t* count,
                                  void* ws, size_t ws_bytes, void* stream) {
  DLP_CHECK_ARG(n >= 0 && count, "flatnonzero: bad arguments");
  DLP_CHECK_ARG(n == 0 || (mask && out), "flatnonzero: null pointer");
  return dlp::compact_u8(n, mask, out, count, ws, ws_bytes, (hipStream_t)stream);
}

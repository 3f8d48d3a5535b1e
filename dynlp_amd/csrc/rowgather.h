// Row-group gather/accumulate shared by the SpMM, frontier Jacobi and residual kernels.
//
// A row is handled by a group of G = 2^g_shift lanes inside one 64-wide wavefront: LC column
// lanes times LE = G/LC edge lanes.  The group loads up to G (column, weight) pairs of the row
// with one coalesced access, then broadcasts them with cross-lane shuffles, so each edge lane
// issues neighbour-row gathers without waiting on the index load; four edges per edge lane
// are unrolled to keep several neighbour rows in flight.  Column chunk k of lane l covers
// column l + k*LC, so one neighbour row X[v, :] is read as LC contiguous elements.
#pragma once
#include "common.h"

namespace dlp {

struct Lanes {
  int G, g_shift, LC, lc_shift, LE, lane, l, e;
  __device__ __forceinline__ Lanes(int gs, int ls) {
    g_shift = gs;
    lc_shift = ls;
    G = 1 << gs;
    LC = 1 << ls;
    LE = G >> ls;
    lane = threadIdx.x & (G - 1);
    l = lane & (LC - 1);
    e = lane >> ls;
  }
};

// acc[k] += sum over entries j in [beg, end) with mask[col_j] != 0 of w_j * X[col_j, l + k*LC].
// All lanes of the group must call this with the same beg/end.
template <typename TW, typename TX, typename TA, int CPL, int U = 4>
__device__ __forceinline__ void row_gather(const Lanes& L, int beg, int end,
                                           const int32_t* __restrict__ indices,
                                           const TW* __restrict__ w,
                                           const uint8_t* __restrict__ mask,
                                           const TX* __restrict__ X, int64_t ldx, int C,
                                           TA (&acc)[CPL]) {
  for (int base = beg; base < end; base += L.G) {
    const int j = base + L.lane;
    int myv = 0;
    TA myw = 0;
    if (j < end) {
      myv = indices[j];
      myw = (TA)w[j];
      if (mask != nullptr && mask[myv] == 0) myw = 0;
    }
    const int cnt = min(L.G, end - base);
    for (int t0 = 0; t0 < cnt; t0 += U * L.LE) {
      int vv[U];
      TA ww[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int t = t0 + u * L.LE + L.e;
        vv[u] = __shfl(myv, t & (L.G - 1), L.G);
        ww[u] = __shfl(myw, t & (L.G - 1), L.G);
        if (t >= cnt) ww[u] = 0;
      }
      TA xv[U][CPL];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const TX* xr = X + (int64_t)vv[u] * ldx;
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
          const int c = L.l + k * L.LC;
          xv[u][k] = (ww[u] != (TA)0 && c < C) ? (TA)xr[c] : (TA)0;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int k = 0; k < CPL; ++k) acc[k] += ww[u] * xv[u][k];
    }
  }
}

// FP32 gather with 16-byte loads: lane l of the column lanes covers columns 4l..4l+3 (X rows
// 16-byte aligned: ldx % 4 == 0), so a 65-column row (ld 68) is 17 lanes of one instruction.
template <int U>
__device__ __forceinline__ void row_gather_v4(const Lanes& L, int beg, int end,
                                              const int32_t* __restrict__ indices,
                                              const float* __restrict__ w,
                                              const uint8_t* __restrict__ mask,
                                              const float* __restrict__ X, int64_t ldx, int C,
                                              float4& acc) {
  const bool lane_on = 4 * L.l < C;
  for (int base = beg; base < end; base += L.G) {
    const int j = base + L.lane;
    int myv = 0;
    float myw = 0.f;
    if (j < end) {
      myv = indices[j];
      myw = w[j];
      if (mask != nullptr && mask[myv] == 0) myw = 0.f;
    }
    const int cnt = min(L.G, end - base);
    for (int t0 = 0; t0 < cnt; t0 += U * L.LE) {
      int vv[U];
      float ww[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int t = t0 + u * L.LE + L.e;
        vv[u] = __shfl(myv, t & (L.G - 1), L.G);
        ww[u] = __shfl(myw, t & (L.G - 1), L.G);
        if (t >= cnt) ww[u] = 0.f;
      }
      float4 xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float4* xr = reinterpret_cast<const float4*>(X + (int64_t)vv[u] * ldx) + L.l;
        xv[u] = (ww[u] != 0.f && lane_on) ? *xr : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        acc.x += ww[u] * xv[u].x;
        acc.y += ww[u] * xv[u].y;
        acc.z += ww[u] * xv[u].z;
        acc.w += ww[u] * xv[u].w;
      }
    }
  }
}

// Process-wide tuning knobs (dlp_tune_set): unroll factor of the scalar gather (4 or 8) and
// whether FP32 operators may use 16-byte loads when the layout allows it.
struct Tuning {
  int unroll = 8;
  int vec4 = 1;
};
Tuning& tuning();

// Sum acc over the edge lanes of the group (xor butterflies inside the group).
template <typename TA, int CPL>
__device__ __forceinline__ void group_reduce(const Lanes& L, TA (&acc)[CPL]) {
  for (int off = L.LC; off < L.G; off <<= 1) {
#pragma unroll
    for (int k = 0; k < CPL; ++k) acc[k] += __shfl_xor(acc[k], off, 64);
  }
}

// Sum of an int over the whole wavefront, result in every lane.
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace dlp

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

// Logical block index: with xcd_per > 0 (grid a multiple of 8, xcd_per = grid / 8) the blocks
// the dispatcher sends to one XCD (blockIdx % 8) take consecutive logical indices, so each
// XCD works through one contiguous eighth of the rows.
__device__ __forceinline__ int64_t block_order(int xcd_per) {
  return xcd_per ? (int64_t)(blockIdx.x & 7) * xcd_per + (blockIdx.x >> 3) : (int64_t)blockIdx.x;
}

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

// Packed 16-byte layout for one row per wavefront: LC = ceil(C/4) column lanes (not rounded
// to a power of two) times LE = 64 / LC edge lanes; lanes e >= LE idle.  For C = 65 this is
// 3 edges x 17 lanes (51 busy lanes instead of 2 x 32 with 34 busy), so a degree-24 row is one
// unrolled step of 8 edges per edge lane.
// With pow2 set, LC is rounded up to a power of two (the xor-butterfly layout); `on` marks the
// lanes that load (edge lane < LE and at least one column < C).
struct LanesP {
  int LC, LE, lane, l, e;
  bool on;
  __device__ __forceinline__ explicit LanesP(int C, bool pow2 = false) {
    const int need = (C + 3) >> 2;
    LC = (pow2 && need > 1) ? 1 << (32 - __clz(need - 1)) : need;
    LE = 64 / LC;
    lane = threadIdx.x & 63;
    e = lane / LC;
    l = lane - e * LC;
    on = e < LE && 4 * l < C;
  }
};

template <int U>
__device__ __forceinline__ void row_gather_v4p(const LanesP& L, int beg, int end,
                                               const int32_t* __restrict__ indices,
                                               const float* __restrict__ w,
                                               const uint8_t* __restrict__ mask,
                                               const float* __restrict__ X, int64_t ldx,
                                               float4& acc) {
  const bool on = L.on;
  for (int base = beg; base < end; base += 64) {
    const int j = base + L.lane;
    int myv = 0;
    float myw = 0.f;
    if (j < end) {
      myv = indices[j];
      myw = w[j];
      if (mask != nullptr && mask[myv] == 0) myw = 0.f;
    }
    const int cnt = min(64, end - base);
    for (int t0 = 0; t0 < cnt; t0 += U * L.LE) {
      int vv[U];
      float ww[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int t = t0 + u * L.LE + L.e;
        vv[u] = __shfl(myv, t & 63, 64);
        ww[u] = __shfl(myw, t & 63, 64);
        if (t >= cnt || !on) ww[u] = 0.f;
      }
      float4 xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float4* xr = reinterpret_cast<const float4*>(X + (int64_t)vv[u] * ldx) + L.l;
        xv[u] = ww[u] != 0.f ? *xr : make_float4(0.f, 0.f, 0.f, 0.f);
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

// Sum of the LE edge-lane partials of column lane l, valid in lanes e == 0.
__device__ __forceinline__ float4 reduce_v4p(const LanesP& L, float4 acc) {
  float4 tot = acc;
  for (int k = 1; k < L.LE; ++k) {
    const int src = L.lane + k * L.LC;
    tot.x += __shfl(acc.x, src, 64);
    tot.y += __shfl(acc.y, src, 64);
    tot.z += __shfl(acc.z, src, 64);
    tot.w += __shfl(acc.w, src, 64);
  }
  return tot;
}

// Flat 16-byte passes over a row-major n x ld FP32 block (ld % 4 == 0) as n*Q4 float4, Q4 =
// ld/4.  A workgroup uses Q4 * (blockDim/Q4) of its threads and the grid stride is a multiple
// of Q4, so every thread keeps one column quad for the whole pass and consecutive threads read
// consecutive 16-byte words.  cnt = columns of the quad below C.
struct Quad {
  bool active;  // threadIdx < Q4 * (256 / Q4); the rest of the workgroup only joins reductions
  int q4;       // column quad of this thread
  int cnt;      // columns of the quad below C (0..4)
  int64_t i0;   // first float4 index
  int64_t r0;   // its row
  int64_t step; // float4 stride
  int64_t rstep;
};

__device__ __forceinline__ Quad quad_of(int Q4, int C) {
  Quad q;
  const int used = Q4 * ((int)blockDim.x / Q4);
  q.active = (int)threadIdx.x < used;
  q.q4 = threadIdx.x % Q4;
  q.cnt = q.active ? max(0, min(4, C - 4 * q.q4)) : 0;
  q.i0 = (int64_t)blockIdx.x * used + threadIdx.x;
  q.r0 = q.i0 / Q4;
  q.step = (int64_t)gridDim.x * used;
  q.rstep = q.step / Q4;
  return q;
}

__device__ __forceinline__ float quad_comp(const float4& v, int j) {
  return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w;
}

// Process-wide tuning knobs (dlp_tune_set).
struct Tuning {
  int unroll = 0;     // 0: per kernel (8 for the scalar and power-of-two gathers, 4 for packed)
  int vec4 = 1;
  int pack = 1;       // packed (non-power-of-two) lane groups for one-row-per-wave FP32 kernels
  int xcd = 1;        // XCD-aware row order: each XCD walks one contiguous eighth of the rows
  int dot_grid = 0;   // blocks of the grid-stride operator with fused dot (0: resident blocks)
  int fuse_dot = 1;   // fuse p^T A p into the operator launch of the global solve
  int pipe = 4;       // gathers per lane in flight in the pipelined global-solve operator (0: off)
  int bf16_word = 4;  // BF16 direction values per lane in the global-solve operator (4 or 8)
  int flat_grid = 0;  // workgroups of the flat PCG vector passes (0: 2048)
  int nt_streams = 1; // non-temporal FP32 streams (Z, R, AP) in the PCG update and direction passes
  int defer_z = 1;    // low-precision directions: Z updated once per captured group of PCG iterations
  int f8_pipe = 0;    // gathers per lane in flight in the FP8 global-solve operator (4, 6 or 8; 0:
                      // 8 when at most 4 edge lanes share a gather, i.e. wide rows, else 4)
  int f8_waves = 6;   // FP8 operator at depth 6 or 8: register budget for 5 or 6 waves per SIMD
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

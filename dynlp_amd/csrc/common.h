// Shared helpers for the gfx950 kernels of the label-propagation hot path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>
#include <string.h>
#include "../../include/dynlp.h"

// Dispatch on the column-chunk count of a row group (template parameter CPL).
#define DLP_DISPATCH_CPL(CPL, MACRO) \
  switch (CPL) {                     \
    case 1: MACRO(1); break;         \
    case 2: MACRO(2); break;         \
    case 3: MACRO(3); break;         \
    default: MACRO(4); break;        \
  }

namespace dlp {

// Thread-local error text returned by dlp_last_error().
void set_error(const char* fmt, ...);

#define DLP_CHECK_ARG(cond, ...)                        \
  do {                                                  \
    if (!(cond)) {                                      \
      ::dlp::set_error(__VA_ARGS__);                    \
      return DLP_ERR_ARG;                               \
    }                                                   \
  } while (0)

#define DLP_CHECK_HIP(expr)                                                    \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::dlp::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                       __FILE__, __LINE__);                                    \
      return DLP_ERR_HIP;                                                      \
    }                                                                          \
  } while (0)

#define DLP_CHECK_LAUNCH() DLP_CHECK_HIP(hipGetLastError())

constexpr int kWave = 64;
constexpr int kBlock = 256;

// A non-negative ratio with NaN mapped to +inf, so that max reductions (fmax drops NaN) and
// threshold tests (NaN compares false) treat a row that is not a number as violating.
__device__ __forceinline__ double nan_to_inf(double x) {
  return x <= 1.79769313486231570e308 ? x : __builtin_huge_val();
}

inline int ceil_log2(int x) {
  int s = 0;
  while ((1 << s) < x) ++s;
  return s;
}
inline int floor_log2(int x) {
  int s = 0;
  while ((2 << s) <= x) ++s;
  return s;
}
inline unsigned grid_for(long long threads, int block = kBlock) {
  long long g = (threads + block - 1) / block;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// Lane layout of a row group: G = 2^g_shift lanes per row, LC = 2^lc_shift column lanes,
// LE = G / LC edge lanes.  Column chunk k of a lane covers column l + k*LC.
struct RowMap {
  int g_shift, lc_shift, cpl;
};

// Column lanes: next power of two >= C, at most 64; edge lanes: about the mean degree,
// so that one row group keeps several neighbour rows in flight.
inline RowMap choose_map(int C, double avg_deg, int max_cpl = 4) {
  RowMap m;
  int lc = C <= 1 ? 0 : ceil_log2(C);
  if (lc > 6) lc = 6;
  m.lc_shift = lc;
  m.cpl = (C + (1 << lc) - 1) >> lc;
  if (m.cpl > max_cpl) m.cpl = max_cpl;  // caller checks C <= max_cpl*64
  int le = avg_deg >= 2 ? floor_log2((int)avg_deg) : 0;
  if (le + lc > 6) le = 6 - lc;
  if (le < 0) le = 0;
  // at least 4 lanes per row so index loads stay coalesced
  if (le + lc < 2) le = 2 - lc;
  m.g_shift = le + lc;
  return m;
}

__device__ __forceinline__ float to_acc(float x, float) { return x; }
__device__ __forceinline__ double to_acc(float x, double) { return (double)x; }
__device__ __forceinline__ double to_acc(double x, double) { return x; }

template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int m) { return __shfl_xor(v, m, 64); }

template <typename T>
__device__ __forceinline__ T shfl(T v, int src, int width) { return __shfl(v, src, width); }

// Hardware float atomics on device memory (global_atomic_add_f32/_f64).
__device__ __forceinline__ void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

__device__ __forceinline__ float atomic_take(float* p) {
  return __int_as_float(atomicExch((int*)p, 0));
}
__device__ __forceinline__ double atomic_take(double* p) {
  return __longlong_as_double((long long)atomicExch((unsigned long long*)p, 0ull));
}

// Non-negative float max via integer atomics (IEEE order matches for x >= 0).
__device__ __forceinline__ void atomic_max_nonneg(float* p, float v) {
  atomicMax((int*)p, __float_as_int(v));
}
__device__ __forceinline__ void atomic_max_nonneg(double* p, double v) {
  atomicMax((unsigned long long*)p, (unsigned long long)__double_as_longlong(v));
}

// Appends val to list for every active lane whose add is set, with one atomic per wavefront on
// the shared counter (the lanes' slots follow their order in the wavefront) instead of one per
// lane: lists that most rows of a pass join (changed rows of a batch, violators right after it)
// otherwise serialise every wavefront on the one counter.  Every active lane must call it.
__device__ __forceinline__ void wave_append(bool add, int val, int32_t* __restrict__ list,
                                            int32_t* __restrict__ count) {
  const unsigned long long m = __ballot(add);
  if (m == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __builtin_amdgcn_readfirstlane(__ffsll((long long)m) - 1);
  int base = 0;
  if (lane == leader) base = atomicAdd(count, (int)__popcll(m));
  base = __builtin_amdgcn_readlane(base, leader);
  if (add) list[base + (int)__popcll(m & ((1ull << lane) - 1ull))] = val;
}

// The same for a word in global memory that many workgroups send their maxima to: the atomic is
// skipped when the value already there is at least v (most workgroups, once the first few have
// written), which keeps thousands of atomics off the same L2 line.  A stale read only costs an
// atomic that changes nothing; NaN (any sign bit clear) still goes through, as before.
__device__ __forceinline__ void global_max_nonneg(float* p, float v) {
  if (!(v <= __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) atomic_max_nonneg(p, v);
}
__device__ __forceinline__ void global_max_nonneg(double* p, double v) {
  if (!(v <= __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) atomic_max_nonneg(p, v);
}

}  // namespace dlp

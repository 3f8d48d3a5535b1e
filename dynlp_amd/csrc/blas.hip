// Fused BLAS-1 passes of the column-wise PCG (solvers.py:66-97) over row-major n x C blocks.
//
// Every pass maps column lanes over a row (LC = 2^lc_shift lanes, CPL column chunks per lane)
// and rows over the grid, so each row is one coalesced access.  Column reductions accumulate
// in FP64: per lane, then per workgroup in LDS, then one FP64 atomic per column per workgroup.
#include "blas.h"
#include "rowpipe.h"
#include "scan.h"

namespace dlp {

struct ColMap {
  int lc_shift, cpl;
};

inline ColMap col_map(int C) {
  ColMap m;
  int lc = C <= 1 ? 0 : ceil_log2(C);
  if (lc > 6) lc = 6;
  m.lc_shift = lc;
  m.cpl = (C + (1 << lc) - 1) >> lc;
  return m;
}

constexpr int kMaxCols = 256;

__device__ __forceinline__ void lds_add(double* p, double v) { atomicAdd(p, v); }

// out[c] += sum_r X[r,c] * Y[r,c] * (scale_by_diag ? 1/diag[r] : 1)
template <int CPL>
__global__ __launch_bounds__(256) void k_coldot(int n, int C, int lc_shift, const float* __restrict__ X,
                                                const float* __restrict__ Y, int64_t ld,
                                                double* __restrict__ out) {
  __shared__ double sm[kMaxCols];
  for (int c = threadIdx.x; c < C; c += blockDim.x) sm[c] = 0.0;
  __syncthreads();
  const int LC = 1 << lc_shift;
  const int l = threadIdx.x & (LC - 1);
  double acc[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) acc[k] = 0.0;
  const int64_t groups = ((int64_t)gridDim.x * blockDim.x) >> lc_shift;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> lc_shift; r < n; r += groups) {
    const int64_t o = r * ld;
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = l + k * LC;
      if (c < C) acc[k] += (double)X[o + c] * (double)Y[o + c];
    }
  }
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int c = l + k * LC;
    if (c < C) lds_add(&sm[c], acc[k]);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) atomicAdd(&out[c], sm[c]);
}

// PCG step over all rows: Z += a P, R -= a AP; then Y = R/diag and
// rz_new[c] = sum R*Y, viol = max |R|/(eps_c*diag) (float bits, atomic max).
template <int CPL>
__global__ __launch_bounds__(256) void k_pcg_update(int n, int C, int lc_shift, float* __restrict__ Z,
                                                    float* __restrict__ R, const float* __restrict__ P,
                                                    const float* __restrict__ AP, int64_t ld,
                                                    const double* __restrict__ alpha,
                                                    const float* __restrict__ diag,
                                                    const float* __restrict__ eps,
                                                    double* __restrict__ rz_new, float* __restrict__ viol,
                                                    const int32_t* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double sm[kMaxCols];
  __shared__ float vmax;
  for (int c = threadIdx.x; c < C; c += blockDim.x) sm[c] = 0.0;
  if (threadIdx.x == 0) vmax = 0.f;
  __syncthreads();
  const int LC = 1 << lc_shift;
  const int l = threadIdx.x & (LC - 1);
  float a[CPL], ie[CPL];
  double acc[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int c = l + k * LC;
    a[k] = c < C ? (float)alpha[c] : 0.f;
    ie[k] = c < C ? 1.f / eps[c] : 0.f;
    acc[k] = 0.0;
  }
  float vm = 0.f;
  const int64_t groups = ((int64_t)gridDim.x * blockDim.x) >> lc_shift;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> lc_shift; r < n; r += groups) {
    const int64_t o = r * ld;
    const float inv = 1.f / diag[r];
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = l + k * LC;
      if (c < C) {
        Z[o + c] += a[k] * P[o + c];
        const float rv = R[o + c] - a[k] * AP[o + c];
        R[o + c] = rv;
        acc[k] += (double)rv * (double)rv * (double)inv;
        vm = fmaxf(vm, fabsf(rv) * inv * ie[k]);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int c = l + k * LC;
    if (c < C) lds_add(&sm[c], acc[k]);
  }
  for (int off = 32; off > 0; off >>= 1) vm = fmaxf(vm, __shfl_xor(vm, off, 64));
  if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&vmax, vm);
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) atomicAdd(&rz_new[c], sm[c]);
  if (threadIdx.x == 0) global_max_nonneg(viol, vmax);
}

// P = R/diag + beta P   (P = R/diag when first != 0)
template <int CPL>
__global__ __launch_bounds__(256) void k_pcg_dir(int n, int C, int lc_shift, float* __restrict__ P,
                                                 const float* __restrict__ R, int64_t ld,
                                                 const float* __restrict__ diag,
                                                 const double* __restrict__ beta, int first,
                                                 const int32_t* __restrict__ skip) {
  if (skip && *skip) return;
  const int LC = 1 << lc_shift;
  const int l = threadIdx.x & (LC - 1);
  float b[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int c = l + k * LC;
    b[k] = (c < C && !first) ? (float)beta[c] : 0.f;
  }
  const int64_t groups = ((int64_t)gridDim.x * blockDim.x) >> lc_shift;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> lc_shift; r < n; r += groups) {
    const int64_t o = r * ld;
    const float inv = 1.f / diag[r];
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = l + k * LC;
      if (c < C) P[o + c] = R[o + c] * inv + (first ? 0.f : b[k] * P[o + c]);
    }
  }
}

// ---- Flat 16-byte passes (ld % 4 == 0, 16-byte aligned blocks) ---------------------------
// The n x ld block is a flat array of n*Q4 float4 (Q4 = ld/4).  A workgroup uses Q4 * (256/Q4)
// of its 256 threads and the grid stride is a multiple of Q4, so every thread keeps one column quad
// (threadIdx % Q4) for the whole pass, consecutive threads read consecutive 16-byte words, and
// no lane idles on a power-of-two column layout.  Quad components at columns >= C are left
// unchanged and excluded from the reductions.

// The direction block P of the flat passes (PK): 0 = FP32 (the n x Q4 float4 layout of Z and
// R); 1 = BF16 or 2 = FP16 (rows of H4 groups of 4 values, 8 bytes each; stores round to
// nearest even, FP16 saturating); 3 = FP8 E4M3 (rows of H4 groups of 4 values, 4 bytes each;
// nearest even, saturating).  FP16 and FP8 directions are held scaled by a power of two per
// column (the caller passes the scaled values), which keeps them inside the format's range.
template <int PK>
struct PBlock {
  void* base;
  int64_t H4;
  __device__ __forceinline__ float4 load(int64_t i, int64_t r, int q4) const {
    if constexpr (PK == 1) {
      const uint2 u = reinterpret_cast<const uint2*>(base)[r * H4 + q4];
      return make_float4(bf16_lo(u.x), bf16_hi(u.x), bf16_lo(u.y), bf16_hi(u.y));
    } else if constexpr (PK == 2) {
      const uint2 u = reinterpret_cast<const uint2*>(base)[r * H4 + q4];
      return make_float4(f16_lo(u.x), f16_hi(u.x), f16_lo(u.y), f16_hi(u.y));
    } else if constexpr (PK == 3) {
      return f8_unpack4(reinterpret_cast<const unsigned*>(base)[r * H4 + q4]);
    } else {
      return reinterpret_cast<const float4*>(base)[i];
    }
  }
  // stores v and returns the value held (rounded in a 16-bit block)
  __device__ __forceinline__ float4 store(int64_t i, int64_t r, int q4, const float4& v) const {
    if constexpr (PK == 1) {
      uint2 u;
      u.x = bf16_pack(v.x, v.y);
      u.y = bf16_pack(v.z, v.w);
      reinterpret_cast<uint2*>(base)[r * H4 + q4] = u;
      return make_float4(bf16_lo(u.x), bf16_hi(u.x), bf16_lo(u.y), bf16_hi(u.y));
    } else if constexpr (PK == 2) {
      uint2 u;
      u.x = f16_pack(v.x, v.y);
      u.y = f16_pack(v.z, v.w);
      reinterpret_cast<uint2*>(base)[r * H4 + q4] = u;
      return make_float4(f16_lo(u.x), f16_hi(u.x), f16_lo(u.y), f16_hi(u.y));
    } else if constexpr (PK == 3) {
      const unsigned u = f8_pack4(v.x, v.y, v.z, v.w);
      reinterpret_cast<unsigned*>(base)[r * H4 + q4] = u;
      return f8_unpack4(u);
    } else {
      reinterpret_cast<float4*>(base)[i] = v;
      return v;
    }
  }
};

// Column sums of per-thread quads: LDS, then one FP64 atomic per column per workgroup.
__device__ __forceinline__ void flush_quads(const Quad& Qd, int Q4, int C, const double (&acc)[4], double* sm,
                                            double* __restrict__ out) {
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (j < Qd.cnt) lds_add(&sm[4 * Qd.q4 + j], acc[j]);
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) atomicAdd(&out[c], sm[c]);
}

// Column maxima of non-negative per-thread quad values: LDS max (smx cleared beforehand), then,
// after a barrier the caller places, the workgroup's maxima into out (quad_max_out).
__device__ __forceinline__ void quad_max_lds(int q4, int cnt, const float (&m)[4], float* smx) {
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (j < cnt) atomic_max_nonneg(&smx[4 * q4 + j], m[j]);
}
__device__ __forceinline__ void quad_max_out(int C, const float* smx, float* __restrict__ out) {
  for (int c = threadIdx.x; c < C; c += blockDim.x) global_max_nonneg(&out[c], smx[c]);
}

// Loads and stores of the FP32 vector streams of the flat passes; NT marks them non-temporal so
// the caches keep the direction block the next operator gathers.
template <bool NT>
__device__ __forceinline__ float4 ld4(const float4* p) {
  if constexpr (NT) {
    const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
  } else {
    return *p;
  }
}
template <bool NT>
__device__ __forceinline__ void st4(float4* p, const float4& v) {
  if constexpr (NT) {
    const f4v u = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(u, reinterpret_cast<f4v*>(p));
  } else {
    *p = v;
  }
}

// Two-level preconditioner arguments of the flat passes (na = 0: Jacobi only).
struct CoarseArg {
  int na;
  int64_t ldc;
  double* rc;               // P^T r, na x ldc (restricting passes)
  const float* ec;          // Ac^-1 P^T r, na x ldc (prolonging passes)
  const uint8_t* solved;
  const int32_t* bounds;    // aggregate I = rows [bounds[I], bounds[I+1])
  const uint8_t* aggmap;    // aggregate of every row
};

// Rows of this workgroup in a contiguous split of n rows over the grid, [r_lo, r_hi); a thread
// keeps column quad q4 and walks rows r_lo + roff, stepping rstep (consecutive threads still
// read consecutive 16-byte words).
struct QuadRange {
  bool active;
  int q4, cnt, roff, rstep;
  int64_t r_lo, r_hi;
};

__device__ __forceinline__ QuadRange quad_range(int Q4, int C, int64_t n) {
  QuadRange q;
  q.rstep = (int)blockDim.x / Q4;
  q.active = (int)threadIdx.x < Q4 * q.rstep;
  q.q4 = threadIdx.x % Q4;
  q.roff = threadIdx.x / Q4;
  q.cnt = q.active ? max(0, min(4, C - 4 * q.q4)) : 0;
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  q.r_lo = min(n, (int64_t)blockIdx.x * chunk);
  q.r_hi = min(n, q.r_lo + chunk);
  return q;
}

// Column sums of a workgroup's rows per aggregate: the caller keeps the workgroup's rows within
// two aggregates (rows per workgroup <= the smallest aggregate), so a thread holds two quads.
struct Restrict {
  int64_t A0, bnd;
  double a0[4], a1[4];
  __device__ __forceinline__ void init(int64_t r_lo, int64_t n, const CoarseArg& K) {
    A0 = 0;
    bnd = INT64_MAX;
    if (K.na && r_lo < n) {
      A0 = K.aggmap[r_lo];
      bnd = K.bounds[A0 + 1];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) a0[j] = a1[j] = 0.0;
  }
  __device__ __forceinline__ void add(int64_t r, int j, double v) {
    if (r < bnd) a0[j] += v;
    else a1[j] += v;
  }
  // LDS part (before the workgroup barrier): sm2 holds 2 x 4*Q4 doubles, zeroed beforehand
  __device__ __forceinline__ void to_lds(const QuadRange& q, int Q4, double* sm2) const {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j < q.cnt) {
        lds_add(&sm2[4 * q.q4 + j], a0[j]);
        lds_add(&sm2[4 * Q4 + 4 * q.q4 + j], a1[j]);
      }
    }
  }
  // global part (after the barrier)
  __device__ __forceinline__ void to_global(int Q4, int C, const CoarseArg& K, const double* sm2) const {
    for (int c = threadIdx.x; c < 8 * Q4; c += blockDim.x) {
      const int h = c / (4 * Q4), cc = c - h * 4 * Q4;
      const int64_t A = A0 + h;
      if (cc < C && A < K.na && sm2[c] != 0.0) atomicAdd(&K.rc[A * K.ldc + cc], sm2[c]);
    }
  }
};

// The coarse correction of row r (column quad q4), zero outside the solved set.
__device__ __forceinline__ float4 coarse_quad(const CoarseArg& K, int64_t r, int q4) {
  if (K.na == 0 || !K.solved[r]) return make_float4(0.f, 0.f, 0.f, 0.f);
  return reinterpret_cast<const float4*>(K.ec + (int64_t)K.aggmap[r] * K.ldc)[q4];
}

// pcg start: P = M R (M R = R/diag, plus the coarse correction on solved rows), rz[c] += sum
// R*(M R); with a 16-bit or 8-bit block also rp[c] += sum R*P as held.  FP16 and FP8 blocks hold
// sig[c] * M R (sig from a prior k_pcg_rz4 pass); FP8 blocks also record pmax[c] = max |P| as
// held.
template <int PK>
__global__ __launch_bounds__(256) void k_pcg_start4(int n, int Q4, int C, PBlock<PK> P, const float4* __restrict__ R,
                                                    const float* __restrict__ diag, double* __restrict__ rz,
                                                    double* __restrict__ rp, const float* __restrict__ sig,
                                                    CoarseArg K, float* __restrict__ pmax) {
  __shared__ double sm[kMaxCols + 4];
  __shared__ double sm2[kMaxCols + 4];
  __shared__ float smx[PK == 3 ? kMaxCols + 4 : 1];
  for (int c = threadIdx.x; c < 4 * Q4; c += blockDim.x) {
    sm[c] = sm2[c] = 0.0;
    if constexpr (PK == 3) smx[c] = 0.f;
  }
  __syncthreads();
  const Quad Qd = quad_of(Q4, C);
  float sg[4] = {1.f, 1.f, 1.f, 1.f};
  if constexpr (PK >= 2) {
#pragma unroll
    for (int j = 0; j < 4; ++j) sg[j] = j < Qd.cnt ? sig[4 * Qd.q4 + j] : 1.f;
  }
  const int64_t nq = (int64_t)n * Q4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0}, acc2[4] = {0.0, 0.0, 0.0, 0.0};
  float pm[4] = {0.f, 0.f, 0.f, 0.f};
  if (Qd.active && Qd.r0 < n) {
    int64_t r = Qd.r0;
    for (int64_t i = Qd.i0; i < nq; i += Qd.step, r += Qd.rstep) {
      const float inv = 1.f / diag[r];
      const float4 rv = R[i];
      const float4 ev = coarse_quad(K, r, Qd.q4);
      float4 y = P.load(i, r, Qd.q4);
      y.x = Qd.cnt > 0 ? rv.x * inv + ev.x : y.x;
      y.y = Qd.cnt > 1 ? rv.y * inv + ev.y : y.y;
      y.z = Qd.cnt > 2 ? rv.z * inv + ev.z : y.z;
      y.w = Qd.cnt > 3 ? rv.w * inv + ev.w : y.w;
      const float4 p = P.store(i, r, Qd.q4, make_float4(y.x * sg[0], y.y * sg[1], y.z * sg[2], y.w * sg[3]));
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j < Qd.cnt) {
          acc[j] += (double)quad_comp(rv, j) * (double)quad_comp(y, j);
          if constexpr (PK != 0) acc2[j] += (double)quad_comp(rv, j) * (double)quad_comp(p, j);
          if constexpr (PK == 3) pm[j] = fmaxf(pm[j], fabsf(quad_comp(p, j)));
        }
      }
    }
  }
  if constexpr (PK == 3)
    if (Qd.active) quad_max_lds(Qd.q4, Qd.cnt, pm, smx);
  flush_quads(Qd, Q4, C, acc, sm, rz);   // (its barrier also orders the LDS maxima)
  if constexpr (PK != 0) flush_quads(Qd, Q4, C, acc2, sm2, rp);
  if constexpr (PK == 3) quad_max_out(C, smx, pmax);
}

// rz[c] += sum R*R/diag, and with K the restriction rc += P^T R (the pass before a start that
// needs the column scales or the coarse correction); with rmax also rmax[c] = max |R|/diag.
__global__ __launch_bounds__(256) void k_pcg_rz4(int n, int Q4, int C, const float4* __restrict__ R,
                                                 const float* __restrict__ diag, double* __restrict__ rz, CoarseArg K,
                                                 float* __restrict__ rmax) {
  __shared__ double sm[kMaxCols + 4];
  __shared__ double sm2[2 * (kMaxCols + 4)];
  __shared__ float smx[kMaxCols + 4];
  for (int c = threadIdx.x; c < 4 * Q4; c += blockDim.x) {
    sm[c] = sm2[c] = sm2[4 * Q4 + c] = 0.0;
    smx[c] = 0.f;
  }
  __syncthreads();
  const QuadRange q = quad_range(Q4, C, n);
  Restrict rs;
  rs.init(q.r_lo, n, K);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  float rm[4] = {0.f, 0.f, 0.f, 0.f};
  if (q.active) {
    for (int64_t r = q.r_lo + q.roff; r < q.r_hi; r += q.rstep) {
      const double inv = 1.0 / (double)diag[r];
      const float4 rv = R[r * Q4 + q.q4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j < q.cnt) {
          const double x = (double)quad_comp(rv, j);
          acc[j] += x * x * inv;
          rm[j] = fmaxf(rm[j], (float)(fabs(x) * inv));
          if (K.na) rs.add(r, j, x);
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (j < q.cnt) lds_add(&sm[4 * q.q4 + j], acc[j]);
  if (rmax) quad_max_lds(q.q4, q.cnt, rm, smx);
  if (K.na) rs.to_lds(q, Q4, sm2);
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) atomicAdd(&rz[c], sm[c]);
  if (rmax) quad_max_out(C, smx, rmax);
  if (K.na) rs.to_global(Q4, C, K, sm2);
}

// Z += a P, R -= a AP with the step length a = num/pAp of every column (num = r.p for 16-bit
// directions, r.z otherwise; 0 where pAp <= 0); rz_new[c] += sum R*R/diag; viol = max
// |R|/(eps_c*diag); with K also rc += P^T R (the new residual restricted to the aggregates);
// with FP8 directions (PK 3) also rmax[c] = max |R|/diag (the next direction's scale bound).
// Rows in a contiguous split.
// DEFER: Z and P are not touched; the step lengths go to alpha_out (C floats, written by the
// first workgroup; zeros when the iteration is skipped) for the group's deferred update.
template <int PK, bool NT = false, bool DEFER = false>
__global__ __launch_bounds__(256) void k_pcg_update4(int n, int Q4, int C, float4* __restrict__ Z,
                                                     float4* __restrict__ R, PBlock<PK> P,
                                                     const float4* __restrict__ AP,
                                                     const double* __restrict__ pAp, const double* __restrict__ num,
                                                     const float* __restrict__ diag, const float* __restrict__ eps,
                                                     double* __restrict__ rz_new, float* __restrict__ viol,
                                                     const int32_t* __restrict__ skip, CoarseArg K,
                                                     float* __restrict__ alpha_out, float* __restrict__ rmax) {
  if (skip && *skip) {
    if constexpr (DEFER) {
      if (blockIdx.x == 0)
        for (int c = threadIdx.x; c < C; c += blockDim.x) alpha_out[c] = 0.f;
    }
    return;
  }
  __shared__ double sm[kMaxCols + 4];
  __shared__ double sm2[2 * (kMaxCols + 4)];
  __shared__ float vmax;
  __shared__ float smx[PK == 3 ? kMaxCols + 4 : 1];
  for (int c = threadIdx.x; c < 4 * Q4; c += blockDim.x) {
    sm[c] = sm2[c] = sm2[4 * Q4 + c] = 0.0;
    if constexpr (PK == 3) smx[c] = 0.f;
  }
  if (threadIdx.x == 0) vmax = 0.f;
  __syncthreads();
  const QuadRange q = quad_range(Q4, C, n);
  Restrict rs;
  rs.init(q.r_lo, n, K);
  float a[4], ie[4], rm[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int c = 4 * q.q4 + j;
    const double pa = j < q.cnt ? pAp[c] : 0.0;
    a[j] = pa > 0.0 ? (float)(num[c] / pa) : 0.f;
    ie[j] = j < q.cnt ? 1.f / eps[c] : 0.f;
  }
  if constexpr (DEFER) {
    if (blockIdx.x == 0 && q.active && q.roff == 0) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j < q.cnt) alpha_out[4 * q.q4 + j] = a[j];
    }
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  float vm = 0.f;
  if (q.active) {
    for (int64_t r = q.r_lo + q.roff; r < q.r_hi; r += q.rstep) {
      const int64_t i = r * Q4 + q.q4;
      const float inv = 1.f / diag[r];
      float4 rv = ld4<NT>(R + i);
      const float4 ap = ld4<NT>(AP + i);
      if constexpr (!DEFER) {
        float4 z = ld4<NT>(Z + i);
        const float4 p = P.load(i, r, q.q4);
        if (q.cnt > 0) z.x += a[0] * p.x;
        if (q.cnt > 1) z.y += a[1] * p.y;
        if (q.cnt > 2) z.z += a[2] * p.z;
        if (q.cnt > 3) z.w += a[3] * p.w;
        st4<NT>(Z + i, z);
      }
      if (q.cnt > 0) rv.x -= a[0] * ap.x;
      if (q.cnt > 1) rv.y -= a[1] * ap.y;
      if (q.cnt > 2) rv.z -= a[2] * ap.z;
      if (q.cnt > 3) rv.w -= a[3] * ap.w;
      st4<NT>(R + i, rv);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j < q.cnt) {
          const float x = quad_comp(rv, j);
          acc[j] += (double)x * (double)x * (double)inv;
          const float t = fabsf(x) * inv;
          vm = fmaxf(vm, t * ie[j]);
          if constexpr (PK == 3) rm[j] = fmaxf(rm[j], t);
          if (K.na) rs.add(r, j, (double)x);
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (j < q.cnt) lds_add(&sm[4 * q.q4 + j], acc[j]);
  if (K.na) rs.to_lds(q, Q4, sm2);
  if constexpr (PK == 3) quad_max_lds(q.q4, q.cnt, rm, smx);
  for (int off = 32; off > 0; off >>= 1) vm = fmaxf(vm, __shfl_xor(vm, off, 64));
  if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&vmax, vm);
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) atomicAdd(&rz_new[c], sm[c]);
  if constexpr (PK == 3) quad_max_out(C, smx, rmax);
  if (K.na) rs.to_global(Q4, C, K, sm2);
  if (threadIdx.x == 0) global_max_nonneg(viol, vmax);
}

// Pout = M R + beta P (M R = R/diag plus the coarse correction on solved rows; Pout may be P);
// with a 16-bit block also rp[c] += sum R*Pout as held (the next step length's numerator, exact
// for the rounded direction).  FP16 and FP8 blocks: Pout = sig*M R + beta*P with sig the new
// column scale and beta already rescaled to the held direction; FP8 blocks also record
// pmax[c] = max |Pout| as held (cleared beforehand).
template <int PK, bool NT = false>
__global__ __launch_bounds__(256) void k_pcg_dir4(int n, int Q4, int C, PBlock<PK> P, PBlock<PK> Pout,
                                                  const float4* __restrict__ R,
                                                  const float* __restrict__ diag, const double* __restrict__ beta,
                                                  double* __restrict__ rp, const int32_t* __restrict__ skip,
                                                  const float* __restrict__ sig, CoarseArg K,
                                                  float* __restrict__ pmax) {
  if (skip && *skip) return;
  __shared__ double sm[kMaxCols + 4];
  __shared__ float smx[PK == 3 ? kMaxCols + 4 : 1];
  if constexpr (PK != 0) {
    for (int c = threadIdx.x; c < 4 * Q4; c += blockDim.x) {
      sm[c] = 0.0;
      if constexpr (PK == 3) smx[c] = 0.f;
    }
    __syncthreads();
  }
  const Quad Qd = quad_of(Q4, C);
  float b[4], sg[4] = {1.f, 1.f, 1.f, 1.f};
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = j < Qd.cnt ? (float)beta[4 * Qd.q4 + j] : 0.f;
  if constexpr (PK >= 2) {
#pragma unroll
    for (int j = 0; j < 4; ++j) sg[j] = j < Qd.cnt ? sig[4 * Qd.q4 + j] : 1.f;
  }
  const int64_t nq = (int64_t)n * Q4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  float pm[4] = {0.f, 0.f, 0.f, 0.f};
  if (Qd.active && Qd.r0 < n) {
    int64_t r = Qd.r0;
    for (int64_t i = Qd.i0; i < nq; i += Qd.step, r += Qd.rstep) {
      const float inv = 1.f / diag[r];
      const float4 rv = ld4<NT>(R + i);
      const float4 ev = coarse_quad(K, r, Qd.q4);
      float4 p = P.load(i, r, Qd.q4);
      if constexpr (PK >= 2) {
        if (Qd.cnt > 0) p.x = (rv.x * inv + ev.x) * sg[0] + b[0] * p.x;
        if (Qd.cnt > 1) p.y = (rv.y * inv + ev.y) * sg[1] + b[1] * p.y;
        if (Qd.cnt > 2) p.z = (rv.z * inv + ev.z) * sg[2] + b[2] * p.z;
        if (Qd.cnt > 3) p.w = (rv.w * inv + ev.w) * sg[3] + b[3] * p.w;
      } else {
        if (Qd.cnt > 0) p.x = rv.x * inv + ev.x + b[0] * p.x;
        if (Qd.cnt > 1) p.y = rv.y * inv + ev.y + b[1] * p.y;
        if (Qd.cnt > 2) p.z = rv.z * inv + ev.z + b[2] * p.z;
        if (Qd.cnt > 3) p.w = rv.w * inv + ev.w + b[3] * p.w;
      }
      const float4 held = Pout.store(i, r, Qd.q4, p);
      if constexpr (PK != 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (j < Qd.cnt) {
            acc[j] += (double)quad_comp(rv, j) * (double)quad_comp(held, j);
            if constexpr (PK == 3) pm[j] = fmaxf(pm[j], fabsf(quad_comp(held, j)));
          }
        }
      }
    }
  }
  if constexpr (PK == 3)
    if (Qd.active) quad_max_lds(Qd.q4, Qd.cnt, pm, smx);
  if constexpr (PK != 0) flush_quads(Qd, Q4, C, acc, sm, rp);   // (its barrier also orders the maxima)
  if constexpr (PK == 3) quad_max_out(C, smx, pmax);
}

// The deferred solution update of a group of G <= kMaxGroup iterations: Z += sum_s alpha[s] *
// P_s over the direction slots s < G (slot s = P0 + s*stride), in the order of the iterations,
// with the same FP32 multiply-adds the per-iteration update makes.  Iterations skipped after
// convergence hold zero step lengths and come last: the workgroup reads only the slots up to
// the last one with a nonzero step length, and each thread issues all its slot loads before the
// first multiply-add.  Unless the solve is done (ctl[0]), the next group's first direction
// (slot G) moves to slot 0.

template <int PK, bool NT = false>
__global__ __launch_bounds__(256) void k_pcg_flush4(int n, int Q4, int C, float4* __restrict__ Z, PBlock<PK> P0,
                                                    int64_t stride, int G, const float* __restrict__ alpha,
                                                    const int32_t* __restrict__ ctl) {
  __shared__ int used;
  if (threadIdx.x == 0) used = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < G * C; k += blockDim.x)
    if (alpha[k] != 0.f) atomicMax(&used, k / C + 1);
  __syncthreads();
  const int S = used;
  const Quad Qd = quad_of(Q4, C);
  if (!Qd.active || Qd.r0 >= n) return;
  const bool copy = !(ctl && ctl[0]);
  float a[kMaxGroup][4];
#pragma unroll
  for (int g = 0; g < kMaxGroup; ++g)
#pragma unroll
    for (int j = 0; j < 4; ++j) a[g][j] = (g < S && j < Qd.cnt) ? alpha[g * C + 4 * Qd.q4 + j] : 0.f;
  const int64_t nq = (int64_t)n * Q4;
  int64_t r = Qd.r0;
  for (int64_t i = Qd.i0; i < nq; i += Qd.step, r += Qd.rstep) {
    float4 p[kMaxGroup];
#pragma unroll
    for (int g = 0; g < kMaxGroup; ++g) {
      if (g < S) {
        PBlock<PK> Ps = P0;
        Ps.base = reinterpret_cast<uint16_t*>(P0.base) + g * stride;
        p[g] = Ps.load(i, r, Qd.q4);
      }
    }
    if (S > 0) {
      float4 z = ld4<NT>(Z + i);
#pragma unroll
      for (int g = 0; g < kMaxGroup; ++g) {
        if (g < S) {
          if (Qd.cnt > 0) z.x += a[g][0] * p[g].x;
          if (Qd.cnt > 1) z.y += a[g][1] * p[g].y;
          if (Qd.cnt > 2) z.z += a[g][2] * p[g].z;
          if (Qd.cnt > 3) z.w += a[g][3] * p[g].w;
        }
      }
      st4<NT>(Z + i, z);
    }
    if (copy) {
      PBlock<PK> last = P0;
      last.base = reinterpret_cast<uint16_t*>(P0.base) + G * stride;
      P0.store(i, r, Qd.q4, last.load(i, r, Qd.q4));
    }
  }
}

struct FlatLaunch {
  bool ok;
  int Q4;
  unsigned block, grid;
};

static FlatLaunch flat_launch(const PcgBufs& b) {
  FlatLaunch f{false, 0, 0, 0};
  const uintptr_t al = (uintptr_t)b.Z | (uintptr_t)b.R | (uintptr_t)(b.P16 ? (void*)b.P16 : (void*)b.P) | (uintptr_t)b.AP;
  // the tuning switch for 16-byte loads covers FP32 directions; BF16 directions exist only in
  // this layout (checked when the mode is set)
  if ((b.ld % 4) != 0 || (al & 15) != 0 || b.ld / 4 > 64 || (!tuning().vec4 && !b.P16)) return f;
  f.Q4 = (int)(b.ld / 4);
  f.block = kBlock;
  const long long used = f.Q4 * (kBlock / f.Q4);
  const long long nq = (long long)b.n * f.Q4;
  long long g = (nq + used - 1) / used;
  const long long cap = tuning().flat_grid > 0 ? tuning().flat_grid : 2048;
  if (g > cap) g = cap;
  f.grid = (unsigned)(g < 1 ? 1 : g);
  f.ok = true;
  return f;
}

// Power-of-two scale of an FP16 direction column from its rz = sum R*R/diag: the held values
// sig*R/diag stay within about sqrt(1/diag) of 1 (far inside FP16's range); 1 when rz = 0.
__device__ __forceinline__ float pcg_scale(double rz) {
  if (!(rz > 0.0)) return 1.f;
  double e = rint(-0.5 * log2(rz));
  e = fmin(fmax(e, -100.0), 100.0);
  return (float)ldexp(1.0, (int)e);
}

// Power-of-two scale of an FP8 direction column from a bound on the largest magnitude of the
// unscaled direction: the held values stay within 224, half of E4M3's largest finite value, so
// no value saturates (FP8 has too few exponent bits for the rz-based FP16 scale); 1 when the
// bound is 0 (the column is converged).
__device__ __forceinline__ float f8_scale(double bound) {
  if (!(bound > 0.0)) return 1.f;
  double e = floor(log2(224.0 / bound));
  e = fmin(fmax(e, -100.0), 100.0);
  return (float)ldexp(1.0, (int)e);
}

// The new scale of an FP8 direction column, P_new = M r + beta P_old: max |M r| <= rmax + emax
// (rmax = max |r|/diag, emax = max |coarse correction|) and max |P_old| = pmax / sig_old (pmax =
// the held maximum); the maxima are cleared for the next iteration.
__device__ __forceinline__ float f8_next_scale(int c, double beta, float emax, float* rmax, float* pmax,
                                               const float* sig) {
  const double bound = (double)rmax[c] + (double)emax + fabs(beta) * (double)pmax[c] / (double)sig[c];
  rmax[c] = 0.f;
  pmax[c] = 0.f;
  return f8_scale(bound);
}

// Scalars of one iteration (one workgroup): mode 0: alpha = rz/pAp (0 where pAp <= 0), then
// clear rz_new and viol for the update pass; mode 1: beta = rz_new/rz (0 where rz <= 0),
// rz = rz_new, then clear rz_new, pAp, rp and (after the convergence test, keeping its value in
// viol[1] for the host) viol for the next iteration; with sig (FP16 or FP8 directions) also the
// new column scale, and beta rescaled to the held direction (beta * sig_new / sig_old).  Mode 0
// (the generic passes) also clears rc (nrc doubles).  Mode 2 (before a start that had a scale
// pass): sig from rz (when sig), then rz = 0.  FP8 directions (rmax non-null) take their scale
// from the maxima instead (rmax, pmax; emax in mode 2, the start's coarse correction).
__global__ void k_pcg_scalars(int C, int mode, double* __restrict__ rz, double* __restrict__ rz_new,
                              double* __restrict__ pAp, double* __restrict__ alpha,
                              double* __restrict__ beta, float* __restrict__ viol, int32_t* ctl,
                              double* __restrict__ rp, float* __restrict__ sig, double* __restrict__ rc, int nrc,
                              float* __restrict__ rmax, float* __restrict__ pmax, float* __restrict__ emax) {
  // every thread reads the flag before thread 0 may set it below (after the barrier)
  if (ctl && ctl[0]) return;
  if (mode == 0 && rc)
    for (int i = threadIdx.x; i < nrc; i += blockDim.x) rc[i] = 0.0;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    if (mode == 0) {
      // step length along P as held: r.p / p.Ap (r.p = rz for an exact direction)
      const double p = pAp[c];
      alpha[c] = p > 0.0 ? (rp ? rp[c] : rz[c]) / p : 0.0;
      rz_new[c] = 0.0;
    } else if (mode == 1) {
      const double r = rz[c];
      double bt = r > 0.0 ? rz_new[c] / r : 0.0;
      if (sig) {
        const float s = rmax ? f8_next_scale(c, bt, 0.f, rmax, pmax, sig) : pcg_scale(rz_new[c]);
        bt *= (double)s / (double)sig[c];
        sig[c] = s;
      }
      beta[c] = bt;
      rz[c] = rz_new[c];
      rz_new[c] = 0.0;
      pAp[c] = 0.0;
      if (rp) rp[c] = 0.0;
    } else {
      if (sig && rmax) {
        sig[c] = f8_scale((double)rmax[c] + (double)(emax ? emax[c] : 0.f));
        rmax[c] = 0.f;
        pmax[c] = 0.f;
        if (emax) emax[c] = 0.f;
      } else if (sig) {
        sig[c] = pcg_scale(rz[c]);
      }
      rz[c] = 0.0;
    }
  }
  if (mode == 0 && threadIdx.x == 0) *viol = 0.f;
  if (mode == 1) {
    __syncthreads();
    if (threadIdx.x == 0) {
      if (ctl) {
        ctl[1] += 1;
        if (viol[0] <= 1.f) ctl[0] = 1;
      }
      viol[1] = viol[0];   // the last violation, for the host
      viol[0] = 0.f;
    }
  }
}

// ---- two-level preconditioner: coarse operator, its inverse, and the coarse correction ----

// ec[I][c] = sum_J Ainv[J][I] rcs[J] for thread I < na (Ainv symmetric, rcs = rc[:, c] in LDS);
// returns sum_I rcs[I] ec[I][c] in thread 0 (the workgroup's threads all call it).
// With emax, thread 0 also gets max_I |ec[I][c]| there.
__device__ __forceinline__ double coarse_column(int na, int c, int64_t ldc, const double* __restrict__ Ainv,
                                                const double* rcs, float* __restrict__ ec, double* part,
                                                float* partm = nullptr, float* emax = nullptr) {
  const int I = threadIdx.x;
  double v = 0.0;
  float m = 0.f;
  if (I < na) {
    double e = 0.0;
#pragma unroll 8
    for (int J = 0; J < na; ++J) e += Ainv[(int64_t)J * na + I] * rcs[J];
    const float ef = (float)e;
    ec[(int64_t)I * ldc + c] = ef;
    v = rcs[I] * e;
    m = fabsf(ef);
  }
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  if (emax) {
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) partm[threadIdx.x >> 6] = m;
  }
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    for (int k = 0; k < ((int)blockDim.x >> 6); ++k) t += part[k];
    if (emax) {
      float mm = 0.f;
      for (int k = 0; k < ((int)blockDim.x >> 6); ++k) mm = fmaxf(mm, partm[k]);
      *emax = mm;
    }
  }
  return t;
}

// ec[:, c] = Ainv rc[:, c] (one workgroup of kCoarseMax threads per column c < C, thread I =
// aggregate, rc[:, c] staged in LDS, then cleared), and rz[c] += rc[:, c] . ec[:, c] (the
// coarse part of r^T M r) -- the start's correction; with emax also emax[c] = max_I |ec[I][c]|.
__global__ __launch_bounds__(kCoarseMax) void k_coarse_apply(int na, int C, int64_t ldc,
                                                             const double* __restrict__ Ainv,
                                                             double* __restrict__ rc, float* __restrict__ ec,
                                                             double* __restrict__ rz, const int32_t* __restrict__ ctl,
                                                             float* __restrict__ emax) {
  if (ctl && ctl[0]) return;
  __shared__ double rcs[kCoarseMax];
  __shared__ double part[kCoarseMax / 64];
  __shared__ float partm[kCoarseMax / 64];
  const int c = blockIdx.x, I = threadIdx.x;
  rcs[I] = I < na ? rc[(int64_t)I * ldc + c] : 0.0;
  __syncthreads();
  if (I < na) rc[(int64_t)I * ldc + c] = 0.0;   // ready for the next restriction
  const double v = coarse_column(na, c, ldc, Ainv, rcs, ec, part, partm, emax ? emax + c : nullptr);
  if (threadIdx.x == 0) atomicAdd(&rz[c], v);
}

// One iteration's scalars with the coarse correction (one workgroup of kCoarseMax threads per
// column c): ec[:, c] = Ainv rc[:, c]; rz_c = rz_new[c] + rc . ec (r^T M r); beta = rz_c/rz[c]
// (with FP16 or FP8 directions the new column scale and beta rescaled to the held direction;
// FP8 scales come from the maxima rmax and pmax and the correction's own maximum);
// rz[c] = rz_c; rz_new, pAp, rp and rc of the column cleared for the next iteration.
// Workgroup 0 counts the iteration and sets the done flag when viol <= 1, then moves viol to
// viol[1] (read by the host) and clears it.
__global__ __launch_bounds__(kCoarseMax) void k_coarse_scalars(int na, int C, int64_t ldc,
                                                               const double* __restrict__ Ainv,
                                                               double* __restrict__ rc, float* __restrict__ ec,
                                                               double* __restrict__ rz, double* __restrict__ rz_new,
                                                               double* __restrict__ pAp, double* __restrict__ rp,
                                                               double* __restrict__ beta, float* __restrict__ sig,
                                                               float* __restrict__ viol, int32_t* ctl,
                                                               float* __restrict__ rmax, float* __restrict__ pmax) {
  if (ctl && ctl[0]) return;
  __shared__ double rcs[kCoarseMax];
  __shared__ double part[kCoarseMax / 64];
  __shared__ float partm[kCoarseMax / 64];
  const int c = blockIdx.x, I = threadIdx.x;
  rcs[I] = I < na ? rc[(int64_t)I * ldc + c] : 0.0;
  __syncthreads();
  if (I < na) rc[(int64_t)I * ldc + c] = 0.0;
  float em = 0.f;
  const double v = coarse_column(na, c, ldc, Ainv, rcs, ec, part, partm, rmax ? &em : nullptr);
  if (threadIdx.x == 0) {
    const double rzc = rz_new[c] + v;
    const double r = rz[c];
    double bt = r > 0.0 ? rzc / r : 0.0;
    if (sig) {
      const float sn = rmax ? f8_next_scale(c, bt, em, rmax, pmax, sig) : pcg_scale(rzc);
      bt *= (double)sn / (double)sig[c];
      sig[c] = sn;
    }
    beta[c] = bt;
    rz[c] = rzc;
    rz_new[c] = 0.0;
    pAp[c] = 0.0;
    if (rp) rp[c] = 0.0;
    if (c == 0) {
      if (ctl) {
        ctl[1] += 1;
        if (viol[0] <= 1.f) ctl[0] = 1;
      }
      viol[1] = viol[0];   // the last violation, for the host
      viol[0] = 0.f;
    }
  }
}

// Ac += P^T A P over a share of aggregate I's rows (workgroup I * parts + part): for every
// solved row r of aggregate I, diag_r at column I and -w_rs at column aggmap[s] for every solved
// neighbour s.  Entries of aggregate I itself (most of them: neighbours are mostly near in row
// order) are summed in registers; the others go through an LDS row; the row is added to Ac with
// one atomic per nonzero entry.  Ac must be zero beforehand.
__global__ __launch_bounds__(256) void k_coarse_build(int n, int na, int parts, const int32_t* __restrict__ bounds,
                                                      const uint8_t* __restrict__ aggmap,
                                                      const int32_t* __restrict__ indptr,
                                                      const int32_t* __restrict__ rowlen,
                                                      const int32_t* __restrict__ col, const float* __restrict__ w,
                                                      const uint8_t* __restrict__ solved,
                                                      const double* __restrict__ diag, double* __restrict__ Ac) {
  __shared__ double row[kCoarseMax];
  for (int t = threadIdx.x; t < na; t += blockDim.x) row[t] = 0.0;
  __syncthreads();
  const int I = blockIdx.x / parts, part = blockIdx.x - I * parts;
  const int64_t lo0 = bounds[I], hi0 = bounds[I + 1];
  const int64_t span = (hi0 - lo0 + parts - 1) / parts;
  const int64_t lo = min(hi0, lo0 + part * span), hi = min(hi0, lo + span);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  double own = 0.0;
  for (int64_t r = lo + wave; r < hi; r += nw) {
    if (!solved[r]) continue;
    if (lane == 0) own += diag[r];
    const int beg = indptr[r], end = beg + rowlen[r];
    for (int j = beg + lane; j < end; j += 64) {
      const int sc = col[j];
      if (sc >= 0 && sc < n && solved[sc]) {
        const int J = aggmap[sc];
        if (J == I) own -= (double)w[j];
        else lds_add(&row[J], -(double)w[j]);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) own += __shfl_xor(own, off, 64);
  if (lane == 0 && own != 0.0) lds_add(&row[I], own);
  __syncthreads();
  for (int t = threadIdx.x; t < na; t += blockDim.x)
    if (row[t] != 0.0) atomicAdd(&Ac[(int64_t)I * na + t], row[t]);
}

// diff[u] += (live entries v > u of row u), diff[v] -= 1 for each of them: the exclusive prefix
// sum of diff is the number of edges crossing each row boundary.
__global__ void k_cross_diff(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ rowlen,
                             const int32_t* __restrict__ col, int32_t* __restrict__ diff) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const int beg = indptr[u], end = beg + rowlen[u];
  int cnt = 0;
  for (int j = beg; j < end; ++j) {
    const int v = col[j];
    if (v > u && v < n) {
      ++cnt;
      atomicSub(diff + v, 1);
    }
  }
  if (cnt) atomicAdd(diff + u, cnt);
}

// cand[4(k-1) ..] = the boundary b with the fewest crossing edges in target k's window (ties:
// nearest to the target, then smaller; -1 when the window is empty), its crossing count, and
// the counts half a spacing to either side of it (clamped to the row range).  One workgroup per
// target.
__global__ __launch_bounds__(256) void k_window_argmin(int n, int S0, int min_agg, const int32_t* __restrict__ cross,
                                                       int32_t* __restrict__ cand) {
  __shared__ long long bc[256], bd[256], bb[256];
  const int k = blockIdx.x + 1;
  const long long t = (long long)k * S0;
  const long long lo = max((long long)min_agg, t - S0 / 2), hi = min((long long)n - min_agg, t + S0 / 2);
  long long c0 = LLONG_MAX, d0 = LLONG_MAX, b0 = -1;
  for (long long b = lo + threadIdx.x; b < hi; b += blockDim.x) {
    const long long c = cross[b], d = b > t ? b - t : t - b;
    if (c < c0 || (c == c0 && (d < d0 || (d == d0 && b < b0)))) {
      c0 = c;
      d0 = d;
      b0 = b;
    }
  }
  bc[threadIdx.x] = c0;
  bd[threadIdx.x] = d0;
  bb[threadIdx.x] = b0;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)blockDim.x; ++i) {
      if (bb[i] < 0) continue;
      if (b0 < 0 || bc[i] < c0 || (bc[i] == c0 && (bd[i] < d0 || (bd[i] == d0 && bb[i] < b0)))) {
        c0 = bc[i];
        d0 = bd[i];
        b0 = bb[i];
      }
    }
    int32_t* o = cand + 4 * (k - 1);
    o[0] = (int32_t)b0;
    if (b0 >= 0) {
      const long long bl = max(0LL, b0 - S0 / 2), br = min((long long)n - 1, b0 + S0 / 2);
      o[1] = cross[b0];
      o[2] = cross[bl];
      o[3] = cross[br];
    }
  }
}

__global__ void k_aggmap(int n, int na, const int32_t* __restrict__ bounds, uint8_t* __restrict__ aggmap) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int lo = 0, hi = na;   // bounds[lo] <= r < bounds[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (bounds[mid] <= r) lo = mid;
    else hi = mid;
  }
  aggmap[r] = (uint8_t)lo;
}

// Aggregates without solved rows get a unit diagonal (decoupled from the rest).
__global__ void k_coarse_fixdiag(int na, double* __restrict__ Ac) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I < na && Ac[(int64_t)I * na + I] == 0.0) Ac[(int64_t)I * na + I] = 1.0;
}

// Cholesky factor L (lower, row-major na x na) of Ac, one workgroup with the matrix in LDS.  A
// pivot that is not positive (a coarse row without coupling) is replaced by 1.
__global__ __launch_bounds__(256) void k_coarse_factor(int na, const double* Ac, double* L) {   // L may be Ac
  extern __shared__ double Ls[];
  for (int i = threadIdx.x; i < na * na; i += blockDim.x) Ls[i] = Ac[i];
  __syncthreads();
  for (int k = 0; k < na; ++k) {
    if (threadIdx.x == 0) {
      const double d = Ls[k * na + k];
      Ls[k * na + k] = d > 0.0 ? sqrt(d) : 1.0;
    }
    __syncthreads();
    const double pk = Ls[k * na + k];
    for (int i = k + 1 + threadIdx.x; i < na; i += blockDim.x) Ls[i * na + k] /= pk;
    __syncthreads();
    const int m = na - k - 1;
    for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
      const int i = k + 1 + idx / m, j = k + 1 + idx % m;
      if (j <= i) Ls[i * na + j] -= Ls[i * na + k] * Ls[j * na + k];
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < na * na; i += blockDim.x) {
    const int r = i / na, c = i - r * na;
    L[i] = c <= r ? Ls[i] : 0.0;
  }
}

// Column j of Ainv = L^-T L^-1 e_j (one workgroup per column): forward, then back substitution,
// each step's dot product reduced over the workgroup.
__global__ __launch_bounds__(256) void k_coarse_inverse(int na, const double* __restrict__ L, double* __restrict__ Ainv) {
  __shared__ double y[kCoarseMax];
  __shared__ double part[4];
  const int j = blockIdx.x, t = threadIdx.x;
  for (int i = t; i < na; i += blockDim.x) y[i] = 0.0;
  __syncthreads();
  auto reduce = [&](double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((t & 63) == 0) part[t >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < ((int)blockDim.x >> 6); ++k) s += part[k];
    __syncthreads();
    return s;
  };
  for (int i = j; i < na; ++i) {   // L y = e_j (y[i] = 0 for i < j)
    double v = 0.0;
    for (int k = j + t; k < i; k += blockDim.x) v += L[(int64_t)i * na + k] * y[k];
    const double s = reduce(v);
    if (t == 0) y[i] = ((i == j ? 1.0 : 0.0) - s) / L[(int64_t)i * na + i];
    __syncthreads();
  }
  for (int i = na - 1; i >= 0; --i) {   // L^T x = y, in place
    double v = 0.0;
    for (int k = i + 1 + t; k < na; k += blockDim.x) v += L[(int64_t)k * na + i] * y[k];
    const double s = reduce(v);
    if (t == 0) y[i] = (y[i] - s) / L[(int64_t)i * na + i];
    __syncthreads();
  }
  for (int i = t; i < na; i += blockDim.x) Ainv[(int64_t)i * na + j] = y[i];
}

// Clears n doubles (a fill kernel, so it can sit in a captured graph).
__global__ void k_zero_f64(int n, double* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

}  // namespace dlp

namespace dlp {

static unsigned pass_grid(int n, int lc_shift) {
  long long threads = (long long)n << lc_shift;
  unsigned g = grid_for(threads);
  return g > 2048 ? 2048 : g;
}

int pcg_flat_chunk(const PcgBufs& b) {
  const FlatLaunch f = flat_launch(b);
  if (!f.ok) return 0;
  return (int)(((int64_t)b.n + f.grid - 1) / f.grid);
}

// Lane words per row of the low-precision direction block: 8-byte groups for BF16/FP16, 4-byte
// groups for FP8.
static int64_t pwords(const PcgBufs& b) {
  return !b.P16 ? 0 : b.p16 == 3 ? b.ldh / 2 : b.ldh / 4;
}

static CoarseArg coarse_arg(const PcgBufs& b) {
  if (!b.na) return CoarseArg{0, 0, nullptr, nullptr, nullptr, nullptr, nullptr};
  return CoarseArg{b.na, b.ldc, b.rc, b.ec, b.solved, b.bounds, b.aggmap};
}

// The restriction of the flat passes keeps a workgroup's rows within two aggregates.
static int check_coarse(const PcgBufs& b, const FlatLaunch& f) {
  if (!b.na) return DLP_OK;
  DLP_CHECK_ARG(f.ok && b.na <= kCoarseMax && b.rc && b.ec && b.Ainv && b.solved && b.bounds && b.aggmap &&
                    b.ldc >= 4 * f.Q4 && ((int64_t)b.n + f.grid - 1) / f.grid <= b.min_agg,
                "pcg: coarse layout (na=%d min_agg=%d n=%d grid=%u)", b.na, b.min_agg, b.n, f.grid);
  return DLP_OK;
}

// The arrays a solve starts from zero, cleared by one launch (a launch per array costs more
// than its work): up to kClearMax (pointer, count of doubles) pairs.
constexpr int kClearMax = 12;
struct ClearList {
  double* p[kClearMax];
  int n[kClearMax];
  int k;
  void add(void* ptr, int count) {
    p[k] = reinterpret_cast<double*>(ptr);
    n[k++] = count;
  }
};

__global__ void k_clear_many(ClearList L) {
  for (int j = 0; j < L.k; ++j)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < L.n[j]; i += gridDim.x * blockDim.x) L.p[j][i] = 0.0;
}

static void clear_many(const ClearList& L, hipStream_t st) {
  int most = 1;
  for (int j = 0; j < L.k; ++j) most = L.n[j] > most ? L.n[j] : most;
  const int grid = (most + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_clear_many, dim3(grid < 16 ? grid : 16), dim3(kBlock), 0, st, L);
}

int pcg_start(const PcgBufs& b, hipStream_t st) {
  DLP_CHECK_ARG(b.C >= 1 && b.C <= kMaxCols, "pcg: C=%d outside [1, %d]", b.C, kMaxCols);
  const ColMap m = col_map(b.C);
  const unsigned g = pass_grid(b.n, m.lc_shift);
  const FlatLaunch f = flat_launch(b);
  DLP_CHECK_ARG(f.ok || !b.P16, "pcg: 16-bit directions need 16-byte FP32 rows");
  DLP_CHECK_ARG(!b.P16 || b.p16 == 1 || (b.p16 == 2 && b.sig) || (b.p16 == 3 && b.sig && b.rmax && b.pmax && b.emax),
                "pcg: low-precision direction kind %d", b.p16);
  DLP_CHECK_ARG(f.ok || !b.na, "pcg: the coarse correction needs 16-byte FP32 rows");
  int rc = check_coarse(b, f);
  if (rc) return rc;
  const bool f8 = f.ok && b.P16 && b.p16 == 3;
  ClearList L{};
  L.add(b.rz, b.C);
  L.add(b.rz_new, b.C);
  L.add(b.pAp, b.C);
  L.add(b.viol, 1);   // viol and the pad word after it (the iterations clear viol after each test)
  if (b.ctl) L.add(b.ctl, 1);
  if (f.ok && b.P16) L.add(b.rp, b.C);
  if (f8)   // the maxima accumulate from zero (C floats each fit in C/2 doubles)
    for (float* mx : {b.rmax, b.pmax, b.emax}) L.add(mx, (b.C + 1) / 2);
  // step lengths of the deferred updates (held as doubles' worth of float pairs)
  if (f.ok && b.P16 && b.slot_stride > 0 && b.alpha_hist) L.add(b.alpha_hist, (b.group * b.C + 1) / 2);
  if (f.ok && b.na) L.add(b.rc, b.na * (int)b.ldc);
  clear_many(L, st);
  if (f.ok) {
    const CoarseArg K = coarse_arg(b);
    float* sig = (b.P16 && b.p16 >= 2) ? b.sig : nullptr;
    if (sig || b.na) {
      // scale pass: rz (for the FP16 column scales), the FP8 maxima and the restricted
      // residual, then the coarse correction of the start
      hipLaunchKernelGGL(k_pcg_rz4, dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C, (const float4*)b.R, b.diag,
                         b.rz, K, f8 ? b.rmax : nullptr);
      if (b.na)
        hipLaunchKernelGGL(k_coarse_apply, dim3(b.C), dim3(kCoarseMax), 0, st, b.na, b.C, b.ldc, b.Ainv, b.rc, b.ec,
                           b.rz, nullptr, f8 ? b.emax : nullptr);
      hipLaunchKernelGGL(k_pcg_scalars, dim3(1), dim3(kBlock), 0, st, b.C, 2, b.rz, b.rz_new, b.pAp, b.alpha,
                         b.beta, b.viol, nullptr, nullptr, sig, nullptr, 0, f8 ? b.rmax : nullptr,
                         f8 ? b.pmax : nullptr, f8 ? b.emax : nullptr);
    }
    if (!b.P16)
      hipLaunchKernelGGL(k_pcg_start4<0>, dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C,
                         PBlock<0>{b.P, 0}, (const float4*)b.R, b.diag, b.rz, nullptr, nullptr, K, nullptr);
    else if (b.p16 == 3)
      hipLaunchKernelGGL(k_pcg_start4<3>, dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C,
                         PBlock<3>{b.P16, pwords(b)}, (const float4*)b.R, b.diag, b.rz, b.rp, b.sig, K, b.pmax);
    else if (b.p16 == 2)
      hipLaunchKernelGGL(k_pcg_start4<2>, dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C,
                         PBlock<2>{b.P16, pwords(b)}, (const float4*)b.R, b.diag, b.rz, b.rp, b.sig, K, nullptr);
    else
      hipLaunchKernelGGL(k_pcg_start4<1>, dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C,
                         PBlock<1>{b.P16, pwords(b)}, (const float4*)b.R, b.diag, b.rz, b.rp, nullptr, K, nullptr);
    DLP_CHECK_LAUNCH();
    return DLP_OK;
  }
#define K(CPLV)                                                                                     \
  hipLaunchKernelGGL((k_pcg_dir<CPLV>), dim3(g), dim3(kBlock), 0, st, b.n, b.C, m.lc_shift, b.P, b.R, \
                     b.ld, b.diag, b.beta, 1, nullptr);                                                     \
  hipLaunchKernelGGL((k_coldot<CPLV>), dim3(g), dim3(kBlock), 0, st, b.n, b.C, m.lc_shift, b.R, b.P,  \
                     b.ld, b.rz);
  DLP_DISPATCH_CPL(m.cpl, K)
#undef K
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// Makes A exactly symmetric: A[i][j] = A[j][i] = (A[i][j] + A[j][i]) / 2.
__global__ void k_symmetrize(int na, double* __restrict__ A) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= na * na) return;
  const int i = idx / na, j = idx - i * na;
  if (j < i) {
    const double v = 0.5 * (A[i * na + j] + A[j * na + i]);
    A[i * na + j] = v;
    A[j * na + i] = v;
  }
}

int coarse_build(int n, int na, const int32_t* bounds, const uint8_t* aggmap, const int32_t* indptr,
                 const int32_t* rowlen, const int32_t* col, const float* w, const uint8_t* solved,
                 const double* diag, double* Ac, double* Ainv, hipStream_t st) {
  DLP_CHECK_ARG(n >= 1 && na >= 1 && na <= kCoarseMax && bounds && aggmap, "coarse_build: na=%d n=%d", na, n);
  // each workgroup (4 wavefronts) takes about 256 rows of one aggregate: a wavefront's rows
  // are walked one after another, each a chain of dependent loads
  int parts = (int)(((long long)n / na + 255) / 256);
  parts = parts < 8 ? 8 : (parts > 256 ? 256 : parts);
  hipLaunchKernelGGL(k_zero_f64, dim3(grid_for((long long)na * na)), dim3(kBlock), 0, st, na * na, Ac);
  hipLaunchKernelGGL(k_coarse_build, dim3(na * parts), dim3(kBlock), 0, st, n, na, parts, bounds, aggmap, indptr,
                     rowlen, col, w, solved, diag, Ac);
  hipLaunchKernelGGL(k_coarse_fixdiag, dim3(1), dim3(kCoarseMax), 0, st, na, Ac);
  hipLaunchKernelGGL(k_coarse_factor, dim3(1), dim3(kBlock), sizeof(double) * na * na, st, na, Ac, Ac);
  hipLaunchKernelGGL(k_coarse_inverse, dim3(na), dim3(kBlock), 0, st, na, Ac, Ainv);
  hipLaunchKernelGGL(k_symmetrize, dim3(grid_for((long long)na * na)), dim3(kBlock), 0, st, na, Ainv);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

int coarse_boundaries(int n, int nk, int S0, int min_agg, const int32_t* indptr, const int32_t* rowlen,
                      const int32_t* col, int32_t* diff, int32_t* cross, void* ws, size_t ws_bytes,
                      int32_t* cand, hipStream_t st) {
  DLP_CHECK_ARG(n >= 1 && nk >= 1 && S0 >= 1 && min_agg >= 1, "coarse_boundaries: bad sizes");
  DLP_CHECK_HIP(hipMemsetAsync(diff, 0, sizeof(int32_t) * n, st));
  hipLaunchKernelGGL(k_cross_diff, dim3(grid_for(n)), dim3(kBlock), 0, st, n, indptr, rowlen, col, diff);
  int rc = exclusive_scan_i32(n, diff, cross, nullptr, ws, ws_bytes, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_window_argmin, dim3(nk), dim3(kBlock), 0, st, n, S0, min_agg, cross, cand);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

int coarse_aggmap(int n, int na, const int32_t* bounds, uint8_t* aggmap, hipStream_t st) {
  DLP_CHECK_ARG(n >= 1 && na >= 1 && na <= 256, "coarse_aggmap: na=%d", na);
  hipLaunchKernelGGL(k_aggmap, dim3(grid_for(n)), dim3(kBlock), 0, st, n, na, bounds, aggmap);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

int pcg_finish_iteration(const PcgBufs& b, hipStream_t st, bool have_pap) {
  const ColMap m = col_map(b.C);
  const unsigned g = pass_grid(b.n, m.lc_shift);
  const FlatLaunch f = flat_launch(b);
  if (f.ok) {
    if (!have_pap) {
#define K(CPLV)                                                                                       \
  hipLaunchKernelGGL((k_coldot<CPLV>), dim3(g), dim3(kBlock), 0, st, b.n, b.C, m.lc_shift, b.P, b.AP, b.ld, \
                     b.pAp);
      DLP_DISPATCH_CPL(m.cpl, K)
#undef K
    }
    int rc = check_coarse(b, f);
    if (rc) return rc;
    const CoarseArg K = coarse_arg(b);
    double* rp = b.P16 ? b.rp : nullptr;
    float* sig = (b.P16 && b.p16 >= 2) ? b.sig : nullptr;
    const bool f8 = b.P16 && b.p16 == 3;
    float* rmax = f8 ? b.rmax : nullptr;
    float* pmax = f8 ? b.pmax : nullptr;
    // the update pass takes its step lengths from pAp and r.p (16-bit directions) or r.z
    const double* num = b.P16 ? b.rp : b.rz;
    const bool nt = tuning().nt_streams != 0;
    // deferred solution updates: this iteration gathers slot b.slot, the next direction goes to
    // the slot after it, and the step lengths to the group's history
    const bool defer = b.P16 && b.slot_stride > 0;
    DLP_CHECK_ARG(!defer || (b.alpha_hist && b.slot >= 0 && b.slot < b.group), "pcg: deferred update slot %d",
                  b.slot);
    uint16_t* Pcur = b.P16 ? b.P16 + (defer ? b.slot * b.slot_stride : 0) : nullptr;
    uint16_t* Pnext = defer ? Pcur + b.slot_stride : Pcur;
    float* aout = defer ? b.alpha_hist + (int64_t)b.slot * b.C : nullptr;
#define DLP_UPDATE_NT(PKV, PB, NTV, DV)                                                                      \
  hipLaunchKernelGGL((k_pcg_update4<PKV, NTV, DV>), dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C,     \
                     (float4*)b.Z, (float4*)b.R, PBlock<PKV>{PB, pwords(b)}, (const float4*)b.AP,            \
                     b.pAp, num, b.diag, b.eps, b.rz_new, b.viol, b.ctl, K, aout, rmax)
#define DLP_UPDATE(PKV, PB)                                   \
  if (defer && nt) DLP_UPDATE_NT(PKV, PB, true, true);       \
  else if (defer) DLP_UPDATE_NT(PKV, PB, false, true);       \
  else if (nt) DLP_UPDATE_NT(PKV, PB, true, false);          \
  else DLP_UPDATE_NT(PKV, PB, false, false)
#define DLP_DIR_NT(PKV, PB, PO, NTV)                                                                         \
  hipLaunchKernelGGL((k_pcg_dir4<PKV, NTV>), dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C,           \
                     PBlock<PKV>{PB, pwords(b)}, PBlock<PKV>{PO, pwords(b)},                                 \
                     (const float4*)b.R, b.diag, b.beta, rp, b.ctl, sig, K, pmax)
#define DLP_DIR(PKV, PB, PO)              \
  if (nt) DLP_DIR_NT(PKV, PB, PO, true); \
  else DLP_DIR_NT(PKV, PB, PO, false)
    if (!b.P16) { DLP_UPDATE(0, b.P); }
    else if (b.p16 == 3) { DLP_UPDATE(3, Pcur); }
    else if (b.p16 == 2) { DLP_UPDATE(2, Pcur); }
    else { DLP_UPDATE(1, Pcur); }
    // beta (from r^T M r: the update's r^T D^-1 r plus the coarse part), the convergence test,
    // and the coarse correction the direction uses
    if (b.na)
      hipLaunchKernelGGL(k_coarse_scalars, dim3(b.C), dim3(kCoarseMax), 0, st, b.na, b.C, b.ldc, b.Ainv, b.rc, b.ec,
                         b.rz, b.rz_new, b.pAp, rp, b.beta, sig, b.viol, b.ctl, rmax, pmax);
    else
      hipLaunchKernelGGL(k_pcg_scalars, dim3(1), dim3(kBlock), 0, st, b.C, 1, b.rz, b.rz_new, b.pAp, b.alpha,
                         b.beta, b.viol, b.ctl, rp, sig, nullptr, 0, rmax, pmax, nullptr);
    if (!b.P16) { DLP_DIR(0, b.P, b.P); }
    else if (b.p16 == 3) { DLP_DIR(3, Pcur, Pnext); }
    else if (b.p16 == 2) { DLP_DIR(2, Pcur, Pnext); }
    else { DLP_DIR(1, Pcur, Pnext); }
#undef DLP_UPDATE
#undef DLP_DIR
#undef DLP_UPDATE_NT
#undef DLP_DIR_NT
    DLP_CHECK_LAUNCH();
    return DLP_OK;
  }
  DLP_CHECK_ARG(!b.P16 && !b.na, "pcg: 16-bit directions and the coarse correction need 16-byte FP32 rows");
#define K(CPLV)                                                                                       \
  if (!have_pap)                                                                                      \
    hipLaunchKernelGGL((k_coldot<CPLV>), dim3(g), dim3(kBlock), 0, st, b.n, b.C, m.lc_shift, b.P, b.AP, \
                       b.ld, b.pAp);                                                                 \
  hipLaunchKernelGGL(k_pcg_scalars, dim3(1), dim3(kBlock), 0, st, b.C, 0, b.rz, b.rz_new, b.pAp,        \
                     b.alpha, b.beta, b.viol, b.ctl, nullptr, nullptr, nullptr, 0, nullptr, nullptr, nullptr); \
  hipLaunchKernelGGL((k_pcg_update<CPLV>), dim3(g), dim3(kBlock), 0, st, b.n, b.C, m.lc_shift, b.Z, b.R, \
                     b.P, b.AP, b.ld, b.alpha, b.diag, b.eps, b.rz_new, b.viol, b.ctl);              \
  hipLaunchKernelGGL(k_pcg_scalars, dim3(1), dim3(kBlock), 0, st, b.C, 1, b.rz, b.rz_new, b.pAp,        \
                     b.alpha, b.beta, b.viol, b.ctl, nullptr, nullptr, nullptr, 0, nullptr, nullptr, nullptr); \
  hipLaunchKernelGGL((k_pcg_dir<CPLV>), dim3(g), dim3(kBlock), 0, st, b.n, b.C, m.lc_shift, b.P, b.R,   \
                     b.ld, b.diag, b.beta, 0, b.ctl);
  DLP_DISPATCH_CPL(m.cpl, K)
#undef K
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

int pcg_flush(const PcgBufs& b, hipStream_t st) {
  if (!b.P16 || b.slot_stride <= 0) return DLP_OK;
  const FlatLaunch f = flat_launch(b);
  DLP_CHECK_ARG(f.ok && b.alpha_hist && b.group >= 1 && b.group <= kMaxGroup,
                "pcg: deferred updates need the flat passes and at most %d iterations per group", kMaxGroup);
  const bool nt = tuning().nt_streams != 0;
#define DLP_FLUSH(PKV, NTV)                                                                                 \
  hipLaunchKernelGGL((k_pcg_flush4<PKV, NTV>), dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C, (float4*)b.Z, \
                     PBlock<PKV>{b.P16, pwords(b)}, b.slot_stride, b.group, b.alpha_hist, b.ctl)
  if (b.p16 == 3) {
    if (nt) DLP_FLUSH(3, true);
    else DLP_FLUSH(3, false);
  } else if (b.p16 == 2) {
    if (nt) DLP_FLUSH(2, true);
    else DLP_FLUSH(2, false);
  } else {
    if (nt) DLP_FLUSH(1, true);
    else DLP_FLUSH(1, false);
  }
#undef DLP_FLUSH
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

}  // namespace dlp

// ---------------------------------------------------------------------------------------
// Generic column-wise BLAS-1 over row-major blocks (FP32 or FP64), for the reference-API
// solvers (solvers.py:66-97): column dots, column-scaled axpy, P = X + beta P, X / diag.

namespace dlp {

template <typename T, int CPL>
__global__ __launch_bounds__(256) void k_coldot_g(int n, int C, int lc_shift, const T* __restrict__ X, int64_t ldx,
                                                  const T* __restrict__ Y, int64_t ldy,
                                                  const T* __restrict__ d, double* __restrict__ out) {
  __shared__ double sm[kMaxCols];
  for (int c = threadIdx.x; c < C; c += blockDim.x) sm[c] = 0.0;
  __syncthreads();
  const int LC = 1 << lc_shift;
  const int l = threadIdx.x & (LC - 1);
  double acc[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) acc[k] = 0.0;
  const int64_t groups = ((int64_t)gridDim.x * blockDim.x) >> lc_shift;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> lc_shift; r < n; r += groups) {
    const double s = d ? 1.0 / (double)d[r] : 1.0;
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = l + k * LC;
      if (c < C) acc[k] += (double)X[r * ldx + c] * (double)Y[r * ldy + c] * s;
    }
  }
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int c = l + k * LC;
    if (c < C) atomicAdd(&sm[c], acc[k]);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) atomicAdd(&out[c], sm[c]);
}

// mode 0: Y += sign*a_c*X;  mode 1: Y = X + a_c*Y;  mode 2: Y = X / d[r]
template <typename T, int CPL>
__global__ __launch_bounds__(256) void k_colop_g(int n, int C, int lc_shift, int mode, double sign,
                                                 const double* __restrict__ a, const T* __restrict__ X,
                                                 int64_t ldx, const T* __restrict__ d, T* __restrict__ Y,
                                                 int64_t ldy) {
  const int LC = 1 << lc_shift;
  const int l = threadIdx.x & (LC - 1);
  T av[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int c = l + k * LC;
    av[k] = (a && c < C) ? (T)(sign * a[c]) : (T)0;
  }
  const int64_t groups = ((int64_t)gridDim.x * blockDim.x) >> lc_shift;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> lc_shift; r < n; r += groups) {
    const T inv = (mode >= 2) ? (T)1 / d[r] : (T)0;
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = l + k * LC;
      if (c < C) {
        const T x = X[r * ldx + c];
        T* y = Y + r * ldy + c;
        if (mode == 0) *y += av[k] * x;
        else if (mode == 1) *y = x + av[k] * *y;
        else if (mode == 2) *y = x * inv;
        else *y += av[k] * (x * inv);
      }
    }
  }
}

}  // namespace dlp

using namespace dlp;

extern "C" int dlp_coldot(int n, int C, const void* X, int64_t ldx, const void* Y, int64_t ldy,
                          const void* d, double* out, int dtype, void* stream) {
  DLP_CHECK_ARG(n >= 0 && C >= 1 && C <= kMaxCols && out, "coldot: bad arguments");
  DLP_CHECK_ARG(n == 0 || (X && Y), "coldot: null pointer");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(kBlock), 0, st, C, out);
  if (n == 0) return DLP_OK;
  const ColMap m = col_map(C);
  const unsigned g = pass_grid(n, m.lc_shift);
#define K(CPLV)                                                                                       \
  if (dtype == DLP_F32)                                                                               \
    hipLaunchKernelGGL((k_coldot_g<float, CPLV>), dim3(g), dim3(kBlock), 0, st, n, C, m.lc_shift,       \
                       (const float*)X, ldx, (const float*)Y, ldy, (const float*)d, out);              \
  else                                                                                                \
    hipLaunchKernelGGL((k_coldot_g<double, CPLV>), dim3(g), dim3(kBlock), 0, st, n, C, m.lc_shift,      \
                       (const double*)X, ldx, (const double*)Y, ldy, (const double*)d, out);
  DLP_DISPATCH_CPL(m.cpl, K)
#undef K
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_colop(int n, int C, int mode, double sign, const double* a, const void* X, int64_t ldx,
                         const void* d, void* Y, int64_t ldy, int dtype, void* stream) {
  DLP_CHECK_ARG(n >= 0 && C >= 1 && C <= kMaxCols && mode >= 0 && mode <= 3, "colop: bad arguments");
  if (n == 0) return DLP_OK;
  const bool needs_a = mode != 2, needs_d = mode >= 2;
  DLP_CHECK_ARG(X && Y && (!needs_a || a != nullptr) && (!needs_d || d != nullptr), "colop: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const ColMap m = col_map(C);
  const unsigned g = pass_grid(n, m.lc_shift);
#define K(CPLV)                                                                                        \
  if (dtype == DLP_F32)                                                                                \
    hipLaunchKernelGGL((k_colop_g<float, CPLV>), dim3(g), dim3(kBlock), 0, st, n, C, m.lc_shift, mode,   \
                       sign, a, (const float*)X, ldx, (const float*)d, (float*)Y, ldy);                \
  else                                                                                                 \
    hipLaunchKernelGGL((k_colop_g<double, CPLV>), dim3(g), dim3(kBlock), 0, st, n, C, m.lc_shift, mode,  \
                       sign, a, (const double*)X, ldx, (const double*)d, (double*)Y, ldy);
  DLP_DISPATCH_CPL(m.cpl, K)
#undef K
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

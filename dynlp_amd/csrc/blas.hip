// Fused BLAS-1 passes of the column-wise PCG (solvers.py:66-97) over row-major n x C blocks.
//
// Every pass maps column lanes over a row (LC = 2^lc_shift lanes, CPL column chunks per lane)
// and rows over the grid, so each row is one coalesced access.  Column reductions accumulate
// in FP64: per lane, then per workgroup in LDS, then one FP64 atomic per column per workgroup.
#include "blas.h"
#include "rowpipe.h"

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
  if (threadIdx.x == 0) atomic_max_nonneg(viol, vmax);
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
// nearest even, FP16 saturating).  FP16 directions are held scaled by a power of two per column
// (the caller passes the scaled values), which keeps them inside FP16's range.
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

// pcg start: P = R/diag, rz[c] += sum R*R/diag; with a 16-bit block also rp[c] += sum R*P as
// held.  FP16 blocks hold sig[c] * R/diag (sig from the rz of a prior k_pcg_rz4 pass).
template <int PK>
__global__ __launch_bounds__(256) void k_pcg_start4(int n, int Q4, int C, PBlock<PK> P, const float4* __restrict__ R,
                                                    const float* __restrict__ diag, double* __restrict__ rz,
                                                    double* __restrict__ rp, const float* __restrict__ sig) {
  __shared__ double sm[kMaxCols + 4];
  __shared__ double sm2[kMaxCols + 4];
  for (int c = threadIdx.x; c < 4 * Q4; c += blockDim.x) sm[c] = sm2[c] = 0.0;
  __syncthreads();
  const Quad Qd = quad_of(Q4, C);
  float sg[4] = {1.f, 1.f, 1.f, 1.f};
  if constexpr (PK == 2) {
#pragma unroll
    for (int j = 0; j < 4; ++j) sg[j] = j < Qd.cnt ? sig[4 * Qd.q4 + j] : 1.f;
  }
  const int64_t nq = (int64_t)n * Q4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0}, acc2[4] = {0.0, 0.0, 0.0, 0.0};
  if (Qd.active && Qd.r0 < n) {
    int64_t r = Qd.r0;
    for (int64_t i = Qd.i0; i < nq; i += Qd.step, r += Qd.rstep) {
      const float inv = 1.f / diag[r];
      const float4 rv = R[i];
      float4 y = P.load(i, r, Qd.q4);
      y.x = Qd.cnt > 0 ? rv.x * inv : y.x;
      y.y = Qd.cnt > 1 ? rv.y * inv : y.y;
      y.z = Qd.cnt > 2 ? rv.z * inv : y.z;
      y.w = Qd.cnt > 3 ? rv.w * inv : y.w;
      const float4 p = P.store(i, r, Qd.q4, make_float4(y.x * sg[0], y.y * sg[1], y.z * sg[2], y.w * sg[3]));
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j < Qd.cnt) {
          acc[j] += (double)quad_comp(rv, j) * (double)quad_comp(y, j);
          if constexpr (PK != 0) acc2[j] += (double)quad_comp(rv, j) * (double)quad_comp(p, j);
        }
      }
    }
  }
  flush_quads(Qd, Q4, C, acc, sm, rz);
  if constexpr (PK != 0) flush_quads(Qd, Q4, C, acc2, sm2, rp);
}

// rz[c] += sum R*R/diag (the scale pass before an FP16 start).
__global__ __launch_bounds__(256) void k_pcg_rz4(int n, int Q4, int C, const float4* __restrict__ R,
                                                 const float* __restrict__ diag, double* __restrict__ rz) {
  __shared__ double sm[kMaxCols + 4];
  for (int c = threadIdx.x; c < 4 * Q4; c += blockDim.x) sm[c] = 0.0;
  __syncthreads();
  const Quad Qd = quad_of(Q4, C);
  const int64_t nq = (int64_t)n * Q4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (Qd.active && Qd.r0 < n) {
    int64_t r = Qd.r0;
    for (int64_t i = Qd.i0; i < nq; i += Qd.step, r += Qd.rstep) {
      const double inv = 1.0 / (double)diag[r];
      const float4 rv = R[i];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j < Qd.cnt) acc[j] += (double)quad_comp(rv, j) * (double)quad_comp(rv, j) * inv;
    }
  }
  flush_quads(Qd, Q4, C, acc, sm, rz);
}

// Z += a P, R -= a AP; rz_new[c] += sum R*R/diag; viol = max |R|/(eps_c*diag).
template <int PK>
__global__ __launch_bounds__(256) void k_pcg_update4(int n, int Q4, int C, float4* __restrict__ Z,
                                                     float4* __restrict__ R, PBlock<PK> P,
                                                     const float4* __restrict__ AP,
                                                     const double* __restrict__ alpha,
                                                     const float* __restrict__ diag, const float* __restrict__ eps,
                                                     double* __restrict__ rz_new, float* __restrict__ viol,
                                                     const int32_t* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ double sm[kMaxCols + 4];
  __shared__ float vmax;
  for (int c = threadIdx.x; c < 4 * Q4; c += blockDim.x) sm[c] = 0.0;
  if (threadIdx.x == 0) vmax = 0.f;
  __syncthreads();
  const Quad Qd = quad_of(Q4, C);
  float a[4], ie[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int c = 4 * Qd.q4 + j;
    a[j] = j < Qd.cnt ? (float)alpha[c] : 0.f;
    ie[j] = j < Qd.cnt ? 1.f / eps[c] : 0.f;
  }
  const int64_t nq = (int64_t)n * Q4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  float vm = 0.f;
  if (Qd.active && Qd.r0 < n) {
    int64_t r = Qd.r0;
    for (int64_t i = Qd.i0; i < nq; i += Qd.step, r += Qd.rstep) {
      const float inv = 1.f / diag[r];
      float4 z = Z[i], rv = R[i];
      const float4 p = P.load(i, r, Qd.q4), ap = AP[i];
      if (Qd.cnt > 0) { z.x += a[0] * p.x; rv.x -= a[0] * ap.x; }
      if (Qd.cnt > 1) { z.y += a[1] * p.y; rv.y -= a[1] * ap.y; }
      if (Qd.cnt > 2) { z.z += a[2] * p.z; rv.z -= a[2] * ap.z; }
      if (Qd.cnt > 3) { z.w += a[3] * p.w; rv.w -= a[3] * ap.w; }
      Z[i] = z;
      R[i] = rv;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j < Qd.cnt) {
          const float x = quad_comp(rv, j);
          acc[j] += (double)x * (double)x * (double)inv;
          vm = fmaxf(vm, fabsf(x) * inv * ie[j]);
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (j < Qd.cnt) lds_add(&sm[4 * Qd.q4 + j], acc[j]);
  for (int off = 32; off > 0; off >>= 1) vm = fmaxf(vm, __shfl_xor(vm, off, 64));
  if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&vmax, vm);
  __syncthreads();
  for (int c = threadIdx.x; c < C; c += blockDim.x) atomicAdd(&rz_new[c], sm[c]);
  if (threadIdx.x == 0) atomic_max_nonneg(viol, vmax);
}

// P = R/diag + beta P; with a 16-bit block also rp[c] += sum R*P as held (the next step
// length's numerator, exact for the rounded direction).  FP16 blocks: P = sig*R/diag + beta*P
// with sig the new column scale and beta already rescaled to the held direction.
template <int PK>
__global__ __launch_bounds__(256) void k_pcg_dir4(int n, int Q4, int C, PBlock<PK> P, const float4* __restrict__ R,
                                                  const float* __restrict__ diag, const double* __restrict__ beta,
                                                  double* __restrict__ rp, const int32_t* __restrict__ skip,
                                                  const float* __restrict__ sig) {
  if (skip && *skip) return;
  __shared__ double sm[kMaxCols + 4];
  if constexpr (PK != 0) {
    for (int c = threadIdx.x; c < 4 * Q4; c += blockDim.x) sm[c] = 0.0;
    __syncthreads();
  }
  const Quad Qd = quad_of(Q4, C);
  float b[4], sg[4] = {1.f, 1.f, 1.f, 1.f};
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = j < Qd.cnt ? (float)beta[4 * Qd.q4 + j] : 0.f;
  if constexpr (PK == 2) {
#pragma unroll
    for (int j = 0; j < 4; ++j) sg[j] = j < Qd.cnt ? sig[4 * Qd.q4 + j] : 1.f;
  }
  const int64_t nq = (int64_t)n * Q4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (Qd.active && Qd.r0 < n) {
    int64_t r = Qd.r0;
    for (int64_t i = Qd.i0; i < nq; i += Qd.step, r += Qd.rstep) {
      const float inv = 1.f / diag[r];
      const float4 rv = R[i];
      float4 p = P.load(i, r, Qd.q4);
      if constexpr (PK == 2) {
        if (Qd.cnt > 0) p.x = rv.x * inv * sg[0] + b[0] * p.x;
        if (Qd.cnt > 1) p.y = rv.y * inv * sg[1] + b[1] * p.y;
        if (Qd.cnt > 2) p.z = rv.z * inv * sg[2] + b[2] * p.z;
        if (Qd.cnt > 3) p.w = rv.w * inv * sg[3] + b[3] * p.w;
      } else {
        if (Qd.cnt > 0) p.x = rv.x * inv + b[0] * p.x;
        if (Qd.cnt > 1) p.y = rv.y * inv + b[1] * p.y;
        if (Qd.cnt > 2) p.z = rv.z * inv + b[2] * p.z;
        if (Qd.cnt > 3) p.w = rv.w * inv + b[3] * p.w;
      }
      const float4 held = P.store(i, r, Qd.q4, p);
      if constexpr (PK != 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (j < Qd.cnt) acc[j] += (double)quad_comp(rv, j) * (double)quad_comp(held, j);
      }
    }
  }
  if constexpr (PK != 0) flush_quads(Qd, Q4, C, acc, sm, rp);
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

// Scalars of one iteration (one workgroup): mode 0: alpha = rz/pAp (0 where pAp <= 0), then
// clear rz_new and viol for the update pass; mode 1: beta = rz_new/rz (0 where rz <= 0),
// rz = rz_new, then clear pAp for the next iteration; with sig (FP16 directions) also the new
// column scale, and beta rescaled to the held direction (beta * sig_new / sig_old).  Mode 2
// (before an FP16 start): sig from rz, then rz = 0.
__global__ void k_pcg_scalars(int C, int mode, double* __restrict__ rz, double* __restrict__ rz_new,
                              double* __restrict__ pAp, double* __restrict__ alpha,
                              double* __restrict__ beta, float* __restrict__ viol, int32_t* ctl,
                              double* __restrict__ rp, float* __restrict__ sig) {
  // every thread reads the flag before thread 0 may set it below (after the barrier)
  if (ctl && ctl[0]) return;
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
        const float s = pcg_scale(rz_new[c]);
        bt *= (double)s / (double)sig[c];
        sig[c] = s;
      }
      beta[c] = bt;
      rz[c] = rz_new[c];
      pAp[c] = 0.0;
      if (rp) rp[c] = 0.0;
    } else {
      sig[c] = pcg_scale(rz[c]);
      rz[c] = 0.0;
    }
  }
  if (mode == 0 && threadIdx.x == 0) *viol = 0.f;
  if (mode == 1 && ctl) {
    __syncthreads();
    if (threadIdx.x == 0) {
      ctl[1] += 1;
      if (*viol <= 1.f) ctl[0] = 1;
    }
  }
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

int pcg_start(const PcgBufs& b, hipStream_t st) {
  DLP_CHECK_ARG(b.C >= 1 && b.C <= kMaxCols, "pcg: C=%d outside [1, %d]", b.C, kMaxCols);
  const ColMap m = col_map(b.C);
  const unsigned g = pass_grid(b.n, m.lc_shift);
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(kBlock), 0, st, b.C, b.rz);
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(kBlock), 0, st, b.C, b.pAp);
  if (b.ctl) hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(kBlock), 0, st, 1, (double*)b.ctl);
  const FlatLaunch f = flat_launch(b);
  DLP_CHECK_ARG(f.ok || !b.P16, "pcg: 16-bit directions need 16-byte FP32 rows");
  DLP_CHECK_ARG(!b.P16 || b.p16 == 1 || (b.p16 == 2 && b.sig), "pcg: 16-bit direction kind %d", b.p16);
  if (b.P16) {
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(kBlock), 0, st, b.C, b.rp);
    if (b.p16 == 2) {
      hipLaunchKernelGGL(k_pcg_rz4, dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C, (const float4*)b.R, b.diag,
                         b.rz);
      hipLaunchKernelGGL(k_pcg_scalars, dim3(1), dim3(kBlock), 0, st, b.C, 2, b.rz, b.rz_new, b.pAp, b.alpha,
                         b.beta, b.viol, nullptr, nullptr, b.sig);
      hipLaunchKernelGGL(k_pcg_start4<2>, dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C,
                         PBlock<2>{b.P16, b.ldh / 4}, (const float4*)b.R, b.diag, b.rz, b.rp, b.sig);
    } else {
      hipLaunchKernelGGL(k_pcg_start4<1>, dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C,
                         PBlock<1>{b.P16, b.ldh / 4}, (const float4*)b.R, b.diag, b.rz, b.rp, nullptr);
    }
    DLP_CHECK_LAUNCH();
    return DLP_OK;
  }
  if (f.ok) {
    hipLaunchKernelGGL(k_pcg_start4<0>, dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C,
                       PBlock<0>{b.P, 0}, (const float4*)b.R, b.diag, b.rz, nullptr, nullptr);
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
    double* rp = b.P16 ? b.rp : nullptr;
    float* sig = (b.P16 && b.p16 == 2) ? b.sig : nullptr;
    hipLaunchKernelGGL(k_pcg_scalars, dim3(1), dim3(kBlock), 0, st, b.C, 0, b.rz, b.rz_new, b.pAp, b.alpha,
                       b.beta, b.viol, b.ctl, rp, sig);
#define DLP_UPDATE(PKV, PB)                                                                                  \
  hipLaunchKernelGGL(k_pcg_update4<PKV>, dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C, (float4*)b.Z, \
                     (float4*)b.R, PBlock<PKV>{PB, b.P16 ? b.ldh / 4 : 0}, (const float4*)b.AP, b.alpha, b.diag, \
                     b.eps, b.rz_new, b.viol, b.ctl)
#define DLP_DIR(PKV, PB)                                                                                     \
  hipLaunchKernelGGL(k_pcg_dir4<PKV>, dim3(f.grid), dim3(f.block), 0, st, b.n, f.Q4, b.C,                  \
                     PBlock<PKV>{PB, b.P16 ? b.ldh / 4 : 0}, (const float4*)b.R, b.diag, b.beta, rp, b.ctl, sig)
    if (!b.P16) DLP_UPDATE(0, b.P);
    else if (b.p16 == 2) DLP_UPDATE(2, b.P16);
    else DLP_UPDATE(1, b.P16);
    hipLaunchKernelGGL(k_pcg_scalars, dim3(1), dim3(kBlock), 0, st, b.C, 1, b.rz, b.rz_new, b.pAp, b.alpha,
                       b.beta, b.viol, b.ctl, rp, sig);
    if (!b.P16) DLP_DIR(0, b.P);
    else if (b.p16 == 2) DLP_DIR(2, b.P16);
    else DLP_DIR(1, b.P16);
#undef DLP_UPDATE
#undef DLP_DIR
    DLP_CHECK_LAUNCH();
    return DLP_OK;
  }
  DLP_CHECK_ARG(!b.P16, "pcg: 16-bit directions need 16-byte FP32 rows");
#define K(CPLV)                                                                                       \
  if (!have_pap)                                                                                      \
    hipLaunchKernelGGL((k_coldot<CPLV>), dim3(g), dim3(kBlock), 0, st, b.n, b.C, m.lc_shift, b.P, b.AP, \
                       b.ld, b.pAp);                                                                 \
  hipLaunchKernelGGL(k_pcg_scalars, dim3(1), dim3(kBlock), 0, st, b.C, 0, b.rz, b.rz_new, b.pAp,        \
                     b.alpha, b.beta, b.viol, b.ctl, nullptr, nullptr);                              \
  hipLaunchKernelGGL((k_pcg_update<CPLV>), dim3(g), dim3(kBlock), 0, st, b.n, b.C, m.lc_shift, b.Z, b.R, \
                     b.P, b.AP, b.ld, b.alpha, b.diag, b.eps, b.rz_new, b.viol, b.ctl);              \
  hipLaunchKernelGGL(k_pcg_scalars, dim3(1), dim3(kBlock), 0, st, b.C, 1, b.rz, b.rz_new, b.pAp,        \
                     b.alpha, b.beta, b.viol, b.ctl, nullptr, nullptr);                              \
  hipLaunchKernelGGL((k_pcg_dir<CPLV>), dim3(g), dim3(kBlock), 0, st, b.n, b.C, m.lc_shift, b.P, b.R,   \
                     b.ld, b.diag, b.beta, 0, b.ctl);
  DLP_DISPATCH_CPL(m.cpl, K)
#undef K
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
    const T inv = (mode == 2) ? (T)1 / d[r] : (T)0;
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = l + k * LC;
      if (c < C) {
        const T x = X[r * ldx + c];
        T* y = Y + r * ldy + c;
        if (mode == 0) *y += av[k] * x;
        else if (mode == 1) *y = x + av[k] * *y;
        else *y = x * inv;
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
  DLP_CHECK_ARG(n >= 0 && C >= 1 && C <= kMaxCols && mode >= 0 && mode <= 2, "colop: bad arguments");
  if (n == 0) return DLP_OK;
  DLP_CHECK_ARG(X && Y && (mode == 2 ? d != nullptr : a != nullptr), "colop: null pointer");
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

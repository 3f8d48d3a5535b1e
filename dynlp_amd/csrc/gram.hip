// Block PCG building blocks: the tall-skinny FP64 Gram product G += X^T diag(s) Y on the
// matrix cores (v_mfma_f64_16x16x4f64), an FP64 Cholesky solve of the k x k systems in one
// workgroup, and the row-local block updates (n x k times k x k).
//
// Gram: a workgroup stages 32 rows of X and Y in LDS (converted to FP64, columns padded to a
// multiple of 16), and each of its four wavefronts accumulates a share of the 16 x 16 output
// tiles with MFMA over 4-row k-steps.  Operand layout of the 16x16x4 FP64 MFMA: lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16]; the result tile has 4 values per lane,
// D[l / 16 + 4 * v][l % 16], v = 0..3 (pinned by the Gram test).  Partial tiles are added to G
// with FP64 atomics.
#include "common.h"

namespace dlp {

typedef double d4_t __attribute__((ext_vector_type(4)));

constexpr int kGramRows = 32;   // rows per LDS stage (2 x 32 x 80 doubles = 40 KB)
constexpr int kGramMaxT = 5;    // up to 80 columns (5 tiles of 16)

template <typename T>
__global__ __launch_bounds__(256) void k_gram_mfma(int n, int C, const T* __restrict__ X, int64_t ldx,
                                                   const T* __restrict__ Y, int64_t ldy,
                                                   const T* __restrict__ rowdiv, double* __restrict__ G) {
  const int nt = (C + 15) >> 4;          // tiles per side
  const int CP = nt << 4;                // padded columns
  __shared__ double xs[kGramRows * kGramMaxT * 16];
  __shared__ double ys[kGramRows * kGramMaxT * 16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ntiles = nt * nt;
  // each wave owns tiles wave, wave+4, ... (at most 7 of 25)
  d4_t acc[7];
#pragma unroll
  for (int q = 0; q < 7; ++q) acc[q] = (d4_t){0.0, 0.0, 0.0, 0.0};
  const int li = lane & 15, lk = lane >> 4;
  for (int64_t r0 = (int64_t)blockIdx.x * kGramRows; r0 < n; r0 += (int64_t)gridDim.x * kGramRows) {
    __syncthreads();
    for (int t = threadIdx.x; t < kGramRows * CP; t += blockDim.x) {
      const int rr = t / CP, c = t - rr * CP;
      const int64_t r = r0 + rr;
      double xv = 0.0, yv = 0.0;
      if (r < n && c < C) {
        xv = (double)X[r * ldx + c];
        yv = (double)Y[r * ldy + c];
        if (rowdiv) yv /= (double)rowdiv[r];
      }
      xs[rr * CP + c] = xv;
      ys[rr * CP + c] = yv;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 7; ++q) {
      const int tile = wave + 4 * q;
      if (tile < ntiles) {
        const int ti = tile / nt, tj = tile - ti * nt;
        d4_t a4 = acc[q];
        for (int k = 0; k < kGramRows; k += 4) {
          const double a = xs[(k + lk) * CP + ti * 16 + li];
          const double b = ys[(k + lk) * CP + tj * 16 + li];
          a4 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, a4, 0, 0, 0);
        }
        acc[q] = a4;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const int tile = wave + 4 * q;
    if (tile < ntiles) {
      const int ti = tile / nt, tj = tile - ti * nt;
      const int j = tj * 16 + li;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = ti * 16 + lk + 4 * v;
        if (i < C && j < C) atomicAdd(&G[i * C + j], acc[q][v]);
      }
    }
  }
}

// Register-tiled Gram for C = 16*NT (+1 with EXTRA): every wavefront keeps all NT x NT output
// tiles in accumulator registers and streams its own 16-row chunks of X and Y straight from
// memory (4 k-steps of 4 rows per iteration, all loads issued before the multiplies); an
// operand fragment is read once and feeds NT MFMAs.  The extra column (h of the label block)
// goes to VALU FP64 FMAs instead of a fifth, mostly padded tile row.  SYM: Y is X (the B
// fragments are the A fragments scaled by 1/rowdiv).  Partials are summed in LDS per workgroup,
// then added to G with one FP64 atomic per element per workgroup.
// Operand fragments of D k-steps (4 rows each) starting at k-step k0: lane (li, lk) holds
// column 16t + li of row 4*k + lk (zero beyond n).
template <typename T, int NT, bool EXTRA, bool SYM, int D>
struct GramFrag {
  T xa[D][NT], yb[SYM ? 1 : D][NT], xe[D], ye[D], dv[D];
  __device__ __forceinline__ void load(int64_t k0, int n, int li, int lk, const T* __restrict__ X, int64_t ldx,
                                       const T* __restrict__ Y, int64_t ldy, const T* __restrict__ rowdiv) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int64_t r = ((k0 + d) << 2) + lk;
      const bool ok = r < n;
      const T* xr = X + (ok ? r : 0) * ldx;
      const T* yr = Y + (ok ? r : 0) * ldy;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        xa[d][t] = ok ? xr[16 * t + li] : (T)0;
        if constexpr (!SYM) yb[d][t] = ok ? yr[16 * t + li] : (T)0;
      }
      if constexpr (EXTRA) {
        xe[d] = ok ? xr[16 * NT] : (T)0;
        ye[d] = ok ? yr[16 * NT] : (T)0;
      }
      dv[d] = (rowdiv && ok) ? rowdiv[r] : (T)1;
    }
  }
};

// Register-tiled Gram for C = 16*NT (+1 with EXTRA): every wavefront keeps all NT x NT output
// tiles in accumulator registers and streams its own 16-row chunks of X and Y straight from
// memory; an operand fragment is read once and feeds NT MFMAs, and the fragments of the next
// chunk are loaded while the current one is multiplied.  The extra column (h of the label
// block) goes to VALU FP64 FMAs instead of a fifth, mostly padded tile row.  SYM: Y is X (the B
// fragments are the A fragments scaled by 1/rowdiv).  Partials are summed in LDS per workgroup,
// then added to G with one FP64 atomic per element per workgroup.
template <typename T, int NT, bool EXTRA, bool SYM>
__global__ __launch_bounds__(sizeof(T) == 4 ? 512 : 256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1))) void k_gram_reg(int n, int C, const T* __restrict__ X, int64_t ldx,
                                                  const T* __restrict__ Y, int64_t ldy, const T* __restrict__ rowdiv,
                                                  double* __restrict__ G) {
  constexpr int CT = 16 * NT;
  constexpr int C1 = CT + (EXTRA ? 1 : 0);
  // k-steps per chunk: 4, or 2 when the two operands' double-buffered fragments would cost a
  // wave of occupancy
  constexpr int D = (SYM || NT < 4) ? 4 : 2;
  __shared__ double red[C1 * C1];
  for (int i = threadIdx.x; i < C1 * C1; i += blockDim.x) red[i] = 0.0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  d4_t acc[NT][NT];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int u = 0; u < NT; ++u) acc[t][u] = (d4_t){0.0, 0.0, 0.0, 0.0};
  double er[NT], ec[NT], ed = 0.0;
#pragma unroll
  for (int t = 0; t < NT; ++t) er[t] = ec[t] = 0.0;
  const int64_t nks = ((int64_t)n + 3) >> 2;
  const int64_t nwb = blockDim.x >> 6;
  const int64_t wid = (int64_t)blockIdx.x * nwb + wave, nw = (int64_t)gridDim.x * nwb;
  GramFrag<T, NT, EXTRA, SYM, D> cur, nxt;
  int64_t k0 = wid * D;
  if (k0 < nks) cur.load(k0, n, li, lk, X, ldx, Y, ldy, rowdiv);
  for (; k0 < nks; k0 += nw * D) {
    if (k0 + nw * D < nks) nxt.load(k0 + nw * D, n, li, lk, X, ldx, Y, ldy, rowdiv);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double inv = 1.0 / (double)cur.dv[d];
      double a[NT], b[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        a[t] = (double)cur.xa[d][t];
        if constexpr (SYM) b[t] = a[t] * inv;
        else b[t] = (double)cur.yb[d][t] * inv;
      }
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int u = 0; u < NT; ++u) acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[u], acc[t][u], 0, 0, 0);
      if constexpr (EXTRA) {
        const double xd = (double)cur.xe[d], yd = (double)cur.ye[d] * inv;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          er[t] += xd * b[t];   // G[CT][16t + li]
          ec[t] += a[t] * yd;   // G[16t + li][CT]
        }
        if (li == 0) ed += xd * yd;
      }
    }
    cur = nxt;
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int u = 0; u < NT; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) atomicAdd(&red[(16 * t + lk + 4 * v) * C1 + 16 * u + li], acc[t][u][v]);
  if constexpr (EXTRA) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      double x = er[t], y = ec[t];
      x += __shfl_xor(x, 16, 64);
      x += __shfl_xor(x, 32, 64);
      y += __shfl_xor(y, 16, 64);
      y += __shfl_xor(y, 32, 64);
      if (lk == 0) {
        atomicAdd(&red[CT * C1 + 16 * t + li], x);
        atomicAdd(&red[(16 * t + li) * C1 + CT], y);
      }
    }
    for (int off = 32; off > 0; off >>= 1) ed += __shfl_xor(ed, off, 64);
    if (lane == 0) atomicAdd(&red[CT * C1 + CT], ed);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < C1 * C1; i += blockDim.x) atomicAdd(&G[i], red[i]);
}

// S = A^{-1} B for SPD A (k x k, FP64, row-major), B k x k; one workgroup.  The system is
// symmetrically scaled to unit diagonal; a pivot below tol (relative) marks a direction that
// has collapsed (its column of the block has converged or become dependent): that row and
// column are dropped (deflation), and the matching rows of S are zero.
__global__ __launch_bounds__(256) void k_chol_solve(int k, const double* __restrict__ A, const double* __restrict__ B,
                                                    double* __restrict__ S, double tol) {
  extern __shared__ double sh[];
  double* L = sh;              // k*k
  double* X = sh + k * k;      // k*k (right-hand sides, then solution)
  double* dsc = X + k * k;     // k scaling
  int* alive = (int*)(dsc + k);
  for (int t = threadIdx.x; t < k; t += blockDim.x) {
    const double d = A[t * k + t];
    dsc[t] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
    alive[t] = d > 0.0;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < k * k; t += blockDim.x) {
    const int i = t / k, j = t - (t / k) * k;
    L[t] = A[t] * dsc[i] * dsc[j];
    X[t] = B[t] * dsc[i];
  }
  __syncthreads();
  // right-looking Cholesky, lower triangle
  for (int c = 0; c < k; ++c) {
    __shared__ double piv;
    if (threadIdx.x == 0) {
      double p = alive[c] ? L[c * k + c] : 0.0;
      if (!(p > tol)) {
        alive[c] = 0;
        piv = 0.0;
      } else {
        piv = sqrt(p);
      }
    }
    __syncthreads();
    const double pv = piv;
    for (int i = c + threadIdx.x; i < k; i += blockDim.x) {
      if (pv > 0.0) L[i * k + c] = (i == c) ? pv : L[i * k + c] / pv;
      else L[i * k + c] = (i == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (pv > 0.0) {
      for (int t = threadIdx.x; t < (k - c - 1) * (k - c - 1); t += blockDim.x) {
        const int i = c + 1 + t / (k - c - 1), j = c + 1 + t % (k - c - 1);
        if (j <= i) L[i * k + j] -= L[i * k + c] * L[j * k + c];
      }
    }
    __syncthreads();
  }
  // forward and back substitution, one thread per right-hand side column
  for (int col = threadIdx.x; col < k; col += blockDim.x) {
    for (int i = 0; i < k; ++i) {
      if (!alive[i]) {
        X[i * k + col] = 0.0;
        continue;
      }
      double s = X[i * k + col];
      for (int j = 0; j < i; ++j) s -= L[i * k + j] * X[j * k + col];
      X[i * k + col] = s / L[i * k + i];
    }
    for (int i = k - 1; i >= 0; --i) {
      if (!alive[i]) {
        X[i * k + col] = 0.0;
        continue;
      }
      double s = X[i * k + col];
      for (int j = i + 1; j < k; ++j) s -= L[j * k + i] * X[j * k + col];
      X[i * k + col] = s / L[i * k + i];
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < k * k; t += blockDim.x) {
    const int i = t / k;
    S[t] = X[t] * dsc[i];
  }
}

// Row-local block updates on 64-row tiles (S = FP64 C x C step matrix, C <= 80):
// mode 0: Z += P S, R -= AP S, viol = max |R|/(eps*diag);  mode 1: P = Y + P S, with Y the
// preconditioned residual (Y null: Y = R/diag, the Jacobi preconditioner).  The step matrix
// (FP32 copy, row pitch 80) and the tile's rows of P (and AP) are staged in LDS with coalesced
// loads; a thread computes 20 consecutive output
// columns of one row (four threads a row), the products read S as 16-byte words; the results go
// back through the LDS tiles so that Z, R and P are read and written row-contiguously.
constexpr int kBuRows = 64, kBuPitchS = 80, kBuCols = 20;
template <int MODE>
__global__ __launch_bounds__(256) void k_block_update_tiled(int n, int C, float* __restrict__ Z, float* __restrict__ R,
                                                            float* __restrict__ P, const float* __restrict__ AP,
                                                            const float* __restrict__ Y, int64_t ld,
                                                            const double* __restrict__ S,
                                                            const float* __restrict__ diag,
                                                            const float* __restrict__ eps, float* __restrict__ viol) {
  __shared__ __attribute__((aligned(16))) float st[kGramMaxT * 16 * kBuPitchS];
  __shared__ float pt[kBuRows * (kGramMaxT * 16 + 1)];
  __shared__ float at[MODE == 0 ? kBuRows * (kGramMaxT * 16 + 1) : 1];
  __shared__ float ie[kGramMaxT * 16];
  __shared__ float vmax;
  const int PT = C + 1;   // tile row pitch: the 16 rows a wavefront reads fall in distinct banks
  for (int t = threadIdx.x; t < C * kBuPitchS; t += blockDim.x) {
    const int i = t / kBuPitchS, j = t - i * kBuPitchS;
    st[t] = j < C ? (float)S[i * C + j] : 0.f;
  }
  if (MODE == 0) {
    for (int j = threadIdx.x; j < C; j += blockDim.x) ie[j] = 1.f / eps[j];
    if (threadIdx.x == 0) vmax = 0.f;
  }
  const int row = threadIdx.x >> 2, j0 = (threadIdx.x & 3) * kBuCols;
  float vm = 0.f;
  for (int64_t r0 = (int64_t)blockIdx.x * kBuRows; r0 < n; r0 += (int64_t)gridDim.x * kBuRows) {
    const int rows = (int)min<int64_t>(kBuRows, n - r0);
    __syncthreads();   // st and ie ready; the previous tile's write-back has finished with pt/at
    for (int t = threadIdx.x; t < rows * C; t += blockDim.x) {
      const int rr = t / C, c = t - rr * C;
      const int64_t g = (r0 + rr) * ld + c;
      pt[rr * PT + c] = P[g];
      if (MODE == 0) at[rr * PT + c] = AP[g];
    }
    __syncthreads();
    float az[kBuCols], ar[kBuCols];
#pragma unroll
    for (int k = 0; k < kBuCols; ++k) az[k] = ar[k] = 0.f;
    if (row < rows && j0 < C) {
      for (int i = 0; i < C; ++i) {
        const float pv = pt[row * PT + i];
        const float av = MODE == 0 ? at[row * PT + i] : 0.f;
        const float4* sr = reinterpret_cast<const float4*>(st + i * kBuPitchS + j0);
#pragma unroll
        for (int q = 0; q < kBuCols / 4; ++q) {
          const float4 s4 = sr[q];
          const float sv[4] = {s4.x, s4.y, s4.z, s4.w};
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            az[4 * q + u] = __builtin_fmaf(pv, sv[u], az[4 * q + u]);
            if (MODE == 0) ar[4 * q + u] = __builtin_fmaf(av, sv[u], ar[4 * q + u]);
          }
        }
      }
    }
    __syncthreads();   // every thread has read the tiles: they now take the products
    if (row < rows) {
#pragma unroll
      for (int k = 0; k < kBuCols; ++k) {
        const int j = j0 + k;
        if (j < C) {
          pt[row * PT + j] = az[k];
          if (MODE == 0) at[row * PT + j] = ar[k];
        }
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < rows * C; t += blockDim.x) {
      const int rr = t / C, c = t - rr * C;
      const int64_t g = (r0 + rr) * ld + c;
      if (MODE == 0) {
        Z[g] += pt[rr * PT + c];
        const float rv = R[g] - at[rr * PT + c];
        R[g] = rv;
        vm = fmaxf(vm, fabsf(rv) / diag[r0 + rr] * ie[c]);
      } else {
        P[g] = (Y ? Y[g] : R[g] / diag[r0 + rr]) + pt[rr * PT + c];
      }
    }
  }
  if (MODE == 0) {
    for (int off = 32; off > 0; off >>= 1) vm = fmaxf(vm, __shfl_xor(vm, off, 64));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&vmax, vm);
    __syncthreads();
    if (threadIdx.x == 0) global_max_nonneg(viol, vmax);
  }
}

__global__ void k_zero_f64b(int n, double* __restrict__ p) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.0;
}

__global__ void k_zero_f32b(float* __restrict__ p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0.f;
}

__global__ void k_copy_f64(int n, const double* __restrict__ a, double* __restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = a[i];
}

}  // namespace dlp

using namespace dlp;

// G = X^T diag(1/rowdiv) Y (FP64, C x C row-major, overwritten).  X, Y: n x C, FP32 or FP64.
extern "C" int dlp_gram(int n, int C, const void* X, int64_t ldx, const void* Y, int64_t ldy, const void* rowdiv,
                        double* G, int dtype, void* stream) {
  DLP_CHECK_ARG(n >= 0 && C >= 1 && C <= 16 * kGramMaxT && G, "gram: C=%d outside [1, %d]", C, 16 * kGramMaxT);
  DLP_CHECK_ARG(n == 0 || (X && Y), "gram: null pointer");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_f64b, dim3(8), dim3(kBlock), 0, st, C * C, G);
  const int nt = C / 16, rem = C % 16;
  if (n > 0 && nt >= 1 && nt <= 4 && rem <= 1) {
    // register-tiled form: C = 16, 17, 32, 33, 48, 49, 64, 65
    const bool sym = X == Y && ldx == ldy;
    // FP32 operands: workgroups of 8 wavefronts (the two per SIMD the kernel keeps resident),
    // one per CU, so the per-workgroup partials cost one FP64 atomic per element of G per CU;
    // FP64 operands (more registers per fragment): 4 wavefronts, two workgroups per CU
    int cus = 256, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    const int wg = dtype == DLP_F32 ? 512 : 256;
    const long long need = ((long long)n + 15) / 16 / (wg / 64);
    long long grid = 2LL * cus * 256 / wg;
    if (grid > need) grid = need < 1 ? 1 : need;
#define DLP_GRAM(TV, NTV, EV, SV)                                                                        \
  hipLaunchKernelGGL((k_gram_reg<TV, NTV, EV, SV>), dim3((unsigned)grid), dim3(wg), 0, st, n, C,         \
                     (const TV*)X, ldx, (const TV*)Y, ldy, (const TV*)rowdiv, G)
#define DLP_GRAM_NT(TV, EV, SV)            \
  switch (nt) {                            \
    case 1: DLP_GRAM(TV, 1, EV, SV); break; \
    case 2: DLP_GRAM(TV, 2, EV, SV); break; \
    case 3: DLP_GRAM(TV, 3, EV, SV); break; \
    default: DLP_GRAM(TV, 4, EV, SV); break; \
  }
#define DLP_GRAM_T(TV)                                   \
  if (rem) {                                             \
    if (sym) DLP_GRAM_NT(TV, true, true) else DLP_GRAM_NT(TV, true, false) \
  } else {                                               \
    if (sym) DLP_GRAM_NT(TV, false, true) else DLP_GRAM_NT(TV, false, false) \
  }
    if (dtype == DLP_F32) {
      DLP_GRAM_T(float)
    } else {
      DLP_GRAM_T(double)
    }
#undef DLP_GRAM_T
#undef DLP_GRAM_NT
#undef DLP_GRAM
    DLP_CHECK_LAUNCH();
    return DLP_OK;
  }
  if (n > 0) {
    unsigned grid = (unsigned)((n + kGramRows - 1) / kGramRows);
    if (grid > 1024) grid = 1024;
    if (dtype == DLP_F32)
      hipLaunchKernelGGL(k_gram_mfma<float>, dim3(grid), dim3(kBlock), 0, st, n, C, (const float*)X, ldx,
                         (const float*)Y, ldy, (const float*)rowdiv, G);
    else
      hipLaunchKernelGGL(k_gram_mfma<double>, dim3(grid), dim3(kBlock), 0, st, n, C, (const double*)X, ldx,
                         (const double*)Y, ldy, (const double*)rowdiv, G);
  }
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// S = A^{-1} B (k x k FP64, device), A SPD; deflates collapsed directions (relative pivot tol).
extern "C" int dlp_chol_solve(int k, const double* A, const double* B, double* S, double tol, void* stream) {
  DLP_CHECK_ARG(k >= 1 && k <= 128 && A && B && S, "chol_solve: bad arguments (k=%d)", k);
  const size_t shm = sizeof(double) * (2 * k * k + k) + sizeof(int) * k;
  hipLaunchKernelGGL(k_chol_solve, dim3(1), dim3(kBlock), shm, (hipStream_t)stream, k, A, B, S, tol);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

namespace dlp {
int copy_f64(int n, const double* a, double* b, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_f64, dim3(8), dim3(kBlock), 0, st, n, a, b);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

int block_update(int n, int C, int mode, float* Z, float* R, float* P, const float* AP, const float* Y, int64_t ld,
                 const double* S, const float* diag, const float* eps, float* viol, hipStream_t st) {
  static_assert(4 * kBuCols == kGramMaxT * 16 && kBuPitchS >= kGramMaxT * 16, "tile covers every column count");
  DLP_CHECK_ARG(C >= 1 && C <= kGramMaxT * 16, "block_update: C=%d outside [1, %d]", C, kGramMaxT * 16);
  if (mode == 0) hipLaunchKernelGGL(k_zero_f32b, dim3(1), dim3(1), 0, st, viol);
  long long tiles = ((long long)n + kBuRows - 1) / kBuRows;
  if (tiles < 1) return DLP_OK;
  if (tiles > 4096) tiles = 4096;
  if (mode == 0)
    hipLaunchKernelGGL(k_block_update_tiled<0>, dim3((unsigned)tiles), dim3(kBlock), 0, st, n, C, Z, R, P, AP, Y, ld,
                       S, diag, eps, viol);
  else
    hipLaunchKernelGGL(k_block_update_tiled<1>, dim3((unsigned)tiles), dim3(kBlock), 0, st, n, C, Z, R, P, AP, Y, ld,
                       S, diag, eps, viol);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}}  // namespace dlp

// The block PCG's row-local updates (FP32 n x C blocks of leading dimension ld, S = FP64 C x C
// on the device): mode 0: Z += P S, R -= AP S and viol[0] = max |R|/(eps*diag) over the rows
// (cleared first); mode 1: P = Y + P S (Y = NULL: Y = R/diag).
extern "C" int dlp_block_update(int n, int C, int mode, float* Z, float* R, float* P, const float* AP, const float* Y,
                                int64_t ld, const double* S, const float* diag, const float* eps, float* viol,
                                void* stream) {
  DLP_CHECK_ARG(n >= 0 && C >= 1 && C <= 16 * kGramMaxT && ld >= C && (mode == 0 || mode == 1),
                "block_update: bad n=%d, C=%d, ld=%lld or mode=%d", n, C, (long long)ld, mode);
  DLP_CHECK_ARG(R && P && S && diag && (mode == 1 || (Z && AP && eps && viol)), "block_update: null pointer");
  if (n == 0) return DLP_OK;
  return block_update(n, C, mode, Z, R, P, AP, Y, ld, S, diag, eps, viol, (hipStream_t)stream);
}

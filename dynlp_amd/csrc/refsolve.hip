// Device side of the reference-API solvers on a batch system (local ids, plain CSR, FP32 or
// FP64 state): the column-wise PCG recurrences with their scalars kept on the device, the
// residual push rounds with device-resident candidate lists, and the parking of rows outside
// the solved set.  The host drives the loops and reads one scalar (or one count) per check.
//
// Column PCG (solvers.py:66-97 is the algorithm; the data flow is this file's): scal holds, per
// column, rz (r.z of the current direction), pAp and a spare slot, then the violation
// max |R_uc| / (eps_c diag_u).  An iteration is one operator pass (AP = diag*P - W P), one
// column-dot pass (pAp) and one fused update pass that forms alpha = rz/pAp per column, moves Z
// and R and reduces the violation; a direction is one column-dot pass (r.z_new) and one pass
// that forms beta and P = Y + beta*P, then r.z moves to its slot.
//
// Push rounds (solvers.py:100-125 is the algorithm): a round takes the residual of every
// candidate (D = R/diag, Z += D, R = 0), scatters W*D to the neighbours with atomics, and
// selects the next candidates among the vertices this round touched (first touches are
// listed through an int32 stamp per vertex) that violate |R_uc| > eps_c diag_u.  Counts stay
// on the device, so rounds are launched back to back; kernels loop over the count they read.
#include "rowgather.h"

namespace dlp {

// scal layout: [0, C) rz, [C, 2C) pAp, [2C, 3C) r.z of the new direction, [3C] violation
__device__ __forceinline__ double safe_ratio(double num, double den) { return den > 0.0 ? num / den : 0.0; }

__device__ __forceinline__ void atomic_max_pos(double* p, double v) {
  // non-negative doubles order like their bit patterns (+inf included)
  atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

template <typename T>
__device__ __forceinline__ void block_max_to(double local, double* out) {
  __shared__ double sm[kBlock / kWave];
  for (int off = 32; off > 0; off >>= 1) local = fmax(local, __shfl_xor(local, off, 64));
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) sm[wv] = local;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) m = fmax(m, sm[i]);
    atomic_max_pos(out, m);
  }
}

__global__ void k_fill_f64(int count, double* __restrict__ p, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// Z += alpha P, R -= alpha AP with alpha_c = rz_c / pAp_c (0 when pAp_c <= 0), and the
// violation max |R_uc| / (eps_c diag_u) of the updated residual.
template <typename T>
__global__ __launch_bounds__(256) void k_cg_update(int n, int C, int64_t ld, const double* __restrict__ scal,
                                                   T* __restrict__ Z, const T* __restrict__ P, T* __restrict__ R,
                                                   const T* __restrict__ AP, const T* __restrict__ diag,
                                                   const double* __restrict__ eps, double* __restrict__ viol) {
  const int64_t total = (int64_t)n * C;
  double vm = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / C;
    const int c = (int)(t - r * C);
    const T a = (T)safe_ratio(scal[c], scal[C + c]);
    const int64_t o = r * ld + c;
    Z[o] += a * P[o];
    const T rv = R[o] - a * AP[o];
    R[o] = rv;
    vm = fmax(vm, nan_to_inf(fabs((double)rv) / ((double)diag[r] * eps[c])));
  }
  block_max_to<T>(vm, viol);
}

// P = Y + beta P with beta_c = rzn_c / rz_c (0 when rz_c <= 0; P = Y on the first direction).
template <typename T>
__global__ __launch_bounds__(256) void k_cg_direction(int n, int C, int64_t ld, const double* __restrict__ scal,
                                                      const T* __restrict__ Y, T* __restrict__ P, int first) {
  const int64_t total = (int64_t)n * C;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / C;
    const int c = (int)(t - r * C);
    const int64_t o = r * ld + c;
    P[o] = first ? Y[o] : Y[o] + (T)safe_ratio(scal[2 * C + c], scal[c]) * P[o];
  }
}

__global__ void k_cg_commit_rz(int C, double* __restrict__ scal) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C) scal[c] = scal[2 * C + c];
}

// R = rhs - AZ (AZ already in R) and the violation of the result.
template <typename T>
__global__ __launch_bounds__(256) void k_cg_residual(int n, int C, int64_t ld, const T* __restrict__ rhs,
                                                     int64_t ldr, T* __restrict__ R, const T* __restrict__ diag,
                                                     const double* __restrict__ eps, double* __restrict__ viol) {
  const int64_t total = (int64_t)n * C;
  double vm = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / C;
    const int c = (int)(t - r * C);
    const int64_t o = r * ld + c;
    const T rv = rhs[r * ldr + c] - R[o];
    R[o] = rv;
    vm = fmax(vm, nan_to_inf(fabs((double)rv) / ((double)diag[r] * eps[c])));
  }
  block_max_to<T>(vm, viol);
}

static unsigned flat_grid(int64_t total) {
  long long g = (long long)grid_for(total);
  return (unsigned)(g > 4096 ? 4096 : (g < 1 ? 1 : g));
}

// ---------------------------------------------------------------------------------------------
// Push rounds.

__device__ __forceinline__ void first_touch(int v, int32_t* __restrict__ stamp, int32_t* __restrict__ touched,
                                            int32_t* __restrict__ ntouched) {
  wave_append(atomicExch(stamp + v, 1) == 0, v, touched, ntouched);
}

template <typename T>
__global__ __launch_bounds__(256) void k_round_take(const int32_t* __restrict__ count, const int32_t* __restrict__ cand,
                                                    int C, int64_t ld, const T* __restrict__ diag,
                                                    T* __restrict__ Z, T* __restrict__ R, T* __restrict__ D,
                                                    int32_t* __restrict__ stamp, int32_t* __restrict__ touched,
                                                    int32_t* __restrict__ ntouched) {
  const int64_t total = (int64_t)*count * C;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / C;
    const int c = (int)(t - i * C);
    const int v = cand[i];
    const int64_t o = (int64_t)v * ld + c;
    const T dv = R[o] / diag[v];
    D[i * C + c] = dv;
    Z[o] += dv;
    R[o] = (T)0;
    if (c == 0) first_touch(v, stamp, touched, ntouched);
  }
}

// R[col] += w * D[i] over the entries of every candidate row: one wavefront per row, its lanes
// split into column lanes (a power of two covering C, at most 64) and edge lanes.
template <typename T>
__global__ __launch_bounds__(256) void k_round_scatter(const int32_t* __restrict__ count, const int32_t* __restrict__ cand,
                                                       int C, int64_t ld, int lc_shift, const int32_t* __restrict__ indptr,
                                                       const int32_t* __restrict__ indices, const T* __restrict__ w,
                                                       const T* __restrict__ D, T* __restrict__ R,
                                                       int32_t* __restrict__ stamp, int32_t* __restrict__ touched,
                                                       int32_t* __restrict__ ntouched,
                                                       unsigned long long* __restrict__ edges) {
  const int nc = *count;
  const int LC = 1 << lc_shift, LE = 64 >> lc_shift;
  const int lane = threadIdx.x & 63, l = lane & (LC - 1), e = lane >> lc_shift;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < nc; i += waves) {
    const int v = cand[i];
    const int beg = indptr[v], end = indptr[v + 1];
    if (lane == 0 && edges) atomicAdd(edges, (unsigned long long)(end - beg));
    for (int j = beg + e; j < end; j += LE) {
      const int u = indices[j];
      const T wv = w[j];
      for (int c = l; c < C; c += LC) atomicAdd(R + (int64_t)u * ld + c, wv * D[i * C + c]);
      if (l == 0) first_touch(u, stamp, touched, ntouched);
    }
  }
}

// Next candidates: touched vertices that violate (stamps cleared, the touched count reset for
// the next round by the launch that follows).
template <typename T>
__global__ __launch_bounds__(256) void k_round_select(const int32_t* __restrict__ ntouched, const int32_t* __restrict__ touched,
                                                      int C, int64_t ld, const T* __restrict__ R,
                                                      const T* __restrict__ diag, const double* __restrict__ eps,
                                                      int32_t* __restrict__ stamp, int32_t* __restrict__ next,
                                                      int32_t* __restrict__ nnext) {
  const int nt = *ntouched;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nt; i += (int64_t)gridDim.x * blockDim.x) {
    const int v = touched[i];
    stamp[v] = 0;
    const double dv = (double)diag[v];
    bool viol = false;
    for (int c = 0; c < C && !viol; ++c) viol = !(fabs((double)R[(int64_t)v * ld + c]) <= eps[c] * dv);
    wave_append(viol, v, next, nnext);
  }
}

// counts: [0] current candidates, [1] next candidates, [2] touched; the round ends by moving
// [1] to [0] and clearing [1], [2] (one thread).
__global__ void k_round_rotate(int32_t* __restrict__ counts) {
  if (threadIdx.x == 0) {
    counts[0] = counts[1];
    counts[1] = 0;
    counts[2] = 0;
  }
}

// Rows outside the solved set (mask 0) take the exact values of an inert row: F = 0, h = 1
// (column C1 - 1), residual 0 (solvers.py:365-368).
template <typename T>
__global__ void k_park(int n, const uint8_t* __restrict__ solved, int C1, int64_t ld, T* __restrict__ Z,
                       T* __restrict__ R) {
  const int64_t total = (int64_t)n * C1;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / C1;
    if (solved[r]) continue;
    const int c = (int)(t - r * C1);
    Z[r * ld + c] = c == C1 - 1 ? (T)1 : (T)0;
    if (R) R[r * ld + c] = (T)0;
  }
}

}  // namespace dlp

using namespace dlp;

#define DLP_TYPED(dtype, CALL)                 \
  if (dtype == DLP_F32) {                      \
    using T = float;                           \
    CALL;                                      \
  } else if (dtype == DLP_F64) {               \
    using T = double;                          \
    CALL;                                      \
  } else {                                     \
    DLP_CHECK_ARG(false, "bad dtype %d", dtype); \
  }

extern "C" int dlp_cg_iteration(int n, int C, int64_t ld, int dtype, const int32_t* indptr, const int32_t* rowlen,
                                const int32_t* indices, const void* w, int w_dtype, const uint8_t* mask,
                                double avg_deg, const void* diag, void* Z, void* R, void* P, void* AP, double* scal,
                                const double* eps, void* stream) {
  DLP_CHECK_ARG(n >= 0 && C >= 1 && C <= 256 && ld >= C && scal && eps, "cg_iteration: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(kBlock), 0, st, 1, scal + 3 * C, 0.0);
  if (n == 0) {
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(kBlock), 0, st, C, scal + C, 0.0);
    DLP_CHECK_LAUNCH();
    return DLP_OK;
  }
  DLP_CHECK_ARG(indptr && indices && w && diag && Z && R && P && AP, "cg_iteration: null pointer");
  // AP = diag * P - W P, then pAp per column
  int rc = dlp_row_operator(n, nullptr, C, indptr, rowlen, indices, w, w_dtype, mask, P, ld, dtype, diag, -1.0,
                            nullptr, 0, AP, ld, dtype, 0, avg_deg, stream);
  if (rc) return rc;
  rc = dlp_coldot(n, C, P, ld, AP, ld, nullptr, scal + C, dtype, stream);
  if (rc) return rc;
  const unsigned g = flat_grid((int64_t)n * C);
  DLP_TYPED(dtype, hipLaunchKernelGGL(k_cg_update<T>, dim3(g), dim3(kBlock), 0, st, n, C, ld, scal, (T*)Z,
                                      (const T*)P, (T*)R, (const T*)AP, (const T*)diag, eps, scal + 3 * C));
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_cg_direction(int n, int C, int64_t ld, int dtype, const void* R, const void* Y, void* P,
                                double* scal, int first, void* stream) {
  DLP_CHECK_ARG(n >= 0 && C >= 1 && C <= 256 && ld >= C && scal, "cg_direction: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  int rc = dlp_coldot(n, C, R, ld, Y, ld, nullptr, scal + 2 * C, dtype, stream);
  if (rc) return rc;
  if (n > 0) {
    DLP_CHECK_ARG(R && Y && P, "cg_direction: null pointer");
    const unsigned g = flat_grid((int64_t)n * C);
    DLP_TYPED(dtype, hipLaunchKernelGGL(k_cg_direction<T>, dim3(g), dim3(kBlock), 0, st, n, C, ld, scal,
                                        (const T*)Y, (T*)P, first));
  }
  hipLaunchKernelGGL(k_cg_commit_rz, dim3(1), dim3(kBlock), 0, st, C, scal);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_cg_residual(int n, int C, int64_t ld, int dtype, const int32_t* indptr, const int32_t* rowlen,
                               const int32_t* indices, const void* w, int w_dtype, const uint8_t* mask,
                               double avg_deg, const void* diag, const void* rhs, int64_t ldr, const void* Z, void* R,
                               const double* eps, double* viol, void* stream) {
  DLP_CHECK_ARG(n >= 0 && C >= 1 && C <= 256 && ld >= C && ldr >= C && eps && viol, "cg_residual: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(kBlock), 0, st, 1, viol, 0.0);
  DLP_CHECK_LAUNCH();
  if (n == 0) return DLP_OK;
  DLP_CHECK_ARG(indptr && indices && w && diag && rhs && Z && R, "cg_residual: null pointer");
  int rc = dlp_row_operator(n, nullptr, C, indptr, rowlen, indices, w, w_dtype, mask, Z, ld, dtype, diag, -1.0,
                            nullptr, 0, R, ld, dtype, 0, avg_deg, stream);
  if (rc) return rc;
  const unsigned g = flat_grid((int64_t)n * C);
  DLP_TYPED(dtype, hipLaunchKernelGGL(k_cg_residual<T>, dim3(g), dim3(kBlock), 0, st, n, C, ld, (const T*)rhs, ldr,
                                      (T*)R, (const T*)diag, eps, viol));
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_push_rounds(int n, int C, int64_t ld, int dtype, const int32_t* indptr, const int32_t* indices,
                               const void* w, const void* diag, const double* eps, void* Z, void* R, void* D,
                               int32_t* cand, int32_t* next, int32_t* counts, int32_t* stamp, int32_t* touched,
                               unsigned long long* edges, int rounds, void* stream) {
  DLP_CHECK_ARG(n >= 1 && C >= 1 && C <= 256 && ld >= C && rounds >= 1, "push_rounds: bad arguments");
  DLP_CHECK_ARG(indptr && indices && w && diag && eps && Z && R && D && cand && next && counts && stamp && touched,
                "push_rounds: null pointer");
  hipStream_t st = (hipStream_t)stream;
  int lc = 0;
  while ((1 << lc) < C && lc < 6) ++lc;
  const unsigned g_rows = flat_grid((int64_t)n);
  const unsigned g_elems = flat_grid((int64_t)n * C);
  const unsigned g_waves = flat_grid((int64_t)n * 64);
  for (int t = 0; t < rounds; ++t) {
    int32_t* cur = (t & 1) ? next : cand;
    int32_t* nxt = (t & 1) ? cand : next;
    DLP_TYPED(dtype, {
      hipLaunchKernelGGL(k_round_take<T>, dim3(g_elems), dim3(kBlock), 0, st, counts, cur, C, ld, (const T*)diag,
                         (T*)Z, (T*)R, (T*)D, stamp, touched, counts + 2);
      hipLaunchKernelGGL(k_round_scatter<T>, dim3(g_waves), dim3(kBlock), 0, st, counts, cur, C, ld, lc, indptr,
                         indices, (const T*)w, (const T*)D, (T*)R, stamp, touched, counts + 2, edges);
      hipLaunchKernelGGL(k_round_select<T>, dim3(g_rows), dim3(kBlock), 0, st, counts + 2, touched, C, ld,
                         (const T*)R, (const T*)diag, eps, stamp, nxt, counts + 1);
    });
    hipLaunchKernelGGL(k_round_rotate, dim3(1), dim3(64), 0, st, counts);
    DLP_CHECK_LAUNCH();
  }
  return DLP_OK;
}

extern "C" int dlp_park_rows(int n, const uint8_t* solved, int C1, int64_t ld, int dtype, void* Z, void* R,
                             void* stream) {
  DLP_CHECK_ARG(n >= 0 && C1 >= 1 && ld >= C1, "park_rows: bad arguments");
  if (n == 0) return DLP_OK;
  DLP_CHECK_ARG(solved && Z, "park_rows: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned g = flat_grid((int64_t)n * C1);
  DLP_TYPED(dtype, hipLaunchKernelGGL(k_park<T>, dim3(g), dim3(kBlock), 0, st, n, solved, C1, ld, (T*)Z, (T*)R));
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// ---------------------------------------------------------------------------------------------
// Row copies between the global store and a batch system, supernode averages, and the commit
// step of the frontier Jacobi.

namespace dlp {

template <typename TS, typename TD>
__global__ void k_rows_copy(int nrows, const int64_t* __restrict__ rows, int C, const TS* __restrict__ src,
                            int64_t lds, TD* __restrict__ dst, int64_t ldd, int gather) {
  const int64_t total = (int64_t)nrows * C;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / C;
    const int c = (int)(t - i * C);
    const int64_t r = rows ? rows[i] : i;
    if (gather) dst[i * ldd + c] = (TD)src[r * lds + c];
    else dst[r * ldd + c] = (TD)src[i * lds + c];
  }
}

// Per new vertex i (one wavefront each): S_i[c] = base_S + sum over entries j with known[col_j]
// of w_j X[col_j, c], T_i = base_T + sum of those w_j; base (mode 1) S = the label columns of B,
// T = the row sum of B; (mode 2) S = 0, T = row sum of B; (mode 0) both 0.  Then S and T of the
// vertex's component accumulate atomically.
template <typename T>
__global__ __launch_bounds__(256) void k_supernode_sums(int m, const int32_t* __restrict__ newrows,
                                                        const int32_t* __restrict__ lab,
                                                        const int32_t* __restrict__ indptr,
                                                        const int32_t* __restrict__ indices, const T* __restrict__ w,
                                                        const uint8_t* __restrict__ known, int C,
                                                        const T* __restrict__ X, int64_t ldx, const T* __restrict__ B,
                                                        int K, int base_mode, T* __restrict__ S, T* __restrict__ Tt) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves) {
    const int r = newrows[i];
    const int l = lab[i];
    T tsum = (T)0;
    if (known) {
      const int beg = indptr[r], end = indptr[r + 1];
      for (int j = beg + lane; j < end; j += 64) {
        const int u = indices[j];
        if (known[u]) tsum += w[j];
      }
    }
    for (int off = 32; off > 0; off >>= 1) tsum += __shfl_xor(tsum, off, 64);
    T brow = (T)0;
    if (base_mode) for (int k = 0; k < K; ++k) brow += B[(int64_t)r * K + k];
    if (lane == 0) atomicAdd(Tt + l, tsum + brow);
    for (int c = lane; c < C; c += 64) {
      T acc = (T)0;
      if (base_mode == 1) acc = K == 2 ? B[(int64_t)r * K + 1] : B[(int64_t)r * K + c];
      if (known) {
        const int beg = indptr[r], end = indptr[r + 1];
        for (int j = beg; j < end; ++j) {
          const int u = indices[j];
          if (known[u]) acc += w[j] * X[(int64_t)u * ldx + c];
        }
      }
      atomicAdd(S + (int64_t)l * C + c, acc);
    }
  }
}

template <typename T>
__global__ void k_supernode_write(int m, const int32_t* __restrict__ newrows, const int32_t* __restrict__ lab, int C,
                                  const T* __restrict__ S, const T* __restrict__ Tt, const T* __restrict__ fallback,
                                  T offset, T* __restrict__ out, int64_t ldo) {
  const int64_t total = (int64_t)m * C;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / C;
    const int c = (int)(t - i * C);
    const int l = lab[i];
    const T tt = Tt[l];
    out[(int64_t)newrows[i] * ldo + c] = (tt > (T)0 ? S[(int64_t)l * C + c] / tt : fallback[c]) + offset;
  }
}

template <typename T>
__global__ void k_zero_t(int64_t count, T* __restrict__ p) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (int64_t)gridDim.x * blockDim.x)
    p[t] = (T)0;
}

// Rows of the frontier whose value changed take it (X[f_g] = Y[g]) and mark their neighbours;
// *nchanged counts them.
template <typename T>
__global__ __launch_bounds__(256) void k_frontier_commit(int nf, const int32_t* __restrict__ frontier, int C,
                                                         int64_t ldx, const T* __restrict__ Y,
                                                         const uint8_t* __restrict__ changed, T* __restrict__ X,
                                                         const int32_t* __restrict__ indptr,
                                                         const int32_t* __restrict__ indices,
                                                         uint8_t* __restrict__ mark, int32_t* __restrict__ nchanged,
                                                         unsigned long long* __restrict__ edges) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t g = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); g < nf; g += waves) {
    if (!changed[g]) continue;
    const int r = frontier[g];
    for (int c = lane; c < C; c += 64) X[(int64_t)r * ldx + c] = Y[g * C + c];
    for (int j = indptr[r] + lane; j < indptr[r + 1]; j += 64) mark[indices[j]] = 1;
    if (lane == 0) {
      atomicAdd(nchanged, 1);
      if (edges) atomicAdd(edges, (unsigned long long)(indptr[r + 1] - indptr[r]));
    }
  }
}

}  // namespace dlp

extern "C" int dlp_rows_copy(int nrows, const int64_t* rows, int C, const void* src, int64_t lds, int src_dtype,
                             void* dst, int64_t ldd, int dst_dtype, int gather, void* stream) {
  DLP_CHECK_ARG(nrows >= 0 && C >= 1 && lds >= C && ldd >= C, "rows_copy: bad arguments");
  if (nrows == 0) return DLP_OK;
  DLP_CHECK_ARG(src && dst, "rows_copy: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned g = flat_grid((int64_t)nrows * C);
#define DLP_RC(TS, TD) \
  hipLaunchKernelGGL((k_rows_copy<TS, TD>), dim3(g), dim3(kBlock), 0, st, nrows, rows, C, (const TS*)src, lds, (TD*)dst, ldd, gather)
  if (src_dtype == DLP_F32 && dst_dtype == DLP_F32) DLP_RC(float, float);
  else if (src_dtype == DLP_F32 && dst_dtype == DLP_F64) DLP_RC(float, double);
  else if (src_dtype == DLP_F64 && dst_dtype == DLP_F32) DLP_RC(double, float);
  else if (src_dtype == DLP_F64 && dst_dtype == DLP_F64) DLP_RC(double, double);
  else DLP_CHECK_ARG(false, "rows_copy: bad dtypes %d %d", src_dtype, dst_dtype);
#undef DLP_RC
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" size_t dlp_supernode_workspace_bytes(int nc, int C, int dtype) {
  const size_t e = dtype == DLP_F64 ? 8 : 4;
  return (((size_t)nc * C * e + 255) & ~(size_t)255) + (size_t)nc * e + 256;
}

extern "C" int dlp_supernode_average(int m, const int32_t* newrows, const int32_t* lab, int nc, const int32_t* indptr,
                                     const int32_t* indices, const void* w, const uint8_t* known, int C,
                                     const void* X, int64_t ldx, const void* B, int K, int base_mode,
                                     const void* fallback, double offset, void* out, int64_t ldo, int dtype, void* ws,
                                     size_t ws_bytes, void* stream) {
  DLP_CHECK_ARG(m >= 0 && nc >= 0 && C >= 1 && base_mode >= 0 && base_mode <= 2, "supernode_average: bad arguments");
  if (m == 0) return DLP_OK;
  DLP_CHECK_ARG(newrows && lab && fallback && out && ws && (!known || (indptr && indices && w && X)) &&
                    (base_mode == 0 || (B && K >= 1 && (base_mode != 1 || K == 2 || K == C))),
                "supernode_average: null pointer or bad shape");
  DLP_CHECK_ARG(ws_bytes >= dlp_supernode_workspace_bytes(nc, C, dtype), "supernode_average: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  const size_t e = dtype == DLP_F64 ? 8 : 4;
  char* S = (char*)ws;
  char* Tt = S + ((((size_t)nc * C * e) + 255) & ~(size_t)255);
  const unsigned gz = flat_grid((int64_t)nc * (C + 1));
  const unsigned gw = flat_grid((int64_t)m * 64);
  const unsigned ge = flat_grid((int64_t)m * C);
  DLP_TYPED(dtype, {
    hipLaunchKernelGGL(k_zero_t<T>, dim3(gz), dim3(kBlock), 0, st, (int64_t)nc * C, (T*)S);
    hipLaunchKernelGGL(k_zero_t<T>, dim3(gz), dim3(kBlock), 0, st, (int64_t)nc, (T*)Tt);
    hipLaunchKernelGGL(k_supernode_sums<T>, dim3(gw), dim3(kBlock), 0, st, m, newrows, lab, indptr, indices,
                       (const T*)w, known, C, (const T*)X, ldx, (const T*)B, K, base_mode, (T*)S, (T*)Tt);
    hipLaunchKernelGGL(k_supernode_write<T>, dim3(ge), dim3(kBlock), 0, st, m, newrows, lab, C, (const T*)S,
                       (const T*)Tt, (const T*)fallback, (T)offset, (T*)out, ldo);
  });
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_frontier_commit(int nf, const int32_t* frontier, int C, int64_t ldx, const void* Y,
                                   const uint8_t* changed, void* X, const int32_t* indptr, const int32_t* indices,
                                   uint8_t* mark, int32_t* nchanged, unsigned long long* edges, int dtype,
                                   void* stream) {
  DLP_CHECK_ARG(nf >= 0 && C >= 1 && ldx >= C && nchanged, "frontier_commit: bad arguments");
  if (nf == 0) return DLP_OK;
  DLP_CHECK_ARG(frontier && Y && changed && X && indptr && indices && mark, "frontier_commit: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned g = flat_grid((int64_t)nf * 64);
  DLP_TYPED(dtype, hipLaunchKernelGGL(k_frontier_commit<T>, dim3(g), dim3(kBlock), 0, st, nf, frontier, C, ldx,
                                      (const T*)Y, changed, (T*)X, indptr, indices, mark, nchanged, edges));
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// Aggregation AMG on the device (reference amg.py:1-134): heavy-edge handshake matching,
// Galerkin coarsening by aggregates (W_c = off-diagonal part of P^T W P, s_c = P^T s) and the
// restriction / prolongation of the V-cycle.
//
// Matching choices are made row-locally over CSR rows (a row is scanned by one thread), which
// gives the reference's scatter_max/scatter_min result (heaviest weight, ties -> smallest
// neighbour id, amg.py:38-44) without atomics.  Coarse rows are built by a counting sort on
// the aggregate id, a segmented radix sort of every row's columns (rocPRIM: coarse rows of an
// expander graph hold thousands of entries after a few levels) and a duplicate merge per row.
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "common.h"
#include "scan.h"

namespace dlp {

// mode 0: entries with both ends unmatched (amg.py:47-50); mode 1: r unmatched, c matched
// (amg.py:56-58).  choice[r] = best column or n.
template <typename T>
__global__ void k_best_choice(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                              const T* __restrict__ w, const uint8_t* __restrict__ matched, int mode,
                              int32_t* __restrict__ choice) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int best_c = n;
  T best_w = 0;
  const bool mr = matched[r] != 0;
  if (!mr) {
    for (int j = indptr[r]; j < indptr[r + 1]; ++j) {
      const int c = indices[j];
      const bool ok = mode == 0 ? !matched[c] : matched[c] != 0;
      if (!ok) continue;
      const T wv = w[j];
      if (best_c == n || wv > best_w || (wv == best_w && c < best_c)) {
        best_w = wv;
        best_c = c;
      }
    }
  }
  choice[r] = best_c;
}

__global__ void k_mutual(int n, const int32_t* __restrict__ choice, uint8_t* __restrict__ matched,
                         int32_t* __restrict__ agg) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int p = choice[r];
  if (p < n && choice[p] == r) {
    matched[r] = 1;
    agg[r] = min(r, p);
  }
}

__global__ void k_attach(int n, const int32_t* __restrict__ choice, int32_t* __restrict__ agg) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int p = choice[r];
  if (p < n) agg[r] = agg[p];
}

__global__ void k_agg_init(int n, int32_t* __restrict__ agg, uint8_t* __restrict__ matched) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  agg[r] = -1;
  matched[r] = 0;
}

__global__ void k_left_flag(int n, int32_t* __restrict__ agg, int32_t* __restrict__ flag) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int a = agg[r];
  if (a < 0) {
    a = r;
    agg[r] = r;
  }
  flag[a] = 1;
}

__global__ void k_zero_i32(int n, int32_t* __restrict__ p) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) p[r] = 0;
}

__global__ void k_relabel(int n, int32_t* __restrict__ agg, const int32_t* __restrict__ rank) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) agg[r] = rank[agg[r]];
}

__global__ void k_compose(int n, const int32_t* __restrict__ prev, const int32_t* __restrict__ a2,
                          int32_t* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) out[r] = a2[prev[r]];
}

__global__ void k_coarse_count(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                               const int32_t* __restrict__ agg, int32_t* __restrict__ cnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int ar = agg[r];
  int k = 0;
  for (int j = indptr[r]; j < indptr[r + 1]; ++j) k += agg[indices[j]] != ar;
  if (k) atomicAdd(cnt + ar, k);
}

template <typename T>
__global__ void k_coarse_fill(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                              const T* __restrict__ w, const int32_t* __restrict__ agg,
                              int32_t* __restrict__ cursor, int32_t* __restrict__ ci, T* __restrict__ cw) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int ar = agg[r];
  const int beg = indptr[r], end = indptr[r + 1];
  // every entry of fine row r lands in coarse row agg[r]: one reservation for all of them (the
  // few coarse rows of a deep level would otherwise take one atomic per fine entry)
  int k = 0;
  for (int j = beg; j < end; ++j) k += agg[indices[j]] != ar;
  if (k == 0) return;
  int pos = atomicAdd(cursor + ar, k);
  for (int j = beg; j < end; ++j) {
    const int ac = agg[indices[j]];
    if (ac != ar) {
      ci[pos] = ac;
      cw[pos] = w[j];
      ++pos;
    }
  }
}

// Duplicate merge of row-sorted entries (sk, sv), one wavefront per row: every run of equal
// columns becomes one entry (weights summed) at the front of the row's segment of ci/cw, unique
// count into ucnt.  The row is read in chunks of 64 entries; inside a chunk a segmented
// inclusive scan over the lanes sums each run up to every lane, the run's last lane writes it
// at the position the earlier run ends of the chunk give (ballot + count), and a run still open
// at the chunk's end is carried into the next chunk (added to its first run when that run has
// the same column, else written before it).
template <typename T>
__global__ __launch_bounds__(256) void k_coarse_merge_sorted(int nc, const int32_t* __restrict__ off,
                                                             const uint32_t* __restrict__ sk,
                                                             const T* __restrict__ sv, int32_t* __restrict__ ci,
                                                             T* __restrict__ cw, int32_t* __restrict__ ucnt) {
  const int r = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= nc) return;   // wave-uniform
  const int beg = off[r], end = off[r + 1];
  const unsigned long long below = (1ull << lane) - 1ull;
  int o = beg;
  int carry_key = -1;   // the run left open by the previous chunk (-1: none)
  T carry_val = T(0);
  for (int base = beg; base < end; base += 64) {
    const int cnt = min(64, end - base);
    const bool valid = lane < cnt;
    const int key = valid ? (int)sk[base + lane] : -2;
    T val = valid ? sv[base + lane] : T(0);
    // a carried run that this chunk does not continue is complete: write it first
    if (carry_key >= 0 && __shfl(key, 0, 64) != carry_key) {
      if (lane == 0) {
        ci[o] = carry_key;
        cw[o] = carry_val;
      }
      ++o;
      carry_key = -1;
    }
    const int prev = __shfl_up(key, 1, 64);
    const bool head = valid && (lane == 0 ? key != carry_key : key != prev);
    // segmented inclusive scan: f marks a run start at or below this lane within the window
    bool f = head;
    for (int d = 1; d < 64; d <<= 1) {
      const T up_val = __shfl_up(val, d, 64);
      const int up_f = __shfl_up((int)f, d, 64);
      if (lane >= d && !f) val += up_val;
      f = f || (lane >= d && up_f);
    }
    // lanes of the chunk's first run continue the carried run when it has the same key
    if (valid && !f && carry_key >= 0) val += carry_val;
    const int next = __shfl_down(key, 1, 64);
    const bool last_lane = lane == cnt - 1;
    const bool open_end = last_lane && base + cnt < end;   // may continue in the next chunk
    const bool tail = valid && (last_lane ? !open_end : next != key);
    const unsigned long long tails = __ballot(tail);
    if (tail) {
      const int pos = o + __popcll(tails & below);
      ci[pos] = key;
      cw[pos] = val;
    }
    o += __popcll(tails);
    // the open run (if any) moves on as the carry
    const int last_key = __shfl(key, cnt - 1, 64);
    const T last_val = __shfl(val, cnt - 1, 64);
    if (base + cnt < end) {
      carry_key = last_key;
      carry_val = last_val;
    } else {
      carry_key = -1;
    }
  }
  if (lane == 0) ucnt[r] = o - beg;
}

template <typename T>
__global__ void k_coarse_compact(int nc, const int32_t* __restrict__ off, const int32_t* __restrict__ ci,
                                 const T* __restrict__ cw, const int32_t* __restrict__ optr,
                                 int32_t* __restrict__ oi, T* __restrict__ ow) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nc) return;
  const int b = off[r], o = optr[r], k = optr[r + 1] - o;
  for (int i = 0; i < k; ++i) {
    oi[o + i] = ci[b + i];
    ow[o + i] = cw[b + i];
  }
}

template <typename T>
__global__ void k_scatter_add_rows(int n, int C, const int32_t* __restrict__ agg, const T* __restrict__ X,
                                   int64_t ldx, T* __restrict__ Y, int64_t ldy) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * C) return;
  const int64_t r = t / C;
  const int c = (int)(t - r * C);
  atomic_add(Y + (int64_t)agg[r] * ldy + c, X[r * ldx + c]);
}

template <typename T>
__global__ void k_gather_add_rows(int n, int C, const int32_t* __restrict__ agg, const T* __restrict__ Xc,
                                  int64_t ldc, T* __restrict__ X, int64_t ldx) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * C) return;
  const int64_t r = t / C;
  const int c = (int)(t - r * C);
  X[r * ldx + c] += Xc[(int64_t)agg[r] * ldc + c];
}

}  // namespace dlp

using namespace dlp;

// Handshake matching (amg.py:32-64) over the CSR of one level: agg (out, int32[n]) holds the
// renumbered aggregate of every row; *nc_out (host) the aggregate count.  Workspace ints:
// choice(n) + flag(n) + scan workspace; matched bytes(n).
extern "C" int dlp_amg_match(int n, const int32_t* indptr, const int32_t* indices, const void* w, int dtype,
                             int rounds, int32_t* agg, int32_t* choice, int32_t* flag, uint8_t* matched,
                             void* scan_ws, size_t scan_bytes, int32_t* nc_dev, void* stream) {
  DLP_CHECK_ARG(n >= 1 && indptr && indices && w && agg && choice && flag && matched && nc_dev,
                "amg_match: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const dim3 g(grid_for(n)), b(kBlock);
  hipLaunchKernelGGL(k_agg_init, g, b, 0, st, n, agg, matched);
  for (int r = 0; r < rounds; ++r) {
    if (dtype == DLP_F32)
      hipLaunchKernelGGL(k_best_choice<float>, g, b, 0, st, n, indptr, indices, (const float*)w, matched, 0, choice);
    else
      hipLaunchKernelGGL(k_best_choice<double>, g, b, 0, st, n, indptr, indices, (const double*)w, matched, 0,
                         choice);
    hipLaunchKernelGGL(k_mutual, g, b, 0, st, n, choice, matched, agg);
  }
  if (dtype == DLP_F32)
    hipLaunchKernelGGL(k_best_choice<float>, g, b, 0, st, n, indptr, indices, (const float*)w, matched, 1, choice);
  else
    hipLaunchKernelGGL(k_best_choice<double>, g, b, 0, st, n, indptr, indices, (const double*)w, matched, 1, choice);
  hipLaunchKernelGGL(k_attach, g, b, 0, st, n, choice, agg);
  hipLaunchKernelGGL(k_zero_i32, g, b, 0, st, n, flag);
  hipLaunchKernelGGL(k_left_flag, g, b, 0, st, n, agg, flag);
  DLP_CHECK_LAUNCH();
  int rc = exclusive_scan_i32(n, flag, flag, nc_dev, scan_ws, scan_bytes, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_relabel, g, b, 0, st, n, agg, flag);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_amg_compose(int n, const int32_t* prev, const int32_t* a2, int32_t* out, void* stream) {
  DLP_CHECK_ARG(n >= 0, "amg_compose: bad n");
  if (n == 0) return DLP_OK;
  hipLaunchKernelGGL(k_compose, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream, n, prev, a2, out);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// Coarse operator, phase 1: per coarse row counts (cnt, nc+1 ints zeroed here) and the
// unmerged entries (ci/cw of capacity nnz) in row segments off (exclusive scan, nc+1).
// Phase 2 (dlp_amg_coarsen_merge): sort + merge per row, unique counts into ucnt.
extern "C" int dlp_amg_coarsen_fill(int n, int nc, const int32_t* indptr, const int32_t* indices, const void* w,
                                    int dtype, const int32_t* agg, int32_t* off, int32_t* cursor, int32_t* ci,
                                    void* cw, void* scan_ws, size_t scan_bytes, void* stream) {
  DLP_CHECK_ARG(n >= 1 && nc >= 1, "amg_coarsen_fill: bad sizes");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(nc + 1)), dim3(kBlock), 0, st, nc + 1, off);
  hipLaunchKernelGGL(k_coarse_count, dim3(grid_for(n)), dim3(kBlock), 0, st, n, indptr, indices, agg, off);
  DLP_CHECK_LAUNCH();
  int rc = exclusive_scan_i32(nc + 1, off, off, nullptr, scan_ws, scan_bytes, st);
  if (rc) return rc;
  DLP_CHECK_HIP(hipMemcpyAsync(cursor, off, sizeof(int32_t) * (nc + 1), hipMemcpyDeviceToDevice, st));
  if (dtype == DLP_F32)
    hipLaunchKernelGGL(k_coarse_fill<float>, dim3(grid_for(n)), dim3(kBlock), 0, st, n, indptr, indices,
                       (const float*)w, agg, cursor, ci, (float*)cw);
  else
    hipLaunchKernelGGL(k_coarse_fill<double>, dim3(grid_for(n)), dim3(kBlock), 0, st, n, indptr, indices,
                       (const double*)w, agg, cursor, ci, (double*)cw);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

template <typename T>
static int coarsen_merge_sorted(int nc, const int32_t* off, int32_t* ci, T* cw, int32_t* ucnt, hipStream_t st) {
  int total = 0;
  DLP_CHECK_HIP(hipMemcpyAsync(&total, off + nc, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  DLP_CHECK_HIP(hipStreamSynchronize(st));
  if (total <= 0) {
    hipLaunchKernelGGL(k_zero_i32, dim3(grid_for(nc)), dim3(kBlock), 0, st, nc, ucnt);
    DLP_CHECK_LAUNCH();
    return DLP_OK;
  }
  unsigned bits = 1;
  while (bits < 31 && (1u << bits) < (unsigned)nc) ++bits;   // columns are coarse ids < nc
  // sorted copies (keys, values) and the sort's temporary storage, stream-ordered
  size_t tmp_bytes = 0;
  const uint32_t* keys = reinterpret_cast<const uint32_t*>(ci);
  DLP_CHECK_HIP(rocprim::segmented_radix_sort_pairs(nullptr, tmp_bytes, keys, (uint32_t*)nullptr, cw, (T*)nullptr,
                                                    (unsigned)total, (unsigned)nc, off, off + 1, 0, bits, st));
  const size_t kb = ((size_t)total * sizeof(uint32_t) + 255) & ~(size_t)255;
  const size_t vb = ((size_t)total * sizeof(T) + 255) & ~(size_t)255;
  void* buf = nullptr;
  DLP_CHECK_HIP(hipMallocAsync(&buf, kb + vb + tmp_bytes, st));
  uint32_t* sk = reinterpret_cast<uint32_t*>(buf);
  T* sv = reinterpret_cast<T*>(reinterpret_cast<char*>(buf) + kb);
  void* tmp = reinterpret_cast<char*>(buf) + kb + vb;
  hipError_t e = rocprim::segmented_radix_sort_pairs(tmp, tmp_bytes, keys, sk, cw, sv, (unsigned)total, (unsigned)nc,
                                                     off, off + 1, 0, bits, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_coarse_merge_sorted<T>, dim3(grid_for((long long)nc * 64)), dim3(kBlock), 0, st, nc, off,
                       sk, sv, ci, cw, ucnt);
    e = hipGetLastError();
  }
  const hipError_t ef = hipFreeAsync(buf, st);
  DLP_CHECK_HIP(e);
  DLP_CHECK_HIP(ef);
  return DLP_OK;
}

extern "C" int dlp_amg_coarsen_merge(int nc, const int32_t* off, int32_t* ci, void* cw, int dtype, int32_t* ucnt,
                                     void* scan_ws, size_t scan_bytes, int32_t* total_dev, void* stream) {
  DLP_CHECK_ARG(nc >= 1, "amg_coarsen_merge: bad nc");
  hipStream_t st = (hipStream_t)stream;
  int rc = dtype == DLP_F32 ? coarsen_merge_sorted<float>(nc, off, ci, (float*)cw, ucnt, st)
                            : coarsen_merge_sorted<double>(nc, off, ci, (double*)cw, ucnt, st);
  if (rc) return rc;
  // ucnt has nc+1 slots: scan to row pointers, total in total_dev
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, 1, ucnt + nc);
  DLP_CHECK_LAUNCH();
  return exclusive_scan_i32(nc + 1, ucnt, ucnt, total_dev, scan_ws, scan_bytes, st);
}

extern "C" int dlp_amg_coarsen_compact(int nc, const int32_t* off, const int32_t* ci, const void* cw, int dtype,
                                       const int32_t* optr, int32_t* oi, void* ow, void* stream) {
  DLP_CHECK_ARG(nc >= 1, "amg_coarsen_compact: bad nc");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == DLP_F32)
    hipLaunchKernelGGL(k_coarse_compact<float>, dim3(grid_for(nc)), dim3(kBlock), 0, st, nc, off, ci,
                       (const float*)cw, optr, oi, (float*)ow);
  else
    hipLaunchKernelGGL(k_coarse_compact<double>, dim3(grid_for(nc)), dim3(kBlock), 0, st, nc, off, ci,
                       (const double*)cw, optr, oi, (double*)ow);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// Y[agg[r]] += X[r] (restriction P^T, amg.py:72,129) — Y zeroed by the caller.
extern "C" int dlp_amg_restrict(int n, int C, const int32_t* agg, const void* X, int64_t ldx, void* Y, int64_t ldy,
                                int dtype, void* stream) {
  DLP_CHECK_ARG(n >= 0 && C >= 1, "amg_restrict: bad sizes");
  if (n == 0) return DLP_OK;
  hipStream_t st = (hipStream_t)stream;
  const dim3 g(grid_for((long long)n * C)), b(kBlock);
  if (dtype == DLP_F32)
    hipLaunchKernelGGL(k_scatter_add_rows<float>, g, b, 0, st, n, C, agg, (const float*)X, ldx, (float*)Y, ldy);
  else
    hipLaunchKernelGGL(k_scatter_add_rows<double>, g, b, 0, st, n, C, agg, (const double*)X, ldx, (double*)Y, ldy);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// X[r] += Xc[agg[r]] (prolongation, amg.py:130).
extern "C" int dlp_amg_prolong_add(int n, int C, const int32_t* agg, const void* Xc, int64_t ldc, void* X,
                                   int64_t ldx, int dtype, void* stream) {
  DLP_CHECK_ARG(n >= 0 && C >= 1, "amg_prolong: bad sizes");
  if (n == 0) return DLP_OK;
  hipStream_t st = (hipStream_t)stream;
  const dim3 g(grid_for((long long)n * C)), b(kBlock);
  if (dtype == DLP_F32)
    hipLaunchKernelGGL(k_gather_add_rows<float>, g, b, 0, st, n, C, agg, (const float*)Xc, ldc, (float*)X, ldx);
  else
    hipLaunchKernelGGL(k_gather_add_rows<double>, g, b, 0, st, n, C, agg, (const double*)Xc, ldc, (double*)X, ldx);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// ---------------------------------------------------------------------------------------------
// Coarsest level in FP64: the dense operator, its inverse, and the FP64 application.

namespace dlp {

// A = diag(diag) - W as a dense FP64 n x n matrix (the caller zeroes A); one thread per row.
template <typename T>
__global__ void k_dense_operator(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                                 const T* __restrict__ w, const T* __restrict__ diag, double* __restrict__ A) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double* row = A + (int64_t)r * n;
  for (int j = indptr[r]; j < indptr[r + 1]; ++j) row[indices[j]] -= (double)w[j];
  row[r] += (double)diag[r];
}

__global__ void k_zero_f64n(int64_t count, double* __restrict__ p) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (int64_t)gridDim.x * blockDim.x)
    p[t] = 0.0;
}

// One Gauss-Jordan step on pivot k, out of place (every new entry from the old matrix):
//   i = j = k: 1/p;  i = k: a_kj / p;  j = k: -a_ik / p;  else a_ij - a_ik a_kj / p,
// with p = a_kk.  After the n steps the matrix is its inverse (no pivoting: the operator is
// symmetric positive definite, so every pivot is positive).
__global__ void k_gauss_jordan_step(int n, int k, const double* __restrict__ A, double* __restrict__ B) {
  const int64_t total = (int64_t)n * n;
  const double p = A[(int64_t)k * n + k];
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t / n), j = (int)(t - (int64_t)i * n);
    const double aik = A[(int64_t)i * n + k], akj = A[(int64_t)k * n + j];
    double v;
    if (i == k) v = (j == k) ? 1.0 / p : akj / p;
    else if (j == k) v = -aik / p;
    else v = A[t] - aik * akj / p;
    B[t] = v;
  }
}

// X[r, c] = sum_j A[r, j] R[j, c], FP64 accumulation; one wavefront per output row, its lanes
// over the columns (rows of R read once per wavefront through the lanes).
template <typename T>
__global__ __launch_bounds__(256) void k_dense_apply(int n, int C, const double* __restrict__ A,
                                                     const T* __restrict__ R, int64_t ldr, T* __restrict__ X,
                                                     int64_t ldx) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < n; r += waves) {
    const double* arow = A + r * n;
    for (int c0 = 0; c0 < C; c0 += 64) {
      const int c = c0 + lane;
      double acc = 0.0;
      if (c < C)
        for (int j = 0; j < n; ++j) acc += arow[j] * (double)R[(int64_t)j * ldr + c];
      if (c < C) X[r * ldx + c] = (T)acc;
    }
  }
}

}  // namespace dlp

extern "C" int dlp_dense_operator_f64(int n, const int32_t* indptr, const int32_t* indices, const void* w,
                                      const void* diag, int dtype, double* A, void* stream) {
  DLP_CHECK_ARG(n >= 0 && n <= 16384, "dense_operator: n=%d outside [0, 16384]", n);
  if (n == 0) return DLP_OK;
  DLP_CHECK_ARG(indptr && indices && w && diag && A, "dense_operator: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nn = (int64_t)n * n;
  long long g = (long long)grid_for(nn);
  hipLaunchKernelGGL(k_zero_f64n, dim3((unsigned)(g > 8192 ? 8192 : g)), dim3(kBlock), 0, st, nn, A);
  if (dtype == DLP_F64)
    hipLaunchKernelGGL(k_dense_operator<double>, dim3(grid_for(n)), dim3(kBlock), 0, st, n, indptr, indices,
                       (const double*)w, (const double*)diag, A);
  else
    hipLaunchKernelGGL(k_dense_operator<float>, dim3(grid_for(n)), dim3(kBlock), 0, st, n, indptr, indices,
                       (const float*)w, (const float*)diag, A);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_spd_inverse_f64(int n, double* A, double* scratch, void* stream) {
  DLP_CHECK_ARG(n >= 0 && n <= 16384, "spd_inverse: n=%d outside [0, 16384]", n);
  if (n == 0) return DLP_OK;
  DLP_CHECK_ARG(A && scratch, "spd_inverse: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const int64_t nn = (int64_t)n * n;
  long long g = (long long)grid_for(nn);
  const dim3 grid((unsigned)(g > 8192 ? 8192 : g));
  double* src = A;
  double* dst = scratch;
  for (int k = 0; k < n; ++k) {
    hipLaunchKernelGGL(k_gauss_jordan_step, grid, dim3(kBlock), 0, st, n, k, src, dst);
    double* t = src;
    src = dst;
    dst = t;
  }
  DLP_CHECK_LAUNCH();
  if (src != A) DLP_CHECK_HIP(hipMemcpyAsync(A, src, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
  return DLP_OK;
}

extern "C" int dlp_dense_apply_f64(int n, int C, const double* A, const void* R, int64_t ldr, void* X, int64_t ldx,
                                   int dtype, void* stream) {
  DLP_CHECK_ARG(n >= 0 && C >= 1 && ldr >= C && ldx >= C, "dense_apply: bad arguments");
  if (n == 0) return DLP_OK;
  DLP_CHECK_ARG(A && R && X, "dense_apply: null pointer");
  hipStream_t st = (hipStream_t)stream;
  long long g = (long long)grid_for((int64_t)n * 64);
  const dim3 grid((unsigned)(g > 4096 ? 4096 : g));
  if (dtype == DLP_F64)
    hipLaunchKernelGGL(k_dense_apply<double>, grid, dim3(kBlock), 0, st, n, C, A, (const double*)R, ldr, (double*)X,
                       ldx);
  else
    hipLaunchKernelGGL(k_dense_apply<float>, grid, dim3(kBlock), 0, st, n, C, A, (const float*)R, ldr, (float*)X, ldx);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// Frontier primitives: Jacobi on a vertex subset, residual scatter and neighbour marking.
#include "blas.h"
#include "rowpipe.h"

namespace dlp {

// Y[g] = (rhs[f] + sum w X[col]) / diag[f]; changed[g] = any |Y - X[f]| > delta.
template <typename T, int CPL>
__global__ __launch_bounds__(256) void k_frontier_jacobi(
    int nf, const int32_t* __restrict__ frontier, int C, int g_shift, int lc_shift,
    const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
    const T* __restrict__ w, const T* __restrict__ X, const T* __restrict__ rhs,
    const T* __restrict__ diag, T delta, T* __restrict__ Y, uint8_t* __restrict__ changed) {
  const Lanes L(g_shift, lc_shift);
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> g_shift;
  const bool valid = g < nf;
  const int r = valid ? frontier[g] : 0;
  const int beg = valid ? indptr[r] : 0, end = valid ? indptr[r + 1] : 0;
  T acc[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) acc[k] = 0;
  row_gather<T, T, T, CPL>(L, beg, end, indices, w, nullptr, X, C, C, acc);
  group_reduce<T, CPL>(L, acc);
  bool ch = false;
  if (valid && L.e == 0) {
    const T inv = (T)1 / diag[r];
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = L.l + k * L.LC;
      if (c < C) {
        const T y = (rhs[(int64_t)r * C + c] + acc[k]) * inv;
        Y[g * C + c] = y;
        ch |= fabs(y - X[(int64_t)r * C + c]) > delta;
      }
    }
  }
  const unsigned long long bal = __ballot(ch);
  const int wl = threadIdx.x & 63;
  const unsigned long long gm = (L.G == 64 ? ~0ull : ((1ull << L.G) - 1)) << (wl & ~(L.G - 1));
  if (valid && L.lane == 0) changed[g] = (bal & gm) != 0;
}

// FP32 Jacobi on the pipelined packed-lane gathers (rowpipe.h): row r = frontier[g],
// Y[g] = (rhs[r] + sum w X[col]) / diag[r], changed[g] = any |Y[g] - X[r]| > delta.  X and rhs
// are n x C, Y is nf x C, all contiguous (4-byte aligned rows).
template <Word K>
struct JacobiEpilogue {
  const LanesP& L;
  int C;
  bool owner;
  const float* __restrict__ X;
  const float* __restrict__ rhs;
  const float* __restrict__ diag;
  float delta;
  float* __restrict__ Y;
  uint8_t* __restrict__ changed;
  bool vec_out;
  float xr[4], br[4];
  float inv;
  __device__ __forceinline__ void pre(int, int r) {
    if (owner) {
      load4(X + (int64_t)r * C, 4 * L.l, C, xr);
      load4(rhs + (int64_t)r * C, 4 * L.l, C, br);
    }
    inv = 1.f / diag[r];
  }
  __device__ __forceinline__ void post(int g, int, bool, const float (&acc)[4]) {
    bool ch = false;
    if (owner) {
      float y[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        y[q] = (br[q] + acc[q]) * inv;
        ch |= (4 * L.l + q < C) && fabsf(y[q] - xr[q]) > delta;
      }
      store4(Y + (int64_t)g * C, 4 * L.l, C, vec_out, y);
    }
    const unsigned long long bal = __ballot(ch);
    if ((threadIdx.x & 63) == 0) changed[g] = bal != 0;
  }
};

template <Word K>
__global__ __launch_bounds__(256) void k_frontier_jacobi_rows(int nf, const int32_t* __restrict__ frontier, int C,
                                                              const int32_t* __restrict__ indptr,
                                                              const int32_t* __restrict__ indices,
                                                              const float* __restrict__ w, const float* __restrict__ X,
                                                              int xbytes, const float* __restrict__ rhs,
                                                              const float* __restrict__ diag, float delta,
                                                              float* __restrict__ Y, uint8_t* __restrict__ changed,
                                                              int vec_out) {
  RowPipe G(nf, C, 4, 16, 0, nullptr, indptr, nullptr, indices, w, X, (unsigned)(C * 4), xbytes);
  G.rows = frontier;
  JacobiEpilogue<K> E{G.L, C, G.L.e == 0 && 4 * G.L.l < C, X, rhs, diag, delta, Y, changed, vec_out != 0};
  run_rows<4, K, float>(G, E);
}

// FP32 push scatter, one frontier row per wavefront step: 2^lc_shift lanes cover consecutive
// columns of one edge (CPL chunks of them), 64 >> lc_shift edges per instruction, so every
// atomic instruction adds to consecutive dwords of a row of R.  A row's (column, weight) pairs
// are read 64 at a time, one per lane, and broadcast from registers, so the atomics of
// consecutive edges do not wait on index loads.  Each lane marks the column of its own pair.
template <int CPL>
__global__ __launch_bounds__(256) void k_push_scatter_rows(int nf, const int32_t* __restrict__ frontier, int C,
                                                           int lc_shift, const int32_t* __restrict__ indptr,
                                                           const int32_t* __restrict__ indices,
                                                           const float* __restrict__ w, const float* __restrict__ D,
                                                           float* __restrict__ R, uint8_t* __restrict__ mark) {
  const int LC = 1 << lc_shift, LE = 64 >> lc_shift;
  const int lane = threadIdx.x & 63, l = lane & (LC - 1), e = lane >> lc_shift;
  const int64_t nwx = (int64_t)gridDim.x * 4;
  for (int64_t gg = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); gg < nf; gg += nwx) {
    const int g = __builtin_amdgcn_readfirstlane((int)gg);
    const int r = __builtin_amdgcn_readfirstlane(frontier[g]);
    const int beg = __builtin_amdgcn_readfirstlane(indptr[r]);
    const int end = __builtin_amdgcn_readfirstlane(indptr[r + 1]);
    float d[CPL];
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = l + k * LC;
      d[k] = c < C ? D[(int64_t)g * C + c] : 0.f;
    }
    for (int base = beg; base < end; base += 64) {
      const int cnt = min(64, end - base);
      int v = 0;
      float wv = 0.f;
      if (lane < cnt) {
        v = indices[base + lane];
        wv = w[base + lane];
        mark[v] = 1;
      }
      for (int t0 = 0; t0 < cnt; t0 += LE) {
        const int t = t0 + e;
        const int vv = __shfl(v, t & 63, 64);
        const float ww = __shfl(wv, t & 63, 64);
        if (t < cnt) {
          float* rp = R + (int64_t)vv * C;
#pragma unroll
          for (int k = 0; k < CPL; ++k) {
            const int c = l + k * LC;
            if (c < C) atomicAdd(rp + c, ww * d[k]);
          }
        }
      }
    }
  }
}

// R[col] += w * D[g] and mark[col] = 1 over the rows of the frontier.
template <typename T, int CPL>
__global__ __launch_bounds__(256) void k_push_scatter(
    int nf, const int32_t* __restrict__ frontier, int C, int g_shift, int lc_shift,
    const int32_t* __restrict__ indptr, const int32_t* __restrict__ indices,
    const T* __restrict__ w, const T* __restrict__ D, T* __restrict__ R,
    uint8_t* __restrict__ mark) {
  const Lanes L(g_shift, lc_shift);
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> g_shift;
  const bool valid = g < nf;
  const int r = valid ? frontier[g] : 0;
  const int beg = valid ? indptr[r] : 0, end = valid ? indptr[r + 1] : 0;
  T d[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int c = L.l + k * L.LC;
    d[k] = (valid && c < C) ? D[g * C + c] : (T)0;
  }
  for (int j = beg + L.e; j < end; j += L.LE) {
    const int v = indices[j];
    const T wv = w[j];
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = L.l + k * L.LC;
      if (c < C) atomic_add(&R[(int64_t)v * C + c], wv * d[k]);
    }
    if (L.l == 0) mark[v] = 1;
  }
}

__global__ __launch_bounds__(256) void k_mark_neighbors(int nr, const int32_t* __restrict__ rows,
                                                        const int32_t* __restrict__ indptr,
                                                        const int32_t* __restrict__ indices,
                                                        uint8_t* __restrict__ mark) {
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 3;
  const int lane = threadIdx.x & 7;
  const bool valid = g < nr;
  const int r = valid ? rows[g] : 0;
  const int beg = valid ? indptr[r] : 0, end = valid ? indptr[r + 1] : 0;
  for (int j = beg + lane; j < end; j += 8) mark[indices[j]] = 1;
}

// edges += sum of the row lengths of rows[0..nr): a grid-stride pass with one atomic per
// workgroup (a counter updated by every wavefront of a large frontier serialises at one L2
// address and costs more than the gathers themselves).
__global__ __launch_bounds__(256) void k_count_edges(int nr, const int32_t* __restrict__ rows,
                                                     const int32_t* __restrict__ indptr,
                                                     unsigned long long* __restrict__ edges) {
  __shared__ unsigned long long part[4];
  unsigned long long s = 0;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nr; g += (int64_t)gridDim.x * blockDim.x) {
    const int r = rows[g];
    s += (unsigned long long)(indptr[r + 1] - indptr[r]);
  }
  s = wave_sum_u64(s);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = part[0] + part[1] + part[2] + part[3];
    if (t) atomicAdd(edges, t);
  }
}

static void count_edges(int nr, const int32_t* rows, const int32_t* indptr, unsigned long long* edges,
                        hipStream_t st) {
  if (!edges || nr <= 0) return;
  unsigned g = grid_for(nr);
  if (g > 512) g = 512;
  hipLaunchKernelGGL(k_count_edges, dim3(g), dim3(kBlock), 0, st, nr, rows, indptr, edges);
}

}  // namespace dlp

using namespace dlp;


// The packed form for FP32 (returns false when its limits do not hold).
static bool frontier_jacobi_rows(int nf, const int32_t* frontier, int C, const int32_t* indptr,
                                 const int32_t* indices, const float* w, const float* X, const float* rhs,
                                 const float* diag, float delta, float* Y, uint8_t* changed, int n_hint,
                                 hipStream_t st) {
  const long long xbytes = (long long)n_hint * C * 4;
  if (!tuning().pipe || n_hint <= 0 || n_hint >= (1 << 24) || xbytes >= (1LL << 31) || (((uintptr_t)X) & 3) != 0)
    return false;
  const bool v16 = (C % 4) == 0 && (((uintptr_t)X) & 15) == 0;
  const int vec_out = (C % 4) == 0 && (((uintptr_t)Y) & 15) == 0;
  const void* kfn = v16 ? reinterpret_cast<const void*>(&k_frontier_jacobi_rows<Word::F32x4>)
                        : reinterpret_cast<const void*>(&k_frontier_jacobi_rows<Word::F32x4u>);
  long long grid = resident_grid(kfn);
  const long long need = ((long long)nf + 3) / 4;
  if (grid > need) grid = need;
  if (v16)
    hipLaunchKernelGGL((k_frontier_jacobi_rows<Word::F32x4>), dim3((unsigned)grid), dim3(kBlock), 0, st, nf, frontier,
                       C, indptr, indices, w, X, (int)xbytes, rhs, diag, delta, Y, changed, vec_out);
  else
    hipLaunchKernelGGL((k_frontier_jacobi_rows<Word::F32x4u>), dim3((unsigned)grid), dim3(kBlock), 0, st, nf,
                       frontier, C, indptr, indices, w, X, (int)xbytes, rhs, diag, delta, Y, changed, vec_out);
  return true;
}

template <typename T>
static int frontier_jacobi_t(int nf, const int32_t* frontier, int C, const int32_t* indptr,
                             const int32_t* indices, const void* w, const void* X,
                             const void* rhs, const void* diag, double delta, void* Y,
                             uint8_t* changed, unsigned long long* edges, double avg_deg,
                             hipStream_t st) {
  RowMap m = choose_map(C, avg_deg);
  dim3 grid(grid_for((long long)nf << m.g_shift)), block(kBlock);
#define K(CPLV)                                                                               \
  hipLaunchKernelGGL((k_frontier_jacobi<T, CPLV>), grid, block, 0, st, nf, frontier, C,       \
                     m.g_shift, m.lc_shift, indptr, indices, (const T*)w, (const T*)X,        \
                     (const T*)rhs, (const T*)diag, (T)delta, (T*)Y, changed)
  DLP_DISPATCH_CPL(m.cpl, K)
#undef K
  count_edges(nf, frontier, indptr, edges, st);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_frontier_jacobi(int nf, const int32_t* frontier, int n, int C, const int32_t* indptr,
                                   const int32_t* indices, const void* w, const void* X,
                                   const void* rhs, const void* diag, double delta, void* Y,
                                   uint8_t* changed, unsigned long long* edges, int dtype,
                                   double avg_deg, void* stream) {
  DLP_CHECK_ARG(nf >= 0 && n >= 0 && C >= 1 && C <= 256, "frontier_jacobi: bad nf=%d n=%d C=%d", nf, n, C);
  if (nf == 0) return DLP_OK;
  DLP_CHECK_ARG(frontier && indptr && indices && w && X && rhs && diag && Y && changed,
                "frontier_jacobi: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == DLP_F32 && frontier_jacobi_rows(nf, frontier, C, indptr, indices, (const float*)w, (const float*)X,
                                               (const float*)rhs, (const float*)diag, (float)delta, (float*)Y,
                                               changed, n, st)) {
    count_edges(nf, frontier, indptr, edges, st);
    DLP_CHECK_LAUNCH();
    return DLP_OK;
  }
  if (dtype == DLP_F32)
    return frontier_jacobi_t<float>(nf, frontier, C, indptr, indices, w, X, rhs, diag, delta, Y,
                                    changed, edges, avg_deg, st);
  if (dtype == DLP_F64)
    return frontier_jacobi_t<double>(nf, frontier, C, indptr, indices, w, X, rhs, diag, delta, Y,
                                     changed, edges, avg_deg, st);
  DLP_CHECK_ARG(false, "frontier_jacobi: bad dtype %d", dtype);
  return DLP_ERR_ARG;
}

template <typename T>
static int push_scatter_t(int nf, const int32_t* frontier, int C, const int32_t* indptr,
                          const int32_t* indices, const void* w, const void* D, void* R,
                          uint8_t* mark, unsigned long long* edges, double avg_deg,
                          hipStream_t st) {
  if (std::is_same<T, float>::value && tuning().pipe) {
    int lc = C <= 1 ? 0 : ceil_log2(C);
    if (lc > 6) lc = 6;
    const int cpl = (C + (1 << lc) - 1) >> lc;   // <= 4 for C <= 256
    long long grid = resident_grid(reinterpret_cast<const void*>(&k_push_scatter_rows<1>));
    const long long need = ((long long)nf + 3) / 4;
    if (grid > need) grid = need;
#define K(CPLV)                                                                                             \
  hipLaunchKernelGGL((k_push_scatter_rows<CPLV>), dim3((unsigned)grid), dim3(kBlock), 0, st, nf, frontier, C, lc, \
                     indptr, indices, (const float*)w, (const float*)D, (float*)R, mark)
    DLP_DISPATCH_CPL(cpl, K)
#undef K
    count_edges(nf, frontier, indptr, edges, st);
    DLP_CHECK_LAUNCH();
    return DLP_OK;
  }
  RowMap m = choose_map(C, avg_deg);
  dim3 grid(grid_for((long long)nf << m.g_shift)), block(kBlock);
#define K(CPLV)                                                                              \
  hipLaunchKernelGGL((k_push_scatter<T, CPLV>), grid, block, 0, st, nf, frontier, C,         \
                     m.g_shift, m.lc_shift, indptr, indices, (const T*)w, (const T*)D, (T*)R, \
                     mark)
  DLP_DISPATCH_CPL(m.cpl, K)
#undef K
  count_edges(nf, frontier, indptr, edges, st);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_push_scatter(int nf, const int32_t* frontier, int C, const int32_t* indptr,
                                const int32_t* indices, const void* w, const void* D, void* R,
                                uint8_t* mark, unsigned long long* edges, int dtype,
                                double avg_deg, void* stream) {
  DLP_CHECK_ARG(nf >= 0 && C >= 1 && C <= 256, "push_scatter: bad nf=%d C=%d", nf, C);
  if (nf == 0) return DLP_OK;
  DLP_CHECK_ARG(frontier && indptr && indices && w && D && R && mark, "push_scatter: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == DLP_F32)
    return push_scatter_t<float>(nf, frontier, C, indptr, indices, w, D, R, mark, edges, avg_deg, st);
  if (dtype == DLP_F64)
    return push_scatter_t<double>(nf, frontier, C, indptr, indices, w, D, R, mark, edges, avg_deg, st);
  DLP_CHECK_ARG(false, "push_scatter: bad dtype %d", dtype);
  return DLP_ERR_ARG;
}

extern "C" int dlp_mark_neighbors(int nr, const int32_t* rows, const int32_t* indptr,
                                  const int32_t* indices, uint8_t* mark,
                                  unsigned long long* edges, void* stream) {
  DLP_CHECK_ARG(nr >= 0, "mark_neighbors: bad nr=%d", nr);
  if (nr == 0) return DLP_OK;
  DLP_CHECK_ARG(rows && indptr && indices && mark, "mark_neighbors: null pointer");
  hipLaunchKernelGGL(k_mark_neighbors, dim3(grid_for((long long)nr * 8)), dim3(kBlock), 0,
                     (hipStream_t)stream, nr, rows, indptr, indices, mark);
  count_edges(nr, rows, indptr, edges, (hipStream_t)stream);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

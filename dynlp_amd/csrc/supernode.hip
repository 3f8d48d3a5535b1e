// Unbuffered scatters and the supernode components of new vertices (DynLP step 1).
//
// Scatters: a[idx[i]] op= v[i] for every i (repeated indices all apply), the device form of the
// reference's backend.scatter_add/max/min (backend.py:11-28).
//
// Adaptive (Felzenszwalb-Huttenlocher) components in parallel Boruvka rounds (cc.py:38-83):
// every component takes its lightest outgoing edge and merges across it when the edge's
// dissimilarity is at most min(Int(C1) + k/|C1|, Int(C2) + k/|C2|).  Here a round is three
// passes over device arrays: one packed 64-bit atomic min per edge picks each component's
// lightest outgoing edge (dissimilarity in the high word, order-preserving, the component
// across in the low word, so ties go to the smallest component in the same atomic), one
// thread per component applies the merge test and lists the accepted (component, across)
// pairs, and the connected components of those pairs (the union-find of cc.hip, labels =
// rank of the minimum) renumber the components, whose largest internal dissimilarity, size
// and ground are then merged with atomics.  The host reads two counters per round.
#include "scan.h"
#include "uf.h"

namespace dlp {

// ---------------------------------------------------------------------------------------------
// Scatters.

template <typename T>
__device__ __forceinline__ void atomic_max_any(T* p, T v);
template <typename T>
__device__ __forceinline__ void atomic_min_any(T* p, T v);

// Floating max/min by compare-and-swap on the bit pattern; a NaN on either side wins (as a
// NaN propagates through numpy's maximum.at).
template <typename T, typename U, bool MAX>
__device__ __forceinline__ void atomic_fminmax(T* p, T v) {
  U* q = reinterpret_cast<U*>(p);
  U old = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    T cur;
    __builtin_memcpy(&cur, &old, sizeof(T));
    if (cur != cur) return;   // NaN stays
    const bool take = (v != v) || (MAX ? v > cur : v < cur);
    if (!take) return;
    U nv;
    __builtin_memcpy(&nv, &v, sizeof(T));
    if (__hip_atomic_compare_exchange_strong(q, &old, nv, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return;
  }
}
template <>
__device__ __forceinline__ void atomic_max_any<float>(float* p, float v) { atomic_fminmax<float, uint32_t, true>(p, v); }
template <>
__device__ __forceinline__ void atomic_min_any<float>(float* p, float v) { atomic_fminmax<float, uint32_t, false>(p, v); }
template <>
__device__ __forceinline__ void atomic_max_any<double>(double* p, double v) {
  atomic_fminmax<double, unsigned long long, true>(p, v);
}
template <>
__device__ __forceinline__ void atomic_min_any<double>(double* p, double v) {
  atomic_fminmax<double, unsigned long long, false>(p, v);
}
template <>
__device__ __forceinline__ void atomic_max_any<int32_t>(int32_t* p, int32_t v) { atomicMax(p, v); }
template <>
__device__ __forceinline__ void atomic_min_any<int32_t>(int32_t* p, int32_t v) { atomicMin(p, v); }
template <typename T, bool MAX>
__device__ __forceinline__ void atomic_i64(T* p, T v) {
  unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
  unsigned long long old = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    const long long cur = (long long)old;
    if (!(MAX ? (long long)v > cur : (long long)v < cur)) return;
    if (__hip_atomic_compare_exchange_strong(q, &old, (unsigned long long)v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return;
  }
}
template <>
__device__ __forceinline__ void atomic_max_any<int64_t>(int64_t* p, int64_t v) { atomic_i64<int64_t, true>(p, v); }
template <>
__device__ __forceinline__ void atomic_min_any<int64_t>(int64_t* p, int64_t v) { atomic_i64<int64_t, false>(p, v); }

template <typename T>
__device__ __forceinline__ void atomic_add_any(T* p, T v) {
  atomicAdd(p, v);
}
template <>
__device__ __forceinline__ void atomic_add_any<int64_t>(int64_t* p, int64_t v) {
  atomicAdd(reinterpret_cast<unsigned long long*>(p), (unsigned long long)v);
}

// Element t of the m x cols update: row i = t / cols, column j = t % cols; v_stride = cols
// (one row of values per index) or 0 (one value for every element).
template <typename T, typename I, int OP>
__global__ void k_scatter(int64_t n_out, int64_t cols, T* __restrict__ a, int64_t m, const I* __restrict__ idx,
                          const T* __restrict__ v, int64_t v_stride, int32_t* __restrict__ bad) {
  const int64_t total = m * cols;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / cols, j = t - i * cols;
    const int64_t r = (int64_t)idx[i];
    if (r < 0 || r >= n_out) {
      if (bad && j == 0) atomicAdd(bad, 1);
      continue;
    }
    const T x = v_stride ? v[i * v_stride + j] : v[0];
    T* p = a + r * cols + j;
    if (OP == 0) atomic_add_any<T>(p, x);
    else if (OP == 1) atomic_max_any<T>(p, x);
    else atomic_min_any<T>(p, x);
  }
}

// ---------------------------------------------------------------------------------------------
// Adaptive components.

constexpr unsigned long long kNoEdge = ~0ull;

// Order-preserving map of an FP32 value to 32 unsigned bits (negative values reversed).
__device__ __forceinline__ uint32_t f32_key(float x) {
  const uint32_t b = __float_as_uint(x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float f32_of_key(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void k_fh_init(int m, int32_t* __restrict__ lab, float* __restrict__ Int, float* __restrict__ size,
                          float* __restrict__ g, const float* __restrict__ ground) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= m) return;
  lab[v] = v;
  Int[v] = 0.f;
  size[v] = 1.f;
  g[v] = ground ? ground[v] : 1.f;
}

__global__ void k_fh_clear(int nc, unsigned long long* __restrict__ best, float* __restrict__ Int2,
                           float* __restrict__ size2, float* __restrict__ g2, int32_t* __restrict__ ctl) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nc) {
    best[c] = kNoEdge;
    Int2[c] = 0.f;
    size2[c] = 0.f;
    g2[c] = 0.f;
  }
  if (c < 2) ctl[c] = 0;   // [0] components with an outgoing edge, [1] accepted merges
}

// Every edge between two components (only from components without ground in the grounding
// pass) offers itself to its first endpoint's component.
__global__ void k_fh_pick(int64_t ne, const int32_t* __restrict__ rows, const int32_t* __restrict__ cols,
                          const float* __restrict__ d, const int32_t* __restrict__ lab, const float* __restrict__ g,
                          int ground_pass, unsigned long long* __restrict__ best) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    const int cu = lab[rows[e]], cv = lab[cols[e]];
    if (cu == cv) continue;
    if (ground_pass && g[cu] != 0.f) continue;
    const unsigned long long key = ((unsigned long long)f32_key(d[e]) << 32) | (uint32_t)cv;
    if (key < best[cu]) atomicMin(best + cu, key);
  }
}

// The merge test of each component with an outgoing edge (FP32, as the reference keeps
// Int and size): d <= min(Int[c] + k/size[c], Int[t] + k/size[t]); every edge passes in the
// grounding pass.  Accepted pairs are listed; dacc[c] keeps the accepted edge's value.
__global__ void k_fh_accept(int nc, const unsigned long long* __restrict__ best, const float* __restrict__ Int,
                            const float* __restrict__ size, float k, int ground_pass, int32_t* __restrict__ ec,
                            int32_t* __restrict__ et, float* __restrict__ dacc, int32_t* __restrict__ ctl) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const unsigned long long b = best[c];
  dacc[c] = -__builtin_huge_valf();
  if (b == kNoEdge) return;
  atomicAdd(ctl, 1);
  const int t = (int)(uint32_t)(b & 0xffffffffu);
  const float dc = f32_of_key((uint32_t)(b >> 32));
  const bool ok = ground_pass || dc <= fminf(Int[c] + k / size[c], Int[t] + k / size[t]);
  if (!ok) return;
  const int i = atomicAdd(ctl + 1, 1);
  ec[i] = c;
  et[i] = t;
  dacc[c] = dc;
}

// Merged state under the new numbering root[] (Int starts at 0, as the reference's
// scatter_max into zeros).
__global__ void k_fh_merge(int nc, const int32_t* __restrict__ root, const float* __restrict__ Int,
                           const float* __restrict__ size, const float* __restrict__ g,
                           const float* __restrict__ dacc, float* __restrict__ Int2, float* __restrict__ size2,
                           float* __restrict__ g2) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int r = root[c];
  atomic_max_any<float>(Int2 + r, fmaxf(Int[c], dacc[c]));
  atomicAdd(size2 + r, size[c]);
  atomicAdd(g2 + r, g[c]);
}

__global__ void k_fh_relabel(int m, int32_t* __restrict__ lab, const int32_t* __restrict__ root) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < m) lab[v] = root[lab[v]];
}

// Dissimilarities of the adaptive components (cc.py:77-83):
//   d_e = (1 - w_e/wmax) / sqrt(sr_e * sc_e)   in the weights' precision, then FP32,
// with sr_e = s_r[ir[e]] (or s_r[e] when ir is null), likewise sc_e; and per-block partial sums
// of d (FP64, unrounded) for the mean k.
template <typename T>
__global__ __launch_bounds__(256) void k_fh_dissim(int64_t ne, const T* __restrict__ w, const float* __restrict__ s_r,
                                                   const float* __restrict__ s_c, const int32_t* __restrict__ ir,
                                                   const int32_t* __restrict__ ic, T wmax, float* __restrict__ d,
                                                   double* __restrict__ part) {
  __shared__ double sm[256];
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    const float sr = ir ? s_r[ir[e]] : s_r[e];
    const float sc = ic ? s_c[ic[e]] : s_c[e];
    const float ss = sqrtf(sr * sc);   // FP32, as the scales are kept
    const T x = ((T)1 - w[e] / wmax) / (T)ss;
    d[e] = (float)x;
    acc += (double)x;
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sm[threadIdx.x] += sm[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// Mean dissimilarity of each new vertex over its live entries (stream.py:67-81):
//   s_v = sum_{w>0} (1 - w/wmax) / max(count, 1) + 1e-6 over the entries of row v.
template <typename T>
__global__ void k_fh_scale(int64_t ne, const int32_t* __restrict__ lr, const T* __restrict__ w, T wmax,
                           float* __restrict__ s, float* __restrict__ cnt) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    if (w[e] > (T)0) {
      atomicAdd(s + lr[e], (float)((T)1 - w[e] / wmax));
      atomicAdd(cnt + lr[e], 1.f);
    }
  }
}
__global__ void k_fh_scale_clear(int m, float* __restrict__ s, float* __restrict__ cnt) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < m) s[v] = cnt[v] = 0.f;
}
__global__ void k_fh_scale_finish(int m, float* __restrict__ s, const float* __restrict__ cnt) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < m) s[v] = s[v] / fmaxf(cnt[v], 1.f) + 1e-6f;
}

struct FHWork {
  int32_t *lab, *root, *ec, *et, *ctl;
  float *Int, *size, *g, *Int2, *size2, *g2, *dacc;
  unsigned long long* best;
  void* cc_ws;
  size_t cc_bytes;
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static size_t fh_bytes(int m) {
  const size_t a = align256((size_t)m * 4);
  return 11 * a + align256((size_t)m * 8) + 256 + dlp_cc_workspace_bytes(m);
}

}  // namespace dlp

using namespace dlp;

extern "C" int dlp_scatter(int op, int dtype, int64_t n_out, int64_t cols, void* a, int64_t m, const void* idx,
                           int idx64, const void* v, int64_t v_stride, int32_t* bad, void* stream) {
  DLP_CHECK_ARG(op >= 0 && op <= 2, "scatter: op %d (0 add, 1 max, 2 min)", op);
  DLP_CHECK_ARG(n_out >= 0 && cols >= 1 && m >= 0 && (v_stride == 0 || v_stride == cols), "scatter: bad sizes");
  if (m == 0) return DLP_OK;
  DLP_CHECK_ARG(a && idx && v, "scatter: null pointer");
  hipStream_t st = (hipStream_t)stream;
  long long g = (long long)grid_for(m * cols);
  const dim3 grid((unsigned)(g > 65536 ? 65536 : g)), block(kBlock);
#define DLP_SCAT(T, I)                                                                                      \
  switch (op) {                                                                                             \
    case 0: hipLaunchKernelGGL((k_scatter<T, I, 0>), grid, block, 0, st, n_out, cols, (T*)a, m, (const I*)idx, \
                               (const T*)v, v_stride, bad); break;                                          \
    case 1: hipLaunchKernelGGL((k_scatter<T, I, 1>), grid, block, 0, st, n_out, cols, (T*)a, m, (const I*)idx, \
                               (const T*)v, v_stride, bad); break;                                          \
    default: hipLaunchKernelGGL((k_scatter<T, I, 2>), grid, block, 0, st, n_out, cols, (T*)a, m, (const I*)idx, \
                                (const T*)v, v_stride, bad); break;                                         \
  }
#define DLP_SCAT_I(T) if (idx64) { DLP_SCAT(T, int64_t) } else { DLP_SCAT(T, int32_t) }
  switch (dtype) {
    case DLP_F32: DLP_SCAT_I(float); break;
    case DLP_F64: DLP_SCAT_I(double); break;
    case DLP_I32: DLP_SCAT_I(int32_t); break;
    case DLP_I64: DLP_SCAT_I(int64_t); break;
    default: DLP_CHECK_ARG(false, "scatter: dtype %d", dtype);
  }
#undef DLP_SCAT_I
#undef DLP_SCAT
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" size_t dlp_fh_workspace_bytes(int m) { return fh_bytes(m < 1 ? 1 : m); }

extern "C" int dlp_fh_components(int m, int64_t ne, const int32_t* rows, const int32_t* cols, const float* d,
                                 double k, const float* ground, int32_t* labels, int32_t* n_comp, void* ws,
                                 size_t ws_bytes, void* stream) {
  DLP_CHECK_ARG(m >= 0 && ne >= 0 && n_comp, "fh_components: bad arguments");
  if (m == 0) {
    *n_comp = 0;
    return DLP_OK;
  }
  DLP_CHECK_ARG(labels && (ne == 0 || (rows && cols && d)), "fh_components: null pointer");
  DLP_CHECK_ARG(ws && ws_bytes >= fh_bytes(m), "fh_components: workspace too small (%zu < %zu)", ws_bytes,
                fh_bytes(m));
  hipStream_t st = (hipStream_t)stream;
  const size_t a = align256((size_t)m * 4);
  char* p = (char*)ws;
  FHWork W;
  W.lab = labels;
  W.root = (int32_t*)p; p += a;
  W.ec = (int32_t*)p; p += a;
  W.et = (int32_t*)p; p += a;
  W.Int = (float*)p; p += a;
  W.size = (float*)p; p += a;
  W.g = (float*)p; p += a;
  W.Int2 = (float*)p; p += a;
  W.size2 = (float*)p; p += a;
  W.g2 = (float*)p; p += a;
  W.dacc = (float*)p; p += a;
  p += a;   // spare
  W.best = (unsigned long long*)p; p += align256((size_t)m * 8);
  W.ctl = (int32_t*)p; p += 256;
  W.cc_ws = p;
  W.cc_bytes = dlp_cc_workspace_bytes(m);
  hipLaunchKernelGGL(k_fh_init, dim3(grid_for(m)), dim3(kBlock), 0, st, m, W.lab, W.Int, W.size, W.g, ground);
  DLP_CHECK_LAUNCH();
  int32_t nc = m;
  bool fh = true;
  const float kf = (float)k;
  long long ge = (long long)grid_for(ne);
  const dim3 egrid((unsigned)(ge > 65536 ? 65536 : (ge < 1 ? 1 : ge)));
  int32_t host[3];
  while (ne > 0) {
    hipLaunchKernelGGL(k_fh_clear, dim3(grid_for(nc > 2 ? nc : 2)), dim3(kBlock), 0, st, nc, W.best, W.Int2,
                       W.size2, W.g2, W.ctl);
    hipLaunchKernelGGL(k_fh_pick, egrid, dim3(kBlock), 0, st, ne, rows, cols, d, W.lab, W.g, fh ? 0 : 1, W.best);
    hipLaunchKernelGGL(k_fh_accept, dim3(grid_for(nc)), dim3(kBlock), 0, st, nc, W.best, W.Int, W.size, kf,
                       fh ? 0 : 1, W.ec, W.et, W.dacc, W.ctl);
    DLP_CHECK_LAUNCH();
    DLP_CHECK_HIP(hipMemcpyAsync(host, W.ctl, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    DLP_CHECK_HIP(hipStreamSynchronize(st));
    if (host[0] == 0) break;          // no edge leaves any (eligible) component
    if (host[1] == 0) {               // no edge passes the merge test
      if (fh && ground) {
        fh = false;                   // components without ground join their lightest neighbour
        continue;
      }
      break;
    }
    int rc = dlp_cc_edges(nc, host[1], W.ec, W.et, W.root, W.ctl + 2, W.cc_ws, W.cc_bytes, stream);
    if (rc) return rc;
    DLP_CHECK_HIP(hipMemcpyAsync(host + 2, W.ctl + 2, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    hipLaunchKernelGGL(k_fh_merge, dim3(grid_for(nc)), dim3(kBlock), 0, st, nc, W.root, W.Int, W.size, W.g, W.dacc,
                       W.Int2, W.size2, W.g2);
    hipLaunchKernelGGL(k_fh_relabel, dim3(grid_for(m)), dim3(kBlock), 0, st, m, W.lab, W.root);
    DLP_CHECK_LAUNCH();
    DLP_CHECK_HIP(hipStreamSynchronize(st));
    nc = host[2];
    float* t;
    t = W.Int; W.Int = W.Int2; W.Int2 = t;
    t = W.size; W.size = W.size2; W.size2 = t;
    t = W.g; W.g = W.g2; W.g2 = t;
  }
  *n_comp = nc;
  return DLP_OK;
}

extern "C" int dlp_fh_scale(int m, int64_t ne, const int32_t* lr, const void* w, int dtype, double wmax, float* s,
                            float* cnt, void* stream) {
  DLP_CHECK_ARG(m >= 0 && ne >= 0 && wmax > 0, "fh_scale: bad arguments");
  if (m == 0) return DLP_OK;
  DLP_CHECK_ARG(s && cnt && (ne == 0 || (lr && w)), "fh_scale: null pointer");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fh_scale_clear, dim3(grid_for(m)), dim3(kBlock), 0, st, m, s, cnt);
  if (ne > 0) {
    long long g = (long long)grid_for(ne);
    const dim3 grid((unsigned)(g > 65536 ? 65536 : g));
    if (dtype == DLP_F64)
      hipLaunchKernelGGL(k_fh_scale<double>, grid, dim3(kBlock), 0, st, ne, lr, (const double*)w, wmax, s, cnt);
    else
      hipLaunchKernelGGL(k_fh_scale<float>, grid, dim3(kBlock), 0, st, ne, lr, (const float*)w, (float)wmax, s, cnt);
  }
  hipLaunchKernelGGL(k_fh_scale_finish, dim3(grid_for(m)), dim3(kBlock), 0, st, m, s, cnt);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// Partial sums: at most kDissimBlocks blocks, summed on the host in block order (deterministic).
constexpr int kDissimBlocks = 1024;

extern "C" size_t dlp_fh_dissimilarity_workspace_bytes(void) { return sizeof(double) * kDissimBlocks; }

extern "C" int dlp_fh_dissimilarity(int64_t ne, const void* w, int dtype, const float* s_r, const float* s_c,
                                    const int32_t* ir, const int32_t* ic, double wmax, float* d, double* sum_out,
                                    void* ws, size_t ws_bytes, void* stream) {
  DLP_CHECK_ARG(ne >= 0 && wmax > 0 && sum_out, "fh_dissimilarity: bad arguments");
  *sum_out = 0.0;
  if (ne == 0) return DLP_OK;
  DLP_CHECK_ARG(w && s_r && s_c && d && ws, "fh_dissimilarity: null pointer");
  DLP_CHECK_ARG(ws_bytes >= sizeof(double) * kDissimBlocks, "fh_dissimilarity: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  long long g = (long long)grid_for(ne);
  const int nb = (int)(g > kDissimBlocks ? kDissimBlocks : g);
  double* part = (double*)ws;
  if (dtype == DLP_F64)
    hipLaunchKernelGGL(k_fh_dissim<double>, dim3(nb), dim3(kBlock), 0, st, ne, (const double*)w, s_r, s_c, ir, ic,
                       wmax, d, part);
  else
    hipLaunchKernelGGL(k_fh_dissim<float>, dim3(nb), dim3(kBlock), 0, st, ne, (const float*)w, s_r, s_c, ir, ic,
                       (float)wmax, d, part);
  DLP_CHECK_LAUNCH();
  double host[kDissimBlocks];
  DLP_CHECK_HIP(hipMemcpyAsync(host, part, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
  DLP_CHECK_HIP(hipStreamSynchronize(st));
  double sum = 0.0;
  for (int b = 0; b < nb; ++b) sum += host[b];
  *sum_out = sum;
  return DLP_OK;
}

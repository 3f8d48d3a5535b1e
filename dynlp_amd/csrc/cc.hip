// Connected components by lock-free union-find, canonical labels by a device scan.
//
// Hooking always links the larger root under the smaller one (compare-and-swap on the
// larger root's parent), so every tree's root is its minimum vertex id and parent ids
// strictly decrease along any path.  After all unions a read-only pass gives each vertex its
// root, the component minimum; flags (root or not) are scanned so that a component's label
// is the rank of its minimum among all component minima, which is the labelling the
// reference's hook/jump + unique produces (cc.py:12-23).
#include "common.h"
#include "scan.h"

namespace dlp {

__device__ __forceinline__ int ld_relaxed(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(int32_t* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving (every stored value is an ancestor, so races are benign).
__device__ __forceinline__ int uf_find(int32_t* P, int x) {
  while (true) {
    const int p = ld_relaxed(P + x);
    if (p == x) return x;
    const int gp = ld_relaxed(P + p);
    if (gp == p) return p;
    st_relaxed(P + x, gp);
    x = gp;
  }
}

// Each failed CAS means another thread hooked a root, so the number of roots strictly
// decreases over failures: the loop ends.
__device__ __forceinline__ void uf_union(int32_t* P, int a, int b) {
  while (true) {
    a = uf_find(P, a);
    b = uf_find(P, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(P + a, a, b) == a) return;
  }
}

__global__ void k_cc_init(int n, int32_t* __restrict__ P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) P[i] = i;
}

__global__ void k_cc_hook_edges(int64_t m, const int32_t* __restrict__ rows,
                                const int32_t* __restrict__ cols, int n, int32_t* P) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int a = rows[e], b = cols[e];
    if (a != b && (unsigned)a < (unsigned)n && (unsigned)b < (unsigned)n) uf_union(P, a, b);
  }
}

// One group of 8 lanes per row; entry (r, c) with c > r joins r and c (each undirected
// pair once when both directions are stored; a one-direction entry with c < r is handled
// by the lane that sees it, so unsymmetric inputs are also covered).
__global__ __launch_bounds__(256) void k_cc_hook_csr(int n, const int32_t* __restrict__ indptr,
                                                     const int32_t* __restrict__ rowlen,
                                                     const int32_t* __restrict__ indices,
                                                     const uint8_t* __restrict__ mask,
                                                     int32_t* P) {
  const int64_t r64 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 3;
  const int lane = threadIdx.x & 7;
  if (r64 >= n) return;
  const int r = (int)r64;
  if (mask && !mask[r]) return;
  const int beg = indptr[r];
  const int end = rowlen ? beg + rowlen[r] : indptr[r + 1];
  for (int j = beg + lane; j < end; j += 8) {
    const int c = indices[j];
    if (c == r || (unsigned)c >= (unsigned)n) continue;
    if (mask && !mask[c]) continue;
    uf_union(P, r, c);
  }
}

// Root of x without path compression.  The final pass must not store into other vertices'
// entries: a halving store by one thread could overwrite the root another thread has just
// written for its own vertex, leaving a non-root there.
__device__ __forceinline__ int uf_root(const int32_t* P, int x) {
  int p = ld_relaxed(P + x);
  while (p != x) {
    x = p;
    p = ld_relaxed(P + x);
  }
  return x;
}

// roots[i] = root of i (its component's minimum id), flag[i] = (i is a root).
__global__ void k_cc_compress(int n, const int32_t* P, int32_t* __restrict__ roots, int32_t* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int r = uf_root(P, i);
    roots[i] = r;
    flag[i] = (r == i) ? 1 : 0;
  }
}

// labels[i] = rank[roots[i]]; labels may alias roots (each thread reads then writes its own i).
__global__ void k_cc_label(int n, const int32_t* roots, const int32_t* __restrict__ rank, int32_t* labels) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) labels[i] = rank[roots[i]];
}

struct CCWork {
  int32_t* parent;
  int32_t* flag;   // flags, scanned in place into ranks
  void* scan_ws;
  size_t scan_bytes;
};

static size_t cc_bytes(int n) {
  const size_t a = ((size_t)n * 4 + 255) & ~(size_t)255;
  return 2 * a + scan_workspace_bytes(n);
}

static int cc_carve(int n, void* ws, size_t ws_bytes, CCWork& w) {
  DLP_CHECK_ARG(ws != nullptr && ws_bytes >= cc_bytes(n), "cc: workspace too small (%zu < %zu)",
                ws_bytes, cc_bytes(n));
  const size_t a = ((size_t)n * 4 + 255) & ~(size_t)255;
  char* p = (char*)ws;
  w.parent = (int32_t*)p;
  w.flag = (int32_t*)(p + a);
  w.scan_ws = p + 2 * a;
  w.scan_bytes = ws_bytes - 2 * a;
  return DLP_OK;
}

static int cc_finish(int n, CCWork& w, int32_t* labels, int32_t* n_comp, hipStream_t st) {
  hipLaunchKernelGGL(k_cc_compress, dim3(grid_for(n)), dim3(kBlock), 0, st, n, w.parent, labels, w.flag);
  DLP_CHECK_LAUNCH();
  int rc = exclusive_scan_i32(n, w.flag, w.flag, n_comp, w.scan_ws, w.scan_bytes, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_cc_label, dim3(grid_for(n)), dim3(kBlock), 0, st, n, labels, w.flag, labels);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

}  // namespace dlp

using namespace dlp;

extern "C" size_t dlp_cc_workspace_bytes(int n) { return cc_bytes(n < 1 ? 1 : n); }

extern "C" int dlp_cc_edges(int n, int64_t m, const int32_t* rows, const int32_t* cols,
                            int32_t* labels, int32_t* n_comp, void* ws, size_t ws_bytes,
                            void* stream) {
  DLP_CHECK_ARG(n >= 0 && m >= 0, "cc_edges: bad sizes n=%d m=%lld", n, (long long)m);
  hipStream_t st = (hipStream_t)stream;
  if (n == 0) {
    DLP_CHECK_HIP(hipMemsetAsync(n_comp, 0, sizeof(int32_t), st));
    return DLP_OK;
  }
  DLP_CHECK_ARG(labels && n_comp && (m == 0 || (rows && cols)), "cc_edges: null pointer");
  CCWork w;
  int rc = cc_carve(n, ws, ws_bytes, w);
  if (rc) return rc;
  hipLaunchKernelGGL(k_cc_init, dim3(grid_for(n)), dim3(kBlock), 0, st, n, w.parent);
  DLP_CHECK_LAUNCH();
  if (m > 0) {
    unsigned g = grid_for(m);
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_cc_hook_edges, dim3(g), dim3(kBlock), 0, st, m, rows, cols, n, w.parent);
    DLP_CHECK_LAUNCH();
  }
  return cc_finish(n, w, labels, n_comp, st);
}

extern "C" int dlp_cc_csr(int n, const int32_t* indptr, const int32_t* rowlen,
                          const int32_t* indices, const uint8_t* mask, int32_t* labels,
                          int32_t* n_comp, void* ws, size_t ws_bytes, void* stream) {
  DLP_CHECK_ARG(n >= 0, "cc_csr: bad n=%d", n);
  hipStream_t st = (hipStream_t)stream;
  if (n == 0) {
    DLP_CHECK_HIP(hipMemsetAsync(n_comp, 0, sizeof(int32_t), st));
    return DLP_OK;
  }
  DLP_CHECK_ARG(indptr && indices && labels && n_comp, "cc_csr: null pointer");
  CCWork w;
  int rc = cc_carve(n, ws, ws_bytes, w);
  if (rc) return rc;
  hipLaunchKernelGGL(k_cc_init, dim3(grid_for(n)), dim3(kBlock), 0, st, n, w.parent);
  DLP_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_cc_hook_csr, dim3(grid_for((long long)n * 8)), dim3(kBlock), 0, st, n,
                     indptr, rowlen, indices, mask, w.parent);
  DLP_CHECK_LAUNCH();
  return cc_finish(n, w, labels, n_comp, st);
}

// Connected components by lock-free union-find (uf.h), canonical labels by a device scan.
//
// A read-only pass gives each vertex its root, the component minimum; flags (root or not) are
// scanned so that a component's label is the rank of its minimum among all component minima,
// which is the labelling the reference's hook/jump + unique produces (cc.py:12-23).
#include <mutex>
#include "scan.h"
#include "uf.h"

namespace dlp {

__global__ void k_uf_init(int n, int32_t* __restrict__ P, const int32_t* __restrict__ need) {
  if (need && *need == 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) P[i] = i;
}

// Round r of the neighbour links: every filtered vertex joins its r-th filtered entry.
__global__ void k_uf_link_round(int n, const int32_t* __restrict__ indptr, const int32_t* __restrict__ rowlen,
                                const int32_t* __restrict__ indices, const uint8_t* __restrict__ filt, int keep,
                                int32_t* P, int r, const int32_t* __restrict__ need) {
  if (need && *need == 0) return;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n || !uf_in(filt, keep, v)) return;
  const int len = rowlen ? rowlen[v] : indptr[v + 1] - indptr[v];
  if (r >= len) return;
  const int u = indices[indptr[v] + r];
  if (u != v && (unsigned)u < (unsigned)n && uf_in(filt, keep, u)) uf_union(P, v, u);
}

// P[v] = root of v for every vertex (each thread stores only its own entry).
__global__ void k_uf_flatten(int n, int32_t* P, const int32_t* __restrict__ need) {
  if (need && *need == 0) return;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) uf_store(P + v, uf_root(P, v));
}

// giant[0] = the most frequent root among kRootSamples hashed vertices of a flattened forest
// (ties -> smaller root).  One workgroup of 256.
__global__ __launch_bounds__(256) void k_uf_sample(int n, const int32_t* __restrict__ P, int32_t* __restrict__ giant,
                                                   const int32_t* __restrict__ need) {
  if (need && *need == 0) return;
  __shared__ int32_t s[kRootSamples];
  __shared__ unsigned long long best;
  if (threadIdx.x == 0) best = ~0ull;
  for (int i = threadIdx.x; i < kRootSamples; i += blockDim.x) {
    const unsigned v = (unsigned)(((unsigned long long)i * 2654435761ull + 12345ull) % (unsigned long long)n);
    s[i] = P[v];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kRootSamples; i += blockDim.x) {
    int cnt = 0;
    for (int j = 0; j < kRootSamples; ++j) cnt += s[j] == s[i];
    // smallest key = largest count, then smallest root
    const unsigned long long key = ((unsigned long long)(kRootSamples - cnt) << 32) | (unsigned)s[i];
    atomicMin(&best, key);
  }
  __syncthreads();
  if (threadIdx.x == 0) giant[0] = (int32_t)(best & 0xffffffffu);
}

// Remaining links: vertices outside the sampled giant component join every filtered entry past
// the first kNeighbourRounds (8 lanes per row).
__global__ __launch_bounds__(256) void k_uf_link_rest(int n, const int32_t* __restrict__ indptr,
                                                      const int32_t* __restrict__ rowlen,
                                                      const int32_t* __restrict__ indices,
                                                      const uint8_t* __restrict__ filt, int keep, int32_t* P,
                                                      const int32_t* __restrict__ giant,
                                                      const int32_t* __restrict__ need) {
  if (need && *need == 0) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * 8) return;
  const int v = (int)(t >> 3);
  const int lane = (int)(t & 7);
  if (!uf_in(filt, keep, v) || uf_load(P + v) == *giant) return;
  const int beg = indptr[v];
  const int end = rowlen ? beg + rowlen[v] : indptr[v + 1];
  for (int j = beg + kNeighbourRounds + lane; j < end; j += 8) {
    const int u = indices[j];
    if (u != v && (unsigned)u < (unsigned)n && uf_in(filt, keep, u)) uf_union(P, v, u);
  }
}

int uf_components_csr(int n, const int32_t* indptr, const int32_t* rowlen, const int32_t* indices,
                      const uint8_t* filt, int keep, int32_t* P, int32_t* giant, const int32_t* need,
                      hipStream_t st) {
  if (n <= 0) return DLP_OK;
  const dim3 g1(grid_for(n)), b(kBlock);
  hipLaunchKernelGGL(k_uf_init, g1, b, 0, st, n, P, need);
  // unions compress no paths, so every round starts from flattened trees
  for (int r = 0; r < kNeighbourRounds; ++r) {
    hipLaunchKernelGGL(k_uf_link_round, g1, b, 0, st, n, indptr, rowlen, indices, filt, keep, P, r, need);
    hipLaunchKernelGGL(k_uf_flatten, g1, b, 0, st, n, P, need);
  }
  hipLaunchKernelGGL(k_uf_sample, dim3(1), b, 0, st, n, P, giant, need);
  hipLaunchKernelGGL(k_uf_link_rest, dim3(grid_for((long long)n * 8)), b, 0, st, n, indptr, rowlen, indices, filt,
                     keep, P, giant, need);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// Edge-list links over [e0, e1): with giant non-null an edge whose two endpoints already sit in
// the sampled giant component is skipped.
__global__ void k_uf_link_edges(int64_t e0, int64_t e1, const int32_t* __restrict__ rows,
                                const int32_t* __restrict__ cols, int n, int32_t* P,
                                const int32_t* __restrict__ giant) {
  const int gc = giant ? *giant : -1;
  for (int64_t e = e0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < e1; e += (int64_t)gridDim.x * blockDim.x) {
    const int a = rows[e], c = cols[e];
    if (a == c || (unsigned)a >= (unsigned)n || (unsigned)c >= (unsigned)n) continue;
    if (giant && uf_load(P + a) == gc && uf_load(P + c) == gc) continue;
    uf_union(P, a, c);
  }
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
  int32_t* giant;
  void* scan_ws;
  size_t scan_bytes;
};

constexpr size_t cc_bytes(int n) {
  const size_t a = ((size_t)n * 4 + 255) & ~(size_t)255;
  return 2 * a + 256 + scan_workspace_bytes(n);
}

static int cc_carve(int n, void* ws, size_t ws_bytes, CCWork& w) {
  DLP_CHECK_ARG(ws != nullptr && ws_bytes >= cc_bytes(n), "cc: workspace too small (%zu < %zu)",
                ws_bytes, cc_bytes(n));
  const size_t a = ((size_t)n * 4 + 255) & ~(size_t)255;
  char* p = (char*)ws;
  w.parent = (int32_t*)p;
  w.flag = (int32_t*)(p + a);
  w.giant = (int32_t*)(p + 2 * a);
  w.scan_ws = p + 2 * a + 256;
  w.scan_bytes = ws_bytes - 2 * a - 256;
  return DLP_OK;
}

// *out = 0 (a kernel rather than a memset, so that the no-workspace entry point can be
// captured into a graph).
__global__ void k_cc_zero(int32_t* out) {
  if (threadIdx.x == 0) *out = 0;
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
    DLP_CHECK_ARG(n_comp, "cc: null pointer");
    hipLaunchKernelGGL(k_cc_zero, dim3(1), dim3(64), 0, st, n_comp);
    DLP_CHECK_LAUNCH();
    return DLP_OK;
  }
  DLP_CHECK_ARG(labels && n_comp && (m == 0 || (rows && cols)), "cc_edges: null pointer");
  CCWork w;
  int rc = cc_carve(n, ws, ws_bytes, w);
  if (rc) return rc;
  hipLaunchKernelGGL(k_uf_init, dim3(grid_for(n)), dim3(kBlock), 0, st, n, w.parent, nullptr);
  DLP_CHECK_LAUNCH();
  if (m > 0) {
    // a prefix of the edges first (about kNeighbourRounds per vertex), then the rest with the
    // edges inside the sampled giant component skipped
    const int64_t m1 = m < (int64_t)kNeighbourRounds * n ? m : (int64_t)kNeighbourRounds * n;
    auto grid_of = [](int64_t cnt) {
      unsigned g = grid_for(cnt);
      return g > 65536 ? 65536u : (g < 1 ? 1u : g);
    };
    // the prefix in chunks of n edges, each followed by a flattening pass
    for (int64_t e0 = 0; e0 < m1; e0 += n) {
      const int64_t e1 = e0 + n < m1 ? e0 + n : m1;
      hipLaunchKernelGGL(k_uf_link_edges, dim3(grid_of(e1 - e0)), dim3(kBlock), 0, st, e0, e1, rows, cols, n,
                         w.parent, nullptr);
      hipLaunchKernelGGL(k_uf_flatten, dim3(grid_for(n)), dim3(kBlock), 0, st, n, w.parent, nullptr);
    }
    if (m > m1) {
      hipLaunchKernelGGL(k_uf_sample, dim3(1), dim3(kBlock), 0, st, n, w.parent, w.giant, nullptr);
      hipLaunchKernelGGL(k_uf_link_edges, dim3(grid_of(m - m1)), dim3(kBlock), 0, st, m1, m, rows, cols, n,
                         w.parent, w.giant);
    }
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
    DLP_CHECK_ARG(n_comp, "cc: null pointer");
    hipLaunchKernelGGL(k_cc_zero, dim3(1), dim3(64), 0, st, n_comp);
    DLP_CHECK_LAUNCH();
    return DLP_OK;
  }
  DLP_CHECK_ARG(indptr && indices && labels && n_comp, "cc_csr: null pointer");
  CCWork w;
  int rc = cc_carve(n, ws, ws_bytes, w);
  if (rc) return rc;
  rc = uf_components_csr(n, indptr, rowlen, indices, mask, 1, w.parent, w.giant, nullptr, st);
  if (rc) return rc;
  return cc_finish(n, w, labels, n_comp, st);
}

// ---- Connected components with no workspace and no allocation ------------------------------
// Scratch of dlp_cc_csr_static: a file-scope device array, one copy per GPU (each at its own
// address), sized for kCCStaticMaxN vertices and held until the process ends.  Its address is
// looked up once per device, with that device current, on the entry point's first call there;
// the array's name is only a host-side stand-in and is never passed to a kernel.
namespace {
constexpr int kCCStaticMaxN = 1 << 21;
constexpr size_t kCCStaticBytes = dlp::cc_bytes(kCCStaticMaxN);
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kCCStaticBytes];

static int static_scratch(void** out) {
  int dev = 0;
  DLP_CHECK_HIP(hipGetDevice(&dev));
  DLP_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "cc_csr_static: device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    DLP_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DLP_CHECK_ARG(p != nullptr, "cc_csr_static: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return DLP_OK;
}

extern "C" int dlp_cc_csr_static(int n, const int32_t* indptr, const int32_t* rowlen, const int32_t* indices,
                                 const uint8_t* mask, int32_t* labels, int32_t* n_comp, void* stream) {
  DLP_CHECK_ARG(n >= 0 && n <= kCCStaticMaxN,
                "cc_csr_static: n=%d outside [0, %d], the vertices its fixed scratch holds (use dlp_cc_csr)", n,
                kCCStaticMaxN);
  void* ws = nullptr;
  const int rc = static_scratch(&ws);
  if (rc) return rc;
  return dlp_cc_csr(n, indptr, rowlen, indices, mask, labels, n_comp, ws, kCCStaticBytes, stream);
}

extern "C" int dlp_cc_static_max_n(void) { return kCCStaticMaxN; }

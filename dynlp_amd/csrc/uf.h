// Lock-free union-find shared by connected components and grounding.
//
// Hooking always links the larger root under the smaller one (compare-and-swap on the larger
// root's parent), so every tree's root is its minimum vertex id and parent ids strictly
// decrease along any path.  Components are found Afforest-style: each vertex is first linked to
// its first kNeighbourRounds neighbours (trees flattened after each round), the largest component
// is found from
// a sample of roots, and only vertices outside it link their remaining neighbours (an edge with
// one endpoint outside is seen from that endpoint, since both directions are stored).  The
// resulting partition, and its minimum-id roots, are those of a plain union of every edge.
#pragma once
#include "common.h"

namespace dlp {

constexpr int kNeighbourRounds = 2;
constexpr int kRootSamples = 512;

__device__ __forceinline__ int uf_load(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void uf_store(int32_t* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Joins the trees of a and b.  Only parent pointers are followed (one hop for the larger
// candidate, two for the smaller per step), never rewritten: compressing paths while a million
// threads link makes the hot entries near the roots a store target of every thread (5-14x slower
// on 1M-vertex graphs), and the trees are flattened between passes instead.  The larger root
// is hooked under the smaller by compare-and-swap; the loop ends on success, or when the larger
// candidate's parent is already the smaller, or when both candidates meet.  A failed CAS means
// another thread hooked that root, so the number of roots decreases over failures.
__device__ __forceinline__ void uf_union(int32_t* P, int a, int b) {
  int p1 = uf_load(P + a), p2 = uf_load(P + b);
  while (p1 != p2) {
    const int hi = p1 > p2 ? p1 : p2;
    const int lo = p1 > p2 ? p2 : p1;
    const int ph = uf_load(P + hi);
    if (ph == lo) return;
    if (ph == hi && atomicCAS(P + hi, hi, lo) == hi) return;
    p1 = uf_load(P + ph);
    p2 = uf_load(P + lo);
  }
}

// Root of x without stores.  Passes that run after the unions write only their own vertex's
// entry: a halving store by one thread could overwrite the root another thread has just
// written, leaving a non-root there.
__device__ __forceinline__ int uf_root(const int32_t* P, int x) {
  int p = uf_load(P + x);
  while (p != x) {
    x = p;
    p = uf_load(P + x);
  }
  return x;
}

// A vertex takes part when filt is null or (filt[v] != 0) == (keep != 0).
__device__ __forceinline__ bool uf_in(const uint8_t* filt, int keep, int v) {
  return filt == nullptr || ((filt[v] != 0) == (keep != 0));
}

// Components of the graph of a (slack) CSR restricted to the filtered vertices, in P (roots =
// component minima; vertices outside the filter stay singletons).  giant: one device int of
// scratch.  need: optional device flag; when it reads 0 every pass returns at once.
int uf_components_csr(int n, const int32_t* indptr, const int32_t* rowlen, const int32_t* indices,
                      const uint8_t* filt, int keep, int32_t* P, int32_t* giant, const int32_t* need,
                      hipStream_t st);

}  // namespace dlp

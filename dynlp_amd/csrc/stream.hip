// Resident incremental state of one graph: edge-batch apply on a slack CSR, grounding,
// changed-row refresh of the FP64 residual, and the round-synchronous push.
//
// Graph store: every vertex owns a row segment [indptr[r], indptr[r+1]) of which the first
// rowlen[r] entries are live (column, FP32 weight).  Deletions tombstone (col -1), a per-row
// compaction closes the gaps, insertions upsert in place or append into the slack; an append
// that finds the row full is recorded as overflow and retried after the host regrows the
// store.  Labeled vertices stay in the store; `solved` (U: unlabeled and grounded) masks the
// operator, so W_UU never has to be materialised.
#include <cstring>
#include "rowpipe.h"
#include "scan.h"
#include "blas.h"
#include "uf.h"
#include <algorithm>
#include <vector>

namespace dlp {

// ---------------------------------------------------------------------------------------
// Edge-batch apply.

__device__ __forceinline__ void stamp_append(int v, int32_t* __restrict__ stamp, int stampv,
                                             int32_t* __restrict__ list, int32_t* __restrict__ count) {
  wave_append(atomicExch(stamp + v, stampv) != stampv, v, list, count);
}

__device__ __forceinline__ bool id_ok(int v, int n) { return (unsigned)v < (unsigned)n; }
// An insertion weight must be finite and positive (NaN fails the first test).
__device__ __forceinline__ bool weight_ok(float x) { return x > 0.f && x <= 3.402823466e38f; }

// |res| / diag with NaN mapped to +inf: a row whose residual or diagonal is not a number
// fails every threshold test and reaches the column maxima, so it is never certified.
__device__ __forceinline__ double ratio_or_inf(double res, double dg) { return nan_to_inf(fabs(res) / dg); }

// Row a loses its entry b (tombstone by compare-and-swap; not found counts as missing).
__device__ __forceinline__ void delete_entry(int a, int b, const int32_t* __restrict__ indptr,
                                             const int32_t* __restrict__ rowlen, int32_t* __restrict__ col,
                                             float* __restrict__ w, const uint8_t* __restrict__ labeled,
                                             int32_t* __restrict__ labcnt, int32_t* __restrict__ stamp, int stampv,
                                             int32_t* __restrict__ rows, int32_t* __restrict__ nrows,
                                             int32_t* __restrict__ missing) {
  const int beg = indptr[a], end = beg + rowlen[a];
  bool found = false;
  for (int j = beg; j < end; ++j) {
    if (col[j] == b) {
      if (atomicCAS(col + j, b, -1) == b) {
        w[j] = 0.f;
        found = true;
      }
      break;
    }
  }
  if (!found) {
    atomicAdd(missing, 1);
    return;
  }
  if (labeled[b]) atomicSub(labcnt + a, 1);
  stamp_append(a, stamp, stampv, rows, nrows);
}

// Directed deletion e < 2*nd: row a loses entry b.  Rows are stamped into the changed list;
// the labeled-neighbour counter of a drops when b is labeled.  A pair with an id outside
// [0, n) is skipped and counted once (invalid); the tombstone is a compare-and-swap, so a pair
// deleted twice in one call is removed once (the second counts as missing).
__global__ void k_apply_delete(int n, int nd, const int32_t* __restrict__ du, const int32_t* __restrict__ dv,
                               const int32_t* __restrict__ indptr, const int32_t* __restrict__ rowlen,
                               int32_t* __restrict__ col, float* __restrict__ w,
                               const uint8_t* __restrict__ labeled, int32_t* __restrict__ labcnt,
                               int32_t* __restrict__ stamp, int stampv, int32_t* __restrict__ rows,
                               int32_t* __restrict__ nrows, int32_t* __restrict__ missing,
                               int32_t* __restrict__ invalid) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * (int64_t)nd) return;
  const int a = e < nd ? du[e] : dv[e - nd];
  const int b = e < nd ? dv[e] : du[e - nd];
  if (!id_ok(a, n) || !id_ok(b, n)) {
    if (e < nd) atomicAdd(invalid, 1);
    return;
  }
  if (a == b) return;
  delete_entry(a, b, indptr, rowlen, col, w, labeled, labcnt, stamp, stampv, rows, nrows, missing);
}

// Close the tombstone gaps of the listed rows (stable order of the live entries).
__global__ void k_compact_rows(const int32_t* __restrict__ nrows, const int32_t* __restrict__ rows,
                               const int32_t* __restrict__ indptr, int32_t* __restrict__ rowlen,
                               int32_t* __restrict__ col, float* __restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *nrows) return;
  const int r = rows[i];
  const int beg = indptr[r], end = beg + rowlen[r];
  int o = beg;
  for (int j = beg; j < end; ++j) {
    const int c = col[j];
    if (c >= 0) {
      if (o != j) {
        col[o] = c;
        w[o] = w[j];
      }
      ++o;
    }
  }
  for (int j = o; j < end; ++j) {
    col[j] = -1;
    w[j] = 0.f;
  }
  rowlen[r] = o - beg;
}

// Insertions of one call are reduced to the last occurrence of each unordered pair before
// any row changes, so both directions of a pair come from the same occurrence (the stored
// matrix stays symmetric) and no two wavefronts insert the same entry: an open-addressing
// table keyed by min*n+max holds the largest occurrence index of every pair (its capacity is
// a power of two at least twice the insertions, so every probe sequence meets an empty slot).
constexpr unsigned long long kHashEmpty = ~0ull;
constexpr int kPairTableCap = 1 << 20;   // 2 x the 2^19 insertions one call may hold

__device__ __forceinline__ unsigned pair_slot(unsigned long long key, int bits) {
  return (unsigned)((key * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

__device__ __forceinline__ bool insertion_ok(int a, int b, float wt, int n) {
  return id_ok(a, n) && id_ok(b, n) && weight_ok(wt);
}

__global__ void k_pair_table_clear(int cap, unsigned long long* __restrict__ keys, int32_t* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cap) {
    keys[i] = kHashEmpty;
    vals[i] = -1;
  }
}

__global__ void k_pair_table_fill(int n, int ni, const int32_t* __restrict__ iu, const int32_t* __restrict__ iv,
                                  const float* __restrict__ iw, int bits, unsigned long long* __restrict__ keys,
                                  int32_t* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ni) return;
  const int a = iu[i], b = iv[i];
  if (!insertion_ok(a, b, iw[i], n) || a == b) return;
  const unsigned long long key = (unsigned long long)min(a, b) * (unsigned)n + (unsigned)max(a, b);
  const unsigned mask = (1u << bits) - 1u;
  for (unsigned h = pair_slot(key, bits);; h = (h + 1) & mask) {
    const unsigned long long old = atomicCAS(keys + h, kHashEmpty, key);
    if (old == kHashEmpty || old == key) {
      atomicMax(vals + h, i);
      return;
    }
  }
}

// Largest occurrence index of the pair (present: filled by k_pair_table_fill).
__device__ __forceinline__ int pair_last(unsigned long long key, int bits, const unsigned long long* __restrict__ keys,
                                         const int32_t* __restrict__ vals) {
  const unsigned mask = (1u << bits) - 1u;
  for (unsigned h = pair_slot(key, bits);; h = (h + 1) & mask) {
    const unsigned long long k = keys[h];
    if (k == key) return vals[h];
    if (k == kHashEmpty) return -1;
  }
}

// Upsert of the directed entry (a -> b, weight wt) by one whole wavefront: the lanes read the
// row's live entries in parallel; an existing entry takes the new weight, else lane 0 claims a
// slot (a full row records `tag` as overflow).  Callers never upsert one entry twice at once.
__device__ __forceinline__ void upsert_entry(int a, int b, float wt, int lane, const int32_t* __restrict__ indptr,
                                             int32_t* __restrict__ rowlen, int32_t* __restrict__ col,
                                             float* __restrict__ w, const uint8_t* __restrict__ labeled,
                                             int32_t* __restrict__ labcnt, int32_t* __restrict__ stamp,
                                             int stampv, int32_t* __restrict__ rows, int32_t* __restrict__ nrows,
                                             int32_t* __restrict__ overflow, int32_t* __restrict__ noverflow,
                                             int32_t tag) {
  const int beg = indptr[a], cap = indptr[a + 1] - beg;
  // a concurrent insertion that overflows the row raises rowlen past cap for a moment
  const int len = min(cap, __hip_atomic_load(rowlen + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  int hit = -1;
  for (int j0 = 0; j0 < len; j0 += 64) {
    const int j = j0 + lane;
    const unsigned long long m = __ballot(j < len && col[beg + j] == b);
    if (m) {
      hit = j0 + __ffsll((long long)m) - 1;
      break;
    }
  }
  if (lane != 0) return;
  if (hit >= 0) {   // existing pair: new weight
    w[beg + hit] = wt;
    stamp_append(a, stamp, stampv, rows, nrows);
    return;
  }
  const int slot = atomicAdd(rowlen + a, 1);
  if (slot < cap) {
    w[beg + slot] = wt;
    col[beg + slot] = b;
    if (labeled[b]) atomicAdd(labcnt + a, 1);
    stamp_append(a, stamp, stampv, rows, nrows);
  } else {
    atomicSub(rowlen + a, 1);
    overflow[atomicAdd(noverflow, 1)] = tag;
  }
}

// Directed insertion e < 2*ni: upsert (a, b, wt), one wavefront each (2*ni of them), when this
// occurrence is the pair's last in the call (earlier ones are counted in *ndup).  A pair with
// an id outside [0, n) or a weight that is not finite and positive is skipped and counted once
// (invalid).
__global__ __launch_bounds__(256) void k_apply_insert(int n, int ni, const int32_t* __restrict__ iu,
                                                      const int32_t* __restrict__ iv, const float* __restrict__ iw,
                                                      int bits, const unsigned long long* __restrict__ keys,
                                                      const int32_t* __restrict__ vals,
                                                      const int32_t* __restrict__ indptr,
                                                      int32_t* __restrict__ rowlen, int32_t* __restrict__ col,
                                                      float* __restrict__ w, const uint8_t* __restrict__ labeled,
                                                      int32_t* __restrict__ labcnt, int32_t* __restrict__ stamp,
                                                      int stampv, int32_t* __restrict__ rows,
                                                      int32_t* __restrict__ nrows, int32_t* __restrict__ overflow,
                                                      int32_t* __restrict__ noverflow, int32_t* __restrict__ invalid,
                                                      int32_t* __restrict__ ndup) {
  const int64_t e = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (e >= 2 * (int64_t)ni) return;   // wave-uniform
  const int i = (int)(e < ni ? e : e - ni);
  const int a = e < ni ? iu[i] : iv[i];
  const int b = e < ni ? iv[i] : iu[i];
  const float wt = iw[i];
  if (!insertion_ok(a, b, wt, n)) {
    if (e < ni && lane == 0) atomicAdd(invalid, 1);
    return;
  }
  if (a == b) return;
  const unsigned long long key = (unsigned long long)min(a, b) * (unsigned)n + (unsigned)max(a, b);
  if (pair_last(key, bits, keys, vals) != i) {
    if (e < ni && lane == 0) atomicAdd(ndup, 1);
    return;
  }
  upsert_entry(a, b, wt, lane, indptr, rowlen, col, w, labeled, labcnt, stamp, stampv, rows, nrows, overflow,
               noverflow, (int32_t)e);
}

// ---------------------------------------------------------------------------------------
// Ordered operations: one sequence of pair operations (u[s], v[s], w[s]), s = 0..nops-1, with
// w[s] == 0 a deletion and w[s] > 0 an insertion (upsert) of weight w[s].  Applying the
// sequence in order leaves every unordered pair as its LAST operation made it, so each pair
// takes that one operation only: the pair table holds the largest s per pair, deletions whose
// s is the largest tombstone the pair, insertions whose s is the largest upsert it, the rest
// change nothing.  A group of requests is thereby applied as one call, with their order kept.

__device__ __forceinline__ bool op_ok(int a, int b, float wt, int n) {
  return id_ok(a, n) && id_ok(b, n) && (wt == 0.f || weight_ok(wt));
}

__global__ void k_ops_table_fill(int n, int nops, const int32_t* __restrict__ u, const int32_t* __restrict__ v,
                                 const float* __restrict__ wt, int bits, unsigned long long* __restrict__ keys,
                                 int32_t* __restrict__ vals, int32_t* __restrict__ invalid) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nops) return;
  const int a = u[s], b = v[s];
  if (!op_ok(a, b, wt[s], n)) {
    atomicAdd(invalid, 1);
    return;
  }
  if (a == b) return;
  const unsigned long long key = (unsigned long long)min(a, b) * (unsigned)n + (unsigned)max(a, b);
  const unsigned mask = (1u << bits) - 1u;
  for (unsigned h = pair_slot(key, bits);; h = (h + 1) & mask) {
    const unsigned long long old = atomicCAS(keys + h, kHashEmpty, key);
    if (old == kHashEmpty || old == key) {
      atomicMax(vals + h, s);
      return;
    }
  }
}

// Directed entry e < 2*nops of operation s = e mod nops (e < nops: row u[s] gets v[s]).
__device__ __forceinline__ bool op_entry(int64_t e, int nops, int n, const int32_t* __restrict__ u,
                                         const int32_t* __restrict__ v, const float* __restrict__ wt, int bits,
                                         const unsigned long long* __restrict__ keys,
                                         const int32_t* __restrict__ vals, int& a, int& b, float& x) {
  const int s = (int)(e < nops ? e : e - nops);
  a = e < nops ? u[s] : v[s];
  b = e < nops ? v[s] : u[s];
  x = wt[s];
  if (!op_ok(a, b, x, n) || a == b) return false;
  const unsigned long long key = (unsigned long long)min(a, b) * (unsigned)n + (unsigned)max(a, b);
  return pair_last(key, bits, keys, vals) == s;
}

__global__ void k_ops_delete(int n, int nops, const int32_t* __restrict__ u, const int32_t* __restrict__ v,
                             const float* __restrict__ wt, int bits, const unsigned long long* __restrict__ keys,
                             const int32_t* __restrict__ vals, const int32_t* __restrict__ indptr,
                             const int32_t* __restrict__ rowlen, int32_t* __restrict__ col, float* __restrict__ w,
                             const uint8_t* __restrict__ labeled, int32_t* __restrict__ labcnt,
                             int32_t* __restrict__ stamp, int stampv, int32_t* __restrict__ rows,
                             int32_t* __restrict__ nrows, int32_t* __restrict__ missing) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * (int64_t)nops) return;
  int a, b;
  float x;
  if (!op_entry(e, nops, n, u, v, wt, bits, keys, vals, a, b, x) || x != 0.f) return;
  delete_entry(a, b, indptr, rowlen, col, w, labeled, labcnt, stamp, stampv, rows, nrows, missing);
}

__global__ __launch_bounds__(256) void k_ops_insert(int n, int nops, const int32_t* __restrict__ u,
                                                    const int32_t* __restrict__ v, const float* __restrict__ wt,
                                                    int bits, const unsigned long long* __restrict__ keys,
                                                    const int32_t* __restrict__ vals,
                                                    const int32_t* __restrict__ indptr, int32_t* __restrict__ rowlen,
                                                    int32_t* __restrict__ col, float* __restrict__ w,
                                                    const uint8_t* __restrict__ labeled,
                                                    int32_t* __restrict__ labcnt, int32_t* __restrict__ stamp,
                                                    int stampv, int32_t* __restrict__ rows,
                                                    int32_t* __restrict__ nrows, int32_t* __restrict__ overflow,
                                                    int32_t* __restrict__ noverflow) {
  const int64_t e = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (e >= 2 * (int64_t)nops) return;   // wave-uniform
  int a, b;
  float x;
  if (!op_entry(e, nops, n, u, v, wt, bits, keys, vals, a, b, x) || x == 0.f) return;   // wave-uniform
  upsert_entry(a, b, x, lane, indptr, rowlen, col, w, labeled, labcnt, stamp, stampv, rows, nrows, overflow,
               noverflow, (int32_t)e);
}

// Regrow: copy every row into a store with new offsets (rows keep their live prefix).
__global__ void k_regrow_copy(int n, const int32_t* __restrict__ old_ptr, const int32_t* __restrict__ rowlen,
                              const int32_t* __restrict__ old_col, const float* __restrict__ old_w,
                              const int32_t* __restrict__ new_ptr, int32_t* __restrict__ new_col,
                              float* __restrict__ new_w) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(t >> 3);
  const int lane = (int)(t & 7);
  if (r >= n) return;
  const int ob = old_ptr[r], nb = new_ptr[r], len = rowlen[r], ncap = new_ptr[r + 1] - nb;
  for (int j = lane; j < ncap; j += 8) {
    new_col[nb + j] = j < len ? old_col[ob + j] : -1;
    new_w[nb + j] = j < len ? old_w[ob + j] : 0.f;
  }
}

__global__ void k_regrow_caps(int n, const int32_t* __restrict__ rowlen, int extra, float growth,
                              int32_t* __restrict__ cap) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int len = rowlen[r];
  cap[r] = len + max(extra, (int)(len * growth));
}

// ---------------------------------------------------------------------------------------
// Grounding: U = unlabeled vertices whose component in the unlabeled subgraph holds a vertex
// with a labeled neighbour (stream.py:149-158).

__global__ void k_ground_init(int n, uint8_t* __restrict__ grounded, const int32_t* __restrict__ need) {
  if (need && *need == 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) grounded[i] = 0;
}

// Roots by a read-only walk (uf_root: a store in this pass could overwrite another vertex's
// final root); roots[i] is written for unlabeled i only.
__global__ void k_ground_roots(int n, const int32_t* P, const uint8_t* __restrict__ labeled,
                               const int32_t* __restrict__ labcnt, uint8_t* __restrict__ grounded,
                               int32_t* __restrict__ roots, const int32_t* __restrict__ need) {
  if (need && *need == 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || labeled[i]) return;
  const int x = uf_root(P, i);
  roots[i] = x;
  if (labcnt[i] > 0) grounded[x] = 1;
}

// New solved flag; vertices that toggle are listed (for the neighbour pass) and stamped changed.
__global__ void k_ground_toggle(int n, const int32_t* __restrict__ roots, const uint8_t* __restrict__ labeled,
                                const uint8_t* __restrict__ grounded, uint8_t* __restrict__ solved,
                                int32_t* __restrict__ toggled, int32_t* __restrict__ ntoggled,
                                int32_t* __restrict__ stamp, int stampv, int32_t* __restrict__ rows,
                                int32_t* __restrict__ nrows, const int32_t* __restrict__ need,
                                int32_t* __restrict__ ninert) {
  if (need && *need == 0) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint8_t now = (!labeled[i] && grounded[roots[i]]) ? 1 : 0;
  if (!labeled[i] && !now) atomicAdd(ninert, 1);
  if (now != solved[i]) {
    solved[i] = now;
    toggled[atomicAdd(ntoggled, 1)] = i;
    stamp_append(i, stamp, stampv, rows, nrows);
  }
}

// Exact local grounding test (no inert vertex before the batch): a component that loses its
// labeled contact must contain an endpoint of a deleted edge, so it suffices that every
// unlabeled deletion endpoint still reaches, through unlabeled vertices of the updated graph,
// a vertex with a labeled neighbour.  One wavefront per endpoint runs a bounded BFS with its
// queue and visited set in LDS; an endpoint that exhausts the budget (or any insertion while
// inert vertices exist) requests the full connected-components pass (need[0] = 1).
constexpr int kBallCap = 256;      // vertices a check may visit
constexpr int kBallHash = 1024;    // open-addressing slots (power of two)

__global__ __launch_bounds__(256) void k_ground_ball(int n, int nd, const int32_t* __restrict__ du, const int32_t* __restrict__ dv,
                                                     int ni, const int32_t* __restrict__ indptr,
                                                     const int32_t* __restrict__ rowlen, const int32_t* __restrict__ col,
                                                     const uint8_t* __restrict__ labeled,
                                                     const int32_t* __restrict__ labcnt, int32_t* __restrict__ flags) {
  __shared__ int32_t queue[4][kBallCap];
  __shared__ int32_t hset[4][kBallHash];
  __shared__ int32_t qlen[4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    atomicAdd(flags + 2, 1);
    if (ni > 0 && flags[1] > 0) flags[0] = 1;
  }
  const int64_t e = (int64_t)blockIdx.x * 4 + wave;
  if (e >= 2 * (int64_t)nd) return;
  const int x = e < nd ? du[e] : dv[e - nd];
  if (!id_ok(x, n) || !id_ok(e < nd ? dv[e] : du[e - nd], n)) return;   // refused by the apply
  if (labeled[x]) return;
  if (labcnt[x] > 0) return;              // grounded directly
  for (int t = lane; t < kBallHash; t += 64) hset[wave][t] = -1;
  if (lane == 0) {
    queue[wave][0] = x;
    qlen[wave] = 1;
    hset[wave][(x * 2654435761u) & (kBallHash - 1)] = x;
  }
  // wavefront-local LDS hand-off (each wave owns its own queue and set; waves may have left)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  bool found = false;
  int head = 0;
  while (!found && head < qlen[wave]) {
    const int v = queue[wave][head++];
    const int beg = indptr[v], end = beg + rowlen[v];
    for (int j0 = beg; j0 < end && !found; j0 += 64) {
      const int j = j0 + lane;
      bool hit = false;
      if (j < end) {
        const int c = col[j];
        if (c >= 0 && !labeled[c]) {
          if (labcnt[c] > 0) {
            hit = true;
          } else {
            unsigned hsh = (unsigned)(c * 2654435761u) & (kBallHash - 1);
            for (int probe = 0; probe < kBallHash; ++probe) {
              const int old = atomicCAS(&hset[wave][hsh], -1, c);
              if (old == -1) {
                const int q = atomicAdd(&qlen[wave], 1);
                if (q < kBallCap) queue[wave][q] = c;
                break;
              }
              if (old == c) break;
              hsh = (hsh + 1) & (kBallHash - 1);
            }
          }
        }
      }
      found = __ballot(hit) != 0ull;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (qlen[wave] > kBallCap) break;
  }
  if (!found && lane == 0) atomicExch(flags, 1);
}

__global__ void k_mark_neighbours_changed(const int32_t* __restrict__ list, const int32_t* __restrict__ count,
                                          const int32_t* __restrict__ indptr,
                                          const int32_t* __restrict__ rowlen,
                                          const int32_t* __restrict__ col, int32_t* __restrict__ stamp,
                                          int stampv, int32_t* __restrict__ rows,
                                          int32_t* __restrict__ nrows) {
  const int nl = *count;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int i = (int)(t >> 3);
  const int lane = (int)(t & 7);
  if (i >= nl) return;
  const int r = list[i];
  const int beg = indptr[r], end = beg + rowlen[r];
  for (int j = beg + lane; j < end; j += 8) {
    const int c = col[j];
    if (c >= 0) stamp_append(c, stamp, stampv, rows, nrows);
  }
}

// ---------------------------------------------------------------------------------------
// Row refresh: the equation and FP64 residual of a row computed from scratch,
//   s = sum_{labeled nbrs} w + eta,  diag = s + sum_{solved nbrs} w,
//   rhs_c = sum_{labeled nbrs l with column(y_l) = c} w + eta*prior_c,  rhs_h = diag,
//   R = rhs - diag*Z + sum_{solved nbrs} w Z[nbr]     (stream.py:164-167, solvers.py:404-407)
// Inert rows (not solved) are parked: diag 1, Z = (0, .., 0, 1), R = 0 (solvers.py:365-368).
// Rows violating |R_c| > eps_c*diag are appended to viol; rho holds column maxima of |R|/diag.

struct RefreshArgs {
  int nrows;
  const int32_t* rows;     // listed rows (device) or NULL with rows_count
  const int32_t* rows_count;  // device count when rows is a device list
  int C1;                  // columns incl. h
  int K2;                  // binary problem: only label 1 feeds column 0
  double eta, prior;
  const int32_t* indptr;
  const int32_t* rowlen;
  const int32_t* col;
  const float* w;
  const uint8_t* labeled;
  const int32_t* y;
  const uint8_t* solved;
  int64_t ld;
  double* diag64;
  float* diag32;
  const double* eps;       // FP64 thresholds (C1)
  int32_t* viol;
  int32_t* nviol;
  double* rho;             // C1 column maxima, or NULL
  double* hmax;            // the verification pass: max over solved rows of h (column C1-1), or NULL
  double* mass;            // sum over violating rows of max_c |R_c|/(eps_c*diag), or NULL
};

// Rows outside the solved set are parked so that an unmasked neighbour sum over every entry of
// a row is exactly what its residual needs: a labeled row holds the indicator of its class
// column (so W Z adds the labeled-neighbour weights of the right-hand side) and h = 0; any
// other row holds zeros.  R = 0 and diag = 1 on all of them.  Returns the column that holds 1,
// or -1.
__device__ __forceinline__ int parked_class(const RefreshArgs& A, int r) {
  if (!A.labeled[r]) return -1;
  const int yl = A.y[r];
  return A.K2 ? (yl == 1 ? 0 : -1) : yl;
}

// One FP64 atomic per wavefront: the wave's sum of per-row mass terms.
__device__ __forceinline__ void flush_mass(double* dst, double v) {
  if (!dst) return;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v > 0.0) atomicAdd(dst, v);
}

template <typename TZ, typename TR, int CPL>
__device__ __forceinline__ void refresh_row(const RefreshArgs& A, const Lanes& L, int64_t i, int nr,
                                            TZ* __restrict__ Z, TR* __restrict__ R, double (&rmax)[CPL],
                                            double& mass) {
  const bool valid = i < nr;
  const int r = valid ? (A.rows ? A.rows[i] : (int)i) : 0;
  const bool live = valid && A.solved[r];
  const int C = A.C1 - 1;  // label columns
  int beg = 0, end = 0;
  if (live) {
    beg = A.indptr[r];
    end = beg + A.rowlen[r];
  }
  double acc[CPL], rhs[CPL];
  double s = 0.0, rs = 0.0;
#pragma unroll
  for (int k = 0; k < CPL; ++k) acc[k] = rhs[k] = 0.0;
  for (int base = beg; base < end; base += L.G) {
    const int j = base + L.lane;
    int myv = 0, kind = 0, mycol = -1;  // kind 1: labeled, 2: solved
    float myw = 0.f;
    if (j < end) {
      myv = A.col[j];
      myw = A.w[j];
      if (myv >= 0) {
        if (A.labeled[myv]) {
          kind = 1;
          const int yl = A.y[myv];
          mycol = A.K2 ? (yl == 1 ? 0 : -1) : yl;
        } else if (A.solved[myv]) {
          kind = 2;
        }
      }
    }
    // scalar sums (each entry counted by one lane of the group)
    if (kind == 1) s += myw;
    if (kind == 2) rs += myw;
    const int cnt = min(L.G, end - base);
    for (int t0 = 0; t0 < cnt; t0 += 4 * L.LE) {
      int vv[4], kk[4], cc[4];
      double ww[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = t0 + u * L.LE + L.e;
        const int src = t & (L.G - 1);
        vv[u] = __shfl(myv, src, L.G);
        kk[u] = __shfl(kind, src, L.G);
        cc[u] = __shfl(mycol, src, L.G);
        ww[u] = (double)__shfl(myw, src, L.G);
        if (t >= cnt) kk[u] = 0;
      }
      TZ xv[4][CPL];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const TZ* zr = Z + (int64_t)vv[u] * A.ld;
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
          const int c = L.l + k * L.LC;
          xv[u][k] = (kk[u] == 2 && c < A.C1) ? zr[c] : (TZ)0;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
          const int c = L.l + k * L.LC;
          acc[k] += ww[u] * (double)xv[u][k];
          if (kk[u] == 1 && cc[u] == c) rhs[k] += ww[u];
        }
    }
  }
  // reduce: acc/rhs over edge lanes; s/rs over the whole group
  for (int off = L.LC; off < L.G; off <<= 1) {
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      acc[k] += __shfl_xor(acc[k], off, 64);
      rhs[k] += __shfl_xor(rhs[k], off, 64);
    }
  }
  for (int off = 1; off < L.G; off <<= 1) {
    s += __shfl_xor(s, off, 64);
    rs += __shfl_xor(rs, off, 64);
  }
  double m = 0.0;  // max_c ratio/eps_c of this lane
  bool viol = false;
  if (valid && L.e == 0) {
    const int64_t zo = (int64_t)r * A.ld;
    if (!live) {
      if (L.l == 0) {
        A.diag64[r] = 1.0;
        A.diag32[r] = 1.f;
      }
      const int pc = parked_class(A, r);
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        const int c = L.l + k * L.LC;
        if (c < A.C1) {
          Z[zo + c] = (TZ)(c == pc ? 1 : 0);
          R[zo + c] = (TR)0;
        }
      }
    } else {
      const double sg = s + A.eta;
      const double dg = sg + rs;
      if (L.l == 0) {
        A.diag64[r] = dg;
        A.diag32[r] = (float)dg;
      }
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        const int c = L.l + k * L.LC;
        if (c < A.C1) {
          const double rh = c == C ? dg : rhs[k] + A.eta * A.prior;
          const double res = rh - dg * (double)Z[zo + c] + acc[k];
          R[zo + c] = (TR)res;
          const double ratio = ratio_or_inf(res, dg);
          viol |= ratio > A.eps[c];
          rmax[k] = fmax(rmax[k], ratio);
          m = fmax(m, ratio / A.eps[c]);
        }
      }
    }
  }
  const unsigned long long bal = __ballot(viol);
  const int wl = threadIdx.x & 63;
  const unsigned long long gm = (L.G == 64 ? ~0ull : ((1ull << L.G) - 1)) << (wl & ~(L.G - 1));
  wave_append(valid && L.lane == 0 && (bal & gm) && A.viol, r, A.viol, A.nviol);
  if (A.mass) {
    for (int off = 1; off < L.G; off <<= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if (valid && L.lane == 0 && m > 1.0) mass += m;
  }
}

// FP32 state with 16-byte rows (ld % 4 == 0), one row per wavefront: column lane l covers
// columns 4l..4l+3 with double accumulators, LE edge lanes (LanesP: packed ceil(C1/4) column
// lanes, or a power of two).  Labeled neighbours (rare) add their weight to the rhs of the lane
// owning their class column.
__device__ __forceinline__ void refresh_row_v4(const RefreshArgs& A, const LanesP& L, int64_t i, int nr,
                                               float* __restrict__ Z, float* __restrict__ R, double (&rmax)[4],
                                               double& mass) {
  const bool valid = i < nr;
  const int r = valid ? (A.rows ? A.rows[i] : (int)i) : 0;
  const bool live = valid && A.solved[r];
  const int C = A.C1 - 1;
  int beg = 0, end = 0;
  if (live) {
    beg = A.indptr[r];
    end = beg + A.rowlen[r];
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0}, rhs[4] = {0.0, 0.0, 0.0, 0.0};
  double s = 0.0, rs = 0.0;
  for (int base = beg; base < end; base += 64) {
    const int j = base + L.lane;
    int myv = 0, kind = 0, mycol = -1;
    float myw = 0.f;
    if (j < end) {
      myv = A.col[j];
      myw = A.w[j];
      if (myv >= 0) {
        if (A.labeled[myv]) {
          kind = 1;
          const int yl = A.y[myv];
          mycol = A.K2 ? (yl == 1 ? 0 : -1) : yl;
        } else if (A.solved[myv]) {
          kind = 2;
        }
      }
    }
    if (kind == 1) s += myw;
    if (kind == 2) rs += myw;
    const int cnt = min(64, end - base);
    for (int t0 = 0; t0 < cnt; t0 += 4 * L.LE) {
      int vv[4], kk[4], cc[4];
      float ww[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = t0 + u * L.LE + L.e;
        const int src = t & 63;
        vv[u] = __shfl(myv, src, 64);
        kk[u] = __shfl(kind, src, 64);
        cc[u] = __shfl(mycol, src, 64);
        ww[u] = __shfl(myw, src, 64);
        if (t >= cnt || !L.on) kk[u] = 0;
      }
      float4 xv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float4* zr = reinterpret_cast<const float4*>(Z + (int64_t)vv[u] * A.ld) + L.l;
        xv[u] = kk[u] == 2 ? *zr : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double wd = (double)ww[u];
        acc[0] += wd * (double)xv[u].x;
        acc[1] += wd * (double)xv[u].y;
        acc[2] += wd * (double)xv[u].z;
        acc[3] += wd * (double)xv[u].w;
        if (kk[u] == 1 && (cc[u] >> 2) == L.l) rhs[cc[u] & 3] += wd;
      }
    }
  }
  {
    // edge-lane partials of column lane l are at lanes l + k*LC
    double tacc[4], trhs[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      tacc[q] = acc[q];
      trhs[q] = rhs[q];
    }
    for (int k = 1; k < L.LE; ++k) {
      const int src = L.lane + k * L.LC;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        tacc[q] += __shfl(acc[q], src, 64);
        trhs[q] += __shfl(rhs[q], src, 64);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      acc[q] = tacc[q];
      rhs[q] = trhs[q];
    }
  }
  for (int off = 1; off < 64; off <<= 1) {
    s += __shfl_xor(s, off, 64);
    rs += __shfl_xor(rs, off, 64);
  }
  double m = 0.0;  // max_c ratio/eps_c of this lane
  bool viol = false;
  if (valid && L.e == 0 && 4 * L.l < A.C1) {
    const int64_t zo = (int64_t)r * A.ld;
    if (!live) {
      if (L.l == 0) {
        A.diag64[r] = 1.0;
        A.diag32[r] = 1.f;
      }
      const int pc = parked_class(A, r);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = 4 * L.l + q;
        if (c < A.C1) {
          Z[zo + c] = (c == pc) ? 1.f : 0.f;
          R[zo + c] = 0.f;
        }
      }
    } else {
      const double dg = s + A.eta + rs;
      if (L.l == 0) {
        A.diag64[r] = dg;
        A.diag32[r] = (float)dg;
      }
      const float4 zr = *(reinterpret_cast<const float4*>(Z + zo) + L.l);
      const float zv[4] = {zr.x, zr.y, zr.z, zr.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = 4 * L.l + q;
        if (c < A.C1) {
          const double rh = c == C ? dg : rhs[q] + A.eta * A.prior;
          const double res = rh - dg * (double)zv[q] + acc[q];
          R[zo + c] = (float)res;
          const double ratio = ratio_or_inf(res, dg);
          viol |= ratio > A.eps[c];
          rmax[q] = fmax(rmax[q], ratio);
          m = fmax(m, ratio / A.eps[c]);
        }
      }
    }
  }
  const unsigned long long bal = __ballot(viol);
  wave_append(valid && L.lane == 0 && bal && A.viol, r, A.viol, A.nviol);
  if (A.mass) {
    for (int off = 1; off < 64; off <<= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if (valid && L.lane == 0 && m > 1.0) mass += m;
  }
}

__global__ __launch_bounds__(256) void k_refresh_v4(RefreshArgs A, int pow2, int xcd, float* __restrict__ Z,
                                                    float* __restrict__ R) {
  const LanesP L(A.C1, pow2 != 0);
  const int nr = A.rows_count ? *A.rows_count : A.nrows;
  // with xcd (grid a multiple of 8) the blocks one XCD runs walk one contiguous eighth of the
  // (listed) rows
  int64_t lo = 0, hi = nr, w0, waves;
  const int wave = threadIdx.x >> 6;
  if (xcd) {
    const int64_t chunk = ((int64_t)nr + 7) >> 3;
    lo = (blockIdx.x & 7) * chunk;
    hi = min((int64_t)nr, lo + chunk);
    w0 = (int64_t)(blockIdx.x >> 3) * 4 + wave;
    waves = (int64_t)(gridDim.x >> 3) * 4;
  } else {
    w0 = (int64_t)blockIdx.x * 4 + wave;
    waves = (int64_t)gridDim.x * 4;
  }
  double rmax[4] = {0.0, 0.0, 0.0, 0.0};
  double mass = 0.0;
  for (int64_t i = lo + w0; i < hi; i += waves) refresh_row_v4(A, L, i, nr, Z, R, rmax, mass);
  flush_mass(A.mass, mass);
  if (!A.rho) return;
  __shared__ double sm[256];
  for (int c = threadIdx.x; c < A.C1; c += blockDim.x) sm[c] = 0.0;
  __syncthreads();
  if (L.e == 0 && 4 * L.l < A.C1) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = 4 * L.l + q;
      if (c < A.C1) atomic_max_nonneg(&sm[c], rmax[q]);
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < A.C1; c += blockDim.x) global_max_nonneg(A.rho + c, sm[c]);
}

// Residual verification of every solved row after a global solve (refresh mode 3): the row
// equations (diag, parking) are unchanged since the last refresh of the changed rows, so
// R = rhs - diag*Z + W Z needs no per-entry look-up of the neighbour's kind: parked rows hold
// exactly the values an unmasked neighbour sum needs (class indicator on labeled rows, zero
// elsewhere), and the labeled-neighbour part of rhs arrives through W Z.  FP64 accumulation on
// the pipelined gathers of rowpipe.h; writes R (FP32), the violator list and column maxima.
struct VerifyEpilogue {
  const LanesP& L;
  const RefreshArgs& A;
  const float* __restrict__ Z;
  float* __restrict__ R;
  double rmax[4] = {0.0, 0.0, 0.0, 0.0};
  float hm = 0.f;          // max h over this lane's solved rows (the lane holding column C1-1)
  float4 zr;
  __device__ __forceinline__ bool owner() const { return L.e == 0 && 4 * L.l < A.C1; }
  __device__ __forceinline__ void pre(int, int r) {
    zr = owner() ? *(reinterpret_cast<const float4*>(Z + (int64_t)r * A.ld) + L.l) : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  // rows outside the solved set keep their parked state (R = 0)
  __device__ __forceinline__ void post(int, int r, bool live, const double (&acc)[4]) {
    if (!live) return;  // wave-uniform
    bool viol = false;
    if (owner()) {
      const double dg = A.diag64[r];
      const int C = A.C1 - 1;
      const float zv[4] = {zr.x, zr.y, zr.z, zr.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = 4 * L.l + q;
        if (c < A.C1) {
          const double rh = c == C ? dg : A.eta * A.prior;
          const double res = rh - dg * (double)zv[q] + acc[q];
          R[(int64_t)r * A.ld + c] = (float)res;
          const double ratio = ratio_or_inf(res, dg);
          viol |= ratio > A.eps[c];
          rmax[q] = fmax(rmax[q], ratio);
          if (c == C) hm = fmaxf(hm, zv[q]);
        }
      }
    }
    const unsigned long long bal = __ballot(viol);
    if (L.lane == 0 && bal && A.viol) A.viol[atomicAdd(A.nviol, 1)] = r;
  }
};

template <int S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_verify(RefreshArgs A, int xcd, int xbytes, const float* __restrict__ Z,
                                                float* __restrict__ R) {
  const RowPipe G(A.nrows, A.C1, 4, 16, xcd, A.solved, A.indptr, A.rowlen, A.col, A.w, Z, (unsigned)(A.ld * 4),
                  xbytes);
  VerifyEpilogue E{G.L, A, Z, R};
  run_rows<S, Word::F32x4, double>(G, E);
  if (!A.rho) return;
  __shared__ double sm[256];
  __shared__ double shm;
  for (int c = threadIdx.x; c < A.C1; c += blockDim.x) sm[c] = 0.0;
  if (threadIdx.x == 0) shm = 0.0;
  __syncthreads();
  if (A.hmax) {
    float hm = E.hm;
    for (int off = 32; off > 0; off >>= 1) hm = fmaxf(hm, __shfl_xor(hm, off, 64));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&shm, (double)fmaxf(hm, 0.f));
  }
  if (E.owner()) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = 4 * G.L.l + q;
      if (c < A.C1) atomic_max_nonneg(&sm[c], E.rmax[q]);
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < A.C1; c += blockDim.x) global_max_nonneg(A.rho + c, sm[c]);
  if (A.hmax && threadIdx.x == 0) global_max_nonneg(A.hmax, shm);
}

// Grid-stride over the listed rows in wave-uniform steps (the row count may live on the device).
template <typename TZ, typename TR, int CPL>
__global__ __launch_bounds__(256) void k_refresh(RefreshArgs A, int g_shift, int lc_shift, TZ* __restrict__ Z,
                                                 TR* __restrict__ R) {
  const Lanes L(g_shift, lc_shift);
  const int nr = A.rows_count ? *A.rows_count : A.nrows;
  const int64_t groups = ((int64_t)gridDim.x * blockDim.x) >> g_shift;
  const int64_t g0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> g_shift;
  const int64_t wave_first = ((int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63)) >> g_shift;
  double rmax[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) rmax[k] = 0.0;
  double mass = 0.0;
  for (int64_t base = wave_first; base < nr; base += groups)
    refresh_row<TZ, TR, CPL>(A, L, base + (g0 - wave_first), nr, Z, R, rmax, mass);
  flush_mass(A.mass, mass);
  if (!A.rho) return;
  // column maxima: across the wave's row groups, then the workgroup (LDS), then one atomic
  // per column per workgroup
  __shared__ double sm[256];
  for (int c = threadIdx.x; c < A.C1; c += blockDim.x) sm[c] = 0.0;
  for (int off = L.G; off < 64; off <<= 1) {
#pragma unroll
    for (int k = 0; k < CPL; ++k) rmax[k] = fmax(rmax[k], __shfl_xor(rmax[k], off, 64));
  }
  __syncthreads();
  if ((threadIdx.x & 63) < L.G && L.e == 0) {
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = L.l + k * L.LC;
      if (c < A.C1) atomic_max_nonneg(&sm[c], rmax[k]);
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < A.C1; c += blockDim.x) global_max_nonneg(A.rho + c, sm[c]);
}

// ---------------------------------------------------------------------------------------
// Asynchronous scatter push (the service's repair step; chaotic relaxation as FusedPush,
// kernels.py:269-351, with the compaction fused into the next round).
//
// Round t reads candidate list t%3 (vertices whose residual changed in round t-1, each listed
// once).  A candidate that violates |R_c| > eps_c*diag takes its residual atomically
// (concurrent pushes landing later are kept), adds R/diag to Z and scatters w*R/diag into its
// solved neighbours' residuals with hardware FP32 atomics; every neighbour whose residual it
// changed is appended once to list (t+1)%3.  A round with an empty list sets `done`.  The
// round number lives on the device (rnd[node parity]) so one captured graph of rounds replays
// any number of times; counts are a ring of three (read t%3, append (t+1)%3, reset (t+2)%3).
// All list appends are aggregated per wavefront (one atomic per wave).

struct PushState {
  int n, C1;
  int64_t ld;
  const int32_t* indptr;
  const int32_t* rowlen;
  const int32_t* col;
  const float* w;
  const uint8_t* solved;
  const float* diag;     // FP32 diag of the push
  const float* eps;      // FP32 thresholds, C1
  float* Z;
  float* R;
  int32_t* ctag;         // candidate stamp: round of the list a vertex was last appended to
  int32_t* list[3];
  int32_t* count;        // [3]
  int32_t* rnd;          // [2] round number by node parity
  int32_t* done;         // set when a round finds no candidates
  int32_t* btag;         // batch stamp: vertex pushed during the current batch
  int32_t* blist;
  int32_t* bcount;
  const int32_t* bstamp; // device: current batch stamp
  unsigned long long* edges;
  unsigned long long* stats;  // [0] candidates, [1] pushes, [2] scatter entries, [3] unused
};

template <int CPL>
__device__ __forceinline__ void push_round_body(const PushState& S, int parity, int g_shift, int lc_shift) {
  const Lanes L(g_shift, lc_shift);
  const int t = S.rnd[parity];
  const int nc = S.count[t % 3];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.count[(t + 2) % 3] = 0;
    S.rnd[parity ^ 1] = t + 1;
    if (nc == 0) *S.done = 1;
  }
  if (nc == 0) return;
  const int tq = t + 1;
  const int32_t* cand = S.list[t % 3];
  int32_t* next = S.list[tq % 3];
  int32_t* ncount = S.count + (tq % 3);
  const int bs = *S.bstamp;
  const int64_t groups = ((int64_t)gridDim.x * blockDim.x) >> g_shift;
  const int64_t g0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> g_shift;
  const int64_t wave_first = ((int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63)) >> g_shift;
  const int wl = threadIdx.x & 63;
  const unsigned long long gm = (L.G == 64 ? ~0ull : ((1ull << L.G) - 1)) << (wl & ~(L.G - 1));
  unsigned long long my_cand = 0, my_push = 0, my_sc = 0;
  for (int64_t base = wave_first; base < nc; base += groups) {
    const int64_t gi = base + (g0 - wave_first);
    const bool valid = gi < nc;
    const int v = valid ? cand[gi] : 0;
    const int64_t o = (int64_t)v * S.ld;
    // violation test on the current residual (edge lane 0 of the group reads the row)
    bool viol = false;
    if (valid && L.e == 0) {
      const float dv = S.diag[v];
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        const int c = L.l + k * L.LC;
        if (c < S.C1) viol |= fabsf(S.R[o + c]) > S.eps[c] * dv;
      }
    }
    const bool push = valid && (__ballot(viol) & gm) != 0ull;
    if (valid && L.lane == 0) ++my_cand;
    // take the residual and update Z
    float d[CPL];
    int beg = 0, end = 0;
#pragma unroll
    for (int k = 0; k < CPL; ++k) d[k] = 0.f;
    if (push) {
      if (L.e == 0) {
        const float inv = 1.f / S.diag[v];
#pragma unroll
        for (int k = 0; k < CPL; ++k) {
          const int c = L.l + k * L.LC;
          if (c < S.C1) {
            d[k] = atomic_take(S.R + o + c) * inv;
            S.Z[o + c] += d[k];
          }
        }
      }
      beg = S.indptr[v];
      end = beg + S.rowlen[v];
      if (L.lane == 0) {
        ++my_push;
        my_sc += (unsigned long long)(end - beg);
      }
    }
#pragma unroll
    for (int k = 0; k < CPL; ++k) d[k] = __shfl(d[k], L.l, L.G);
    // scatter w*d into the solved neighbours (no-return atomics)
    for (int j = beg + L.e; j < end; j += L.LE) {
      const int x = S.col[j];
      if (x < 0 || !S.solved[x]) continue;
      const float wv = S.w[j];
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        const int c = L.l + k * L.LC;
        if (c < S.C1) atomic_add(S.R + (int64_t)x * S.ld + c, wv * d[k]);
      }
    }
    // neighbours join the next list once (all lanes of the group over the entries)
    for (int j0 = beg; j0 < end; j0 += L.G) {
      const int j = j0 + L.lane;
      bool app = false;
      int x = 0;
      if (j < end) {
        x = S.col[j];
        app = x >= 0 && S.solved[x] && atomicExch(S.ctag + x, tq) != tq;
      }
      wave_append(app, x, next, ncount);
    }
    // every candidate's residual changed in this batch: it joins the FP64 verification list
    const bool first = valid && L.lane == 0 && atomicExch(S.btag + v, bs) != bs;
    wave_append(first, v, S.blist, S.bcount);
  }
  const unsigned long long sc = wave_sum_u64(my_cand);
  const unsigned long long sp = wave_sum_u64(my_push);
  const unsigned long long ss = wave_sum_u64(my_sc);
  if (wl == 0) {
    if (sc) atomicAdd(S.stats + 0, sc);
    if (sp) atomicAdd(S.stats + 1, sp);
    if (ss) {
      atomicAdd(S.stats + 2, ss);
      atomicAdd(S.edges, ss);
    }
  }
}

template <int CPL>
__global__ __launch_bounds__(256) void k_push_round(PushState S, int parity, int g_shift, int lc_shift) {
  push_round_body<CPL>(S, parity, g_shift, lc_shift);
}

// The same round for several graphs in one launch: graph blockIdx.y takes gridDim.x blocks
// (its lane layout from gshift/lcshift); graphs with active[y] == 0 are left alone.
template <int CPL>
__global__ __launch_bounds__(256) void k_push_round_multi(const PushState* __restrict__ states,
                                                          const int32_t* __restrict__ gshift,
                                                          const int32_t* __restrict__ lcshift,
                                                          const int32_t* __restrict__ active, int parity) {
  const int y = blockIdx.y;
  if (!active[y]) return;
  const PushState S = states[y];
  push_round_body<CPL>(S, parity, gshift[y], lcshift[y]);
}

// done flag and edge count of every graph of a multi push: out[2y] = done, out[2y+1] = edges
// (a graph outside the call has a zeroed descriptor and reads as done).
__global__ void k_push_probe_multi(const PushState* __restrict__ states, int ns, int64_t* __restrict__ out) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= ns) return;
  const PushState S = states[y];
  out[2 * y] = S.done ? (int64_t)*S.done : 1;
  out[2 * y + 1] = S.edges ? (int64_t)*S.edges : 0;
}

// New push phase: the current violators become the candidate list of round t0 = rnd[0].
__global__ void k_push_begin(PushState S, const int32_t* __restrict__ viol, const int32_t* __restrict__ nviol) {
  const int t0 = S.rnd[0];
  const int nv = *nviol;
  int32_t* dst = S.list[t0 % 3];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += gridDim.x * blockDim.x) {
    const int v = viol[i];
    dst[i] = v;
    S.ctag[v] = t0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.count[t0 % 3] = nv;
    S.count[(t0 + 1) % 3] = 0;
    S.count[(t0 + 2) % 3] = 0;
    S.rnd[1] = t0;
    *S.done = 0;
  }
}

// ---------------------------------------------------------------------------------------
// Host-driven FP64 push round of the plain per-batch path (solvers.py:100-125 with the
// FrontierOps scatter): D = R[c]/diag, Z[c] += D, R[c] = 0, R[N(c)] += W D (FP64 atomics),
// mark c and N(c).  The host compacts the marks and filters violators between rounds.
template <int CPL>
__global__ __launch_bounds__(256) void k_push64_round(int nc, const int32_t* __restrict__ cand, int C1,
                                                      int64_t ld, int g_shift, int lc_shift,
                                                      const int32_t* __restrict__ indptr,
                                                      const int32_t* __restrict__ rowlen,
                                                      const int32_t* __restrict__ col,
                                                      const float* __restrict__ w,
                                                      const uint8_t* __restrict__ solved,
                                                      const double* __restrict__ diag, double* __restrict__ Z,
                                                      double* __restrict__ R, const double* __restrict__ Dsnap,
                                                      uint8_t* __restrict__ mark,
                                                      unsigned long long* __restrict__ edges) {
  const Lanes L(g_shift, lc_shift);
  const int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> g_shift;
  if (g >= nc) return;
  const int u = cand[g];
  const int64_t o = (int64_t)u * ld;
  double d[CPL];
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int c = L.l + k * L.LC;
    d[k] = c < C1 ? Dsnap[g * (int64_t)C1 + c] : 0.0;
  }
  const int beg = indptr[u], end = beg + rowlen[u];
  for (int j = beg + L.e; j < end; j += L.LE) {
    const int v = col[j];
    if (v < 0 || !solved[v]) continue;
    const double wv = (double)w[j];
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = L.l + k * L.LC;
      if (c < C1) atomic_add(&R[(int64_t)v * ld + c], wv * d[k]);
    }
    if (L.l == 0) mark[v] = 1;
  }
  if (L.lane == 0) {
    mark[u] = 1;
    if (edges) atomicAdd(edges, (unsigned long long)(end - beg));
  }
  (void)o;
}

// D[g] = R[c_g]/diag, Z[c_g] += D[g], R[c_g] = 0 (before the scatter, as the reference does).
__global__ void k_take64(int nc, const int32_t* __restrict__ cand, int C1, int64_t ld,
                         const double* __restrict__ diag, double* __restrict__ Z, double* __restrict__ R,
                         double* __restrict__ Dsnap) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t g = t / C1;
  if (g >= nc) return;
  const int c = (int)(t - g * C1);
  const int u = cand[g];
  const int64_t o = (int64_t)u * ld + c;
  const double d = R[o] / diag[u];
  Dsnap[t] = d;
  Z[o] += d;
  R[o] = 0.0;
}

}  // namespace dlp

// ===========================================================================================
// Handle and C ABI.

using namespace dlp;

namespace {

struct LP {
  dlp_lp_desc d;
  int stampv = 0;       // changed-row stamp (per batch)
  int bstamp_host = 0;  // batch-touched stamp
  // scratch
  int32_t* stamp = nullptr;
  int32_t* rows = nullptr;      // changed rows
  int32_t* cnt = nullptr;       // device counters (see CNT_*)
  int32_t* overflow = nullptr;
  int32_t* toggled = nullptr;
  int32_t* parent = nullptr;
  int32_t* roots = nullptr;
  uint8_t* grounded = nullptr;
  int32_t* viol = nullptr;
  double* eps64 = nullptr;
  double* rho = nullptr;
  double* mass = nullptr;       // refresh of the changed rows: sum of violators' max ratio/eps
  uint8_t* mark = nullptr;      // FP64 path marks
  double* dsnap = nullptr;      // FP64 path increments of the current candidates
  int32_t* cand64 = nullptr;
  void* scan_ws = nullptr;
  size_t scan_bytes = 0;
  PushState S;
  int64_t* host = nullptr;      // pinned mirror of stat_dev (kStatWords words)
  int64_t* stat_dev = nullptr;  // status words gathered on the device for one copy
  hipGraphExec_t gexec = nullptr;
  int graph_rounds = 0;
  hipStream_t graph_stream = nullptr;
  RowMap map;                   // lane layout for C1 columns
  RowMap map_seed;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // global solve (column-wise PCG on the FP32 state)
  float* P = nullptr;
  float* AP = nullptr;
  uint16_t* P16 = nullptr;      // low-precision directions of the column PCG (pcg_p16 != 0)
  int64_t ldh = 0;              // its row length in 16-bit units (direction_ldh)
  int p16_slots = 0;            // direction slots allocated in P16 (slot_units each)
  int pcg_group = 0;            // iterations per deferred solution update (0: every iteration)
  float* pcg_ahist = nullptr;   // their step lengths, pcg_group x C1 (rounded up to even)
  int ahist_len = 0;
  int pcg_p16 = 0;              // 0 FP32 directions, 1 BF16, 2 FP16 / 3 FP8 with a per-column scale
  // rz, rz_new, pAp, alpha, beta, rp (C1 doubles each), then sig, rmax, pmax, emax (C1 floats
  // each, in slots of C1 doubles)
  double* pcg_scal = nullptr;
  // two-level preconditioner of the column PCG: coarse_na aggregates requested (0: Jacobi
  // only), rebuilt every coarse_refresh global solves; cg_* = the current layout and buffers
  int coarse_na = 0, coarse_refresh = 64;
  int cg_na = 0, cg_min = 0, cg_age = -1;
  // fewest iterations of a global solve since the coarse operator was built, and whether a later
  // solve needed a quarter more (the operator has gone stale: rebuild before the next solve)
  int64_t cg_base_iters = 0;
  int cg_stale = 0;
  int cg_builds = 0;            // coarse operators built; the boundaries are searched every 4th
  double* cg_Ac = nullptr;      // buffers sized for kCoarseMax aggregates
  double* cg_Ainv = nullptr;
  double* cg_rc = nullptr;
  float* cg_ec = nullptr;
  int32_t* cg_bounds = nullptr; // kCoarseMax + 1
  uint8_t* cg_aggmap = nullptr; // n
  int32_t* cg_diff = nullptr;   // n, then the crossing counts (n)
  int32_t* cg_cand = nullptr;   // kCoarseMax candidate boundaries, 4 ints each
  void* cg_ws = nullptr;        // scan workspace
  size_t cg_ws_bytes = 0;
  float* pcg_viol = nullptr;
  hipGraphExec_t pcg_exec = nullptr;
  int pcg_graph_iters = 0;
  double pcg_ms = 0.0;
  int64_t pcg_iters = 0;
  double* bk = nullptr;         // block PCG k x k matrices: G, PAP, Gn, alpha, beta
  int32_t* gflags = nullptr;    // [0] need full CC, [1] inert unlabeled vertices, [2] checks run,
                                // [4] giant component root (CC scratch)
  hipGraphExec_t bpcg_exec = nullptr;
  int bpcg_graph_iters = 0;
  unsigned long long* pair_keys = nullptr;  // insertion pair table (kPairTableCap slots)
  int32_t* pair_vals = nullptr;
  int64_t dev_bytes = 0;        // device memory this handle holds (its own buffers)
  int last_refresh_kernel = -1; // kernel of the last refresh: 0 k_refresh, 1 k_refresh_v4, 2 k_verify
  int applies = 0;              // apply calls since the batch began (the changed-row list grows)
  long long batch_ops = 0;      // pair operations applied since the batch began
  double push_ms = 0.0;         // summed duration of push-round launches (events)
  int64_t push_launched = 0;    // rounds launched
};

enum { CNT_ROWS = 0, CNT_MISSING, CNT_OVERFLOW, CNT_TOGGLED, CNT_VIOL, CNT_CAND64, CNT_N };
// counters past the status block's first words: pairs refused by the apply (ids outside
// [0, n), insertion weights not finite and positive), insertion occurrences superseded by a
// later one of the same pair; CNT_ALL counters are cleared per batch
enum { CNT_INVALID = CNT_N, CNT_DUP, CNT_ALL };
// status words: [0, 32) counters, [32, 32 + 256) rho as FP64 bits
constexpr int kStatRho = 32;
constexpr int kStatWords = kStatRho + 256;

// Gathers the host-visible status into one buffer so the host reads it with a single copy:
// counters, push list count and round, edge visits, push statistics, and (if rho) rho[C1].
__global__ void k_status_gather(const int32_t* __restrict__ cnt, const int32_t* __restrict__ bcount,
                                const int32_t* __restrict__ rnd, const unsigned long long* __restrict__ edges,
                                const unsigned long long* __restrict__ stats, const double* __restrict__ mass,
                                const double* __restrict__ rho, int C1, int64_t* __restrict__ out) {
  const int t = threadIdx.x;
  if (t < CNT_N) out[t] = cnt[t];
  if (t == 0) {
    out[CNT_N] = *bcount;
    out[CNT_N + 1] = rnd[0];
    out[CNT_N + 2] = (int64_t)*edges;
    out[CNT_N + 9] = __double_as_longlong(*mass);
    out[CNT_N + 10] = cnt[CNT_INVALID];
    out[CNT_N + 11] = cnt[CNT_DUP];
  }
  if (t < 3) out[CNT_N + 3 + t] = (int64_t)stats[t];
  if (t == 3) out[CNT_N + 6] = rho ? __double_as_longlong(rho[C1]) : 0;   // the verification's h maximum
  if (rho)
    for (int c = t; c < C1; c += blockDim.x) reinterpret_cast<double*>(out + kStatRho)[c] = rho[c];
}

// out[0] = *done, out[1] = *edges (the push loop's per-launch check).
__global__ void k_push_probe(const int32_t* __restrict__ done, const unsigned long long* __restrict__ edges,
                             int64_t* __restrict__ out) {
  if (threadIdx.x == 0) {
    out[0] = *done;
    out[1] = (int64_t)*edges;
  }
}

template <typename T>
int dev_alloc(T** p, size_t count, int64_t* acc) {
  const size_t bytes = sizeof(T) * (count ? count : 1);
  DLP_CHECK_HIP(hipMalloc((void**)p, bytes));
  *acc += (int64_t)bytes;
  return DLP_OK;
}

// A set of device buffers that is allocated whole or not at all: commit() allocates every
// request, and on a failure frees what it took and leaves every destination pointer untouched,
// so a handle never holds part of a group (a later call simply tries again).
struct AllocGroup {
  void** dst[16];
  size_t bytes[16];
  int k = 0;
  void add(void** d, size_t b) {
    dst[k] = d;
    bytes[k] = b ? b : 1;
    ++k;
  }
  int commit(int64_t* acc, const char* what) {
    void* got[16];
    for (int i = 0; i < k; ++i) {
      if (hipMalloc(&got[i], bytes[i]) != hipSuccess) {
        (void)hipGetLastError();
        for (int j = 0; j < i; ++j) (void)hipFree(got[j]);
        set_error("%s: device allocation of %zu bytes failed", what, bytes[i]);
        return DLP_ERR_HIP;
      }
    }
    for (int i = 0; i < k; ++i) {
      *dst[i] = got[i];
      *acc += (int64_t)bytes[i];
    }
    return DLP_OK;
  }
};

int lp_free(LP* h) {
  void* ptrs[] = {h->stamp, h->rows, h->cnt, h->overflow, h->toggled, h->parent, h->roots, h->grounded,
                  h->viol, h->eps64, h->rho, h->mark, h->dsnap, h->cand64, h->scan_ws,
                  h->S.ctag, h->S.list[0],
                  h->S.list[1], h->S.list[2], h->S.count, h->S.rnd, h->S.done, h->S.btag,
                  h->S.blist, h->S.bcount, (void*)h->S.bstamp, (void*)h->S.eps, h->S.edges,
                  h->S.stats};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (h->host) (void)hipHostFree(h->host);
  if (h->stat_dev) (void)hipFree(h->stat_dev);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->P) (void)hipFree(h->P);
  if (h->AP) (void)hipFree(h->AP);
  if (h->P16) (void)hipFree(h->P16);
  if (h->pcg_ahist) (void)hipFree(h->pcg_ahist);
  if (h->cg_Ac) (void)hipFree(h->cg_Ac);
  if (h->cg_Ainv) (void)hipFree(h->cg_Ainv);
  if (h->cg_rc) (void)hipFree(h->cg_rc);
  if (h->cg_ec) (void)hipFree(h->cg_ec);
  if (h->cg_bounds) (void)hipFree(h->cg_bounds);
  if (h->cg_aggmap) (void)hipFree(h->cg_aggmap);
  if (h->cg_diff) (void)hipFree(h->cg_diff);
  if (h->cg_cand) (void)hipFree(h->cg_cand);
  if (h->cg_ws) (void)hipFree(h->cg_ws);
  if (h->pcg_scal) (void)hipFree(h->pcg_scal);
  if (h->pcg_viol) (void)hipFree(h->pcg_viol);
  if (h->pcg_exec) (void)hipGraphExecDestroy(h->pcg_exec);
  if (h->bk) (void)hipFree(h->bk);
  if (h->gflags) (void)hipFree(h->gflags);
  if (h->pair_keys) (void)hipFree(h->pair_keys);
  if (h->pair_vals) (void)hipFree(h->pair_vals);
  if (h->mass) (void)hipFree(h->mass);
  if (h->bpcg_exec) (void)hipGraphExecDestroy(h->bpcg_exec);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  delete h;
  return DLP_OK;
}

RefreshArgs refresh_args(LP* h, int which) {
  RefreshArgs A;
  A.nrows = h->d.n;
  A.rows = nullptr;
  A.rows_count = nullptr;
  if (which == 0) {
    A.rows = h->rows;
    A.rows_count = h->cnt + CNT_ROWS;
  } else if (which == 1) {
    A.rows = h->S.blist;
    A.rows_count = h->S.bcount;
  }
  A.C1 = h->d.C1;
  A.K2 = h->d.K2;
  A.eta = h->d.eta;
  A.prior = h->d.prior;
  A.indptr = h->d.indptr;
  A.rowlen = h->d.rowlen;
  A.col = h->d.col;
  A.w = h->d.w;
  A.labeled = h->d.labeled;
  A.y = h->d.y;
  A.solved = h->d.solved;
  A.ld = h->d.ld;
  A.diag64 = h->d.diag64;
  A.diag32 = h->d.diag32;
  A.eps = h->eps64;
  A.viol = h->viol;
  A.nviol = h->cnt + CNT_VIOL;
  A.rho = h->rho;
  A.hmax = nullptr;
  A.mass = which == 0 ? h->mass : nullptr;
  return A;
}

unsigned persistent_grid(int n_groups, int g_shift) {
  long long threads = (long long)n_groups << g_shift;
  unsigned g = grid_for(threads);
  return g > 4096 ? 4096 : g;
}

}  // namespace

extern "C" int dlp_lp_create(const dlp_lp_desc* d, void** handle) {
  DLP_CHECK_ARG(d && handle, "lp_create: null argument");
  DLP_CHECK_ARG(d->n >= 1 && d->C1 >= 1 && d->C1 <= 256 && d->ld >= d->C1, "lp_create: bad sizes");
  DLP_CHECK_ARG(d->indptr && d->rowlen && d->col && d->w && d->labeled && d->y && d->solved &&
                    d->labcnt && d->Z32 && d->R32 && d->diag64 && d->diag32,
                "lp_create: null pointer in descriptor");
  LP* h = new LP();
  memset(&h->S, 0, sizeof(h->S));
  h->d = *d;
  const int n = d->n;
  const size_t nz = (size_t)n * d->ld;
  int rc = DLP_OK;
#define A_(p, c) if (!rc) rc = dev_alloc(&(p), (c), &h->dev_bytes)
  A_(h->stamp, n); A_(h->rows, n); A_(h->cnt, CNT_ALL); A_(h->overflow, 1 << 20);
  A_(h->toggled, n); A_(h->parent, n); A_(h->roots, n); A_(h->grounded, n); A_(h->viol, n);
  A_(h->eps64, d->C1); A_(h->rho, d->C1 + 1); A_(h->mark, n); A_(h->cand64, n);
  A_(h->S.ctag, n); A_(h->S.list[0], n); A_(h->S.list[1], n); A_(h->S.list[2], n);
  A_(h->S.count, 3); A_(h->S.rnd, 2); A_(h->S.done, 1); A_(h->S.btag, n); A_(h->S.blist, n);
  A_(h->S.bcount, 1); A_(h->S.edges, 1); A_(h->S.stats, 4);
  int32_t* bst = nullptr;
  float* eps32 = nullptr;
  A_(bst, 1); A_(eps32, d->C1); A_(h->gflags, 8); A_(h->mass, 1);
  A_(h->pair_keys, kPairTableCap); A_(h->pair_vals, kPairTableCap);
#undef A_
  h->S.bstamp = bst;
  h->S.eps = eps32;
  if (!rc && d->Z64) rc = dev_alloc(&h->dsnap, nz, &h->dev_bytes);
  h->scan_bytes = dlp_scan_workspace_bytes(n);
  if (!rc) rc = dev_alloc((uint8_t**)&h->scan_ws, h->scan_bytes, &h->dev_bytes);
  if (!rc && hipHostMalloc((void**)&h->host, kStatWords * sizeof(int64_t)) != hipSuccess) rc = DLP_ERR_HIP;
  if (!rc) rc = dev_alloc(&h->stat_dev, kStatWords, &h->dev_bytes);
  if (rc) {
    set_error("lp_create: device allocation failed");
    lp_free(h);
    return DLP_ERR_HIP;
  }
  // initial contents: stamps/tags -1 (never equal to a live stamp), counters 0
  (void)hipMemset(h->stamp, 0xff, sizeof(int32_t) * n);
  (void)hipMemset(h->S.ctag, 0xff, sizeof(int32_t) * n);
  (void)hipMemset(h->S.btag, 0xff, sizeof(int32_t) * n);
  (void)hipMemset(h->cnt, 0, sizeof(int32_t) * CNT_ALL);
  (void)hipMemset(h->S.count, 0, sizeof(int32_t) * 3);
  (void)hipMemset(h->S.rnd, 0, sizeof(int32_t) * 2);
  (void)hipMemset(h->S.done, 0, sizeof(int32_t));
  (void)hipMemset(h->S.bcount, 0, sizeof(int32_t));
  (void)hipMemset((void*)h->S.bstamp, 0, sizeof(int32_t));
  (void)hipMemset(h->S.edges, 0, sizeof(unsigned long long));
  (void)hipMemset(h->S.stats, 0, 4 * sizeof(unsigned long long));
  (void)hipMemset(h->mark, 0, n);
  (void)hipMemset(h->gflags, 0, 8 * sizeof(int32_t));
  h->S.n = n;
  h->S.C1 = d->C1;
  h->S.ld = d->ld;
  h->S.indptr = d->indptr;
  h->S.rowlen = d->rowlen;
  h->S.col = d->col;
  h->S.w = d->w;
  h->S.solved = d->solved;
  h->S.diag = d->diag32;
  h->S.Z = d->Z32;
  h->S.R = d->R32;
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
    set_error("lp_create: event creation failed");
    lp_free(h);
    return DLP_ERR_HIP;
  }
  h->map = choose_map(d->C1, 20.0);
  h->map_seed = choose_map(d->C1, 1.0);
  if (hipDeviceSynchronize() != hipSuccess) {
    set_error("lp_create: initialisation failed");
    lp_free(h);
    return DLP_ERR_HIP;
  }
  *handle = h;
  return DLP_OK;
}

extern "C" int dlp_lp_destroy(void* handle) {
  if (!handle) return DLP_OK;
  (void)hipDeviceSynchronize();
  return lp_free((LP*)handle);
}

extern "C" int dlp_lp_set_store(void* handle, const int32_t* indptr, int32_t* rowlen, int32_t* col,
                                float* w) {
  DLP_CHECK_ARG(handle && indptr && rowlen && col && w, "lp_set_store: null argument");
  LP* h = (LP*)handle;
  h->d.indptr = indptr;
  h->d.rowlen = rowlen;
  h->d.col = col;
  h->d.w = w;
  h->S.indptr = indptr;
  h->S.rowlen = rowlen;
  h->S.col = col;
  h->S.w = w;
  if (h->gexec) {  // the captured rounds hold the old pointers
    (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  if (h->pcg_exec) {
    (void)hipGraphExecDestroy(h->pcg_exec);
    h->pcg_exec = nullptr;
  }
  if (h->bpcg_exec) {
    (void)hipGraphExecDestroy(h->bpcg_exec);
    h->bpcg_exec = nullptr;
  }
  return DLP_OK;
}

extern "C" int dlp_lp_set_eps(void* handle, const double* eps, double push_scale, void* stream) {
  DLP_CHECK_ARG(handle && eps && push_scale > 0, "lp_set_eps: bad argument");
  LP* h = (LP*)handle;
  const int C1 = h->d.C1;
  float e32[256];
  for (int c = 0; c < C1; ++c) e32[c] = (float)(eps[c] * push_scale);
  hipStream_t st = (hipStream_t)stream;
  DLP_CHECK_HIP(hipMemcpyAsync(h->eps64, eps, sizeof(double) * C1, hipMemcpyHostToDevice, st));
  DLP_CHECK_HIP(hipMemcpyAsync((void*)h->S.eps, e32, sizeof(float) * C1, hipMemcpyHostToDevice, st));
  DLP_CHECK_HIP(hipStreamSynchronize(st));
  return DLP_OK;
}

__global__ void k_begin_batch(int32_t* cnt, int32_t* bstamp, int32_t* bcount, unsigned long long* edges) {
  if (threadIdx.x < CNT_ALL) cnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    *bstamp = *bstamp + 1;
    *bcount = 0;
    *edges = 0;
  }
}

extern "C" int dlp_lp_begin_batch(void* handle, void* stream) {
  DLP_CHECK_ARG(handle, "lp_begin_batch: null handle");
  LP* h = (LP*)handle;
  h->stampv += 1;
  h->applies = 0;
  h->batch_ops = 0;
  hipLaunchKernelGGL(k_begin_batch, dim3(1), dim3(64), 0, (hipStream_t)stream, h->cnt,
                     (int32_t*)h->S.bstamp, h->S.bcount, h->S.edges);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_lp_apply(void* handle, int nd, const int32_t* du, const int32_t* dv, int ni,
                            const int32_t* iu, const int32_t* iv, const float* iw, void* stream) {
  DLP_CHECK_ARG(handle && nd >= 0 && ni >= 0, "lp_apply: bad argument");
  DLP_CHECK_ARG((nd == 0 || (du && dv)) && (ni == 0 || (iu && iv && iw)), "lp_apply: null arrays");
  DLP_CHECK_ARG(ni <= (1 << 19), "lp_apply: at most 524288 insertions per call");
  LP* h = (LP*)handle;
  hipStream_t st = (hipStream_t)stream;
  const int n = h->d.n;
  if (nd > 0) {
    hipLaunchKernelGGL(k_apply_delete, dim3(grid_for(2LL * nd)), dim3(kBlock), 0, st, n, nd, du, dv,
                       h->d.indptr, h->d.rowlen, h->d.col, h->d.w, h->d.labeled, h->d.labcnt,
                       h->stamp, h->stampv, h->rows, h->cnt + CNT_ROWS, h->cnt + CNT_MISSING,
                       h->cnt + CNT_INVALID);
    DLP_CHECK_LAUNCH();
    // the changed-row list holds this call's rows after those of earlier calls of the batch
    hipLaunchKernelGGL(k_compact_rows, dim3(grid_for(h->applies ? (long long)n : 2LL * nd)), dim3(kBlock), 0, st,
                       h->cnt + CNT_ROWS, h->rows, h->d.indptr, h->d.rowlen, h->d.col, h->d.w);
    DLP_CHECK_LAUNCH();
  }
  if (ni > 0) {
    int bits = 1;
    while ((1LL << bits) < 2LL * ni) ++bits;
    const int cap = 1 << bits;   // <= kPairTableCap (ni <= 2^19)
    hipLaunchKernelGGL(k_pair_table_clear, dim3(grid_for(cap)), dim3(kBlock), 0, st, cap, h->pair_keys,
                       h->pair_vals);
    hipLaunchKernelGGL(k_pair_table_fill, dim3(grid_for(ni)), dim3(kBlock), 0, st, n, ni, iu, iv, iw, bits,
                       h->pair_keys, h->pair_vals);
    hipLaunchKernelGGL(k_apply_insert, dim3(grid_for(2LL * ni * 64)), dim3(kBlock), 0, st, n, ni, iu, iv, iw, bits,
                       (const unsigned long long*)h->pair_keys, (const int32_t*)h->pair_vals, h->d.indptr,
                       h->d.rowlen, h->d.col, h->d.w, h->d.labeled, h->d.labcnt, h->stamp, h->stampv, h->rows,
                       h->cnt + CNT_ROWS, h->overflow, h->cnt + CNT_OVERFLOW, h->cnt + CNT_INVALID,
                       h->cnt + CNT_DUP);
    DLP_CHECK_LAUNCH();
  }
  h->applies += 1;
  h->batch_ops += (long long)nd + ni;
  return DLP_OK;
}

// The ordered operation sequence (see k_ops_table_fill): deletions (w == 0) and insertions
// (w > 0) of a group of requests in one call, each pair taking its last operation.
extern "C" int dlp_lp_apply_ops(void* handle, int nops, const int32_t* u, const int32_t* v, const float* w,
                                void* stream) {
  DLP_CHECK_ARG(handle && nops >= 0, "lp_apply_ops: bad argument");
  DLP_CHECK_ARG(nops == 0 || (u && v && w), "lp_apply_ops: null arrays");
  DLP_CHECK_ARG(nops <= (1 << 19), "lp_apply_ops: at most 524288 operations per call");
  if (nops == 0) return DLP_OK;
  LP* h = (LP*)handle;
  hipStream_t st = (hipStream_t)stream;
  const int n = h->d.n;
  int bits = 1;
  while ((1LL << bits) < 2LL * nops) ++bits;
  const int cap = 1 << bits;   // <= kPairTableCap
  hipLaunchKernelGGL(k_pair_table_clear, dim3(grid_for(cap)), dim3(kBlock), 0, st, cap, h->pair_keys, h->pair_vals);
  hipLaunchKernelGGL(k_ops_table_fill, dim3(grid_for(nops)), dim3(kBlock), 0, st, n, nops, u, v, w, bits,
                     h->pair_keys, h->pair_vals, h->cnt + CNT_INVALID);
  DLP_CHECK_LAUNCH();
  const unsigned long long* keys = h->pair_keys;
  const int32_t* vals = h->pair_vals;
  hipLaunchKernelGGL(k_ops_delete, dim3(grid_for(2LL * nops)), dim3(kBlock), 0, st, n, nops, u, v, w, bits, keys,
                     vals, h->d.indptr, h->d.rowlen, h->d.col, h->d.w, h->d.labeled, h->d.labcnt, h->stamp,
                     h->stampv, h->rows, h->cnt + CNT_ROWS, h->cnt + CNT_MISSING);
  DLP_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_compact_rows, dim3(grid_for(h->applies ? (long long)n : 2LL * nops)), dim3(kBlock), 0, st,
                     h->cnt + CNT_ROWS, h->rows,
                     h->d.indptr, h->d.rowlen, h->d.col, h->d.w);
  DLP_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ops_insert, dim3(grid_for(2LL * nops * 64)), dim3(kBlock), 0, st, n, nops, u, v, w, bits,
                     keys, vals, h->d.indptr, h->d.rowlen, h->d.col, h->d.w, h->d.labeled, h->d.labcnt, h->stamp,
                     h->stampv, h->rows, h->cnt + CNT_ROWS, h->overflow, h->cnt + CNT_OVERFLOW);
  DLP_CHECK_LAUNCH();
  h->applies += 1;
  h->batch_ops += nops;
  return DLP_OK;
}

__global__ void k_retry_gather(const int32_t* __restrict__ ovf, const int32_t* __restrict__ novf, int ni,
                               const int32_t* __restrict__ iu, const int32_t* __restrict__ iv,
                               const float* __restrict__ iw, int32_t* __restrict__ ru,
                               int32_t* __restrict__ rv, float* __restrict__ rw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *novf) return;
  const int e = ovf[i];
  // directed entry e: row a gets b; as a one-direction insertion (b, a) is stored already
  ru[i] = e < ni ? iu[e] : iv[e - ni];
  rv[i] = e < ni ? iv[e] : iu[e - ni];
  rw[i] = e < ni ? iw[e] : iw[e - ni];
}

// Directed re-insertion (each overflowed entry is one direction only), one wavefront each,
// by the same upsert as the batch insertion (the entries are distinct pairs).
__global__ __launch_bounds__(256) void k_apply_insert_directed(
    const int32_t* __restrict__ nent, const int32_t* __restrict__ ru, const int32_t* __restrict__ rv,
    const float* __restrict__ rw, const int32_t* __restrict__ indptr, int32_t* __restrict__ rowlen,
    int32_t* __restrict__ col, float* __restrict__ w, const uint8_t* __restrict__ labeled,
    int32_t* __restrict__ labcnt, int32_t* __restrict__ stamp, int stampv, int32_t* __restrict__ rows,
    int32_t* __restrict__ nrows, int32_t* __restrict__ overflow, int32_t* __restrict__ noverflow) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (i >= *nent) return;   // wave-uniform
  upsert_entry(ru[i], rv[i], rw[i], lane, indptr, rowlen, col, w, labeled, labcnt, stamp, stampv, rows, nrows,
               overflow, noverflow, (int32_t)i);
}

extern "C" int dlp_lp_regrow_plan(void* handle, int extra, double growth, int32_t* new_indptr,
                                  int64_t* total, void* stream) {
  DLP_CHECK_ARG(handle && new_indptr && total && extra >= 1, "lp_regrow_plan: bad argument");
  LP* h = (LP*)handle;
  const int n = h->d.n;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_regrow_caps, dim3(grid_for(n)), dim3(kBlock), 0, st, n, h->d.rowlen, extra,
                     (float)growth, new_indptr);
  DLP_CHECK_LAUNCH();
  int rc = exclusive_scan_i32(n, new_indptr, new_indptr, new_indptr + n, h->scan_ws, h->scan_bytes, st);
  if (rc) return rc;
  int32_t t = 0;
  DLP_CHECK_HIP(hipMemcpyAsync(&t, new_indptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  DLP_CHECK_HIP(hipStreamSynchronize(st));
  *total = t;
  return DLP_OK;
}

// Copies the store into the new arrays, switches the handle to them and re-inserts the
// overflowed entries of the last apply (iu/iv/iw: that call's insertion arrays).
extern "C" int dlp_lp_regrow_commit(void* handle, const int32_t* new_indptr, int32_t* new_col,
                                    float* new_w, int ni, const int32_t* iu, const int32_t* iv,
                                    const float* iw, int32_t* scratch3, void* stream) {
  DLP_CHECK_ARG(handle && new_indptr && new_col && new_w, "lp_regrow_commit: null argument");
  LP* h = (LP*)handle;
  const int n = h->d.n;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_regrow_copy, dim3(grid_for((long long)n * 8)), dim3(kBlock), 0, st, n, h->d.indptr,
                     h->d.rowlen, h->d.col, h->d.w, new_indptr, new_col, new_w);
  DLP_CHECK_LAUNCH();
  int rc = dlp_lp_set_store(handle, new_indptr, h->d.rowlen, new_col, new_w);
  if (rc) return rc;
  if (ni > 0 && scratch3) {
    // scratch3: 3 * (1<<20) int32 (ru, rv, rw) of caller memory
    int32_t* ru = scratch3;
    int32_t* rv = scratch3 + (1 << 20);
    float* rw = (float*)(scratch3 + 2 * (1 << 20));
    int32_t* cnt2 = h->cnt + CNT_CAND64;  // reuse as the count of entries being retried
    DLP_CHECK_HIP(hipMemcpyAsync(cnt2, h->cnt + CNT_OVERFLOW, sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_retry_gather, dim3(grid_for(1 << 20)), dim3(kBlock), 0, st, h->overflow, cnt2, ni,
                       iu, iv, iw, ru, rv, rw);
    DLP_CHECK_LAUNCH();
    DLP_CHECK_HIP(hipMemsetAsync(h->cnt + CNT_OVERFLOW, 0, sizeof(int32_t), st));
    hipLaunchKernelGGL(k_apply_insert_directed, dim3(grid_for((1LL << 20) * 64)), dim3(kBlock), 0, st, cnt2, ru,
                       rv, rw, new_indptr, h->d.rowlen, new_col, new_w, h->d.labeled, h->d.labcnt, h->stamp,
                       h->stampv, h->rows, h->cnt + CNT_ROWS, h->overflow, h->cnt + CNT_OVERFLOW);
    DLP_CHECK_LAUNCH();
  }
  return DLP_OK;
}

__global__ void k_ground_prep(int32_t* flags, int force, int32_t* toggled_count) {
  if (threadIdx.x == 0) {
    flags[0] = force ? 1 : 0;
    *toggled_count = 0;
  }
}
__global__ void k_ground_inert_reset(int32_t* flags) {
  if (threadIdx.x == 0 && flags[0]) flags[1] = 0;
}

static int ground_impl(LP* h, int force, int nd, const int32_t* du, const int32_t* dv, int ni, hipStream_t st) {
  const int n = h->d.n;
  int32_t* need = h->gflags;
  hipLaunchKernelGGL(k_ground_prep, dim3(1), dim3(64), 0, st, h->gflags, force, h->cnt + CNT_TOGGLED);
  if (!force && nd > 0) {
    hipLaunchKernelGGL(k_ground_ball, dim3((unsigned)((2LL * nd + 3) / 4)), dim3(kBlock), 0, st, n, nd, du, dv, ni,
                       h->d.indptr, h->d.rowlen, h->d.col, h->d.labeled, h->d.labcnt, h->gflags);
  } else if (!force) {
    // insertions only: they matter when inert vertices exist
    hipLaunchKernelGGL(k_ground_ball, dim3(1), dim3(kBlock), 0, st, n, 0, du, dv, ni, h->d.indptr, h->d.rowlen,
                       h->d.col, h->d.labeled, h->d.labcnt, h->gflags);
  }
  DLP_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ground_inert_reset, dim3(1), dim3(64), 0, st, h->gflags);
  hipLaunchKernelGGL(k_ground_init, dim3(grid_for(n)), dim3(kBlock), 0, st, n, h->grounded, need);
  DLP_CHECK_LAUNCH();
  // components of the unlabeled subgraph (labeled vertices excluded)
  int rc = uf_components_csr(n, h->d.indptr, h->d.rowlen, h->d.col, h->d.labeled, 0, h->parent, h->gflags + 4,
                             need, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ground_roots, dim3(grid_for(n)), dim3(kBlock), 0, st, n, h->parent, h->d.labeled,
                     h->d.labcnt, h->grounded, h->roots, need);
  DLP_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ground_toggle, dim3(grid_for(n)), dim3(kBlock), 0, st, n, h->roots, h->d.labeled,
                     h->grounded, h->d.solved, h->toggled, h->cnt + CNT_TOGGLED, h->stamp, h->stampv,
                     h->rows, h->cnt + CNT_ROWS, need, h->gflags + 1);
  DLP_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_mark_neighbours_changed, dim3(grid_for((long long)n * 8)), dim3(kBlock), 0, st,
                     h->toggled, h->cnt + CNT_TOGGLED, h->d.indptr, h->d.rowlen, h->d.col, h->stamp,
                     h->stampv, h->rows, h->cnt + CNT_ROWS);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_lp_ground(void* handle, void* stream) {
  DLP_CHECK_ARG(handle, "lp_ground: null handle");
  return ground_impl((LP*)handle, 1, 0, nullptr, nullptr, 0, (hipStream_t)stream);
}

// Incremental grounding after an apply of nd deletions (du, dv: that call's arrays) and ni
// insertions: the exact local test decides on the device whether the full pass runs.
extern "C" int dlp_lp_ground_incremental(void* handle, int nd, const int32_t* du, const int32_t* dv, int ni,
                                         void* stream) {
  DLP_CHECK_ARG(handle && nd >= 0 && ni >= 0 && (nd == 0 || (du && dv)), "lp_ground_incremental: bad argument");
  LP* h = (LP*)handle;
  // Past n/16 endpoints the bounded searches (one wavefront each) cost more than the full
  // connected-components pass: run that instead.
  const int force = 2LL * nd > (long long)h->d.n / 16 ? 1 : 0;
  return ground_impl(h, force, nd, du, dv, ni, (hipStream_t)stream);
}

// [0] full pass requested by the last grounding, [1] inert unlabeled vertices, [2] local checks run.
extern "C" int dlp_lp_ground_flags(void* handle, int32_t* out, void* stream) {
  DLP_CHECK_ARG(handle && out, "lp_ground_flags: null argument");
  LP* h = (LP*)handle;
  DLP_CHECK_HIP(hipMemcpyAsync(out, h->gflags, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, (hipStream_t)stream));
  DLP_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return DLP_OK;
}

// Clears the refresh outputs (violator count, column maxima, push mass) in one launch.
__global__ void k_refresh_reset(int32_t* __restrict__ nviol, double* __restrict__ rho, int C1,
                                double* __restrict__ mass) {
  for (int c = threadIdx.x; c <= C1; c += blockDim.x) rho[c] = 0.0;   // C1 maxima and the h maximum
  if (threadIdx.x == 0) {
    *nviol = 0;
    if (mass) *mass = 0.0;
  }
}

// ---------------------------------------------------------------------------------------
// Refresh of listed rows on the pipelined gathers (modes 0 and 1).  The parking trick of the
// verification pass applies once every row outside the solved set holds its parked state: a
// first pass parks the listed rows that left it (or are labeled), then the pipelined pass
// computes R = rhs - diag*Z + W Z over all neighbours of the listed solved rows, with the
// row's diagonal recomputed (mode 0: the sum of its weights to labeled and solved neighbours,
// plus eta) or kept (mode 1: the push changed no equation).

__global__ void k_park_listed(RefreshArgs A, float* __restrict__ Z, float* __restrict__ R) {
  const int nr = *A.rows_count;
  const int lane = threadIdx.x & 63;
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < nr;
       i += ((int64_t)gridDim.x * blockDim.x) >> 6) {
    const int r = A.rows[i];
    if (A.solved[r]) continue;   // wave-uniform
    const int pc = parked_class(A, r);
    const int64_t zo = (int64_t)r * A.ld;
    for (int c = lane; c < A.C1; c += 64) {
      Z[zo + c] = c == pc ? 1.f : 0.f;
      R[zo + c] = 0.f;
    }
    if (lane == 0) {
      A.diag64[r] = 1.0;
      A.diag32[r] = 1.f;
    }
  }
}

struct ListedRefreshEpilogue {
  const LanesP& L;
  const RefreshArgs& A;
  const float* __restrict__ Z;
  float* __restrict__ R;
  bool eqs;                // recompute the diagonal (mode 0)
  double rmax[4] = {0.0, 0.0, 0.0, 0.0};
  double mass = 0.0;       // lane 0: this wave's rows' push mass
  float4 zr;
  double dg = 1.0;
  __device__ __forceinline__ bool owner() const { return L.e == 0 && 4 * L.l < A.C1; }
  __device__ __forceinline__ void pre(int, int r) {
    zr = owner() ? *(reinterpret_cast<const float4*>(Z + (int64_t)r * A.ld) + L.l) : make_float4(0.f, 0.f, 0.f, 0.f);
    if (!A.solved[r]) return;   // wave-uniform; parked by k_park_listed
    if (!eqs) {
      dg = A.diag64[r];
      return;
    }
    const int beg = __builtin_amdgcn_readfirstlane(A.indptr[r]);
    const int end = beg + __builtin_amdgcn_readfirstlane(A.rowlen[r]);
    double sw = 0.0;
    for (int j = beg + L.lane; j < end; j += 64) {
      const int v = A.col[j];
      if (v >= 0 && (A.labeled[v] || A.solved[v])) sw += (double)A.w[j];
    }
    for (int off = 32; off > 0; off >>= 1) sw += __shfl_xor(sw, off, 64);
    dg = sw + A.eta;
  }
  __device__ __forceinline__ void post(int, int r, bool live, const double (&acc)[4]) {
    if (!live) return;  // wave-uniform
    bool viol = false;
    double m = 0.0;
    if (owner()) {
      if (eqs && L.l == 0) {
        A.diag64[r] = dg;
        A.diag32[r] = (float)dg;
      }
      const int C = A.C1 - 1;
      const float zv[4] = {zr.x, zr.y, zr.z, zr.w};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = 4 * L.l + q;
        if (c < A.C1) {
          const double rh = c == C ? dg : A.eta * A.prior;
          const double res = rh - dg * (double)zv[q] + acc[q];
          R[(int64_t)r * A.ld + c] = (float)res;
          const double ratio = ratio_or_inf(res, dg);
          viol |= ratio > A.eps[c];
          rmax[q] = fmax(rmax[q], ratio);
          m = fmax(m, ratio / A.eps[c]);
        }
      }
    }
    const unsigned long long bal = __ballot(viol);
    if (bal && A.viol) {
      // most listed rows violate right after a batch: the wavefront keeps its violating rows in
      // a register (lane k: the k-th) and reserves list space once per 64, instead of one
      // atomic per row on the one counter every wavefront shares
      if (L.lane == nbuf) vbuf = r;
      if (++nbuf == 64) flush_viol();
    }
    if (A.mass) {
      for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
      if (L.lane == 0 && m > 1.0) mass += m;
    }
  }
  int vbuf = 0, nbuf = 0;  // violating rows held for the list (nbuf wave-uniform)
  __device__ __forceinline__ void flush_viol() {
    if (nbuf == 0) return;
    int base = 0;
    if (L.lane == 0) base = atomicAdd(A.nviol, nbuf);
    base = __shfl(base, 0, 64);
    if (L.lane < nbuf) A.viol[base + L.lane] = vbuf;
    nbuf = 0;
  }
};

template <int S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_refresh_pipe(
    RefreshArgs A, int eqs, int xcd, int xbytes, float* __restrict__ Z, float* __restrict__ R) {
  const int nr = *A.rows_count;
  RowPipe G(nr, A.C1, 4, 16, xcd, A.solved, A.indptr, A.rowlen, A.col, A.w, Z, (unsigned)(A.ld * 4), xbytes);
  G.rows = A.rows;
  ListedRefreshEpilogue E{G.L, A, Z, R, eqs != 0};
  run_rows<S, Word::F32x4, double>(G, E);
  E.flush_viol();
  flush_mass(A.mass, E.mass);
  if (!A.rho) return;
  __shared__ double sm[256];
  for (int c = threadIdx.x; c < A.C1; c += blockDim.x) sm[c] = 0.0;
  __syncthreads();
  if (E.owner()) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = 4 * G.L.l + q;
      if (c < A.C1) atomic_max_nonneg(&sm[c], E.rmax[q]);
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < A.C1; c += blockDim.x) global_max_nonneg(A.rho + c, sm[c]);
}

extern "C" int dlp_lp_refresh(void* handle, int which, int prec, void* stream) {
  DLP_CHECK_ARG(handle && which >= 0 && which <= 3 && (prec == 0 || prec == 1), "lp_refresh: bad argument");
  LP* h = (LP*)handle;
  DLP_CHECK_ARG(prec == 0 || (h->d.Z64 && h->d.R64), "lp_refresh: FP64 state not attached");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_refresh_reset, dim3(1), dim3(kBlock), 0, st, h->cnt + CNT_VIOL, h->rho, h->d.C1,
                     which == 0 ? h->mass : nullptr);
  const long long xbytes = (long long)h->d.n * h->d.ld * 4;
  if (which == 3) {
    // verification of every solved row with unchanged equations: the pipelined kernel when the
    // FP32 layout allows it, else the full refresh
    if (prec == 0 && (h->d.ld % 4) == 0 && h->d.C1 <= 256 && tuning().vec4 && tuning().pipe &&
        h->d.n < (1 << 24) && xbytes < (1LL << 31)) {
      RefreshArgs A = refresh_args(h, 2);
      A.hmax = h->rho + h->d.C1;
      const int xcd = tuning().xcd != 0;
      const void* kfn = reinterpret_cast<const void*>(&k_verify<4>);
      long long g = resident_grid(kfn);
      const long long need = ((long long)h->d.n + 3) / 4;
      if (g > need) g = need;
      if (xcd) g = g < 8 ? 8 : g & ~7LL;
      hipLaunchKernelGGL((k_verify<4>), dim3((unsigned)g), dim3(kBlock), 0, st, A, xcd, (int)xbytes, h->d.Z32,
                         h->d.R32);
      DLP_CHECK_LAUNCH();
      h->last_refresh_kernel = 2;
      return DLP_OK;
    }
    which = 2;
  }
  if ((which == 0 || which == 1) && prec == 0 && (h->d.ld % 4) == 0 && h->d.C1 <= 256 && tuning().vec4 &&
      tuning().pipe && h->d.n < (1 << 24) && xbytes < (1LL << 31)) {
    // listed rows on the pipelined gathers: park the listed rows outside the solved set, then
    // the residual (and, for the changed rows, the diagonal) of the listed solved rows
    RefreshArgs A = refresh_args(h, which);
    if (which == 0) {
      hipLaunchKernelGGL(k_park_listed, dim3(1024), dim3(kBlock), 0, st, A, h->d.Z32, h->d.R32);
      DLP_CHECK_LAUNCH();
    }
    const int xcd = tuning().xcd != 0;
    long long g = resident_grid(reinterpret_cast<const void*>(&k_refresh_pipe<4>));
    const long long need = (which == 0 ? 2 * h->batch_ops : (long long)h->d.n) / 4 + 8;
    if (g > need) g = need;
    if (xcd) g = g < 8 ? 8 : g & ~7LL;
    hipLaunchKernelGGL((k_refresh_pipe<4>), dim3((unsigned)g), dim3(kBlock), 0, st, A, which == 0 ? 1 : 0, xcd,
                       (int)xbytes, h->d.Z32, h->d.R32);
    DLP_CHECK_LAUNCH();
    h->last_refresh_kernel = 3;
    return DLP_OK;
  }
  RefreshArgs A = refresh_args(h, which);
  const RowMap& m = h->map;
  // listed rows: the changed list is O(batch) (two rows per pair operation, four rows a block
  // in the one-row-per-wavefront kernel); the touched list can reach n
  unsigned cap = which == 0 ? 512u : (which == 1 ? 2048u : 4096u);
  if (which == 0) cap = (unsigned)std::max<long long>(cap, std::min<long long>(1LL << 20, (2 * h->batch_ops + 3) / 4));
  unsigned grid = persistent_grid(h->d.n, m.g_shift);
  if (grid > cap) grid = cap;
  if (prec == 0 && (h->d.ld % 4) == 0 && tuning().vec4 && h->d.C1 <= 256) {
    // one row per wavefront: the grid is capped at the resident blocks (no tail of late blocks)
    const unsigned res = (unsigned)resident_grid(reinterpret_cast<const void*>(&k_refresh_v4));
    unsigned g4 = grid_for((long long)h->d.n << 6);
    if (g4 > res) g4 = res;
    if (g4 > cap) g4 = cap;
    const int xcd = tuning().xcd != 0;
    if (xcd) g4 = g4 < 8 ? 8 : g4 & ~7u;
    hipLaunchKernelGGL(k_refresh_v4, dim3(g4), dim3(kBlock), 0, st, A, tuning().pack ? 0 : 1, xcd, h->d.Z32,
                       h->d.R32);
    DLP_CHECK_LAUNCH();
    h->last_refresh_kernel = 1;
    return DLP_OK;
  }
#define K(CPLV)                                                                                    \
  if (prec == 0)                                                                                   \
    hipLaunchKernelGGL((k_refresh<float, float, CPLV>), dim3(grid), dim3(kBlock), 0, st, A, m.g_shift, \
                       m.lc_shift, h->d.Z32, h->d.R32);                                            \
  else                                                                                             \
    hipLaunchKernelGGL((k_refresh<double, double, CPLV>), dim3(grid), dim3(kBlock), 0, st, A,     \
                       m.g_shift, m.lc_shift, h->d.Z64, h->d.R64);
  DLP_DISPATCH_CPL(m.cpl, K)
#undef K
  DLP_CHECK_LAUNCH();
  h->last_refresh_kernel = 0;
  return DLP_OK;
}

static int capture_rounds(LP* h, int rounds, hipStream_t st) {
  if (h->gexec && h->graph_rounds == rounds) return DLP_OK;
  if (h->gexec) {
    (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  hipStream_t cs;
  DLP_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  const RowMap& m = h->map;
  const unsigned grid = persistent_grid(h->d.n, m.g_shift);
  hipGraph_t graph;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int r = 0; r < rounds; ++r) {
#define K(CPLV)                                                                              \
  hipLaunchKernelGGL((k_push_round<CPLV>), dim3(grid), dim3(kBlock), 0, cs, h->S, r & 1, m.g_shift, \
                     m.lc_shift);
      DLP_DISPATCH_CPL(m.cpl, K)
#undef K
    }
    e = hipStreamEndCapture(cs, &graph);
  }
  if (e == hipSuccess) {
    e = hipGraphInstantiate(&h->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
  }
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess) {
    set_error("push graph capture failed: %s", hipGetErrorString(e));
    h->gexec = nullptr;
    return DLP_ERR_HIP;
  }
  h->graph_rounds = rounds;
  (void)st;
  return DLP_OK;
}

// Round numbers are even at the start of every capture replay: the graph alternates node
// parity 0,1,0,1,... so it must start on rnd[0].  A push phase begins with rnd[1] = rnd[0]
// (k_push_begin) and a replay of an even number of rounds ends with the round in rnd[0].
extern "C" int dlp_lp_push(void* handle, int max_rounds, int rounds_per_launch, int use_graph,
                           int64_t budget_edges, int64_t* rounds_out, int32_t* converged_out,
                           void* stream) {
  DLP_CHECK_ARG(handle && max_rounds >= 0 && rounds_per_launch >= 2 && (rounds_per_launch % 2) == 0,
                "lp_push: bad argument (rounds_per_launch must be even)");
  LP* h = (LP*)handle;
  hipStream_t st = (hipStream_t)stream;
  const RowMap& m = h->map;
  hipLaunchKernelGGL(k_push_begin, dim3(256), dim3(kBlock), 0, st, h->S, h->viol, h->cnt + CNT_VIOL);
  DLP_CHECK_LAUNCH();
  if (use_graph) {
    int rc = capture_rounds(h, rounds_per_launch, st);
    if (rc) return rc;
  }
  const unsigned grid = persistent_grid(h->d.n, m.g_shift);
  int64_t rounds = 0;
  int32_t converged = 0;
  while (rounds < max_rounds) {
    DLP_CHECK_HIP(hipEventRecord(h->ev0, st));
    if (use_graph) {
      DLP_CHECK_HIP(hipGraphLaunch(h->gexec, st));
    } else {
      for (int r = 0; r < rounds_per_launch; ++r) {
#define K(CPLV)                                                                              \
  hipLaunchKernelGGL((k_push_round<CPLV>), dim3(grid), dim3(kBlock), 0, st, h->S, r & 1, m.g_shift, \
                     m.lc_shift);
        DLP_DISPATCH_CPL(m.cpl, K)
#undef K
      }
      DLP_CHECK_LAUNCH();
    }
    rounds += rounds_per_launch;
    DLP_CHECK_HIP(hipEventRecord(h->ev1, st));
    hipLaunchKernelGGL(k_push_probe, dim3(1), dim3(64), 0, st, h->S.done, h->S.edges, h->stat_dev);
    DLP_CHECK_LAUNCH();
    DLP_CHECK_HIP(hipMemcpyAsync(h->host, h->stat_dev, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    DLP_CHECK_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->ev0, h->ev1) == hipSuccess) h->push_ms += ms;
    h->push_launched += rounds_per_launch;
    if (((volatile int64_t*)h->host)[0]) {
      converged = 1;
      break;
    }
    if (budget_edges > 0 && ((volatile int64_t*)h->host)[1] > budget_edges) break;
  }
  if (rounds_out) *rounds_out = rounds;
  if (converged_out) *converged_out = converged;
  return DLP_OK;
}

// ---------------------------------------------------------------------------------------
// One push launch across graphs.  A plan holds the graphs' push descriptors (re-uploaded at
// every call, so a regrown store needs no new plan) grouped by column-chunk class, and a
// captured hipGraph of rounds in which each round is one launch per class covering every
// graph of the class (grid x = blocks per graph, y = graph).

namespace {
struct PushPlan {
  std::vector<LP*> hs;
  struct Class {
    int cpl = 0;
    std::vector<int> members;   // plan indices
    PushState* states = nullptr;
    int32_t* gshift = nullptr;
    int32_t* lcshift = nullptr;
    int32_t* active = nullptr;
  };
  std::vector<Class> classes;
  PushState* all = nullptr;     // every graph in plan order (probe)
  int64_t* probe = nullptr;
  int64_t* probe_host = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_rounds = 0;
  unsigned gx = 512;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

int plan_free(PushPlan* p) {
  for (auto& c : p->classes) {
    if (c.states) (void)hipFree(c.states);
    if (c.gshift) (void)hipFree(c.gshift);
    if (c.lcshift) (void)hipFree(c.lcshift);
    if (c.active) (void)hipFree(c.active);
  }
  if (p->all) (void)hipFree(p->all);
  if (p->probe) (void)hipFree(p->probe);
  if (p->probe_host) (void)hipHostFree(p->probe_host);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->ev0) (void)hipEventDestroy(p->ev0);
  if (p->ev1) (void)hipEventDestroy(p->ev1);
  delete p;
  return DLP_OK;
}

void launch_multi_round(PushPlan* p, int parity, hipStream_t st) {
  for (auto& c : p->classes) {
    const dim3 grid(p->gx, (unsigned)c.members.size());
#define K(CPLV)                                                                                          \
  hipLaunchKernelGGL((k_push_round_multi<CPLV>), grid, dim3(kBlock), 0, st, c.states, c.gshift, c.lcshift, \
                     c.active, parity);
    DLP_DISPATCH_CPL(c.cpl, K)
#undef K
  }
}
}  // namespace

extern "C" int dlp_push_plan_create(void* const* handles, int nh, void** plan) {
  DLP_CHECK_ARG(handles && plan && nh >= 1 && nh <= 4096, "push_plan_create: bad argument");
  PushPlan* p = new PushPlan();
  for (int i = 0; i < nh; ++i) {
    if (!handles[i]) {
      delete p;
      DLP_CHECK_ARG(false, "push_plan_create: null handle %d", i);
    }
    p->hs.push_back((LP*)handles[i]);
  }
  for (int i = 0; i < nh; ++i) {
    const int cpl = p->hs[i]->map.cpl;
    PushPlan::Class* c = nullptr;
    for (auto& k : p->classes)
      if (k.cpl == cpl) c = &k;
    if (!c) {
      p->classes.emplace_back();
      c = &p->classes.back();
      c->cpl = cpl;
    }
    c->members.push_back(i);
  }
  bool ok = true;
  for (auto& c : p->classes) {
    const size_t k = c.members.size();
    ok = ok && hipMalloc((void**)&c.states, sizeof(PushState) * k) == hipSuccess &&
         hipMalloc((void**)&c.gshift, sizeof(int32_t) * k) == hipSuccess &&
         hipMalloc((void**)&c.lcshift, sizeof(int32_t) * k) == hipSuccess &&
         hipMalloc((void**)&c.active, sizeof(int32_t) * k) == hipSuccess;
    if (!ok) break;
    std::vector<int32_t> gs(k), ls(k);
    for (size_t j = 0; j < k; ++j) {
      gs[j] = p->hs[c.members[j]]->map.g_shift;
      ls[j] = p->hs[c.members[j]]->map.lc_shift;
    }
    ok = hipMemcpy(c.gshift, gs.data(), sizeof(int32_t) * k, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(c.lcshift, ls.data(), sizeof(int32_t) * k, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemset(c.active, 0, sizeof(int32_t) * k) == hipSuccess;
    if (!ok) break;
  }
  ok = ok && hipMalloc((void**)&p->all, sizeof(PushState) * nh) == hipSuccess &&
       hipMalloc((void**)&p->probe, sizeof(int64_t) * 2 * nh) == hipSuccess &&
       hipHostMalloc((void**)&p->probe_host, sizeof(int64_t) * 2 * nh) == hipSuccess &&
       hipEventCreate(&p->ev0) == hipSuccess && hipEventCreate(&p->ev1) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    plan_free(p);
    set_error("push_plan_create: device allocation failed");
    return DLP_ERR_HIP;
  }
  *plan = p;
  return DLP_OK;
}

extern "C" int dlp_push_plan_destroy(void* plan) {
  if (!plan) return DLP_OK;
  (void)hipDeviceSynchronize();
  return plan_free((PushPlan*)plan);
}

// Push rounds for the graphs with active[i] != 0 (each from its last refresh's violators), one
// launch per round and column class for all of them, until each converges, exceeds its edge
// budget or max_rounds pass.  Synchronises once per launch of rounds_per_launch rounds.
extern "C" int dlp_lp_push_multi(void* plan, const int32_t* active, int max_rounds, int rounds_per_launch,
                                 int use_graph, const int64_t* budget_edges, int64_t* rounds_out,
                                 int32_t* converged_out, void* stream) {
  DLP_CHECK_ARG(plan && active && max_rounds >= 0 && rounds_per_launch >= 2 && (rounds_per_launch % 2) == 0,
                "lp_push_multi: bad argument (rounds_per_launch must be even)");
  PushPlan* p = (PushPlan*)plan;
  hipStream_t st = (hipStream_t)stream;
  const int nh = (int)p->hs.size();
  std::vector<int32_t> on(active, active + nh);
  // graphs outside this call are not touched at all (another thread may be using them)
  std::vector<PushState> all(nh);
  for (int i = 0; i < nh; ++i) {
    if (on[i]) all[i] = p->hs[i]->S;
    else std::memset(&all[i], 0, sizeof(PushState));
    if (rounds_out) rounds_out[i] = 0;
    if (converged_out) converged_out[i] = on[i] ? 0 : 1;
  }
  DLP_CHECK_HIP(hipMemcpyAsync(p->all, all.data(), sizeof(PushState) * nh, hipMemcpyHostToDevice, st));
  auto upload_active = [&]() -> int {
    for (auto& c : p->classes) {
      std::vector<PushState> cs(c.members.size());
      std::vector<int32_t> ca(c.members.size());
      for (size_t j = 0; j < c.members.size(); ++j) {
        cs[j] = all[c.members[j]];
        ca[j] = on[c.members[j]] ? 1 : 0;
      }
      DLP_CHECK_HIP(hipMemcpyAsync(c.states, cs.data(), sizeof(PushState) * cs.size(), hipMemcpyHostToDevice, st));
      DLP_CHECK_HIP(hipMemcpyAsync(c.active, ca.data(), sizeof(int32_t) * ca.size(), hipMemcpyHostToDevice, st));
      // the host vectors go out of scope: wait for the copies
      DLP_CHECK_HIP(hipStreamSynchronize(st));
    }
    return DLP_OK;
  };
  int left = 0;
  for (int i = 0; i < nh; ++i) {
    if (!on[i]) continue;
    LP* h = p->hs[i];
    hipLaunchKernelGGL(k_push_begin, dim3(256), dim3(kBlock), 0, st, h->S, h->viol, h->cnt + CNT_VIOL);
    ++left;
  }
  DLP_CHECK_LAUNCH();
  int rc = upload_active();
  if (rc) return rc;
  if (use_graph && (!p->exec || p->graph_rounds != rounds_per_launch)) {
    if (p->exec) {
      (void)hipGraphExecDestroy(p->exec);
      p->exec = nullptr;
    }
    hipStream_t cs;
    DLP_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t graph;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      for (int r = 0; r < rounds_per_launch; ++r) launch_multi_round(p, r & 1, cs);
      e = hipStreamEndCapture(cs, &graph);
    }
    if (e == hipSuccess) {
      e = hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
    }
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) {
      p->exec = nullptr;
      set_error("multi push graph capture failed: %s", hipGetErrorString(e));
      return DLP_ERR_HIP;
    }
    p->graph_rounds = rounds_per_launch;
  }
  int64_t rounds = 0;
  while (left > 0 && rounds < max_rounds) {
    DLP_CHECK_HIP(hipEventRecord(p->ev0, st));
    if (use_graph) {
      DLP_CHECK_HIP(hipGraphLaunch(p->exec, st));
    } else {
      for (int r = 0; r < rounds_per_launch; ++r) launch_multi_round(p, r & 1, st);
      DLP_CHECK_LAUNCH();
    }
    rounds += rounds_per_launch;
    DLP_CHECK_HIP(hipEventRecord(p->ev1, st));
    hipLaunchKernelGGL(k_push_probe_multi, dim3(grid_for(nh)), dim3(kBlock), 0, st, p->all, nh, p->probe);
    DLP_CHECK_LAUNCH();
    DLP_CHECK_HIP(hipMemcpyAsync(p->probe_host, p->probe, sizeof(int64_t) * 2 * nh, hipMemcpyDeviceToHost, st));
    DLP_CHECK_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, p->ev0, p->ev1);
    bool changed = false;
    const volatile int64_t* hp = p->probe_host;
    for (int i = 0; i < nh; ++i) {
      if (!on[i]) continue;
      LP* h = p->hs[i];
      h->push_ms += ms;
      h->push_launched += rounds_per_launch;
      if (rounds_out) rounds_out[i] = rounds;
      if (hp[2 * i]) {
        if (converged_out) converged_out[i] = 1;
        on[i] = 0;
      } else if (budget_edges && budget_edges[i] > 0 && hp[2 * i + 1] > budget_edges[i]) {
        on[i] = 0;
      } else {
        continue;
      }
      changed = true;
      --left;
    }
    if (changed && left > 0) {
      rc = upload_active();
      if (rc) return rc;
    }
  }
  return DLP_OK;
}

// Host-driven FP64 round (plain path): candidates cand[0..nc) in device memory.
extern "C" int dlp_lp_push64_round(void* handle, int nc, const int32_t* cand, void* stream) {
  DLP_CHECK_ARG(handle && nc >= 0, "lp_push64_round: bad argument");
  LP* h = (LP*)handle;
  DLP_CHECK_ARG(h->d.Z64 && h->d.R64 && h->dsnap, "lp_push64_round: FP64 state not attached");
  if (nc == 0) return DLP_OK;
  hipStream_t st = (hipStream_t)stream;
  const int C1 = h->d.C1;
  hipLaunchKernelGGL(k_take64, dim3(grid_for((long long)nc * C1)), dim3(kBlock), 0, st, nc, cand, C1,
                     h->d.ld, h->d.diag64, h->d.Z64, h->d.R64, h->dsnap);
  DLP_CHECK_LAUNCH();
  RowMap m = choose_map(C1, 1.0);
  m.g_shift = m.lc_shift < 2 ? 2 : m.lc_shift;
  const unsigned grid = grid_for((long long)nc << m.g_shift);
#define K(CPLV)                                                                                      \
  hipLaunchKernelGGL((k_push64_round<CPLV>), dim3(grid), dim3(kBlock), 0, st, nc, cand, C1, h->d.ld,    \
                     m.g_shift, m.lc_shift, h->d.indptr, h->d.rowlen, h->d.col, h->d.w, h->d.solved,   \
                     h->d.diag64, h->d.Z64, h->d.R64, h->dsnap, h->mark, h->S.edges);
  DLP_DISPATCH_CPL(m.cpl, K)
#undef K
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// Marked vertices (in order) that violate in FP64 become the next candidates; marks cleared.
__global__ void k_filter64(const int32_t* __restrict__ idx, const int32_t* __restrict__ nidx, int C1, int64_t ld,
                           const double* __restrict__ R, const double* __restrict__ diag,
                           const double* __restrict__ eps, uint8_t* __restrict__ mark, int32_t* __restrict__ out,
                           int32_t* __restrict__ nout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *nidx) return;
  const int v = idx[i];
  mark[v] = 0;
  const double d = diag[v];
  bool viol = false;
  for (int c = 0; c < C1; ++c) viol |= !(fabs(R[(int64_t)v * ld + c]) <= eps[c] * d);
  wave_append(viol, v, out, nout);
}

extern "C" int dlp_lp_push64_next(void* handle, int32_t* cand_out, int32_t* count_out_dev, void* stream) {
  DLP_CHECK_ARG(handle && cand_out && count_out_dev, "lp_push64_next: bad argument");
  LP* h = (LP*)handle;
  hipStream_t st = (hipStream_t)stream;
  const int n = h->d.n;
  int rc = compact_u8(n, h->mark, h->cand64, h->cnt + CNT_CAND64, h->scan_ws, h->scan_bytes, st);
  if (rc) return rc;
  DLP_CHECK_HIP(hipMemsetAsync(count_out_dev, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_filter64, dim3(grid_for(n)), dim3(kBlock), 0, st, h->cand64, h->cnt + CNT_CAND64,
                     h->d.C1, h->d.ld, h->d.R64, h->d.diag64, h->eps64, h->mark, cand_out, count_out_dev);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_lp_violators(void* handle, int32_t** list, int32_t** count) {
  DLP_CHECK_ARG(handle && list && count, "lp_violators: null argument");
  LP* h = (LP*)handle;
  *list = h->viol;
  *count = h->cnt + CNT_VIOL;
  return DLP_OK;
}

extern "C" int dlp_lp_status(void* handle, int64_t* out, double* rho, void* stream) {
  DLP_CHECK_ARG(handle && out, "lp_status: null argument");
  LP* h = (LP*)handle;
  hipStream_t st = (hipStream_t)stream;
  const int C1 = h->d.C1;
  hipLaunchKernelGGL(k_status_gather, dim3(1), dim3(kBlock), 0, st, h->cnt, h->S.bcount, h->S.rnd, h->S.edges,
                     h->S.stats, h->mass, rho ? h->rho : nullptr, C1, h->stat_dev);
  DLP_CHECK_LAUNCH();
  const size_t words = rho ? (size_t)kStatRho + C1 : (size_t)CNT_N + 12;
  DLP_CHECK_HIP(hipMemcpyAsync(h->host, h->stat_dev, words * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  DLP_CHECK_HIP(hipStreamSynchronize(st));
  const volatile int64_t* hs = h->host;
  for (int i = 0; i < CNT_N + 7; ++i) out[i] = hs[i];
  out[CNT_N + 9] = hs[CNT_N + 9];
  out[CNT_N + 10] = hs[CNT_N + 10];
  out[CNT_N + 11] = hs[CNT_N + 11];
  if (rho) std::memcpy(rho, (const void*)(h->host + kStatRho), sizeof(double) * C1);
  out[CNT_N + 7] = h->push_launched;
  out[CNT_N + 8] = (int64_t)(h->push_ms * 1000.0);  // microseconds
  out[CNT_N + 12] = h->last_refresh_kernel;
  return DLP_OK;
}

// ---------------------------------------------------------------------------------------
// Global solve: column-wise PCG with the Jacobi preconditioner on the FP32 state, from the
// kept residual R32 (solvers.py:66-97, the DynLP+ global path solvers.py:268-288).

// 16-bit units per direction slot: n rows of ldh, rounded up to 256 bytes, so that every slot
// starts as aligned as the block (FP8 rows of ld bytes alone leave n*ld a multiple of 4 only).
static int64_t slot_units(int64_t n, int64_t ldh) {
  return (n * ldh + 127) / 128 * 128;
}

static PcgBufs pcg_bufs(LP* h) {
  PcgBufs b{};
  b.n = h->d.n;
  b.C = h->d.C1;
  b.ld = h->d.ld;
  b.Z = h->d.Z32;
  b.R = h->d.R32;
  b.P = h->P;
  b.AP = h->AP;
  b.diag = h->d.diag32;
  b.eps = h->S.eps;
  const int C1 = h->d.C1;
  b.rz = h->pcg_scal;
  b.rz_new = h->pcg_scal + C1;
  b.pAp = h->pcg_scal + 2 * C1;
  b.alpha = h->pcg_scal + 3 * C1;
  b.beta = h->pcg_scal + 4 * C1;
  b.viol = h->pcg_viol;
  b.ctl = h->pcg_viol ? reinterpret_cast<int32_t*>(h->pcg_viol) + 2 : nullptr;
  b.P16 = h->pcg_p16 ? h->P16 : nullptr;
  b.ldh = h->ldh;
  b.rp = h->pcg_scal + 5 * C1;
  b.p16 = h->pcg_p16;
  b.sig = reinterpret_cast<float*>(h->pcg_scal + 6 * C1);
  b.rmax = reinterpret_cast<float*>(h->pcg_scal + 7 * C1);
  b.pmax = reinterpret_cast<float*>(h->pcg_scal + 8 * C1);
  b.emax = reinterpret_cast<float*>(h->pcg_scal + 9 * C1);
  if (b.P16 && h->pcg_group > 0) {
    b.slot_stride = slot_units(h->d.n, h->ldh);
    b.group = h->pcg_group;
    b.alpha_hist = h->pcg_ahist;
  }
  if (h->cg_na && h->cg_age >= 0) {
    b.na = h->cg_na;
    b.min_agg = h->cg_min;
    b.bounds = h->cg_bounds;
    b.aggmap = h->cg_aggmap;
    b.ldc = h->d.ld;
    b.rc = h->cg_rc;
    b.ec = h->cg_ec;
    b.Ainv = h->cg_Ainv;
    b.solved = h->d.solved;
  }
  return b;
}

static int pcg_iteration(LP* h, const PcgBufs& b, hipStream_t st) {
  // AP = diag*P - W_UU P, with p^T A p accumulated in the same pass when the layout allows
  if (b.P16 || ((b.ld % 4) == 0 && tuning().vec4 && tuning().fuse_dot)) {
    const uint16_t* P16 = b.P16 && b.slot_stride > 0 ? b.P16 + b.slot * b.slot_stride : b.P16;
    int rc = pcg_operator(b.n, b.C, h->d.indptr, h->d.rowlen, h->d.col, h->d.w, h->d.solved, b.P, b.ld, P16,
                          b.ldh, b.p16, b.diag, -1.f, b.AP, b.ld, b.pAp, st, b.ctl);
    if (rc) return rc;
    return pcg_finish_iteration(b, st, true);
  }
  int rc = dlp_row_operator(b.n, nullptr, b.C, h->d.indptr, h->d.rowlen, h->d.col, h->d.w, DLP_F32,
                            h->d.solved, b.P, b.ld, DLP_F32, b.diag, -1.0, nullptr, 0, b.AP, b.ld, DLP_F32,
                            0, 20.0, st);
  if (rc) return rc;
  return pcg_finish_iteration(b, st);
}

// The coarse operator of the two-level preconditioner, rebuilt from the current graph and
// equations every coarse_refresh global solves, or before the next one once a solve needed a
// quarter more iterations than the fewest since the build (a stale Ac still gives an SPD
// preconditioner, only a weaker one: the equations of a few percent of the rows change between
// rebuilds, more on a graph that fragments).
// Aggregates are row ranges: coarse_na targets spaced S0 rows apart, each moved to the row
// boundary with the fewest crossing edges within half a spacing; a moved boundary is kept when it
// cuts fewer edges than the rows half a spacing to either side of it (on a graph whose ids follow
// its communities these are the community edges, where the coarse space works best; a target
// inside a community is dropped), all of them when under a quarter qualify (no such structure);
// ranges hold at least the rows one workgroup of the vector passes covers.  A new
// aggregate count drops the captured iterations, which hold it as a launch argument.
static int coarse_alloc(LP* h, hipStream_t st);

static int coarse_prepare(LP* h, hipStream_t st) {
  const int n = h->d.n;
  int chunk = 0;
  if (h->coarse_na > 0 && n >= 64) {
    PcgBufs probe = pcg_bufs(h);
    probe.na = 0;
    chunk = pcg_flat_chunk(probe);
  }
  if (chunk <= 0) {
    if (h->cg_na) {
      if (h->pcg_exec) {
        (void)hipGraphExecDestroy(h->pcg_exec);
        h->pcg_exec = nullptr;
      }
      h->cg_na = 0;
    }
    h->cg_age = -1;
    return DLP_OK;
  }
  {
    const int rc = coarse_alloc(h, st);
    if (rc) return rc;
  }
  if (h->cg_age >= 0 && h->cg_age < h->coarse_refresh && !h->cg_stale) {
    h->cg_age += 1;
    return DLP_OK;
  }
  h->cg_stale = 0;
  const int min_agg = chunk > 32 ? chunk : 32;
  if (h->cg_age >= 0 && h->cg_na > 0 && min_agg == h->cg_min && (h->cg_builds % 4) != 0) {
    // a batch moves few edges across the row ranges: keep them, rebuild A_c from the current
    // equations (the boundary search, a pass with an atomic per edge, runs every 4th build)
    int rc = coarse_build(n, h->cg_na, h->cg_bounds, h->cg_aggmap, h->d.indptr, h->d.rowlen, h->d.col, h->d.w,
                          h->d.solved, h->d.diag64, h->cg_Ac, h->cg_Ainv, st);
    if (rc) return rc;
    h->cg_builds += 1;
    h->cg_age = 1;
    return DLP_OK;
  }
  // boundaries: targets k*S0, each moved to the lowest-cut row boundary near it
  const int want = h->coarse_na < kCoarseMax ? h->coarse_na : kCoarseMax;
  int S0 = (n + want - 1) / want;
  if (S0 < min_agg) S0 = min_agg;
  const int nk = (n - 1) / S0;   // targets k*S0 < n, k >= 1
  std::vector<int32_t> bounds(1, 0);
  if (nk >= 1) {
    int rc = coarse_boundaries(n, nk, S0, min_agg, h->d.indptr, h->d.rowlen, h->d.col, h->cg_diff,
                               h->cg_diff + n, h->cg_ws, h->cg_ws_bytes, h->cg_cand, st);
    if (rc) return rc;
    std::vector<int32_t> c4(4 * (size_t)nk);
    DLP_CHECK_HIP(hipMemcpyAsync(c4.data(), h->cg_cand, sizeof(int32_t) * c4.size(), hipMemcpyDeviceToHost, st));
    DLP_CHECK_HIP(hipStreamSynchronize(st));
    // keep the boundaries that cut fewer edges than the rows half a spacing to either side (a
    // community edge, not a slope the window ran down); without such structure keep them all
    std::vector<int32_t> dip, all;
    for (int k = 0; k < nk; ++k) {
      const int32_t* o = &c4[4 * (size_t)k];
      if (o[0] < 0) continue;
      all.push_back(o[0]);
      if (o[1] < o[2] && o[1] < o[3]) dip.push_back(o[0]);
    }
    std::vector<int32_t>& cand = (int)dip.size() * 4 >= nk ? dip : all;
    std::sort(cand.begin(), cand.end());
    for (const int32_t b : cand) {
      if (b - bounds.back() >= min_agg && n - b >= min_agg && (int)bounds.size() < kCoarseMax) bounds.push_back(b);
    }
  }
  bounds.push_back(n);
  const int na = (int)bounds.size() - 1;
  DLP_CHECK_HIP(hipMemcpyAsync(h->cg_bounds, bounds.data(), sizeof(int32_t) * bounds.size(), hipMemcpyHostToDevice,
                               st));
  int rc = coarse_aggmap(n, na, h->cg_bounds, h->cg_aggmap, st);
  if (rc) return rc;
  rc = coarse_build(n, na, h->cg_bounds, h->cg_aggmap, h->d.indptr, h->d.rowlen, h->d.col, h->d.w, h->d.solved,
                    h->d.diag64, h->cg_Ac, h->cg_Ainv, st);
  if (rc) return rc;
  // the host vector goes out of scope: finish the copy first
  DLP_CHECK_HIP(hipStreamSynchronize(st));
  if (na != h->cg_na || h->cg_age < 0 || min_agg != h->cg_min) {
    if (h->pcg_exec) {
      (void)hipGraphExecDestroy(h->pcg_exec);
      h->pcg_exec = nullptr;
    }
  }
  h->cg_na = na;
  h->cg_min = min_agg;
  h->cg_age = 1;
  h->cg_builds += 1;
  return DLP_OK;
}

// One group of iterations: iteration i gathers direction slot i (deferred updates) and the
// group ends with the deferred solution update.
static int pcg_group_launch(LP* h, const PcgBufs& b, int iters, hipStream_t st) {
  int rc = DLP_OK;
  for (int i = 0; i < iters && rc == DLP_OK; ++i) {
    PcgBufs bi = b;
    bi.slot = i;
    rc = pcg_iteration(h, bi, st);
  }
  if (rc == DLP_OK) rc = pcg_flush(b, st);
  return rc;
}

static int capture_pcg(LP* h, int iters) {
  if (h->pcg_exec && h->pcg_graph_iters == iters) return DLP_OK;
  if (h->pcg_exec) {
    (void)hipGraphExecDestroy(h->pcg_exec);
    h->pcg_exec = nullptr;
  }
  hipStream_t cs;
  DLP_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  const PcgBufs b = pcg_bufs(h);
  hipGraph_t graph;
  int rc = DLP_OK;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = pcg_group_launch(h, b, iters, cs);
    e = hipStreamEndCapture(cs, &graph);
  }
  if (e == hipSuccess && rc == DLP_OK) {
    e = hipGraphInstantiate(&h->pcg_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
  }
  (void)hipStreamDestroy(cs);
  if (rc != DLP_OK) {   // a launch refused during capture: its own message stands
    h->pcg_exec = nullptr;
    return rc;
  }
  if (e != hipSuccess) {
    set_error("pcg graph capture failed: %s", hipGetErrorString(e));
    h->pcg_exec = nullptr;
    return DLP_ERR_HIP;
  }
  h->pcg_graph_iters = iters;
  return DLP_OK;
}

// Iterations per deferred solution update for captured groups of iters_per_launch (0: the
// solution is updated by every iteration).
static int pcg_group_for(const LP* h, int iters_per_launch) {
  return (h->pcg_p16 && tuning().defer_z && (h->d.ld % 4) == 0 && iters_per_launch <= kMaxGroup)
             ? iters_per_launch : 0;
}

// Slots of C1 doubles in pcg_scal.
constexpr int kScalSlots = 10;

// Row length (16-bit units) of the low-precision direction block for leading dimension ld: BF16
// and FP16 rows of 8-value groups covering ld columns, FP8 rows of exactly ld bytes (the flat
// passes address column quad q < ld/4 of every row).
static int64_t direction_ldh(int kind, int64_t ld) {
  return kind == 3 ? ld / 2 : (ld + 7) / 8 * 8;
}

// Buffers of the coarse correction (sized for kCoarseMax aggregates), whole or not at all.
static int coarse_alloc(LP* h, hipStream_t st) {
  if (h->cg_Ac) return DLP_OK;
  const int n = h->d.n;
  const size_t ldc = (size_t)h->d.ld;
  const size_t ws = scan_workspace_bytes(n);
  AllocGroup g;
  g.add((void**)&h->cg_Ac, sizeof(double) * kCoarseMax * kCoarseMax);
  g.add((void**)&h->cg_Ainv, sizeof(double) * kCoarseMax * kCoarseMax);
  g.add((void**)&h->cg_rc, sizeof(double) * kCoarseMax * ldc);
  g.add((void**)&h->cg_ec, sizeof(float) * kCoarseMax * ldc);
  g.add((void**)&h->cg_bounds, sizeof(int32_t) * (kCoarseMax + 1));
  g.add((void**)&h->cg_aggmap, (size_t)n);
  g.add((void**)&h->cg_diff, sizeof(int32_t) * 2 * (size_t)n);
  g.add((void**)&h->cg_cand, sizeof(int32_t) * 4 * kCoarseMax);
  g.add(&h->cg_ws, ws);
  int rc = g.commit(&h->dev_bytes, "coarse operator");
  if (rc) return rc;
  h->cg_ws_bytes = ws;
  DLP_CHECK_HIP(hipMemsetAsync(h->cg_rc, 0, sizeof(double) * kCoarseMax * ldc, st));
  DLP_CHECK_HIP(hipMemsetAsync(h->cg_ec, 0, sizeof(float) * kCoarseMax * ldc, st));
  return DLP_OK;
}

// Every buffer a global solve with captured groups of iters_per_launch iterations uses (the
// block PCG's k x k matrices when block), so that a caller can take them at admission; the
// solves call it too and then allocate nothing new.  Each set of buffers is allocated whole or
// not at all (a failed call leaves the handle as it was).
static int pcg_reserve(LP* h, int iters_per_launch, bool block, hipStream_t st) {
  const int C1 = h->d.C1;
  const size_t nz = (size_t)h->d.n * h->d.ld;
  if (!h->P) {
    AllocGroup g;
    g.add((void**)&h->P, nz * sizeof(float));
    g.add((void**)&h->AP, nz * sizeof(float));
    g.add((void**)&h->pcg_scal, kScalSlots * C1 * sizeof(double));
    g.add((void**)&h->pcg_viol, 4 * sizeof(float));   // viol, last viol, done, iterations
    int rc = g.commit(&h->dev_bytes, "global solve");
    if (rc) return rc;
    DLP_CHECK_HIP(hipMemsetAsync(h->P, 0, nz * sizeof(float), st));
    DLP_CHECK_HIP(hipMemsetAsync(h->AP, 0, nz * sizeof(float), st));
  }
  if (block) {   // FP32 directions, Jacobi preconditioner, and the k x k matrices
    if (!h->bk) {
      AllocGroup g;
      g.add((void**)&h->bk, 5 * C1 * C1 * sizeof(double));
      return g.commit(&h->dev_bytes, "block PCG");
    }
    return DLP_OK;
  }
  // deferred solution updates keep a group's directions: iters_per_launch + 1 slots
  const int group = pcg_group_for(h, iters_per_launch);
  const int slots = group > 0 ? group + 1 : 1;
  const int64_t ldh = direction_ldh(h->pcg_p16, h->d.ld);
  if (h->pcg_p16 && (!h->P16 || h->p16_slots < slots || h->ldh != ldh)) {
    const size_t nh = (size_t)slot_units(h->d.n, ldh) * slots;
    uint16_t* fresh = nullptr;
    AllocGroup g;
    g.add((void**)&fresh, nh * sizeof(uint16_t));
    int rc = g.commit(&h->dev_bytes, "global solve directions");
    if (rc) return rc;
    if (h->P16) {
      (void)hipFree(h->P16);
      h->dev_bytes -= (int64_t)((size_t)slot_units(h->d.n, h->ldh) * h->p16_slots * sizeof(uint16_t));
    }
    h->P16 = fresh;
    h->ldh = ldh;
    h->p16_slots = slots;
    DLP_CHECK_HIP(hipMemsetAsync(h->P16, 0, nh * sizeof(uint16_t), st));
    if (h->pcg_exec) {   // captured iterations address the old block
      (void)hipGraphExecDestroy(h->pcg_exec);
      h->pcg_exec = nullptr;
    }
  }
  if (group > 0 && h->ahist_len < group * C1 + 1) {
    const int len = (group * C1 + 2) & ~1;
    float* fresh = nullptr;
    AllocGroup g;
    g.add((void**)&fresh, len * sizeof(float));
    int rc = g.commit(&h->dev_bytes, "global solve step lengths");
    if (rc) return rc;
    if (h->pcg_ahist) {
      (void)hipFree(h->pcg_ahist);
      h->dev_bytes -= (int64_t)h->ahist_len * (int64_t)sizeof(float);
      if (h->pcg_exec) {   // captured iterations address the old history
        (void)hipGraphExecDestroy(h->pcg_exec);
        h->pcg_exec = nullptr;
      }
    }
    h->pcg_ahist = fresh;
    h->ahist_len = len;
  }
  if (h->coarse_na > 0 && h->d.n >= 64) return coarse_alloc(h, st);
  return DLP_OK;
}

// Runs PCG iterations (in captured groups of iters_per_launch when use_graph) until the FP32
// recurrence residual meets the push thresholds (max |R|/(eps*diag) <= 1), the violation
// stalls for `patience` launches, or max_iter.  *iters_out: iterations run; *viol_out: last
// violation ratio.
// The column PCG; with verify, the FP64 verification pass (refresh mode 3) is enqueued behind
// every group of iterations, before the host reads the group's status, so a solve that ends
// after a group has its certificate without another host round trip (a group that does not
// end the solve costs one extra pass, which also replaces the kept residual by the FP64 one).
static int lp_pcg(LP* h, int max_iter, int iters_per_launch, int use_graph, int patience, int64_t* iters_out,
                  double* viol_out, hipStream_t st, bool verify) {
  int rc = pcg_reserve(h, iters_per_launch, false, st);
  if (rc) return rc;
  const int group = pcg_group_for(h, iters_per_launch);
  if (group != h->pcg_group && h->pcg_exec) {   // captured with the other update scheme
    (void)hipGraphExecDestroy(h->pcg_exec);
    h->pcg_exec = nullptr;
  }
  h->pcg_group = group;
  rc = coarse_prepare(h, st);
  if (rc) return rc;
  const PcgBufs b = pcg_bufs(h);
  rc = pcg_start(b, st);
  if (rc) return rc;
  if (use_graph) {
    rc = capture_pcg(h, iters_per_launch);
    if (rc) return rc;
  }
  // Iterations after the one that meets the thresholds return at once (device flag), so a
  // group may be longer than the solve needs; the device counts the iterations that ran.
  int64_t launched = 0, it = 0;
  float v = 0.f, best = 3.4e38f;
  int stall = 0;
  while (launched < max_iter) {
    DLP_CHECK_HIP(hipEventRecord(h->ev0, st));
    if (use_graph) {
      DLP_CHECK_HIP(hipGraphLaunch(h->pcg_exec, st));
    } else {
      rc = pcg_group_launch(h, b, iters_per_launch, st);
      if (rc) return rc;
    }
    launched += iters_per_launch;
    DLP_CHECK_HIP(hipEventRecord(h->ev1, st));
    if (verify) {
      rc = dlp_lp_refresh(h, 3, 0, st);
      if (rc) return rc;
    }
    DLP_CHECK_HIP(hipMemcpyAsync(h->host, h->pcg_viol, 4 * sizeof(float), hipMemcpyDeviceToHost, st));
    DLP_CHECK_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->ev0, h->ev1) == hipSuccess) h->pcg_ms += ms;
    const volatile float* hv = reinterpret_cast<volatile float*>(h->host);
    const volatile int32_t* hc = reinterpret_cast<volatile int32_t*>(h->host) + 2;
    v = hv[1];   // the last completed iteration's violation
    const int done = hc[0];
    const int64_t dev_it = hc[1];
    h->pcg_iters += dev_it - it;
    it = dev_it;
    if (done || v <= 1.f) break;
    if (v < 0.9f * best) {
      best = v;
      stall = 0;
    } else if (++stall >= patience) {
      break;
    }
  }
  // a coarse operator that has gone stale (on a fragmenting graph the equations drift from the
  // ones it was built from) shows as a solve needing a quarter more iterations (and 2) than the
  // fewest any solve needed since its build: rebuild it before the next solve instead of
  // waiting for its age
  if (h->cg_na && (h->cg_age == 1 || it < h->cg_base_iters)) {
    h->cg_base_iters = it;
  } else if (h->cg_na && h->cg_age > 1 && h->cg_base_iters > 0 && it >= h->cg_base_iters * 5 / 4 + 2) {
    h->cg_stale = 1;
  }
  if (iters_out) *iters_out = it;
  if (viol_out) *viol_out = v;
  return DLP_OK;
}

extern "C" int dlp_lp_pcg(void* handle, int max_iter, int iters_per_launch, int use_graph, int patience,
                          int64_t* iters_out, double* viol_out, void* stream) {
  DLP_CHECK_ARG(handle && max_iter >= 0 && iters_per_launch >= 1, "lp_pcg: bad argument");
  return lp_pcg((LP*)handle, max_iter, iters_per_launch, use_graph, patience, iters_out, viol_out,
                (hipStream_t)stream, false);
}

extern "C" int dlp_lp_pcg_verify(void* handle, int max_iter, int iters_per_launch, int use_graph, int patience,
                                 int64_t* iters_out, double* viol_out, void* stream) {
  DLP_CHECK_ARG(handle && max_iter >= 0 && iters_per_launch >= 1, "lp_pcg_verify: bad argument");
  return lp_pcg((LP*)handle, max_iter, iters_per_launch, use_graph, patience, iters_out, viol_out,
                (hipStream_t)stream, true);
}

// Reported scores [F, h] in FP64: solved rows as held, rows outside the solved set as F = 0,
// h = 1 (whatever their parked state).
template <typename TZ>
__global__ void k_scores(int n, int C1, const TZ* __restrict__ Z, int64_t ld, const uint8_t* __restrict__ solved,
                         double* __restrict__ out, int64_t ldo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * C1) return;
  const int r = (int)(i / C1), c = (int)(i - (int64_t)r * C1);
  out[(int64_t)r * ldo + c] = solved[r] ? (double)Z[(int64_t)r * ld + c] : (c == C1 - 1 ? 1.0 : 0.0);
}

extern "C" int dlp_lp_scores(void* handle, double* out, int64_t ldo, void* stream) {
  DLP_CHECK_ARG(handle && out, "lp_scores: null argument");
  LP* h = (LP*)handle;
  DLP_CHECK_ARG(ldo >= h->d.C1, "lp_scores: ldo < C1");
  hipStream_t st = (hipStream_t)stream;
  const long long total = (long long)h->d.n * h->d.C1;
  if (h->d.Z64)
    hipLaunchKernelGGL(k_scores<double>, dim3(grid_for(total)), dim3(kBlock), 0, st, h->d.n, h->d.C1, h->d.Z64,
                       h->d.ld, h->d.solved, out, ldo);
  else
    hipLaunchKernelGGL(k_scores<float>, dim3(grid_for(total)), dim3(kBlock), 0, st, h->d.n, h->d.C1, h->d.Z32,
                       h->d.ld, h->d.solved, out, ldo);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

// [0] aggregates of the current coarse operator (0: none), [1] their minimum rows, [2] global
// solves since its build; bounds_out (kCoarseMax + 1 ints, may be NULL): the row ranges.
extern "C" int dlp_lp_coarse_info(void* handle, int32_t* out, int32_t* bounds_out, void* stream) {
  DLP_CHECK_ARG(handle && out, "lp_coarse_info: null argument");
  LP* h = (LP*)handle;
  out[0] = h->cg_age >= 0 ? h->cg_na : 0;
  out[1] = h->cg_min;
  out[2] = h->cg_age;
  if (bounds_out && out[0] > 0) {
    DLP_CHECK_HIP(hipMemcpyAsync(bounds_out, h->cg_bounds, sizeof(int32_t) * (out[0] + 1), hipMemcpyDeviceToHost,
                                 (hipStream_t)stream));
    DLP_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
  }
  return DLP_OK;
}

extern "C" int dlp_lp_set_coarse(void* handle, int na, int refresh) {
  DLP_CHECK_ARG(handle && na >= 0 && refresh >= 1, "lp_set_coarse: bad argument (na=%d refresh=%d)", na, refresh);
  LP* h = (LP*)handle;
  h->coarse_na = na < kCoarseMax ? na : kCoarseMax;
  h->coarse_refresh = refresh;
  return DLP_OK;
}

extern "C" int dlp_lp_set_pcg_directions(void* handle, int kind) {
  DLP_CHECK_ARG(handle, "lp_set_pcg_directions: null handle");
  DLP_CHECK_ARG(kind >= 0 && kind <= 3, "lp_set_pcg_directions: kind %d (0 FP32, 1 BF16, 2 FP16, 3 FP8)", kind);
  LP* h = (LP*)handle;
  DLP_CHECK_ARG(!kind || ((h->d.ld % 4) == 0 && h->d.C1 <= 256), "lp_set_pcg_directions: needs 16-byte FP32 rows");
  if (h->pcg_p16 != kind && h->pcg_exec) {   // captured iterations use the other layout
    (void)hipGraphExecDestroy(h->pcg_exec);
    h->pcg_exec = nullptr;
  }
  h->pcg_p16 = kind;
  return DLP_OK;
}

extern "C" int dlp_lp_reserve(void* handle, int iters_per_launch, int block, void* stream) {
  DLP_CHECK_ARG(handle && iters_per_launch >= 1, "lp_reserve: bad argument");
  LP* h = (LP*)handle;
  if (block) DLP_CHECK_ARG(h->d.C1 <= 80, "lp_reserve: block PCG takes at most 80 columns (C1=%d)", h->d.C1);
  hipStream_t st = (hipStream_t)stream;
  const int rc = pcg_reserve(h, iters_per_launch, block != 0, st);
  if (rc) return rc;
  DLP_CHECK_HIP(hipStreamSynchronize(st));
  return DLP_OK;
}

extern "C" int dlp_lp_device_bytes(void* handle, int64_t* out) {
  DLP_CHECK_ARG(handle && out, "lp_device_bytes: null argument");
  *out = ((LP*)handle)->dev_bytes;
  return DLP_OK;
}

extern "C" int64_t dlp_lp_handle_bytes(int n, int C1, int64_t ld, int fp64, int dirs, int iters_per_launch,
                                       int coarse, int block) {
  if (n < 1 || C1 < 1 || ld < C1 || iters_per_launch < 1) return -1;
  const int64_t N = n, nz = N * ld;
  int64_t b = 0;
  b += 13 * 4 * N + 2 * N;                                   // int32 lists and stamps, byte masks
  b += 4 * (int64_t)(1 << 20);                               // overflow list
  b += 4 * CNT_ALL + 8 * C1 + 8 * (C1 + 1);                  // counters, eps64, rho
  b += 4 * 3 + 4 * 2 + 4 + 4 + 8 + 8 * 4 + 4 + 4 * C1 + 4 * 8 + 8;   // push state words
  b += (8 + 4) * (int64_t)kPairTableCap;                     // insertion pair table
  if (fp64) b += 8 * nz;                                     // FP64 increments
  b += (int64_t)dlp_scan_workspace_bytes(n) + 8 * kStatWords;
  if (fp64) return b;   // the plain FP64 path repairs by push only: no global-solve buffers
  // global solve
  b += 2 * 4 * nz + kScalSlots * 8 * C1 + 16;
  if (block) return b + 5 * 8 * (int64_t)C1 * C1;
  const bool p16 = dirs != 0;
  const int group =
      (p16 && tuning().defer_z && (ld % 4) == 0 && iters_per_launch <= kMaxGroup) ? iters_per_launch : 0;
  if (p16) b += 2 * slot_units(N, direction_ldh(dirs, ld)) * (group > 0 ? group + 1 : 1);
  if (group > 0) b += 4 * ((group * C1 + 2) & ~1);
  if (coarse > 0 && n >= 64)
    b += 2 * 8 * kCoarseMax * kCoarseMax + (8 + 4) * kCoarseMax * ld + 4 * (kCoarseMax + 1) + N + 8 * N +
         16 * kCoarseMax + (int64_t)scan_workspace_bytes(n);
  return b;
}

// Cumulative timing of the global solve: [0] iterations, [1] microseconds (HIP events).
extern "C" int dlp_lp_pcg_stats(void* handle, int64_t* out) {
  DLP_CHECK_ARG(handle && out, "lp_pcg_stats: null argument");
  LP* h = (LP*)handle;
  out[0] = h->pcg_iters;
  out[1] = (int64_t)(h->pcg_ms * 1000.0);
  return DLP_OK;
}

// Average duration (ms, HIP events on `stream`) of the global solve's operator launch
// AP = diag*P - W_UU P on the live state, over reps launches.
extern "C" int dlp_lp_time_operator(void* handle, int reps, double* ms_out, void* stream) {
  DLP_CHECK_ARG(handle && reps >= 1 && ms_out, "lp_time_operator: bad argument");
  LP* h = (LP*)handle;
  DLP_CHECK_ARG(h->P && h->AP, "lp_time_operator: run dlp_lp_pcg first");
  hipStream_t st = (hipStream_t)stream;
  const PcgBufs b = pcg_bufs(h);
  // the launch the global solve makes: the 16-byte-row operator with the fused p^T A p
  // epilogue when the layout allows it, else the plain row operator
  const bool fused = (b.ld % 4) == 0 && tuning().vec4 && tuning().fuse_dot;
  auto launch = [&]() -> int {
    if (fused || b.P16)
      return pcg_operator(b.n, b.C, h->d.indptr, h->d.rowlen, h->d.col, h->d.w, h->d.solved, b.P, b.ld, b.P16,
                          b.ldh, b.p16, b.diag, -1.f, b.AP, b.ld, b.pAp, st, nullptr);
    return dlp_row_operator(b.n, nullptr, b.C, h->d.indptr, h->d.rowlen, h->d.col, h->d.w, DLP_F32, h->d.solved,
                            b.P, b.ld, DLP_F32, b.diag, -1.0, nullptr, 0, b.AP, b.ld, DLP_F32, 0, 20.0, st);
  };
  int rc = launch();
  if (rc) return rc;
  DLP_CHECK_HIP(hipEventRecord(h->ev0, st));
  for (int i = 0; i < reps; ++i) {
    rc = launch();
    if (rc) return rc;
  }
  DLP_CHECK_HIP(hipEventRecord(h->ev1, st));
  DLP_CHECK_HIP(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  DLP_CHECK_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *ms_out = ms / reps;
  return DLP_OK;
}

// ---------------------------------------------------------------------------------------
// Block PCG (one search space shared by all C1 columns): alpha = (P^T A P)^{-1} (R^T D^{-1} R),
// beta = (R^T D^{-1} R)_old^{-1} (R^T D^{-1} R)_new, with the k x k Grams on the FP64 matrix
// cores and their Cholesky solves in FP64 (gram.hip); vectors FP32.

extern "C" int dlp_gram(int n, int C, const void* X, int64_t ldx, const void* Y, int64_t ldy, const void* rowdiv,
                        double* G, int dtype, void* stream);
extern "C" int dlp_chol_solve(int k, const double* A, const double* B, double* S, double tol, void* stream);
namespace dlp {
int block_update(int n, int C, int mode, float* Z, float* R, float* P, const float* AP, const float* Y, int64_t ld,
                 const double* S, const float* diag, const float* eps, float* viol, hipStream_t st);
int copy_f64(int n, const double* a, double* b, hipStream_t st);
}

static int bpcg_iteration(LP* h, const PcgBufs& b, hipStream_t st) {
  const int C = b.C;
  double* G = h->bk;
  double* PAP = G + C * C;
  double* Gn = PAP + C * C;
  double* alpha = Gn + C * C;
  double* beta = alpha + C * C;
  int rc = dlp_row_operator(b.n, nullptr, C, h->d.indptr, h->d.rowlen, h->d.col, h->d.w, DLP_F32, h->d.solved,
                            b.P, b.ld, DLP_F32, b.diag, -1.0, nullptr, 0, b.AP, b.ld, DLP_F32, 0, 20.0, st);
  if (!rc) rc = dlp_gram(b.n, C, b.P, b.ld, b.AP, b.ld, nullptr, PAP, DLP_F32, st);
  if (!rc) rc = dlp_chol_solve(C, PAP, G, alpha, 1e-8, st);
  if (!rc) rc = block_update(b.n, C, 0, b.Z, b.R, b.P, b.AP, nullptr, b.ld, alpha, b.diag, b.eps, b.viol, st);
  if (!rc) rc = dlp_gram(b.n, C, b.R, b.ld, b.R, b.ld, b.diag, Gn, DLP_F32, st);
  if (!rc) rc = dlp_chol_solve(C, G, Gn, beta, 1e-8, st);
  if (!rc) rc = block_update(b.n, C, 1, b.Z, b.R, b.P, b.AP, nullptr, b.ld, beta, b.diag, b.eps, b.viol, st);
  if (!rc) rc = copy_f64(C * C, Gn, G, st);
  return rc;
}

static int capture_bpcg(LP* h, int iters) {
  if (h->bpcg_exec && h->bpcg_graph_iters == iters) return DLP_OK;
  if (h->bpcg_exec) {
    (void)hipGraphExecDestroy(h->bpcg_exec);
    h->bpcg_exec = nullptr;
  }
  hipStream_t cs;
  DLP_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  const PcgBufs b = pcg_bufs(h);
  hipGraph_t graph;
  int rc = DLP_OK;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (int i = 0; i < iters && rc == DLP_OK; ++i) rc = bpcg_iteration(h, b, cs);
    e = hipStreamEndCapture(cs, &graph);
  }
  if (e == hipSuccess && rc == DLP_OK) {
    e = hipGraphInstantiate(&h->bpcg_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
  }
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess || rc != DLP_OK) {
    set_error("block pcg graph capture failed: %s", hipGetErrorString(e));
    h->bpcg_exec = nullptr;
    return DLP_ERR_HIP;
  }
  h->bpcg_graph_iters = iters;
  return DLP_OK;
}

extern "C" int dlp_lp_bpcg(void* handle, int max_iter, int iters_per_launch, int use_graph, int patience,
                           int64_t* iters_out, double* viol_out, void* stream) {
  DLP_CHECK_ARG(handle && max_iter >= 0 && iters_per_launch >= 1, "lp_bpcg: bad argument");
  LP* h = (LP*)handle;
  DLP_CHECK_ARG(h->d.C1 <= 80, "lp_bpcg: block PCG takes at most 80 columns (C1=%d)", h->d.C1);
  hipStream_t st = (hipStream_t)stream;
  const int C1 = h->d.C1;
  int rc = pcg_reserve(h, iters_per_launch, true, st);
  if (rc) return rc;
  PcgBufs b = pcg_bufs(h);
  b.P16 = nullptr;                                 // block PCG keeps FP32 directions
  b.p16 = 0;
  b.slot_stride = 0;
  b.na = 0;                                        // and the Jacobi preconditioner
  rc = pcg_start(b, st);                           // P = R/diag
  if (!rc) rc = dlp_gram(b.n, C1, b.R, b.ld, b.R, b.ld, b.diag, h->bk, DLP_F32, st);   // G = R^T D^-1 R
  if (rc) return rc;
  if (use_graph) {
    rc = capture_bpcg(h, iters_per_launch);
    if (rc) return rc;
  }
  int64_t it = 0;
  float v = 0.f, best = 3.4e38f;
  int stall = 0;
  while (it < max_iter) {
    DLP_CHECK_HIP(hipEventRecord(h->ev0, st));
    if (use_graph) {
      DLP_CHECK_HIP(hipGraphLaunch(h->bpcg_exec, st));
    } else {
      for (int i = 0; i < iters_per_launch; ++i) {
        rc = bpcg_iteration(h, b, st);
        if (rc) return rc;
      }
    }
    it += iters_per_launch;
    DLP_CHECK_HIP(hipEventRecord(h->ev1, st));
    DLP_CHECK_HIP(hipMemcpyAsync(h->host, h->pcg_viol, sizeof(float), hipMemcpyDeviceToHost, st));
    DLP_CHECK_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->ev0, h->ev1) == hipSuccess) h->pcg_ms += ms;
    h->pcg_iters += iters_per_launch;
    v = *(volatile float*)h->host;
    if (v <= 1.f) break;
    if (v < 0.9f * best) {
      best = v;
      stall = 0;
    } else if (++stall >= patience) {
      break;
    }
  }
  if (iters_out) *iters_out = it;
  if (viol_out) *viol_out = v;
  return DLP_OK;
}

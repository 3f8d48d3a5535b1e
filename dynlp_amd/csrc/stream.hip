// Resident incremental state of one graph: edge-batch apply on a slack CSR, grounding,
// changed-row refresh of the FP64 residual, and the round-synchronous push.
//
// Graph store: every vertex owns a row segment [indptr[r], indptr[r+1]) of which the first
// rowlen[r] entries are live (column, FP32 weight).  Deletions tombstone (col -1), a per-row
// compaction closes the gaps, insertions upsert in place or append into the slack; an append
// that finds the row full is recorded as overflow and retried after the host regrows the
// store.  Labeled vertices stay in the store; `solved` (U: unlabeled and grounded) masks the
// operator, so W_UU never has to be materialised.
#include "rowgather.h"
#include "scan.h"

namespace dlp {

// ---------------------------------------------------------------------------------------
// Edge-batch apply.

__device__ __forceinline__ void stamp_append(int v, int32_t* __restrict__ stamp, int stampv,
                                             int32_t* __restrict__ list, int32_t* __restrict__ count) {
  if (atomicExch(stamp + v, stampv) != stampv) list[atomicAdd(count, 1)] = v;
}

// Directed deletion e < 2*nd: row a loses entry b.  Rows are stamped into the changed list;
// the labeled-neighbour counter of a drops when b is labeled.
__global__ void k_apply_delete(int nd, const int32_t* __restrict__ du, const int32_t* __restrict__ dv,
                               const int32_t* __restrict__ indptr, const int32_t* __restrict__ rowlen,
                               int32_t* __restrict__ col, float* __restrict__ w,
                               const uint8_t* __restrict__ labeled, int32_t* __restrict__ labcnt,
                               int32_t* __restrict__ stamp, int stampv, int32_t* __restrict__ rows,
                               int32_t* __restrict__ nrows, int32_t* __restrict__ missing) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * (int64_t)nd) return;
  const int a = e < nd ? du[e] : dv[e - nd];
  const int b = e < nd ? dv[e] : du[e - nd];
  if (a == b) return;
  const int beg = indptr[a], end = beg + rowlen[a];
  bool found = false;
  for (int j = beg; j < end; ++j) {
    if (col[j] == b) {
      col[j] = -1;
      w[j] = 0.f;
      found = true;
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

// Directed insertion e < 2*ni: upsert (a, b, wt).  A full row records e as overflow.
__global__ void k_apply_insert(int ni, const int32_t* __restrict__ iu, const int32_t* __restrict__ iv,
                               const float* __restrict__ iw, const int32_t* __restrict__ indptr,
                               int32_t* __restrict__ rowlen, int32_t* __restrict__ col,
                               float* __restrict__ w, const uint8_t* __restrict__ labeled,
                               int32_t* __restrict__ labcnt, int32_t* __restrict__ stamp, int stampv,
                               int32_t* __restrict__ rows, int32_t* __restrict__ nrows,
                               int32_t* __restrict__ overflow, int32_t* __restrict__ noverflow) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * (int64_t)ni) return;
  const int a = e < ni ? iu[e] : iv[e - ni];
  const int b = e < ni ? iv[e] : iu[e - ni];
  const float wt = e < ni ? iw[e] : iw[e - ni];
  if (a == b) return;
  const int beg = indptr[a], cap = indptr[a + 1] - beg;
  const int len = __hip_atomic_load(rowlen + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (int j = beg; j < beg + len; ++j) {
    if (col[j] == b) {  // existing pair: new weight
      w[j] = wt;
      stamp_append(a, stamp, stampv, rows, nrows);
      return;
    }
  }
  const int slot = atomicAdd(rowlen + a, 1);
  if (slot < cap) {
    w[beg + slot] = wt;
    col[beg + slot] = b;
    if (labeled[b]) atomicAdd(labcnt + a, 1);
    stamp_append(a, stamp, stampv, rows, nrows);
  } else {
    atomicSub(rowlen + a, 1);
    overflow[atomicAdd(noverflow, 1)] = (int32_t)e;
  }
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

__device__ __forceinline__ int ld_rlx(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx(int32_t* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int uf_find2(int32_t* P, int x) {
  while (true) {
    const int p = ld_rlx(P + x);
    if (p == x) return x;
    const int gp = ld_rlx(P + p);
    if (gp == p) return p;
    st_rlx(P + x, gp);
    x = gp;
  }
}
__device__ __forceinline__ void uf_union2(int32_t* P, int a, int b) {
  while (true) {
    a = uf_find2(P, a);
    b = uf_find2(P, b);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(P + a, a, b) == a) return;
  }
}

__global__ void k_ground_init(int n, int32_t* __restrict__ P, uint8_t* __restrict__ grounded) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    P[i] = i;
    grounded[i] = 0;
  }
}

__global__ __launch_bounds__(256) void k_ground_hook(int n, const int32_t* __restrict__ indptr,
                                                     const int32_t* __restrict__ rowlen,
                                                     const int32_t* __restrict__ col,
                                                     const uint8_t* __restrict__ labeled, int32_t* P) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(t >> 3);
  const int lane = (int)(t & 7);
  if (r >= n || labeled[r]) return;
  const int beg = indptr[r], end = beg + rowlen[r];
  for (int j = beg + lane; j < end; j += 8) {
    const int c = col[j];
    if (c > r && !labeled[c]) uf_union2(P, r, c);
  }
}

__global__ void k_ground_roots(int n, int32_t* P, const uint8_t* __restrict__ labeled,
                               const int32_t* __restrict__ labcnt, uint8_t* __restrict__ grounded) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || labeled[i]) return;
  const int r = uf_find2(P, i);
  P[i] = r;
  if (labcnt[i] > 0) grounded[r] = 1;
}

// New solved flag; vertices that toggle are listed (for the neighbour pass) and stamped changed.
__global__ void k_ground_toggle(int n, const int32_t* __restrict__ P, const uint8_t* __restrict__ labeled,
                                const uint8_t* __restrict__ grounded, uint8_t* __restrict__ solved,
                                int32_t* __restrict__ toggled, int32_t* __restrict__ ntoggled,
                                int32_t* __restrict__ stamp, int stampv, int32_t* __restrict__ rows,
                                int32_t* __restrict__ nrows) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint8_t now = (!labeled[i] && grounded[P[i]]) ? 1 : 0;
  if (now != solved[i]) {
    solved[i] = now;
    toggled[atomicAdd(ntoggled, 1)] = i;
    stamp_append(i, stamp, stampv, rows, nrows);
  }
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
};

template <typename TZ, typename TR, int CPL>
__device__ __forceinline__ void refresh_row(const RefreshArgs& A, const Lanes& L, int64_t i, int nr,
                                            TZ* __restrict__ Z, TR* __restrict__ R) {
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
  bool viol = false;
  if (valid && L.e == 0) {
    const int64_t zo = (int64_t)r * A.ld;
    if (!live) {
      if (L.l == 0) {
        A.diag64[r] = 1.0;
        A.diag32[r] = 1.f;
      }
#pragma unroll
      for (int k = 0; k < CPL; ++k) {
        const int c = L.l + k * L.LC;
        if (c < A.C1) {
          Z[zo + c] = (TZ)(c == C ? 1 : 0);
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
          const double ratio = fabs(res) / dg;
          viol |= ratio > A.eps[c];
          if (A.rho) atomic_max_nonneg(A.rho + c, ratio);
        }
      }
    }
  }
  const unsigned long long bal = __ballot(viol);
  const int wl = threadIdx.x & 63;
  const unsigned long long gm = (L.G == 64 ? ~0ull : ((1ull << L.G) - 1)) << (wl & ~(L.G - 1));
  if (valid && L.lane == 0 && (bal & gm) && A.viol) A.viol[atomicAdd(A.nviol, 1)] = r;
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
  for (int64_t base = wave_first; base < nr; base += groups)
    refresh_row<TZ, TR, CPL>(A, L, base + (g0 - wave_first), nr, Z, R);
}

// ---------------------------------------------------------------------------------------
// Round-synchronous push (the service's push; deterministic).
//
// Round t: F_t (stamped tag[t%2][u] == t, D[t%2][u] = its increment) has been pushed
// (Z_u += D_u, R_u = 0).  Candidates N(F_t) (list t%3, each once) gather
//   R_v += sum_{u in N(v), tag[t%2][u] == t} w_vu D_u
// and, if now violating, are pushed at once for round t+1 (stamp t+1, D[(t+1)%2]) and append
// their neighbours to list (t+1)%3.  Each wave owns one candidate's R/Z rows, and every value
// read from other vertices belongs to the other parity, so no two waves race.  The round
// number lives on the device (rnd[node parity]) so one captured graph replays any number of
// times; counts are a ring of three (read t%3, append (t+1)%3, reset (t+2)%3).

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
  float* D[2];
  int32_t* tag[2];
  int32_t* ctag;         // candidate stamp (round of the list a vertex was appended to)
  int32_t* list[3];
  int32_t* count;        // [3]
  int32_t* rnd;          // [2] round number by node parity
  int32_t* done;         // set when a round finds no candidates
  int32_t* btag;         // batch stamp: vertex touched during the current batch
  int32_t* blist;
  int32_t* bcount;
  const int32_t* bstamp; // device: current batch stamp
  unsigned long long* edges;
  unsigned long long* stats;  // [0] candidates, [1] pushes, [2] frontier entries gathered, [3] entries scanned by pushes
};

template <int CPL>
__device__ __forceinline__ void push_vertex(const PushState& S, const Lanes& L, int v, int tq,
                                            float (&r)[CPL], int32_t* next, int32_t* ncount) {
  // r holds R_v (columns of this lane); push it: D = R/diag, Z += D, R = 0
  const float inv = 1.f / S.diag[v];
  const int64_t o = (int64_t)v * S.ld;
  const int q = tq & 1;
  if (L.e == 0) {
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = L.l + k * L.LC;
      if (c < S.C1) {
        const float d = r[k] * inv;
        S.Z[o + c] += d;
        S.R[o + c] = 0.f;
        S.D[q][o + c] = d;
      }
    }
  }
  if (L.lane == 0) S.tag[q][v] = tq;
  const int beg = S.indptr[v], end = beg + S.rowlen[v];
  if (L.lane == 0) {
    atomicAdd(S.stats + 1, 1ull);
    atomicAdd(S.stats + 3, (unsigned long long)(end - beg));
  }
  for (int j = beg + L.lane; j < end; j += L.G) {
    const int x = S.col[j];
    if (x >= 0 && S.solved[x]) {
      if (atomicExch(S.ctag + x, tq) != tq) next[atomicAdd(ncount, 1)] = x;
    }
  }
}

template <int CPL>
__device__ __forceinline__ bool violates(const PushState& S, const Lanes& L, int v, const float (&r)[CPL]) {
  const float dv = S.diag[v];
  bool viol = false;
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int c = L.l + k * L.LC;
    if (c < S.C1) viol |= fabsf(r[k]) > S.eps[c] * dv;
  }
  const unsigned long long bal = __ballot(viol);
  const int wl = threadIdx.x & 63;
  const unsigned long long gm = (L.G == 64 ? ~0ull : ((1ull << L.G) - 1)) << (wl & ~(L.G - 1));
  return (bal & gm) != 0;
}

// Seed: push the listed violators as F_{t0} (t0 = rnd[0]), candidates into list t0%3.
template <int CPL>
__device__ __forceinline__ void seed_one(const PushState& S, const Lanes& L, int v, int t0) {
  float r[CPL];
  const int64_t o = (int64_t)v * S.ld;
#pragma unroll
  for (int k = 0; k < CPL; ++k) {
    const int c = L.l + k * L.LC;
    r[k] = c < S.C1 ? S.R[o + c] : 0.f;
  }
  const int bs = *S.bstamp;
  if (L.lane == 0 && atomicExch(S.btag + v, bs) != bs) S.blist[atomicAdd(S.bcount, 1)] = v;
  push_vertex<CPL>(S, L, v, t0, r, S.list[t0 % 3], S.count + (t0 % 3));
}

template <int CPL>
__global__ __launch_bounds__(256) void k_push_seed(PushState S, const int32_t* __restrict__ seeds,
                                                   const int32_t* __restrict__ nseeds, int g_shift,
                                                   int lc_shift) {
  const Lanes L(g_shift, lc_shift);
  const int t0 = S.rnd[0];
  const int ns = *nseeds;
  const int64_t groups = ((int64_t)gridDim.x * blockDim.x) >> g_shift;
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> g_shift; i < ns; i += groups)
    seed_one<CPL>(S, L, seeds[i], t0);
}

// Reset for a new push phase: counts, done flag; rnd[1] mirrors rnd[0].
__global__ void k_push_begin(PushState S) {
  if (threadIdx.x == 0) {
    const int t0 = S.rnd[0];
    S.count[0] = S.count[1] = S.count[2] = 0;
    S.rnd[1] = t0;
    *S.done = 0;
  }
}

template <int CPL>
__global__ __launch_bounds__(256) void k_push_round(PushState S, int parity, int g_shift, int lc_shift) {
  const Lanes L(g_shift, lc_shift);
  const int t = S.rnd[parity];
  const int nc = S.count[t % 3];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.count[(t + 2) % 3] = 0;
    S.rnd[parity ^ 1] = t + 1;
    if (nc == 0) *S.done = 1;
  }
  if (nc == 0) return;
  const int p = t & 1;
  const int tq = t + 1;
  const int32_t* cand = S.list[t % 3];
  int32_t* next = S.list[tq % 3];
  int32_t* ncount = S.count + (tq % 3);
  const int bs = *S.bstamp;
  const int64_t groups = ((int64_t)gridDim.x * blockDim.x) >> g_shift;
  const int64_t g0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> g_shift;
  const int64_t wave_first = ((int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63)) >> g_shift;
  unsigned long long my_edges = 0, my_cand = 0, my_dg = 0;
  for (int64_t base = wave_first; base < nc; base += groups) {
    const int64_t gi = base + (g0 - wave_first);
    const bool valid = gi < nc;
    const int v = valid ? cand[gi] : 0;
    if (valid && L.lane == 0) ++my_cand;
    int beg = 0, end = 0;
    if (valid) {
      beg = S.indptr[v];
      end = beg + S.rowlen[v];
      if (L.lane == 0) {
        my_edges += (unsigned long long)(end - beg);
        if (atomicExch(S.btag + v, bs) != bs) S.blist[atomicAdd(S.bcount, 1)] = v;
      }
    }
    // gather the increments of this round's frontier neighbours
    float acc[CPL];
#pragma unroll
    for (int k = 0; k < CPL; ++k) acc[k] = 0.f;
    for (int b0 = beg; b0 < end; b0 += L.G) {
      const int j = b0 + L.lane;
      int myv = 0;
      float myw = 0.f;
      if (j < end) {
        myv = S.col[j];
        if (myv >= 0 && S.tag[p][myv] == t && S.solved[myv]) myw = S.w[j];
        if (myw != 0.f) ++my_dg;
      }
      const int cnt = min(L.G, end - b0);
      for (int t0 = 0; t0 < cnt; t0 += 4 * L.LE) {
        int vv[4];
        float ww[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int tt = t0 + u * L.LE + L.e;
          vv[u] = __shfl(myv, tt & (L.G - 1), L.G);
          ww[u] = __shfl(myw, tt & (L.G - 1), L.G);
          if (tt >= cnt) ww[u] = 0.f;
        }
        float xv[4][CPL];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float* dr = S.D[p] + (int64_t)vv[u] * S.ld;
#pragma unroll
          for (int k = 0; k < CPL; ++k) {
            const int c = L.l + k * L.LC;
            xv[u][k] = (ww[u] != 0.f && c < S.C1) ? dr[c] : 0.f;
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int k = 0; k < CPL; ++k) acc[k] += ww[u] * xv[u][k];
      }
    }
    for (int off = L.LC; off < L.G; off <<= 1) {
#pragma unroll
      for (int k = 0; k < CPL; ++k) acc[k] += __shfl_xor(acc[k], off, 64);
    }
    // R_v += acc (edge lane 0 holds the sums; broadcast to the group)
    float r[CPL];
    const int64_t o = (int64_t)v * S.ld;
#pragma unroll
    for (int k = 0; k < CPL; ++k) {
      const int c = L.l + k * L.LC;
      float val = 0.f;
      if (valid && L.e == 0 && c < S.C1) {
        val = S.R[o + c] + acc[k];
        S.R[o + c] = val;
      }
      r[k] = __shfl(val, L.l, L.G);
    }
    const bool vio = violates<CPL>(S, L, valid ? v : 0, r) && valid;
    if (vio) push_vertex<CPL>(S, L, v, tq, r, next, ncount);
  }
  const unsigned long long se = wave_sum_u64(my_edges);
  const unsigned long long sc = wave_sum_u64(my_cand);
  const unsigned long long sd = wave_sum_u64(my_dg);
  if ((threadIdx.x & 63) == 0) {
    if (se) atomicAdd(S.edges, se);
    if (sc) atomicAdd(S.stats + 0, sc);
    if (sd) atomicAdd(S.stats + 2, sd);
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
  uint8_t* grounded = nullptr;
  int32_t* viol = nullptr;
  double* eps64 = nullptr;
  double* rho = nullptr;
  uint8_t* mark = nullptr;      // FP64 path marks
  double* dsnap = nullptr;      // FP64 path increments of the current candidates
  int32_t* cand64 = nullptr;
  void* scan_ws = nullptr;
  size_t scan_bytes = 0;
  PushState S;
  int64_t* host = nullptr;      // pinned status mirror
  hipGraphExec_t gexec = nullptr;
  int graph_rounds = 0;
  hipStream_t graph_stream = nullptr;
  RowMap map;                   // lane layout for C1 columns
  RowMap map_seed;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double push_ms = 0.0;         // summed duration of push-round launches (events)
  int64_t push_launched = 0;    // rounds launched
};

enum { CNT_ROWS = 0, CNT_MISSING, CNT_OVERFLOW, CNT_TOGGLED, CNT_VIOL, CNT_CAND64, CNT_N };

template <typename T>
int dev_alloc(T** p, size_t count) {
  DLP_CHECK_HIP(hipMalloc((void**)p, sizeof(T) * (count ? count : 1)));
  return DLP_OK;
}

int lp_free(LP* h) {
  void* ptrs[] = {h->stamp, h->rows, h->cnt, h->overflow, h->toggled, h->parent, h->grounded,
                  h->viol, h->eps64, h->rho, h->mark, h->dsnap, h->cand64, h->scan_ws,
                  h->S.D[0], h->S.D[1], h->S.tag[0], h->S.tag[1], h->S.ctag, h->S.list[0],
                  h->S.list[1], h->S.list[2], h->S.count, h->S.rnd, h->S.done, h->S.btag,
                  h->S.blist, h->S.bcount, (void*)h->S.bstamp, (void*)h->S.eps, h->S.edges,
                  h->S.stats};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (h->host) (void)hipHostFree(h->host);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
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
#define A_(p, c) if (!rc) rc = dev_alloc(&(p), (c))
  A_(h->stamp, n); A_(h->rows, n); A_(h->cnt, CNT_N); A_(h->overflow, 1 << 20);
  A_(h->toggled, n); A_(h->parent, n); A_(h->grounded, n); A_(h->viol, n);
  A_(h->eps64, d->C1); A_(h->rho, d->C1); A_(h->mark, n); A_(h->cand64, n);
  A_(h->S.D[0], nz); A_(h->S.D[1], nz); A_(h->S.tag[0], n); A_(h->S.tag[1], n);
  A_(h->S.ctag, n); A_(h->S.list[0], n); A_(h->S.list[1], n); A_(h->S.list[2], n);
  A_(h->S.count, 3); A_(h->S.rnd, 2); A_(h->S.done, 1); A_(h->S.btag, n); A_(h->S.blist, n);
  A_(h->S.bcount, 1); A_(h->S.edges, 1); A_(h->S.stats, 4);
  int32_t* bst = nullptr;
  float* eps32 = nullptr;
  A_(bst, 1); A_(eps32, d->C1);
#undef A_
  h->S.bstamp = bst;
  h->S.eps = eps32;
  if (!rc && d->Z64) rc = dev_alloc(&h->dsnap, nz);
  h->scan_bytes = dlp_scan_workspace_bytes(n);
  if (!rc && hipMalloc(&h->scan_ws, h->scan_bytes) != hipSuccess) rc = DLP_ERR_HIP;
  if (!rc && hipHostMalloc((void**)&h->host, 64 * sizeof(int64_t)) != hipSuccess) rc = DLP_ERR_HIP;
  if (rc) {
    set_error("lp_create: device allocation failed");
    lp_free(h);
    return DLP_ERR_HIP;
  }
  // initial contents: stamps/tags -1 (never equal to a live stamp), counters 0
  (void)hipMemset(h->stamp, 0xff, sizeof(int32_t) * n);
  (void)hipMemset(h->S.tag[0], 0xff, sizeof(int32_t) * n);
  (void)hipMemset(h->S.tag[1], 0xff, sizeof(int32_t) * n);
  (void)hipMemset(h->S.ctag, 0xff, sizeof(int32_t) * n);
  (void)hipMemset(h->S.btag, 0xff, sizeof(int32_t) * n);
  (void)hipMemset(h->cnt, 0, sizeof(int32_t) * CNT_N);
  (void)hipMemset(h->S.count, 0, sizeof(int32_t) * 3);
  (void)hipMemset(h->S.rnd, 0, sizeof(int32_t) * 2);
  (void)hipMemset(h->S.done, 0, sizeof(int32_t));
  (void)hipMemset(h->S.bcount, 0, sizeof(int32_t));
  (void)hipMemset((void*)h->S.bstamp, 0, sizeof(int32_t));
  (void)hipMemset(h->S.edges, 0, sizeof(unsigned long long));
  (void)hipMemset(h->S.stats, 0, 4 * sizeof(unsigned long long));
  (void)hipMemset(h->mark, 0, n);
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
  if (threadIdx.x < CNT_N) cnt[threadIdx.x] = 0;
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
  if (nd > 0) {
    hipLaunchKernelGGL(k_apply_delete, dim3(grid_for(2LL * nd)), dim3(kBlock), 0, st, nd, du, dv,
                       h->d.indptr, h->d.rowlen, h->d.col, h->d.w, h->d.labeled, h->d.labcnt,
                       h->stamp, h->stampv, h->rows, h->cnt + CNT_ROWS, h->cnt + CNT_MISSING);
    DLP_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_compact_rows, dim3(grid_for(2LL * nd)), dim3(kBlock), 0, st,
                       h->cnt + CNT_ROWS, h->rows, h->d.indptr, h->d.rowlen, h->d.col, h->d.w);
    DLP_CHECK_LAUNCH();
  }
  if (ni > 0) {
    hipLaunchKernelGGL(k_apply_insert, dim3(grid_for(2LL * ni)), dim3(kBlock), 0, st, ni, iu, iv, iw,
                       h->d.indptr, h->d.rowlen, h->d.col, h->d.w, h->d.labeled, h->d.labcnt,
                       h->stamp, h->stampv, h->rows, h->cnt + CNT_ROWS, h->overflow,
                       h->cnt + CNT_OVERFLOW);
    DLP_CHECK_LAUNCH();
  }
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

// Directed re-insertion (each overflowed entry is one direction only).
__global__ void k_apply_insert_directed(const int32_t* __restrict__ nent, const int32_t* __restrict__ ru,
                                        const int32_t* __restrict__ rv, const float* __restrict__ rw,
                                        const int32_t* __restrict__ indptr, int32_t* __restrict__ rowlen,
                                        int32_t* __restrict__ col, float* __restrict__ w,
                                        const uint8_t* __restrict__ labeled, int32_t* __restrict__ labcnt,
                                        int32_t* __restrict__ stamp, int stampv, int32_t* __restrict__ rows,
                                        int32_t* __restrict__ nrows, int32_t* __restrict__ overflow,
                                        int32_t* __restrict__ noverflow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *nent) return;
  const int a = ru[i], b = rv[i];
  const int beg = indptr[a], cap = indptr[a + 1] - beg;
  const int slot = atomicAdd(rowlen + a, 1);
  if (slot < cap) {
    w[beg + slot] = rw[i];
    col[beg + slot] = b;
    if (labeled[b]) atomicAdd(labcnt + a, 1);
    stamp_append(a, stamp, stampv, rows, nrows);
  } else {
    atomicSub(rowlen + a, 1);
    overflow[atomicAdd(noverflow, 1)] = i;
  }
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
    hipLaunchKernelGGL(k_apply_insert_directed, dim3(grid_for(1 << 20)), dim3(kBlock), 0, st, cnt2, ru, rv,
                       rw, new_indptr, h->d.rowlen, new_col, new_w, h->d.labeled, h->d.labcnt, h->stamp,
                       h->stampv, h->rows, h->cnt + CNT_ROWS, h->overflow, h->cnt + CNT_OVERFLOW);
    DLP_CHECK_LAUNCH();
  }
  return DLP_OK;
}

extern "C" int dlp_lp_ground(void* handle, void* stream) {
  DLP_CHECK_ARG(handle, "lp_ground: null handle");
  LP* h = (LP*)handle;
  const int n = h->d.n;
  hipStream_t st = (hipStream_t)stream;
  DLP_CHECK_HIP(hipMemsetAsync(h->cnt + CNT_TOGGLED, 0, sizeof(int32_t), st));
  hipLaunchKernelGGL(k_ground_init, dim3(grid_for(n)), dim3(kBlock), 0, st, n, h->parent, h->grounded);
  DLP_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ground_hook, dim3(grid_for((long long)n * 8)), dim3(kBlock), 0, st, n, h->d.indptr,
                     h->d.rowlen, h->d.col, h->d.labeled, h->parent);
  DLP_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ground_roots, dim3(grid_for(n)), dim3(kBlock), 0, st, n, h->parent, h->d.labeled,
                     h->d.labcnt, h->grounded);
  DLP_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ground_toggle, dim3(grid_for(n)), dim3(kBlock), 0, st, n, h->parent, h->d.labeled,
                     h->grounded, h->d.solved, h->toggled, h->cnt + CNT_TOGGLED, h->stamp, h->stampv,
                     h->rows, h->cnt + CNT_ROWS);
  DLP_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_mark_neighbours_changed, dim3(grid_for((long long)n * 8)), dim3(kBlock), 0, st,
                     h->toggled, h->cnt + CNT_TOGGLED, h->d.indptr, h->d.rowlen, h->d.col, h->stamp,
                     h->stampv, h->rows, h->cnt + CNT_ROWS);
  DLP_CHECK_LAUNCH();
  return DLP_OK;
}

extern "C" int dlp_lp_refresh(void* handle, int which, int prec, void* stream) {
  DLP_CHECK_ARG(handle && which >= 0 && which <= 2 && (prec == 0 || prec == 1), "lp_refresh: bad argument");
  LP* h = (LP*)handle;
  DLP_CHECK_ARG(prec == 0 || (h->d.Z64 && h->d.R64), "lp_refresh: FP64 state not attached");
  hipStream_t st = (hipStream_t)stream;
  DLP_CHECK_HIP(hipMemsetAsync(h->cnt + CNT_VIOL, 0, sizeof(int32_t), st));
  DLP_CHECK_HIP(hipMemsetAsync(h->rho, 0, sizeof(double) * h->d.C1, st));
  RefreshArgs A = refresh_args(h, which);
  const RowMap& m = h->map;
  const unsigned grid = persistent_grid(h->d.n, m.g_shift);
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
                           int64_t* rounds_out, void* stream) {
  DLP_CHECK_ARG(handle && max_rounds >= 0 && rounds_per_launch >= 2 && (rounds_per_launch % 2) == 0,
                "lp_push: bad argument (rounds_per_launch must be even)");
  LP* h = (LP*)handle;
  hipStream_t st = (hipStream_t)stream;
  const RowMap& m = h->map;
  const RowMap& ms = h->map;
  hipLaunchKernelGGL(k_push_begin, dim3(1), dim3(64), 0, st, h->S);
  DLP_CHECK_LAUNCH();
  {
    const unsigned grid = persistent_grid(h->d.n, ms.g_shift);
#define K(CPLV)                                                                                   \
  hipLaunchKernelGGL((k_push_seed<CPLV>), dim3(grid), dim3(kBlock), 0, st, h->S, h->viol, h->cnt + CNT_VIOL, \
                     ms.g_shift, ms.lc_shift);
    DLP_DISPATCH_CPL(ms.cpl, K)
#undef K
    DLP_CHECK_LAUNCH();
  }
  if (use_graph) {
    int rc = capture_rounds(h, rounds_per_launch, st);
    if (rc) return rc;
  }
  const unsigned grid = persistent_grid(h->d.n, m.g_shift);
  int64_t rounds = 0;
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
    DLP_CHECK_HIP(hipMemcpyAsync(h->host, h->S.done, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    DLP_CHECK_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->ev0, h->ev1) == hipSuccess) h->push_ms += ms;
    h->push_launched += rounds_per_launch;
    if (*(volatile int32_t*)h->host) break;
  }
  if (rounds_out) *rounds_out = rounds;
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
  for (int c = 0; c < C1; ++c) viol |= fabs(R[(int64_t)v * ld + c]) > eps[c] * d;
  if (viol) out[atomicAdd(nout, 1)] = v;
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
  int32_t tmp[CNT_N + 8];
  DLP_CHECK_HIP(hipMemcpyAsync(tmp, h->cnt, sizeof(int32_t) * CNT_N, hipMemcpyDeviceToHost, st));
  DLP_CHECK_HIP(hipMemcpyAsync(tmp + CNT_N, h->S.bcount, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  DLP_CHECK_HIP(hipMemcpyAsync(tmp + CNT_N + 1, h->S.rnd, sizeof(int32_t) * 2, hipMemcpyDeviceToHost, st));
  unsigned long long edges = 0, pst[4] = {0, 0, 0, 0};
  DLP_CHECK_HIP(hipMemcpyAsync(&edges, h->S.edges, sizeof(edges), hipMemcpyDeviceToHost, st));
  DLP_CHECK_HIP(hipMemcpyAsync(pst, h->S.stats, sizeof(pst), hipMemcpyDeviceToHost, st));
  if (rho) DLP_CHECK_HIP(hipMemcpyAsync(rho, h->rho, sizeof(double) * h->d.C1, hipMemcpyDeviceToHost, st));
  DLP_CHECK_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < CNT_N; ++i) out[i] = tmp[i];
  out[CNT_N] = tmp[CNT_N];
  out[CNT_N + 1] = tmp[CNT_N + 1];
  out[CNT_N + 2] = (int64_t)edges;
  for (int i = 0; i < 4; ++i) out[CNT_N + 3 + i] = (int64_t)pst[i];
  out[CNT_N + 7] = h->push_launched;
  out[CNT_N + 8] = (int64_t)(h->push_ms * 1000.0);  // microseconds
  return DLP_OK;
}

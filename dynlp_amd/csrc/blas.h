// Column-wise PCG passes (blas.hip) used by the resident state's global solve.
#pragma once
#include "common.h"

namespace dlp {

struct PcgBufs {
  int n, C;
  int64_t ld;
  float* Z;
  float* R;
  float* P;
  float* AP;
  const float* diag;
  const float* eps;
  double* rz;
  double* rz_new;
  double* pAp;
  double* alpha;
  double* beta;
  float* viol;   // [0] the iteration's violation, [1] the last completed one (for the host)
  // [0] done (set once the violation is <= 1: every later pass of a captured group returns at
  // once), [1] iterations completed; null disables the device-side exit
  int32_t* ctl;
  // low-precision directions (null: P in FP32): rows of ldh 16-bit units; rp = r.p per column;
  // p16 = 1 BF16 or 2 FP16 (ldh % 8 == 0), 3 FP8 E4M3 (ldh even: 4-byte groups of 4 values);
  // FP16 and FP8 are held scaled by sig[c] (a power of two per column, C floats).  FP8 scales
  // come from the column maxima rmax (max |r|/diag), pmax (max |P| as held) and emax (max |e|
  // of the start's coarse correction), C floats each, zero between uses
  uint16_t* P16;
  int64_t ldh;
  double* rp;
  int p16;
  float* sig;
  float* rmax;
  float* pmax;
  float* emax;
  // deferred solution updates (16-bit directions on the flat passes; slot_stride 0: Z is updated
  // by every iteration): a group of `group` iterations keeps its directions in slots of
  // slot_stride values (slot s at P16 + s*slot_stride); iteration `slot` gathers slot `slot`,
  // writes the next direction to slot `slot`+1 and its step lengths to alpha_hist[slot*C ..]
  // (group*C floats, rounded up to an even count); pcg_flush then adds sum_s alpha_s P_s to Z
  int64_t slot_stride;
  int slot, group;
  float* alpha_hist;
  // two-level preconditioner M = D^-1 + P Ac^-1 P^T (na = 0: Jacobi only): aggregate I is the
  // row range [bounds[I], bounds[I+1]) (at least min_agg rows), aggmap[r] = the aggregate of row
  // r, P[r, aggmap[r]] = 1 on solved rows; rc = P^T r (na x ldc doubles), ec = the coarse
  // correction Ac^-1 rc (na x ldc floats), Ainv = Ac^-1 (na x na), solved = row mask
  int na, min_agg;
  const int32_t* bounds;
  const uint8_t* aggmap;
  int64_t ldc;
  double* rc;
  float* ec;
  const double* Ainv;
  const uint8_t* solved;
};

// Longest group of iterations with deferred solution updates.
constexpr int kMaxGroup = 8;

// The group's deferred update Z += sum_s alpha_s P_s (no-op without deferred updates), and the
// next group's first direction moved from slot `group` to slot 0 unless the solve is done.
int pcg_flush(const PcgBufs& b, hipStream_t st);

// Rows per workgroup of the flat PCG passes for this layout (0 when the flat form does not
// apply): the coarse aggregates must be at least this long.
int pcg_flat_chunk(const PcgBufs& b);

// Ac = P^T A P of the solved rows (aggregates = row ranges of bounds, aggmap) into Ac (na x na),
// then Ainv = Ac^-1 (FP64 Cholesky in one workgroup, na <= kCoarseMax).
constexpr int kCoarseMax = 128;
int coarse_build(int n, int na, const int32_t* bounds, const uint8_t* aggmap, const int32_t* indptr,
                 const int32_t* rowlen, const int32_t* col, const float* w, const uint8_t* solved,
                 const double* diag, double* Ac, double* Ainv, hipStream_t st);

// Candidate aggregate boundaries: cross[b] = edges (u < v) with u < b <= v, then for every
// k in [1, nk] the b within [k*S0 - S0/2, k*S0 + S0/2) (clamped to [min_agg, n - min_agg)) with
// the fewest crossing edges (ties: nearest to k*S0) into cand[4(k-1)] (-1: empty window),
// followed by cross[b], cross[b - S0/2] and cross[b + S0/2].  diff and cross hold n ints each;
// ws/ws_bytes: scan workspace (scan_workspace_bytes(n)); cand holds 4 nk ints.
int coarse_boundaries(int n, int nk, int S0, int min_agg, const int32_t* indptr, const int32_t* rowlen,
                      const int32_t* col, int32_t* diff, int32_t* cross, void* ws, size_t ws_bytes,
                      int32_t* cand, hipStream_t st);
// aggmap[r] = I with bounds[I] <= r < bounds[I+1].
int coarse_aggmap(int n, int na, const int32_t* bounds, uint8_t* aggmap, hipStream_t st);

// P = R/diag, rz = sum R*R/diag, pAp = 0.
int pcg_start(const PcgBufs& b, hipStream_t st);
// After AP = A P has been formed: pAp (unless have_pap: the operator accumulated it), alpha,
// Z/R update (rz_new, viol), beta, new P.
int pcg_finish_iteration(const PcgBufs& b, hipStream_t st, bool have_pap = false);

// Operator with the fused p^T A p epilogue (spmm.hip).
int row_operator_dot(int n, int C, const int32_t* indptr, const int32_t* rowlen, const int32_t* indices,
                     const float* w, const uint8_t* mask, const float* X, int64_t ldx, const float* a, float b,
                     float* out, int64_t ldo, double* dot, hipStream_t st, const int32_t* skip = nullptr);

// The global solve's operator: pipelined row gathers, rows outside rowmask skipped, neighbour
// rows unmasked (P = 0 outside the solved set), fused p^T A p; skip = device exit flag;
// x16_kind = 1 (BF16) or 2 (FP16) for X16 (spmm.hip).
int pcg_operator(int n, int C, const int32_t* indptr, const int32_t* rowlen, const int32_t* indices, const float* w,
                 const uint8_t* rowmask, const float* X, int64_t ldx, const uint16_t* X16, int64_t ldh, int x16_kind,
                 const float* a, float b, float* out, int64_t ldo, double* dot, hipStream_t st, const int32_t* skip);

// Blocks of a 256-thread kernel that are resident on the current device at once (spmm.hip).
int resident_grid(const void* kernel);

}  // namespace dlp

/*
 * dynlp.h — C ABI of the MI355X (gfx950) label-propagation hot path.
 *
 * Every entry point takes device pointers owned by the caller, sizes as int32/int64,
 * dtype codes, and a hipStream_t passed as void* (NULL = default stream).  Kernels are
 * enqueued on that stream; no entry point synchronises the host unless its comment says so.
 * Return value: 0 on success, a negative DLP_ERR_* code otherwise; dlp_last_error() gives
 * the message (thread-local).  Scratch memory comes from caller workspaces whose size the
 * matching *_workspace_bytes() function reports; no entry point allocates device memory
 * except the dlp_push_plan_* / dlp_graph_* handles, which own their scratch until destroyed.
 *
 * CSR: indptr int32[n+1] holds row starts; when rowlen (int32[n]) is non-NULL the row r
 * holds rowlen[r] live entries from indptr[r] (slack CSR: capacity indptr[r+1]-indptr[r]),
 * otherwise indptr[r+1]-indptr[r].  Column indices int32.  Weights FP32 or FP64.
 * mask (uint8[n], may be NULL): vertices with mask 0 are outside the solved set; their rows
 * carry no off-diagonal term and they contribute nothing to other rows (W_UU restriction).
 * Dense blocks are row-major with a leading dimension (elements).
 *
 * Each declaration cites the reference interface it replaces (dynlp/<file>:<line>).
 */
#ifndef DYNLP_H
#define DYNLP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { DLP_F32 = 0, DLP_F64 = 1, DLP_I32 = 2, DLP_I64 = 3 };
enum { DLP_OK = 0, DLP_ERR_ARG = -1, DLP_ERR_HIP = -2, DLP_ERR_CAPACITY = -3, DLP_ERR_STATE = -4 };

/* Last error message of the calling thread (empty string if none). */
const char* dlp_last_error(void);
/* ABI version (incremented on any signature change). */
int dlp_abi_version(void);

/* ---- Row-major SpMM -------------------------------------------------------------------
 * Y = a .* X + b * (W X), all n rows.  Replaces kernels.spmm_axpy (kernels.py:377-393; CUDA
 * kernels.py:140-166, 240-266).  X, Y: n x C, leading dims ldx, ldy.  a: length n or NULL
 * (treated as 0).  dtype applies to w, X, a and Y alike (the reference requires W and X to
 * share a dtype, solvers.py:36-37).  avg_deg picks the lane layout (0 = compute it). */
int dlp_spmm_axpy(int n, int C, const int32_t* indptr, const int32_t* indices, const void* w,
                  const void* X, int64_t ldx, const void* a, double b, void* Y, int64_t ldy,
                  int dtype, double avg_deg, void* stream);

/* ---- Masked row operator (the building block of apply_A / residual refresh) -----------
 * For each listed row r (rows[i], or r = i when rows is NULL, i < nrows):
 *   out[o] = add[r] + a[r] * X[r] + b * sum_{j in row r, mask[col_j]} w_j X[col_j]
 * where o = i when compact_out != 0, else o = r; rows with mask[r] == 0 skip the sum.
 * add and a may be NULL.  Types: w_dtype for w; x_dtype for X; out_dtype for out, add and
 * the accumulation; a is FP64 when out is FP64, else FP32.
 * Replaces solvers.apply_A (solvers.py:46-49) and the changed-row refresh
 * R[Q] = rhs[Q] - diag[Q] Z[Q] + W[Q] Z (solvers.py:404-407). */
int dlp_row_operator(int nrows, const int32_t* rows, int C, const int32_t* indptr,
                     const int32_t* rowlen, const int32_t* indices, const void* w, int w_dtype,
                     const uint8_t* mask, const void* X, int64_t ldx, int x_dtype,
                     const void* a, double b, const void* add, int64_t ldadd, void* out,
                     int64_t ldo, int out_dtype, int compact_out, double avg_deg, void* stream);

/* FP32 operator over all rows with the column dot fused into its epilogue:
 *   out[r] = a[r] * X[r] + b * sum_{j in row r, mask[col_j]} w_j X[col_j]   (every row r < n)
 *   dot[c] += sum_r X[r,c] * out[r,c]   (FP64, accumulated: dot is not cleared)
 * With a = diag, b = -1 this is AP = apply_A(P) and pAp = (P * AP).sum(axis=0) of the
 * reference PCG loop (solvers.py:78-79) in one pass.  Needs ldx % 4 == 0, ldo % 4 == 0 and
 * 16-byte aligned X and out (DLP_ERR_ARG otherwise); C <= 256. */
int dlp_row_operator_dot(int n, int C, const int32_t* indptr, const int32_t* rowlen,
                         const int32_t* indices, const float* w, const uint8_t* mask,
                         const float* X, int64_t ldx, const float* a, float b, float* out,
                         int64_t ldo, double* dot, void* stream);

/* Process-wide tuning of the row kernels (no reference counterpart; the reference templated
 * lanes per row at JIT time, kernels.py:365-374):
 *   key 0 = unroll factor of the gather loops (0 default = per kernel, 4 or 8);
 *   key 1 = 16-byte FP32 loads when the row layout allows them (1 default, 0 off);
 *   key 2 = packed lane groups, ceil(C/4) column lanes per row (1 default) or a power of two;
 *   key 3 = XCD-aware row order (each XCD walks one contiguous eighth of the rows; 1 default);
 *   key 4 = blocks of the fused-dot operator (0 default = the resident blocks of the device);
 *   key 5 = fuse p^T A p into the global solve's operator launch (1 default);
 *   key 6 = gathers in flight per lane in the global solve's pipelined operator (4, 6 or 8, 4
 *           default; 0 = the plain fused operator);
 *   key 7 = BF16 direction values per lane of that operator (4 default: 8-byte words; 8);
 *   key 8 = workgroups of the flat PCG vector passes (0 default = 2048);
 *   key 9 = non-temporal FP32 streams (Z, R, AP) in the PCG update and direction passes
 *           (1 default: the caches keep the direction block the next operator gathers);
 *   key 10 = deferred solution updates with 16-bit or FP8 directions (1 default: each captured
 *           group of at most 8 global-solve iterations keeps its directions, one per slot, and
 *           adds sum alpha_s P_s to Z once at its end; 0, or longer groups: Z updated by every
 *           iteration);
 *   key 11 = waves per SIMD the FP8 operator is built for at depth 6 or 8 (6 default; 5);
 *   key 12 = gathers in flight per lane in the FP8 operator (0 default: 8 when at most 4 edge
 *           lanes share a gather instruction, i.e. wide rows, else 4; or 4, 6, 8). */
int dlp_tune_set(int key, int value);

/* ---- Column residual ratios ------------------------------------------------------------
 * rho[c] = max_{listed rows r} |R[r,c]| / diag[r]  (FP64 out[C], overwritten).  R dtype
 * r_dtype, diag FP64.  Replaces solvers.rho_cols (solvers.py:52-54). */
int dlp_rho_cols(int nrows, const int32_t* rows, int C, const void* R, int64_t ldr, int r_dtype,
                 const double* diag, double* rho, void* stream);

/* Rows violating |R[r,c]| > eps[c] * diag[r] for some c, appended to out (order arbitrary,
 * set exact); *count (device int32) receives the number.  Scans the listed rows (or all).
 * Replaces the `violating` filter of solvers.push (solvers.py:105-107, 115-116). */
int dlp_violating_rows(int nrows, const int32_t* rows, int C, const void* R, int64_t ldr,
                       int r_dtype, const double* diag, const double* eps, int32_t* out,
                       int32_t* count, void* stream);

/* ---- Column-wise BLAS-1 of the reference PCG (solvers.py:66-97) ---------------------
 * out[c] = sum_r X[r,c]*Y[r,c] (/ d[r] when d != NULL), FP64 accumulation, out overwritten
 * (the (R*Y).sum(axis=0), (P*AP).sum(axis=0) of solvers.py:76,79,94). */
int dlp_coldot(int n, int C, const void* X, int64_t ldx, const void* Y, int64_t ldy, const void* d,
               double* out, int dtype, void* stream);
/* mode 0: Y += sign*a[c]*X (solvers.py:81-82); mode 1: Y = X + a[c]*Y (:95); mode 2:
 * Y = X / d[r] (jacobi_M, :62-63); mode 3: Y += sign*a[c]*X / d[r] (the weighted-Jacobi
 * correction of the AMG smoother, amg.py:114-117).  a: device FP64[C]. */
int dlp_colop(int n, int C, int mode, double sign, const double* a, const void* X, int64_t ldx,
              const void* d, void* Y, int64_t ldy, int dtype, void* stream);

/* ---- Aggregation AMG (amg.AMG, amg.py:77-134) ------------------------------------------
 * Handshake matching of one level (amg.py:32-64): agg[r] = renumbered aggregate, *nc_dev =
 * count.  choice/flag: int32[n] workspaces, matched: uint8[n]; scan workspace
 * dlp_scan_workspace_bytes(n). */
int dlp_amg_match(int n, const int32_t* indptr, const int32_t* indices, const void* w, int dtype,
                  int rounds, int32_t* agg, int32_t* choice, int32_t* flag, uint8_t* matched,
                  void* scan_ws, size_t scan_bytes, int32_t* nc_dev, void* stream);
/* out[r] = a2[prev[r]] (composition of two matching passes, amg.py:92). */
int dlp_amg_compose(int n, const int32_t* prev, const int32_t* a2, int32_t* out, void* stream);
/* Galerkin coarse operator W_c = offdiag(P^T W P) (amg.py:67-74) in three steps: fill
 * (counting sort by aggregate into segments off[nc+1], entries ci/cw of capacity nnz), merge
 * (per-row sort + duplicate sum; optr[nc+1] row pointers, *total_dev entries), compact. */
int dlp_amg_coarsen_fill(int n, int nc, const int32_t* indptr, const int32_t* indices, const void* w,
                         int dtype, const int32_t* agg, int32_t* off, int32_t* cursor, int32_t* ci,
                         void* cw, void* scan_ws, size_t scan_bytes, void* stream);
int dlp_amg_coarsen_merge(int nc, const int32_t* off, int32_t* ci, void* cw, int dtype, int32_t* optr,
                          void* scan_ws, size_t scan_bytes, int32_t* total_dev, void* stream);
int dlp_amg_coarsen_compact(int nc, const int32_t* off, const int32_t* ci, const void* cw, int dtype,
                            const int32_t* optr, int32_t* oi, void* ow, void* stream);
/* Y[agg[r]] += X[r] (restriction P^T, amg.py:72,129; Y zeroed by the caller) and
 * X[r] += Xc[agg[r]] (prolongation, amg.py:130). */
int dlp_amg_restrict(int n, int C, const int32_t* agg, const void* X, int64_t ldx, void* Y, int64_t ldy,
                     int dtype, void* stream);
int dlp_amg_prolong_add(int n, int C, const int32_t* agg, const void* Xc, int64_t ldc, void* X,
                        int64_t ldx, int dtype, void* stream);
/* One level pass of the V-cycle (amg.py:114-134, A = diag - W on a plain FP32 CSR): mode 0 the
 * residual out = R - A X (the Level residual of the restriction step), mode 1 one weighted-
 * Jacobi sweep out = X + omega (R - A X) / diag (amg.py:114-117's smoother with its residual in
 * the same pass).  FP32 n x C blocks of leading dimension ld, a multiple of 4, 16-byte aligned,
 * n * ld * 4 < 2^31, out a separate block; DLP_ERR_ARG otherwise. */
int dlp_amg_level(int n, int C, const int32_t* indptr, const int32_t* indices, const float* w, const float* diag,
                  const float* X, const float* R, int64_t ld, double omega, int mode, float* out, void* stream);

/* Coarsest AMG level in FP64 on the device (amg.py:104-108, where the reference inverts with
 * cuSOLVER and casts the inverse to the working dtype): A = diag(diag) - W as a dense FP64
 * n x n matrix (n <= 16384); its inverse in place by Gauss-Jordan steps without pivoting (the
 * operator is symmetric positive definite; scratch: another n x n FP64 buffer); and X = A R with
 * FP64 accumulation for an n x C block R (FP32 or FP64 by dtype, X likewise). */
int dlp_dense_operator_f64(int n, const int32_t* indptr, const int32_t* indices, const void* w, const void* diag,
                           int dtype, double* A, void* stream);
int dlp_spd_inverse_f64(int n, double* A, double* scratch, void* stream);
int dlp_dense_apply_f64(int n, int C, const double* A, const void* R, int64_t ldr, void* X, int64_t ldx, int dtype,
                        void* stream);

/* ---- Frontier primitives (FrontierOps, kernels.py:436-492) ----------------------------- */
/* Y[g] = (rhs[f_g] + sum_j w_j X[col_j]) / diag[f_g], changed[g] = any_c |Y[g,c]-X[f_g,c]| > delta
 * for g < nf (f = frontier); X and rhs are n x C, Y is nf x C (contiguous); adds the frontier's
 * edge count to *edges (device uint64, may be NULL).  dtype for w, X, rhs, diag, Y.  Replaces
 * FrontierOps.jacobi (kernels.py:464-474; CUDA kernels.py:31-102, 186-220). */
int dlp_frontier_jacobi(int nf, const int32_t* frontier, int n, int C, const int32_t* indptr,
                        const int32_t* indices, const void* w, const void* X, const void* rhs,
                        const void* diag, double delta, void* Y, uint8_t* changed,
                        unsigned long long* edges, int dtype, double avg_deg, void* stream);

/* R[col_j] += w_j * D[g] (atomic) for every entry of row f_g, and mark[col_j] = 1.
 * Replaces FrontierOps.push (kernels.py:476-486; CUDA kernels.py:105-126, 222-238). */
int dlp_push_scatter(int nf, const int32_t* frontier, int C, const int32_t* indptr,
                     const int32_t* indices, const void* w, const void* D, void* R,
                     uint8_t* mark, unsigned long long* edges, int dtype, double avg_deg,
                     void* stream);

/* mark[col_j] = 1 for every entry of the listed rows.  Replaces FrontierOps.mark_neighbors
 * (kernels.py:488-492; CUDA kernels.py:128-138). */
int dlp_mark_neighbors(int nr, const int32_t* rows, const int32_t* indptr, const int32_t* indices,
                       uint8_t* mark, unsigned long long* edges, void* stream);

/* Indices i < n with mask[i] != 0, in increasing order, to out; *count (device int32).
 * Replaces xp.flatnonzero (solvers.py:115, 227, 238).  Workspace: dlp_scan_workspace_bytes(n). */
int dlp_flatnonzero_u8(int n, const uint8_t* mask, int32_t* out, int32_t* count, void* ws,
                       size_t ws_bytes, void* stream);
size_t dlp_scan_workspace_bytes(int n);

/* ---- Residual push rounds with device-resident lists (solvers.push, FusedPush) ----------
 * `rounds` push rounds on the batch system (n rows, plain CSR, state Z, R: n x C with leading
 * dimension ld, dtype FP32/FP64 for w, diag, Z, R, D): a round takes the residual of every
 * current candidate (D = R/diag, Z += D, R = 0), scatters w*D to the neighbours (atomics) and
 * lists as next candidates the vertices it touched (candidates included) that violate
 * |R_uc| > eps_c * diag_u (eps: device FP64[C]).  counts (device int32[3]): [0] current
 * candidates, [1] and [2] zero between rounds; candidates alternate between cand and next
 * (after an odd number of rounds the current list is in next).  stamp: int32[n] zeros between
 * rounds; touched: int32[n]; D: n x C scratch; edges (device uint64, may be NULL) += entries
 * scattered.  No host synchronisation: the caller polls counts[0].  Replaces the host-driven
 * rounds of solvers.push (solvers.py:100-125) and the FusedPush kernels (kernels.py:287-350). */
int dlp_push_rounds(int n, int C, int64_t ld, int dtype, const int32_t* indptr, const int32_t* indices,
                    const void* w, const void* diag, const double* eps, void* Z, void* R, void* D, int32_t* cand,
                    int32_t* next, int32_t* counts, int32_t* stamp, int32_t* touched, unsigned long long* edges,
                    int rounds, void* stream);

/* ---- Column-wise PCG on a batch system (solvers.pcg, solvers.py:66-97) -------------------
 * Z, R, P, AP (and Y) are n x C with leading dimension ld, dtype FP32/FP64 (with diag, rhs);
 * the operator is diag - W over a (slack, rowlen non-NULL) CSR with weights of w_dtype
 * (FP32 weights with FP64 state allowed), restricted to mask != 0 rows and columns when mask
 * is non-NULL (the W_UU of the handle's solved set); scal: device FP64[3C + 1] = [rz | pAp |
 * r.z of the new direction | violation]; eps: device FP64[C].  dlp_cg_residual: R = rhs -
 * (diag Z - W Z) and *viol = max |R_uc| / (eps_c diag_u) (NaN counts as +inf).
 * dlp_cg_direction: r.z_new = R.Y per column, then P = Y (first != 0) or Y + beta P with beta =
 * r.z_new / rz (0 when rz <= 0), and rz = r.z_new.  dlp_cg_iteration: AP = diag P - W P, pAp =
 * P.AP, alpha = rz/pAp (0 when pAp <= 0), Z += alpha P, R -= alpha AP, violation of the new R in
 * scal[3C].  No host synchronisation. */
int dlp_cg_residual(int n, int C, int64_t ld, int dtype, const int32_t* indptr, const int32_t* rowlen,
                    const int32_t* indices, const void* w, int w_dtype, const uint8_t* mask, double avg_deg,
                    const void* diag, const void* rhs, int64_t ldr, const void* Z, void* R, const double* eps,
                    double* viol, void* stream);
int dlp_cg_direction(int n, int C, int64_t ld, int dtype, const void* R, const void* Y, void* P, double* scal,
                     int first, void* stream);
int dlp_cg_iteration(int n, int C, int64_t ld, int dtype, const int32_t* indptr, const int32_t* rowlen,
                     const int32_t* indices, const void* w, int w_dtype, const uint8_t* mask, double avg_deg,
                     const void* diag, void* Z, void* R, void* P, void* AP, double* scal, const double* eps,
                     void* stream);

/* Rows with solved[r] == 0 take the values of an inert row: Z = (0, .., 0, 1) over C1 columns
 * (h last), R = 0 (R may be NULL).  IncrementalLP._park (solvers.py:365-368). */
int dlp_park_rows(int n, const uint8_t* solved, int C1, int64_t ld, int dtype, void* Z, void* R, void* stream);

/* ---- Connected components (cc.connected_components, cc.py:8-23) ----------------------
 * labels[v] = rank of the minimum vertex id of v's component (int32, bit-exact with the
 * reference's hook/jump + unique); *n_comp (device int32).  Edges (rows[e], cols[e]), any
 * direction, duplicates and self-loops allowed.  Workspace: dlp_cc_workspace_bytes(n). */
int dlp_cc_edges(int n, int64_t m, const int32_t* rows, const int32_t* cols, int32_t* labels,
                 int32_t* n_comp, void* ws, size_t ws_bytes, void* stream);
/* Same over a (slack) CSR; with mask non-NULL only vertices with mask != 0 and the entries
 * between them take part (others become singleton components). */
int dlp_cc_csr(int n, const int32_t* indptr, const int32_t* rowlen, const int32_t* indices,
               const uint8_t* mask, int32_t* labels, int32_t* n_comp, void* ws, size_t ws_bytes,
               void* stream);
size_t dlp_cc_workspace_bytes(int n);
/* dlp_cc_csr with no workspace and no device allocation: its scratch is a fixed device array
 * of the library (one copy per GPU, about 17 MB, held until the process ends), so n is at most
 * dlp_cc_static_max_n() (2^21; larger n: DLP_ERR_ARG before any device work).  Calls no
 * hipMalloc/hipFree and no memset, so it can be captured into a HIP graph once it has run
 * outside capture on that device (the first call there looks up the array's address).  Every
 * call on one GPU shares the array: calls that overlap on one GPU (different streams or
 * threads) race.  No counterpart in the reference; for callers that cannot allocate (graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). */
int dlp_cc_csr_static(int n, const int32_t* indptr, const int32_t* rowlen, const int32_t* indices,
                      const uint8_t* mask, int32_t* labels, int32_t* n_comp, void* stream);
/* The largest n dlp_cc_csr_static takes. */
int dlp_cc_static_max_n(void);


/* Row copies with a precision change (FP32/FP64 either side): gather != 0: dst[i] = src[rows[i]],
 * else dst[rows[i]] = src[i], C values per row, rows int64 (NULL: row i), distinct when scattering.
 * The global-store reads and writes of Solver._warm/_store (solvers.py:176-184). */
int dlp_rows_copy(int nrows, const int64_t* rows, int C, const void* src, int64_t lds, int src_dtype, void* dst,
                  int64_t ldd, int dst_dtype, int gather, void* stream);

/* Supernode averages of the m new vertices newrows[i] (components lab[i] < nc): per vertex
 * S_i[c] = base + sum over its entries j with known[col_j] of w_j X[col_j, c] and T_i = base' +
 * sum of those w_j (known NULL: no neighbour terms); base_mode 1: S base = the label columns of B
 * (n x K; column 1 when K = 2) and T base = B's row sum, 2: S base 0 and T base = B's row sum, 0:
 * none.  Per component S and T add up; out[newrows[i], c] = (T > 0 ? S/T : fallback[c]) + offset
 * (fallback: device, C values).  dtype for w, X, B, fallback, out.  Workspace:
 * dlp_supernode_workspace_bytes(nc, C, dtype).  solvers.supernode_init / init_h / _comp_avg
 * (solvers.py:128-163). */
int dlp_supernode_average(int m, const int32_t* newrows, const int32_t* lab, int nc, const int32_t* indptr,
                          const int32_t* indices, const void* w, const uint8_t* known, int C, const void* X, int64_t ldx,
                          const void* B, int K, int base_mode, const void* fallback, double offset, void* out,
                          int64_t ldo, int dtype, void* ws, size_t ws_bytes, void* stream);
size_t dlp_supernode_workspace_bytes(int nc, int C, int dtype);

/* Frontier rows whose Jacobi value changed (changed[g] != 0) take it, X[frontier[g]] = Y[g]
 * (Y: nf x C contiguous), and mark their neighbours; *nchanged (device) += their count and
 * *edges (device, may be NULL) += their entries.  The update and next-frontier marking of
 * DynLP (solvers.py:228-238). */
int dlp_frontier_commit(int nf, const int32_t* frontier, int C, int64_t ldx, const void* Y, const uint8_t* changed,
                        void* X, const int32_t* indptr, const int32_t* indices, uint8_t* mark, int32_t* nchanged,
                        unsigned long long* edges, int dtype, void* stream);

/* ---- Unbuffered scatters (backend.scatter_add/max/min, backend.py:11-28) ---------------
 * a[idx[i], j] op= v[i * v_stride + j] for i < m, j < cols (a: n_out x cols row-major; v_stride
 * = cols, or 0 for one value v[0] everywhere); op 0 add, 1 max, 2 min; repeated indices all
 * apply (atomics).  dtype DLP_F32/F64/I32/I64 for a and v; idx int32 (idx64 = 0) or int64.
 * Floating max/min let a NaN win (as numpy's maximum.at).  Indices outside [0, n_out) are
 * skipped and counted in *bad (device int32, may be NULL). */
int dlp_scatter(int op, int dtype, int64_t n_out, int64_t cols, void* a, int64_t m, const void* idx, int idx64,
                const void* v, int64_t v_stride, int32_t* bad, void* stream);

/* ---- Supernodes of new vertices (DynLP step 1, stream.py:67-81, cc.py:26-83) -------------
 * Adaptive (Felzenszwalb-Huttenlocher) components of m vertices over the directed edge list
 * (rows[e], cols[e]) with FP32 dissimilarities d[e], in Boruvka rounds: every component takes
 * its lightest outgoing edge (ties: the smallest component across) and merges across it when
 * d <= min(Int(C1) + k/|C1|, Int(C2) + k/|C2|) (Int = largest internal dissimilarity, FP32 as
 * the reference keeps it); when no edge qualifies and ground (FP32 per vertex, >= 0, may be
 * NULL) is given, components whose ground sums to 0 then join across their lightest outgoing
 * edge until none can.  labels (int32[m]) = rank of the component's first numbering as the
 * reference's connected_components gives it; *n_comp (host).  Synchronises once per round.
 * Replaces cc.fh_components and _lightest (cc.py:26-74). */
int dlp_fh_components(int m, int64_t ne, const int32_t* rows, const int32_t* cols, const float* d, double k,
                      const float* ground, int32_t* labels, int32_t* n_comp, void* ws, size_t ws_bytes,
                      void* stream);
size_t dlp_fh_workspace_bytes(int m);
/* s[v] = mean over the entries (lr[e] = v) with w > 0 of (1 - w/wmax), plus 1e-6 (FP32; w FP32 or
 * FP64 by dtype): the local scale of new_components (stream.py:72-74). cnt: FP32[m] scratch. */
int dlp_fh_scale(int m, int64_t ne, const int32_t* lr, const void* w, int dtype, double wmax, float* s, float* cnt,
                 void* stream);
/* d[e] = (1 - w[e]/wmax) / sqrt(sr_e * sc_e) (in w's precision, stored FP32), with sr_e =
 * s_r[ir[e]] when ir is non-NULL, else s_r[e] (likewise sc_e with ic, s_c), and *sum_out = the
 * sum of d (host, FP64, deterministic order; synchronises): the dissimilarities and k = mean of
 * cc.adaptive_components (cc.py:77-83).  Workspace: dlp_fh_dissimilarity_workspace_bytes(). */
int dlp_fh_dissimilarity(int64_t ne, const void* w, int dtype, const float* s_r, const float* s_c, const int32_t* ir,
                         const int32_t* ic, double wmax, float* d, double* sum_out, void* ws, size_t ws_bytes,
                         void* stream);
size_t dlp_fh_dissimilarity_workspace_bytes(void);

/* ---- Resident incremental state of one graph (the service's per-graph handle) ----------
 * Replaces IncrementalStream's in-place batch apply (stream.py:180-346) and IncrementalLP's
 * resident Z/R with changed-row refresh and push (solvers.py:346-434), for edge-level updates
 * on a slack CSR.  All arrays are caller-owned device memory except the handle's scratch.
 * Columns: C1 = C + 1 (label columns then h, the absorption column of solvers.py:10-12). */
typedef struct dlp_lp_desc {
  int32_t n;              /* vertices (global ids; rows outside U are inert) */
  int32_t C1;             /* solved columns including h */
  int32_t K2;             /* 1: binary problem, only label 1 feeds column 0 (stream.py:167) */
  int32_t pad_;
  double eta;             /* dongle weight added to s (stream.py:165) */
  double prior;           /* prior value per label column (eta*prior joins rhs) */
  const int32_t* indptr;  /* slack store: row segments */
  int32_t* rowlen;        /* live entries per row */
  int32_t* col;           /* column of each entry, -1 in unused slots */
  float* w;               /* FP32 weight of each entry (the exact operator weight) */
  const uint8_t* labeled; /* labeled vertices (never solved) */
  const int32_t* y;       /* label of each vertex (read where labeled) */
  uint8_t* solved;        /* U mask: unlabeled and grounded (maintained by dlp_lp_ground) */
  int32_t* labcnt;        /* labeled-neighbour count per vertex (maintained by dlp_lp_apply) */
  float* Z32;             /* n x ld working solution [F, h] */
  float* R32;             /* n x ld kept residual */
  int64_t ld;
  double* diag64;         /* diag = s + rowsum(W_UU), FP64 (inert rows 1) */
  float* diag32;
  double* Z64;            /* FP64 state of the plain per-batch path (may be NULL) */
  double* R64;
} dlp_lp_desc;

int dlp_lp_create(const dlp_lp_desc* desc, void** handle);
int dlp_lp_destroy(void* handle);
/* Switch the handle to another store (after a regrow). */
int dlp_lp_set_store(void* handle, const int32_t* indptr, int32_t* rowlen, int32_t* col, float* w);
/* Stop thresholds: |R_c| <= eps[c]*diag (host array of C1, FP64); the FP32 push uses
 * eps*push_scale.  Synchronises the stream. */
int dlp_lp_set_eps(void* handle, const double* eps, double push_scale, void* stream);
/* New batch: clears the changed-row list, counters and the batch-touched list. */
int dlp_lp_begin_batch(void* handle, void* stream);
/* Deletions (du, dv) then insertions (iu, iv, iw) of undirected pairs (device arrays).
 * Self-loops are skipped; an insertion of a stored pair replaces its weight.  Both stored
 * directions change; endpoints join the changed-row list.  The kernels check every pair before
 * indexing a row by it: a pair with an id outside [0, n), or an insertion whose weight is not
 * finite and positive, is skipped and counted as invalid (dlp_lp_status out[16]).  A pair
 * repeated within the deletions is deleted once (the repeats count as missing); of a pair
 * repeated within the insertions, in either orientation, only the last occurrence is applied,
 * to both directions (the earlier ones are counted, out[17]).  Missing deletions and
 * insertions that found a full row are counted (dlp_lp_status); overflowed insertions are
 * re-inserted by dlp_lp_regrow_commit.  Replaces IncrementalStream._toggle/_set_w
 * (stream.py:202-228) and the weight/row-sum/class-weight updates (stream.py:221-235). */
int dlp_lp_apply(void* handle, int nd, const int32_t* du, const int32_t* dv, int ni,
                 const int32_t* iu, const int32_t* iv, const float* iw, void* stream);
/* A sequence of nops <= 2^19 pair operations (u[s], v[s], w[s]) applied as if one after the
 * other: w[s] == 0 deletes the pair, w[s] > 0 inserts it (or replaces its weight).  Each
 * unordered pair ends as its last operation in the sequence leaves it, so the requests of a
 * group are applied in arrival order by one call (their deletions and insertions interleaved)
 * with no host-side merge.  Invalid operations (ids outside [0, n), a weight that is negative,
 * NaN or infinite) are skipped and counted (out[16]); a deletion that is its pair's last
 * operation and finds no stored edge counts as missing; superseded operations are not counted.
 * Overflowed insertions are re-inserted by dlp_lp_regrow_commit with ni = nops and (u, v, w).
 * Successive calls in one batch apply in order.  Replaces the same part of the reference as
 * dlp_lp_apply, for a group of update batches (stream.py:202-235 applied per batch in order). */
int dlp_lp_apply_ops(void* handle, int nops, const int32_t* u, const int32_t* v, const float* w, void* stream);
/* New row offsets with room to grow (rowlen + max(extra, growth*rowlen)); *total = capacity
 * (host; synchronises). */
int dlp_lp_regrow_plan(void* handle, int extra, double growth, int32_t* new_indptr, int64_t* total,
                       void* stream);
/* Copy into the new store, switch to it, re-insert the last apply's overflowed entries
 * (scratch3: 3 * 2^20 int32 of device memory). */
int dlp_lp_regrow_commit(void* handle, const int32_t* new_indptr, int32_t* new_col, float* new_w,
                         int ni, const int32_t* iu, const int32_t* iv, const float* iw,
                         int32_t* scratch3, void* stream);
/* Grounding: U = unlabeled vertices whose component in the unlabeled subgraph has a vertex
 * with a labeled neighbour (stream.py:148-158, replacing the ball/full ground checks of
 * stream.py:237-294); toggled vertices and their neighbours join the changed-row list. */
int dlp_lp_ground(void* handle, void* stream);
/* Incremental grounding after an apply (du, dv: its deletion arrays; ni its insertion count):
 * when no inert vertex existed, a component can only lose its labeled contact if it contains an
 * endpoint of a deleted edge, so a bounded BFS from every unlabeled deletion endpoint to a
 * vertex with a labeled neighbour decides it exactly; otherwise (or when a BFS exhausts its
 * budget, or insertions meet inert vertices) the full pass of dlp_lp_ground runs, decided on
 * the device; with more than n/16 endpoints (2*nd > n/16) the full pass runs directly.  Replaces the ball check of stream.py:237-273 and the full check :275-280. */
int dlp_lp_ground_incremental(void* handle, int nd, const int32_t* du, const int32_t* dv, int ni,
                              void* stream);
/* out[0] = full pass requested by the last grounding, out[1] = inert unlabeled vertices (after
 * the last full pass), out[2] = grounding calls (host copy; synchronises). */
int dlp_lp_ground_flags(void* handle, int32_t* out, void* stream);
/* Row refresh from scratch: s, diag, rhs and R = rhs - diag*Z + W_UU Z in FP64 for the rows of
 * `which` (0 changed list, 1 rows touched by pushes in this batch, 2 all rows, 3 all solved
 * rows whose equations are unchanged since their last refresh: the verification after a
 * global solve, same result as 2 without re-deriving s and diag); rows outside the solved set
 * are parked with diag 1, R 0 and Z = the indicator of the class column on labeled rows, zeros
 * elsewhere (h 0) -- dlp_lp_scores reports them as F 0, h 1.  prec 0: Z32 -> R32 (rounded),
 * prec 1: Z64 -> R64.  Violators (|R_c| > eps_c*diag) are listed for the next push; column
 * maxima of |R|/diag go to the status.  Replaces solvers.py:404-407 (refresh) and :370-373
 * (full residual). */
int dlp_lp_refresh(void* handle, int which, int prec, void* stream);
/* Scores [F, h] of every vertex in FP64 (n x C1, leading dimension ldo): solved rows from the
 * state (Z64 when attached, else Z32), every other row F = 0, h = 1 -- the layout the oracle's
 * solve_system reports.  The reference reports F on the solved rows (Solver.scores,
 * solvers.py:186-187). */
int dlp_lp_scores(void* handle, double* out, int64_t ldo, void* stream);
/* Asynchronous FP32 push (chaotic relaxation, hardware float atomics) from the current
 * violators until no vertex violates or max_rounds; one kernel per round (the violation test
 * of round t+1 is fused with its pushes), replayed from a captured hipGraph of
 * rounds_per_launch rounds (even) when use_graph != 0.  Synchronises once per launch.  Stops
 * early once this batch's scattered entries exceed budget_edges (> 0), the push budget of
 * solvers.py:283 (DynLP+ then falls back to the global solve).  *converged_out = 1 when no
 * vertex violates.  Replaces FusedPush.run (kernels.py:415-433) as the service's repair step. */
int dlp_lp_push(void* handle, int max_rounds, int rounds_per_launch, int use_graph,
                int64_t budget_edges, int64_t* rounds_out, int32_t* converged_out, void* stream);
/* Global solve: column-wise PCG (Jacobi preconditioner, FP32 vectors, FP64 column dots) on
 * the FP32 state from the kept residual, until max |R|/(eps*diag) <= 1 in the recurrence,
 * a stall of `patience` launches, or max_iter; iterations are replayed from a captured
 * hipGraph of iters_per_launch when use_graph != 0.  Synchronises once per launch.  The
 * caller verifies with an FP64 refresh.  Replaces solvers.pcg (solvers.py:66-97) as used by
 * DynLPPlus._global (solvers.py:268-270). */
int dlp_lp_pcg(void* handle, int max_iter, int iters_per_launch, int use_graph, int patience,
               int64_t* iters_out, double* viol_out, void* stream);
/* dlp_lp_pcg with the FP64 verification pass of every solved row (dlp_lp_refresh mode 3) enqueued
 * behind each launch of iterations, before the host reads their status: a solve that ends after
 * a launch returns with its certificate computed (dlp_lp_status) and one host round trip fewer.
 * A launch that does not end the solve costs one extra pass, which replaces the kept residual by
 * its FP64 value.  The certification step of DynLPPlus._global (solvers.py:322-337) fused with
 * the solve. */
int dlp_lp_pcg_verify(void* handle, int max_iter, int iters_per_launch, int use_graph, int patience,
                      int64_t* iters_out, double* viol_out, void* stream);
/* Block PCG variant of dlp_lp_pcg: one search space for all C1 (<= 80) columns, alpha =
 * (P^T A P)^{-1} (R^T D^{-1} R), beta from the ratio of successive R^T D^{-1} R, the k x k
 * Grams in FP64 on the matrix cores and their Cholesky solves in FP64 (collapsed directions
 * deflated).  The block form of solvers.py:66-97 (the reference uses only the Gram diagonals). */
int dlp_lp_bpcg(void* handle, int max_iter, int iters_per_launch, int use_graph, int patience,
                int64_t* iters_out, double* viol_out, void* stream);
/* G = X^T diag(1/rowdiv) Y (C x C FP64, overwritten; rowdiv may be NULL), C <= 80, on the FP64
 * matrix cores (v_mfma_f64_16x16x4f64). */
int dlp_gram(int n, int C, const void* X, int64_t ldx, const void* Y, int64_t ldy, const void* rowdiv,
             double* G, int dtype, void* stream);
/* S = A^{-1} B for SPD A (k x k FP64, k <= 128) by Cholesky in one workgroup, scaled to unit
 * diagonal; pivots below tol deflate their direction (row of S = 0). */
int dlp_chol_solve(int k, const double* A, const double* B, double* S, double tol, void* stream);
/* Row-local updates of the block PCG (the block form of solvers.py:66-97, its Z/R/P updates
 * :79-80, :93-94 with the k x k step matrices): FP32 n x C blocks of leading dimension ld, S the
 * C x C FP64 step matrix on the device (C <= 80).  mode 0: Z += P S, R -= AP S and viol[0] =
 * max |R_uc| / (eps_c diag_u) (viol cleared first); mode 1: P = Y + P S, Y the preconditioned
 * residual M(R) (Y = NULL: R/diag, the Jacobi preconditioner). */
int dlp_block_update(int n, int C, int mode, float* Z, float* R, float* P, const float* AP, const float* Y,
                     int64_t ld, const double* S, const float* diag, const float* eps, float* viol, void* stream);
/* Takes now every device buffer the global solve needs for captured groups of
 * iters_per_launch iterations (block != 0: the block PCG of dlp_lp_bpcg), with the current
 * direction storage and coarse setting, so that a service can allocate at admission and a
 * request allocates nothing; dlp_lp_pcg / dlp_lp_bpcg reserve what is missing themselves.
 * Each set of buffers is allocated whole or not at all: after a failure (DLP_ERR_HIP) the
 * handle holds what it held before and stays usable.  Synchronises the stream.  No
 * reference counterpart (CuPy allocates inside solvers.pcg, solvers.py:66-97). */
int dlp_lp_reserve(void* handle, int iters_per_launch, int block, void* stream);
/* Device bytes the handle holds in its own buffers (caller-owned arrays not included). */
int dlp_lp_device_bytes(void* handle, int64_t* out);
/* Device bytes a handle for n vertices and C1 columns (row length ld) holds once its global
 * solve is reserved (fp64 != 0: the plain FP64 path, which repairs by push and reserves no
 * global-solve buffers): dirs = direction storage kind of
 * dlp_lp_set_pcg_directions, coarse = aggregates of dlp_lp_set_coarse, block = block PCG.
 * Equals dlp_lp_device_bytes after dlp_lp_create + the same settings + dlp_lp_reserve; -1 on
 * bad sizes.  For admission by memory budget (no reference counterpart). */
int64_t dlp_lp_handle_bytes(int n, int C1, int64_t ld, int fp64, int dirs, int iters_per_launch, int coarse,
                            int block);
/* Cumulative global-solve timing: out[0] iterations, out[1] microseconds (HIP events). */
int dlp_lp_pcg_stats(void* handle, int64_t* out);
/* Average duration in ms (HIP events on `stream`, synchronises) of the global solve's operator
 * launch AP = diag*P - W_UU P on the live state, over reps launches (after dlp_lp_pcg). */
int dlp_lp_time_operator(void* handle, int reps, double* ms_out, void* stream);
/* Storage of the column PCG directions: kind 0 = FP32 (default), 1 = BF16, 2 = FP16 held
 * scaled by a power of two per column (chosen from r.z so the values sit near 1; saturating
 * stores), 3 = FP8 E4M3 held scaled by a power of two per column (chosen from a bound on the
 * new direction's largest magnitude, max |r|/diag + max |coarse correction| + |beta| max |P|,
 * so every value stays within 224; saturating stores), rows of ld bytes.  Z, R and A*P stay
 * FP32 with FP64 column dots; the step length is r.p / p.Ap for the direction as held, so the
 * residual recurrence stays that of the FP32 path and the FP64 verification certifies as
 * before.  16-bit directions halve the bytes of every neighbour-row gather of the global
 * solve, FP8 quarters them; FP16 keeps 11 significant bits (BF16: 8, FP8: 4), and coarser
 * directions can cost iterations, never accuracy (the Python layer's default picks FP8 from 48
 * columns up, FP16 below).  No reference counterpart (solvers.py:66-97 keeps P in the working
 * dtype). */
int dlp_lp_set_pcg_directions(void* handle, int kind);
/* Host-side check of an edge batch before it is applied (no device work): ids (id_bytes 4 =
 * int32, 8 = int64) in [0, n), insertion weights finite and > 0.  Returns 0 (valid, no pair
 * repeated), a mask of 1 (an unordered pair repeats within the deletions) and 2 (within the
 * insertions, either orientation) -- the caller collapses repeats before dlp_lp_apply, which
 * takes distinct pairs -- or -1 (an id out of range; *where = its position in del_u, del_v,
 * ins_u, ins_v taken one after the other), -2 (a bad weight; *where = its index), -3 (bad
 * arguments).  The batch contract of the reference's update path (stream.py:180-230 builds
 * each batch's system from distinct edges; check_batch / normalize_batch in graphs.py). */
int dlp_check_edge_batch(int64_t n, int64_t nd, const void* del_u, const void* del_v, int64_t ni,
                         const void* ins_u, const void* ins_v, const float* ins_w, int id_bytes,
                         int64_t* where);
/* Two-level preconditioner of the column PCG, M = D^-1 + P Ac^-1 P^T: P aggregates the solved
 * rows into up to na (<= 128) runs of consecutive rows (at least the rows one workgroup of the vector
 * passes spans; each of the na evenly spaced boundaries moves to the row boundary with the fewest
 * crossing edges within half a spacing), Ac = P^T A P is factored in FP64 and inverted on the
 * device, and rebuilt from
 * the current equations every `refresh` global solves (a stale Ac is still an SPD preconditioner,
 * only a weaker one; 0 aggregates: Jacobi only, the default of a new handle).  The coarse correction removes the slow, nearly constant modes of the
 * label-grounded system (12 -> 6-7 iterations at SBM-1M, k = 64).  No reference counterpart:
 * solvers.pcg takes any SPD preconditioner M (solvers.py:66-97; jacobi_M :62-63, AMG amg.py). */
int dlp_lp_set_coarse(void* handle, int na, int refresh);
/* The current coarse operator: out[0] = aggregates (0: none built), out[1] = their minimum rows,
 * out[2] = global solves since it was built; bounds_out (host, na + 1 ints, may be NULL) = the
 * row ranges [bounds[I], bounds[I+1]) (synchronises).  No reference counterpart. */
int dlp_lp_coarse_info(void* handle, int32_t* out, int32_t* bounds_out, void* stream);
/* One push across several resident graphs (the service's fused repair step; no reference
 * counterpart: the reference repairs one system at a time, kernels.py:396-433).  A plan
 * (dlp_push_plan_create, which allocates its own small device arrays; destroy with
 * dlp_push_plan_destroy) groups the handles' push descriptors by column class; then
 * dlp_lp_push_multi runs the asynchronous push of dlp_lp_push for every handle i with
 * active[i] != 0, starting from that handle's last refresh's violators, as one launch per
 * round and column class covering all of them (grid y = graph), replayed from one captured
 * hipGraph of rounds_per_launch (even) rounds when use_graph != 0.  A graph leaves the launch
 * when it converges (converged_out[i] = 1), when its scattered entries in this batch exceed
 * budget_edges[i] (> 0; NULL: no budget) or after max_rounds; rounds_out[i] = rounds launched
 * while it was active.  Host arrays of the plan's size; synchronises once per launch.  Calls
 * on one plan, or on handles of an active plan, must not overlap. */
int dlp_push_plan_create(void* const* handles, int nh, void** plan);
int dlp_push_plan_destroy(void* plan);
int dlp_lp_push_multi(void* plan, const int32_t* active, int max_rounds, int rounds_per_launch, int use_graph,
                      const int64_t* budget_edges, int64_t* rounds_out, int32_t* converged_out, void* stream);
/* Plain FP64 path, one host-driven round of solvers.push (solvers.py:108-116): candidates
 * take D = R/diag, Z += D, R = 0, scatter W D with FP64 atomics and mark N(cand) and cand. */
int dlp_lp_push64_round(void* handle, int nc, const int32_t* cand, void* stream);
/* Marked vertices (ascending) that violate in FP64 -> cand_out, count (device); marks cleared. */
int dlp_lp_push64_next(void* handle, int32_t* cand_out, int32_t* count_out_dev, void* stream);
/* Device list and count of the last refresh's violators. */
int dlp_lp_violators(void* handle, int32_t** list, int32_t** count);
/* Counters (host; synchronises): [0] changed rows, [1] missing deletions, [2] overflowed
 * insertions, [3] toggled vertices, [4] violators, [5] FP64-path candidates, [6] rows touched
 * by pushes, [7] round number, [8] edge visits of this batch, cumulative push work [9]
 * candidates, [10] pushes, [11] entries scattered by pushes, [12] the bits of an FP64 h
 * maximum: max over the solved rows of h (column C1-1) from the last refresh when it was a
 * verification of every solved row (mode 3 on the 16-byte FP32 layout), else 0,
 * [13] push rounds launched, [14] their summed duration in microseconds (HIP events),
 * [15] the bits of an FP64 "push mass": over the rows the last changed-row refresh found
 * violating, the sum of max_c |R_c|/(eps_c*diag) (about the row pushes a repair by push alone
 * would take), [16] pairs of this batch refused by dlp_lp_apply (ids outside [0, n) or
 * insertion weights not finite and positive), [17] earlier occurrences of pairs repeated in its
 * insertions (not applied), [18] the kernel of the last dlp_lp_refresh (0 the generic refresh,
 * 1 the 16-byte-row refresh, 2 the pipelined verification of mode 3, 3 the pipelined refresh of
 * listed rows of modes 0 and 1; -1 none yet);
 * rho (C1, may be NULL): column maxima of the last refresh, in which a NaN residual counts as
 * +inf (so it never certifies).  `out` holds 19 entries. */
int dlp_lp_status(void* handle, int64_t* out, double* rho, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* DYNLP_H */

"""Label-propagation solvers for one batch system, on the device.

The public surface is the reference's (dynlp/solvers.py): `Stats`, `spmm`, `apply_A`,
`rho_cols`, `violation`, `jacobi_M`, `pcg` (:66-97; `block_pcg` is its block form),
`push` (:100-125), `supernode_init` / `init_h` (:128-163), `predict`, `ItLP` (:190-210), `DynLP` (:213-241), `DynLPPlus`
(:244-343), `IncrementalLP` (:346-434), `Reference` (:437-445), `make` (:448-458) and
`SOLVERS`.  Behind it the work is device work:

- the column PCG keeps its recurrences on the device: a `_ColumnCG` workspace holds the
  direction, its product and the FP64 column scalars (r.z, pAp, violation), and each iteration
  is three C-ABI passes (operator, column dots, fused update with the violation reduced in the
  same pass); the host reads one FP64 violation per iteration;
- the residual push runs rounds back to back on device-resident candidate lists
  (kernels.FusedPush, dlp_push_rounds) and the host polls the candidate count every few rounds;
- DynLP+ is a pipeline of stages (start, h bootstrap, working-precision repair, FP64
  certification with scaled refinement), whose only host arithmetic is on C-long vectors of
  column maxima read back once per stage;
- the incremental solver parks, refreshes and repairs its resident FP64 state with the same
  kernels (dlp_park_rows, the row operator, the push rounds).

Torch tensors are only buffers here (allocation, views, dtype copies); every sum, product,
scatter and compaction is a HIP kernel of the C ABI.  A `System` holds the batch system on the
active unlabeled set U in local ids (stream.py:17-58).
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np
import torch

from . import _lib
from .backend import CSR, Timer, device, dtype_code, ptr, stream
from .cc import connected_components
from .kernels import FrontierOps, FusedPush, spmm_axpy


@dataclass
class Stats:
    iters: int = 0
    edges: int = 0              # edge visits: hardware-independent work
    ms: float = 0.0
    setup_ms: float = 0.0
    path: str = ""
    cert: float = float("nan")
    hmax: float = float("nan")


@dataclass
class System:
    """Batch system on U (stream.py:17-58); device tensors, local row ids."""
    U: torch.Tensor           # global ids of the rows (sorted, int64)
    W: CSR
    s: torch.Tensor
    diag: torch.Tensor
    B: torch.Tensor           # n_u x K class weights to labeled neighbours
    rhs: torch.Tensor         # n_u x C
    prior: torch.Tensor       # C
    K: int
    C: int
    eta: float = 0.0
    is_new: torch.Tensor | None = None
    seed: torch.Tensor | None = None
    t: int = 0                # batch number of a stream
    comps: dict | None = None  # new vertices' supernode components {"tau"|"fh": (labels, count)}
    n_inserted: int = 0
    n_deleted: int = 0
    n_active: int = 0
    n_ungrounded: int = 0
    solved: torch.Tensor | None = None    # global-index systems: rows really solved (others inert)
    changed: torch.Tensor | None = None   # global-index systems: rows whose equation changed
    _conv: dict = field(default_factory=dict, repr=False, compare=False)

    n_u = property(lambda self: int(self.U.shape[0]))
    nnz = property(lambda self: self.W.nnz)
    n_solved = property(lambda self: self.n_u if self.solved is None else int(self.solved.sum()))

    def label_cols(self, M):
        return M[:, 1:2] if self.K == 2 else M

    def astype(self, dt) -> "System":
        """Same system in another precision (cached; index arrays shared) (stream.py:49-58)."""
        dt = torch.float64 if str(dt) in ("float64", "torch.float64") else torch.float32
        if dt == self.W.data.dtype:
            return self
        if dt not in self._conv:
            self._conv[dt] = replace(self, W=self.W.astype(dt), s=self.s.to(dt), diag=self.diag.to(dt),
                                     B=self.B.to(dt), rhs=self.rhs.to(dt), prior=self.prior.to(dt), _conv={})
        return self._conv[dt]


def build_system(g, dtype="float64", eta=0.0, active=None, labeled=None) -> System:
    """The batch system of the graph's current edge set (Stream._build, stream.py:142-177): U =
    active, unlabeled vertices whose component of the unlabeled subgraph touches a labeled one
    (device CC); W = the graph restricted to U; s = weight to labeled vertices + eta; diag = s +
    row sums of W; rhs = class weights (+ eta * prior).  The edge list is ingested on the host
    (FP64 sums) and uploaded once."""
    dt = torch.float64 if dtype in ("float64", torch.float64) else torch.float32
    n, K = g.n, g.K
    C = 1 if K == 2 else K
    act = np.ones(n, dtype=bool) if active is None else np.asarray(active, dtype=bool)
    lab = (g.labeled if labeled is None else np.asarray(labeled, dtype=bool)) & act
    unl = act & ~(g.labeled if labeled is None else np.asarray(labeled, dtype=bool))
    u, v, w = g.u, g.v, g.w.astype(np.float64)
    # class weights of every vertex to its labeled neighbours (both edge directions)
    classw = np.zeros((n, K))
    for a, b in ((u, v), (v, u)):
        hit = lab[b] & unl[a]
        np.add.at(classw, (a[hit], g.y[b[hit]]), w[hit])
    # grounding: components of the unlabeled subgraph with any labeled contact
    both = unl[u] & unl[v]
    comp, ncomp = connected_components(n, np.r_[u[both], v[both]], np.r_[v[both], u[both]])
    comp = comp.cpu().numpy()
    contact = np.zeros(ncomp, dtype=bool)
    contact[comp[unl & (classw.sum(axis=1) > 0)]] = True
    inU = unl & contact[comp]
    U = np.flatnonzero(inU)
    n_u = U.shape[0]
    local = np.full(n, -1, dtype=np.int64)
    local[U] = np.arange(n_u)
    kept = inU[u] & inU[v]
    rr = np.concatenate([local[u[kept]], local[v[kept]]])
    cc = np.concatenate([local[v[kept]], local[u[kept]]])
    ww = np.concatenate([w[kept], w[kept]])
    order = np.lexsort((cc, rr))
    rr, cc, ww = rr[order], cc[order], ww[order]
    indptr = np.zeros(n_u + 1, dtype=np.int64)
    np.cumsum(np.bincount(rr, minlength=n_u), out=indptr[1:])
    Bm = classw[U]
    s = Bm.sum(axis=1) + eta
    diag = s + np.bincount(rr, weights=ww, minlength=n_u)
    prior = np.full(C, 0.5 if K == 2 else 1.0 / K)
    rhs = (Bm[:, 1:2] if K == 2 else Bm) + eta * prior[None, :]
    dev = device()
    up = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev).to(dt)  # noqa: E731
    W = CSR(up(ww), torch.as_tensor(cc.astype(np.int32), device=dev),
            torch.as_tensor(indptr.astype(np.int32), device=dev), (n_u, n_u))
    every = torch.ones(n_u, dtype=torch.bool, device=dev)
    return System(torch.as_tensor(U, device=dev), W, up(s), up(diag), up(Bm), up(rhs), up(prior), K, C, eta,
                  is_new=every, seed=every)


# ---------------------------------------------------------------------------------------------
# Small device helpers.

def _f64(values):
    """Device FP64 vector of host values."""
    return torch.as_tensor(np.array(values, dtype=np.float64).reshape(-1), device=device())


_ONES = {}


def _ones64(n):
    t = _ONES.get(n)
    if t is None:
        t = _ONES[n] = torch.ones(max(n, 1), dtype=torch.float64, device=device())
    return t


def _col_max(X, c):
    """max_r X[r, c] of a non-negative column (one column-maxima pass)."""
    out = torch.zeros(1, dtype=torch.float64, device=X.device)
    if X.shape[0]:
        _lib.call("dlp_rho_cols", int(X.shape[0]), None, 1, ptr(X) + c * X.element_size(), X.stride(0),
                  dtype_code(X.dtype), ptr(_ones64(int(X.shape[0]))), ptr(out), stream())
    return float(out.item())


def _axpy_cols(a, X, Y, sign=1.0):
    """Y += sign * a[c] * X column-wise (dlp_colop mode 0); a on the host."""
    if X.shape[0]:
        _lib.call("dlp_colop", int(X.shape[0]), int(X.shape[1]), 0, float(sign), ptr(_f64(a)), ptr(X), X.stride(0),
                  None, ptr(Y), Y.stride(0), dtype_code(X.dtype), stream())
    return Y


def _count_violating(sys, R, eps):
    """Rows of R with some |R_uc| > eps_c * diag_u (dlp_violating_rows; the list is discarded)."""
    n = int(R.shape[0])
    if not n:
        return 0
    out = torch.empty(n, dtype=torch.int32, device=R.device)
    cnt = torch.zeros(1, dtype=torch.int32, device=R.device)
    d = sys.diag.double().contiguous()
    _lib.call("dlp_violating_rows", n, None, int(R.shape[1]), ptr(R), R.stride(0), dtype_code(R.dtype), ptr(d),
              ptr(_f64(eps)), ptr(out), ptr(cnt), stream())
    return int(cnt.item())


def _nonzero_rows(mask):
    """Indices (device int32, ascending) of the non-zero entries of a bool/uint8 vector."""
    n = int(mask.shape[0])
    m = mask.to(torch.uint8).contiguous()
    out = torch.empty(max(n, 1), dtype=torch.int32, device=m.device)
    cnt = torch.zeros(1, dtype=torch.int32, device=m.device)
    nb = int(_lib.load().dlp_scan_workspace_bytes(max(n, 1)))
    ws = torch.empty(nb, dtype=torch.uint8, device=m.device)
    _lib.call("dlp_flatnonzero_u8", n, ptr(m), ptr(out), ptr(cnt), ptr(ws), nb, stream())
    return out[: int(cnt.item())]


# ---------------------------------------------------------------------------------------------
# The reference's BLAS surface.

def spmm(W: CSR, X):
    """W @ X for row-major X (solvers.py:40-43)."""
    return spmm_axpy(W, X.contiguous())


def apply_A(sys: System, X):
    """diag * X - W X in one pass (solvers.py:46-49)."""
    return spmm_axpy(sys.W, X.contiguous(), a=sys.diag, b=-1.0)


def rho_cols(sys: System, R):
    """max_u |R_uc| / diag_u per column (solvers.py:52-54); FP64 device vector."""
    C = R.shape[1]
    out = torch.zeros(C, dtype=torch.float64, device=R.device)
    if R.shape[0]:
        d = sys.diag.double().contiguous()
        _lib.call("dlp_rho_cols", int(R.shape[0]), None, C, ptr(R.contiguous()), C, dtype_code(R.dtype), ptr(d),
                  ptr(out), stream())
    return out


def violation(sys: System, R, eps_d):
    """max |R_uc| / (eps_c * diag_u) (solvers.py:57-59): column maxima on the device, their C
    ratios on the host."""
    if not R.shape[0]:
        return 0.0
    rc = rho_cols(sys, R).cpu().numpy()
    return float((rc / np.broadcast_to(np.asarray(eps_d, dtype=np.float64), rc.shape)).max())


def jacobi_M(sys: System):
    """M(R) = R / diag (solvers.py:62-63)."""
    def apply(R):
        R = R.contiguous()
        Y = torch.empty_like(R)
        if R.shape[0]:
            _lib.call("dlp_colop", int(R.shape[0]), int(R.shape[1]), 2, 1.0, None, ptr(R), R.shape[1],
                      ptr(sys.diag.to(R.dtype).contiguous()), ptr(Y), R.shape[1], dtype_code(R.dtype), stream())
        return Y
    return apply


def predict(F, K):
    """solvers.py:166-167 (the decision of every row, int32)."""
    a = F.detach().cpu().numpy()
    lab = (a[:, 0] > 0.5) if K == 2 else a.argmax(axis=1)
    return torch.as_tensor(lab.astype(np.int32), device=F.device)


# ---------------------------------------------------------------------------------------------
# Column PCG on the device.

class _ColumnCG:
    """Device workspace of one column-wise PCG solve of A Z = rhs: R, P, AP, the FP64 scalars
    [rz | pAp | r.z_new | violation] and the thresholds.  `residual`, `direction` and `iterate`
    are single C-ABI calls; `violation` reads one FP64 value.  The operator is diag - W of a
    batch system, or (`operator=`) any (indptr, rowlen, indices, w, mask, avg_deg, diag) tuple,
    e.g. a resident graph's slack store restricted to its solved set."""

    def __init__(self, sys: System, Z, rhs, eps, operator=None):
        self.Z, self.rhs = Z, rhs
        self.n, self.C = int(Z.shape[0]), int(Z.shape[1])
        self.ld = int(Z.stride(0))
        self.dt = dtype_code(Z.dtype)
        block = lambda: torch.empty((self.n, self.ld), dtype=Z.dtype, device=Z.device)[:, : self.C]  # noqa: E731
        self.R, self.P, self.AP = block(), block(), block()
        self.scal = torch.zeros(3 * self.C + 1, dtype=torch.float64, device=Z.device)
        self.eps = _f64(np.broadcast_to(np.asarray(eps, dtype=np.float64), (self.C,)))
        if operator is None:
            W = sys.W
            operator = (W.indptr, None, W.indices, W.data, None, float(W.avg_deg), sys.diag)
        ip, rl, ix, w, mask, deg, d = operator
        self._op = (ptr(ip), ptr(rl), ptr(ix), ptr(w), dtype_code(w.dtype), ptr(mask), float(deg), ptr(d))

    def residual(self):
        """R = rhs - A Z; returns its violation max |R|/(eps diag)."""
        ip, rl, ix, w, wdt, mask, deg, d = self._op
        _lib.call("dlp_cg_residual", self.n, self.C, self.ld, self.dt, ip, rl, ix, w, wdt, mask, deg, d,
                  ptr(self.rhs), self.rhs.stride(0), ptr(self.Z), ptr(self.R), ptr(self.eps),
                  ptr(self.scal) + 24 * self.C, stream())
        return self.violation()

    def direction(self, Y, first):
        _lib.call("dlp_cg_direction", self.n, self.C, self.ld, self.dt, ptr(self.R), ptr(Y), ptr(self.P),
                  ptr(self.scal), int(first), stream())

    def iterate(self):
        ip, rl, ix, w, wdt, mask, deg, d = self._op
        _lib.call("dlp_cg_iteration", self.n, self.C, self.ld, self.dt, ip, rl, ix, w, wdt, mask, deg, d, ptr(self.Z),
                  ptr(self.R), ptr(self.P), ptr(self.AP), ptr(self.scal), ptr(self.eps), stream())

    def violation(self):
        return float(self.scal[3 * self.C].item())


class _Progress:
    """Stall detection: a check that fails to beat the best value by `gain` counts against
    `patience` consecutive checks."""

    def __init__(self, first, gain, patience):
        self.best, self.gain, self.patience, self.flat = first, gain, patience, 0

    def stalled(self, v):
        if v < self.gain * self.best:
            self.best, self.flat = v, 0
        else:
            self.flat += 1
        return self.flat >= self.patience


def cg_solve(cg: _ColumnCG, M, max_iter, st: Stats, nnz, work_M=0, patience=400):
    """The column PCG loop on a prepared workspace: to max |R_uc|/(eps_c diag_u) <= 1, the true
    residual confirming convergence, a stall exit after `patience` iterations without a 10 %
    gain.  M(R) returns the preconditioned residual.  Updates cg.Z in place."""
    st.edges += nnz
    v = cg.residual()
    if v <= 1:
        return
    cg.direction(M(cg.R), first=True)
    track = _Progress(v, 0.9, patience)
    for _ in range(max_iter):
        cg.iterate()
        st.iters, st.edges = st.iters + 1, st.edges + nnz + work_M
        v = cg.violation()
        if v <= 1:
            st.edges += nnz
            v = cg.residual()              # the recurrence converged: check the true residual
            if v <= 1:
                break
        if track.stalled(v):
            break
        cg.direction(M(cg.R), first=False)


def pcg(sys: System, Z, rhs, eps, M, max_iter, st: Stats, work_M=0, patience=400):
    """Column-wise preconditioned CG (solvers.py:66-97) of A Z = rhs on a batch system, on the
    device (`cg_solve`).  Z (n_u x C, row-major) is updated in place and returned."""
    Z = Z.contiguous()
    cg_solve(_ColumnCG(sys, Z, rhs.contiguous(), eps), M, max_iter, st, sys.nnz, work_M, patience)
    return Z


def block_pcg(sys: System, Z, rhs, eps, M, max_iter, st: Stats, work_M=0, patience=400):
    """Block PCG of A Z = rhs: one search space for all C columns (the block form of
    solvers.py:66-97), the k x k Grams G = R^T M(R) and P^T A P on the FP64 matrix cores
    (dlp_gram), their Cholesky solves in FP64 (dlp_chol_solve; collapsed directions deflate),
    the row-local updates Z += P S, R -= AP S, P = M(R) + P B on FP32 blocks (dlp_block_update).
    Same stop tests as `pcg`.  FP32 blocks of at most 80 columns; anything else goes to `pcg`."""
    Z = Z.contiguous()
    if Z.dtype != torch.float32 or not 1 <= Z.shape[1] <= 80:
        return pcg(sys, Z, rhs, eps, M, max_iter, st, work_M, patience)
    cg = _ColumnCG(sys, Z, rhs.contiguous(), eps)     # R, P and the true-residual pass
    st.edges += sys.nnz
    v = cg.residual()
    if v <= 1:
        return Z
    n, C, ld = cg.n, cg.C, cg.ld
    dev = Z.device
    mats = torch.zeros((4, C, C), dtype=torch.float64, device=dev)   # G, P^T A P, G_new, step
    G, PAP, Gn, S = mats[0], mats[1], mats[2], mats[3]
    viol = torch.zeros(1, dtype=torch.float32, device=dev)
    diag = sys.diag.to(torch.float32).contiguous()
    eps32 = cg.eps.to(torch.float32)
    f32 = _lib.F32

    def gram(X, ldx, Y, ldy, out):
        _lib.call("dlp_gram", n, C, ptr(X), ldx, ptr(Y), ldy, None, ptr(out), f32, stream())

    Y = M(cg.R).contiguous()
    cg.P.copy_(Y)
    gram(cg.R, ld, Y, C, G)
    track = _Progress(v, 0.9, patience)
    for _ in range(max_iter):
        AP = apply_A(sys, cg.P)
        gram(cg.P, ld, AP, C, PAP)
        _lib.call("dlp_chol_solve", C, ptr(PAP), ptr(G), ptr(S), 1e-8, stream())
        _lib.call("dlp_block_update", n, C, 0, ptr(cg.Z), ptr(cg.R), ptr(cg.P), ptr(AP), None, ld, ptr(S), ptr(diag),
                  ptr(eps32), ptr(viol), stream())
        st.iters, st.edges = st.iters + 1, st.edges + sys.nnz + work_M
        v = float(viol.item())
        if v <= 1:
            st.edges += sys.nnz
            v = cg.residual()              # the recurrence converged: check the true residual
            if v <= 1:
                break
        if track.stalled(v):
            break
        Y = M(cg.R).contiguous()
        gram(cg.R, ld, Y, C, Gn)
        _lib.call("dlp_chol_solve", C, ptr(G), ptr(Gn), ptr(S), 1e-8, stream())
        _lib.call("dlp_block_update", n, C, 1, None, ptr(cg.R), ptr(cg.P), None, ptr(Y), ld, ptr(S), ptr(diag), None,
                  None, stream())
        G.copy_(Gn)
    return Z


def push(sys: System, Z, R, eps, ops: FrontierOps, st: Stats, budget_edges=None, max_iter=10**7):
    """Residual push on the violating set (solvers.py:100-125) in device rounds (FusedPush):
    candidates take D = R/diag into Z and scatter W D into R, the next candidates are the
    touched rows that still violate.  Every 32 rounds the edge budget and the progress of the
    violation are checked (8 checks without a 1 % gain stop it).  R stays exact up to rounding.
    Returns (Z, R, converged)."""
    n, C = int(R.shape[0]), int(R.shape[1])
    if not n:
        return Z, R, True
    rounds = FusedPush(sys.W, C)
    diag = sys.diag.to(R.dtype).contiguous()
    rounds.seed(R, diag, np.broadcast_to(np.asarray(eps, dtype=np.float64), (C,)).copy())
    start = int(ops.edges.item())
    track = _Progress(float("inf"), 0.99, 8)
    done, left = 0, rounds.pending()
    while left and done < max_iter:
        step = int(min(32, max_iter - done))
        rounds.launch(Z, R, diag, step, edges=ops.edges)
        done += step
        left = rounds.pending()
        if left and budget_edges is not None and int(ops.edges.item()) - start > budget_edges:
            break
        if left and track.stalled(violation(sys, R, eps)):
            break
    st.iters += done
    return Z, R, left == 0


# ---------------------------------------------------------------------------------------------
# Supernode starts of new rows.

def _supernode_fill(sys, X, col0, C, known, base_mode, fallback, offset, comp):
    new = _nonzero_rows(sys.is_new)
    m = int(new.shape[0])
    if not m:
        return
    lab, nc = sys.comps[comp]
    lab = lab.to(torch.int32).contiguous()
    W = sys.W
    dt = dtype_code(X.dtype)
    nb = int(_lib.load().dlp_supernode_workspace_bytes(max(int(nc), 1), C, dt))
    ws = torch.empty(nb, dtype=torch.uint8, device=X.device)
    kmask = None if known is None else known.to(torch.uint8).contiguous()
    B = sys.B.to(X.dtype).contiguous()
    fb = torch.as_tensor(np.broadcast_to(np.asarray(fallback, dtype=np.float64), (C,)).copy(), device=X.device)
    fb = fb.to(X.dtype)
    base = ptr(X) + col0 * X.element_size()
    _lib.call("dlp_supernode_average", m, ptr(new), ptr(lab), int(nc), ptr(W.indptr), ptr(W.indices),
              ptr(W.data.to(X.dtype)), ptr(kmask), C, base, X.stride(0), ptr(B), int(sys.K), base_mode, ptr(fb),
              float(offset), base, X.stride(0), dt, ptr(ws), nb, stream())


def supernode_init(sys: System, X, known=None, cols=None, comp="tau"):
    """New rows take their supernode's value (solvers.py:138-153): with known None, the share
    of the component's weight to each class (DynLP's rule); with a known-row mask, the weighted
    average over the component of label weights and known neighbours' values.  Components from
    sys.comps[comp]; without them (or without new rows) X is left as it is.  On the device
    (dlp_supernode_average), in place on X[:, cols]."""
    if sys.is_new is None or sys.comps is None or comp not in sys.comps:
        return X
    c0 = 0 if cols is None else (cols.start or 0)
    C = sys.C if cols is None else (cols.stop - c0)
    prior = sys.prior.double().cpu().numpy()
    _supernode_fill(sys, X, c0, C, known, 1, prior, 0.0, comp)
    return X


def init_h(sys: System, H, known, fallback, comp="fh"):
    """New rows: h = 1 + the component's weighted average of h over known neighbours, labeled
    neighbours counting h = 0 (solvers.py:156-163); fallback where the component has no such
    weight.  On the device, in place on H[:, 0]."""
    if sys.is_new is None or sys.comps is None or comp not in sys.comps:
        return H
    _supernode_fill(sys, H, 0, 1, known, 2, [float(fallback) - 1.0], 1.0, comp)
    return H


# ---------------------------------------------------------------------------------------------
# Solvers.

class Solver:
    """Holds the scores of every vertex seen so far (n_total x C, store precision); a batch
    system reads and writes its rows through dlp_rows_copy."""
    name = "base"

    def __init__(self, n_total, K, dtype="float32", store_dtype=None):
        self.n_total, self.K = n_total, K
        self.dtype = torch.float64 if dtype == "float64" else torch.float32
        sd = store_dtype or dtype
        self.store_dtype = torch.float64 if sd == "float64" else torch.float32
        self.C, self.F = 1 if K == 2 else K, None

    def _rows(self, sys, store, work_dtype):
        out = torch.empty((sys.n_u, store.shape[1]), dtype=work_dtype, device=store.device)
        _lib.call("dlp_rows_copy", sys.n_u, ptr(sys.U), int(store.shape[1]), ptr(store), store.stride(0),
                  dtype_code(store.dtype), ptr(out), out.stride(0), dtype_code(work_dtype), 1, stream())
        return out

    def _put_rows(self, sys, store, X):
        X = X.contiguous()
        _lib.call("dlp_rows_copy", sys.n_u, ptr(sys.U), int(store.shape[1]), ptr(X), X.stride(0), dtype_code(X.dtype),
                  ptr(store), store.stride(0), dtype_code(store.dtype), 0, stream())

    def _warm(self, sys: System):
        if self.F is None:
            self.F = torch.empty((self.n_total, self.C), dtype=self.store_dtype, device=device())
            self.F.copy_(sys.prior.to(self.store_dtype).expand(self.n_total, self.C))
        return self._rows(sys, self.F, self.dtype)

    def _store(self, sys: System, X):
        self._put_rows(sys, self.F, X)

    def scores(self, sys: System):
        return self._rows(sys, self.F, self.store_dtype)


class ItLP(Solver):
    """Jacobi from the prior on every row until no value moves by more than delta
    (solvers.py:190-210): each sweep is the frontier-Jacobi kernel over all rows."""
    name = "itlp"

    def __init__(self, n_total, K, dtype="float32", delta=1e-4, max_iter=200_000):
        super().__init__(n_total, K, dtype)
        self.delta, self.max_iter = delta, max_iter

    def solve(self, sys):
        st, sw = Stats(path="jacobi"), sys.astype(self.dtype)
        with Timer() as tm:
            X = self._warm(sw)
            X.copy_(sw.prior.expand_as(X))
            ops = FrontierOps(sw.W, self.C)
            every = torch.arange(sw.n_u, dtype=torch.int32, device=X.device)
            while sw.n_u and st.iters < self.max_iter:
                Y, moved = ops.jacobi(every, X, sw.rhs, sw.diag, self.delta)
                X, st.iters = Y, st.iters + 1
                if not int(_nonzero_rows(moved).shape[0]):
                    break
            self._store(sw, X)
        st.edges, st.ms = int(ops.edges), tm.ms
        return st


class DynLP(Solver):
    """DynLP Algorithm 2 (solvers.py:213-241): supernode start of the new rows, then Jacobi on a
    frontier that starts at the seed rows; rows whose value moves by more than delta take it
    and put their neighbours on the next frontier (dlp_frontier_commit)."""

    def __init__(self, n_total, K, dtype="float32", delta=1e-4, group="auto", comp="tau", max_iter=10**7):
        super().__init__(n_total, K, dtype)
        self.delta, self.group, self.comp, self.max_iter = delta, group, comp, max_iter
        self.name = "dynlp" if comp == "tau" else f"dynlp-{comp}"

    def solve(self, sys):
        st, sw = Stats(path="frontier-jacobi"), sys.astype(self.dtype)
        with Timer() as tm:
            X = supernode_init(sw, self._warm(sw), comp=self.comp)
            ops = FrontierOps(sw.W, self.C, self.group)
            n, W = sw.n_u, sw.W
            seed = sw.seed if sw.seed is not None else torch.ones(n, dtype=torch.bool, device=X.device)
            frontier = _nonzero_rows(seed)
            mark = torch.zeros(max(n, 1), dtype=torch.uint8, device=X.device)
            moved = torch.zeros(1, dtype=torch.int32, device=X.device)
            while frontier.shape[0] and st.iters < self.max_iter:
                Y, ch = ops.jacobi(frontier, X, sw.rhs, sw.diag, self.delta)
                st.iters += 1
                mark.zero_()
                moved.zero_()
                _lib.call("dlp_frontier_commit", int(frontier.shape[0]), ptr(frontier), self.C, X.stride(0), ptr(Y),
                          ptr(ch.to(torch.uint8)), ptr(X), ptr(W.indptr), ptr(W.indices), ptr(mark), ptr(moved),
                          ptr(ops.edges), dtype_code(X.dtype), stream())
                if not int(moved.item()):
                    break
                frontier = _nonzero_rows(mark[:n])
            self._store(sw, X)
        st.edges, st.ms = int(ops.edges), tm.ms
        return st


class DynLPPlus(Solver):
    """Certified DynLP+ (solvers.py:244-343) as a pipeline:
    1. start: stored scores and h, new rows from their supernodes;
    2. h bootstrap (first batch): A h = diag to eps_h gives hmax;
    3. working-precision repair to eps_c = max(tol / (1.25 hmax), floor): push when few rows
       violate ("auto"), else the global PCG (Jacobi or AMG preconditioner);
    4. FP64 certification: the FP64 residual, its column maxima and max h give the bound
       max_c rho_c * hmax (hmax = max h / (1 - rho_h)); while it exceeds tol, a repair of the
       column-scaled residual refines the FP64 solution (10 rounds FP32, 2 FP64)."""

    def __init__(self, n_total, K, dtype="float32", method="auto", tol=1e-3, eps_h=0.05, group="auto",
                 push_frac=0.05, push_budget=30.0, amg_hmax=2000.0, max_iter=100_000, amg_opts=None,
                 block=False):
        super().__init__(n_total, K, dtype, store_dtype="float64")
        self.method, self.tol, self.eps_h, self.group = method, tol, eps_h, group
        self.push_frac, self.push_budget, self.amg_hmax = push_frac, push_budget, amg_hmax
        self.max_iter, self.amg_opts = max_iter, amg_opts or {}
        # block: the global solves run as one block PCG over all columns (block_pcg)
        self.block = bool(block)
        self.name, self.H, self.hmax = f"dynlp+{method}" + ("-block" if block else ""), None, None

    # -- preconditioner and working repair --------------------------------------------------
    def _wants_amg(self):
        if self.method == "amg":
            return True
        return self.method == "auto" and (self.hmax if self.hmax else float("inf")) > self.amg_hmax

    def _preconditioner(self, ctx):
        """(M, work per application): Jacobi, or the AMG built once per batch system."""
        if not self._wants_amg():
            return jacobi_M(ctx["sys"]), 0
        if ctx["amg"] is None:
            from .amg import AMG
            with Timer() as t:
                ctx["amg"] = AMG(ctx["sys"].W, ctx["sys"].s, **self.amg_opts)
            ctx["st"].setup_ms += t.ms
            ctx["st"].edges += 10 * ctx["sys"].nnz
        return ctx["amg"], ctx["amg"].work_per_apply

    def _global_solve(self, ctx, Z, rhs, eps):
        M, work = self._preconditioner(ctx)
        solve = block_pcg if self.block else pcg
        return solve(ctx["sys"], Z, rhs, eps, M, self.max_iter, ctx["st"], work)

    def _repair(self, ctx, Z, rhs, eps):
        """Working-precision repair of A Z = rhs to eps (solvers.py:272-288)."""
        sw, st = ctx["sys"], ctx["st"]
        tag = "amg" if self._wants_amg() else "pcg"
        if self.method != "auto":
            st.path = self.method
            return self._global_solve(ctx, Z, rhs, eps)
        cg = _ColumnCG(sw, Z, rhs, eps)
        cg.residual()
        st.edges += sw.nnz
        if sw.n_u and _count_violating(sw, cg.R, eps) > self.push_frac * sw.n_u:
            st.path = tag
            return self._global_solve(ctx, Z, rhs, eps)
        Z, _, ok = push(sw, Z, cg.R, eps, ctx["ops"], st, self.push_budget * max(sw.nnz, 1))
        st.path = "push" if ok else "push+" + tag
        return Z if ok else self._global_solve(ctx, Z, rhs, eps)

    # -- certificate -----------------------------------------------------------------------
    def _floor(self, scale=1.0):
        """Smallest residual ratio the working precision resolves at values of size scale."""
        unit = 1e-7 if self.dtype == torch.float32 else 2e-16
        return min(0.5, max(20 * unit, 10 * unit * scale))

    def _bound(self, rc, h_est):
        """(hmax, bound) from column maxima rc of |R|/diag (h last) and max h (solvers.py:295-301)."""
        rho_h = float(rc[self.C])
        if rho_h > self.eps_h:
            return h_est, float("inf")
        hmax = h_est / (1.0 - rho_h)
        return hmax, float(np.max(rc[: self.C])) * hmax

    # -- pipeline --------------------------------------------------------------------------
    def _start(self, sw):
        """[F, h] of the batch rows (new rows from their supernodes) and [rhs, diag]."""
        if self.H is None:
            self.H = torch.ones((self.n_total, 1), dtype=self.store_dtype, device=device())
        known = None if sw.is_new is None else ~sw.is_new
        X = supernode_init(sw, self._warm(sw), known, comp="fh")
        H = self._rows(sw, self.H, self.dtype)
        if known is not None and sw.comps is not None:
            H = init_h(sw, H, known, self.hmax or 1.0)
        Z = torch.cat([X, H], dim=1).contiguous()
        rhs = torch.cat([sw.rhs, sw.diag[:, None]], dim=1).contiguous()
        return Z, rhs

    def _bootstrap_h(self, ctx, Z, rhs):
        """First batch: A h = diag to eps_h (Jacobi PCG, or AMG when the method is amg)."""
        sw, C = ctx["sys"], self.C
        M, work = self._preconditioner(ctx) if self.method == "amg" else (jacobi_M(sw), 0)
        h = pcg(sw, Z[:, C:].contiguous(), rhs[:, C:].contiguous(), [self.eps_h], M, self.max_iter, ctx["st"], work)
        Z[:, C:] = h
        return _col_max(Z, C) / (1.0 - self.eps_h)

    def _certify(self, ctx, s64, Z):
        """FP64 residual, bound, scaled refinement; returns (Z64, bound, hmax)."""
        sw, st, C, tol = ctx["sys"], ctx["st"], self.C, self.tol
        Z64 = Z.double().contiguous()
        rhs64 = torch.cat([s64.rhs, s64.diag[:, None]], dim=1).contiguous()
        bound, hmax = 0.0, self.hmax
        rounds = 10 if self.dtype == torch.float32 else 2
        for _ in range(rounds if sw.n_u else 0):
            cg = _ColumnCG(s64, Z64, rhs64, np.ones(C + 1))
            cg.residual()
            st.edges += s64.nnz
            rc = rho_cols(s64, cg.R).cpu().numpy()
            h_est = _col_max(Z64, C)
            hmax, bound = self._bound(rc, h_est)
            if bound <= tol:
                break
            scale = np.where(rc > 0, rc, 1.0)
            base = np.array([tol / (1.25 * hmax)] * C + [self.eps_h])
            eps_c = np.clip(base / scale, self._floor(h_est), 1.0)
            Rs = _axpy_cols(1.0 / scale, cg.R, torch.zeros_like(cg.R)).to(self.dtype).contiguous()
            D = self._repair(ctx, torch.zeros_like(Rs), Rs, eps_c)
            _axpy_cols(scale, D.double().contiguous(), Z64)
        return Z64, bound, hmax

    def solve(self, sys):
        st, C = Stats(), self.C
        sw, s64 = sys.astype(self.dtype), sys.astype(torch.float64)
        with Timer() as tm:
            Z, rhs = self._start(sw)
            ctx = {"sys": sw, "st": st, "amg": None, "ops": FrontierOps(sw.W, C + 1, self.group)}
            if self.hmax is None and sw.n_u:
                self.hmax = self._bootstrap_h(ctx, Z, rhs)
            self.hmax = self.hmax or 1.0
            eps = [max(self.tol / (1.25 * self.hmax), self._floor())] * C + [max(self.eps_h, self._floor(self.hmax))]
            Z = self._repair(ctx, Z, rhs, eps)
            Z64, bound, hmax = self._certify(ctx, s64, Z)
            st.cert, st.hmax, self.hmax = bound, hmax, hmax
            self._store(sw, Z64[:, :C])
            self._put_rows(sw, self.H, Z64[:, C:])
        st.edges += int(ctx["ops"].edges)
        st.ms = tm.ms
        return st


class IncrementalLP(DynLPPlus):
    """Certified DynLP+ with resident FP64 state (solvers.py:346-434) for the global-index
    systems of stream.IncrementalStream: Z = [F, h] and its residual R persist across batches.
    A batch parks the rows outside the solved set (dlp_park_rows), starts its new rows from
    their supernodes, refreshes R on the changed rows (one row-operator pass; all rows every
    full_every batches), and, when few rows violate, repairs them with device push rounds under
    up to four threshold tightenings from the kept residual's certificate; otherwise (and on
    the first batch) it solves like DynLP+ and recomputes R."""

    def __init__(self, n_total, K, dtype="float32", tol=1e-3, full_every=16, max_rounds=50_000, **kw):
        super().__init__(n_total, K, dtype, method="auto", tol=tol, **kw)
        self.name, self.full_every, self.max_rounds = "dynlp+inc", full_every, max_rounds
        self.Z = self.R = self._fused = None
        self.since_full = 0

    def _rhs64(self, s64):
        return torch.cat([s64.rhs, s64.diag[:, None]], dim=1).contiguous()

    def _resync(self, s64, st):
        """R = rhs - A Z over every row (the periodic full refresh, solvers.py:370-373)."""
        cg = _ColumnCG(s64, self.Z, self._rhs64(s64), np.ones(self.C + 1))
        cg.residual()
        self.R = cg.R
        st.edges += s64.nnz
        self.since_full = 0

    def _refresh(self, s64, rows, st):
        """R[Q] = rhs[Q] - A[Q] Z for the listed rows (solvers.py:404-407): the row operator with
        a = -diag, b = +1 and rhs added."""
        C1 = self.C + 1
        Q = rows.to(torch.int32).contiguous()
        W = s64.W
        neg = _axpy_cols([1.0], s64.diag[:, None].contiguous(), torch.zeros((s64.n_u, 1), dtype=torch.float64,
                                                                              device=self.Z.device), sign=-1.0)
        _lib.call("dlp_row_operator", int(Q.shape[0]), ptr(Q), C1, ptr(W.indptr), None, ptr(W.indices), ptr(W.data),
                  _lib.F64, None, ptr(self.Z), C1, _lib.F64, ptr(neg), 1.0, ptr(self._rhs64(s64)), C1, ptr(self.R), C1,
                  _lib.F64, 0, W.avg_deg, stream())
        st.edges += int(W.nnz * Q.shape[0] / max(s64.n_u, 1))   # mean-degree estimate

    def _thresholds(self, hmax):
        return [self.tol / (1.25 * hmax)] * self.C + [self.eps_h]

    def _local_repair(self, s64, st):
        """Push rounds on the violating rows under up to four tightenings of the thresholds from
        the kept residual's bound; True when it certifies tol."""
        C = self.C
        if self._fused is None or self._fused.W is not s64.W:
            self._fused = FusedPush(s64.W, C + 1, self.group)
        hmax, bound = self.hmax or 1.0, float("inf")
        diag = s64.diag.contiguous()
        for _ in range(4):
            eps = self._thresholds(hmax)
            rounds, ok, edges = self._fused.run(self.Z, self.R, diag, eps, None, self.max_rounds)
            st.iters, st.edges = st.iters + rounds, st.edges + edges
            hmax, bound = self._bound(rho_cols(s64, self.R).cpu().numpy(), _col_max(self.Z, C))
            if ok and bound <= self.tol:
                break
        st.cert, st.hmax, self.hmax = bound, hmax, hmax
        return bound <= self.tol

    def solve(self, sys):
        if sys.solved is None:
            raise ValueError("dynlp+inc needs a global-index system of stream.IncrementalStream")
        s64, C, st = sys.astype("float64"), self.C, Stats(path="inc-push")
        with Timer() as tm:
            solved = sys.solved.to(torch.uint8).contiguous()
            local = self.Z is not None
            if local:
                _lib.call("dlp_park_rows", s64.n_u, ptr(solved), C + 1, C + 1, _lib.F64, ptr(self.Z), ptr(self.R),
                          stream())
                if s64.is_new is not None and s64.comps is not None and "fh" in s64.comps:
                    known = s64.solved.bool() & ~s64.is_new
                    supernode_init(s64, self.Z, known, cols=slice(0, C), comp="fh")
                    _supernode_fill(s64, self.Z, C, 1, known, 2, [(self.hmax or 1.0) - 1.0], 1.0, "fh")
                if sys.changed is not None and sys.changed.numel():
                    self._refresh(s64, sys.changed, st)
                self.since_full += 1
                if self.since_full >= self.full_every:   # bound the rounding drift of the kept residual
                    self._resync(s64, st)
                few = _count_violating(s64, self.R, self._thresholds(self.hmax or 1.0))
                local = few <= self.push_frac * max(sys.n_solved, 1) and self._local_repair(s64, st)
            if not local:   # first batch, large change, or a push that did not certify
                if self.Z is not None:
                    self.F, self.H = self.Z[:, :C].clone(), self.Z[:, C:].clone()
                inner = DynLPPlus.solve(self, sys)
                st.iters, st.edges, st.setup_ms = st.iters + inner.iters, st.edges + inner.edges, inner.setup_ms
                st.cert, st.hmax, st.path = inner.cert, inner.hmax, "inc-" + inner.path
                self.Z = torch.cat([self.F, self.H], dim=1).double().contiguous()
                _lib.call("dlp_park_rows", s64.n_u, ptr(solved), C + 1, C + 1, _lib.F64, ptr(self.Z), None, stream())
                self._resync(s64, st)
            self.F, self.H = self.Z[:, :C], self.Z[:, C:]
        st.ms = tm.ms
        return st


class Reference:
    """High-accuracy FP64 solution F* and absorption bound (solvers.py:437-445): DynLP+ in FP64
    with the AMG preconditioner, tol 1e-8, eps_h 1e-6."""

    def __init__(self, n_total, K, tol=1e-8):
        self.inner = DynLPPlus(n_total, K, "float64", method="amg", tol=tol, eps_h=1e-6, max_iter=100_000)

    def solve(self, sys):
        st = self.inner.solve(sys)
        return self.inner.scores(sys), self.inner.hmax, st


_FACTORIES = {
    "itlp": lambda n, K, dt, delta, tol, group, kw: ItLP(n, K, dt, delta, **kw),
    "dynlp": lambda n, K, dt, delta, tol, group, kw: DynLP(n, K, dt, delta, group, comp="tau", **kw),
    "dynlp-fh": lambda n, K, dt, delta, tol, group, kw: DynLP(n, K, dt, delta, group, comp="fh", **kw),
    "dynlp+inc": lambda n, K, dt, delta, tol, group, kw: IncrementalLP(n, K, dt, tol=tol, group=group, **kw),
}


def make(name, n_total, K, *, dtype, delta, tol, group, max_iter=None):
    """Solver by name (solvers.py:448-458): the SOLVERS list; `dynlp+<method>` for any DynLP+
    method (pcg, amg, auto), dynlp+inc on IncrementalStream systems."""
    kw = {} if max_iter is None else {"max_iter": max_iter}
    if name in _FACTORIES:
        return _FACTORIES[name](n_total, K, dtype, delta, tol, group, kw)
    head, _, method = name.partition("+")
    if head == "dynlp" and method:
        return DynLPPlus(n_total, K, dtype, method=method, tol=tol, group=group, **kw)
    raise ValueError(f"unknown solver {name!r}")


SOLVERS = ["itlp", "dynlp", "dynlp-fh", "dynlp+pcg", "dynlp+amg", "dynlp+auto", "dynlp+inc"]

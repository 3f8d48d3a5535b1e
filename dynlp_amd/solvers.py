"""Label-propagation solvers for one batch system, on the device.

Same surface as dynlp/solvers.py of the reference: `Stats`, `spmm`, `apply_A`, `rho_cols`,
`violation`, `jacobi_M`, `pcg` (:66-97), `push` (:100-125), `predict`, `ItLP` (:190-210),
`DynLP` (:213-241), `DynLPPlus` (:244-343), `Reference` (:437-445), `make` (:448-458) and
`SOLVERS`.  Every sweep, product and reduction runs in the HIP kernels of the C ABI; Python
keeps the reference's control flow (and its host synchronisations).  New rows of a stream's
systems start from their supernode components (`supernode_init`/`init_h`, :128-163).  The
resident incremental method (`dynlp+inc`, IncrementalLP :346-434) runs on the
global-index systems of `stream.IncrementalStream`; for edge-level streams the resident path
is `stream.EdgeStreamLP`.

A `System` holds the batch system on the active unlabeled set U in local ids (stream.py:17-58):
W (CSR U x U), s, diag, B, rhs, prior.
"""
from __future__ import annotations

import ctypes as ct
import time
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
    """Stream._build (stream.py:142-177) for the graph's current edge set: U = active, unlabeled
    and grounded (components of the unlabeled subgraph with labeled contact, found with the
    device CC); W = A restricted to U; s, diag, B, rhs.  labeled overrides the graph's labeled
    set (a stream draws its own).  Host ingest of the edge list, device grounding and products."""
    dt = torch.float64 if dtype in ("float64", torch.float64) else torch.float32
    n, K = g.n, g.K
    act = np.ones(n, dtype=bool) if active is None else np.asarray(active, dtype=bool)
    glab = g.labeled if labeled is None else np.asarray(labeled, dtype=bool)
    lab = glab & act
    unl = act & ~glab
    u, v, w = g.u, g.v, g.w.astype(np.float64)
    # class weights to labeled neighbours
    Bfull = np.zeros((n, K))
    m1 = lab[v] & unl[u]
    np.add.at(Bfull, (u[m1], g.y[v[m1]]), w[m1])
    m2 = lab[u] & unl[v]
    np.add.at(Bfull, (v[m2], g.y[u[m2]]), w[m2])
    # grounding on the device
    uu = unl[u] & unl[v]
    labs, nc = connected_components(n, np.r_[u[uu], v[uu]], np.r_[v[uu], u[uu]])
    labs = labs.cpu().numpy()
    gw = np.zeros(nc)
    np.add.at(gw, labs[unl], Bfull[unl].sum(axis=1))
    solved = unl & (gw[labs] > 0)
    U = np.flatnonzero(solved)
    loc = np.full(n, -1, dtype=np.int64)
    loc[U] = np.arange(U.shape[0])
    keep = solved[u] & solved[v]
    r = np.r_[loc[u[keep]], loc[v[keep]]]
    c = np.r_[loc[v[keep]], loc[u[keep]]]
    ww = np.r_[w[keep], w[keep]]
    order = np.lexsort((c, r))
    r, c, ww = r[order], c[order], ww[order]
    n_u = U.shape[0]
    indptr = np.zeros(n_u + 1, dtype=np.int64)
    np.cumsum(np.bincount(r, minlength=n_u), out=indptr[1:])
    dev = device()
    W = CSR(torch.as_tensor(ww, device=dev).to(dt), torch.as_tensor(c.astype(np.int32), device=dev),
            torch.as_tensor(indptr.astype(np.int32), device=dev), (n_u, n_u))
    C = 1 if K == 2 else K
    prior = torch.full((C,), 0.5 if K == 2 else 1.0 / K, dtype=dt, device=dev)
    Bm = torch.as_tensor(Bfull[U], device=dev).to(dt)
    s = Bm.sum(dim=1) + eta
    rowsum = W @ torch.ones(n_u, dtype=dt, device=dev)
    diag = s + rowsum
    rhs = (Bm[:, 1:2] if K == 2 else Bm) + eta * prior[None, :]
    ones = torch.ones(n_u, dtype=torch.bool, device=dev)
    return System(torch.as_tensor(U, device=dev), W, s, diag, Bm, rhs.contiguous(), prior, K, C, eta,
                  is_new=ones, seed=ones)


def _dev64(x):
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=device())


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
    """max |R_uc| / (eps_c * diag_u) (solvers.py:57-59)."""
    if not R.shape[0]:
        return 0.0
    rc = rho_cols(sys, R)
    return float((rc / torch.as_tensor(eps_d, dtype=torch.float64, device=rc.device)).max())


def _coldot(X, Y, d=None):
    C = X.shape[1]
    out = torch.empty(C, dtype=torch.float64, device=X.device)
    _lib.call("dlp_coldot", int(X.shape[0]), C, ptr(X), C, ptr(Y), C, ptr(d), ptr(out), dtype_code(X.dtype),
              stream())
    return out


def _colop(mode, a, X, Y, d=None, sign=1.0):
    C = X.shape[1]
    _lib.call("dlp_colop", int(X.shape[0]), C, mode, float(sign), ptr(a), ptr(X), C, ptr(d), ptr(Y), C,
              dtype_code(X.dtype), stream())
    return Y


def jacobi_M(sys: System):
    """R / diag (solvers.py:62-63)."""
    def M(R):
        Y = torch.empty_like(R)
        return _colop(2, None, R.contiguous(), Y, d=sys.diag.to(R.dtype).contiguous())
    return M


def _safe_div(num, den):
    return torch.where(den > 0, num / torch.where(den > 0, den, torch.ones_like(den)), torch.zeros_like(num))


def pcg(sys: System, Z, rhs, eps, M, max_iter, st: Stats, work_M=0, patience=400):
    """One CG per column until |r_c| <= eps_c * diag; true-residual re-check; stall exit
    (solvers.py:66-97).  Z (n_u x C, row-major) is updated in place and returned."""
    eps_d = [float(e) for e in (eps if hasattr(eps, "__len__") else [eps])]
    Z = Z.contiguous()
    R = (rhs - apply_A(sys, Z)).contiguous()
    st.edges += sys.nnz
    best = violation(sys, R, eps_d)
    if best <= 1:
        return Z
    Y = M(R).contiguous()
    P = Y.clone()
    rz, stall = _coldot(R, Y), 0
    for _ in range(max_iter):
        AP = apply_A(sys, P)
        pAp = _coldot(P, AP)
        alpha = _safe_div(rz, pAp)
        _colop(0, alpha, P, Z)
        _colop(0, alpha, AP, R, sign=-1.0)
        st.iters, st.edges = st.iters + 1, st.edges + sys.nnz + work_M
        v = violation(sys, R, eps_d)
        if v <= 1:
            R = (rhs - apply_A(sys, Z)).contiguous()
            st.edges += sys.nnz
            v = violation(sys, R, eps_d)
            if v <= 1:
                break
        best, stall = (v, 0) if v < 0.9 * best else (best, stall + 1)
        if stall >= patience:
            break
        Y = M(R).contiguous()
        rz_new = _coldot(R, Y)
        beta = _safe_div(rz_new, rz)
        _colop(1, beta, Y, P)          # P = Y + beta * P
        rz = rz_new
    return Z


def push(sys: System, Z, R, eps, ops: FrontierOps, st: Stats, budget_edges=None, max_iter=10**7):
    """Residual push on the violating set (solvers.py:100-125): D = R[c]/diag, Z[c] += D,
    R[c] = 0, R[N(c)] += W D, next candidates = touched vertices that violate.  Returns
    (Z, R, converged)."""
    n = sys.n_u
    eps64 = _dev64(eps)
    diag64 = sys.diag.double().contiguous()
    C = R.shape[1]
    dt = dtype_code(R.dtype)
    cand = torch.empty(n, dtype=torch.int32, device=R.device)
    cnt = torch.zeros(1, dtype=torch.int32, device=R.device)
    _lib.call("dlp_violating_rows", n, None, C, ptr(R), C, dt, ptr(diag64), ptr(eps64), ptr(cand), ptr(cnt), stream())
    nc = int(cnt.item())
    cand = torch.sort(cand[:nc]).values
    e0, it, best, stall = int(ops.edges), 0, float("inf"), 0
    ws_bytes = int(_lib.load().dlp_scan_workspace_bytes(max(n, 1)))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=R.device)
    touched = torch.empty(n, dtype=torch.int32, device=R.device)
    tcnt = torch.zeros(1, dtype=torch.int32, device=R.device)
    while nc and it < max_iter:
        D = R[cand.long()] / sys.diag[cand.long()][:, None]
        Z[cand.long()] += D
        R[cand.long()] = 0
        mark = torch.zeros(n, dtype=torch.uint8, device=R.device)
        mark[cand.long()] = 1
        ops.push(cand, D, R, mark)
        _lib.call("dlp_flatnonzero_u8", n, ptr(mark), ptr(touched), ptr(tcnt), ptr(ws), ws_bytes, stream())
        nt = int(tcnt.item())
        _lib.call("dlp_violating_rows", nt, ptr(touched), C, ptr(R), C, dt, ptr(diag64), ptr(eps64), ptr(cand),
                  ptr(cnt), stream())
        nc = int(cnt.item())
        cand = torch.sort(cand[:nc]).values
        it += 1
        if it % 32 == 0:
            if budget_edges is not None and int(ops.edges) - e0 > budget_edges:
                break
            v = violation(sys, R, eps)
            best, stall = (v, 0) if v < 0.99 * best else (best, stall + 1)
            if stall >= 8:
                break
    st.iters += it
    return Z, R, nc == 0


def _comp_avg(lab, nc, S_rows, T_rows, fallback):
    """Per component of the new vertices: sum S / sum T, or fallback where sum T = 0
    (solvers.py:128-135)."""
    from .backend import scatter_add
    lab = lab.long()
    S = torch.zeros((nc,) + tuple(S_rows.shape[1:]), dtype=S_rows.dtype, device=S_rows.device)
    T = torch.zeros(nc, dtype=S_rows.dtype, device=S_rows.device)
    scatter_add(S, lab, S_rows)
    scatter_add(T, lab, T_rows.to(S_rows.dtype))
    ok = T > 0
    Tn = torch.where(ok, T, torch.ones_like(T))
    fb = torch.as_tensor(fallback, dtype=S.dtype, device=S.device)
    if S.dim() == 2:
        return torch.where(ok[:, None], S / Tn[:, None], fb.expand_as(S))
    return torch.where(ok, S / Tn, fb.expand_as(S))


def _rows_product(W: CSR, rows, Y):
    """(W @ Y)[rows] (the product of the row subset W[rows])."""
    return spmm(W, Y.contiguous())[rows]


def supernode_init(sys: System, X, known=None, cols=None, comp="tau"):
    """New vertices take their supernode's value (solvers.py:138-153): with known None, the
    share of the component's weight to each class (DynLP's rule); with a known-row mask, the
    weighted average over all known neighbours (labels and the last solution).  Components from
    sys.comps[comp]; a system without them (or without new rows) leaves X as it is."""
    if sys.is_new is None or sys.comps is None or comp not in sys.comps:
        return X
    new = torch.nonzero(sys.is_new).flatten()
    if not new.shape[0]:
        return X
    cols = slice(0, sys.C) if cols is None else cols
    B = sys.B.to(X.dtype)
    s_, t = sys.label_cols(B)[new], B[new].sum(dim=1)
    if known is not None:
        kf = known.to(X.dtype)
        s_ = s_ + _rows_product(sys.W, new, X[:, cols] * kf[:, None])
        t = t + _rows_product(sys.W, new, kf[:, None])[:, 0]
    lab, nc = sys.comps[comp]
    X[new, cols] = _comp_avg(lab, nc, s_, t, sys.prior.to(X.dtype)[None, :])[lab.long()]
    return X


def init_h(sys: System, H, known, fallback, comp="fh"):
    """New vertices: h ~ 1 + the component's weighted average of h over known neighbours
    (labeled ones count h = 0) (solvers.py:156-163)."""
    if sys.is_new is None or sys.comps is None or comp not in sys.comps:
        return H
    new = torch.nonzero(sys.is_new).flatten()
    if new.shape[0]:
        kf = known.to(H.dtype)
        lab, nc = sys.comps[comp]
        num = _rows_product(sys.W, new, (H[:, 0] * kf)[:, None])[:, 0]
        den = sys.B.to(H.dtype)[new].sum(dim=1) + _rows_product(sys.W, new, kf[:, None])[:, 0]
        hc = _comp_avg(lab, nc, num, den, fallback - 1) + 1
        H[new, 0] = hc[lab.long()]
    return H


def predict(F, K):
    """solvers.py:166-167."""
    return (F[:, 0] > 0.5).to(torch.int32) if K == 2 else F.argmax(dim=1).to(torch.int32)


class Solver:
    name = "base"

    def __init__(self, n_total, K, dtype="float32", store_dtype=None):
        self.n_total, self.K = n_total, K
        self.dtype = torch.float64 if dtype == "float64" else torch.float32
        sd = store_dtype or dtype
        self.store_dtype = torch.float64 if sd == "float64" else torch.float32
        self.C, self.F = 1 if K == 2 else K, None

    def _warm(self, sys: System):
        if self.F is None:
            self.F = torch.empty((self.n_total, self.C), dtype=self.store_dtype, device=device())
            self.F[:] = sys.prior.to(self.store_dtype)
        return self.F[sys.U].to(self.dtype).contiguous()

    def _store(self, sys: System, X):
        self.F[sys.U] = X.to(self.store_dtype)

    def scores(self, sys: System):
        return self.F[sys.U]


class ItLP(Solver):
    """Jacobi from the prior on all rows until max |dX| <= delta (solvers.py:190-210): each
    sweep is the frontier-Jacobi kernel over every row, stopping when no row changed."""
    name = "itlp"

    def __init__(self, n_total, K, dtype="float32", delta=1e-4, max_iter=200_000):
        super().__init__(n_total, K, dtype)
        self.delta, self.max_iter = delta, max_iter

    def solve(self, sys):
        st, sys = Stats(path="jacobi"), sys.astype(self.dtype)
        with Timer() as tm:
            X = self._warm(sys)
            X[:] = sys.prior
            ops = FrontierOps(sys.W, self.C)
            rows = torch.arange(sys.n_u, dtype=torch.int32, device=X.device)
            for _ in range(self.max_iter):
                Y, ch = ops.jacobi(rows, X, sys.rhs, sys.diag, self.delta)
                X, st.iters = Y, st.iters + 1
                if not sys.n_u or not bool(ch.any()):
                    break
            self._store(sys, X)
        st.edges, st.ms = int(ops.edges), tm.ms
        return st


class DynLP(Solver):
    """DynLP Algorithm 2 frontier Jacobi from the seed set (solvers.py:213-241)."""

    def __init__(self, n_total, K, dtype="float32", delta=1e-4, group="auto", comp="tau", max_iter=10**7):
        super().__init__(n_total, K, dtype)
        self.delta, self.group, self.comp, self.max_iter = delta, group, comp, max_iter
        self.name = "dynlp" if comp == "tau" else f"dynlp-{comp}"

    def solve(self, sys):
        st, sys = Stats(path="frontier-jacobi"), sys.astype(self.dtype)
        with Timer() as tm:
            X = supernode_init(sys, self._warm(sys), comp=self.comp)
            ops = FrontierOps(sys.W, self.C, self.group)
            n = sys.n_u
            seed = sys.seed if sys.seed is not None else torch.ones(n, dtype=torch.bool, device=X.device)
            frontier = torch.nonzero(seed).flatten().to(torch.int32)
            ws_bytes = int(_lib.load().dlp_scan_workspace_bytes(max(n, 1)))
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=X.device)
            nxt = torch.empty(n, dtype=torch.int32, device=X.device)
            cnt = torch.zeros(1, dtype=torch.int32, device=X.device)
            while frontier.shape[0] and st.iters < self.max_iter:
                Y, ch = ops.jacobi(frontier, X, sys.rhs, sys.diag, self.delta)
                st.iters += 1
                upd = frontier[ch]
                if not upd.shape[0]:
                    break
                X[upd.long()] = Y[ch]
                mark = torch.zeros(n, dtype=torch.uint8, device=X.device)
                mark[upd.long()] = 1
                ops.mark_neighbors(upd, mark)
                _lib.call("dlp_flatnonzero_u8", n, ptr(mark), ptr(nxt), ptr(cnt), ptr(ws), ws_bytes, stream())
                frontier = nxt[: int(cnt.item())].clone()
            self._store(sys, X)
        st.edges, st.ms = int(ops.edges), tm.ms
        return st


class DynLPPlus(Solver):
    """Certified DynLP+ (solvers.py:244-343): working-precision solve (push, PCG or AMG-PCG),
    certificate from the FP64 residual (bound = max rho_c * hmax), FP64 refinement rounds."""

    def __init__(self, n_total, K, dtype="float32", method="auto", tol=1e-3, eps_h=0.05, group="auto",
                 push_frac=0.05, push_budget=30.0, amg_hmax=2000.0, max_iter=100_000, amg_opts=None):
        super().__init__(n_total, K, dtype, store_dtype="float64")
        self.method, self.tol, self.eps_h, self.group = method, tol, eps_h, group
        self.push_frac, self.push_budget, self.amg_hmax = push_frac, push_budget, amg_hmax
        self.max_iter, self.amg_opts = max_iter, amg_opts or {}
        self.name, self.H, self.hmax = f"dynlp+{method}", None, None

    def _use_amg(self):
        return self.method == "amg" or (self.method == "auto" and (self.hmax or float("inf")) > self.amg_hmax)

    def _precond(self, sys, st, amg_box):
        if not self._use_amg():
            return jacobi_M(sys), 0
        if amg_box[0] is None:
            from .amg import AMG
            with Timer() as t:
                amg_box[0] = AMG(sys.W, sys.s, **self.amg_opts)
            st.setup_ms += t.ms
            st.edges += 10 * sys.nnz
        return amg_box[0], amg_box[0].work_per_apply

    def _global(self, sys, Z, rhs, eps, st, amg_box):
        M, wm = self._precond(sys, st, amg_box)
        return pcg(sys, Z, rhs, eps, M, self.max_iter, st, wm)

    def _run(self, sys, Z, rhs, eps, ops, st, amg_box):
        if self.method != "auto":
            st.path = self.method
            return self._global(sys, Z, rhs, eps, st, amg_box)
        R = (rhs - apply_A(sys, Z)).contiguous()
        st.edges += sys.nnz
        rc = (R.abs() / (torch.as_tensor(eps, dtype=R.dtype, device=R.device)[None, :] * sys.diag[:, None]))
        if sys.n_u and float((rc > 1).any(dim=1).float().mean()) > self.push_frac:
            st.path = "amg" if self._use_amg() else "pcg"
            return self._global(sys, Z, rhs, eps, st, amg_box)
        Z, R, ok = push(sys, Z, R, eps, ops, st, self.push_budget * max(sys.nnz, 1))
        st.path = "push"
        if ok:
            return Z
        st.path = "push+amg" if self._use_amg() else "push+pcg"
        return self._global(sys, Z, rhs, eps, st, amg_box)

    def _floor(self, scale=1.0):
        u = 1e-7 if self.dtype == torch.float32 else 2e-16
        return min(0.5, max(20 * u, 10 * u * scale))

    def _bound(self, rc, h_est):
        C = self.C
        if float(rc[C]) > self.eps_h:
            return h_est, float("inf")
        hmax = h_est / (1 - float(rc[C]))
        return hmax, float(rc[:C].max()) * hmax

    def solve(self, sys):
        st, C, tol = Stats(), self.C, self.tol
        sw, s64 = sys.astype(self.dtype), sys.astype(torch.float64)
        with Timer() as tm:
            if self.H is None:
                self.H = torch.ones((self.n_total, 1), dtype=self.store_dtype, device=device())
            # new rows start from their supernodes (systems of a stream carry the components;
            # other systems keep the stored values: the prior and h = 1 on first sight)
            known = ~sw.is_new if sw.is_new is not None else None
            X = supernode_init(sw, self._warm(sw), known, comp="fh")
            Hc = self.H[sw.U].to(self.dtype)
            if known is not None and sw.comps is not None:
                Hc = init_h(sw, Hc, known, self.hmax or 1.0)
            Z = torch.cat([X, Hc], dim=1).contiguous()
            rhs = torch.cat([sw.rhs, sw.diag[:, None]], dim=1).contiguous()
            ops, amg_box, hmax = FrontierOps(sw.W, C + 1, self.group), [None], self.hmax
            if hmax is None and sw.n_u:
                M, wm = self._precond(sw, st, amg_box) if self.method == "amg" else (jacobi_M(sw), 0)
                h = pcg(sw, Z[:, C:].contiguous(), rhs[:, C:].contiguous(), [self.eps_h], M, self.max_iter, st, wm)
                Z[:, C:] = h
                hmax = float(Z[:, C].max()) / (1 - self.eps_h)
            self.hmax = hmax = hmax or 1.0
            eps = [max(tol / (1.25 * hmax), self._floor())] * C + [max(self.eps_h, self._floor(hmax))]
            Z = self._run(sw, Z, rhs, eps, ops, st, amg_box)
            Z64 = Z.double()
            rhs64 = torch.cat([s64.rhs, s64.diag[:, None]], dim=1).contiguous()
            bound = 0.0
            for _ in range(10 if self.dtype == torch.float32 else 2):
                if not sw.n_u:
                    break
                R64 = (rhs64 - apply_A(s64, Z64)).contiguous()
                st.edges += s64.nnz
                rc, h_est = rho_cols(s64, R64), float(Z64[:, C].max())
                hmax, bound = self._bound(rc, h_est)
                if bound <= tol:
                    break
                scale = torch.where(rc > 0, rc, torch.ones_like(rc))
                base = torch.as_tensor([tol / (1.25 * hmax)] * C + [self.eps_h], dtype=torch.float64, device=rc.device)
                eps_c = torch.clamp(base / scale, self._floor(h_est), 1.0)
                Rs = (R64 / scale[None, :]).to(self.dtype).contiguous()
                D = self._run(sw, torch.zeros_like(Rs), Rs, eps_c.tolist(), ops, st, amg_box)
                Z64 += D.double() * scale[None, :]
            st.cert, st.hmax, self.hmax = bound, hmax, hmax
            self._store(sw, Z64[:, :C])
            self.H[sw.U] = Z64[:, C:].to(self.store_dtype)
        st.edges += int(ops.edges)
        st.ms = tm.ms
        return st


class IncrementalLP(DynLPPlus):
    """Certified DynLP+ with resident state (solvers.py:346-434), for the global-index systems
    of stream.IncrementalStream: Z = [F, h] and its FP64 residual R are kept between batches; a
    batch parks the rows outside the solved set (F = 0, h = 1, R = 0), refreshes R on its
    changed rows, and when few rows violate repairs them with the fused asynchronous push (up to
    4 rounds, the thresholds re-tightened from the bound each time), the certificate read off
    the kept residual; large changes and the first batch go through DynLP+ and a full residual.
    R is recomputed in full every full_every batches to bound rounding drift.  New rows start
    from their supernode's averages (supernode_init, init_h).  The edge-update service path of
    this package is stream.EdgeStreamLP."""

    def __init__(self, n_total, K, dtype="float32", tol=1e-3, full_every=16, max_rounds=50_000, **kw):
        super().__init__(n_total, K, dtype, method="auto", tol=tol, **kw)
        self.name, self.full_every, self.max_rounds = "dynlp+inc", full_every, max_rounds
        self.Z = self.R = self._fused = None
        self.since_full = 0

    @staticmethod
    def _rhs64(s64):
        return torch.cat([s64.rhs, s64.diag[:, None]], dim=1).contiguous()

    def _park(self, Z, out):
        """Exact values of inert rows (diag 1, rhs 0; h's rhs is diag): F = 0, h = 1."""
        Z[out, : self.C] = 0
        Z[out, self.C] = 1

    def _full_residual(self, s64, st):
        self.R = (self._rhs64(s64) - apply_A(s64, self.Z)).contiguous()
        st.edges += s64.nnz
        self.since_full = 0

    def _refresh_rows(self, s64, Q, st):
        """R[Q] = rhs[Q] - diag[Q] Z[Q] + W[Q] Z (solvers.py:404-407), one row operator launch."""
        C1 = self.C + 1
        rhs = self._rhs64(s64)
        neg_diag = (-s64.diag).contiguous()
        W = s64.W
        Q = Q.to(torch.int32).contiguous()
        _lib.call("dlp_row_operator", int(Q.shape[0]), ptr(Q), C1, ptr(W.indptr), None, ptr(W.indices), ptr(W.data),
                  _lib.F64, None, ptr(self.Z), C1, _lib.F64, ptr(neg_diag), 1.0, ptr(rhs), C1, ptr(self.R), C1,
                  _lib.F64, 0, W.avg_deg, stream())
        rows = Q.long()
        st.edges += int((W.indptr[rows + 1] - W.indptr[rows]).sum())

    def _push(self, s64, eps, st):
        """Repair the violating rows of R; True when every row satisfies eps."""
        Z, R = self.Z, self.R
        eps_d = torch.as_tensor(eps, dtype=torch.float64, device=R.device)
        if self._fused is None or self._fused.W is not s64.W:
            self._fused = FusedPush(s64.W, Z.shape[1], self.group)
        viol = (R.abs() > eps_d[None, :] * s64.diag[:, None]).any(dim=1)
        cand = torch.nonzero(viol).flatten().to(torch.int32)
        rounds, ok, edges = self._fused.run(Z, R, s64.diag, eps_d, cand, self.max_rounds)
        st.iters, st.edges = st.iters + rounds, st.edges + edges
        return ok

    def solve(self, sys):
        if sys.solved is None:
            raise ValueError("dynlp+inc needs an IncrementalStream system (bench --incremental)")
        s64, C, tol, st = sys.astype("float64"), self.C, self.tol, Stats(path="inc-push")
        with Timer() as tm:
            out = ~sys.solved.bool()
            small = self.Z is not None
            if small:
                Z = self.Z
                self._park(Z, out)
                self.R[out] = 0
                if s64.is_new is not None:
                    known = s64.solved.bool() & ~s64.is_new
                    Z[:, :C] = supernode_init(s64, Z[:, :C].clone(), known, comp="fh")
                    Z[:, C:] = init_h(s64, Z[:, C:].clone(), known, self.hmax or 1.0)
                if sys.changed is not None and sys.changed.numel():
                    self._refresh_rows(s64, sys.changed, st)
                self.since_full += 1
                if self.since_full >= self.full_every:  # bound rounding drift in the kept residual
                    self._full_residual(s64, st)
                eps = [tol / (1.25 * (self.hmax or 1.0))] * C + [self.eps_h]
                eps_d = torch.as_tensor(eps, dtype=torch.float64, device=self.R.device)
                viol = (self.R.abs() > eps_d[None, :] * s64.diag[:, None]).any(dim=1)
                small = float(viol.sum()) / max(sys.n_solved, 1) <= self.push_frac
            if small:
                bound, hmax = float("inf"), self.hmax
                for _ in range(4):
                    ok = self._push(s64, eps, st)
                    hmax, bound = self._bound(rho_cols(s64, self.R), float(self.Z[:, C].max()))
                    if ok and bound <= tol:
                        break
                    eps = [tol / (1.25 * hmax)] * C + [self.eps_h]
                small, st.cert, st.hmax, self.hmax = bound <= tol, bound, hmax, hmax
            if not small:  # first or large batch, or the push did not certify: global solve
                if self.Z is not None:
                    self.F, self.H = self.Z[:, :C].clone(), self.Z[:, C:].clone()
                inner = DynLPPlus.solve(self, sys)
                st.iters, st.edges, st.setup_ms = st.iters + inner.iters, st.edges + inner.edges, inner.setup_ms
                st.cert, st.hmax, st.path = inner.cert, inner.hmax, "inc-" + inner.path
                self.Z = torch.cat([self.F, self.H], dim=1).double().contiguous()
                self._park(self.Z, out)
                self._full_residual(s64, st)
            self.F, self.H = self.Z[:, :C], self.Z[:, C:]
        st.ms = tm.ms
        return st


class Reference:
    """High-accuracy FP64 solution F* and absorption bound (solvers.py:437-445)."""

    def __init__(self, n_total, K, tol=1e-8):
        self.inner = DynLPPlus(n_total, K, "float64", method="amg", tol=tol, eps_h=1e-6, max_iter=100_000)

    def solve(self, sys):
        st = self.inner.solve(sys)
        return self.inner.scores(sys), self.inner.hmax, st


def make(name, n_total, K, *, dtype, delta, tol, group, max_iter=None):
    """solvers.py:448-458 (dynlp+inc on IncrementalStream systems; for edge-update streams the
    resident path is stream.EdgeStreamLP)."""
    kw = {} if max_iter is None else {"max_iter": max_iter}
    if name == "itlp":
        return ItLP(n_total, K, dtype, delta, **kw)
    if name in ("dynlp", "dynlp-fh"):
        return DynLP(n_total, K, dtype, delta, group, comp="fh" if name == "dynlp-fh" else "tau", **kw)
    if name == "dynlp+inc":
        return IncrementalLP(n_total, K, dtype, tol=tol, group=group, **kw)
    if name.startswith("dynlp+"):
        return DynLPPlus(n_total, K, dtype, method=name.split("+", 1)[1], tol=tol, group=group, **kw)
    raise ValueError(f"unknown solver {name!r}")


SOLVERS = ["itlp", "dynlp", "dynlp-fh", "dynlp+pcg", "dynlp+amg", "dynlp+auto", "dynlp+inc"]

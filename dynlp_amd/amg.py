"""Aggregation-based algebraic multigrid for grounded graph Laplacians, on the device.

Same construction and interface as dynlp/amg.py of the reference (`Level` :20-29, `AMG`
:77-134): A = diag(s + rowsum(W)) - W; aggregates by heavy-edge handshake matching (two
passes of up to three rounds per level, about 4x coarsening); W_c = off-diagonal part of
P^T W P with duplicates summed, s_c = P^T s; stop at max_coarse rows or when a level shrinks
by less than min_ratio; exact FP64 inverse at the coarsest level (n <= 4*max_coarse), formed
by Gauss-Jordan steps and applied with FP64 accumulation on the device (the reference casts it
to the working dtype); weighted-Jacobi smoothing (omega 0.7, nu 1) in a symmetric V-cycle.
Matching, coarsening, restriction and prolongation are HIP kernels (csrc/amg.hip); level
residuals R - A X are one row-operator launch each and the smoother's correction one column
pass (csrc/spmm.hip, csrc/blas.hip).
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _lib
from .backend import CSR, device, dtype_code, ptr, stream
from .kernels import spmm_axpy


# Levels of at most this many stored entries per row run their smoothing and residual passes
# on the pipelined row gathers; above it the row operator's lane groups are faster (SBM-1M
# levels: 20 entries a row 0.81 against 0.97 ms a residual pass, 106 a row 0.66 against 0.55).
PIPE_MAX_DEG = 48


@dataclass
class Level:
    W: CSR
    s: torch.Tensor
    diag: torch.Tensor
    agg: torch.Tensor | None = None   # fine -> coarse map (int32)
    nc: int = 0
    neg_diag: torch.Tensor | None = None

    def apply(self, X):
        return spmm_axpy(self.W, X.contiguous(), a=self.diag, b=-1.0)

    def pipelined(self, X):
        """True when the level's passes run on the pipelined row gathers (dlp_amg_level): FP32
        blocks of a multiple of 4 columns on a level of at most PIPE_MAX_DEG entries per row
        (denser coarse levels run faster on the row operator's lane groups)."""
        return (X.dtype == torch.float32 and self.W.data.dtype == torch.float32 and X.shape[1] % 4 == 0
                and X.is_contiguous() and self.W.avg_deg <= PIPE_MAX_DEG and X.numel() * 4 < (1 << 31))

    def level_pass(self, R, X, mode, omega=0.0):
        """mode 0: R - A X; mode 1: X + omega (R - A X) / diag (one pipelined pass)."""
        W, C = self.W, X.shape[1]
        out = torch.empty_like(X)
        _lib.call("dlp_amg_level", int(W.shape[0]), C, ptr(W.indptr), ptr(W.indices), ptr(W.data), ptr(self.diag),
                  ptr(X), ptr(R.contiguous()), C, float(omega), int(mode), ptr(out), stream())
        return out

    def residual(self, R, X):
        """R - (diag X - W X) in one row-operator launch."""
        if self.pipelined(X):
            return self.level_pass(R, X, 0)
        if self.neg_diag is None:
            d = self.diag.contiguous()
            self.neg_diag = torch.zeros_like(d)
            _lib.call("dlp_colop", int(d.shape[0]), 1, 0, -1.0, ptr(torch.ones(1, dtype=torch.float64, device=d.device)),
                      ptr(d), 1, None, ptr(self.neg_diag), 1, dtype_code(d.dtype), stream())
        W, C = self.W, X.shape[1]
        out = torch.empty_like(R)
        dt = dtype_code(X.dtype)
        _lib.call("dlp_row_operator", int(W.shape[0]), None, C, ptr(W.indptr), None, ptr(W.indices), ptr(W.data),
                  dtype_code(W.data.dtype), None, ptr(X), C, dt, ptr(self.neg_diag), 1.0, ptr(R), C, ptr(out), C, dt,
                  0, W.avg_deg, stream())
        return out


def _diag_of(W: CSR, s):
    """diag = s + W 1 in one row-operator launch (the row sums added to s)."""
    n = W.shape[0]
    dt = W.data.dtype
    ones = torch.ones((n, 1), dtype=dt, device=device())
    out = torch.empty((n, 1), dtype=dt, device=device())
    _lib.call("dlp_row_operator", n, None, 1, ptr(W.indptr), None, ptr(W.indices), ptr(W.data), dtype_code(dt), None,
              ptr(ones), 1, dtype_code(dt), None, 1.0, ptr(s.contiguous()), 1, ptr(out), 1, dtype_code(dt), 0,
              W.avg_deg, stream())
    return out[:, 0].contiguous()


def _jacobi_add(omega, T, X, d):
    """X += omega * T / d (rows), in place: the weighted-Jacobi correction (dlp_colop mode 3)."""
    C = T.shape[1]
    a = torch.full((C,), float(omega), dtype=torch.float64, device=T.device)
    _lib.call("dlp_colop", int(T.shape[0]), C, 3, 1.0, ptr(a), ptr(T), C, ptr(d), ptr(X), C, dtype_code(T.dtype),
              stream())
    return X


def _scan_ws(n):
    nb = int(_lib.load().dlp_scan_workspace_bytes(max(int(n), 1)))
    return torch.empty(nb, dtype=torch.uint8, device=device()), nb


def _handshake(W: CSR, rounds):
    """(agg int32[n] renumbered, n_aggregates) (amg.py:32-64)."""
    n = W.shape[0]
    dev = device()
    agg = torch.empty(n, dtype=torch.int32, device=dev)
    choice = torch.empty(n, dtype=torch.int32, device=dev)
    flag = torch.empty(n, dtype=torch.int32, device=dev)
    matched = torch.empty(n, dtype=torch.uint8, device=dev)
    nc = torch.zeros(1, dtype=torch.int32, device=dev)
    ws, nb = _scan_ws(n)
    _lib.call("dlp_amg_match", n, ptr(W.indptr), ptr(W.indices), ptr(W.data), dtype_code(W.data.dtype), int(rounds),
              ptr(agg), ptr(choice), ptr(flag), ptr(matched), ptr(ws), nb, ptr(nc), stream())
    return agg, int(nc.item())


def _coarsen_W(W: CSR, agg, nc):
    """Off-diagonal part of P^T W P, duplicates summed, columns sorted (amg.py:67-74)."""
    n, nnz = W.shape[0], int(W.indptr[-1].item())
    dev, dt = device(), W.data.dtype
    off = torch.empty(nc + 1, dtype=torch.int32, device=dev)
    cursor = torch.empty(nc + 1, dtype=torch.int32, device=dev)
    ci = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    cw = torch.empty(max(nnz, 1), dtype=dt, device=dev)
    ws, nb = _scan_ws(nc + 1)
    _lib.call("dlp_amg_coarsen_fill", n, nc, ptr(W.indptr), ptr(W.indices), ptr(W.data), dtype_code(dt), ptr(agg),
              ptr(off), ptr(cursor), ptr(ci), ptr(cw), ptr(ws), nb, stream())
    optr = torch.empty(nc + 1, dtype=torch.int32, device=dev)
    total = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.call("dlp_amg_coarsen_merge", nc, ptr(off), ptr(ci), ptr(cw), dtype_code(dt), ptr(optr), ptr(ws), nb,
              ptr(total), stream())
    T = int(total.item())
    oi = torch.empty(max(T, 1), dtype=torch.int32, device=dev)
    ow = torch.empty(max(T, 1), dtype=dt, device=dev)
    _lib.call("dlp_amg_coarsen_compact", nc, ptr(off), ptr(ci), ptr(cw), dtype_code(dt), ptr(optr), ptr(oi),
              ptr(ow), stream())
    return CSR(ow[:T], oi[:T], optr, (nc, nc))


def restrict(agg, X, nc):
    """P^T X: rows summed into their aggregates."""
    X = X.contiguous()
    C = 1 if X.dim() == 1 else X.shape[1]
    Y = torch.zeros((nc, C) if X.dim() == 2 else (nc,), dtype=X.dtype, device=X.device)
    _lib.call("dlp_amg_restrict", int(X.shape[0]), C, ptr(agg), ptr(X), C, ptr(Y), C, dtype_code(X.dtype), stream())
    return Y


def prolong_add(agg, Xc, X):
    """X += P Xc (in place)."""
    C = X.shape[1]
    _lib.call("dlp_amg_prolong_add", int(X.shape[0]), C, ptr(agg), ptr(Xc.contiguous()), C, ptr(X), C,
              dtype_code(X.dtype), stream())
    return X


class AMG:
    """Symmetric V-cycle preconditioner (weighted-Jacobi smoothing, exact coarse solve)."""

    def __init__(self, W: CSR, s, *, max_coarse=1500, passes=2, rounds=3, omega=0.7, nu=1,
                 max_levels=25, min_ratio=0.85, coarse_sweeps=20):
        self.omega, self.nu, self.coarse_sweeps = omega, nu, coarse_sweeps
        self.levels: list[Level] = []
        dt = W.data.dtype
        n = W.shape[0]
        s = s.to(dt).contiguous()
        cur = Level(W, s, _diag_of(W, s))
        while n > max_coarse and len(self.levels) < max_levels:
            agg = torch.arange(n, dtype=torch.int32, device=device())
            nc, Wp = n, cur.W
            for _ in range(passes):
                a2, nc2 = _handshake(Wp, rounds)
                comp = torch.empty_like(agg)
                _lib.call("dlp_amg_compose", n, ptr(agg), ptr(a2), ptr(comp), stream())
                agg = comp
                Wp = _coarsen_W(Wp, a2, nc2)
                nc = nc2
            if nc > min_ratio * n:
                break
            Wc = _coarsen_W(cur.W, agg, nc)
            sc = restrict(agg, cur.s, nc)
            cur.agg, cur.nc = agg, nc
            self.levels.append(cur)
            cur = Level(Wc, sc, _diag_of(Wc, sc))
            n = nc
        self.coarse = cur
        self.inv = None
        if n <= 4 * max_coarse:
            # the exact coarse inverse, formed and kept in FP64 on the device
            self.inv = torch.empty((n, n), dtype=torch.float64, device=device())
            scratch = torch.empty_like(self.inv)
            _lib.call("dlp_dense_operator_f64", n, ptr(cur.W.indptr), ptr(cur.W.indices), ptr(cur.W.data),
                      ptr(cur.diag), dtype_code(dt), ptr(self.inv), stream())
            _lib.call("dlp_spd_inverse_f64", n, ptr(self.inv), ptr(scratch), stream())
        self.work_per_apply = sum(2 * nu * L.W.nnz + 2 * L.W.shape[0] for L in self.levels)
        self.work_per_apply += n * n if self.inv is not None else coarse_sweeps * cur.W.nnz
        self.sizes = [L.W.shape[0] for L in self.levels] + [n]

    def _smooth(self, L: Level, X, R, sweeps):
        """sweeps of X += omega (R - A X) / diag (amg.py:114-117): one pipelined pass a sweep into
        a new block, or the residual pass then the correction in place on X."""
        for _ in range(sweeps):
            if L.pipelined(X):
                X = L.level_pass(R, X, 1, self.omega)
            else:
                _jacobi_add(self.omega, L.residual(R, X), X, L.diag)
        return X

    def _first(self, L: Level, R):
        """The first sweep from X = 0: X = omega R / diag."""
        return _jacobi_add(self.omega, R, torch.zeros_like(R), L.diag)

    def _cycle(self, lvl, R):
        if lvl == len(self.levels):
            L = self.coarse
            if self.inv is not None:        # FP64 product with the exact coarse inverse
                X = torch.empty_like(R)
                _lib.call("dlp_dense_apply_f64", int(R.shape[0]), int(R.shape[1]), ptr(self.inv), ptr(R), R.stride(0),
                          ptr(X), X.stride(0), dtype_code(R.dtype), stream())
                return X
            return self._smooth(L, self._first(L, R), R, self.coarse_sweeps - 1)
        L = self.levels[lvl]
        X = self._smooth(L, self._first(L, R), R, self.nu - 1)
        Rc = restrict(L.agg, L.residual(R, X), L.nc)
        X = prolong_add(L.agg, self._cycle(lvl + 1, Rc), X)
        return self._smooth(L, X, R, self.nu)

    def _apply_block(self, R):
        """One V-cycle on a block of at most 256 columns.  FP32 blocks run padded to a multiple
        of 4 columns (zero columns stay zero: the cycle is linear), so the level operators read
        16-byte words."""
        C = R.shape[1]
        Cp = (C + 3) // 4 * 4
        if R.dtype != torch.float32 or Cp == C or Cp > 256:
            return self._cycle(0, R)
        Rp = torch.zeros((R.shape[0], Cp), dtype=R.dtype, device=R.device)
        Rp[:, :C] = R
        return self._cycle(0, Rp)[:, :C].contiguous()

    def __call__(self, R):
        R = R.contiguous()
        if R.shape[1] <= 256:
            return self._apply_block(R)
        # the row kernels take at most 256 columns: apply the (linear) cycle per column block
        return torch.cat([self._apply_block(R[:, j:j + 256].contiguous()) for j in range(0, R.shape[1], 256)], dim=1)

"""Frontier kernels, row-major SpMM and the fused asynchronous push on gfx950.

Same Python surface as dynlp/kernels.py of the reference (`spmm_axpy` kernels.py:377-393,
`FusedPush` :396-433, `FrontierOps` :436-492, `auto_group` :365-370); the kernels behind it
are hand-written HIP (dynlp_amd/csrc) reached through the C ABI.  One 64-wide wavefront is
split into row groups of column lanes x edge lanes (any column count, no 32-column limit), so
there is no separate "rows"/"cols" kernel family and no vectorised fallback: `use_kernel` is
always True.
"""
from __future__ import annotations

import torch

from . import _lib
from .backend import CSR, device, dtype_code, ptr, stream

MAX_KERNEL_COLS = 256  # columns a row group covers (4 column chunks of 64 lanes)


def auto_group(avg_deg: float) -> int:
    """Power-of-two edge lanes close to the mean degree, in [2, 32] (kernels.py:365-370)."""
    g = 2
    while g < 32 and g * 2 <= max(avg_deg, 1.0):
        g *= 2
    return g


def _avg_deg(W: CSR, group):
    if group == "auto" or group is None:
        return W.avg_deg
    return float(group)


def _check_cols(C):
    if not 1 <= C <= MAX_KERNEL_COLS:
        raise ValueError(f"column count {C} outside [1, {MAX_KERNEL_COLS}]")


def spmm_axpy(W: CSR, X, a=None, b=1.0, group="auto"):
    """a * X + b * (W @ X) for row-major X (n x C) in one pass (kernels.py:377-393)."""
    n, C = W.shape[0], X.shape[1]
    _check_cols(C)
    if X.dtype != W.data.dtype:
        raise TypeError("W and X must share a dtype (solvers.py:36-37)")
    X = X.contiguous()
    Y = torch.empty((n, C), dtype=X.dtype, device=X.device)
    if a is not None:
        a = a.to(X.dtype).contiguous()
    _lib.call("dlp_spmm_axpy", n, C, ptr(W.indptr), ptr(W.indices), ptr(W.data), ptr(X), C, ptr(a),
              float(b), ptr(Y), C, dtype_code(X.dtype), _avg_deg(W, group), stream())
    return Y


class FusedPush:
    """Residual push with device-resident candidate lists (kernels.py:396-433 is the interface):
    rounds of take (D = R/diag, Z += D, R = 0), scatter (R[N(u)] += w D, atomics) and select
    (touched rows that still violate |R_uc| > eps_c diag_u) run back to back through
    dlp_push_rounds; the host reads the candidate count every `check_every` rounds.  R stays
    exact up to rounding; any column count up to 256; FP32 or FP64 (W, Z, R and diag share it).
    threads and blocks_per_sm are accepted for the reference's signature; the launch shapes are
    the library's."""

    def __init__(self, W: CSR, C, group="auto", threads=256, blocks_per_sm=8):
        _check_cols(C)
        n = W.shape[0]
        dev = device()
        self.W, self.C, self.n = W, C, n
        self.lists = [torch.empty(max(n, 1), dtype=torch.int32, device=dev) for _ in range(2)]
        self.counts = torch.zeros(3, dtype=torch.int32, device=dev)
        self.stamp = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
        self.touched = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        self.edges = torch.zeros(1, dtype=torch.int64, device=dev)
        self._D = {}
        self.parity = 0
        self.eps = None

    def _scratch(self, dtype):
        D = self._D.get(dtype)
        if D is None:
            D = self._D[dtype] = torch.empty((max(self.n, 1), self.C), dtype=dtype, device=device())
        return D

    def seed(self, R, diag, eps, cand=None):
        """Current candidates: the given rows, or (cand None) every row that violates."""
        self.counts.zero_()
        self.parity = 0
        self.eps = torch.as_tensor(eps, dtype=torch.float64, device=R.device).reshape(-1).contiguous()
        if cand is None:
            d64 = diag.double().contiguous()
            _lib.call("dlp_violating_rows", self.n, None, self.C, ptr(R), R.stride(0), dtype_code(R.dtype), ptr(d64),
                      ptr(self.eps), ptr(self.lists[0]), ptr(self.counts), stream())
        else:
            k = int(cand.shape[0])
            self.lists[0][:k] = cand.to(torch.int32)
            self.counts[0] = k

    def launch(self, Z, R, diag, rounds, edges=None):
        """Enqueue `rounds` rounds (no host synchronisation); edge visits go to `edges` (a device
        int64 counter) or this object's own."""
        W = self.W
        if W.data.dtype != Z.dtype or R.dtype != Z.dtype or diag.dtype != Z.dtype:
            raise TypeError("W, Z, R and diag must share a dtype")
        cur, nxt = self.lists[self.parity], self.lists[1 - self.parity]
        _lib.call("dlp_push_rounds", self.n, self.C, Z.stride(0), dtype_code(Z.dtype), ptr(W.indptr), ptr(W.indices),
                  ptr(W.data), ptr(diag), ptr(self.eps), ptr(Z), ptr(R), ptr(self._scratch(Z.dtype)), ptr(cur),
                  ptr(nxt), ptr(self.counts), ptr(self.stamp), ptr(self.touched),
                  ptr(self.edges if edges is None else edges), int(rounds), stream())
        self.parity ^= int(rounds) & 1

    def pending(self):
        """Candidates of the next round (reads one count)."""
        return int(self.counts[0].item())

    def run(self, Z, R, diag, eps, cand, max_rounds, check_every=8):
        """Push until no vertex violates or max_rounds; returns (rounds, converged, edge visits).
        cand None: start from every violating row."""
        self.edges.zero_()
        if not self.n:
            return 0, True, 0
        self.seed(R, diag, eps, cand)
        rounds, left = 0, self.pending()
        while left and rounds < max_rounds:
            step = int(min(check_every, max_rounds - rounds))
            self.launch(Z, R, diag, step)
            rounds += step
            left = self.pending()
        return rounds, left == 0, int(self.edges.item())


class FrontierOps:
    """Frontier primitives bound to one CSR matrix and column count (kernels.py:436-492)."""

    def __init__(self, W: CSR, C, group="auto", threads=256, mapping="auto"):
        _check_cols(C)
        self.W, self.C, self.threads = W, C, threads
        self.deg = _avg_deg(W, group)
        self.real = W.data.dtype
        self.mapping = mapping
        self.use_kernel = True
        self.edges = torch.zeros((), dtype=torch.int64, device=device())  # device-side edge-visit counter

    def jacobi(self, frontier, X, rhs, diag, delta):
        """New values for frontier rows and a changed-by-more-than-delta flag (kernels.py:464-474)."""
        f = frontier.to(torch.int32).contiguous()
        nf = int(f.shape[0])
        Y = torch.empty((nf, self.C), dtype=X.dtype, device=X.device)
        ch = torch.empty(nf, dtype=torch.uint8, device=X.device)
        W = self.W
        _lib.call("dlp_frontier_jacobi", nf, ptr(f), int(X.shape[0]), self.C, ptr(W.indptr), ptr(W.indices), ptr(W.data),
                  ptr(X.contiguous()), ptr(rhs.contiguous()), ptr(diag), float(delta), ptr(Y), ptr(ch),
                  ptr(self.edges), dtype_code(X.dtype), self.deg, stream())
        return Y, ch.bool()

    def push(self, frontier, D, R, mark):
        """R += W[:, frontier] @ D (W symmetric) and mark touched vertices (kernels.py:476-486)."""
        f = frontier.to(torch.int32).contiguous()
        W = self.W
        _lib.call("dlp_push_scatter", int(f.shape[0]), ptr(f), self.C, ptr(W.indptr), ptr(W.indices),
                  ptr(W.data), ptr(D.contiguous()), ptr(R), ptr(mark), ptr(self.edges), dtype_code(R.dtype),
                  self.deg, stream())

    def mark_neighbors(self, rows, mark):
        """mark[N(rows)] = 1 (kernels.py:488-492)."""
        r = rows.to(torch.int32).contiguous()
        W = self.W
        _lib.call("dlp_mark_neighbors", int(r.shape[0]), ptr(r), ptr(W.indptr), ptr(W.indices), ptr(mark),
                  ptr(self.edges), stream())

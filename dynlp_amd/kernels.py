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
    """Asynchronous residual push with resident buffers (kernels.py:396-433).

    Takes the residual of every violating vertex (|r_u| > eps_c * diag_u), adds r_u / diag_u
    to z_u and scatters it to the neighbours with hardware float atomics, keeping r exact up
    to rounding (chaotic relaxation; converges for this M-matrix).  Any column count.
    """

    def __init__(self, W: CSR, C, group="auto", threads=256, blocks_per_sm=8):
        _check_cols(C)
        n = W.shape[0]
        self.W, self.C = W, C
        z = lambda *s: torch.zeros(s, dtype=torch.int32, device=device())  # noqa: E731
        self.front, self.touched, self.mark, self.fcnt, self.tcnt = z(2, n), z(n), z(n), z(2), z(2)
        self.edges = torch.zeros(1, dtype=torch.int64, device=device())
        sms = torch.cuda.get_device_properties(device()).multi_processor_count
        self.grid = int(sms * blocks_per_sm)
        if threads != 256:
            raise ValueError("FusedPush runs 256-thread workgroups")

    def run(self, Z, R, diag, eps, cand, max_rounds, check_every=8):
        """Push until no vertex violates; returns (rounds, converged, edge visits)."""
        W = self.W
        n0 = int(cand.shape[0])
        self.edges.zero_()
        if not n0:
            return 0, True, 0
        dt = dtype_code(Z.dtype)
        if W.data.dtype != Z.dtype or R.dtype != Z.dtype or diag.dtype != Z.dtype:
            raise TypeError("W, Z, R and diag must share a dtype")
        self.front[0, :n0] = cand.to(torch.int32)
        self.fcnt.zero_()
        self.tcnt.zero_()
        self.fcnt[0] = n0
        eps = torch.as_tensor(eps, dtype=Z.dtype, device=Z.device).contiguous()
        rounds, s = 0, stream()
        fc, tc = self.fcnt.data_ptr(), self.tcnt.data_ptr()
        while rounds < max_rounds:
            p, q = rounds & 1, 1 - (rounds & 1)
            _lib.call("dlp_fused_push_round", ptr(W.indptr), ptr(W.indices), ptr(W.data), ptr(diag),
                      ptr(self.front[p]), fc + 4 * p, fc + 4 * q, self.C, ptr(Z), ptr(R), ptr(self.mark),
                      ptr(self.touched), tc + 4 * p, ptr(self.edges), dt, self.grid, s)
            _lib.call("dlp_compact_frontier", ptr(self.touched), tc + 4 * p, tc + 4 * q, ptr(self.mark),
                      self.C, ptr(R), ptr(diag), ptr(eps), ptr(self.front[q]), fc + 4 * q, dt, self.grid, s)
            rounds += 1
            if rounds % check_every == 0 and int(self.fcnt[q]) == 0:
                break
        return rounds, int(self.fcnt[rounds & 1]) == 0, int(self.edges[0])


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

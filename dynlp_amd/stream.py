"""Streaming protocol and the per-batch linear system, resident on the device.

Per batch, the harmonic system on the active unlabeled set U (reference stream.py:1-6):
    (diag - W_UU) F_U = rhs,   diag = s + rowsum(W_UU),   s = w(u, L) + eta
with h (A h = diag) riding along as one extra column for the certificate |F - F*| <= rho*h
(solvers.py:10-12).

The reference streams vertex arrivals over a fixed graph and patches edge weights in place
(`IncrementalStream`, stream.py:180-346).  This module streams edge insertions and deletions:
`EdgeStreamLP` keeps the graph in a slack CSR on the device (global vertex ids, rows outside
U inert), applies each batch in place, re-grounds, refreshes the changed rows' FP64 residual
and repairs it by the round-synchronous push; the FP64 plain path (`fp64=True`) repairs with
the reference's host-driven push loop (solvers.py:100-125) instead.
"""
from __future__ import annotations

import ctypes as ct
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .backend import device, ptr, stream
from .graphs import EdgeBatch, Graph, check_batch, normalize_batch, pack_ops
from .solvers import System  # noqa: F401  (the batch system, reference stream.py:17-58)

_OPS_PER_CALL = 1 << 19     # operations one dlp_lp_apply_ops call takes (its pair table's capacity / 2)
# storage of the global solve's search directions -> dlp_lp_set_pcg_directions kind
DIRECTION_KINDS = {"fp32": 0, "bf16": 1, "fp16": 2, "fp8": 3}
# "auto": FP8 from this many columns (label columns + h) up, FP16 below.  Measured on one MI355X:
# 65 columns (SBM-1M, k = 64) FP8 6-7 % more edge updates per second; 17 and 33 columns (ER-1M,
# k = 16; SBM-300k, k = 32) FP16 as fast or 6-7 % faster (narrow rows gain little from fewer
# bytes per gather and FP8's step takes a few more iterations)
FP8_MIN_COLUMNS = 48


def resolve_directions(kind: str, C1: int) -> str:
    """The concrete direction storage for a kind name ("auto" by column count)."""
    if kind == "auto":
        return "fp8" if C1 >= FP8_MIN_COLUMNS else "fp16"
    if kind not in DIRECTION_KINDS:
        raise ValueError(f"pcg_directions {kind!r}: 'auto', 'fp32', 'bf16', 'fp16' or 'fp8'")
    return kind


@dataclass
class Stats:
    """Per-batch statistics (solvers.py:25-33 `Stats`, plus latency parts)."""
    iters: int = 0              # push rounds
    edges: int = 0              # edge visits
    ms: float = 0.0
    setup_ms: float = 0.0
    path: str = ""
    cert: float = float("nan")
    hmax: float = float("nan")
    rho: float = float("nan")   # max_c<C |R_uc|/diag_u over rows refreshed at the end of the batch
    refine: int = 0             # FP64 verification rounds that found violators
    touched: int = 0            # rows touched by pushes
    changed: int = 0            # rows whose equation changed
    toggled: int = 0            # vertices that changed groundedness
    extra: dict = field(default_factory=dict)


def slack_layout(indptr, indices, data, slack_min=8, slack_frac=0.25):
    """Host ingest: a CSR expanded to per-row capacity deg + max(slack_min, slack_frac*deg)."""
    n = indptr.shape[0] - 1
    deg = np.diff(indptr).astype(np.int64)
    cap = deg + np.maximum(slack_min, np.ceil(slack_frac * deg).astype(np.int64))
    new_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(cap, out=new_ptr[1:])
    if new_ptr[-1] >= 2**31:
        raise ValueError("slack store exceeds int32 offsets")
    rows = np.repeat(np.arange(n, dtype=np.int64), deg)
    pos = new_ptr[rows] + (np.arange(indices.shape[0], dtype=np.int64) - indptr[:-1].astype(np.int64)[rows])
    col = np.full(int(new_ptr[-1]), -1, dtype=np.int32)
    w = np.zeros(int(new_ptr[-1]), dtype=np.float32)
    col[pos] = indices
    w[pos] = data
    return new_ptr.astype(np.int32), deg.astype(np.int32), col, w


class EdgeStreamLP:
    """One graph's resident label-propagation state on the device.

    eps: stop threshold on the label columns, |R_uc| <= eps * diag_u (rho of solvers.py:52-54);
    eps_h: the h column's (solvers.py:247 default 0.05).  fp64=False is the mixed-precision
    service path (FP32 Z/R and push, FP64 residual verification and refinement); fp64=True keeps
    Z/R in FP64 and repairs with the host-driven push rounds of the reference.  pcg_directions:
    storage of the global solve's search directions: "auto" (default: "fp8" from 48 columns up,
    "fp16" below), "fp16" (half the gather bytes, a power-of-two scale per column, 11 significant
    bits, step lengths exact for the rounded direction), "fp8" (E4M3: a quarter of the bytes, 4
    significant bits, the scale from a bound on each column's largest value), "bf16" (half the
    bytes, 8 significant bits) or "fp32".  pcg_coarse: aggregates of
    the two-level preconditioner of the global solve (runs of consecutive rows; 0 = Jacobi only),
    its coarse operator rebuilt from the current equations every coarse_refresh global solves.
    pcg_iters_per_launch: global-solve iterations per replayed graph (iterations past
    convergence return at once; with 16-bit directions each group adds its directions to Z once,
    at its end), 7 by default: a 10,000-edge repair of SBM-1M takes 6-7."""

    def __init__(self, g: Graph, eps=1e-6, eps_h=0.05, fp64=False, eta=0.0, slack_min=8,
                 slack_frac=0.25, push_scale=0.9, rounds_per_launch=4, use_graph=True, max_rounds=200_000,
                 push_frac=0.05, push_max_rows=4096, push_budget=1.0, push_mass=0.15, pcg_iters_per_launch=7,
                 pcg_max_iter=20_000, pcg_mode="column", incremental_ground=True, pcg_directions="auto", pcg_coarse=128,
                 coarse_refresh=64, fp64_repair="push"):
        self.g, self.n, self.K = g, g.n, g.K
        self.C = 1 if g.K == 2 else g.K
        self.C1 = self.C + 1
        self.eps, self.eps_h, self.fp64, self.eta = float(eps), float(eps_h), bool(fp64), float(eta)
        self.rounds_per_launch, self.use_graph, self.max_rounds = rounds_per_launch, use_graph, max_rounds
        self.push_scale = push_scale
        self.push_frac, self.push_budget, self.push_max_rows = push_frac, push_budget, push_max_rows
        self.push_mass = float(push_mass)
        self.pcg_iters_per_launch, self.pcg_max_iter = pcg_iters_per_launch, pcg_max_iter
        if pcg_mode not in ("column", "block"):
            raise ValueError(f"pcg_mode {pcg_mode!r}: 'column' or 'block'")
        self.pcg_mode = pcg_mode
        self.incremental_ground = bool(incremental_ground)
        if fp64_repair not in ("push", "pcg"):
            raise ValueError(f"fp64_repair {fp64_repair!r}: 'push' or 'pcg'")
        self.fp64_repair = fp64_repair
        dev = device()
        ip, ix, d = g.csr()
        sp_, rl, col, w = slack_layout(ip, ix, d, slack_min, slack_frac)
        labcnt = np.zeros(self.n, dtype=np.int32)
        np.add.at(labcnt, np.repeat(np.arange(self.n), np.diff(ip)), g.labeled[ix].astype(np.int32))
        t = lambda a, dt: torch.as_tensor(a).to(dev, dt).contiguous()  # noqa: E731
        self.indptr, self.rowlen = t(sp_, torch.int32), t(rl, torch.int32)
        self.col, self.w = t(col, torch.int32), t(w, torch.float32)
        self.labeled = t(g.labeled.astype(np.uint8), torch.uint8)
        self.y = t(g.y, torch.int32)
        self.solved = torch.zeros(self.n, dtype=torch.uint8, device=dev)
        self.labcnt = t(labcnt, torch.int32)
        self.ld = (self.C1 + 3) // 4 * 4    # 16-byte aligned rows: 128-bit neighbour-row gathers
        self.prior = 0.5 if self.K == 2 else 1.0 / self.K
        Z = torch.zeros((self.n, self.ld), dtype=torch.float32, device=dev)
        Z[:, : self.C] = self.prior
        Z[:, self.C] = 1.0
        self.Z32 = Z
        self.R32 = torch.zeros((self.n, self.ld), dtype=torch.float32, device=dev)
        self.diag64 = torch.ones(self.n, dtype=torch.float64, device=dev)
        self.diag32 = torch.ones(self.n, dtype=torch.float32, device=dev)
        self.Z64 = self.R64 = None
        if self.fp64:
            self.Z64 = Z.double()
            self.R64 = torch.zeros((self.n, self.ld), dtype=torch.float64, device=dev)
            self.cand = torch.zeros(self.n, dtype=torch.int32, device=dev)
            self.cand_n = torch.zeros(1, dtype=torch.int32, device=dev)
            self._rho_dev = torch.zeros(self.C1, dtype=torch.float64, device=dev)
        self._desc = _lib.LPDesc(n=self.n, C1=self.C1, K2=int(self.K == 2), pad_=0, eta=self.eta,
                                 prior=self.prior, indptr=ptr(self.indptr), rowlen=ptr(self.rowlen),
                                 col=ptr(self.col), w=ptr(self.w), labeled=ptr(self.labeled), y=ptr(self.y),
                                 solved=ptr(self.solved), labcnt=ptr(self.labcnt), Z32=ptr(self.Z32),
                                 R32=ptr(self.R32), ld=self.ld, diag64=ptr(self.diag64), diag32=ptr(self.diag32),
                                 Z64=ptr(self.Z64), R64=ptr(self.R64))
        h = ct.c_void_p()
        _lib.call("dlp_lp_create", ct.byref(self._desc), ct.byref(h))
        self.h = h
        self._set_eps()
        pcg_directions = resolve_directions(pcg_directions, self.C1)
        self.pcg_directions = pcg_directions
        _lib.call("dlp_lp_set_pcg_directions", self.h, 0 if self.fp64 else DIRECTION_KINDS[pcg_directions])
        self.pcg_coarse = int(pcg_coarse)
        _lib.call("dlp_lp_set_coarse", self.h, self.pcg_coarse, int(coarse_refresh))
        if not self.fp64:
            # every buffer of the global solve now, so that no request allocates device memory
            _lib.call("dlp_lp_reserve", self.h, int(pcg_iters_per_launch), int(pcg_mode == "block"), stream())
        self._status = np.zeros(19, dtype=np.int64)
        self._rho = np.zeros(self.C1, dtype=np.float64)
        self.batches = 0
        self.scratch3 = None
        self._stage = self._stage_dev = None
        self.nnz = int(ix.shape[0])

    @staticmethod
    def footprint(g: Graph, fp64=False, slack_min=8, slack_frac=0.25, pcg_directions="auto",
                  pcg_coarse=128, pcg_iters_per_launch=7, pcg_mode="column", **_):
        """Device bytes of a resident graph built with these settings, before its first
        request: the arrays this object allocates (slack store, masks, labels, FP32 state, FP64
        diagonal; the FP64 state when fp64) plus the library handle with its global solve
        reserved (dlp_lp_handle_bytes, equal to what the handle then reports).  A batch of up
        to 2^16 words stages without growing; a regrow after an overflow allocates a new store
        beside the old one for the duration of the request."""
        n = int(g.n)
        C1 = (1 if g.K == 2 else g.K) + 1
        ld = (C1 + 3) // 4 * 4
        deg = np.bincount(g.u, minlength=n) + np.bincount(g.v, minlength=n)
        cap = int((deg + np.maximum(slack_min, np.ceil(slack_frac * deg))).sum())
        b = 4 * (n + 1) + 4 * n + 8 * cap                  # indptr, rowlen, col + w
        b += n + 4 * n + n + 4 * n                         # labeled, y, solved, labcnt
        b += 2 * 4 * n * ld + 8 * n + 4 * n                # Z32, R32, diag64, diag32
        b += 4 * (1 << 16)                                 # batch staging
        if fp64:
            b += 2 * 8 * n * ld + 4 * n + 4 + 8 * C1       # Z64, R64, candidates, rho
        kinds = DIRECTION_KINDS
        pcg_directions = resolve_directions(pcg_directions, C1)
        hb = _lib.load().dlp_lp_handle_bytes(n, C1, ld, int(fp64), kinds[pcg_directions], int(pcg_iters_per_launch),
                                              int(pcg_coarse), int(pcg_mode == "block"))
        if hb < 0:
            raise ValueError("dlp_lp_handle_bytes: bad sizes")
        return int(b + hb)

    def device_bytes(self):
        """Device bytes the library handle holds (its own buffers)."""
        out = ct.c_int64(0)
        _lib.call("dlp_lp_device_bytes", self.h, ct.byref(out))
        return int(out.value)

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and self.h.value:
                _lib.call("dlp_lp_destroy", self.h)
                self.h = None
        except Exception:
            pass

    def _set_eps(self):
        eps = np.array([self.eps] * self.C + [self.eps_h], dtype=np.float64)
        _lib.call("dlp_lp_set_eps", self.h, eps.ctypes.data, float(self.push_scale), stream())

    # -- device steps ---------------------------------------------------------------------------
    def status(self):
        _lib.call("dlp_lp_status", self.h, self._status.ctypes.data, self._rho.ctypes.data, stream())
        s = self._status
        return dict(changed=int(s[0]), missing=int(s[1]), overflow=int(s[2]), toggled=int(s[3]), viol=int(s[4]),
                    touched=int(s[6]), round=int(s[7]), edges=int(s[8]), cand_total=int(s[9]),
                    push_total=int(s[10]), scatter_total=int(s[11]),
                    rounds_launched=int(s[13]), push_us=int(s[14]), mass=float(s[15:16].view(np.float64)[0]),
                    hmax=float(s[12:13].view(np.float64)[0]), invalid=int(s[16]), duplicates=int(s[17]),
                    refresh_kernel=int(s[18]))

    def _push(self, budget=0):
        r, ok = ct.c_int64(0), ct.c_int32(0)
        _lib.call("dlp_lp_push", self.h, self.max_rounds, self.rounds_per_launch, int(self.use_graph),
                  int(budget), ct.byref(r), ct.byref(ok), stream())
        return int(r.value), bool(ok.value)

    def _pcg(self, verify=False):
        """The global solve; with verify (column PCG only) the library also runs the FP64
        verification pass behind its last group of iterations, before it reads their status.
        Returns (iterations, last violation, whether the verification ran)."""
        it, v = ct.c_int64(0), ct.c_double(0.0)
        verify = verify and self.pcg_mode != "block"
        fn = "dlp_lp_bpcg" if self.pcg_mode == "block" else ("dlp_lp_pcg_verify" if verify else "dlp_lp_pcg")
        _lib.call(fn, self.h, self.pcg_max_iter, self.pcg_iters_per_launch, int(self.use_graph), 50,
                  ct.byref(it), ct.byref(v), stream())
        return int(it.value), float(v.value), verify

    def coarse_info(self):
        """The global solve's coarse operator: aggregates (0: none), their minimum rows, solves
        since it was built, and the row ranges (bounds, na + 1 ints)."""
        out = np.zeros(3, dtype=np.int32)
        bounds = np.zeros(129, dtype=np.int32)
        _lib.call("dlp_lp_coarse_info", self.h, out.ctypes.data, bounds.ctypes.data, stream())
        return dict(aggregates=int(out[0]), min_rows=int(out[1]), age=int(out[2]), bounds=bounds[: out[0] + 1].copy())

    def pcg_stats(self):
        out = np.zeros(2, dtype=np.int64)
        _lib.call("dlp_lp_pcg_stats", self.h, out.ctypes.data)
        return dict(pcg_iters=int(out[0]), pcg_us=int(out[1]))

    def _push64(self):
        """Host-driven FP64 rounds (solvers.py:108-116); one host sync per round."""
        lst, cnt = ct.c_void_p(), ct.c_void_p()
        _lib.call("dlp_lp_violators", self.h, ct.byref(lst), ct.byref(cnt))
        st = self.status()
        nc, src, rounds = st["viol"], lst.value, 0
        bufs = [self.cand, torch.empty_like(self.cand)]
        while nc and rounds < self.max_rounds:
            _lib.call("dlp_lp_push64_round", self.h, nc, src, stream())
            dst = bufs[rounds & 1]
            _lib.call("dlp_lp_push64_next", self.h, ptr(dst), ptr(self.cand_n), stream())
            nc = int(self.cand_n.item())
            src = dst.data_ptr()
            rounds += 1
        return rounds

    def _pcg64(self):
        """The plain FP64 path's global repair: column PCG (Jacobi preconditioner, no graphs) on
        the correction A D = R64 over the solved rows, then Z64 += D; one FP64 violation read
        per iteration.  Returns the iterations."""
        from .solvers import Stats as _S, _ColumnCG, cg_solve
        n, C1, ld, dev = self.n, self.C1, self.ld, device()
        D = torch.zeros((n, ld), dtype=torch.float64, device=dev)[:, :C1]
        eps = np.array([self.eps] * self.C + [self.eps_h])
        op = (self.indptr, self.rowlen, self.col, self.w, self.solved, self.nnz / max(n, 1), self.diag64)
        cg = _ColumnCG(None, D, self.R64[:, :C1], eps, operator=op)

        def jacobi(R):
            Y = torch.empty((n, ld), dtype=torch.float64, device=dev)[:, :C1]
            _lib.call("dlp_colop", n, C1, 2, 1.0, None, ptr(R), ld, ptr(self.diag64), ptr(Y), ld, _lib.F64, stream())
            return Y

        st = _S()
        cg_solve(cg, jacobi, self.pcg_max_iter, st, self.nnz)
        ones = torch.ones(C1, dtype=torch.float64, device=dev)
        _lib.call("dlp_colop", n, C1, 0, 1.0, ptr(ones), ptr(D), ld, None, ptr(self.Z64), ld, _lib.F64, stream())
        return st.iters

    def push_planned(self, s):
        """Whether the mixed path's repair of the current batch starts with a push: few
        violators (at most push_frac*n and push_max_rows) and a residual mass (the violators'
        summed max |R|/(eps*diag), about the row pushes a push-only repair takes) small next to
        the graph (a global solve costs a few dozen passes over all rows).  A batch of more
        edges than push_max_rows changes more rows than a push may start from: it goes to the
        global solve without reading the violator count (s may then be None)."""
        if self.fp64 or getattr(self, "_batch_edges", 0) > self.push_max_rows or s is None:
            return False
        return s["viol"] <= min(self.push_frac * self.n, self.push_max_rows) and s["mass"] <= self.push_mass * self.n

    def push_budget_edges(self):
        return int(self.push_budget * max(self.nnz, 1))

    def _repair(self, st: Stats, force_global=False, pushed=None, s0=None):
        """Mixed path (DynLP+ policy, solvers.py:272-288 with refinement :322-337): push while
        the violating set is small and within the edge budget, else the global PCG; then FP64
        verification of every row the repair touched, pushing again until it certifies.
        pushed = (rounds, converged) of a push of this batch already run (the service's push
        across graphs); s0 = the status after the changed-row refresh, when already read.  FP64
        path: the reference's incremental push to completion
        (solvers.py:375-389), or (fp64_repair="pcg") an FP64 global column PCG first."""
        if self.fp64:
            if self.fp64_repair == "pcg":
                st.path = "fp64-pcg"
                st.extra["pcg_iters"] = st.extra.get("pcg_iters", 0) + self._pcg64()
                _lib.call("dlp_lp_refresh", self.h, 2, 1, stream())   # FP64 residual of every row
            st.iters += self._push64()
            s = self.status()
            # the FP64 kept residual is exact up to rounding: read its column maxima
            _lib.call("dlp_rho_cols", self.n, None, self.C1, ptr(self.R64), self.ld, _lib.F64, ptr(self.diag64),
                      ptr(self._rho_dev), stream())
            self._rho[:] = self._rho_dev.cpu().numpy()
            return s
        self._rho_all = False
        ok, tried = False, pushed is not None
        if pushed is not None:
            st.iters += pushed[0]
            ok = pushed[1]
        elif not force_global:
            s = s0
            if s is None and getattr(self, "_batch_edges", 0) <= self.push_max_rows:
                s = self.status()
            if self.push_planned(s):
                tried = True
                r, ok = self._push(budget=self.push_budget_edges())
                st.iters += r
        if not ok:
            st.path = "push+pcg" if tried else "pcg"
            it, _, verified = self._pcg(verify=True)
            st.extra["pcg_iters"] = st.extra.get("pcg_iters", 0) + it
            # FP64 residual of every solved row; their equations are those of the changed-row
            # refresh, so the verification pass (mode 3) suffices (the column PCG ran it already)
            if not verified:
                _lib.call("dlp_lp_refresh", self.h, 3, 0, stream())
            s = self.status()
            if s["viol"] == 0:
                # certified on every row: the pass's column maxima are rho over all rows
                self._rho_all = True
                return s
            r, ok = self._push()
            st.iters += r
        for it in range(8):
            _lib.call("dlp_lp_refresh", self.h, 1, 0, stream())
            s = self.status()
            if s["viol"] == 0:
                return s
            st.refine += 1
            r, _ = self._push()
            st.iters += r
        return self.status()

    def _upload(self, b: EdgeBatch):
        """The batch's five arrays packed into one pinned staging buffer and sent with one copy
        (the previous request has finished with both buffers: every request ends in a sync).
        Returns the device views and the batch as applied (repeated pairs collapsed)."""
        b = normalize_batch(b, self.n, max_insertions=1 << 19)   # the device apply's per-call limit
        du, dv, iu, iv, iw = self._stage_arrays(b.del_u, b.del_v, b.ins_u, b.ins_v,
                                                np.asarray(b.ins_w, dtype=np.float32).view(np.int32))
        return (du, dv, iu, iv, iw.view(torch.float32)), b

    def _stage_arrays(self, *arrays):
        """int32 host arrays packed into the pinned staging buffer and sent with one copy;
        returns their device views (valid until the next staging: callers synchronise first)."""
        total = sum(len(a) for a in arrays)
        if self._stage is None or self._stage.numel() < total:
            cap = max(total, 1 << 16)
            self._stage = torch.empty(cap, dtype=torch.int32, pin_memory=True)
            self._stage_dev = torch.empty(cap, dtype=torch.int32, device=device())
        h = self._stage.numpy()
        offs = [0]
        for a in arrays:
            h[offs[-1]:offs[-1] + len(a)] = a
            offs.append(offs[-1] + len(a))
        d = self._stage_dev
        d[:total].copy_(self._stage[:total], non_blocking=True)
        return [d[offs[i]:offs[i + 1]] for i in range(len(arrays))]

    def _regrow(self, ins, extra):
        iu, iv, iw = ins
        new_ptr = torch.empty(self.n + 1, dtype=torch.int32, device=device())
        total = ct.c_int64(0)
        _lib.call("dlp_lp_regrow_plan", self.h, int(extra), 0.25, ptr(new_ptr), ct.byref(total), stream())
        col = torch.empty(int(total.value), dtype=torch.int32, device=device())
        w = torch.empty(int(total.value), dtype=torch.float32, device=device())
        if self.scratch3 is None:
            self.scratch3 = torch.empty(3 << 20, dtype=torch.int32, device=device())
        _lib.call("dlp_lp_regrow_commit", self.h, ptr(new_ptr), ptr(col), ptr(w), int(iu.shape[0]), ptr(iu),
                  ptr(iv), ptr(iw), ptr(self.scratch3), stream())
        old = (self.indptr, self.col, self.w)
        self.indptr, self.col, self.w = new_ptr, col, w
        # The copy above still reads the old store.  Its memory returns to the caching allocator
        # of the stream that allocated it, which may be another thread's: that thread could take
        # it before the copy has run.  Free it only once this stream is done with it.
        stream_obj = torch.cuda.current_stream()
        stream_obj.synchronize()
        del old

    def resident_bytes(self):
        """Device bytes this state holds now: its arrays (the store as regrown) and the
        library handle's buffers."""
        arrs = [self.indptr, self.rowlen, self.col, self.w, self.labeled, self.y, self.solved, self.labcnt,
                self.Z32, self.R32, self.diag64, self.diag32, self.Z64, self.R64, self._stage_dev, self.scratch3]
        return sum(int(a.numel() * a.element_size()) for a in arrs if a is not None) + self.device_bytes()

    # -- public ---------------------------------------------------------------------------------
    def initial_solve(self):
        """Ground, refresh every row, solve from F = prior, h = 1: global PCG then push (mixed
        path), or the FP64 push (FP64 path); finishes with an FP64 residual of every row."""
        st = Stats(path="initial")
        t0 = time.perf_counter()
        _lib.call("dlp_lp_begin_batch", self.h, stream())
        _lib.call("dlp_lp_ground", self.h, stream())
        prec = 1 if self.fp64 else 0
        _lib.call("dlp_lp_refresh", self.h, 2, prec, stream())
        s = self._repair(st, force_global=not self.fp64)
        if not self.fp64:
            for tries in range(21):
                _lib.call("dlp_lp_refresh", self.h, 2, 0, stream())
                s = self.status()
                if not s["viol"]:
                    break
                if tries == 20:
                    raise RuntimeError(f"initial solve: {s['viol']} rows still violate after refinement")
                st.iters += self._push()[0]
        st.ms = (time.perf_counter() - t0) * 1e3
        self._finish(st, s)
        return st

    def load_state(self, Z):
        """Start from a given [F, h] block (host or device, n x C1) and re-solve to eps: the FP64
        residual of every row, then the repair.  Used to hand one path's solution to another."""
        Zt = torch.as_tensor(Z).to(device())[:, : self.C1]
        if self.fp64:
            self.Z64[:, : self.C1].copy_(Zt.double())
        else:
            self.Z32[:, : self.C1].copy_(Zt.float())
        st = Stats(path="load")
        _lib.call("dlp_lp_begin_batch", self.h, stream())
        _lib.call("dlp_lp_ground", self.h, stream())
        _lib.call("dlp_lp_refresh", self.h, 2, 1 if self.fp64 else 0, stream())
        s = self._repair(st)
        self._finish(st, s)
        return st

    def apply_and_solve(self, b: EdgeBatch) -> Stats:
        """Apply one batch of edge updates and repair the solution to eps (certified in FP64).
        Ragged arrays, more than 2^19 insertions, insertion weights that are not finite and
        positive (ValueError) and ids outside [0, n) (IndexError) are refused before any device
        work; a pair repeated within the deletions or within the insertions counts once (the
        last insertion's weight)."""
        req = self.begin_request(b)
        return self.end_request(req)

    def begin_request(self, b: EdgeBatch):
        """First half of apply_and_solve: upload, apply on the slack store, re-ground, refresh
        the changed rows' FP64 residual.  Returns the request's state for end_request; its
        "status" holds the refresh's counters when a push could repair it (push_planned)."""
        st = Stats(path="fp64-push" if self.fp64 else "push")
        t0 = time.perf_counter()
        (du, dv, iu, iv, iw), b = self._upload(b)
        self._batch_edges = int(du.shape[0]) + int(iu.shape[0])
        _lib.call("dlp_lp_begin_batch", self.h, stream())
        _lib.call("dlp_lp_apply", self.h, int(du.shape[0]), ptr(du), ptr(dv), int(iu.shape[0]), ptr(iu), ptr(iv),
                  ptr(iw), stream())
        if self.incremental_ground:
            _lib.call("dlp_lp_ground_incremental", self.h, int(du.shape[0]), ptr(du), ptr(dv), int(iu.shape[0]),
                      stream())
        else:
            _lib.call("dlp_lp_ground", self.h, stream())
        _lib.call("dlp_lp_refresh", self.h, 0, 1 if self.fp64 else 0, stream())
        req = {"st": st, "t0": t0, "ins": (iu, iv, iw), "ins_rows": (b.ins_u, b.ins_v), "status": None}
        if not (self.fp64 or self._batch_edges > self.push_max_rows):
            req["status"] = self._settle(req)
        return req

    def _settle(self, req):
        """The status after the first half, with an overflowed store regrown first: the
        overflowed insertions re-inserted, the grounding and the changed-row refresh redone, so
        that the repair sees the whole batch (one regrow makes room for every insertion of the
        call in any one row)."""
        s = self.status()
        if s["overflow"]:
            self._regrow(req["ins"], self._regrow_extra(req["ins_rows"]))
            _lib.call("dlp_lp_ground", self.h, stream())
            _lib.call("dlp_lp_refresh", self.h, 0, 1 if self.fp64 else 0, stream())
            s = self.status()
        return s

    def begin_group(self, batches):
        """begin_request for several batches applied in order, as one repair: their operations
        go to the device as one ordered sequence (dlp_lp_apply_ops; each pair ends as its last
        operation leaves it, no host-side merge), in calls of at most 2^19 operations (a call
        that overflows a row is regrown before the next one).  Each batch is checked as
        apply_and_solve checks it; a pair repeated within a batch is fine."""
        st = Stats(path="fp64-push" if self.fp64 else "push")
        t0 = time.perf_counter()
        for b in batches:
            check_batch(b, self.n)
        u, v, w = pack_ops(batches)
        nops = int(u.shape[0])
        self._batch_edges = nops
        _lib.call("dlp_lp_begin_batch", self.h, stream())
        lo = 0
        ops = None
        while True:
            hi = min(nops, lo + _OPS_PER_CALL)
            du, dv, dw = self._stage_arrays(u[lo:hi], v[lo:hi], w[lo:hi].view(np.int32))
            ops = (du, dv, dw.view(torch.float32))
            _lib.call("dlp_lp_apply_ops", self.h, hi - lo, ptr(ops[0]), ptr(ops[1]), ptr(ops[2]), stream())
            ins_rows = (u[lo:hi][w[lo:hi] > 0], v[lo:hi][w[lo:hi] > 0])
            if hi >= nops:
                break
            # the next call's operations may touch this call's overflowed pairs, and the overflow
            # list names entries of this call's arrays: regrow before going on (the status read
            # also frees the staging buffer)
            if self.status()["overflow"]:
                self._regrow(ops, self._regrow_extra(ins_rows))
            lo = hi
        if self.incremental_ground and nops <= _OPS_PER_CALL:
            # every endpoint is checked (a superset of the deletions' endpoints)
            _lib.call("dlp_lp_ground_incremental", self.h, nops, ptr(ops[0]), ptr(ops[1]), nops, stream())
        else:
            _lib.call("dlp_lp_ground", self.h, stream())
        _lib.call("dlp_lp_refresh", self.h, 0, 1 if self.fp64 else 0, stream())
        req = {"st": st, "t0": t0, "ins": ops, "ins_rows": ins_rows, "status": None}
        if not (self.fp64 or nops > self.push_max_rows):
            req["status"] = self._settle(req)
        return req

    def apply_group_and_solve(self, batches) -> Stats:
        """Several batches applied in order and repaired once (begin_group + end_request)."""
        return self.end_request(self.begin_group(batches))

    @staticmethod
    def _regrow_extra(ins_rows):
        """Slack per row that holds every insertion of the call in any one row."""
        r = np.concatenate([np.asarray(ins_rows[0]), np.asarray(ins_rows[1])])
        return max(8, int(np.bincount(r).max()) if r.size else 0)

    def end_request(self, req, pushed=None):
        """Second half: the repair (with pushed = (rounds, converged) of a push already run on
        this batch), a regrow after overflowed insertions, and the request's statistics."""
        st = req["st"]
        prec = 1 if self.fp64 else 0
        if req["status"] is None:
            # a large batch: settled here, in the thread that repairs it (no status was read
            # after the first half): an overflowed store is regrown before the repair, which
            # then goes straight to the global solve
            self._settle(req)
        s = self._repair(st, pushed=pushed, s0=req["status"])
        if s["invalid"] or s["duplicates"]:
            raise RuntimeError(f"device apply refused {s['invalid']} pairs and merged {s['duplicates']} "
                               "repeated entries that the host checks had passed")
        if s["overflow"]:
            # (the apply's overflow is regrown before the repair; this only guards the counter)
            self._regrow(req["ins"], self._regrow_extra(req["ins_rows"]))
            _lib.call("dlp_lp_ground", self.h, stream())   # re-inserted edges can re-ground rows
            _lib.call("dlp_lp_refresh", self.h, 0, prec, stream())
            s = self._repair(st)
        st.ms = (time.perf_counter() - req["t0"]) * 1e3
        if s["missing"]:
            st.extra["missing_deletions"] = s["missing"]
        self._finish(st, s)
        self.batches += 1
        return st

    def _finish(self, st, s):
        """Request statistics, with the certificate of solvers.py:295-301 as the reference's
        incremental solver reports it per batch (solvers.py:421-422): column maxima rho_c of
        |R_uc|/diag_u over every row, h_est = max h over every row (the reference parks rows
        outside the solved set at h = 1, solvers.py:365-368; solved rows have h >= 1),
        hmax = h_est / (1 - rho_h) and cert = max_c<C rho_c * hmax (inf while rho_h > eps_h).
        After a clean verification of every solved row both come with its status; otherwise one
        pass over the kept residual and one over the h column, read back together."""
        st.edges, st.touched, st.changed, st.toggled = s["edges"], s["touched"], s["changed"], s["toggled"]
        C, C1 = self.C, self.C1
        if getattr(self, "_rho_all", False) and not self.fp64 and s.get("hmax", 0.0) > 0.0:
            rc, h_est = self._rho, s["hmax"]
        else:
            dev = device()
            if getattr(self, "_cert_buf", None) is None:
                self._cert_buf = torch.zeros(C1 + 1, dtype=torch.float64, device=dev)
                self._ones64 = torch.ones(self.n, dtype=torch.float64, device=dev)
            buf = self._cert_buf
            # the FP64 path's column maxima come from a pass over its exact kept residual
            if not self.fp64:
                _lib.call("dlp_rho_cols", self.n, None, C1, ptr(self.R32), self.ld, _lib.F32, ptr(self.diag64),
                          ptr(buf), stream())
            Zf = self.Z64 if self.fp64 else self.Z32
            _lib.call("dlp_rho_cols", self.n, None, 1, ptr(Zf) + C * Zf.element_size(), self.ld,
                      _lib.F64 if self.fp64 else _lib.F32, ptr(self._ones64), ptr(buf) + 8 * C1, stream())
            host = buf.cpu().numpy()
            if not self.fp64:
                self._rho[:] = host[:C1]
            rc, h_est = self._rho, float(host[C1])
        h_est = max(h_est, 1.0)
        st.rho = float(rc[:C].max())
        if rc[C] > self.eps_h:
            st.hmax, st.cert = h_est, float("inf")
        else:
            st.hmax = h_est / (1.0 - float(rc[C]))
            st.cert = float(rc[:C].max()) * st.hmax

    def rho(self):
        """max over label columns and all rows of |R_uc|/diag_u of the kept residual (every row
        of it was last written by an FP64 refresh or, on the FP64 path, kept in FP64)."""
        R = self.R64 if self.fp64 else self.R32
        if not hasattr(self, "_rho_dev"):
            self._rho_dev = torch.zeros(self.C1, dtype=torch.float64, device=device())
        _lib.call("dlp_rho_cols", self.n, None, self.C1, ptr(R), self.ld, _lib.F64 if self.fp64 else _lib.F32,
                  ptr(self.diag64), ptr(self._rho_dev), stream())
        return float(self._rho_dev[: self.C].max().item())

    def error_report(self):
        """Errors of the held state, from device quantities (no exact solution needed):
        rel_l2_residual = ||R||_2 / ||rhs||_2 over the label columns of the solved rows (R the
        kept FP64-refreshed residual, rhs = R + A Z); cert_bound = max_c rho_c * hmax with hmax =
        max h / (1 - rho_h), which bounds max |F - F*| (solvers.py:295-301); argmax_certified_frac
        = solved rows whose prediction (solvers.py:166-167) cannot differ from that of F*: top two
        scores more than 2 * cert_bound apart (K > 2), |F - 0.5| > cert_bound (K = 2)."""
        C, dev = self.C, device()
        Zf = self.Z64 if self.fp64 else self.Z32
        AZ = torch.empty((self.n, C), dtype=torch.float64, device=dev)
        _lib.call("dlp_row_operator", self.n, None, C, ptr(self.indptr), ptr(self.rowlen), ptr(self.col), ptr(self.w),
                  _lib.F32, ptr(self.solved), ptr(Zf), self.ld, _lib.F64 if self.fp64 else _lib.F32, ptr(self.diag64),
                  -1.0, None, 0, ptr(AZ), C, _lib.F64, 0, self.nnz / max(self.n, 1), stream())
        if not hasattr(self, "_rho_dev"):
            self._rho_dev = torch.zeros(self.C1, dtype=torch.float64, device=dev)
        Rfull = self.R64 if self.fp64 else self.R32
        _lib.call("dlp_rho_cols", self.n, None, self.C1, ptr(Rfull), self.ld, _lib.F64 if self.fp64 else _lib.F32,
                  ptr(self.diag64), ptr(self._rho_dev), stream())
        rc = self._rho_dev.cpu().numpy()
        # the rest on the host, in FP64, over the solved rows
        sm = self.solved.cpu().numpy().astype(bool)
        R = Rfull[:, :C].cpu().numpy().astype(np.float64)[sm]
        rhs = R + AZ.cpu().numpy()[sm]
        rel = float(np.linalg.norm(R) / max(np.linalg.norm(rhs), 1e-300))
        F = Zf[:, : self.C1].cpu().numpy().astype(np.float64)[sm]
        h_est = float(F[:, C].max()) if F.shape[0] else 1.0
        if rc[C] > self.eps_h:
            hmax, cert = h_est, float("inf")
        else:
            hmax = h_est / (1 - float(rc[C]))
            cert = float(rc[:C].max()) * hmax
        if F.shape[0] == 0:
            frac = 1.0
        elif C == 1:
            frac = float((np.abs(F[:, 0] - 0.5) > cert).mean())
        else:
            top2 = -np.partition(-F[:, :C], 1, axis=1)[:, :2]
            frac = float(((top2[:, 0] - top2[:, 1]) > 2 * cert).mean())
        return dict(rho=float(rc[:C].max()), rho_h=float(rc[C]), hmax=hmax, cert_bound=cert,
                    rel_l2_residual=rel, argmax_certified_frac=frac, solved_rows=int(F.shape[0]))

    def time_operator(self, reps=20):
        """Average ms of the global solve's A*P launch (HIP events on the current stream)."""
        ms = ct.c_double(0.0)
        _lib.call("dlp_lp_time_operator", self.h, int(reps), ct.byref(ms), stream())
        return float(ms.value)

    def ground_flags(self):
        out = np.zeros(3, dtype=np.int32)
        _lib.call("dlp_lp_ground_flags", self.h, out.ctypes.data, stream())
        return dict(full_pass=int(out[0]), inert=int(out[1]), calls=int(out[2]))

    def stored_entries(self):
        return int(self.rowlen.sum().item())

    def edge_graph(self) -> Graph:
        """The graph as the device store holds it now (every live entry u < v with its weight),
        with this graph's labels: a checkpoint from which a new state can be built."""
        import dataclasses
        ip = self.indptr.cpu().numpy().astype(np.int64)
        rl = self.rowlen.cpu().numpy().astype(np.int64)
        col, w = self.col.cpu().numpy(), self.w.cpu().numpy()
        caps = np.diff(ip)
        rows = np.repeat(np.arange(self.n, dtype=np.int64), caps)
        live = (np.arange(ip[-1], dtype=np.int64) - np.repeat(ip[:-1], caps)) < np.repeat(rl, caps)
        r, c, x = rows[live], col[live].astype(np.int64), w[live]
        keep = c > r
        return dataclasses.replace(self.g, u=r[keep], v=c[keep], w=x[keep].astype(np.float32))

    def full_check(self):
        """FP64 residual of every row: (rho per column, violators) — the periodic full refresh
        of solvers.py:370-373, which also resyncs the kept residual."""
        prec = 1 if self.fp64 else 0
        _lib.call("dlp_lp_refresh", self.h, 2, prec, stream())
        s = self.status()
        return self._rho.copy(), s["viol"]

    def scores_with_h(self):
        """[F, h] of every vertex (FP64, host): rows outside the solved set as F = 0, h = 1."""
        out = torch.empty((self.n, self.C1), dtype=torch.float64, device=device())
        _lib.call("dlp_lp_scores", self.h, ptr(out), self.C1, stream())
        return out.cpu().numpy()

    def solved_mask(self):
        return self.solved.cpu().numpy().astype(bool)

    def components(self):
        """Canonical CC labels (int32, cc.py:8-23) of the current whole graph."""
        from . import cc as _cc
        from .backend import CSR
        W = CSR(self.w, self.col, self.indptr, (self.n, self.n), rowlen=self.rowlen)
        lab, _ = _cc.connected_components_csr(W)
        return lab.cpu().numpy()


# ---------------------------------------------------------------------------------------------
# The reference's vertex-arrival protocol (stream.py:84-346), producing systems for the
# reference-API solvers of solvers.py.  Host schedule; systems built on the device.

def new_components(m, lr, lc, w, tau, wmax, ground):
    """DynLP step 1 for the m new vertices (stream.py:67-81), from every stored entry (lr -> lc,
    w) of their rows (lc = -1 when the neighbour is not new): "tau" = connected components over
    the edges heavier than the graph's mean weight tau (the paper's supernodes); "fh" = adaptive
    components (cc.adaptive_components), whose scale is each vertex's mean dissimilarity
    1 - w/wmax over all its live entries; ground = each vertex's weight to labeled vertices."""
    from .cc import adaptive_components, connected_components, vertex_scales
    dev = device()
    lr, lc, w = np.asarray(lr), np.asarray(lc), np.asarray(w)    # host arrays of the stored entries
    s = vertex_scales(m, torch.as_tensor(lr, device=dev), torch.as_tensor(w, device=dev), wmax)
    e = (w > 0) & (lc >= 0)
    r, c, we = lr[e], lc[e], w[e]
    heavy = we > tau
    return {"tau": connected_components(m, r[heavy], c[heavy]),
            "fh": adaptive_components(m, r, c, we, s, s, wmax, np.asarray(ground, dtype=np.float32),
                                      scale_index=(r, c))}


def _arrival_order(g: Graph, seed: int) -> np.ndarray:
    """The graph's arrival order (an edge-list file's `order`), else the permutation numpy's
    RandomState(seed + 1) draws (graphs.from_npz's default)."""
    order = g.meta.get("order")
    if order is None:
        order = np.random.RandomState(seed + 1).permutation(g.n)
    return np.asarray(order, dtype=np.int64)


def _stream_labels(y, K, n, label_frac, rs) -> np.ndarray:
    """The stream's labeled set: the first max(K, label_frac*n) of one random permutation of
    the vertices with known ground truth (y >= 0)."""
    known = np.flatnonzero(y >= 0)
    lab = np.zeros(n, dtype=bool)
    lab[known[rs.permutation(known.shape[0])[: max(K, int(label_frac * n))]]] = True
    return lab


def _seat_classes(lab, y, K, order, n0, rs) -> None:
    """Every class gets a labeled vertex inside the initial snapshot order[:n0] (in place): a
    class without one labels its lowest-id vertex; a class whose labeled vertices all arrive
    later moves its lowest-id labeled vertex into the snapshot, trading places with an unlabeled
    snapshot vertex picked by one randint draw."""
    slot_of = np.empty(order.shape[0], dtype=np.int64)
    slot_of[order] = np.arange(order.shape[0])
    for c in range(K):
        members = np.flatnonzero(lab & (y == c))
        if members.shape[0] == 0:
            members = np.flatnonzero(y == c)[:1]
            lab[members] = True
        if members.shape[0] == 0 or int(slot_of[members].min()) < n0:
            continue
        open_slots = np.flatnonzero(~lab[order[:n0]])
        j = int(open_slots[int(rs.randint(0, open_slots.shape[0]))])
        v = int(members[0])
        u, sv = int(order[j]), int(slot_of[v])
        order[j], order[sv] = v, u
        slot_of[v], slot_of[u] = j, sv


class Stream:
    """A graph revealed in batches by the reference's protocol (stream.py:84-140): an initial
    snapshot of max(init_frac*n, 2K) vertices in arrival order, then n_batches batches of new
    vertices, each with deletions of del_frac x its size random active unlabeled vertices.  The
    labeled set is the stream's own (max(K, label_frac*n) vertices with known ground truth, every
    class present in the snapshot); all draws come from numpy's RandomState(seed + 7), in the
    reference's order.  batches() yields one System per batch (local ids over U, built on the
    device by solvers.build_system) with the reference's bookkeeping: t, is_new (rows not in U
    at the previous batch), seed (is_new, or adjacent to an inserted or deleted vertex),
    n_inserted, n_deleted, n_active, n_ungrounded, and comps, the new vertices' supernode
    components ("tau" and "fh", new_components) that the solvers' warm starts average over."""

    def __init__(self, g: Graph, *, label_frac=0.01, init_frac=0.1, n_batches=10, del_frac=0.1,
                 eta_rel=0.0, seed=0, dtype="float32"):
        self.g, self.dtype, self.del_frac = g, dtype, float(del_frac)
        n, K, y = g.n, g.K, np.asarray(g.y)
        self.rs = np.random.RandomState(seed + 7)
        self.labeled = _stream_labels(y, K, n, label_frac, self.rs)
        self.order = _arrival_order(g, seed).copy()
        self.n0 = max(int(init_frac * n), 2 * K)
        _seat_classes(self.labeled, y, K, self.order, self.n0, self.rs)
        self.y = y
        self.chunks = np.array_split(self.order[self.n0:], n_batches) if n_batches > 0 else []
        self.indptr, self.indices, self.data = g.csr()
        self.entry_row = np.repeat(np.arange(n, dtype=np.int64), np.diff(self.indptr))
        self.tau = float(self.data.mean()) if self.data.size else 0.0
        self.wmax = float(self.data.max()) if self.data.size else 1.0
        # eta_rel times the mean weighted degree (both directions of every edge)
        self.eta = float(eta_rel * 2.0 * float(np.sum(g.w, dtype=np.float64)) / n)
        self.schedule = self._plan()

    def _plan(self):
        """(inserted, deleted) vertices of every batch: batch 0 inserts the snapshot; batch t
        inserts chunk t-1 and deletes min(del_frac*|chunk|, |pool| - 1) vertices of the pool of
        active unlabeled ones (one permutation draw of the pool when it deletes any)."""
        n = self.g.n
        active = np.zeros(n, dtype=bool)
        none = np.zeros(0, dtype=np.int64)
        plan = []
        for t in range(1 + len(self.chunks)):
            ins = self.order[: self.n0] if t == 0 else self.chunks[t - 1]
            dele = none
            if t:
                pool = np.flatnonzero(active & ~self.labeled)
                nd = min(int(self.del_frac * ins.shape[0]), max(pool.shape[0] - 1, 0))
                if nd > 0:
                    dele = pool[self.rs.permutation(pool.shape[0])[:nd]]
            active[ins] = True
            active[dele] = False
            plan.append((np.asarray(ins, dtype=np.int64), np.asarray(dele, dtype=np.int64)))
        return plan

    def __len__(self):
        return len(self.schedule)

    def _adjacent(self, V) -> np.ndarray:
        """Vertices with a neighbour in V (in the whole graph, active or not)."""
        mark = np.zeros(self.g.n, dtype=bool)
        mark[V] = True
        adj = np.zeros(self.g.n, dtype=bool)
        adj[self.entry_row[mark[self.indices]]] = True
        return adj

    def _components(self, W, is_new, B):
        """The new rows' supernode components over the system's W (rows and columns in the
        system's ids; entries to rows that are not new carry lc = -1)."""
        new = np.flatnonzero(is_new)
        loc = np.full(is_new.shape[0], -1, dtype=np.int64)
        loc[new] = np.arange(new.shape[0])
        ip = W.indptr.cpu().numpy().astype(np.int64)
        rows = np.repeat(np.arange(is_new.shape[0]), np.diff(ip))
        on = is_new[rows]
        cols = W.indices.cpu().numpy().astype(np.int64)[on]
        ground = B.cpu().numpy()[new].sum(axis=1)    # each new vertex's weight to labeled vertices
        return new_components(int(new.shape[0]), loc[rows[on]], loc[cols], W.data.cpu().numpy()[on], self.tau,
                              self.wmax, ground)

    def batches(self):
        n = self.g.n
        active = np.zeros(n, dtype=bool)
        prevU = np.zeros(n, dtype=bool)
        for t, (ins, dele) in enumerate(self.schedule):
            active[ins] = True
            active[dele] = False
            sysm = self._build(t, active, prevU, ins, dele)
            prevU = np.zeros(n, dtype=bool)
            prevU[sysm.U.cpu().numpy()] = True
            yield sysm

    def _build(self, t, active, prevU, ins, dele):
        from .solvers import build_system
        dev = device()
        sysm = build_system(self.g, self.dtype, self.eta, active=active, labeled=self.labeled)
        U = sysm.U.cpu().numpy()
        is_new = ~prevU[U]
        seed = is_new | self._adjacent(np.concatenate([ins, dele]))[U]
        sysm.t = t
        sysm.is_new = torch.as_tensor(is_new, device=dev)
        sysm.seed = torch.as_tensor(seed, device=dev)
        sysm.comps = self._components(sysm.W, is_new, sysm.B)
        sysm.n_inserted, sysm.n_deleted = int(ins.shape[0]), int(dele.shape[0])
        sysm.n_active = int(active.sum())
        sysm.n_ungrounded = int((active & ~self.labeled).sum()) - int(U.shape[0])
        return sysm


class IncrementalStream(Stream):
    """The same stream as global-index systems (stream.py:180-346): every vertex keeps its row;
    rows outside U (inactive, labeled or ungrounded) are inert (diag 1, rhs 0, their entries of
    the graph's fixed sparsity weighted 0), so every solver runs unchanged; `solved` marks U and
    `changed` lists the rows of U whose equation changed in the batch (the vertices that entered
    U and the neighbours of those that entered or left it or were labeled and arrived or left).
    The reference keeps these arrays in place and regrounds with local ball checks backed by a
    full check; both give the exact grounded set, from which each batch is rebuilt here.  The
    class weights, s, diag and rhs are FP64 masters; a float32 stream's system returns them from
    astype("float64"), as the reference's cached FP64 twin does."""

    def __init__(self, g: Graph, *, ball_hops=3, **kw):
        super().__init__(g, **kw)
        self.ball_hops = ball_hops

    def batches(self):
        from .backend import CSR
        from .solvers import System, build_system
        g, n, K, dev = self.g, self.g.n, self.g.K, device()
        C = 1 if K == 2 else K
        rows, cols = self.entry_row, self.indices.astype(np.int64)
        data64 = self.data.astype(np.float64)
        prior64 = np.full(C, 0.5 if K == 2 else 1.0 / K)
        ids = torch.arange(n, dtype=torch.int64, device=dev)
        indptr = torch.as_tensor(self.indptr, device=dev)
        indices = torch.as_tensor(self.indices, device=dev)
        dt = torch.float64 if self.dtype in ("float64", torch.float64) else torch.float32
        active = np.zeros(n, dtype=bool)
        prevU = np.zeros(n, dtype=bool)
        for t, (ins, dele) in enumerate(self.schedule):
            active[ins] = True
            active[dele] = False
            Um = np.zeros(n, dtype=bool)
            Um[build_system(g, "float64", self.eta, active=active, labeled=self.labeled).U.cpu().numpy()] = True
            w64 = np.where(Um[rows] & Um[cols], data64, 0.0)
            labact = active & self.labeled
            B64 = np.zeros((n, K))
            sel = labact[cols]
            np.add.at(B64, (rows[sel], self.y[cols[sel]]), data64[sel])
            rowsum = np.bincount(rows, weights=w64, minlength=n)
            s64 = np.where(Um, B64.sum(axis=1) + self.eta, 1.0)
            diag64 = np.where(Um, s64 + rowsum, 1.0)
            rhs64 = np.where(Um[:, None], (B64[:, 1:2] if K == 2 else B64) + self.eta * prior64[None, :], 0.0)
            entering = Um & ~prevU
            moved = np.flatnonzero(entering | (prevU & ~Um))
            status = np.concatenate([moved, ins[self.labeled[ins]], dele[self.labeled[dele]]])
            changed = np.flatnonzero((self._adjacent(status) | entering) & Um)
            seed = (entering | self._adjacent(np.concatenate([ins, dele]))) & Um
            t64 = lambda a: torch.as_tensor(a, device=dev)  # noqa: E731
            W64 = CSR(t64(w64), indices, indptr, (n, n))
            # components over the system's working-precision weights, as the reference's Wd
            comps = self._components(W64 if dt == torch.float64 else W64.astype(dt), entering, t64(B64))
            sys64 = System(ids, W64, t64(s64), t64(diag64), t64(B64),
                           t64(np.ascontiguousarray(rhs64)), t64(prior64), K, C, self.eta,
                           is_new=t64(entering), seed=t64(seed), t=t, comps=comps, n_inserted=int(ins.shape[0]),
                           n_deleted=int(dele.shape[0]), n_active=int(active.sum()),
                           n_ungrounded=int((active & ~self.labeled & ~Um).sum()), solved=t64(Um),
                           changed=t64(changed))
            sysm = sys64 if dt == torch.float64 else sys64.astype("float32")
            if sysm is not sys64:
                sysm._conv[torch.float64] = sys64
            prevU = Um
            yield sysm

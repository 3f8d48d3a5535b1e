"""FP64 CPU port of the incremental edge-stream update (test infrastructure and the reported
CPU baseline; never part of the product path).

Per batch, as the device path does it, in NumPy/SciPy FP64 on the host:
  1. edge set update (deletions, then insertions as upserts) on the symmetric CSR;
  2. grounding: connected components of the unlabeled subgraph (stream.py:148-158);
  3. the changed rows' equations and residual R[Q] = rhs[Q] - diag[Q] Z[Q] + W[Q] Z
     (solvers.py:404-407);
  4. repair to |R_uc| <= eps_c diag_u: host push (solvers.py:100-125) when few rows violate,
     else column-wise PCG on the correction (solvers.py:66-97) followed by a true residual.
Single-threaded SciPy sparse kernels ("1 core effective").
"""
from __future__ import annotations

import time

import numpy as np
import scipy.sparse as sp

from . import lp as O


class CPUIncremental:
    def __init__(self, n, u, v, w, labeled, y, K, eps=1e-6, eps_h=0.05, push_max_rows=4096, Z=None):
        self.n, self.K = int(n), int(K)
        self.labeled = np.asarray(labeled, dtype=bool)
        self.y = np.asarray(y)
        self.eps, self.eps_h, self.push_max_rows = eps, eps_h, push_max_rows
        A = sp.coo_matrix((np.r_[w, w].astype(np.float64), (np.r_[u, v], np.r_[v, u])), shape=(n, n)).tocsr()
        A.sum_duplicates()
        self.A = A
        self._rebuild_all()
        C1 = self.sys.C + 1
        if Z is None:
            Z = np.zeros((n, C1))
            Z[:, :self.sys.C] = 1.0 / max(self.sys.C, 2)
            Z[:, self.sys.C] = 1.0
        self.Z = np.array(Z, dtype=np.float64)
        self.Z[~self.sys.solved, : self.sys.C] = 0
        self.Z[~self.sys.solved, self.sys.C] = 1
        self.R = O.residual(self.sys, self.Z)
        self.epsv = np.array([eps] * self.sys.C + [eps_h])

    def _rebuild_all(self):
        coo = sp.triu(self.A, 1).tocoo()
        self.sys = O.build_system(self.n, coo.row, coo.col, coo.data, self.labeled, self.y, self.K)

    def apply(self, ins_u, ins_v, ins_w, del_u, del_v):
        n = self.n
        a = np.r_[del_u, del_v].astype(np.int64)
        b = np.r_[del_v, del_u].astype(np.int64)
        old = np.asarray(self.A[a, b]).ravel()
        iu = np.r_[ins_u, ins_v].astype(np.int64)
        iv = np.r_[ins_v, ins_u].astype(np.int64)
        prev = np.asarray(self.A[iu, iv]).ravel()
        dw = np.r_[-old, np.r_[ins_w, ins_w].astype(np.float64) - prev]
        D = sp.coo_matrix((dw, (np.r_[a, iu], np.r_[b, iv])), shape=(n, n)).tocsr()
        self.A = (self.A + D).tocsr()
        self.A.eliminate_zeros()
        return np.unique(np.r_[a, iu])

    def solve_batch(self, ins_u, ins_v, ins_w, del_u, del_v):
        t0 = time.perf_counter()
        ends = self.apply(ins_u, ins_v, ins_w, del_u, del_v)
        old_solved = self.sys.solved
        self._rebuild_all()    # grounding + W_UU / diag / rhs of the new edge set
        sysm = self.sys
        toggled = np.flatnonzero(old_solved != sysm.solved)
        Q = np.unique(np.r_[ends, toggled, sysm.W[toggled].indices if toggled.size else np.zeros(0, int)])
        C = sysm.C
        self.Z[~sysm.solved, :C] = 0
        self.Z[~sysm.solved, C] = 1
        rhs = np.concatenate([sysm.rhs, sysm.diag[:, None]], axis=1)
        self.R[Q] = rhs[Q] - sysm.diag[Q][:, None] * self.Z[Q] + sysm.W[Q] @ self.Z
        self.R[~sysm.solved] = 0
        viol = (np.abs(self.R) > self.epsv[None, :] * sysm.diag[:, None]).any(axis=1)
        path = "push"
        if viol.sum() <= self.push_max_rows:
            self.Z, self.R, ok, rounds, _ = O.push(sysm.W, sysm.diag, self.Z, self.R, self.epsv)
        else:
            path = "pcg"
            D, _ = O.pcg(sysm.W, sysm.diag, np.zeros_like(self.R), self.R, self.epsv)
            self.Z += D
            self.R = rhs - O.apply_A(sysm.W, sysm.diag, self.Z)
            self.Z, self.R, ok, rounds, _ = O.push(sysm.W, sysm.diag, self.Z, self.R, self.epsv)
        rho = O.rho_cols(sysm.diag, self.R)[:C].max()
        return dict(ms=(time.perf_counter() - t0) * 1e3, rho=float(rho), path=path)


def cpu_case(g, b, eps=1e-6):
    """The reference's CPU-runnable case on the host: Jacobi sweeps X = (rhs + W X) / diag from
    the prior until max |dX| <= eps (so rho <= eps), h by CG to its own threshold 0.05, then the
    batch b repaired by host push (solvers.py:100-125) on the rebuilt system.  Timed pieces in
    ms, single-threaded SciPy."""
    sysm = O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, g.K)
    Wl, dl, rl = sysm.local()
    t = time.perf_counter()
    X = np.full_like(rl, 1.0 / g.K)
    sweeps = 0
    while True:
        Y = (rl + Wl @ X) / dl[:, None]
        moved = float(np.abs(Y - X).max())
        X, sweeps = Y, sweeps + 1
        if moved <= eps:
            break
    jacobi_ms = (time.perf_counter() - t) * 1e3
    hc = dl[:, None].copy()
    H, _ = O.pcg(Wl, dl, np.zeros_like(hc), hc, [0.05])
    Z = np.zeros((g.n, sysm.C + 1))
    Z[:, sysm.C] = 1.0
    Z[sysm.U, : sysm.C] = X
    Z[sysm.U, sysm.C] = H[:, 0]
    port = CPUIncremental(g.n, g.u, g.v, g.w, g.labeled, g.y, g.K, eps=eps, Z=Z, push_max_rows=g.n)
    out = port.solve_batch(b.ins_u, b.ins_v, b.ins_w, b.del_u, b.del_v)
    return {"jacobi_ms": round(jacobi_ms, 1), "jacobi_sweeps": sweeps, "batch_ms": round(out["ms"], 1),
            "batch_path": out["path"], "edge_updates_per_s": round(1000 / (out["ms"] / 1e3), 1),
            "rho": out["rho"]}

"""FP64 CPU oracle for the dynamic label-propagation hot path.

TEST INFRASTRUCTURE ONLY.  This module restates the reference algorithms in NumPy/SciPy so
the HIP path can be checked against them.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import it; the product package (dynlp_amd) never does.

Each function names the reference file:line it restates.  What pins each function (golden
vectors, closed forms, brute force) is listed in `PINS` at the bottom and exercised by
tests/test_oracle.py.

Conventions: CSR triplets (indptr, indices, data) with int32 indices; row-major n x C
float64 blocks; "global-index" systems keep all n vertices as rows, rows outside the solved
set U being inert (diag 1, rhs 0, no entries), as in stream.py:180-188.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph as csgraph
import scipy.sparse.linalg as spla


def csr(indptr, indices, data, n, m=None):
    return sp.csr_matrix((np.asarray(data, dtype=np.float64), np.asarray(indices), np.asarray(indptr)),
                         shape=(n, n if m is None else m))


# ---------------------------------------------------------------------------------------------
# Connected components (cc.py:8-23).  Closed form: hook-larger-onto-smaller plus full pointer
# jumping leaves every vertex labelled with its component's minimum id; `unique(...,
# return_inverse)` then numbers components by increasing minimum id.

def connected_components(n, rows, cols):
    """(int32 labels, n_comp): label = rank of the component's minimum vertex id."""
    if n == 0:
        return np.zeros(0, dtype=np.int32), 0
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    G = sp.coo_matrix((np.ones(rows.shape[0]), (rows, cols)), shape=(n, n)).tocsr()
    nc, lab = csgraph.connected_components(G, directed=True, connection="weak")
    mins = np.full(nc, n, dtype=np.int64)
    np.minimum.at(mins, lab, np.arange(n))
    rank = np.empty(nc, dtype=np.int64)
    rank[np.argsort(mins)] = np.arange(nc)
    return rank[lab].astype(np.int32), int(nc)


def cc_hook_jump(n, rows, cols):
    """Literal restatement of cc.py:12-23 (scatter-min hook + full jump + unique); slow,
    used to pin `connected_components` on small inputs."""
    lab = np.arange(n, dtype=np.int64)
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    while True:
        lr, lc = lab[rows], lab[cols]
        diff = lr != lc
        if not diff.any():
            break
        lr, lc = lr[diff], lc[diff]
        np.minimum.at(lab, np.maximum(lr, lc), np.minimum(lr, lc))
        while True:
            nxt = lab[lab]
            if (nxt == lab).all():
                break
            lab = nxt
    uniq, inv = np.unique(lab, return_inverse=True)
    return inv.astype(np.int32).ravel(), int(uniq.shape[0])


# ---------------------------------------------------------------------------------------------
# The batch system (stream.py:1-6, 142-177; global-index form stream.py:180-188, 324-331).

@dataclass
class System:
    n: int
    K: int
    C: int
    solved: np.ndarray     # bool[n]: rows of U (unlabeled and grounded)
    W: sp.csr_matrix       # n x n, entries only between solved vertices (W_UU in global ids)
    s: np.ndarray          # grounding (inert rows: 1)
    diag: np.ndarray       # s + rowsum(W_UU) (inert rows: 1)
    rhs: np.ndarray        # n x C (inert rows: 0)
    B: np.ndarray          # n x K weight to labeled neighbours per class
    eta: float
    prior: np.ndarray

    @property
    def U(self):
        return np.flatnonzero(self.solved)

    def local(self):
        """(W_UU, diag_U, rhs_U) on the solved rows only."""
        U = self.U
        return self.W[U][:, U].tocsr(), self.diag[U], self.rhs[U]


def build_system(n, u, v, w, labeled, y, K, eta=0.0, active=None):
    """Stream._build (stream.py:142-177) over the current edge set, in global-index form:
    U = active & unlabeled & grounded (component of the U-U graph touches a labeled vertex,
    stream.py:149-158); W = A restricted to U x U; s = B.sum + eta; diag = s + W.1;
    rhs = B (K > 2) or B[:, 1:2] (K == 2), plus eta * prior (stream.py:164-167)."""
    n = int(n)
    active = np.ones(n, dtype=bool) if active is None else np.asarray(active, dtype=bool)
    labeled = np.asarray(labeled, dtype=bool)
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    w = np.asarray(w, dtype=np.float64)
    A = sp.coo_matrix((np.r_[w, w], (np.r_[u, v], np.r_[v, u])), shape=(n, n)).tocsr()
    A.sum_duplicates()
    onehot = np.zeros((n, K))
    li = np.flatnonzero(active & labeled)
    onehot[li, np.asarray(y)[li]] = 1.0
    Bfull = A @ onehot
    Um = active & ~labeled
    coo = A.tocoo()
    keep = Um[coo.row] & Um[coo.col]
    lab, nc = connected_components(n, coo.row[keep], coo.col[keep])
    gw = np.zeros(nc)
    np.add.at(gw, lab[Um], Bfull[Um].sum(axis=1))
    grounded = gw[lab] > 0
    solved = Um & grounded
    keep = solved[coo.row] & solved[coo.col]
    W = sp.coo_matrix((coo.data[keep], (coo.row[keep], coo.col[keep])), shape=(n, n)).tocsr()
    W.sort_indices()
    C = 1 if K == 2 else K
    prior = np.full(C, 0.5 if K == 2 else 1.0 / K)
    Bm = np.where(solved[:, None], Bfull, 0.0)
    s = np.where(solved, Bm.sum(axis=1) + eta, 1.0)
    diag = np.where(solved, s + np.asarray(W.sum(axis=1)).ravel(), 1.0)
    rhs = np.where(solved[:, None], (Bm[:, 1:2] if K == 2 else Bm) + eta * prior[None, :], 0.0)
    return System(n, K, C, solved, W, s, diag, rhs, Bm, eta, prior)


# ---------------------------------------------------------------------------------------------
# Operators (solvers.py:36-59, kernels.py:140-166).

def spmm_axpy(W, X, a=None, b=1.0):
    """Y = a * X + b * (W @ X)  (kernels.py:140-166; a None -> 0)."""
    X = np.asarray(X, dtype=np.float64)
    Y = b * (W @ X)
    if a is not None:
        Y = Y + np.asarray(a, dtype=np.float64)[:, None] * X
    return Y


def apply_A(sys_W, diag, X):
    """A X = diag * X - W X  (solvers.py:46-49)."""
    return spmm_axpy(sys_W, X, a=diag, b=-1.0)


def residual(sys: System, Z):
    """rhs - A Z, with the h column appended when Z has C+1 columns (solvers.py:312-313)."""
    rhs = sys.rhs if Z.shape[1] == sys.C else np.concatenate([sys.rhs, sys.diag[:, None]], axis=1)
    return rhs - apply_A(sys.W, sys.diag, Z)


def rho_cols(diag, R):
    """max_u |R_uc| / diag_u per column (solvers.py:52-54)."""
    if R.shape[0] == 0:
        return np.zeros(R.shape[1])
    return (np.abs(R) / diag[:, None]).max(axis=0)


def violation(diag, R, eps):
    """max |R_uc| / (eps_c * diag_u) (solvers.py:57-59)."""
    if R.shape[0] == 0:
        return 0.0
    return float((np.abs(R) / (np.asarray(eps)[None, :] * diag[:, None])).max())


def predict(F, K):
    """solvers.py:166-167."""
    return (F[:, 0] > 0.5).astype(np.int32) if K == 2 else F.argmax(axis=1).astype(np.int32)


def bound(rc, h_est, C, eps_h):
    """(hmax, certified bound) from column residual ratios (solvers.py:295-301)."""
    if rc[C] > eps_h:
        return h_est, float("inf")
    hmax = h_est / (1 - float(rc[C]))
    return hmax, float(rc[:C].max()) * hmax


# ---------------------------------------------------------------------------------------------
# Exact solution (the reference's tests compare against dense solves, tests/test_dynlp.py:17-19).

def exact_solve(W, diag, rhs, tol=1e-14, dense_max=16000):
    """F* with (diag - W) F* = rhs in FP64: dense solve up to dense_max rows (LAPACK, seconds at
    16,000; sparse LU fills in badly on these expander-like graphs: 168 s at 12,000 rows of an
    SBM), sparse LU beyond."""
    n = W.shape[0]
    rhs = np.asarray(rhs, dtype=np.float64)
    if n == 0:
        return np.zeros_like(rhs)
    A = (sp.diags(diag) - W).tocsc()
    if n <= dense_max:
        return np.linalg.solve(A.toarray(), rhs)
    lu = spla.splu(A)
    X = lu.solve(rhs)
    # one step of iterative refinement in FP64
    X += lu.solve(rhs - A @ X)
    return X


def solve_system(sys: System, with_h=True):
    """Exact [F*, h] on the solved rows (inert rows: F = 0, h = 1), global-index layout."""
    Wl, dl, rl = sys.local()
    U = sys.U
    cols = np.concatenate([rl, dl[:, None]], axis=1) if with_h else rl
    Xl = exact_solve(Wl, dl, cols)
    C = sys.C
    Z = np.zeros((sys.n, C + 1 if with_h else C))
    if with_h:
        Z[:, C] = 1.0
    Z[U] = Xl
    return Z


# ---------------------------------------------------------------------------------------------
# Frontier primitives (kernels.py:31-138 and their fallbacks kernels.py:464-492).

def frontier_jacobi(W, frontier, X, rhs, diag, delta):
    """Y = (rhs[f] + W[f] X) / diag[f]; changed = any_c |Y - X[f]| > delta (kernels.py:31-67)."""
    f = np.asarray(frontier, dtype=np.int64)
    Y = (rhs[f] + W[f] @ X) / diag[f][:, None]
    return Y, (np.abs(Y - X[f]) > delta).any(axis=1)


def push_scatter(W, frontier, D, R, mark):
    """R[col] += w * D[g] over the rows of the frontier; mark[col] = 1 (kernels.py:105-126)."""
    f = np.asarray(frontier, dtype=np.int64)
    sub = W[f].tocoo()
    np.add.at(R, sub.col, sub.data[:, None] * D[sub.row])
    mark[sub.col] = 1


def mark_neighbors(W, rows, mark):
    """mark[indices[j]] = 1 for j in the given rows (kernels.py:128-138)."""
    mark[W[np.asarray(rows, dtype=np.int64)].indices] = 1


def edge_visits(W, rows):
    """Edge-visit counter of FrontierOps._rows (kernels.py:459-462)."""
    ip = W.indptr
    rows = np.asarray(rows, dtype=np.int64)
    return int((ip[rows + 1] - ip[rows]).sum())


# ---------------------------------------------------------------------------------------------
# Residual push (solvers.py:100-125; fused form kernels.py:269-351).

def push(W, diag, Z, R, eps, max_iter=10**7):
    """Host-driven push (solvers.py:100-125): the violating set takes D = R/diag, Z += D,
    R = 0, scatters W D to its neighbours, and the touched set is filtered by the violation
    test.  No budget / stall exits here.  Returns (Z, R, converged, rounds, edge visits)."""
    eps = np.asarray(eps, dtype=np.float64)
    n = W.shape[0]
    viol = lambda idx: (np.abs(R[idx]) > eps[None, :] * diag[idx][:, None]).any(axis=1)  # noqa: E731
    allr = np.arange(n)
    cand = allr[viol(allr)]
    it = edges = 0
    while cand.shape[0] and it < max_iter:
        D = R[cand] / diag[cand][:, None]
        Z[cand] += D
        R[cand] = 0
        mark = np.zeros(n, dtype=np.uint8)
        mark[cand] = 1
        push_scatter(W, cand, D, R, mark)
        edges += edge_visits(W, cand)
        touched = np.flatnonzero(mark)
        cand = touched[viol(touched)]
        it += 1
    return Z, R, cand.shape[0] == 0, it, edges


def sync_push(W, diag, Z, R, eps, max_rounds=10**6):
    """Round-synchronous push on the violating set, the form the service runs on the device:
    round t takes F_t = {u : any_c |R_uc| > eps_c diag_u}, sets D_u = R_u / diag_u,
    Z_u += D_u, R_u = 0, then every v adds sum_{u in F_t} w_vu D_u.  Same fixed point and
    exact-residual invariant as solvers.py:100-125 (it is that loop with the scatter written
    as a gather over the symmetric W).  Returns (Z, R, converged, rounds, pushes)."""
    eps = np.asarray(eps, dtype=np.float64)
    rounds = pushes = 0
    while rounds < max_rounds:
        F = np.flatnonzero((np.abs(R) > eps[None, :] * diag[:, None]).any(axis=1))
        if F.shape[0] == 0:
            return Z, R, True, rounds, pushes
        D = R[F] / diag[F][:, None]
        Z[F] += D
        R[F] = 0
        Dfull = sp.csr_matrix((np.ones(F.shape[0]), (F, np.arange(F.shape[0]))), shape=(W.shape[0], F.shape[0]))
        R += W @ (Dfull @ D)
        rounds += 1
        pushes += F.shape[0]
    return Z, R, False, rounds, pushes


# ---------------------------------------------------------------------------------------------
# Column-wise PCG (solvers.py:62-97).

def pcg(W, diag, Z, rhs, eps, M=None, max_iter=100_000, patience=400):
    """One CG per column until |r_c| <= eps_c * diag; true-residual re-check; stall exit."""
    eps = np.asarray(eps, dtype=np.float64)
    M = (lambda R: R / diag[:, None]) if M is None else M
    R = rhs - apply_A(W, diag, Z)
    best = violation(diag, R, eps)
    iters = 0
    if best <= 1:
        return Z, iters
    P = Y = M(R)
    rz, stall = (R * Y).sum(axis=0), 0
    for _ in range(max_iter):
        AP = apply_A(W, diag, P)
        pAp = (P * AP).sum(axis=0)
        alpha = np.where(pAp > 0, rz / np.where(pAp > 0, pAp, 1), 0)
        Z += alpha[None, :] * P
        R -= alpha[None, :] * AP
        iters += 1
        v = violation(diag, R, eps)
        if v <= 1:
            R = rhs - apply_A(W, diag, Z)
            v = violation(diag, R, eps)
            if v <= 1:
                break
        best, stall = (v, 0) if v < 0.9 * best else (best, stall + 1)
        if stall >= patience:
            break
        Y = M(R)
        rz_new = (R * Y).sum(axis=0)
        P = Y + np.where(rz > 0, rz_new / np.where(rz > 0, rz, 1), 0)[None, :] * P
        rz = rz_new
    return Z, iters


# ---------------------------------------------------------------------------------------------
# Aggregation AMG (amg.py:1-134).

def _handshake(n, rows, cols, w, rounds):
    """Heavy-edge handshake matching, unmatched join the best matched neighbour, renumber
    by unique (amg.py:32-64).  Ties: heaviest weight, then smallest neighbour id."""
    agg = np.full(n, -1, dtype=np.int64)
    matched = np.zeros(n, dtype=bool)
    ar = np.arange(n, dtype=np.int64)

    def best_choice(er, ec, ew):
        best = np.full(n, -np.inf)
        np.maximum.at(best, er, ew)
        tie = ew >= best[er]
        choice = np.full(n, n, dtype=np.int64)
        np.minimum.at(choice, er[tie], ec[tie].astype(np.int64))
        return choice

    for _ in range(rounds):
        ok = ~matched[rows] & ~matched[cols]
        if not ok.any():
            break
        choice = best_choice(rows[ok], cols[ok], w[ok])
        has = choice < n
        partner = np.where(has, choice, 0)
        mutual = has & (choice[partner] == ar)
        matched |= mutual
        agg = np.where(mutual, np.minimum(ar, partner), agg)
    ok = ~matched[rows] & matched[cols]
    if ok.any():
        choice = best_choice(rows[ok], cols[ok], w[ok])
        has = choice < n
        agg[has] = agg[choice[has]]
    left = agg < 0
    agg[left] = ar[left]
    uniq, inv = np.unique(agg, return_inverse=True)
    return inv.ravel().astype(np.int64), int(uniq.shape[0])


def _coarsen(W, s, agg, nc):
    """W_c = off-diagonal part of P^T W P (duplicates summed), s_c = P^T s (amg.py:67-74)."""
    coo = W.tocoo()
    rc, cc = agg[coo.row], agg[coo.col]
    off = rc != cc
    Wc = sp.coo_matrix((coo.data[off], (rc[off], cc[off])), shape=(nc, nc)).tocsr()
    Wc.sum_duplicates()
    Wc.sort_indices()
    sc = np.zeros(nc)
    np.add.at(sc, agg, s)
    return Wc, sc


class AMG:
    """V-cycle preconditioner (amg.py:77-134): weighted-Jacobi smoothing, exact coarse solve."""

    def __init__(self, W, s, max_coarse=1500, passes=2, rounds=3, omega=0.7, nu=1,
                 max_levels=25, min_ratio=0.85, coarse_sweeps=20):
        self.omega, self.nu, self.coarse_sweeps = omega, nu, coarse_sweeps
        self.levels = []
        W = W.tocsr()
        n = W.shape[0]
        cur = dict(W=W, s=s, diag=s + np.asarray(W.sum(axis=1)).ravel())
        while n > max_coarse and len(self.levels) < max_levels:
            agg = np.arange(n, dtype=np.int64)
            nc, Wp = n, cur["W"]
            for _ in range(passes):
                cp = Wp.tocoo()
                a2, nc2 = _handshake(Wp.shape[0], cp.row.astype(np.int64), cp.col.astype(np.int64), cp.data, rounds)
                agg = a2[agg]
                Wp, _ = _coarsen(Wp, np.zeros(Wp.shape[0]), a2, nc2)
                nc = nc2
            if nc > min_ratio * n:
                break
            Wc, sc = _coarsen(cur["W"], cur["s"], agg, nc)
            cur["agg"] = agg
            self.levels.append(cur)
            cur = dict(W=Wc, s=sc, diag=sc + np.asarray(Wc.sum(axis=1)).ravel())
            n = nc
        self.coarse = cur
        self.inv = None
        if n <= 4 * max_coarse:
            Ad = -cur["W"].toarray()
            Ad[np.arange(n), np.arange(n)] += cur["diag"]
            self.inv = np.linalg.inv(Ad)
        self.sizes = [L["W"].shape[0] for L in self.levels] + [n]

    @staticmethod
    def _apply(L, X):
        return L["diag"][:, None] * X - L["W"] @ X

    def _smooth(self, L, X, R, sweeps):
        for _ in range(sweeps):
            X = X + self.omega * (R - self._apply(L, X)) / L["diag"][:, None]
        return X

    def _cycle(self, lvl, R):
        if lvl == len(self.levels):
            L = self.coarse
            if self.inv is not None:
                return self.inv @ R
            X = self.omega * R / L["diag"][:, None]
            return self._smooth(L, X, R, self.coarse_sweeps - 1)
        L = self.levels[lvl]
        X = self.omega * R / L["diag"][:, None]
        X = self._smooth(L, X, R, self.nu - 1)
        Rf = R - self._apply(L, X)
        nc = self.levels[lvl + 1]["W"].shape[0] if lvl + 1 < len(self.levels) else self.coarse["W"].shape[0]
        Rc = np.zeros((nc, R.shape[1]))
        np.add.at(Rc, L["agg"], Rf)
        X = X + self._cycle(lvl + 1, Rc)[L["agg"]]
        return self._smooth(L, X, R, self.nu)

    def __call__(self, R):
        return self._cycle(0, np.asarray(R, dtype=np.float64))


# ---------------------------------------------------------------------------------------------
# Block CG with k x k Gram matrices (the block form of solvers.py:66-97): one search space
# shared by all columns, alpha = (P^T A P)^{-1} (R^T Z), beta = (R^T Z)^{-1} (R_new^T Z_new).

def block_pcg(W, diag, X, rhs, eps, M=None, max_iter=10_000):
    eps = np.asarray(eps, dtype=np.float64)
    M = (lambda R: R / diag[:, None]) if M is None else M
    R = rhs - apply_A(W, diag, X)
    if violation(diag, R, eps) <= 1:
        return X, 0
    Zp = M(R)
    P = Zp.copy()
    G = R.T @ Zp
    it = 0
    for it in range(1, max_iter + 1):
        AP = apply_A(W, diag, P)
        PAP = P.T @ AP
        alpha = np.linalg.lstsq(PAP, G, rcond=1e-12)[0]
        X += P @ alpha
        R -= AP @ alpha
        if violation(diag, R, eps) <= 1:
            R = rhs - apply_A(W, diag, X)
            if violation(diag, R, eps) <= 1:
                break
        Zp = M(R)
        Gn = R.T @ Zp
        beta = np.linalg.lstsq(G, Gn, rcond=1e-12)[0]
        P = Zp + P @ beta
        G = Gn
    return X, it


# ---------------------------------------------------------------------------------------------
# Supernodes of new vertices (cc.py:26-83, stream.py:67-81), NumPy float32 as the reference
# keeps Int, size, d and the scales.

def fh_components(m, rows, cols, d, k, ground=None):
    """Boruvka rounds of Felzenszwalb-Huttenlocher merging (cc.py:38-74): each component's
    lightest outgoing edge (ties: smallest component across, cc.py:26-35) merges when
    d <= min(Int_c + k/size_c, Int_t + k/size_t); then, with ground, components whose ground
    sums to zero join across their lightest outgoing edge until none can."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    d = np.asarray(d, dtype=np.float32)
    gv = None if ground is None else np.asarray(ground, dtype=np.float32)
    kf = np.float32(k)
    lab = np.arange(m, dtype=np.int64)
    Int = np.zeros(m, dtype=np.float32)
    size = np.ones(m, dtype=np.float32)
    nc, fh = m, True
    while rows.shape[0]:
        cu, cv = lab[rows], lab[cols]
        ext = cu != cv
        if not fh:
            g = np.zeros(nc, dtype=np.float32)
            np.add.at(g, lab, gv)
            ext &= g[cu] == 0
        if not ext.any():
            break
        eu, ev, de = cu[ext], cv[ext], d[ext]
        best = np.full(nc, np.inf, dtype=np.float32)
        np.minimum.at(best, eu, de)
        across = np.full(nc, nc, dtype=np.int64)
        sel = de <= best[eu]
        np.minimum.at(across, eu[sel], ev[sel])
        c = np.flatnonzero(across < nc)
        t, dc = across[c], best[c]
        if fh:
            ok = dc <= np.minimum(Int[c] + kf / size[c], Int[t] + kf / size[t])
            if not ok.any():
                if gv is None:
                    break
                fh = False
                continue
            c, t, dc = c[ok], t[ok], dc[ok]
        root, nc = connected_components(nc, c, t)
        root = root.astype(np.int64)
        Int2 = np.zeros(nc, dtype=np.float32)
        size2 = np.zeros(nc, dtype=np.float32)
        np.maximum.at(Int2, root, Int)
        np.maximum.at(Int2, root[c], dc)
        np.add.at(size2, root, size)
        lab, Int, size = root[lab], Int2, size2
    return lab.astype(np.int32), int(nc)


def vertex_scales(m, lr, w, wmax):
    """Mean of 1 - w/wmax over each vertex's live entries, plus 1e-6, in float32
    (stream.py:72-74)."""
    w = np.asarray(w)
    on = w > 0
    s = np.zeros(m, dtype=np.float32)
    cnt = np.zeros(m, dtype=np.float32)
    np.add.at(s, np.asarray(lr)[on], (1 - w[on] / w.dtype.type(wmax)).astype(np.float32))
    np.add.at(cnt, np.asarray(lr)[on], np.float32(1))
    return s / np.maximum(cnt, np.float32(1)) + np.float32(1e-6)


def adaptive_dissimilarity(w, s_r, s_c, wmax):
    """(d float32, k) of cc.py:77-83: d = (1 - w/wmax) / sqrt(s_r s_c) in w's precision (the
    scale product and its root in float32), k = the mean of d before rounding."""
    w = np.asarray(w)
    ss = np.sqrt(np.asarray(s_r, dtype=np.float32) * np.asarray(s_c, dtype=np.float32))
    x = (1 - w / w.dtype.type(wmax)) / ss.astype(w.dtype)
    return x.astype(np.float32), (float(x.astype(np.float64).sum()) / x.shape[0] if x.shape[0] else 0.0)


# ---------------------------------------------------------------------------------------------
# Edge-batch apply: the incremental result must equal the system rebuilt from the updated
# edge set (tests/test_incremental.py:18-41 pattern).

def apply_edge_batch(edges: dict, ins, dels):
    """Edge set {(u, v): w} with u < v; deletions first, then insertions (upsert)."""
    for a, b in dels:
        edges.pop((min(a, b), max(a, b)), None)
    for a, b, w in ins:
        if a != b:
            edges[(min(a, b), max(a, b))] = float(w)
    return edges


PINS = {
    "connected_components": "union-find brute force; literal hook/jump restatement (cc_hook_jump)",
    "cc_hook_jump": "union-find brute force; closed form (rank of component minimum)",
    "build_system": "dense definition (test_system_matches_dense_definition pattern); grounding by brute force",
    "spmm_axpy": "dense matrix product",
    "frontier_jacobi": "dense row formula on a hand-built graph; Jacobi fixed point == exact solution",
    "push_scatter": "dense W^T D",
    "mark_neighbors": "brute-force neighbour sets",
    "push": "exact-residual invariant R == rhs - A Z; convergence to eps; fixed point == exact solve",
    "sync_push": "same invariants; equals push when one round is run",
    "pcg": "exact solve; sum_c F_c = 1 identity (tests/test_dynlp.py:148-156)",
    "block_pcg": "exact solve; single column == scalar CG iterates",
    "AMG": "M(I) equals the dense two-level V-cycle X1 = w D^-1 R, X2 = X1 + P (P^T A P)^-1 P^T (R - A X1), "
           "M R = X2 + w D^-1 (R - A X2) on hand and random 7-8 vertex graphs; M(I) SPD "
           "(tests/test_dynlp.py:159-166); no-coarsening case M = A^-1",
    "fh_components": "hand cases (a weak bridge kept cut, a chain merged, the smallest-component tie, the "
                     "grounding pass) and the reference's two closed-form component cases "
                     "(tests/test_dynlp.py:56-71)",
    "vertex_scales/adaptive_dissimilarity": "hand values on a 3-vertex example",
    "_handshake": "hand cases for the smallest-id tie, the join of unmatched vertices to the heaviest "
                  "matched neighbour (and its tie), singletons, the round limit; a loop-level brute force "
                  "of the matching rules on random graphs, composed over two passes",
    "exact_solve": "numpy.linalg.solve dense; residual ~ 1e-14",
    "rho_cols/bound": "certificate |F - F*| <= rho h for perturbed iterates (tests/test_dynlp.py:136-145)",
}

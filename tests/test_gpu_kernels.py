"""HIP kernels vs the FP64 oracle on seeded inputs (patterns of the reference's
tests/test_dynlp.py:195-222, 246-254 and tests/test_incremental.py:66-81)."""
import numpy as np
import pytest

from dynlp_amd import graphs
from oracle import lp as O

pytestmark = pytest.mark.gpu


def _setup(n=3000, deg=12, seed=8, dtype="float32"):
    import torch
    from dynlp_amd import backend as B
    g = graphs.sbm(n, 2, deg, 0.85, 2, 0.02, seed)
    ip, ix, d = g.csr()
    dt = torch.float32 if dtype == "float32" else torch.float64
    W = B.csr_from_host(ip, ix, d, n, dt)
    Ws = O.csr(ip, ix, d, n)
    return g, W, Ws, dt


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("C", [1, 2, 8, 17, 40, 64, 65, 130])
def test_spmm_axpy_matches_oracle(dtype, C):
    import torch
    from dynlp_amd.kernels import spmm_axpy
    g, W, Ws, dt = _setup(dtype=dtype)
    rs = np.random.default_rng(1)
    X = rs.random((g.n, C)).astype(dtype)
    a = rs.random(g.n).astype(dtype)
    Xd = torch.as_tensor(X, device="cuda")
    ad = torch.as_tensor(a, device="cuda")
    tol = 1e-5 if dtype == "float32" else 1e-12
    Y = spmm_axpy(W, Xd).cpu().numpy()
    np.testing.assert_allclose(Y, O.spmm_axpy(Ws, X), rtol=tol, atol=tol)
    Y2 = spmm_axpy(W, Xd, a=ad, b=-1.0).cpu().numpy()
    np.testing.assert_allclose(Y2, O.spmm_axpy(Ws, X, a, -1.0), rtol=tol, atol=tol)


@pytest.mark.parametrize("group", [1, 2, 8, 32, "auto"])
def test_spmm_lane_layouts_agree(group):
    import torch
    from dynlp_amd.kernels import spmm_axpy
    g, W, Ws, dt = _setup(dtype="float64")
    X = np.random.default_rng(2).random((g.n, 3))
    Y = spmm_axpy(W, torch.as_tensor(X, device="cuda"), group=group).cpu().numpy()
    np.testing.assert_allclose(Y, O.spmm_axpy(Ws, X), rtol=1e-12)


def test_row_operator_masked_subset_mixed_precision():
    import torch
    from dynlp_amd import _lib
    from dynlp_amd.backend import ptr, stream
    g, W, Ws, _ = _setup(dtype="float32")
    n, C = g.n, 5
    rs = np.random.default_rng(3)
    mask = (rs.random(n) < 0.8).astype(np.uint8)
    rows = np.sort(rs.choice(n, 700, replace=False)).astype(np.int32)
    X = rs.random((n, C))                        # FP64 block, FP32 weights
    a = rs.random(n)
    add = rs.random((n, C))
    out = torch.zeros((700, C), dtype=torch.float64, device="cuda")
    Xd, ad, addd = (torch.as_tensor(v, device="cuda") for v in (X, a, add))
    md, rd = torch.as_tensor(mask, device="cuda"), torch.as_tensor(rows, device="cuda")
    _lib.call("dlp_row_operator", 700, ptr(rd), C, ptr(W.indptr), None, ptr(W.indices), ptr(W.data), 0, ptr(md),
              ptr(Xd), C, 1, ptr(ad), -2.0, ptr(addd), C, ptr(out), C, 1, 1, 0.0, stream())
    Wm = Ws.multiply(mask[:, None]).multiply(mask[None, :]).tocsr()
    ref = add[rows] + a[rows, None] * X[rows] - 2.0 * (Wm @ X)[rows]
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-12, atol=1e-12)


TUNE_DEFAULTS = {0: 0, 1: 1, 2: 1, 3: 1, 4: 0, 5: 1, 6: 4}


@pytest.fixture
def tune():
    from dynlp_amd import _lib

    def set_(**kv):
        keys = {"unroll": 0, "vec4": 1, "pack": 2, "xcd": 3, "dot_grid": 4, "fuse_dot": 5, "pipe": 6}
        for k, v in kv.items():
            _lib.call("dlp_tune_set", keys[k], int(v))
    yield set_
    for k, v in TUNE_DEFAULTS.items():
        _lib.call("dlp_tune_set", k, v)


@pytest.mark.parametrize("pack", [0, 1])
@pytest.mark.parametrize("xcd", [0, 1])
@pytest.mark.parametrize("unroll", [4, 8])
@pytest.mark.parametrize("C", [4, 17, 65, 130])
def test_fp32_row_layouts_match_oracle(tune, pack, xcd, unroll, C):
    """16-byte FP32 operator and the fused operator + column dot, every lane layout and row
    order, on a 16-byte-aligned slice (ld = C rounded up to 4) with a solved mask."""
    import torch
    from dynlp_amd import _lib
    from dynlp_amd.backend import ptr, stream
    tune(pack=pack, xcd=xcd, unroll=unroll)
    g, W, Ws, _ = _setup(n=4001, dtype="float32")
    n, ld = g.n, (C + 3) // 4 * 4
    rs = np.random.default_rng(C)
    mask = (rs.random(n) < 0.9).astype(np.uint8)
    X = np.zeros((n, ld), np.float32)
    X[:, :C] = rs.standard_normal((n, C))
    a = (rs.random(n) + 1).astype(np.float32)
    Xd, ad, md = (torch.as_tensor(v, device="cuda") for v in (X, a, mask))
    Wm = Ws.multiply(mask[:, None]).multiply(mask[None, :]).tocsr()
    Xc = X[:, :C].astype(np.float64)
    ref = a[:, None].astype(np.float64) * Xc - (Wm @ Xc)
    ref[mask == 0] = a[mask == 0, None] * Xc[mask == 0]
    out = torch.zeros((n, ld), dtype=torch.float32, device="cuda")
    _lib.call("dlp_row_operator", n, None, C, ptr(W.indptr), None, ptr(W.indices), ptr(W.data), 0, ptr(md),
              ptr(Xd), ld, 0, ptr(ad), -1.0, None, 0, ptr(out), ld, 0, 0, 0.0, stream())
    np.testing.assert_allclose(out.cpu().numpy()[:, :C], ref, rtol=2e-5, atol=2e-5)
    out.zero_()
    dot = torch.full((C,), 0.5, dtype=torch.float64, device="cuda")   # accumulated into
    _lib.call("dlp_row_operator_dot", n, C, ptr(W.indptr), None, ptr(W.indices), ptr(W.data), ptr(md), ptr(Xd), ld,
              ptr(ad), -1.0, ptr(out), ld, ptr(dot), stream())
    o = out.cpu().numpy()
    np.testing.assert_allclose(o[:, :C], ref, rtol=2e-5, atol=2e-5)
    assert not o[:, C:].any()                      # padding columns untouched
    np.testing.assert_allclose(dot.cpu().numpy(), 0.5 + (Xc * o[:, :C].astype(np.float64)).sum(axis=0),
                               rtol=1e-9, atol=1e-9)


def test_row_operator_dot_rejects_unaligned_rows():
    import torch
    from dynlp_amd import _lib
    from dynlp_amd.backend import ptr, stream
    g, W, Ws, _ = _setup(n=200, dtype="float32")
    X = torch.zeros((200, 6), dtype=torch.float32, device="cuda")
    dot = torch.zeros(5, dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.DLPError):
        _lib.call("dlp_row_operator_dot", 200, 5, ptr(W.indptr), None, ptr(W.indices), ptr(W.data), None, ptr(X), 6,
                  None, -1.0, ptr(X), 6, ptr(dot), stream())


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("C", [1, 3, 8, 40, 65])
def test_frontier_kernels_match_oracle(dtype, C):
    import torch
    from dynlp_amd.kernels import FrontierOps
    g, W, Ws, dt = _setup(dtype=dtype)
    n = g.n
    rs = np.random.default_rng(0)
    frontier = np.sort(rs.choice(n, 700, replace=False)).astype(np.int32)
    X, rhs = rs.random((n, C)).astype(dtype), rs.random((n, C)).astype(dtype)
    diag = (np.asarray(Ws.sum(axis=1)).ravel() + 1).astype(dtype)
    ops = FrontierOps(W, C)
    fd = torch.as_tensor(frontier, device="cuda")
    Y, ch = ops.jacobi(fd, torch.as_tensor(X, device="cuda"), torch.as_tensor(rhs, device="cuda"),
                       torch.as_tensor(diag, device="cuda"), 1e-2)
    Yo, cho = O.frontier_jacobi(Ws, frontier, X.astype(np.float64), rhs.astype(np.float64),
                                diag.astype(np.float64), 1e-2)
    rtol = 1e-5 if dtype == "float32" else 1e-12
    np.testing.assert_allclose(Y.cpu().numpy(), Yo, rtol=rtol)
    # flags agree except within rounding of the threshold
    ambiguous = np.abs(np.abs(Yo - X[frontier]).max(axis=1) - 1e-2) < 1e-5
    assert np.array_equal(ch.cpu().numpy()[~ambiguous], cho[~ambiguous])
    D = rs.standard_normal((frontier.size, C)).astype(dtype)
    R = torch.zeros((n, C), dtype=dt, device="cuda")
    mark = torch.zeros(n, dtype=torch.uint8, device="cuda")
    ops.push(fd, torch.as_tensor(D, device="cuda"), R, mark)
    Ro, mo = np.zeros((n, C)), np.zeros(n, dtype=np.uint8)
    O.push_scatter(Ws, frontier, D.astype(np.float64), Ro, mo)
    np.testing.assert_allclose(R.cpu().numpy(), Ro, rtol=rtol, atol=1e-5 if dtype == "float32" else 1e-12)
    assert np.array_equal(mark.cpu().numpy(), mo)
    mark.zero_()
    mo[:] = 0
    ops.mark_neighbors(fd, mark)
    O.mark_neighbors(Ws, frontier, mo)
    assert np.array_equal(mark.cpu().numpy(), mo)
    assert int(ops.edges) == 3 * O.edge_visits(Ws, frontier)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("C", [2, 4, 65])
def test_fused_push_reaches_tolerance(dtype, C):
    """tests/test_incremental.py:66-81: converges, kept R == rhs - A Z, max |R|/diag <= eps."""
    import torch
    from dynlp_amd.kernels import FusedPush
    n = 4000
    g = graphs.sbm(n, 2, 8.0, 0.85, 2, 0.03, 9)
    sysm = O.build_system(n, g.u, g.v, g.w, g.labeled, g.y, 2)
    from dynlp_amd import backend as B
    Wc = sysm.W.tocsr()
    W = B.csr_from_host(Wc.indptr, Wc.indices, Wc.data, n, torch.float32 if dtype == "float32" else torch.float64)
    rs = np.random.default_rng(0)
    rhs = (rs.random((n, C)) * sysm.solved[:, None]).astype(dtype)
    Z = np.zeros((n, C), dtype=dtype)
    diag = sysm.diag.astype(dtype)
    R = rhs.astype(np.float64) - O.apply_A(sysm.W, sysm.diag, Z.astype(np.float64))
    Zd = torch.as_tensor(Z, device="cuda")
    Rd = torch.as_tensor(R.astype(dtype), device="cuda")
    dd = torch.as_tensor(diag, device="cuda")
    eps = np.full(C, 1e-5 if dtype == "float64" else 1e-4)
    cand = torch.as_tensor(np.flatnonzero((np.abs(R) > 0).any(axis=1)).astype(np.int32), device="cuda")
    rounds, ok, edges = FusedPush(W, C).run(Zd, Rd, dd, eps, cand, 100_000)
    assert ok and rounds > 0 and edges > 0
    Zh = Zd.cpu().numpy().astype(np.float64)
    R_true = rhs.astype(np.float64) - O.apply_A(sysm.W, sysm.diag, Zh)
    np.testing.assert_allclose(Rd.cpu().numpy(), R_true, atol=1e-4 if dtype == "float32" else 1e-10)
    assert (np.abs(R_true) / sysm.diag[:, None]).max() / eps[0] <= 1.0 + (0.05 if dtype == "float32" else 1e-6)


def test_connected_components_match_union_find():
    """tests/test_dynlp.py:41-46, plus the closed form (bit-exact labels)."""
    import random
    from dynlp_amd.cc import connected_components
    rnd = random.Random(0)
    for n, m in [(400, 300), (5000, 4000), (20000, 30000), (1, 0)]:
        pairs = [(rnd.randrange(n), rnd.randrange(n)) for _ in range(m)]
        r = np.array([a for a, _ in pairs], dtype=np.int32)
        c = np.array([b for _, b in pairs], dtype=np.int32)
        lab, nc = connected_components(n, r, c)
        ref, ncr = O.connected_components(n, r, c)
        assert nc == ncr
        assert np.array_equal(lab.cpu().numpy(), ref)


def test_cc_csr_masked_matches_oracle():
    import torch
    from dynlp_amd.cc import connected_components_csr
    g, W, Ws, _ = _setup(n=20000, deg=2.5, seed=3)
    mask = (np.random.default_rng(5).random(g.n) < 0.7).astype(np.uint8)
    lab, nc = connected_components_csr(W, torch.as_tensor(mask, device="cuda"))
    coo = Ws.tocoo()
    keep = (mask[coo.row] > 0) & (mask[coo.col] > 0)
    ref, ncr = O.connected_components(g.n, coo.row[keep], coo.col[keep])
    assert nc == ncr and np.array_equal(lab.cpu().numpy(), ref)


@pytest.mark.parametrize("C", [1, 16, 17, 32, 33, 49, 64, 65, 80])
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_gram_on_matrix_cores_matches_numpy(C, dtype):
    """X^T diag(1/d) Y on v_mfma_f64_16x16x4f64 (also pins the operand/result lane layout): the
    register-tiled form (C = 16t or 16t + 1, t <= 4; general and X = Y operands) and the staged
    form (other C), on a row count that is not a multiple of the 16-row chunks."""
    import torch
    from dynlp_amd import _lib
    from dynlp_amd.backend import ptr, stream
    rs = np.random.default_rng(C)
    n = 1037
    X = rs.standard_normal((n, C)).astype(dtype)
    Y = rs.standard_normal((n, C)).astype(dtype)
    d = rs.uniform(1, 3, n).astype(dtype)
    G = torch.empty((C, C), dtype=torch.float64, device="cuda")
    Xd, Yd, dd = (torch.as_tensor(a, device="cuda") for a in (X, Y, d))
    dt = 0 if dtype == "float32" else 1
    _lib.call("dlp_gram", n, C, ptr(Xd), C, ptr(Yd), C, ptr(dd), ptr(G), dt, stream())
    ref = X.astype(np.float64).T @ (Y.astype(np.float64) / d.astype(np.float64)[:, None])
    np.testing.assert_allclose(G.cpu().numpy(), ref, rtol=1e-11, atol=1e-11)
    _lib.call("dlp_gram", n, C, ptr(Xd), C, ptr(Yd), C, None, ptr(G), dt, stream())
    np.testing.assert_allclose(G.cpu().numpy(), X.astype(np.float64).T @ Y.astype(np.float64), rtol=1e-11, atol=1e-11)
    _lib.call("dlp_gram", n, C, ptr(Xd), C, ptr(Xd), C, ptr(dd), ptr(G), dt, stream())
    ref = X.astype(np.float64).T @ (X.astype(np.float64) / d.astype(np.float64)[:, None])
    np.testing.assert_allclose(G.cpu().numpy(), ref, rtol=1e-11, atol=1e-11)


def test_cholesky_solve_and_deflation():
    import torch
    from dynlp_amd import _lib
    from dynlp_amd.backend import ptr, stream
    rs = np.random.default_rng(0)
    k = 65
    M = rs.standard_normal((k, 3 * k))
    A = M @ M.T + k * np.eye(k)
    B = rs.standard_normal((k, k))
    S = torch.empty((k, k), dtype=torch.float64, device="cuda")
    Ad, Bd = torch.as_tensor(A, device="cuda"), torch.as_tensor(B, device="cuda")
    _lib.call("dlp_chol_solve", k, ptr(Ad), ptr(Bd), ptr(S), 1e-12, stream())
    np.testing.assert_allclose(S.cpu().numpy(), np.linalg.solve(A, B), rtol=1e-9, atol=1e-12)
    # a collapsed direction (zero row/column) is dropped; the rest is the reduced solve
    A2 = A.copy()
    A2[7, :] = 0
    A2[:, 7] = 0
    _lib.call("dlp_chol_solve", k, ptr(torch.as_tensor(A2, device="cuda")), ptr(Bd), ptr(S), 1e-12, stream())
    Sh = S.cpu().numpy()
    keep = np.arange(k) != 7
    assert np.all(Sh[7] == 0)
    np.testing.assert_allclose(Sh[keep], np.linalg.solve(A[np.ix_(keep, keep)], B[keep]), rtol=1e-9, atol=1e-12)


def test_cc_labels_stable_over_repeats():
    """Regression for the final-pass race (a halving store overwriting another vertex's root):
    repeated runs on larger random inputs must all give the canonical labels."""
    from dynlp_amd.cc import connected_components
    rs = np.random.default_rng(11)
    n, m = 200_000, 150_000
    r = rs.integers(0, n, m).astype(np.int32)
    c = rs.integers(0, n, m).astype(np.int32)
    ref, ncr = O.connected_components(n, r, c)
    for _ in range(10):
        lab, nc = connected_components(n, r, c)
        assert nc == ncr and np.array_equal(lab.cpu().numpy(), ref)


def test_amg_coarsening_matches_scipy_on_dense_rows():
    """Galerkin coarsening W_c = off-diagonal part of P^T W P against SciPy on rows with
    thousands of entries: 40 random aggregates of an ER graph of mean degree 120, so every
    coarse row spans many 64-entry chunks and each column's run of duplicates crosses chunks."""
    import scipy.sparse as ssp
    import torch
    from dynlp_amd import amg as A
    from dynlp_amd import graphs
    from dynlp_amd.backend import csr_from_host
    rng = np.random.default_rng(5)
    n, nc = 3000, 40
    g = graphs.erdos_renyi(n, 120, 2, seed=5, unit=False)
    ip, ix, d = g.csr()
    W = csr_from_host(ip, ix, d.astype(np.float64), n, torch.float64)
    agg_h = rng.integers(0, nc, n).astype(np.int32)
    Wc = A._coarsen_W(W, torch.as_tensor(agg_h, device="cuda"), nc)
    P = ssp.csr_matrix((np.ones(n), (np.arange(n), agg_h)), shape=(n, nc))
    Ws = ssp.csr_matrix((d.astype(np.float64), ix, ip), shape=(n, n))
    ref = (P.T @ Ws @ P).tolil()
    ref.setdiag(0)
    ref = ref.tocsr()
    ref.eliminate_zeros()
    ref.sort_indices()
    got = Wc.to_scipy()
    assert np.array_equal(got.indptr, ref.indptr) and np.array_equal(got.indices, ref.indices)
    np.testing.assert_allclose(got.data, ref.data, rtol=1e-12)

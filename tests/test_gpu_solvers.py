"""Reference-API solvers and AMG on the device vs the FP64 oracle (patterns of the reference's
tests/test_dynlp.py:97-166, 225-243)."""
import numpy as np
import pytest

from dynlp_amd import graphs
from oracle import lp as O

pytestmark = pytest.mark.gpu


def _sys(g, dtype="float64"):
    from dynlp_amd.solvers import build_system
    return build_system(g, dtype)


def _dense(sys):
    W = sys.W.to_scipy().toarray().astype(np.float64)
    d = sys.diag.double().cpu().numpy()
    return np.diag(d) - W, sys.rhs.double().cpu().numpy(), d


def test_build_system_matches_oracle():
    g = graphs.sbm(1500, 6, 3, 0.8, 3, 0.03, seed=1)
    s = _sys(g)
    o = O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, g.K)
    assert np.array_equal(s.U.cpu().numpy(), o.U)
    Wl, dl, rl = o.local()
    np.testing.assert_allclose(s.W.to_scipy().toarray(), Wl.toarray(), rtol=1e-14)
    np.testing.assert_allclose(s.diag.cpu().numpy(), dl, rtol=1e-13)
    np.testing.assert_allclose(s.rhs.cpu().numpy(), rl, rtol=1e-13)


def test_amg_hierarchy_and_operator_match_oracle():
    """Matching is integer-deterministic (heaviest weight, ties -> smallest id), so the device
    hierarchy equals the oracle's, and M(I) agrees entrywise; M is SPD."""
    import torch
    from dynlp_amd.amg import AMG
    g = graphs.sbm(1200, 12, 6, 0.85, 2, 0.02, seed=7)
    s = _sys(g)
    amg = AMG(s.W, s.s, max_coarse=50)
    o = O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, 2)
    Wl, dl, rl = o.local()
    ref = O.AMG(Wl, o.s[o.U], max_coarse=50)
    assert amg.sizes == ref.sizes and len(amg.levels) >= 2
    for L, Lr in zip(amg.levels, ref.levels):
        assert np.array_equal(L.agg.cpu().numpy(), Lr["agg"])
    n = s.n_u
    Mi = amg(torch.eye(n, dtype=torch.float64, device="cuda")).cpu().numpy()
    np.testing.assert_allclose(Mi, ref(np.eye(n)), atol=1e-10)
    np.testing.assert_allclose(Mi, Mi.T, atol=1e-10)
    assert np.linalg.eigvalsh((Mi + Mi.T) / 2).min() > 0


@pytest.mark.parametrize("K", [2, 4])
def test_reference_matches_direct_solve(K):
    """tests/test_dynlp.py:97-106."""
    from dynlp_amd.solvers import Reference
    g = graphs.sbm(2000, 8, 6, 0.8, K, 0.02, seed=2)
    s = _sys(g)
    ref = Reference(g.n, K)
    F, hmax, _ = ref.solve(s)
    A, rhs, d = _dense(s)
    assert np.abs(F.cpu().numpy() - np.linalg.solve(A, rhs)).max() < 1e-7
    assert hmax >= np.linalg.solve(A, d).max() * (1 - 1e-6)


@pytest.mark.parametrize("K", [2, 3])
@pytest.mark.parametrize("method", ["pcg", "amg", "auto"])
def test_dynlp_plus_certificate_holds(K, method):
    """tests/test_dynlp.py:109-120."""
    from dynlp_amd.solvers import DynLPPlus
    g = graphs.sbm(3000, 6, 8, 0.8, K, 0.02, seed=3)
    s = _sys(g, "float32")
    solver = DynLPPlus(g.n, K, "float32", method=method, tol=1e-3)
    st = solver.solve(s)
    A, rhs, _ = _dense(s.astype("float64"))
    err = np.abs(solver.scores(s).cpu().numpy() - np.linalg.solve(A, rhs)).max()
    assert st.cert <= 1e-3 * 1.0001
    assert err <= st.cert * 1.05 + 1e-6, (err, st.cert)


def test_baselines_run_and_are_close():
    """tests/test_dynlp.py:123-133."""
    from dynlp_amd.solvers import make
    g = graphs.sbm(2000, 4, 8, 0.8, 2, 0.02, seed=4)
    s = _sys(g, "float32")
    A, rhs, _ = _dense(s.astype("float64"))
    Fd = np.linalg.solve(A, rhs)
    for nm in ["itlp", "dynlp"]:
        sv = make(nm, g.n, 2, dtype="float32", delta=1e-5, tol=1e-3, group="auto")
        sv.solve(s)
        assert np.abs(sv.scores(s).cpu().numpy() - Fd).max() < 0.05, nm


def test_multiclass_columns_consistent_with_binary():
    """tests/test_dynlp.py:148-156."""
    import torch
    from dynlp_amd.solvers import Stats, jacobi_M, pcg
    g = graphs.sbm(2000, 4, 6, 0.8, 2, 0.03, seed=6)
    s = _sys(g)
    M = jacobi_M(s)
    Z2 = pcg(s, torch.zeros((s.n_u, 2), dtype=torch.float64, device="cuda"), s.B.clone().contiguous(),
             [1e-12, 1e-12], M, 10_000, Stats())
    Z1 = pcg(s, torch.zeros((s.n_u, 1), dtype=torch.float64, device="cuda"), s.rhs.clone(), [1e-12], M, 10_000,
             Stats())
    np.testing.assert_allclose(Z2[:, 1].cpu().numpy(), Z1[:, 0].cpu().numpy(), atol=1e-8)
    np.testing.assert_allclose(Z2.sum(1).cpu().numpy(), 1.0, atol=1e-8)


@pytest.mark.parametrize("method", ["pcg", "amg", "auto"])
def test_refinement_certifies_beyond_float32(method):
    """tests/test_dynlp.py:225-243: path graph labelled at its ends, h > 1e4."""
    from dynlp_amd.solvers import DynLPPlus
    n = 600
    u = np.arange(n - 1)
    lab = np.zeros(n, dtype=bool)
    lab[[0, n - 1]] = True
    y = (np.arange(n) >= n // 2).astype(np.int32)
    g = graphs.Graph("path", n, u, u + 1, np.ones(n - 1, dtype=np.float32), lab, y, 2)
    s = _sys(g, "float32")
    A, rhs, d = _dense(s.astype("float64"))
    Fd, h = np.linalg.solve(A, rhs), np.linalg.solve(A, d)
    assert h.max() > 1e4
    solver = DynLPPlus(n, 2, "float32", method=method, tol=1e-3)
    st = solver.solve(s)
    assert st.cert <= 1e-3
    assert np.abs(solver.scores(s).cpu().numpy() - Fd).max() <= st.cert * 1.05 + 1e-9


@pytest.mark.parametrize("precond", ["jacobi", "amg"])
def test_block_pcg_matches_oracle_block_pcg(precond):
    """The device block PCG (FP64 Grams on the matrix cores, FP64 Cholesky steps, FP32 blocks)
    against the oracle's FP64 block PCG with the same preconditioner: both reach the stopping
    test, and the device solution is within the stopping bound of the direct solve."""
    import torch
    from dynlp_amd.amg import AMG
    from dynlp_amd.solvers import Stats, block_pcg, jacobi_M
    g = graphs.sbm(3000, 8, 10, 0.8, 16, 0.04, seed=11)
    s = _sys(g, "float32")
    o = O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, g.K)
    Wl, dl, rl = o.local()
    C = rl.shape[1]
    eps = [1e-5] * C
    M = jacobi_M(s) if precond == "jacobi" else AMG(s.W, s.s)
    Mo = None if precond == "jacobi" else O.AMG(Wl, o.s[o.U])
    st = Stats()
    Z = block_pcg(s, torch.zeros((s.n_u, C), dtype=torch.float32, device="cuda"), s.rhs.clone(), eps, M, 500, st)
    Xo, it_o = O.block_pcg(Wl, dl, np.zeros((s.n_u, C)), rl.copy(), eps, M=Mo)
    A, rhs, d = _dense(s.astype("float64"))
    Fd = np.linalg.solve(A, rhs)
    R = rhs - A @ Z.double().cpu().numpy()
    assert (np.abs(R) / d[:, None]).max() <= 1e-5 * 1.0001          # the true residual passes the test
    hmax = np.linalg.solve(A, d).max()
    assert np.abs(Z.double().cpu().numpy() - Fd).max() <= 1e-5 * hmax * 1.05
    assert np.abs(Xo - Fd).max() <= 1e-5 * hmax * 1.05
    assert 1 <= st.iters <= 1.5 * it_o + 3, (st.iters, it_o)           # FP32 blocks: about the FP64 count


def test_dynlp_plus_block_amg_certificate_holds_k64():
    """DynLP+ with the aggregation-AMG preconditioner and the block PCG (the mixed-precision
    configuration's stated method) at k = 64: the certified bound holds against the direct
    solve."""
    from dynlp_amd.solvers import DynLPPlus
    g = graphs.sbm(4000, 64, 12, 0.8, 64, 0.05, seed=12)
    s = _sys(g, "float32")
    solver = DynLPPlus(g.n, 64, "float32", method="amg", tol=1e-6, block=True)
    st = solver.solve(s)
    A, rhs, _ = _dense(s.astype("float64"))
    err = np.abs(solver.scores(s).cpu().numpy() - np.linalg.solve(A, rhs)).max()
    assert solver.name == "dynlp+amg-block" and st.iters >= 1
    assert st.cert <= 1e-6 * 1.0001
    assert err <= st.cert * 1.05 + 1e-9, (err, st.cert)

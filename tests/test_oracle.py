"""Pins for the FP64 oracle (CPU): every oracle function is checked against brute force,
closed forms or dense linear algebra, chosen so a plausible mistake (dropped term, wrong sign,
transposed index) fails.  Patterns follow the reference's tests/test_dynlp.py."""
import random

import numpy as np
import pytest
import scipy.sparse as sp

from dynlp_amd import graphs
from oracle import lp as O


def union_find(n, pairs):
    p = list(range(n))

    def f(x):
        while p[x] != x:
            p[x] = p[p[x]]
            x = p[x]
        return x

    for a, b in pairs:
        p[f(a)] = f(b)
    return [f(x) for x in range(n)]


def canonical_from_roots(roots):
    """Closed form: label = rank of the component's minimum vertex."""
    roots = np.asarray(roots)
    n = roots.shape[0]
    mins = {}
    for v in range(n):
        mins.setdefault(roots[v], v)
    order = sorted(mins.values())
    rank = {m: i for i, m in enumerate(order)}
    return np.array([rank[mins[roots[v]]] for v in range(n)], dtype=np.int32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_cc_bit_exact_vs_union_find_and_hook_jump(seed):
    rnd = random.Random(seed)
    n = 400
    pairs = [(rnd.randrange(n), rnd.randrange(n)) for _ in range(300)]
    r = np.array([a for a, _ in pairs])
    c = np.array([b for _, b in pairs])
    lab, nc = O.connected_components(n, r, c)
    ref = canonical_from_roots(union_find(n, pairs))
    assert np.array_equal(lab, ref) and nc == ref.max() + 1
    lab2, nc2 = O.cc_hook_jump(n, r, c)
    assert np.array_equal(lab2, ref) and nc2 == nc


def test_cc_edge_cases():
    lab, nc = O.connected_components(0, [], [])
    assert nc == 0 and lab.shape == (0,)
    lab, nc = O.connected_components(5, [], [])
    assert nc == 5 and np.array_equal(lab, np.arange(5))
    lab, nc = O.connected_components(4, [3, 2, 2], [3, 1, 1])   # self-loop + duplicate
    assert np.array_equal(lab, [0, 1, 1, 2]) and nc == 3


def _dense_system(g, eta=0.0):
    """Direct dense restatement of stream.py:142-177 with BFS grounding."""
    n, K = g.n, g.K
    A = np.zeros((n, n))
    A[g.u, g.v] = g.w
    A[g.v, g.u] = g.w
    unl = ~g.labeled
    # grounding: BFS over unlabeled vertices from those with a labeled neighbour
    Lnb = (A[:, g.labeled] > 0).any(axis=1) & unl
    grounded = Lnb.copy()
    frontier = list(np.flatnonzero(Lnb))
    while frontier:
        x = frontier.pop()
        for v in np.flatnonzero(A[x] > 0):
            if unl[v] and not grounded[v]:
                grounded[v] = True
                frontier.append(v)
    U = unl & grounded
    B = np.zeros((n, K))
    for l in np.flatnonzero(g.labeled):
        B[:, g.y[l]] += A[:, l]
    W = A * U[:, None] * U[None, :]
    s = B.sum(1) + eta
    diag = s + W.sum(1)
    return U, W, B, s, diag


def test_build_system_matches_dense_definition():
    g = graphs.sbm(300, 6, 2.5, 0.8, 3, 0.05, seed=5)
    sysm = O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, g.K, eta=0.01)
    U, W, B, s, diag = _dense_system(g, eta=0.01)
    assert np.array_equal(sysm.solved, U) and (~U).sum() > 0
    np.testing.assert_allclose(sysm.W.toarray(), W, rtol=1e-14)
    np.testing.assert_allclose(sysm.diag[U], diag[U], rtol=1e-13)
    np.testing.assert_allclose(sysm.rhs[U], B[U] + 0.01 / 3, rtol=1e-13)
    assert np.all(sysm.diag[~U] == 1) and np.all(sysm.rhs[~U] == 0)


def test_build_system_binary_uses_class_one_column():
    g = graphs.sbm(200, 2, 4, 0.8, 2, 0.1, seed=6)
    sysm = O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, 2)
    U, W, B, s, diag = _dense_system(g)
    assert sysm.C == 1
    np.testing.assert_allclose(sysm.rhs[U, 0], B[U, 1], rtol=1e-13)


def test_spmm_and_frontier_ops_vs_dense():
    g = graphs.sbm(500, 4, 8, 0.8, 2, 0.05, seed=7)
    ip, ix, d = g.csr()
    W = O.csr(ip, ix, d, g.n)
    Wd = W.toarray()
    assert np.array_equal(Wd, Wd.T)
    rs = np.random.default_rng(0)
    X, a = rs.random((g.n, 5)), rs.random(g.n)
    np.testing.assert_allclose(O.spmm_axpy(W, X, a, -1.0), a[:, None] * X - Wd @ X, rtol=1e-13)
    f = np.sort(rs.choice(g.n, 50, replace=False))
    rhs, diag = rs.random((g.n, 5)), Wd.sum(1) + 1
    Y, ch = O.frontier_jacobi(W, f, X, rhs, diag, 1e-2)
    np.testing.assert_allclose(Y, (rhs[f] + Wd[f] @ X) / diag[f, None], rtol=1e-13)
    assert np.array_equal(ch, (np.abs(Y - X[f]) > 1e-2).any(1))
    D = rs.standard_normal((50, 5))
    R, mark = np.zeros((g.n, 5)), np.zeros(g.n, dtype=np.uint8)
    O.push_scatter(W, f, D, R, mark)
    np.testing.assert_allclose(R, Wd[f].T @ D, rtol=1e-13, atol=1e-14)
    assert np.array_equal(mark.astype(bool), (Wd[f] != 0).any(0))
    m2 = np.zeros(g.n, dtype=np.uint8)
    O.mark_neighbors(W, f, m2)
    assert np.array_equal(m2, mark)
    assert O.edge_visits(W, f) == int((Wd[f] != 0).sum())


def _small_system(K=4, n=800, seed=11):
    g = graphs.sbm(n, 8, 6, 0.8, K, 0.03, seed=seed)
    return g, O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, g.K)


@pytest.mark.parametrize("which", ["push", "sync_push"])
def test_push_invariant_and_fixed_point(which):
    g, sysm = _small_system()
    C1 = sysm.C + 1
    Z = np.zeros((g.n, C1))
    Z[:, sysm.C] = 1
    R = O.residual(sysm, Z)
    eps = np.array([1e-8] * sysm.C + [1e-6])
    fn = O.push if which == "push" else O.sync_push
    Z, R, ok, rounds, _ = fn(sysm.W, sysm.diag, Z, R, eps)
    assert ok and rounds > 0
    np.testing.assert_allclose(R, O.residual(sysm, Z), atol=1e-12)   # kept residual is exact
    assert O.violation(sysm.diag, R, eps) <= 1
    Fs = O.solve_system(sysm)
    rc = O.rho_cols(sysm.diag, R)
    assert np.abs(Z - Fs)[:, : sysm.C].max() <= rc[: sysm.C].max() * Fs[:, sysm.C].max() * 1.0001


def test_one_sync_round_equals_one_host_round():
    g, sysm = _small_system()
    Z0 = np.zeros((g.n, sysm.C))
    R0 = sysm.rhs - O.apply_A(sysm.W, sysm.diag, Z0)
    eps = np.full(sysm.C, 1e-3)
    Za, Ra, *_ = O.push(sysm.W, sysm.diag, Z0.copy(), R0.copy(), eps, max_iter=1)
    Zb, Rb, *_ = O.sync_push(sysm.W, sysm.diag, Z0.copy(), R0.copy(), eps, max_rounds=1)
    np.testing.assert_allclose(Za, Zb, rtol=1e-14)
    np.testing.assert_allclose(Ra, Rb, rtol=1e-12, atol=1e-15)


def test_pcg_matches_exact_and_rows_sum_to_one():
    """tests/test_dynlp.py:148-156: K-column solve rows sum to 1 when every row is grounded."""
    g, sysm = _small_system(K=3)
    Wl, dl, rl = sysm.local()
    Z, it = O.pcg(Wl, dl, np.zeros_like(rl), rl, [1e-12] * sysm.C)
    Fx = O.exact_solve(Wl, dl, rl)
    np.testing.assert_allclose(Z, Fx, atol=1e-9)
    np.testing.assert_allclose(Z.sum(1), 1.0, atol=1e-9)
    Zb, _ = O.block_pcg(Wl, dl, np.zeros_like(rl), rl, [1e-12] * sysm.C)
    np.testing.assert_allclose(Zb, Fx, atol=1e-9)


def test_exact_solve_sparse_matches_dense():
    g, sysm = _small_system(K=2, n=1200)
    Wl, dl, rl = sysm.local()
    A = np.diag(dl) - Wl.toarray()
    np.testing.assert_allclose(O.exact_solve(Wl, dl, rl, dense_max=10), np.linalg.solve(A, rl), atol=1e-12)


def test_certificate_bound_is_valid_for_any_iterate():
    """tests/test_dynlp.py:136-145: |F - F*| <= rho * h for perturbed F."""
    g, sysm = _small_system(K=2)
    Wl, dl, rl = sysm.local()
    A = np.diag(dl) - Wl.toarray()
    Fd, h = np.linalg.solve(A, rl), np.linalg.solve(A, dl)
    rs = np.random.default_rng(0)
    for scale in (1e-1, 1e-3):
        F = Fd + scale * rs.standard_normal(Fd.shape)
        rho = float((np.abs(rl - A @ F) / dl[:, None]).max())
        assert np.all(np.abs(F - Fd) <= rho * h[:, None] + 1e-12)
    rc = np.array([1e-4, 0.01])
    hmax, bnd = O.bound(rc, 10.0, 1, 0.05)
    assert hmax == pytest.approx(10.0 / 0.99) and bnd == pytest.approx(1e-4 * 10.0 / 0.99)
    assert O.bound(np.array([1e-4, 0.1]), 10.0, 1, 0.05)[1] == float("inf")


def test_amg_is_symmetric_positive_definite():
    """tests/test_dynlp.py:159-166."""
    g = graphs.sbm(1200, 12, 6, 0.85, 2, 0.02, seed=7)
    sysm = O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, 2)
    Wl, dl, rl = sysm.local()
    amg = O.AMG(Wl, sysm.s[sysm.U], max_coarse=50)
    assert len(amg.levels) >= 2
    Mi = amg(np.eye(Wl.shape[0]))
    np.testing.assert_allclose(Mi, Mi.T, atol=1e-10)
    assert np.linalg.eigvalsh((Mi + Mi.T) / 2).min() > 0
    # as a preconditioner it solves the system
    Z, it = O.pcg(Wl, dl, np.zeros_like(rl), rl, [1e-10], M=amg)
    np.testing.assert_allclose(Z, O.exact_solve(Wl, dl, rl), atol=1e-8)


def test_amg_two_level_hand_example():
    """Path 0-1-2-3 with weights 3,1,3: handshake pairs {0,1},{2,3}; W_c = [[0,1],[1,0]]."""
    W = sp.csr_matrix(np.array([[0, 3, 0, 0], [3, 0, 1, 0], [0, 1, 0, 3], [0, 0, 3, 0]], dtype=float))
    coo = W.tocoo()
    agg, nc = O._handshake(4, coo.row, coo.col, coo.data, 3)
    assert nc == 2 and list(agg) == [0, 0, 1, 1]
    Wc, sc = O._coarsen(W, np.array([1.0, 0, 0, 2.0]), agg, nc)
    np.testing.assert_allclose(Wc.toarray(), [[0, 1], [1, 0]])
    np.testing.assert_allclose(sc, [1, 2])


def test_edge_batch_apply_semantics():
    es = {(0, 1): 1.0, (1, 2): 2.0}
    O.apply_edge_batch(es, [(3, 0, 0.5), (2, 1, 7.0), (4, 4, 1.0)], [(1, 0), (5, 6)])
    assert es == {(1, 2): 7.0, (0, 3): 0.5}


def test_planted_sign_error_is_caught():
    """A push with the scatter sign flipped breaks the exact-residual invariant the GPU tests
    rely on; the check must notice."""
    g, sysm = _small_system()
    Z = np.zeros((g.n, sysm.C))
    R = sysm.rhs - O.apply_A(sysm.W, sysm.diag, Z)
    F = np.flatnonzero((np.abs(R) > 0).any(1))
    D = R[F] / sysm.diag[F, None]
    Z[F] += D
    R[F] = 0
    Dfull = np.zeros_like(R)
    Dfull[F] = D
    R -= sysm.W @ Dfull          # wrong sign
    assert np.abs(R - (sysm.rhs - O.apply_A(sysm.W, sysm.diag, Z))).max() > 1e-3


def test_cpu_incremental_port_certifies_each_batch():
    from oracle.incremental import CPUIncremental
    g = graphs.sbm(1500, 6, 8, 0.8, 6, 0.03, seed=12)
    cpu = CPUIncremental(g.n, g.u, g.v, g.w, g.labeled, g.y, g.K, eps=1e-6, push_max_rows=50)
    batches, es = graphs.update_stream(g, 3, 200, seed=12)
    for b in batches:
        out = cpu.solve_batch(b.ins_u, b.ins_v, b.ins_w, b.del_u, b.del_v)
        assert out["rho"] <= 1e-6
    u, v, w = es.pairs()
    sysm = O.build_system(g.n, u, v, w, g.labeled, g.y, g.K)
    R = O.residual(sysm, cpu.Z)
    assert O.rho_cols(sysm.diag, R)[: sysm.C].max() <= 1e-6 * 1.0001

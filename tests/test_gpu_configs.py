"""The CPU-case and ER-1M workloads on the device against the oracle.

ER n = 100,000, degree 10, k = 10, 1 % labelled (run at a fifth of its size for the full oracle
comparison): harmonic label propagation by Jacobi to rho <= 1e-6, then batches of 50/50 edge
updates repaired by residual push.  ER n = 1,000,000, degree 16, k = 16: an update stream with
the solved set, the scores and the connected-component labels checked after every batch (at a
fiftieth of the size against the exact oracle, and at full size through the properties that do
not need a solve: CC labels bit-exact, the solved set, and the FP64 residual the oracle computes
from the device's scores)."""
import numpy as np
import pytest

from dynlp_amd import graphs
from oracle import lp as O

pytestmark = pytest.mark.gpu


def _edges(es):
    u, v, w = es.pairs()
    return u, v, w


def _oracle_check(lp, g, es, eps, label, exact=True):
    """Solved set exact, diag to 1e-12, rho <= eps from the oracle's FP64 residual of the
    device scores, |F - F*| within the certificate (F* by FP64 CG to 1e-13 when `exact`), CC
    labels bit-exact."""
    u, v, w = _edges(es)
    sysm = O.build_system(g.n, u, v, w, g.labeled, g.y, g.K)
    assert np.array_equal(lp.solved_mask(), sysm.solved), label
    np.testing.assert_allclose(lp.diag64.cpu().numpy(), sysm.diag, rtol=1e-12, err_msg=label)
    Z = lp.scores_with_h()
    R = O.residual(sysm, Z)
    rc = O.rho_cols(sysm.diag, R)
    rho = rc[: sysm.C].max()
    assert rho <= eps * 1.0001, (label, rho)
    if exact:
        Wl, dl, rl = sysm.local()
        cols = np.concatenate([rl, dl[:, None]], axis=1)
        Xs, _ = O.pcg(Wl, dl, np.zeros_like(cols), cols, [1e-13] * cols.shape[1])
        Fs = np.zeros_like(Z)
        Fs[:, sysm.C] = 1.0
        Fs[sysm.U] = Xs
        hmax = Fs[:, sysm.C].max()
        err = np.abs(Z[:, : sysm.C] - Fs[:, : sysm.C]).max()
        assert err <= rho * hmax * 1.05 + 1e-12, (label, err, rho * hmax)
    lab = lp.components()
    ref, _ = O.connected_components(g.n, np.r_[u, v], np.r_[v, u])
    assert np.array_equal(lab, ref), label
    return rho


def test_cpu_case_jacobi_reaches_rho():
    """The reference API's plain Jacobi (ItLP) on ER-20k, k = 10: a Jacobi sweep moves every
    value by R/diag, so stopping at max |dX| <= 1e-6 leaves rho <= 1e-6 (checked with the
    oracle's FP64 residual), and F is within rho * hmax of F*."""
    from dynlp_amd.solvers import ItLP, build_system
    g = graphs.config_graph(1, scale=0.2, seed=0)
    sysd = build_system(g, "float64")
    solver = ItLP(g.n, g.K, "float64", delta=1e-6)
    st = solver.solve(sysd)
    F = solver.scores(sysd).cpu().numpy()
    sysm = O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, g.K)
    Wl, dl, rl = sysm.local()
    R = rl - O.apply_A(Wl, dl, F)
    rho = float((np.abs(R) / dl[:, None]).max())
    assert rho <= 1e-6 * 1.0001 and st.iters > 10, (rho, st.iters)
    cols = np.concatenate([rl, dl[:, None]], axis=1)
    Xs, _ = O.pcg(Wl, dl, np.zeros_like(cols), cols, [1e-13] * cols.shape[1])
    assert np.abs(F - Xs[:, :-1]).max() <= rho * Xs[:, -1].max() * 1.05


def test_cpu_case_push_batches_match_oracle():
    """ER-20k, k = 10: the initial solve, then 10 batches of 200 updates (50/50), each repaired
    by residual push alone (the policy set to always push), checked against the oracle."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.config_graph(1, scale=0.2, seed=0)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps, push_frac=1.0, push_max_rows=g.n, push_mass=float("inf"), push_budget=1e9)
    lp.initial_solve()
    batches, _ = graphs.update_stream(g, 10, 200, seed=0)
    es = graphs.EdgeSet(g)
    for t, b in enumerate(batches):
        st = lp.apply_and_solve(b)
        es.apply(b)
        assert st.path == "push", (t, st.path)
        _oracle_check(lp, g, es, eps, f"batch {t}")


def test_er_stream_scaled_matches_oracle():
    """ER-20k, degree 16, k = 16 (the ER-1M workload at 1/50): 10 batches of 200 updates with
    the service's default policy; after every batch solved set, scores and CC labels against the
    oracle."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.config_graph(2, scale=0.02, seed=0)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps)
    lp.initial_solve()
    batches, _ = graphs.update_stream(g, 10, 200, seed=0)
    es = graphs.EdgeSet(g)
    for t, b in enumerate(batches):
        lp.apply_and_solve(b)
        es.apply(b)
        _oracle_check(lp, g, es, eps, f"batch {t}")


def _canonical_cc(n, u, v):
    """The oracle's CC labels by SciPy (the rank of each component's smallest vertex), for
    graphs where the oracle's union-find restatement would take minutes."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    A = sp.coo_matrix((np.ones(u.shape[0]), (u, v)), shape=(n, n))
    nc, lab = connected_components(A, directed=False)
    first = np.full(nc, n, dtype=np.int64)
    np.minimum.at(first, lab, np.arange(n))
    rank = np.empty(nc, dtype=np.int64)
    rank[np.argsort(first)] = np.arange(nc)
    return rank[lab].astype(np.int32)


def test_er_1m_stream_cc_and_residual_at_full_size():
    """ER-1M, degree 16, k = 16, the stated 100 batches of 10,000 updates: after every batch the
    CC labels are bit-exact against SciPy's components ranked as the oracle ranks them and the
    request's certificate holds (rho <= 1e-6); every 25 batches also the solved set against the
    oracle's grounding and the oracle's FP64 residual of the device scores."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.config_graph(2, scale=1.0, seed=0)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps)
    lp.initial_solve()
    batches, _ = graphs.update_stream(g, 100, 10_000, seed=0)
    es = graphs.EdgeSet(g)
    for t, b in enumerate(batches):
        st = lp.apply_and_solve(b)
        es.apply(b)
        assert st.rho <= eps * 1.0001, (t, st.rho)
        u, v, _ = es.pairs()
        assert np.array_equal(lp.components(), _canonical_cc(g.n, u, v)), t
        if t % 25 == 24:
            _oracle_check(lp, g, es, eps, f"ER-1M batch {t}", exact=False)


def test_graph_from_edge_list_file_runs_against_oracle(tmp_path):
    """A graph that enters through the reference's edge-list format (graphs.py:176-204): pairs
    in both directions, repeated (weights summed) and self-loops, written to .npz and read with
    from_npz; the device path (initial solve and two update batches) against the oracle built
    from the file's own arrays coalesced the same way (SciPy COO sums duplicates)."""
    import scipy.sparse as sp
    from dynlp_amd.stream import EdgeStreamLP
    rs = np.random.RandomState(4)
    n, m, K = 3000, 15000, 4
    src = rs.randint(0, n, m)
    dst = rs.randint(0, n, m)
    src[:500], dst[:500] = dst[500:1000], src[500:1000]          # the same pairs reversed
    w = rs.uniform(0.5, 1.5, m).astype(np.float32)
    src[1000:1050] = dst[1000:1050]                               # self-loops (dropped)
    y = rs.randint(0, K, n).astype(np.int32)
    path = tmp_path / "g.npz"
    np.savez(path, src=src, dst=dst, w=w, y=y)
    g = graphs.from_npz(str(path), label_frac=0.02, seed=4)
    # the oracle's reading of the same arrays: undirected, duplicates summed in FP32, loops out
    a, b = np.minimum(src, dst), np.maximum(src, dst)
    keep = a != b
    A = sp.coo_matrix((w[keep].astype(np.float32), (a[keep], b[keep])), shape=(n, n)).tocsr()
    A.sum_duplicates()
    coo = A.tocoo()
    got = {(int(x), int(z)): float(v) for x, z, v in zip(g.u, g.v, g.w)}
    want = {(int(x), int(z)): float(v) for x, z, v in zip(coo.row, coo.col, coo.data)}
    assert got.keys() == want.keys()
    assert max(abs(got[k] - want[k]) / want[k] for k in got) <= 1e-6
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps)
    lp.initial_solve()
    es = graphs.EdgeSet(g)
    _oracle_check(lp, g, es, eps, "from file")
    for t, bt in enumerate(graphs.update_stream(g, 2, 300, seed=4)[0]):
        lp.apply_and_solve(bt)
        es.apply(bt)
        _oracle_check(lp, g, es, eps, f"from file, batch {t}")


def test_fragmenting_sbm300k_stream_at_full_size():
    """The dynamic-connectivity workload at its stated shape: SBM n = 300,000, 32 blocks, degree
    8, 90 % intra-block, unit weights, 1 % labelled, k = 32; 200 batches of 5,000 updates, 70 %
    deletions of inter-block edges first, so the graph fragments.  After every batch the CC
    labels are bit-exact against the oracle, the solved set equals the oracle's grounding, and
    the request's FP64 certificate holds (rho <= 1e-6)."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.config_graph(4, scale=1.0, seed=0)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps)
    lp.initial_solve()
    batches, _ = graphs.update_stream(g, 200, 5_000, ins_frac=0.3, seed=0, inter_first=True)
    es = graphs.EdgeSet(g)
    toggled = 0
    for t, b in enumerate(batches):
        st = lp.apply_and_solve(b)
        es.apply(b)
        toggled += st.toggled
        assert st.rho <= eps * 1.0001, (t, st.rho)
        u, v, _ = es.pairs()
        assert np.array_equal(lp.components(), _canonical_cc(g.n, u, v)), t
        # grounding: unlabeled vertices whose component in the unlabeled subgraph touches a
        # labeled vertex
        un = ~g.labeled
        keep = un[u] & un[v]
        Au = sp.coo_matrix((np.ones(int(keep.sum())), (u[keep], v[keep])), shape=(g.n, g.n))
        _, lu = connected_components(Au, directed=False)
        touch = np.zeros(lu.max() + 1, dtype=bool)
        lab_adj = np.zeros(g.n, dtype=bool)
        lab_adj[u[g.labeled[v]]] = True
        lab_adj[v[g.labeled[u]]] = True
        touch[lu[un & lab_adj]] = True
        solved = un & touch[lu]
        assert np.array_equal(lp.solved_mask(), solved), t
    assert toggled > 0

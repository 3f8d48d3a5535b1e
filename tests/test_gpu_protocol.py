"""The reference's vertex-arrival protocol (stream.Stream, stream.IncrementalStream) feeding
the reference-API solvers, as the reference's tests exercise it (tests/test_dynlp.py:74-135,
tests/test_incremental.py:19-43): each batch system against its dense definition, the
global-index systems against the local ones, and the solvers' certificates against direct
solves of the same systems."""
import numpy as np
import pytest

from dynlp_amd import graphs

pytestmark = pytest.mark.gpu


def same_partition(a, b):
    """Two labelings describe the same partition."""
    a, b = list(a), list(b)
    return len(a) == len(b) and len(set(zip(a, b))) == len(set(a)) == len(set(b))


def dense_A(g):
    ip, ix, d = g.csr()
    A = np.zeros((g.n, g.n))
    A[np.repeat(np.arange(g.n), np.diff(ip)), ix] = d
    return A


def dense_system(sysm):
    W = sysm.W.to_scipy().toarray().astype(np.float64)
    diag = sysm.diag.double().cpu().numpy()
    return np.diag(diag) - W, sysm.rhs.double().cpu().numpy()


def test_system_matches_dense_definition():
    from dynlp_amd.cc import connected_components
    from dynlp_amd.stream import Stream
    g = graphs.sbm(1500, 3, 6, 0.85, 3, 0.02, seed=1)
    st = Stream(g, init_frac=0.2, n_batches=3, label_frac=0.03, del_frac=0.1, dtype="float64")
    A, y, lab = dense_A(g), g.y, st.labeled
    act = np.zeros(g.n, dtype=bool)
    assert len(st) == 4
    for sysm in st.batches():
        ins, dele = st.schedule[sysm.t]
        act[ins] = True
        act[dele] = False
        U = sysm.U.cpu().numpy()
        assert not lab[U].any() and act[U].all()
        assert sysm.n_inserted == len(ins) and sysm.n_deleted == len(dele) and sysm.n_active == act.sum()
        W = sysm.W.to_scipy().toarray()
        np.testing.assert_allclose(W, A[U][:, U], rtol=1e-12)
        al = act & lab
        B = np.stack([A[U][:, al & (y == k)].sum(1) for k in range(3)], axis=1)
        np.testing.assert_allclose(sysm.B.cpu().numpy(), B, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(sysm.diag.cpu().numpy(), sysm.s.cpu().numpy() + W.sum(1), rtol=1e-12)
        # every solved component touches a labeled vertex; every other active unlabeled one does not
        r, c = np.nonzero(W)
        cl, nc = connected_components(len(U), r, c)
        gw = np.zeros(nc.item() if hasattr(nc, "item") else nc)
        np.add.at(gw, cl.cpu().numpy(), sysm.B.sum(1).cpu().numpy())
        assert (gw > 0).all()
        unl = np.flatnonzero(act & ~lab)
        assert sysm.n_ungrounded == len(unl) - len(U)


def test_incremental_stream_matches_rebuilt_system():
    """Sparse graph and heavy deletions (splits, regrounding): the global-index system equals
    the local one on U, its other rows are inert, and `changed` lists every row of U whose
    equation differs from the previous batch."""
    from dynlp_amd.stream import IncrementalStream, Stream
    kw = dict(init_frac=0.3, n_batches=6, del_frac=0.4, label_frac=0.03, dtype="float64", seed=3)
    g = graphs.sbm(4000, 3, 3.0, 0.9, 3, 0.01, seed=5, unit=True)
    prev = None
    for c, s in zip(Stream(g, **kw).batches(), IncrementalStream(g, ball_hops=3, **kw).batches()):
        U = c.U.cpu().numpy()
        solved = s.solved.cpu().numpy()
        assert np.array_equal(np.flatnonzero(solved), U)
        assert s.n_ungrounded == c.n_ungrounded and s.n_active == c.n_active
        Wg = s.W.to_scipy().tocsr()
        np.testing.assert_allclose(Wg[U][:, U].toarray(), c.W.to_scipy().toarray(), rtol=1e-12, atol=1e-12)
        for f in ("s", "diag", "rhs", "B"):
            np.testing.assert_allclose(getattr(s, f).cpu().numpy()[U], getattr(c, f).cpu().numpy(),
                                       rtol=1e-10, atol=1e-12, err_msg=f)
        assert np.array_equal(s.is_new.cpu().numpy()[U], c.is_new.cpu().numpy())
        assert np.array_equal(s.seed.cpu().numpy()[U], c.seed.cpu().numpy())
        for k in ("tau", "fh"):
            assert s.comps[k][1] == c.comps[k][1], k
            assert same_partition(s.comps[k][0].cpu().numpy(), c.comps[k][0].cpu().numpy()), k
        out = ~solved
        assert abs(Wg[out]).sum() == 0 and abs(Wg[:, out]).sum() == 0
        assert (s.diag.cpu().numpy()[out] == 1.0).all() and (s.rhs.cpu().numpy()[out] == 0.0).all()
        diag, rhs = s.diag.cpu().numpy(), s.rhs.cpu().numpy()
        if prev is not None:
            ch = np.zeros(g.n, dtype=bool)
            ch[s.changed.cpu().numpy()] = True
            assert ch[U[c.is_new.cpu().numpy()]].all()
            pW, pdiag, prhs = prev
            differs = (np.abs(diag - pdiag) > 0) | (np.abs(rhs - prhs).max(1) > 0) | \
                (np.abs(Wg - pW).sum(axis=1).A.ravel() > 0)
            assert ch[differs & solved].all()
            assert not ch[~solved].any()
        prev = (Wg, diag, rhs)


@pytest.mark.parametrize("K", [2, 3])
@pytest.mark.parametrize("method", ["pcg", "amg", "auto"])
def test_dynlp_plus_certificate_holds(K, method):
    from dynlp_amd.solvers import DynLPPlus
    from dynlp_amd.stream import Stream
    g = graphs.sbm(3000, K, 8, 0.85, K, 0.02, seed=3)
    st = Stream(g, init_frac=0.2, n_batches=3, label_frac=0.02, dtype="float32")
    solver = DynLPPlus(g.n, K, "float32", method=method, tol=1e-3)
    for sysm in st.batches():
        stats = solver.solve(sysm)
        A, rhs = dense_system(sysm.astype("float64"))
        err = float(np.abs(solver.scores(sysm).double().cpu().numpy() - np.linalg.solve(A, rhs)).max())
        assert stats.cert <= 1e-3 * 1.0001
        assert err <= stats.cert * 1.05 + 1e-6, (err, stats.cert)


def test_solvers_on_global_index_systems():
    """The same solvers on the inert-row systems of IncrementalStream: certified on the solved
    rows, and the inert rows (exactly 0) within the same certificate."""
    from dynlp_amd.solvers import DynLPPlus
    from dynlp_amd.stream import IncrementalStream
    g = graphs.sbm(3000, 3, 6, 0.85, 3, 0.02, seed=6)
    st = IncrementalStream(g, init_frac=0.85, n_batches=4, del_frac=0.1, label_frac=0.02, dtype="float32")
    solver = DynLPPlus(g.n, 3, "float32", method="pcg", tol=1e-3)
    for sysm in st.batches():
        stats = solver.solve(sysm)
        s64 = sysm.astype("float64")
        U = np.flatnonzero(sysm.solved.cpu().numpy())
        A, rhs = dense_system(s64)
        F = solver.scores(sysm).double().cpu().numpy()
        err = float(np.abs(F[U] - np.linalg.solve(A[U][:, U], rhs[U])).max())
        assert stats.cert <= 1e-3 * 1.0001
        assert err <= stats.cert * 1.05 + 1e-6, (sysm.t, err, stats.cert)
        out = np.setdiff1d(np.arange(g.n), U)
        # inert rows (exact value 0, warm-started from earlier batches) fall under the same bound
        assert np.abs(F[out]).max() <= stats.cert * 1.05 + 1e-9


def test_baselines_run_and_are_close():
    from dynlp_amd.solvers import make
    from dynlp_amd.stream import Stream
    g = graphs.sbm(2000, 2, 8, 0.85, 2, 0.02, seed=4)
    st = Stream(g, init_frac=0.3, n_batches=2, label_frac=0.02, dtype="float32")
    solvers = [make(nm, g.n, 2, dtype="float32", delta=1e-5, tol=1e-3, group="auto")
               for nm in ["itlp", "dynlp", "dynlp-fh"]]
    for sysm in st.batches():
        A, rhs = dense_system(sysm.astype("float64"))
        Fd = np.linalg.solve(A, rhs)
        for s in solvers:
            s.solve(sysm)
            assert float(np.abs(s.scores(sysm).double().cpu().numpy() - Fd).max()) < 0.05, s.name


def test_backend_sparse_module():
    """backend.sp.csr_matrix from (data, indices, indptr) and from a SciPy matrix, as the
    reference's callers build W; dtypes follow the data."""
    import scipy.sparse as ssp
    import torch
    from dynlp_amd import backend as B
    m = ssp.random(50, 50, density=0.1, format="csr", random_state=1, dtype=np.float64)
    W = B.sp.csr_matrix(m)
    assert W.data.dtype == torch.float64 and W.shape == (50, 50)
    np.testing.assert_array_equal(W.to_scipy().toarray(), m.toarray())
    W32 = B.sp.csr_matrix((m.data.astype(np.float32), m.indices, m.indptr), shape=(50, 50))
    assert W32.data.dtype == torch.float32 and W32.indices.dtype == torch.int32
    X = torch.rand((50, 3), dtype=torch.float32, device="cuda")
    np.testing.assert_allclose((W32 @ X).cpu().numpy(), m.toarray().astype(np.float32) @ X.cpu().numpy(),
                               rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("K", [2, 3])
def test_incremental_solver_is_certified_on_small_batches(K):
    """dynlp+inc on IncrementalStream systems (tests/test_incremental.py:45-65 of the reference):
    every batch certified to tol, the scores within the certificate of a direct solve, the kept
    residual equal to a fresh one, and small batches repaired by push alone."""
    import torch
    from dynlp_amd.solvers import IncrementalLP, apply_A, make
    from dynlp_amd.stream import IncrementalStream
    tol = 1e-3
    g = graphs.sbm(3000, K, 6, 0.85, K, 0.02, seed=6)
    st = IncrementalStream(g, init_frac=0.85, n_batches=8, del_frac=0.1, label_frac=0.02, dtype="float32")
    solver = IncrementalLP(g.n, K, "float32", tol=tol, full_every=4, push_frac=0.5)
    assert make("dynlp+inc", g.n, K, dtype="float32", delta=1e-4, tol=tol, group="auto").name == "dynlp+inc"
    paths = []
    for sysm in st.batches():
        stats = solver.solve(sysm)
        paths.append(stats.path)
        s64 = sysm.astype("float64")
        U = np.flatnonzero(sysm.solved.cpu().numpy())
        A, rhs = dense_system(s64)
        err = float(np.abs(solver.scores(sysm).double().cpu().numpy()[U] - np.linalg.solve(A[U][:, U], rhs[U])).max())
        assert stats.cert <= tol * 1.0001, (sysm.t, stats)
        assert err <= stats.cert * 1.05 + 1e-9, (sysm.t, err, stats.cert)
        R_true = torch.cat([s64.rhs, s64.diag[:, None]], dim=1) - apply_A(s64, solver.Z)
        drift = float(((R_true - solver.R).abs() / s64.diag[:, None]).max())
        assert drift < 1e-9, drift
    assert "inc-push" in paths, paths


def _components(edges, m, tau, ground=None):
    """new_components on an explicit undirected edge list (all vertices new)."""
    from dynlp_amd.stream import new_components
    a, b, w = zip(*edges)
    gr = np.ones(m) if ground is None else np.asarray(ground, dtype=np.float64)
    return new_components(m, np.asarray(a + b), np.asarray(b + a), np.asarray(w + w, dtype=np.float32), tau, 1.0, gr)


def test_adaptive_components_cut_a_bridge_that_tau_chains():
    """Two tight cliques joined by an edge above the global mean but weak relative to both
    (the reference's tests/test_dynlp.py:56-62)."""
    edges = [(a + o, b + o, 0.95 + 0.01 * ((a + b) % 3)) for o in (0, 6) for a in range(6) for b in range(a + 1, 6)]
    edges.append((5, 6, 0.8))
    comps = _components(edges, 12, tau=0.5)
    assert comps["tau"][1] == 1
    assert comps["fh"][1] == 2 and same_partition(comps["fh"][0].cpu().numpy(), [0] * 6 + [1] * 6)


def test_adaptive_components_merge_a_sparse_cluster():
    """A weak chain (every edge below tau) stays singletons under tau but merges under fh, and a
    component without labeled contact joins its strongest neighbour (tests/test_dynlp.py:65-71)."""
    edges = [(i, i + 1, 0.3 + 0.02 * (i % 3)) for i in range(7)]
    comps = _components(edges, 8, tau=0.5, ground=[1, 0, 0, 0, 0, 0, 0, 0])
    assert comps["tau"][1] == 8
    assert comps["fh"][1] == 1


def test_supernode_init_by_hand():
    """Two supernodes of new rows: DynLP's rule gives each its share of class weight (prior when
    it has none); with known rows, the weighted average over labeled and known neighbours."""
    import torch
    from dynlp_amd.backend import CSR
    from dynlp_amd.solvers import System, init_h, supernode_init
    dev = "cuda"
    # rows 0,1 new (component 0), row 2 new (component 1), row 3 known; edges 0-1 (1), 1-3 (2), 2-3 (1)
    indptr = torch.tensor([0, 1, 3, 4, 6], dtype=torch.int32, device=dev)
    indices = torch.tensor([1, 0, 3, 3, 1, 2], dtype=torch.int32, device=dev)
    data = torch.tensor([1, 1, 2, 1, 2, 1], dtype=torch.float64, device=dev)
    W = CSR(data, indices, indptr, (4, 4))
    B = torch.tensor([[3, 1, 0], [0, 0, 0], [0, 0, 0], [1, 1, 2]], dtype=torch.float64, device=dev)
    prior = torch.full((3,), 1 / 3, dtype=torch.float64, device=dev)
    is_new = torch.tensor([True, True, True, False], device=dev)
    comps = {"tau": (torch.tensor([0, 0, 1], dtype=torch.int32, device=dev), 2)}
    comps["fh"] = comps["tau"]
    sysm = System(torch.arange(4, device=dev), W, B.sum(1), B.sum(1) + 2, B, B.clone(), prior, 3, 3,
                  is_new=is_new, seed=is_new, comps=comps)
    X = torch.zeros((4, 3), dtype=torch.float64, device=dev)
    X[3] = torch.tensor([0.2, 0.3, 0.5], dtype=torch.float64, device=dev)
    Y = supernode_init(sysm, X.clone())
    np.testing.assert_allclose(Y[0].cpu().numpy(), [0.75, 0.25, 0.0])
    np.testing.assert_allclose(Y[1].cpu().numpy(), [0.75, 0.25, 0.0])
    np.testing.assert_allclose(Y[2].cpu().numpy(), [1 / 3] * 3)     # no class weight: the prior
    known = ~is_new
    Y = supernode_init(sysm, X.clone(), known)
    # component 0: class weight (3, 1, 0) + 2 * X[3] = (3.4, 1.6, 1.0) over 4 + 2 = 6
    np.testing.assert_allclose(Y[0].cpu().numpy(), [3.4 / 6, 1.6 / 6, 1.0 / 6])
    np.testing.assert_allclose(Y[2].cpu().numpy(), [0.2, 0.3, 0.5])
    H = torch.zeros((4, 1), dtype=torch.float64, device=dev)
    H[3] = 10.0
    Hn = init_h(sysm, H.clone(), known, 7.0)
    # component 0: (2 * 10) / (4 + 2) + 1; component 1: (1 * 10) / (0 + 1) + 1
    np.testing.assert_allclose(Hn[:3, 0].cpu().numpy(), [20 / 6 + 1, 20 / 6 + 1, 11.0])

"""The oracle reproduces the committed golden fixtures from their stored inputs (CPU)."""
import numpy as np
import pytest

from oracle import lp as O
from tests.golden import fixtures as F


@pytest.mark.parametrize("name", ["sys_er_k2", "sys_er_k16", "sys_sbm_k64"])
def test_system_fixtures(name):
    fx = F.load(name)
    g = F.graph(fx)
    S = O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, g.K)
    assert np.array_equal(S.solved, fx["solved"].astype(bool))
    np.testing.assert_allclose(S.diag, fx["diag"], rtol=1e-14)
    Z = O.solve_system(S)
    np.testing.assert_allclose(Z, fx["Z"], rtol=0, atol=1e-10)
    # the stored solution solves the rebuilt system; labels of a grounded row sum to one
    assert O.rho_cols(S.diag, O.residual(S, fx["Z"])).max() < 1e-10
    if g.K > 2:
        U = S.solved
        np.testing.assert_allclose(fx["Z"][U, : S.C].sum(axis=1), 1.0, atol=1e-9)


@pytest.mark.parametrize("name,scores", [("stream_sbm_k8", True), ("cc_fragmenting", False)])
def test_stream_fixtures(name, scores):
    fx = F.load(name)
    g = F.graph(fx)
    es = {(int(a), int(b)): float(x) for a, b, x in zip(g.u, g.v, g.w)}
    for t, b in enumerate(F.batches(fx)):
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
        u = np.array([k[0] for k in es], dtype=np.int64)
        v = np.array([k[1] for k in es], dtype=np.int64)
        lab, _ = O.connected_components(g.n, np.r_[u, v], np.r_[v, u])
        assert np.array_equal(lab, fx[f"b{t}_cc"])
        lab2, _ = O.cc_hook_jump(g.n, np.r_[u, v], np.r_[v, u])
        assert np.array_equal(lab2, fx[f"b{t}_cc"])
        S = O.build_system(g.n, u, v, np.array(list(es.values())), g.labeled, g.y, g.K)
        assert np.array_equal(S.solved, fx[f"b{t}_solved"].astype(bool))
        if scores:
            np.testing.assert_allclose(O.solve_system(S), fx[f"b{t}_Z"], rtol=0, atol=1e-10)
    if name == "cc_fragmenting":
        assert len(np.unique(fx["b4_cc"])) > len(np.unique(fx["b0_cc"]))   # the stream fragments


def test_spmm_fixture():
    fx = F.load("spmm")
    n = fx["indptr"].shape[0] - 1
    W = O.csr(fx["indptr"], fx["indices"], fx["data"], n)
    for C in (1, 17, 65):
        Y = O.spmm_axpy(W, fx[f"X{C}"].astype(np.float64), fx["a"].astype(np.float64), -1.0)
        np.testing.assert_allclose(Y, fx[f"Y{C}"], rtol=1e-14, atol=1e-14)
        # dense definition
        D = W.toarray().astype(np.float64)
        X = fx[f"X{C}"].astype(np.float64)
        np.testing.assert_allclose(fx[f"Y{C}"], fx["a"].astype(np.float64)[:, None] * X - D @ X, rtol=1e-12, atol=1e-12)


def test_amg_fixture():
    fx = F.load("amg")
    g = F.graph(fx)
    S = O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, g.K)
    Wl, dl, rl = S.local()
    amg = O.AMG(Wl, S.s[S.U], max_coarse=20)
    assert amg.sizes == fx["sizes"].tolist() and len(amg.sizes) >= 3
    assert np.array_equal(amg.levels[0]["agg"], fx["agg0"])
    np.testing.assert_allclose(amg(fx["X"]), fx["MX"], rtol=1e-12, atol=1e-12)

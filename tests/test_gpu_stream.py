"""Resident edge-stream path vs the FP64 oracle rebuilt from the updated edge set
(pattern of tests/test_incremental.py:18-63 in the reference)."""
import numpy as np
import pytest

from dynlp_amd import graphs
from oracle import lp as O

pytestmark = pytest.mark.gpu


def edge_dict(g):
    return {(int(a), int(b)): float(w) for a, b, w in zip(g.u, g.v, g.w)}


def oracle_system(g, es):
    u = np.array([k[0] for k in es], dtype=np.int64)
    v = np.array([k[1] for k in es], dtype=np.int64)
    w = np.array(list(es.values()))
    return O.build_system(g.n, u, v, w, g.labeled, g.y, g.K), (u, v, w)


def check_against_oracle(lp, g, es, eps, label):
    sysm, (u, v, w) = oracle_system(g, es)
    assert np.array_equal(lp.solved_mask(), sysm.solved), label
    np.testing.assert_allclose(lp.diag64.cpu().numpy(), sysm.diag, rtol=1e-12, err_msg=label)
    Z = lp.scores_with_h()
    R = O.residual(sysm, Z)
    rc = O.rho_cols(sysm.diag, R)
    assert rc[: sysm.C].max() <= eps * 1.0001, (label, rc[: sysm.C].max())
    Fs = O.solve_system(sysm)
    hmax = Fs[:, sysm.C].max()
    err = np.abs(Z[:, : sysm.C] - Fs[:, : sysm.C]).max()
    assert err <= rc[: sysm.C].max() * hmax * 1.05 + 1e-12, (label, err)
    # inert rows parked: F = 0, h = 1
    out = ~sysm.solved
    assert np.all(Z[out, : sysm.C] == 0) and np.all(Z[out, sysm.C] == 1), label
    lab = lp.components()
    ref, _ = O.connected_components(g.n, np.r_[u, v], np.r_[v, u])
    assert np.array_equal(lab, ref), label
    return err, rc[: sysm.C].max()


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("K", [2, 8])
def test_stream_matches_rebuilt_system(fp64, K):
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(3000, 8, 10, 0.8, K, 0.02, seed=1)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps, fp64=fp64)
    lp.initial_solve()
    es = edge_dict(g)
    check_against_oracle(lp, g, es, eps, "initial")
    batches, _ = graphs.update_stream(g, 4, 300, seed=1)
    for t, b in enumerate(batches):
        st = lp.apply_and_solve(b)
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
        assert st.extra.get("missing_deletions", 0) == 0
        check_against_oracle(lp, g, es, eps, f"batch {t}")


def test_fragmenting_stream_regrounds():
    """Sparse SBM with deletions of inter-block edges first: components split off and lose
    (and regain) their labeled contact, so rows toggle between solved and inert."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(4000, 16, 3.0, 0.9, 16, 0.01, seed=2, unit=True)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps)
    lp.initial_solve()
    es = edge_dict(g)
    batches, _ = graphs.update_stream(g, 5, 400, ins_frac=0.3, seed=2, inter_first=True)
    toggles = 0
    for t, b in enumerate(batches):
        st = lp.apply_and_solve(b)
        toggles += st.toggled
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
        check_against_oracle(lp, g, es, eps, f"batch {t}")
    assert toggles > 0


def test_overflow_regrow_keeps_results():
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(2000, 4, 6, 0.8, 4, 0.03, seed=3)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps, slack_min=0, slack_frac=0.0)
    lp.initial_solve()
    es = edge_dict(g)
    batches, _ = graphs.update_stream(g, 3, 400, ins_frac=0.8, seed=3)
    cap0 = int(lp.indptr[-1])
    for t, b in enumerate(batches):
        lp.apply_and_solve(b)
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
        check_against_oracle(lp, g, es, eps, f"batch {t}")
    assert int(lp.indptr[-1]) > cap0


def test_repeat_runs_agree_within_certificate():
    """The asynchronous push is order-dependent (float atomics): two runs may differ in the
    last bits, but both are certified, so they agree within twice the certified bound."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(3000, 8, 10, 0.8, 8, 0.02, seed=4)
    batches, _ = graphs.update_stream(g, 2, 300, seed=4)
    outs = []
    for _ in range(2):
        lp = EdgeStreamLP(g, eps=1e-6)
        lp.initial_solve()
        for b in batches:
            lp.apply_and_solve(b)
        outs.append(lp.scores_with_h())
    es = edge_dict(g)
    for b in batches:
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
    sysm, _ = oracle_system(g, es)
    hmax = O.solve_system(sysm)[:, sysm.C].max()
    assert np.abs(outs[0] - outs[1])[:, : sysm.C].max() <= 2 * 1e-6 * hmax * 1.05


@pytest.mark.parametrize("K", [2, 8])
def test_block_pcg_repair_certifies(K):
    """The block PCG global solve (FP64 Grams on the matrix cores) repairs large batches to the
    same certified tolerance."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(3000, 8, 10, 0.8, K, 0.02, seed=7)
    lp = EdgeStreamLP(g, eps=1e-6, pcg_mode="block", push_max_rows=50)
    st0 = lp.initial_solve()
    es = edge_dict(g)
    check_against_oracle(lp, g, es, 1e-6, "initial")
    batches, _ = graphs.update_stream(g, 3, 400, seed=7)
    for t, b in enumerate(batches):
        st = lp.apply_and_solve(b)
        assert "pcg" in st.path
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
        check_against_oracle(lp, g, es, 1e-6, f"batch {t}")

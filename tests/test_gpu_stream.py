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


def check_against_oracle(lp, g, es, eps, label, iterative=False):
    """Solved set, diagonal, certificate and CC labels against the oracle; F* by the oracle's
    direct solve, or (iterative: graphs without community structure, whose factors fill in)
    by its CG to rho <= 1e-13."""
    sysm, (u, v, w) = oracle_system(g, es)
    assert np.array_equal(lp.solved_mask(), sysm.solved), label
    np.testing.assert_allclose(lp.diag64.cpu().numpy(), sysm.diag, rtol=1e-12, err_msg=label)
    Z = lp.scores_with_h()
    R = O.residual(sysm, Z)
    rc = O.rho_cols(sysm.diag, R)
    assert rc[: sysm.C].max() <= eps * 1.0001, (label, rc[: sysm.C].max())
    if iterative:
        Wl, dl, rl = sysm.local()
        cols = np.concatenate([rl, dl[:, None]], axis=1)
        Xs, _ = O.pcg(Wl, dl, np.zeros_like(cols), cols, [1e-13] * cols.shape[1])
        Fs = np.zeros_like(Z)
        Fs[:, sysm.C] = 1.0
        Fs[sysm.U] = Xs
    else:
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


@pytest.mark.parametrize("fp64,dirs", [(False, "fp32"), (True, "fp32"), (False, "bf16"), (False, "fp16"),
                                      (False, "fp8")])
@pytest.mark.parametrize("K", [2, 8])
def test_stream_matches_rebuilt_system(fp64, dirs, K):
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(3000, 8, 10, 0.8, K, 0.02, seed=1)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps, fp64=fp64, pcg_directions=dirs)
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


@pytest.mark.parametrize("pack,xcd,vec4,pipe", [(0, 0, 1, 0), (0, 1, 1, 0), (1, 0, 1, 4), (1, 1, 1, 6),
                                               (1, 1, 0, 8)])
def test_stream_lane_layouts_and_row_orders(pack, xcd, vec4, pipe):
    """The FP32 service path (refresh, global-solve operator, PCG passes) under the power-of-two
    lane layout, the plain row order, scalar loads, the plain fused operator and every depth of
    the pipelined one; the defaults run in the tests above."""
    from dynlp_amd import _lib
    from dynlp_amd.stream import EdgeStreamLP
    _lib.call("dlp_tune_set", 1, vec4)
    _lib.call("dlp_tune_set", 2, pack)
    _lib.call("dlp_tune_set", 3, xcd)
    _lib.call("dlp_tune_set", 6, pipe)
    try:
        g = graphs.sbm(3000, 8, 10, 0.8, 8, 0.02, seed=4)
        lp = EdgeStreamLP(g, eps=1e-6)
        lp.initial_solve()
        es = edge_dict(g)
        batches, _ = graphs.update_stream(g, 3, 300, seed=4)
        for t, b in enumerate(batches):
            lp.apply_and_solve(b)
            O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                               zip(b.del_u.tolist(), b.del_v.tolist()))
        st = lp.load_state(np.zeros((g.n, lp.C1), np.float32))   # a global solve from zero too
        assert "pcg" in st.path, st.path
        check_against_oracle(lp, g, es, 1e-6, f"pack={pack} xcd={xcd} vec4={vec4} pipe={pipe}")
    finally:
        _lib.call("dlp_tune_set", 6, 4)
        _lib.call("dlp_tune_set", 1, 1)
        _lib.call("dlp_tune_set", 2, 1)
        _lib.call("dlp_tune_set", 3, 1)


@pytest.mark.parametrize("dirs,word", [("fp32", 4), ("bf16", 4), ("bf16", 8), ("fp16", 4), ("fp8", 4)])
def test_global_solve_from_zero_certifies(dirs, word):
    """A global solve from a zero state (every row violating) under FP32, BF16 (8- and 16-byte
    lane words), scaled FP16 and scaled FP8 directions: certified against the rebuilt oracle, on
    a fragmenting graph with inert rows."""
    from dynlp_amd import _lib
    from dynlp_amd.stream import EdgeStreamLP
    _lib.call("dlp_tune_set", 7, word)
    g = graphs.sbm(4000, 16, 3.0, 0.9, 16, 0.01, seed=5, unit=True)
    lp = EdgeStreamLP(g, eps=1e-6, pcg_directions=dirs)
    lp.initial_solve()
    es = edge_dict(g)
    batches, _ = graphs.update_stream(g, 2, 400, ins_frac=0.3, seed=5, inter_first=True)
    for b in batches:
        lp.apply_and_solve(b)
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
    try:
        st = lp.load_state(np.zeros((g.n, lp.C1), np.float32))
    finally:
        _lib.call("dlp_tune_set", 7, 4)
    assert "pcg" in st.path
    check_against_oracle(lp, g, es, 1e-6, f"{dirs}/{word}")


@pytest.mark.parametrize("dirs", ["fp32", "bf16", "fp16", "fp8"])
def test_two_level_preconditioner_certifies_and_cuts_iterations(dirs):
    """The coarse correction (runs of consecutive rows as aggregates, Ac = P^T A P factored and
    inverted on the device) certifies against the rebuilt oracle like the Jacobi-preconditioned
    solve, whether Ac is rebuilt for every solve or kept from the first, and needs fewer PCG
    iterations.  Push is switched off so every batch takes the global solve."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(3000, 8, 10, 0.8, 8, 0.01, seed=11)
    batches, _ = graphs.update_stream(g, 3, 200, seed=12)
    iters = {}
    for na, refresh in ((0, 64), (128, 1), (128, 64), (20, 64)):
        lp = EdgeStreamLP(g, eps=1e-6, pcg_directions=dirs, pcg_coarse=na, coarse_refresh=refresh,
                          push_max_rows=0)
        lp.initial_solve()
        es = edge_dict(g)
        s0 = lp.pcg_stats()
        for b in batches:
            st = lp.apply_and_solve(b)
            assert "pcg" in st.path, st.path
            O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                               zip(b.del_u.tolist(), b.del_v.tolist()))
        iters[(na, refresh)] = lp.pcg_stats()["pcg_iters"] - s0["pcg_iters"]
        check_against_oracle(lp, g, es, 1e-6, f"{dirs} coarse={na} refresh={refresh}")
    assert iters[(128, 1)] < iters[(0, 64)] and iters[(128, 64)] < iters[(0, 64)], iters


@pytest.mark.parametrize("target", [6, 12])
def test_coarse_boundaries_follow_communities(target):
    """Aggregate boundaries move to the row boundaries with the fewest crossing edges and stay
    where they cut fewer edges than half a spacing away: on an SBM whose ids follow its blocks
    (1000-row blocks) they are the block edges whether 6 or 12 are asked for, and the solve
    certifies against the oracle."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(6000, 6, 12, 0.9, 4, 0.01, seed=21)
    lp = EdgeStreamLP(g, eps=1e-6, pcg_coarse=target, push_max_rows=0)
    lp.initial_solve()
    info = lp.coarse_info()
    assert info["aggregates"] == 6, info
    assert np.array_equal(info["bounds"], np.arange(0, 6001, 1000)), info["bounds"]
    check_against_oracle(lp, g, edge_dict(g), 1e-6, "communities")


def test_coarse_boundaries_without_communities():
    """On an Erdos-Renyi graph (no structure in the ids) the evenly spaced boundaries are kept
    and the two-level solve still certifies."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.erdos_renyi(4000, 10, 4, 0.02, seed=22)
    lp = EdgeStreamLP(g, eps=1e-6, pcg_coarse=16, push_max_rows=0)
    lp.initial_solve()
    info = lp.coarse_info()
    assert 4 <= info["aggregates"] <= 16, info
    assert np.all(np.diff(info["bounds"]) >= info["min_rows"]), info
    check_against_oracle(lp, g, edge_dict(g), 1e-6, "er")


def test_tiny_graph_solves_without_coarse_layout():
    """Below 64 vertices the global solve runs with the Jacobi preconditioner alone."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(48, 2, 6, 0.8, 2, 0.2, seed=3)
    lp = EdgeStreamLP(g, eps=1e-6, push_max_rows=0)
    lp.initial_solve()
    assert lp.coarse_info()["aggregates"] == 0
    es = edge_dict(g)
    batches, _ = graphs.update_stream(g, 2, 6, seed=4)
    for b in batches:
        lp.apply_and_solve(b)
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
    check_against_oracle(lp, g, es, 1e-6, "tiny")


def test_pcg_group_length_does_not_change_the_solve():
    """Iterations past convergence inside a captured group return at once (device flag): group
    lengths 1 and 7 run the same iterations and give the same state."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(3000, 8, 10, 0.8, 8, 0.02, seed=6)
    out = []
    for per in (1, 7):
        lp = EdgeStreamLP(g, eps=1e-6, pcg_iters_per_launch=per)
        lp.initial_solve()
        before = lp.pcg_stats()["pcg_iters"]
        st = lp.load_state(np.zeros((g.n, lp.C1), np.float32))
        assert "pcg" in st.path
        out.append((lp.pcg_stats()["pcg_iters"] - before, lp.scores_with_h()))
    (i1, Z1), (i7, Z7) = out
    assert abs(i1 - i7) <= 1 and i1 > 1, (i1, i7)
    # both certified to rho <= 1e-6: each within rho * hmax of the exact solution
    C = Z1.shape[1] - 1
    hmax = max(Z1[:, C].max(), Z7[:, C].max())
    np.testing.assert_allclose(Z1[:, :C], Z7[:, :C], rtol=0, atol=2.1e-6 * hmax)


@pytest.mark.parametrize("dirs", ["fp16", "bf16", "fp8"])
@pytest.mark.parametrize("per", [1, 4, 8])
def test_deferred_solution_updates_match_per_iteration_updates(dirs, per):
    """Deferred solution updates (Z += sum_s alpha_s P_s once per captured group of iterations,
    tuning key 10) make the same FP32 multiply-adds, in the same order, as the update inside
    every iteration, for groups of 1, 4 and 8 iterations.  Two runs differ only through the
    order of the FP64 column sums (atomics across workgroups), so the check is: the same
    iteration counts (to one) over an initial solve and three global repairs, both states
    certified, and each within the certificate of the other."""
    from dynlp_amd import _lib
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(4000, 8, 10, 0.8, 8, 0.02, seed=12)
    batches, _ = graphs.update_stream(g, 3, 400, seed=12)
    res = []
    try:
        for defer in (0, 1):
            _lib.call("dlp_tune_set", 10, defer)
            lp = EdgeStreamLP(g, eps=1e-6, pcg_directions=dirs, pcg_iters_per_launch=per, push_max_rows=0)
            lp.initial_solve()
            its = []
            for b in batches:
                st = lp.apply_and_solve(b)
                assert "pcg" in st.path and st.rho <= 1e-6 * 1.0001
                its.append(st.extra.get("pcg_iters", 0))
            res.append((its, lp.scores_with_h()))
            del lp
    finally:
        _lib.call("dlp_tune_set", 10, 1)
    (i0, Z0), (i1, Z1) = res
    assert all(abs(a - b) <= 1 for a, b in zip(i0, i1)) and min(i0) > 0, (i0, i1)
    C = Z0.shape[1] - 1
    hmax = max(Z0[:, C].max(), Z1[:, C].max())
    np.testing.assert_allclose(Z0[:, :C], Z1[:, :C], rtol=0, atol=2.1e-6 * hmax)


def test_launches_without_graphs_match_replayed_graphs():
    """use_graph=False launches the push rounds and the groups of global-solve iterations (with
    their deferred solution updates) one kernel at a time: the same paths, the same iteration
    counts (to one) and both states within each other's certificate, over small requests (push
    rounds first) and large ones (global solve only)."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(30000, 8, 10, 0.8, 8, 0.02, seed=13)
    batches, _ = graphs.update_stream(g, 3, 3000, seed=13)
    small, _ = graphs.update_stream(g, 2, 4, seed=14)
    res = []
    for use_graph in (True, False):
        # no mass limit: the small requests repair by push, the large ones exceed the row limit
        lp = EdgeStreamLP(g, eps=1e-6, use_graph=use_graph, pcg_iters_per_launch=4, push_mass=1e12)
        lp.initial_solve()
        its, paths = [], []
        for b in small + batches:
            st = lp.apply_and_solve(b)
            assert st.rho <= 1e-6 * 1.0001 and np.isfinite(st.cert)
            its.append(st.extra.get("pcg_iters", 0))
            paths.append(st.path)
        res.append((its, paths, lp.scores_with_h()))
        del lp
    (i0, p0, Z0), (i1, p1, Z1) = res
    assert p0 == p1 and any(p.startswith("push") for p in p0) and "pcg" in p0, (p0, p1)
    assert all(abs(a - b) <= 1 for a, b in zip(i0, i1)), (i0, i1)
    C = Z0.shape[1] - 1
    hmax = max(Z0[:, C].max(), Z1[:, C].max())
    np.testing.assert_allclose(Z0[:, :C], Z1[:, :C], rtol=0, atol=2.1e-6 * hmax)


@pytest.mark.parametrize("K", [2, 8])
def test_error_report_against_oracle(K):
    """error_report: relative residual equal to the oracle's on the same state, certified bound
    above the true max |F - F*|, and certified rows predict like F*."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(3000, 8, 10, 0.8, K, 0.02, seed=9)
    lp = EdgeStreamLP(g, eps=1e-6)
    lp.initial_solve()
    es = edge_dict(g)
    b = graphs.update_stream(g, 1, 300, seed=9)[0][0]
    st = lp.apply_and_solve(b)
    O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                       zip(b.del_u.tolist(), b.del_v.tolist()))
    e = lp.error_report()
    # the request's own certificate is the report's (column maxima of the FP64 verification
    # against those of its FP32-kept copy)
    assert st.hmax == pytest.approx(e["hmax"], rel=1e-6)
    assert st.cert == pytest.approx(e["cert_bound"], rel=1e-6)
    assert st.rho == pytest.approx(e["rho"], rel=1e-6)
    sysm, _ = oracle_system(g, es)
    Z = lp.scores_with_h()
    U, C = sysm.solved, sysm.C
    R = O.residual(sysm, Z)
    rel = np.linalg.norm(R[U, :C]) / np.linalg.norm(sysm.rhs[U])
    assert e["rel_l2_residual"] == pytest.approx(rel, rel=1e-3, abs=1e-12)
    Fs = O.solve_system(sysm)
    assert np.abs(Z[U, :C] - Fs[U, :C]).max() <= e["cert_bound"]
    assert e["cert_bound"] <= 1e-6 * Fs[:, C].max() * 1.1
    pz, ps = O.predict(Z[U, :C], K), O.predict(Fs[U, :C], K)
    assert (pz == ps).mean() >= e["argmax_certified_frac"] and e["argmax_certified_frac"] > 0.9


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


def test_local_grounding_checks_match_oracle():
    """Deletion-only batches small enough for the local grounding checks (2 * deletions <=
    n/16): the bounded searches from the deletion endpoints decide the solved set, and when one
    finds no labeled contact the full pass runs; the solved set and solution match the oracle
    after every batch, and the local checks alone decide some batches.  Then a vertex is cut off
    (all its edges deleted): it leaves the solved set."""
    from dynlp_amd.graphs import EdgeBatch
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.erdos_renyi(20000, 6, 4, 0.01, seed=8)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps)
    lp.initial_solve()
    es = edge_dict(g)
    batches, _ = graphs.update_stream(g, 5, 120, ins_frac=0.0, seed=8)
    local_only = 0
    for t, b in enumerate(batches):
        c0 = lp.ground_flags()["calls"]
        lp.apply_and_solve(b)
        f = lp.ground_flags()
        local_only += int(f["calls"] > c0 and not f["full_pass"])
        O.apply_edge_batch(es, [], zip(b.del_u.tolist(), b.del_v.tolist()))
        check_against_oracle(lp, g, es, eps, f"batch {t}", iterative=True)
    assert local_only >= 1
    solved = lp.solved_mask()
    deg = np.bincount(np.array([k[0] for k in es] + [k[1] for k in es]), minlength=g.n)
    x = int(np.flatnonzero(solved & ~g.labeled & (deg >= 2))[0])
    cut = [k for k in es if x in k]
    i32 = lambda z: np.asarray(z, dtype=np.int32)  # noqa: E731
    b = EdgeBatch(i32([]), i32([]), np.zeros(0, np.float32), i32([a for a, _ in cut]), i32([c for _, c in cut]))
    lp.apply_and_solve(b)
    O.apply_edge_batch(es, [], cut)
    check_against_oracle(lp, g, es, eps, "vertex cut off", iterative=True)
    assert not lp.solved_mask()[x]


def test_verify_pass_matches_full_refresh():
    """The verification pass (refresh mode 3, no per-entry kind look-up, parked rows summed
    unmasked) gives the residual, violators and column maxima of the full refresh (mode 2), on a
    fragmenting graph with inert and labeled rows; parked rows report F = 0, h = 1."""
    from dynlp_amd import _lib
    from dynlp_amd.backend import stream
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(4000, 16, 3.0, 0.9, 16, 0.01, seed=3, unit=True)
    lp = EdgeStreamLP(g, eps=1e-6)
    lp.initial_solve()
    batches, _ = graphs.update_stream(g, 3, 400, ins_frac=0.3, seed=3, inter_first=True)
    for t, b in enumerate(batches):
        lp.apply_and_solve(b)
        # perturb the solved rows so the residual is far from zero, then compare the two passes
        import torch
        sm = lp.solved.bool()
        noise = torch.as_tensor(np.random.default_rng(t).normal(0, 1e-3, (g.n, lp.C1)), dtype=torch.float32,
                                device=lp.Z32.device)
        lp.Z32[:, : lp.C1] += noise * sm[:, None]
        out = []
        for mode in (3, 2):
            _lib.call("dlp_lp_refresh", lp.h, mode, 0, stream())
            s = lp.status()
            out.append((lp.R32[:, : lp.C1].double().cpu().numpy(), lp._rho.copy(), s["viol"]))
        (R3, rho3, v3), (R2, rho2, v2) = out
        solved = lp.solved_mask()
        assert (~solved).any()
        np.testing.assert_allclose(R3, R2, rtol=1e-6, atol=1e-12)
        np.testing.assert_allclose(rho3, rho2, rtol=1e-6)
        assert v3 == v2 and v3 > 0
        Zs = lp.scores_with_h()
        assert np.all(Zs[~solved, : lp.C] == 0) and np.all(Zs[~solved, lp.C] == 1)


def test_batch_edge_cases_match_oracle():
    """Requests at the edges of the batch contract, each against the oracle's edge set: an empty
    batch, deletions only, deletions of absent edges (counted missing, no change), insertions of
    present edges (weight replaced), self-loops (ignored), a labeled vertex losing and regaining
    every edge; out-of-range ids and ragged arrays are refused on the host before any device
    work, and the state stays usable."""
    from dynlp_amd.graphs import EdgeBatch
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(2000, 4, 8, 0.8, 4, 0.05, seed=4)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps)
    lp.initial_solve()
    es = edge_dict(g)
    i32, f32 = (lambda x: np.asarray(x, dtype=np.int32)), (lambda x: np.asarray(x, dtype=np.float32))
    keys = sorted(es)
    absent = [(a, b) for a, b in ((0, 1999), (3, 1500), (7, 1234)) if (a, b) not in es]
    lab = int(np.flatnonzero(g.labeled)[0])
    nbrs = [(min(a, b), max(a, b)) for a, b in keys if lab in (a, b)]
    cases = [
        ("empty", [], [], [], [], []),
        ("deletions only", [], [], [], [k[0] for k in keys[:40]], [k[1] for k in keys[:40]]),
        ("absent deletions", [], [], [], [a for a, _ in absent], [b for _, b in absent]),
        ("reweight present", [k[0] for k in keys[100:140]], [k[1] for k in keys[100:140]],
         np.linspace(0.5, 3.0, 40), [], []),
        ("self-loops", [5, 6, 10], [5, 6, 11], [2.0, 2.0, 1.5], [9], [9]),
        ("labeled vertex isolated", [], [], [], [a for a, _ in nbrs], [b for _, b in nbrs]),
        ("labeled vertex back", [a for a, _ in nbrs], [b for _, b in nbrs], np.full(len(nbrs), 0.7), [], []),
    ]
    for name, iu, iv, iw, du, dv in cases:
        b = EdgeBatch(i32(iu), i32(iv), f32(iw), i32(du), i32(dv))
        absent_now = sum((min(a, c), max(a, c)) not in es for a, c in zip(du, dv) if a != c)
        st = lp.apply_and_solve(b)
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
        # both directions of every deletion that finds no stored edge are counted (self-loops
        # are skipped, never stored)
        assert st.extra.get("missing_deletions", 0) == 2 * absent_now, name
        assert st.rho <= eps * 1.0001, name
        check_against_oracle(lp, g, es, eps, name)
    for bad in (EdgeBatch(i32([0]), i32([g.n]), f32([1.0]), i32([]), i32([])),
                EdgeBatch(i32([]), i32([]), f32([]), i32([-1]), i32([3])),
                EdgeBatch(i32([1, 2]), i32([3]), f32([1.0]), i32([]), i32([]))):
        with pytest.raises((IndexError, ValueError)):
            lp.apply_and_solve(bad)
    b = graphs.update_stream(g, 1, 200, seed=4)[0][0]
    lp.apply_and_solve(b)
    O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                       zip(b.del_u.tolist(), b.del_v.tolist()))
    check_against_oracle(lp, g, es, eps, "after refused batches")


@pytest.mark.parametrize("push_max_rows", [None, 0])
def test_overflow_regrow_keeps_results(push_max_rows):
    """Rows without slack overflow on insertion and the store is regrown before the repair, for
    batches a push may repair (status read after the apply) and for large batches that go
    straight to the global solve (push_max_rows 0: settled in the repairing step)."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(2000, 4, 6, 0.8, 4, 0.03, seed=3)
    eps = 1e-6
    kw = {} if push_max_rows is None else dict(push_max_rows=push_max_rows)
    lp = EdgeStreamLP(g, eps=eps, slack_min=0, slack_frac=0.0, **kw)
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


def _stored_entries(lp):
    ip = lp.indptr.cpu().numpy()
    rl = lp.rowlen.cpu().numpy()
    col = lp.col.cpu().numpy()
    w = lp.w.cpu().numpy()
    out = {}
    for r in range(lp.n):
        cs = col[ip[r]: ip[r] + rl[r]]
        assert np.unique(cs).shape == cs.shape, f"row {r} stores a column twice"
        assert np.all(cs >= 0), f"row {r} keeps a tombstone"
        for c, x in zip(cs, w[ip[r]: ip[r] + rl[r]]):
            out[(r, int(c))] = float(x)
    return out


def _check_store(lp, g, es):
    """The device store equals the oracle edge set: both directions, one entry each, same
    weight; labeled-neighbour counts match."""
    st = _stored_entries(lp)
    want = {}
    for (a, b), x in es.items():
        want[(a, b)] = want[(b, a)] = float(np.float32(x))
    assert st == want
    lab = np.zeros(g.n, dtype=np.int64)
    for (a, b) in want:
        lab[a] += int(g.labeled[b])
    assert np.array_equal(lp.labcnt.cpu().numpy(), lab)


def test_repeated_pairs_in_a_batch_match_oracle():
    """Pairs repeated in one request, in both orientations, within the insertions (new and
    stored pairs) and within the deletions: the result equals the oracle's sequential
    application (deletions first, last insertion wins)."""
    from dynlp_amd.graphs import EdgeBatch
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(1500, 4, 8, 0.8, 4, 0.05, seed=8)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps)
    lp.initial_solve()
    es = edge_dict(g)
    keys = sorted(es)
    new = [(a, b) for a, b in ((0, 1400), (3, 1100), (7, 900), (11, 1300)) if (a, b) not in es]
    iu = [new[0][0], new[0][1], new[1][0], keys[5][1], keys[5][0], new[1][1], new[0][0]]
    iv = [new[0][1], new[0][0], new[1][1], keys[5][0], keys[5][1], new[1][0], new[0][1]]
    iw = [float(np.float32(x)) for x in (0.6, 1.2, 0.9, 2.0, 0.7, 1.4, 0.8)]   # the stored FP32 weights
    du = [keys[10][0], keys[10][1], keys[10][0], keys[20][1], keys[20][0]]
    dv = [keys[10][1], keys[10][0], keys[10][1], keys[20][0], keys[20][1]]
    i32 = lambda x: np.asarray(x, dtype=np.int32)  # noqa: E731
    b = EdgeBatch(i32(iu), i32(iv), np.asarray(iw, dtype=np.float32), i32(du), i32(dv))
    st = lp.apply_and_solve(b)
    O.apply_edge_batch(es, zip(iu, iv, iw), zip(du, dv))
    assert st.extra.get("missing_deletions", 0) == 0
    _check_store(lp, g, es)
    check_against_oracle(lp, g, es, eps, "repeated pairs")


def test_device_apply_refuses_bad_pairs_and_keeps_last_occurrence():
    """Straight through the C ABI, with no host check in front: ids outside [0, n) and
    insertion weights that are not finite and positive are skipped and counted (the kernels
    test every id before indexing a row by it), a repeated deletion removes the pair once,
    and of a pair inserted several times, in either orientation, only the last occurrence is
    stored, in both directions.  The state then repairs to the oracle's solution."""
    import torch
    from dynlp_amd import _lib
    from dynlp_amd.backend import device, ptr, stream
    from dynlp_amd.stream import EdgeStreamLP, Stats
    g = graphs.sbm(1500, 4, 8, 0.8, 4, 0.05, seed=9)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps)
    lp.initial_solve()
    es = edge_dict(g)
    keys = sorted(es)
    n = g.n
    new = [(a, b) for a, b in ((2, 1450), (5, 1200), (9, 1000)) if (a, b) not in es]
    nan, inf = float("nan"), float("inf")
    ins = [(new[0][0], new[0][1], 0.5), (new[0][1], new[0][0], 1.5), (new[1][0], new[1][1], 0.75),
           (keys[3][1], keys[3][0], 2.5), (keys[3][0], keys[3][1], 0.25),
           (n, 4, 1.0), (4, -1, 1.0), (1 << 30, 7, 1.0),                 # ids out of range
           (new[2][0], new[2][1], nan), (6, 8, inf), (6, 9, 0.0), (6, 10, -2.0)]   # bad weights
    dels = [(keys[30][0], keys[30][1]), (keys[30][1], keys[30][0]), (n + 5, 3), (-7, 2)]
    dev = device()
    t32 = lambda x: torch.as_tensor(np.asarray(x, dtype=np.int32)).to(dev)  # noqa: E731
    iu, iv = t32([a for a, _, _ in ins]), t32([b for _, b, _ in ins])
    iw = torch.as_tensor(np.asarray([x for _, _, x in ins], dtype=np.float32)).to(dev)
    du, dv = t32([a for a, _ in dels]), t32([b for _, b in dels])
    _lib.call("dlp_lp_begin_batch", lp.h, stream())
    _lib.call("dlp_lp_apply", lp.h, len(dels), ptr(du), ptr(dv), len(ins), ptr(iu), ptr(iv), ptr(iw), stream())
    s = lp.status()
    assert s["invalid"] == 2 + 3 + 4, s      # two deletions, three insertion ids, four weights
    assert s["duplicates"] == 2, s           # the earlier occurrences of new[0] and keys[3]
    assert s["missing"] == 2, s              # the repeat of keys[30], both directions
    O.apply_edge_batch(es, [(a, b, x) for a, b, x in ins[:5]], dels[:2])
    _check_store(lp, g, es)
    _lib.call("dlp_lp_ground", lp.h, stream())
    _lib.call("dlp_lp_refresh", lp.h, 0, 0, stream())
    st = Stats()
    s = lp._repair(st)
    lp._finish(st, s)
    assert st.rho <= eps * 1.0001
    check_against_oracle(lp, g, es, eps, "after the raw apply")


@pytest.mark.parametrize("kw", [dict(), dict(pcg_directions="fp32", pcg_coarse=0), dict(fp64=True),
                                dict(pcg_mode="block"), dict(pcg_iters_per_launch=3), dict(pcg_directions="fp8")])
def test_footprint_is_what_the_graph_holds(kw):
    """The admission estimate equals what the handle holds once built (its own count) and,
    within allocation granularity, the device memory a resident graph takes; a request
    afterwards allocates nothing in the library (the handle's bytes do not change)."""
    import torch
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(20000, 8, 10, 0.8, 8, 0.02, seed=2)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    lp = EdgeStreamLP(g, eps=1e-6, **kw)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    est = EdgeStreamLP.footprint(g, **kw)
    from dynlp_amd import _lib
    n, C1, ld = lp.n, lp.C1, lp.ld
    from dynlp_amd.stream import DIRECTION_KINDS as kinds
    hb = _lib.load().dlp_lp_handle_bytes(n, C1, ld, int(lp.fp64), kinds[lp.pcg_directions],
                                         lp.pcg_iters_per_launch, lp.pcg_coarse, int(lp.pcg_mode == "block"))
    assert lp.device_bytes() == hb
    used = free0 - free1
    # torch's caching allocator rounds its blocks up (2 MiB segments) and hipMalloc to pages
    assert abs(used - est) <= 0.05 * est + (64 << 20), (used, est)
    lp.initial_solve()
    b = graphs.update_stream(g, 1, 2000, seed=2)[0][0]
    lp.apply_and_solve(b)
    assert lp.device_bytes() == hb


@pytest.mark.parametrize("K", [2, 8])
def test_fp64_pcg_repair_matches_rebuilt_system(K):
    """The plain FP64 path with its global repair by FP64 column PCG (no graphs, Jacobi
    preconditioner) instead of push rounds: every batch against the oracle."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(3000, 8, 10, 0.8, K, 0.02, seed=5)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps, fp64=True, fp64_repair="pcg")
    lp.initial_solve()
    es = edge_dict(g)
    check_against_oracle(lp, g, es, eps, "initial")
    batches, _ = graphs.update_stream(g, 3, 400, seed=5)
    for t, b in enumerate(batches):
        st = lp.apply_and_solve(b)
        assert st.path == "fp64-pcg" and st.extra["pcg_iters"] > 0
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
        check_against_oracle(lp, g, es, eps, f"batch {t}")


def _conflicting_batches(g, es, rng):
    """Batches whose operations conflict across and within batches: new pairs inserted and later
    deleted, present edges deleted and later re-inserted with new weights, a pair inserted in
    both orientations, a deletion repeated, a labeled vertex's edges removed and restored."""
    from dynlp_amd.graphs import EdgeBatch
    i32, f32 = (lambda x: np.asarray(x, dtype=np.int32)), (lambda x: np.asarray(x, dtype=np.float32))
    keys = sorted(es)
    new = []
    while len(new) < 60:
        a, b = sorted(rng.integers(0, g.n, 2).tolist())
        if a != b and (a, b) not in es and (a, b) not in new:
            new.append((a, b))
    lab = int(np.flatnonzero(g.labeled)[0])
    nbrs = [k for k in keys if lab in k]
    A = EdgeBatch(i32([a for a, _ in new]), i32([b for _, b in new]), f32(rng.uniform(0.5, 2, 60)),
                  i32([a for a, _ in keys[:30]]), i32([b for _, b in keys[:30]]))
    B = EdgeBatch(i32([b for _, b in keys[:15]] + [new[0][1]]), i32([a for a, _ in keys[:15]] + [new[0][0]]),
                  f32(list(rng.uniform(0.5, 2, 15)) + [3.0]),
                  i32([a for a, _ in new[:20]] + [new[5][1]]), i32([b for _, b in new[:20]] + [new[5][0]]))
    C = EdgeBatch(i32([a for a, _ in new[10:20]] + [new[1][0], new[1][1]]),
                  i32([b for _, b in new[10:20]] + [new[1][1], new[1][0]]),
                  f32(list(rng.uniform(0.5, 2, 10)) + [1.25, 2.5]),
                  i32([a for a, _ in nbrs] + [new[30][0]]), i32([b for _, b in nbrs] + [new[30][1]]))
    D = EdgeBatch(i32([a for a, _ in nbrs[:2]]), i32([b for _, b in nbrs[:2]]), f32([0.9, 1.1]),
                  i32([new[40][0], new[40][0]]), i32([new[40][1], new[40][1]]))
    return [A, B, C, D]


@pytest.mark.parametrize("ops_per_call", [None, 37])
def test_request_group_matches_sequential_oracle(monkeypatch, ops_per_call):
    """A group of conflicting batches applied by dlp_lp_apply_ops (one ordered sequence, or
    calls of 37 operations) and repaired once: the device edge set equals the oracle's after
    applying the batches one after the other (weights bit-exact), and the solution, solved set
    and CC labels match the oracle.  With small calls and no slack the rows overflow and are
    regrown between calls."""
    from dynlp_amd import stream as S
    from dynlp_amd.stream import EdgeStreamLP
    if ops_per_call:
        monkeypatch.setattr(S, "_OPS_PER_CALL", ops_per_call)
    g = graphs.sbm(3000, 4, 8, 0.8, 4, 0.03, seed=12)
    eps = 1e-6
    lp = EdgeStreamLP(g, eps=eps, slack_min=0 if ops_per_call else 4, slack_frac=0.0 if ops_per_call else 0.25)
    lp.initial_solve()
    cap0 = lp.col.numel()
    es = edge_dict(g)
    rng = np.random.default_rng(12)
    for rep in range(2):
        bs = _conflicting_batches(g, es, rng)
        st = lp.apply_group_and_solve(bs)
        for b in bs:
            O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                               zip(b.del_u.tolist(), b.del_v.tolist()))
        assert st.rho <= eps * 1.0001
        ge = lp.edge_graph()
        got = {(int(a), int(b)): float(w) for a, b, w in zip(ge.u, ge.v, ge.w)}
        assert got == {k: float(np.float32(v)) for k, v in es.items()}, rep
        check_against_oracle(lp, g, es, eps, f"group {rep}")
    if ops_per_call:
        assert lp.col.numel() > cap0      # the store was regrown


def test_fp8_directions_take_about_the_fp16_iterations():
    """FP8 (E4M3) directions scaled from a bound on each column's largest value keep the
    iteration count of FP16 directions (the step length is exact for the stored direction, so
    coarser directions only cost conjugacy), and the repaired states are certified, over an
    initial solve from zero and three global repairs on an SBM with 64 label columns (the FP8
    state checked against the oracle's CG solution: the direct solve of this graph fills in)."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(20000, 16, 12, 0.8, 64, 0.01, seed=21)
    batches, _ = graphs.update_stream(g, 3, 1000, seed=21)
    its = {}
    for dirs in ("fp16", "fp8"):
        lp = EdgeStreamLP(g, eps=1e-6, pcg_directions=dirs, push_max_rows=0)
        st0 = lp.initial_solve()
        assert st0.rho <= 1e-6 * 1.0001
        es = edge_dict(g)
        s0 = lp.pcg_stats()["pcg_iters"]
        for b in batches:
            st = lp.apply_and_solve(b)
            assert "pcg" in st.path and st.rho <= 1e-6 * 1.0001, (dirs, st.path, st.rho)
            O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                               zip(b.del_u.tolist(), b.del_v.tolist()))
        its[dirs] = lp.pcg_stats()["pcg_iters"] - s0
        if dirs == "fp8":
            check_against_oracle(lp, g, es, 1e-6, dirs, iterative=True)
        del lp
    assert its["fp8"] <= 1.25 * its["fp16"] + 3, its


@pytest.mark.parametrize("dirs", ["fp16", "fp8"])
def test_low_precision_directions_certify_with_weights_over_six_decades(dirs):
    """Edge weights spread over six decades give direction values of very different sizes within
    a column, the case where a few significant bits and FP8's narrow range (values below 2^-9 of
    the column's bound become zero) cost most: the repairs still end certified against the
    rebuilt oracle."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(4000, 8, 10, 0.8, 8, 0.02, seed=31, wlo=1e-3, whi=1e3)
    lp = EdgeStreamLP(g, eps=1e-6, pcg_directions=dirs, push_max_rows=0)
    st = lp.initial_solve()
    assert st.rho <= 1e-6 * 1.0001
    es = edge_dict(g)
    for b in graphs.update_stream(g, 3, 400, seed=31)[0]:
        st = lp.apply_and_solve(b)
        assert "pcg" in st.path and st.rho <= 1e-6 * 1.0001, (st.path, st.rho)
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
    check_against_oracle(lp, g, es, 1e-6, dirs)


@pytest.mark.parametrize("depth,waves", [(4, 6), (6, 5), (6, 6), (8, 5)])
def test_fp8_operator_depths_certify(depth, waves):
    """The FP8 operator at every depth (gathers per lane in flight) and register budget it is
    built for (the default, depth by row width, runs in the other FP8 tests): a global
    solve from zero and two repairs, certified against the rebuilt oracle."""
    from dynlp_amd import _lib
    from dynlp_amd.stream import EdgeStreamLP
    _lib.call("dlp_tune_set", 12, depth)
    _lib.call("dlp_tune_set", 11, waves)
    try:
        g = graphs.sbm(3000, 8, 10, 0.8, 8, 0.02, seed=41)
        lp = EdgeStreamLP(g, eps=1e-6, pcg_directions="fp8", push_max_rows=0)
        lp.initial_solve()
        es = edge_dict(g)
        for b in graphs.update_stream(g, 2, 300, seed=41)[0]:
            st = lp.apply_and_solve(b)
            assert "pcg" in st.path and st.rho <= 1e-6 * 1.0001
            O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                               zip(b.del_u.tolist(), b.del_v.tolist()))
        st = lp.load_state(np.zeros((g.n, lp.C1), np.float32))
        assert "pcg" in st.path
        check_against_oracle(lp, g, es, 1e-6, f"depth {depth} waves {waves}")
    finally:
        _lib.call("dlp_tune_set", 12, 0)
        _lib.call("dlp_tune_set", 11, 6)


def test_solve_with_enqueued_verification_matches_separate_verification():
    """The global solve with the FP64 verification enqueued behind its launches
    (dlp_lp_pcg_verify, the service's path) against the solve followed by a separate
    verification call: the same paths and iteration counts (to one), both certified, each state
    within the other's certificate, and the enqueued one against the rebuilt oracle.  One case
    with launches of 2 iterations, so that verification also runs behind launches that do not
    end the solve."""
    from dynlp_amd.stream import EdgeStreamLP
    g = graphs.sbm(4000, 8, 10, 0.8, 8, 0.02, seed=51)
    batches, _ = graphs.update_stream(g, 3, 400, seed=51)
    for per in (7, 2):
        res = []
        for fused in (False, True):
            lp = EdgeStreamLP(g, eps=1e-6, push_max_rows=0, pcg_iters_per_launch=per)
            inner = lp._pcg
            lp._pcg = lambda verify=False, _inner=inner, _fused=fused: _inner(verify=_fused)
            lp.initial_solve()
            es = edge_dict(g)
            its = []
            for b in batches:
                st = lp.apply_and_solve(b)
                assert "pcg" in st.path and st.rho <= 1e-6 * 1.0001, (per, fused, st.path, st.rho)
                its.append(st.extra.get("pcg_iters", 0))
                O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                                   zip(b.del_u.tolist(), b.del_v.tolist()))
            if fused:
                check_against_oracle(lp, g, es, 1e-6, f"enqueued verification, {per} per launch")
            res.append((its, lp.scores_with_h()))
            del lp
        (i0, Z0), (i1, Z1) = res
        assert all(abs(a - b) <= 1 for a, b in zip(i0, i1)), (per, i0, i1)
        C = Z0.shape[1] - 1
        hmax = max(Z0[:, C].max(), Z1[:, C].max())
        np.testing.assert_allclose(Z0[:, :C], Z1[:, :C], rtol=0, atol=2.1e-6 * hmax)


@pytest.mark.parametrize("n,K", [(3001, 8), (2999, 16), (4003, 64)])
def test_fp8_directions_any_row_count(n, K):
    """FP8 direction slots of n rows of ld bytes: for n*ld not a multiple of 16 the slots are
    padded so every slot the operator gathers from stays aligned; solves with deferred updates
    (several slots) certify against the rebuilt oracle, and the footprint equals what the
    handle holds."""
    from dynlp_amd import _lib
    from dynlp_amd.stream import DIRECTION_KINDS, EdgeStreamLP
    g = graphs.sbm(n, 8, 10, 0.8, K, 0.02, seed=61)
    lp = EdgeStreamLP(g, eps=1e-6, pcg_directions="fp8", push_max_rows=0)
    hb = _lib.load().dlp_lp_handle_bytes(lp.n, lp.C1, lp.ld, 0, DIRECTION_KINDS["fp8"], lp.pcg_iters_per_launch,
                                         lp.pcg_coarse, 0)
    assert lp.device_bytes() == hb
    lp.initial_solve()
    es = edge_dict(g)
    for b in graphs.update_stream(g, 2, 300, seed=61)[0]:
        st = lp.apply_and_solve(b)
        assert "pcg" in st.path and st.rho <= 1e-6 * 1.0001
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
    check_against_oracle(lp, g, es, 1e-6, f"fp8 n={n} K={K}")
    assert lp.device_bytes() == hb

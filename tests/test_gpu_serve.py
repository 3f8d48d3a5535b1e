"""Serving: each graph's result is the same alone and in the mix (within the certified bound),
admission under an HBM budget with eviction, fused requests."""
import numpy as np
import pytest

from dynlp_amd import graphs
from oracle import lp as O

pytestmark = pytest.mark.gpu

EPS = 1e-6
_ALWAYS_PUSH = dict(push_frac=1.0, push_max_rows=1 << 20, push_mass=float("inf"), push_budget=1e9)


def _graphs():
    return [graphs.sbm(2000, 4, 8, 0.8, 4, 0.03, seed=21),
            graphs.erdos_renyi(5000, 10, 8, 0.02, seed=22, unit=True),
            graphs.sbm(12000, 16, 12, 0.8, 16, 0.01, seed=23)]


def _requests(gs, n=36, rate=4000.0):
    from dynlp_amd.serve import poisson_requests
    return poisson_requests(gs, rate=rate, duration=10.0, size_lo=20, size_hi=400, seed=5, max_requests=n)


def _alone(g, reqs):
    from dynlp_amd.stream import EdgeStreamLP
    lp = EdgeStreamLP(g, eps=EPS)
    lp.initial_solve()
    for r in reqs:
        st = lp.apply_and_solve(r.batch)
        assert st.rho <= EPS * 1.0001
    return lp.scores_with_h()


def _oracle_hmax(g, reqs):
    es = {(int(a), int(b)): float(w) for a, b, w in zip(g.u, g.v, g.w)}
    for r in reqs:
        b = r.batch
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
    u = np.array([k[0] for k in es]); v = np.array([k[1] for k in es]); w = np.array(list(es.values()))
    sysm = O.build_system(g.n, u, v, w, g.labeled, g.y, g.K)
    Fs = O.solve_system(sysm)
    return sysm, Fs


@pytest.mark.parametrize("tight", [False, True])
def test_alone_equals_mixed(tight):
    from dynlp_amd.serve import Service, footprint_bytes
    gs = _graphs()
    reqs = _requests(gs)
    sizes = [footprint_bytes(g) for g in gs]
    budget = sum(sizes) + (1 << 20) if not tight else sizes[2] + max(sizes[0], sizes[1]) + (1 << 20)
    svc = Service(gs, budget, slots=3, eps=EPS)
    done = svc.run(reqs)
    rep = svc.report()
    errs = [(r.gid, r.error) for r in done if r.error is not None]
    assert rep["requests"] == len(reqs), (len(done), errs[:5], {k: rep[k] for k in ("evictions", "failed_repairs")})
    assert rep["max_rho"] <= EPS * 1.0001
    if tight:
        assert rep["evictions"] > 0
    for gid, g in enumerate(gs):
        mine = [r for r in reqs if r.gid == gid]
        sysm, Fs = _oracle_hmax(g, mine)
        hmax = Fs[:, sysm.C].max()
        Zm = svc.graphs[gid].lp.scores_with_h() if svc.graphs[gid].lp is not None else None
        Za = _alone(g, mine)
        # both certified against the exact solution of the final edge set
        assert np.abs(Za[:, : sysm.C] - Fs[:, : sysm.C]).max() <= EPS * hmax * 1.05
        if Zm is not None:
            assert np.abs(Zm[:, : sysm.C] - Fs[:, : sysm.C]).max() <= EPS * hmax * 1.05
            assert np.abs(Zm - Za)[:, : sysm.C].max() <= 2 * EPS * hmax * 1.05


def test_fuse_last_operation_wins():
    from dynlp_amd.graphs import EdgeBatch
    from dynlp_amd.serve import fuse
    a = EdgeBatch(np.array([1, 2]), np.array([5, 6]), np.array([1.0, 2.0], np.float32), np.array([3]), np.array([4]))
    b = EdgeBatch(np.array([4]), np.array([3]), np.array([7.0], np.float32), np.array([5]), np.array([1]))
    f = fuse([a, b])
    ins = {(min(x, y), max(x, y)): w for x, y, w in zip(f.ins_u, f.ins_v, f.ins_w)}
    dels = {(min(x, y), max(x, y)) for x, y in zip(f.del_u, f.del_v)}
    assert ins == {(2, 6): 2.0, (3, 4): 7.0} and dels == {(1, 5)}


def test_malformed_requests_are_refused_and_the_rest_served():
    """Requests with ids outside their graph or ragged arrays fail on their own (error set,
    nothing applied); every other request is served and certified, and each graph ends at the
    exact solution of its well-formed updates."""
    from dynlp_amd.graphs import EdgeBatch
    from dynlp_amd.serve import Request, Service, footprint_bytes
    gs = _graphs()
    reqs = _requests(gs, n=24)
    i32 = lambda x: np.asarray(x, dtype=np.int32)  # noqa: E731
    t_mid = reqs[len(reqs) // 2].arrival
    bad = [Request(0, EdgeBatch(i32([1]), i32([gs[0].n]), np.ones(1, np.float32), i32([]), i32([])), t_mid),
           Request(2, EdgeBatch(i32([1, 2]), i32([3]), np.ones(1, np.float32), i32([]), i32([])), t_mid)]
    svc = Service(gs, sum(footprint_bytes(g) for g in gs) + (1 << 20), slots=3, eps=EPS)
    svc.run(reqs + bad)
    rep = svc.report()
    assert rep["errors"] == 2 and rep["refused"] == 2 and rep["requests"] == len(reqs)
    assert all(r.error and r.error.startswith("refused") for r in bad)
    assert all(r.error is None and r.rho <= EPS * 1.0001 for r in reqs)
    for gid, g in enumerate(gs):
        sysm, Fs = _oracle_hmax(g, [r for r in reqs if r.gid == gid])
        Z = svc.graphs[gid].lp.scores_with_h()
        assert np.abs(Z[:, : sysm.C] - Fs[:, : sysm.C]).max() <= EPS * Fs[:, sysm.C].max() * 1.05


def test_failed_builds_fail_their_requests_and_the_service_drains():
    """A graph whose device state cannot be built (here: an invalid solver option, refused before
    any allocation) fails each of its requests with the reason; the service still drains every
    queue and returns."""
    from dynlp_amd.serve import Service, footprint_bytes
    gs = _graphs()[:2]
    reqs = _requests(gs, n=10)
    svc = Service(gs, sum(footprint_bytes(g) for g in gs) + (1 << 20), slots=2, eps=EPS,
                  lp_kwargs={"pcg_mode": "no-such-mode"})
    done = svc.run(reqs, preload=False)
    assert len(done) == len(reqs)
    assert all(r.error and "repair failed: ValueError" in r.error for r in done)
    rep = svc.report()
    assert rep["requests"] == 0 and rep["errors"] == len(reqs) and rep["failed_repairs"] >= 1
    assert svc.resident == 0


def test_fusion_stops_at_the_group_limit(monkeypatch):
    """Requests queued together are fused only while their operations fit one group (the limit
    lowered here so that two of these requests fit); every request is served and the graph ends
    at the exact solution of all of them."""
    from dynlp_amd import serve
    gs = _graphs()[:1]
    reqs = serve.poisson_requests(gs, rate=1e9, duration=1.0, size_lo=600, size_hi=601, seed=7, max_requests=6)
    ops = [r.batch.size for r in reqs]
    limit = max(a + b for a, b in zip(ops, ops[1:]))
    assert limit < min(ops) * 3
    monkeypatch.setattr(serve, "_MAX_GROUP_OPS", limit)
    svc = serve.Service(gs, serve.footprint_bytes(gs[0]) + (1 << 20), slots=1, eps=EPS)
    svc.run(reqs)
    rep = svc.report()
    assert rep["errors"] == 0 and rep["requests"] == len(reqs)
    assert max(r.fused_with for r in reqs) <= 2 and rep["repairs"] >= 3
    sysm, Fs = _oracle_hmax(gs[0], reqs)
    Z = svc.graphs[0].lp.scores_with_h()
    assert np.abs(Z[:, : sysm.C] - Fs[:, : sysm.C]).max() <= EPS * Fs[:, sysm.C].max() * 1.05


def test_evicted_graph_returns_with_its_updates():
    """Two graphs under a budget that holds one: requests alternate between them, so each
    re-admission rebuilds a graph from its checkpoint (lowered so that checkpoints are taken)
    plus the batches served since; both end at the exact solution of all their updates."""
    from dynlp_amd import serve
    gs = _graphs()[:2]
    reqs = serve.poisson_requests(gs, rate=300.0, duration=10.0, size_lo=50, size_hi=300, seed=11,
                                  max_requests=16)
    sizes = [serve.footprint_bytes(g) for g in gs]
    svc = serve.Service(gs, max(sizes) + (1 << 20), slots=1, eps=EPS)
    orig = svc._served

    def served(s, batches):            # checkpoint after every repair
        orig(s, batches)
        if s.lp is not None and s.history:
            s.base = s.lp.edge_graph()
            s.history, s.hist_ops = [], 0
    svc._served = served
    svc.run(reqs, preload=False)
    rep = svc.report()
    assert rep["errors"] == 0 and rep["requests"] == len(reqs) and rep["evictions"] >= 2
    for gid, g in enumerate(gs):
        mine = [r for r in reqs if r.gid == gid]
        sysm, Fs = _oracle_hmax(g, mine)
        slot = svc.graphs[gid]
        if slot.lp is None:
            slot.lp = svc._new_state(slot)
        Z = slot.lp.scores_with_h()
        assert np.abs(Z[:, : sysm.C] - Fs[:, : sysm.C]).max() <= EPS * Fs[:, sysm.C].max() * 1.05, gid

def test_push_across_graphs_matches_oracle():
    """dlp_lp_push_multi through the C ABI: three graphs of different column counts (so more
    than one column class) take small batches; each call pushes the graphs whose batch a push
    can repair together (graph 1 is left out of every second call and repairs itself), and every
    graph ends certified at the exact solution of its final edge set.  The repair policy is set
    to push whatever the batch's residual mass, so that every batch is a push candidate."""
    import ctypes as ct
    from dynlp_amd import _lib
    from dynlp_amd.stream import EdgeStreamLP
    gs = _graphs()
    lps = [EdgeStreamLP(g, eps=EPS, **_ALWAYS_PUSH) for g in gs]
    for lp in lps:
        lp.initial_solve()
    hs = (ct.c_void_p * 3)(*[lp.h.value for lp in lps])
    plan = ct.c_void_p()
    _lib.call("dlp_push_plan_create", hs, 3, ct.byref(plan))
    streams = [graphs.update_stream(g, 4, 60, seed=30 + i)[0] for i, g in enumerate(gs)]
    sets = [graphs.EdgeSet(g) for g in gs]
    multi = 0
    try:
        for t in range(4):
            reqs = [lp.begin_request(streams[i][t]) for i, lp in enumerate(lps)]
            active = np.array([int(lp.push_planned(r["status"]) and not (i == 1 and t % 2))
                               for i, (lp, r) in enumerate(zip(lps, reqs))], dtype=np.int32)
            multi = max(multi, int(active.sum()))
            budget = np.array([lp.push_budget_edges() for lp in lps], dtype=np.int64)
            rounds = np.zeros(3, dtype=np.int64)
            conv = np.zeros(3, dtype=np.int32)
            _lib.call("dlp_lp_push_multi", plan, active.ctypes.data, 200_000, 4, 1, budget.ctypes.data,
                      rounds.ctypes.data, conv.ctypes.data, None)
            for i, (lp, r) in enumerate(zip(lps, reqs)):
                st = lp.end_request(r, pushed=(int(rounds[i]), bool(conv[i])) if active[i] else None)
                assert st.rho <= EPS * 1.0001, (t, i, st.rho)
                assert active[i] or rounds[i] == 0
                sets[i].apply(streams[i][t])
    finally:
        _lib.call("dlp_push_plan_destroy", plan)
    assert multi >= 2
    for i, g in enumerate(gs):
        u, v, w = sets[i].pairs()
        sysm = O.build_system(g.n, u, v, w, g.labeled, g.y, g.K)
        Fs = O.solve_system(sysm)
        Z = lps[i].scores_with_h()
        assert np.abs(Z[:, : sysm.C] - Fs[:, : sysm.C]).max() <= EPS * Fs[:, sysm.C].max() * 1.05, i


@pytest.mark.parametrize("budget", [1e9, 1e-4])
def test_service_pushes_graphs_together(budget):
    """Eight small graphs whose requests arrive at once: the scheduler repairs them with pushes
    that cover several graphs per call (the repair policy set to push whatever the residual
    mass), and every graph ends certified at the exact solution of its updates.  With an edge
    budget of 1e-4 nnz the pushes stop early and every graph goes on to a global solve in a
    worker."""
    from dynlp_amd.serve import Service, footprint_bytes, poisson_requests
    kw = dict(_ALWAYS_PUSH, push_budget=budget)
    gs = [graphs.sbm(1500 + 300 * i, 4, 8, 0.8, 4, 0.03, seed=40 + i) for i in range(8)]
    reqs = poisson_requests(gs, rate=1e9, duration=1.0, size_lo=10, size_hi=40, seed=9, max_requests=32)
    svc = Service(gs, sum(footprint_bytes(g, **kw) for g in gs) + (1 << 20), slots=2, eps=EPS, lp_kwargs=kw)
    svc.run(reqs)
    rep = svc.report()
    assert rep["errors"] == 0 and rep["requests"] == len(reqs)
    assert rep["push_launches"] >= 1 and rep["graphs_per_push_launch"] > 1, rep
    if budget < 1:
        assert rep["global_repairs"] >= 1, rep
    for gid, g in enumerate(gs):
        mine = [r for r in reqs if r.gid == gid]
        sysm, Fs = _oracle_hmax(g, mine)
        Z = svc.graphs[gid].lp.scores_with_h()
        assert np.abs(Z[:, : sysm.C] - Fs[:, : sysm.C]).max() <= EPS * Fs[:, sysm.C].max() * 1.05, gid


def test_service_regrows_stores_on_many_streams():
    """Stores with no slack, so that nearly every repair regrows its graph's store, served by
    four workers (global solves) and the scheduler, each on its own stream: a regrown store's
    old memory is released only after the copy that reads it (the caching allocator hands
    memory back to the stream that allocated it, which may be another thread's).  Every graph
    ends at the exact solution of its updates."""
    from dynlp_amd.serve import Service, footprint_bytes, stream_requests
    kw = dict(slack_min=0, slack_frac=0.0)
    gs = [graphs.sbm(4000 + 1000 * i, 4, 10, 0.8, 4, 0.02, seed=60 + i) for i in range(6)]
    reqs = stream_requests(gs, rate=3000, duration=0.2, size_lo=200, size_hi=3000, seed=13)
    svc = Service(gs, 4 * sum(footprint_bytes(g, **kw) for g in gs), slots=4, eps=EPS, lp_kwargs=kw)
    svc.run(reqs)
    rep = svc.report()
    assert rep["errors"] == 0 and rep["requests"] == len(reqs)
    for gid, g in enumerate(gs):
        mine = [r for r in reqs if r.gid == gid]
        sysm, Fs = _oracle_hmax(g, mine)
        lp = svc.graphs[gid].lp
        assert lp.col.numel() > 2 * g.m          # regrown past the slack-free store
        Z = lp.scores_with_h()
        assert np.abs(Z[:, : sysm.C] - Fs[:, : sysm.C]).max() <= EPS * Fs[:, sysm.C].max() * 1.05, gid

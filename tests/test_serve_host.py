"""Host logic of the update service (no device): fusing request batches, the ordered operation
sequence of a request group, the serving request stream."""
import numpy as np

from oracle import lp as O


def _apply(es, b):
    return O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                              zip(b.del_u.tolist(), b.del_v.tolist()))


def _random_batches(n, m, count, rng):
    from dynlp_amd.graphs import EdgeBatch
    out = []
    for _ in range(count):
        k = int(rng.integers(1, m))
        # distinct pairs within the deletions and within the insertions of one batch
        pi = np.unique(np.sort(rng.integers(0, n, (k, 2)), axis=1), axis=0)
        pd = np.unique(np.sort(rng.integers(0, n, (k, 2)), axis=1), axis=0)
        out.append(EdgeBatch(pi[:, 0], pi[:, 1], rng.integers(1, 5, len(pi)).astype(np.float32),
                             pd[:, 0], pd[:, 1]))
    return out


def test_fused_history_equals_sequential_application():
    from dynlp_amd.serve import fuse
    rng = np.random.default_rng(3)
    for trial in range(20):
        base = {(a, b): 1.0 for a, b in np.unique(np.sort(rng.integers(0, 40, (60, 2)), axis=1), axis=0).tolist()
                if a != b}
        bs = _random_batches(40, 30, int(rng.integers(1, 6)), rng)
        seq, fused = dict(base), dict(base)
        for b in bs:
            _apply(seq, b)
        _apply(fused, fuse(bs))
        assert seq == fused, trial


def _last_operation(es, u, v, w):
    """The ordered operation sequence as dlp_lp_apply_ops defines it: each unordered pair takes
    only its last operation (w == 0 deletes, w > 0 inserts with that weight)."""
    last = {}
    for s_, (a, b, x) in enumerate(zip(u.tolist(), v.tolist(), w.tolist())):
        if a != b:
            last[(min(a, b), max(a, b))] = x
    for k, x in last.items():
        if x == 0:
            es.pop(k, None)
        else:
            es[k] = float(x)


def test_ordered_operations_equal_sequential_application():
    """pack_ops lays the batches out so that each pair's last operation is what applying the
    batches one after the other leaves: per batch its deletions, then its insertions (a batch
    that deletes and re-inserts a pair keeps it), batches in order.  Pairs repeat within and
    across batches, in both orientations."""
    from dynlp_amd.graphs import EdgeBatch, pack_ops
    rng = np.random.default_rng(4)
    for trial in range(30):
        base = {(a, b): 2.0 for a, b in np.unique(np.sort(rng.integers(0, 30, (60, 2)), axis=1), axis=0).tolist()
                if a != b}
        bs = []
        for _ in range(int(rng.integers(1, 7))):
            ni, nd = int(rng.integers(0, 25)), int(rng.integers(0, 25))
            iu, iv = rng.integers(0, 30, ni), rng.integers(0, 30, ni)
            du, dv = rng.integers(0, 30, nd), rng.integers(0, 30, nd)
            bs.append(EdgeBatch(iu, iv, rng.integers(1, 5, ni).astype(np.float32), du, dv))
        seq, ops = dict(base), dict(base)
        for b in bs:
            # one batch: deletions first, then the last insertion of each pair
            for a, c in zip(b.del_u.tolist(), b.del_v.tolist()):
                seq.pop((min(a, c), max(a, c)), None)
            for a, c, x in zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()):
                if a != c:
                    seq[(min(a, c), max(a, c))] = float(x)
        u, v, w = pack_ops(bs)
        assert u.dtype == np.int32 and v.dtype == np.int32 and w.dtype == np.float32
        assert len(u) == sum(b.size for b in bs)
        _last_operation(ops, u, v, w)
        assert seq == ops, trial


def test_stream_requests_delete_present_edges():
    """The serving stream: arrivals sorted, sizes within bounds, half insertions of pairs u != v,
    and every deletion names an edge present at that point of the graph's stream (checked by
    replaying the stream per graph on a host edge set)."""
    from dynlp_amd import graphs
    from dynlp_amd.serve import stream_requests
    gs = [graphs.erdos_renyi(3000, 8, 4, 0.02, seed=1), graphs.sbm(2000, 4, 10, 0.8, 4, 0.02, seed=2)]
    reqs = stream_requests(gs, rate=500, duration=1.0, size_lo=10, size_hi=3000, seed=3)
    assert len(reqs) > 300
    t = [r.arrival for r in reqs]
    assert t == sorted(t)
    sets = [{(int(a), int(b)) for a, b in zip(g.u, g.v)} for g in gs]
    missing = 0
    for r in reqs:
        b = r.batch
        assert 10 <= b.size <= 3000 and len(b.ins_u) == b.size // 2
        assert np.all(b.ins_u != b.ins_v) and np.all(b.ins_w > 0)
        es = sets[r.gid]
        for a, c in zip(b.del_u.tolist(), b.del_v.tolist()):
            k = (min(a, c), max(a, c))
            if k in es:
                es.remove(k)
            else:
                missing += 1
        for a, c in zip(b.ins_u.tolist(), b.ins_v.tolist()):
            es.add((min(a, c), max(a, c)))
    assert missing <= 0.01 * sum(len(r.batch.del_u) for r in reqs)

def test_check_batch_refuses_ragged_and_out_of_range():
    import pytest
    from dynlp_amd.graphs import EdgeBatch, check_batch
    i = lambda *x: np.asarray(x, dtype=np.int32)  # noqa: E731
    f = lambda *x: np.asarray(x, dtype=np.float32)  # noqa: E731
    e = np.zeros(0, dtype=np.int32)
    check_batch(EdgeBatch(i(0, 9), i(9, 0), f(1, 2), i(3), i(4)), 10)       # valid, ids at both ends
    check_batch(EdgeBatch(e, e, e.astype(np.float32), e, e), 10)           # empty
    check_batch(EdgeBatch(i(5), i(5), f(1), i(6), i(6)), 10)               # self-loops: ignored later
    for bad, exc in ((EdgeBatch(i(0), i(10), f(1), e, e), IndexError),
                     (EdgeBatch(e, e, e.astype(np.float32), i(-1), i(2)), IndexError),
                     (EdgeBatch(i(1, 2), i(3), f(1, 1), e, e), ValueError),
                     (EdgeBatch(i(1), i(3), f(1, 1), e, e), ValueError),
                     (EdgeBatch(e, e, e.astype(np.float32), i(1, 2), i(3)), ValueError)):
        with pytest.raises(exc):
            check_batch(bad, 10)
    two = EdgeBatch(i(1, 2), i(3, 4), f(1, 1), e, e)
    check_batch(two, 10, max_insertions=2)
    with pytest.raises(ValueError, match="at most 1"):
        check_batch(two, 10, max_insertions=1)


def test_check_batch_refuses_bad_weights():
    """Insertion weights must be finite and positive as FP32 (the device operator's weights):
    NaN, +-inf, 0, negatives and values that overflow FP32 are refused."""
    import pytest
    from dynlp_amd.graphs import EdgeBatch, check_batch
    i = lambda *x: np.asarray(x, dtype=np.int32)  # noqa: E731
    e = np.zeros(0, dtype=np.int32)
    check_batch(EdgeBatch(i(1), i(2), np.array([1e-30]), e, e), 10)
    for w in (np.nan, np.inf, -np.inf, 0.0, -1.0, 1e39):
        with pytest.raises(ValueError, match="finite and positive"):
            check_batch(EdgeBatch(i(1, 3), i(2, 4), np.array([1.0, w]), e, e), 10)


def test_normalize_batch_equals_sequential_semantics():
    """A batch with pairs repeated in both orientations, within the deletions and within the
    insertions, normalised to one entry per pair, changes an edge set exactly as the oracle's
    sequential application of the raw batch (deletions first, then insertions in order: the
    last weight wins)."""
    from dynlp_amd.graphs import EdgeBatch, normalize_batch
    from oracle import lp as O
    rs = np.random.RandomState(5)
    n = 12
    for trial in range(200):
        base = {}
        for _ in range(20):
            a, b = rs.randint(0, n, 2)
            if a != b:
                base[(min(a, b), max(a, b))] = float(rs.randint(1, 5))
        k = rs.randint(0, 15)
        iu, iv = rs.randint(0, n, k), rs.randint(0, n, k)
        iw = rs.randint(1, 9, k).astype(np.float32)
        m = rs.randint(0, 10)
        du, dv = rs.randint(0, n, m), rs.randint(0, n, m)
        raw = EdgeBatch(iu.astype(np.int32), iv.astype(np.int32), iw, du.astype(np.int32), dv.astype(np.int32))
        nb = normalize_batch(raw, n)
        ki = np.minimum(nb.ins_u, nb.ins_v).astype(np.int64) * n + np.maximum(nb.ins_u, nb.ins_v)
        kd = np.minimum(nb.del_u, nb.del_v).astype(np.int64) * n + np.maximum(nb.del_u, nb.del_v)
        assert np.unique(ki).shape == ki.shape and np.unique(kd).shape == kd.shape
        seq = O.apply_edge_batch(dict(base), zip(iu.tolist(), iv.tolist(), iw.tolist()), zip(du.tolist(), dv.tolist()))
        one = O.apply_edge_batch(dict(base), zip(nb.ins_u.tolist(), nb.ins_v.tolist(), nb.ins_w.tolist()),
                                 zip(nb.del_u.tolist(), nb.del_v.tolist()))
        assert seq == one, trial
    b = EdgeBatch(np.array([1, 2], np.int32), np.array([2, 3], np.int32), np.array([1.0, 2.0], np.float32),
                  np.array([4], np.int32), np.array([5], np.int32))
    assert normalize_batch(b, n) is b


def test_fuse_of_one_batch_collapses_repeats():
    """A single request goes through the same de-duplication as a fused group, so it changes its
    graph identically whether served alone or fused."""
    from dynlp_amd.graphs import EdgeBatch
    from dynlp_amd.serve import fuse
    from oracle import lp as O
    b = EdgeBatch(np.array([1, 2, 1], np.int32), np.array([2, 1, 3], np.int32), np.array([1.0, 5.0, 2.0], np.float32),
                  np.array([3, 1], np.int32), np.array([1, 3], np.int32))
    f = fuse([b])
    keys = np.minimum(f.ins_u, f.ins_v) * 10 + np.maximum(f.ins_u, f.ins_v)
    assert np.unique(keys).shape == keys.shape
    base = {(1, 3): 4.0, (0, 1): 1.0}
    seq = O.apply_edge_batch(dict(base), zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                             zip(b.del_u.tolist(), b.del_v.tolist()))
    one = O.apply_edge_batch(dict(base), zip(f.ins_u.tolist(), f.ins_v.tolist(), f.ins_w.tolist()),
                             zip(f.del_u.tolist(), f.del_v.tolist()))
    assert seq == one == {(0, 1): 1.0, (1, 2): 5.0, (1, 3): 2.0}


def test_library_batch_check_matches_brute_force():
    """dlp_check_edge_batch (one hashed pass in the library, no device work) against a brute
    force over random batches: the repeat mask (pairs unordered, deletions and insertions
    separately), the first id out of range and the first bad weight, for int32 and int64 ids and
    vertex counts up to 2^31 (keys near 2^62)."""
    import ctypes as ct
    from dynlp_amd import _lib
    lib = _lib.load()
    rs = np.random.RandomState(9)
    P = ct.c_void_p
    for trial in range(400):
        n = int(rs.choice([5, 50, 1 << 20, (1 << 31) - 1]))
        dt = np.int32 if rs.rand() < 0.5 else np.int64
        nd, ni = rs.randint(0, 40), rs.randint(0, 40)
        du, dv, iu, iv = (rs.randint(0, min(n, 1 << 30), m).astype(dt) for m in (nd, nd, ni, ni))
        if n > 100 and nd > 2 and rs.rand() < 0.5:      # plant a repeat (either orientation)
            du[-1], dv[-1] = (dv[0], du[0]) if rs.rand() < 0.5 else (du[0], dv[0])
        if n > 100 and ni > 2 and rs.rand() < 0.5:
            iu[-1], iv[-1] = iv[1], iu[1]
        iw = rs.uniform(0.1, 2, ni).astype(np.float32)
        expect_err = 0
        if rs.rand() < 0.15 and nd + ni:
            arrs = [a for a in (du, dv, iu, iv) if len(a)]
            a = arrs[rs.randint(len(arrs))]
            a[rs.randint(len(a))] = -1 if rs.rand() < 0.5 else (n if dt == np.int64 or n < (1 << 31) - 1 else -2)
            expect_err = -1
        elif rs.rand() < 0.15 and ni:
            iw[rs.randint(ni)] = [np.nan, np.inf, 0.0, -1.0][rs.randint(4)]
            expect_err = -2
        where = ct.c_int64(-7)
        rc = lib.dlp_check_edge_batch(n, nd, P(du.ctypes.data), P(dv.ctypes.data), ni, P(iu.ctypes.data),
                                      P(iv.ctypes.data), P(iw.ctypes.data), np.dtype(dt).itemsize, ct.byref(where))
        flat = np.concatenate([du, dv, iu, iv]).astype(np.int64)
        bad_id = np.flatnonzero((flat < 0) | (flat >= n))
        bad_w = np.flatnonzero(~(np.isfinite(iw) & (iw > 0)))
        if len(bad_id):
            assert rc == -1 and where.value == bad_id[0], (trial, rc, where.value)
        elif len(bad_w):
            assert rc == -2 and where.value == bad_w[0], (trial, rc, where.value)
        else:
            assert expect_err == 0
            rep = lambda u, v: len({(min(a, b), max(a, b)) for a, b in zip(u.tolist(), v.tolist())}) < len(u)  # noqa: E731
            assert rc == (1 if rep(du, dv) else 0) | (2 if rep(iu, iv) else 0), trial
    assert lib.dlp_check_edge_batch(0, 0, None, None, 0, None, None, None, 4, None) == -3


def test_direction_kind_by_column_count():
    """The default direction storage ("auto") is FP8 from 48 columns (label columns + h) up and
    FP16 below; explicit kinds pass through and unknown ones are refused."""
    import pytest
    from dynlp_amd.stream import DIRECTION_KINDS, FP8_MIN_COLUMNS, resolve_directions
    assert FP8_MIN_COLUMNS == 48
    assert resolve_directions("auto", 65) == "fp8" and resolve_directions("auto", 48) == "fp8"
    assert resolve_directions("auto", 47) == "fp16" and resolve_directions("auto", 17) == "fp16"
    for k in DIRECTION_KINDS:
        assert resolve_directions(k, 9) == k
    with pytest.raises(ValueError):
        resolve_directions("fp4", 65)

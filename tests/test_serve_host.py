"""Host logic of the update service (no device): fusing request batches and replaying a fused
history in chunks."""
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


def test_replay_chunks_equal_the_fused_batch():
    from dynlp_amd.serve import fuse, replay_chunks
    rng = np.random.default_rng(4)
    for trial in range(20):
        base = {(a, b): 2.0 for a, b in np.unique(np.sort(rng.integers(0, 50, (80, 2)), axis=1), axis=0).tolist()
                if a != b}
        f = fuse(_random_batches(50, 40, 4, rng))
        one, chunked = dict(base), dict(base)
        _apply(one, f)
        pieces = replay_chunks(f, chunk=int(rng.integers(1, 7)))
        assert sum(p.size for p in pieces) == f.size
        for p in pieces:
            _apply(chunked, p)
        assert one == chunked, trial


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

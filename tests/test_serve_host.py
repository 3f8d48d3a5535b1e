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

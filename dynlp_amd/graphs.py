"""Synthetic graph inputs and host-side CSR conventions.

A graph here is an undirected, loop-free, weighted edge set over vertices 0..n-1, held on the
host as three parallel arrays (u < v, w > 0).  The CSR form stores both directions with sorted
columns per row, duplicates merged and self-loops dropped, following the canonical form of
graphs.py:36-48 in the reference (`coalesce`, `symmetric_from_pairs`), except that a pair
drawn twice keeps its first weight instead of summing both draws (unit-weight ER stays unit).

Weights are drawn in FP64 and rounded to FP32 once, at generation: the FP32 value is THE
weight, so the FP32 working operator and the FP64 certification operator are the same matrix
(an FP32 weight is exact in FP64).

Seeds follow one convention: graph S, labels S+1, update stream S+2, arrivals S+3.
All generators use numpy's PCG64 (`numpy.random.default_rng`).
"""
from __future__ import annotations

import ctypes as ct
import os
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Graph:
    """Undirected weighted graph with labels; edges are unique pairs u < v."""
    name: str
    n: int
    u: np.ndarray          # int64[m], u < v
    v: np.ndarray          # int64[m]
    w: np.ndarray          # float32[m] > 0
    labeled: np.ndarray    # bool[n]
    y: np.ndarray          # int32[n] label of every vertex (ground truth; used where labeled)
    K: int                 # classes
    block: np.ndarray | None = None   # SBM block of each vertex
    meta: dict = field(default_factory=dict)

    @property
    def m(self) -> int:
        return int(self.u.shape[0])

    def csr(self):
        """(indptr int32[n+1], indices int32[2m], data float32[2m]) with sorted columns."""
        return csr_from_pairs(self.n, self.u, self.v, self.w)


def canonical_pairs(n, a, b, w, merge="first"):
    """Unique undirected pairs (u < v), self-loops dropped.  A pair given more than once (in
    either direction) keeps its first occurrence's weight (merge="first", the generators), or
    the FP32 sum of all its weights in the order given (merge="sum", the reference's `coalesce`,
    graphs.py:36-44, which external edge lists go through).  Pairs come out in the order of
    their first occurrence."""
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    w = np.asarray(w, dtype=np.float32)
    keep = a != b
    a, b, w = a[keep], b[keep], w[keep]
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo * n + hi
    _, first, inv = np.unique(key, return_index=True, return_inverse=True)
    if merge == "sum":
        # sequential FP32 sums per pair, in the order the pairs were given
        order = np.argsort(inv, kind="stable")
        starts = np.searchsorted(inv[order], np.arange(first.shape[0]))
        wsum = np.add.reduceat(w[order], starts, dtype=np.float32) if order.size else w[:0]
    elif merge != "first":
        raise ValueError(f"merge must be 'first' or 'sum', not {merge!r}")
    pos = np.argsort(first, kind="stable")
    first = first[pos]
    wout = w[first] if merge == "first" else wsum[pos]
    return lo[first], hi[first], wout


def csr_from_pairs(n, u, v, w):
    """Symmetric CSR of unique pairs: both directions, columns sorted within each row."""
    r = np.concatenate([u, v]).astype(np.int64)
    c = np.concatenate([v, u]).astype(np.int64)
    d = np.concatenate([w, w]).astype(np.float32)
    order = np.argsort(r * max(int(n), 1) + c)      # one int64 key: row-major, columns ascending
    r, c, d = r[order], c[order], d[order]
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(r, minlength=n), out=indptr[1:])
    if indptr[-1] >= 2**31:
        raise ValueError("graph too large for int32 CSR offsets")
    return indptr.astype(np.int32), c.astype(np.int32), d


def csr_rows(indptr):
    """Row index of every stored entry (graphs.py:51-64 `expand_counts`/`csr_rows`)."""
    indptr = np.asarray(indptr)
    return np.repeat(np.arange(indptr.shape[0] - 1, dtype=np.int32), np.diff(indptr))


def _labels(n, frac, seed, min_count=1):
    rng = np.random.default_rng(seed)
    cnt = max(min_count, int(round(frac * n)))
    lab = np.zeros(n, dtype=bool)
    lab[rng.choice(n, size=cnt, replace=False)] = True
    return lab


def erdos_renyi(n, deg, K, label_frac=0.01, seed=0, unit=True, name=None) -> Graph:
    """ER: m = n*deg/2 uniform pairs, loops dropped, duplicates merged; labels uniform over K."""
    rng = np.random.default_rng(seed)
    m = int(n * deg / 2)
    a = rng.integers(0, n, size=m, dtype=np.int64)
    b = rng.integers(0, n, size=m, dtype=np.int64)
    w = np.ones(m, dtype=np.float32) if unit else rng.uniform(0.5, 1.5, size=m).astype(np.float32)
    u, v, w = canonical_pairs(n, a, b, w)
    lab = _labels(n, label_frac, seed + 1, min_count=K)
    y = np.random.default_rng(seed + 1).integers(0, K, size=n).astype(np.int32)
    return Graph(name or f"er-n{n}-d{deg:g}-k{K}", n, u, v, w, lab, y, K,
                 meta=dict(kind="er", deg=deg, seed=seed))


def sbm(n, blocks, deg, p_in, K, label_frac=0.01, seed=0, wlo=0.5, whi=1.5, unit=False, name=None) -> Graph:
    """SBM with equal blocks block(v) = v*blocks // n; a p_in share of the n*deg/2 drawn pairs
    is intra-block (uniform inside the block of a uniform endpoint), the rest joins two
    different blocks uniformly.  Weights U[wlo, whi] rounded to FP32 (or 1).  Labels: block mod K."""
    rng = np.random.default_rng(seed)
    m = int(n * deg / 2)
    blk = (np.arange(n, dtype=np.int64) * blocks) // n
    start = (np.arange(blocks + 1, dtype=np.int64) * n + blocks - 1) // blocks
    # start[b] = first vertex with v*blocks//n == b  (ceil(b*n/blocks))
    size = np.diff(start)
    a = rng.integers(0, n, size=m, dtype=np.int64)
    intra = rng.random(m) < p_in
    ba = blk[a]
    b_in = start[ba] + (rng.random(m) * size[ba]).astype(np.int64)
    other = rng.integers(0, blocks - 1, size=m, dtype=np.int64)
    other = other + (other >= ba)              # a block different from a's
    b_out = start[other] + (rng.random(m) * size[other]).astype(np.int64)
    b = np.where(intra, b_in, b_out)
    w = np.ones(m, dtype=np.float32) if unit else rng.uniform(wlo, whi, size=m).astype(np.float32)
    u, v, w = canonical_pairs(n, a, b, w)
    lab = _labels(n, label_frac, seed + 1, min_count=K)
    y = (blk % K).astype(np.int32)
    return Graph(name or f"sbm-n{n}-b{blocks}-d{deg:g}-k{K}", n, u, v, w, lab, y, K, block=blk,
                 meta=dict(kind="sbm", blocks=blocks, deg=deg, p_in=p_in, seed=seed))


# ---------------------------------------------------------------------------------------------
# On-disk edge lists (the reference's graphs.py:176-204 format).

def from_npz(path, label_frac=0.01, seed=0, name=None) -> Graph:
    """Graph of an edge-list .npz as the reference reads it (`from_npz`, graphs.py:176-184):
    arrays src, dst (int), y (int, one per vertex; n = len(y), K = max(y) + 1) and optionally
    w (default 1) and order (a vertex permutation, the arrival order; default a permutation
    drawn by numpy's legacy RandomState(seed + 1), the generator the reference calls on its
    array module).  Pairs may come in either direction and more than once; duplicates are
    summed and self-loops dropped as in `coalesce`.  Labelled vertices: a label_frac sample
    (seed S+1) of the vertices with y >= 0 (y < 0 marks a vertex without ground truth).  The
    file is read without unpickling."""
    with np.load(path, allow_pickle=False) as z:
        keys = set(z.files)
        missing = {"src", "dst", "y"} - keys
        if missing:
            raise ValueError(f"{path}: missing array(s) {sorted(missing)}")
        src = z["src"].astype(np.int64).ravel()
        dst = z["dst"].astype(np.int64).ravel()
        y = z["y"].astype(np.int32).ravel()
        w = z["w"].astype(np.float32).ravel() if "w" in keys else np.ones(src.shape[0], dtype=np.float32)
        order = z["order"].astype(np.int64).ravel() if "order" in keys else None
    n = int(y.shape[0])
    if src.shape != dst.shape or w.shape != src.shape:
        raise ValueError(f"{path}: src, dst and w must have one entry per edge")
    if src.size and (min(src.min(), dst.min()) < 0 or max(src.max(), dst.max()) >= n):
        raise ValueError(f"{path}: vertex ids must lie in [0, {n})")
    if not np.all(np.isfinite(w)) or (w.size and w.min() <= 0):
        raise ValueError(f"{path}: weights must be finite and positive")
    if order is None:
        order = np.random.RandomState(seed + 1).permutation(n).astype(np.int64)
    elif not np.array_equal(np.sort(order), np.arange(n)):
        raise ValueError(f"{path}: order must be a permutation of 0..{n - 1}")
    known = np.flatnonzero(y >= 0)
    if known.size == 0:
        raise ValueError(f"{path}: no vertex has a label (y >= 0)")
    K = int(y.max()) + 1
    u, v, w = canonical_pairs(n, src, dst, w, merge="sum")
    rng = np.random.default_rng(seed + 1)
    cnt = min(known.size, max(1, int(round(label_frac * known.size))))
    lab = np.zeros(n, dtype=bool)
    lab[rng.choice(known, size=cnt, replace=False)] = True
    return Graph(name or os.path.splitext(os.path.basename(path))[0], n, u, v, w, lab, y, K,
                 meta=dict(kind="npz", path=str(path), seed=seed, order=order))


def save_npz(g: Graph, path) -> None:
    """Writes g in the format `from_npz` reads: its unique pairs as the upper triangle
    (src < dst, the layout of the reference's `build_cached`, graphs.py:187-204), w, y and the
    arrival order when g carries one."""
    arrays = dict(src=g.u.astype(np.int64), dst=g.v.astype(np.int64), w=g.w.astype(np.float32),
                  y=g.y.astype(np.int32))
    if g.meta.get("order") is not None:
        arrays["order"] = np.asarray(g.meta["order"], dtype=np.int64)
    np.savez(path, **arrays)


def load(spec: str, *, n=100_000, K=2, deg=10.0, p_in=0.85, label_frac=0.01, seed=0) -> Graph:
    """The reference's dataset switch (`load`, graphs.py:207-214) over the sources this package
    has: "sbm" (n/K equal blocks, labels = block), "er", or a path to an edge-list .npz.  The
    reference's downloaded and cached corpora (imdb, synthK) are not part of this package."""
    if spec == "sbm":
        return sbm(n, K, deg, p_in, K, label_frac, seed)
    if spec == "er":
        return erdos_renyi(n, deg, K, label_frac, seed)
    if spec.endswith(".npz"):
        return from_npz(spec, label_frac, seed)
    raise ValueError(f"unknown dataset {spec!r}: 'sbm', 'er' or a path to an edge-list .npz")


# ---------------------------------------------------------------------------------------------
# The five workload configurations (sizes as stated; `scale` shrinks n for tests).

def config_graph(cfg: int, scale: float = 1.0, seed: int = 0) -> Graph:
    """Graph of workload configuration cfg (1..4); n multiplied by `scale`."""
    if cfg == 1:   # CPU case: ER n=1e5, deg 10, unit weights, 1% labelled, k=10
        return erdos_renyi(int(100_000 * scale), 10, 10, 0.01, seed, unit=True)
    if cfg == 2:   # ER n=1e6, deg 16, unit, 1%, k=16
        return erdos_renyi(int(1_000_000 * scale), 16, 16, 0.01, seed, unit=True)
    if cfg == 3:   # SBM n=1e6, 64 blocks, deg 20, 80% intra, w~U[0.5,1.5], 0.5% labelled, k=64
        return sbm(int(1_000_000 * scale), 64, 20, 0.8, 64, 0.005, seed)
    if cfg == 4:   # SBM n=3e5, 32 blocks, deg 8, 90% intra, 1% labelled, k=32 (unit weights)
        return sbm(int(300_000 * scale), 32, 8, 0.9, 32, 0.01, seed, unit=True)
    raise ValueError(f"no graph for configuration {cfg}")


def serving_graphs(count=32, seed=0, scale=1.0):
    """The serving workload's resident graphs: graph i is ER or SBM by a coin of seed+i, n
    log-uniform in [50k, 2M] (times `scale`), average degree uniform in [8, 24], k drawn from
    {8, 16, 64}; SBM graphs have k blocks, 80 % intra-block edges and weights U[0.5, 1.5], ER
    graphs unit weights; 1 % labelled."""
    rng = np.random.default_rng(seed + 7)
    out = []
    for i in range(count):
        n = max(1000, int(np.exp(rng.uniform(np.log(50_000), np.log(2_000_000))) * scale))
        deg = float(rng.uniform(8, 24))
        K = int(rng.choice([8, 16, 64]))
        if rng.random() < 0.5:
            out.append(erdos_renyi(n, deg, K, 0.01, seed + 100 + i, unit=True))
        else:
            out.append(sbm(n, K, deg, 0.8, K, 0.01, seed + 100 + i))
    return out


# ---------------------------------------------------------------------------------------------
# Edge-update streams.

@dataclass
class EdgeBatch:
    """One request: edge insertions (u, v, w) and deletions (u, v) of undirected pairs.

    Vertex ids lie in [0, n) and insertion weights are finite and positive (both checked when
    the batch is applied).  Deletions apply before insertions; a deletion of an absent edge is
    counted as missing, an insertion of a present edge replaces its weight, and self-loops are
    ignored, as in the system built from the edge set (its off-diagonal part only).  An
    unordered pair repeated within the deletions is deleted once; one repeated within the
    insertions, in either orientation, is inserted once with the weight of its last occurrence
    (`normalize_batch`)."""
    ins_u: np.ndarray
    ins_v: np.ndarray
    ins_w: np.ndarray
    del_u: np.ndarray
    del_v: np.ndarray

    @property
    def size(self) -> int:
        return int(self.ins_u.shape[0] + self.del_u.shape[0])


def _id_arrays(b: EdgeBatch):
    """The four id arrays, C-contiguous, of one integer width (int32 kept, anything else int64)."""
    arrs = [np.asarray(a) for a in (b.del_u, b.del_v, b.ins_u, b.ins_v)]
    dt = np.int32 if all(a.dtype == np.int32 for a in arrs) else np.int64
    return [np.ascontiguousarray(a, dtype=dt) for a in arrs], np.dtype(dt).itemsize


def _scan_batch(b: EdgeBatch, n: int, max_insertions: int | None) -> int:
    """The checks of check_batch; returns a mask of 1 (a pair repeats within the deletions) and
    2 (within the insertions).  Ids, weights and repeats are scanned in one pass of the library's
    host-side check (dlp_check_edge_batch), no device work."""
    from . import _lib
    nd, ni = len(b.del_u), len(b.ins_u)
    if len(b.del_v) != nd or len(b.ins_v) != ni or len(b.ins_w) != ni:
        raise ValueError(f"batch arrays differ in length: del {nd}/{len(b.del_v)}, "
                         f"ins {ni}/{len(b.ins_v)}/{len(b.ins_w)}")
    if max_insertions is not None and ni > max_insertions:
        raise ValueError(f"{ni} insertions in one batch: at most {max_insertions} (split the batch)")
    ids, width = _id_arrays(b)
    w = np.asarray(b.ins_w)
    with np.errstate(over="ignore", invalid="ignore"):
        w32 = np.ascontiguousarray(w, dtype=np.float32)
    where = ct.c_int64(-1)
    P = ct.c_void_p
    rc = _lib.load().dlp_check_edge_batch(int(n), nd, P(ids[0].ctypes.data), P(ids[1].ctypes.data), ni,
                                           P(ids[2].ctypes.data), P(ids[3].ctypes.data), P(w32.ctypes.data),
                                           width, ct.byref(where))
    if rc == -1:
        a = ids[[0, 0, 1, 2, 3][np.searchsorted(np.cumsum([0, nd, nd, ni, ni]), where.value, side="right")]]
        raise IndexError(f"batch vertex ids must lie in [0, {n}): found {int(a.min())}..{int(a.max())}")
    if rc == -2:
        i = int(where.value)
        raise ValueError(f"insertion weights must be finite and positive (as FP32): ins_w[{i}] = {w[i]!r}")
    if rc < 0:
        raise ValueError(f"batch check refused its arguments (n = {n})")
    return int(rc)


def check_batch(b: EdgeBatch, n: int, max_insertions: int | None = None) -> None:
    """Refuses a batch whose arrays differ in length, that holds more than max_insertions
    insertions or an insertion weight that is not finite and positive (ValueError), or whose
    vertex ids leave [0, n) (IndexError): the device apply indexes rows by these ids."""
    _scan_batch(b, n, max_insertions)


def pack_ops(batches):
    """The ordered operation sequence of `batches` applied one after the other: per batch its
    deletions (weight 0) then its insertions (their weights), as int32 u, v and float32 w (the
    layout of dlp_lp_apply_ops, whose result is each pair's last operation)."""
    us, vs, ws = [], [], []
    for b in batches:
        us += [np.asarray(b.del_u), np.asarray(b.ins_u)]
        vs += [np.asarray(b.del_v), np.asarray(b.ins_v)]
        ws += [np.zeros(len(b.del_u), dtype=np.float32), np.asarray(b.ins_w, dtype=np.float32)]
    if not us:
        return np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float32)
    return (np.concatenate(us).astype(np.int32, copy=False), np.concatenate(vs).astype(np.int32, copy=False),
            np.concatenate(ws))


def _last_unique(key):
    """Position of the last occurrence of each distinct key (ordered by key)."""
    rev = key[::-1]
    _, first_in_rev = np.unique(rev, return_index=True)
    return key.shape[0] - 1 - first_in_rev


def normalize_batch(b: EdgeBatch, n: int, max_insertions: int | None = None) -> EdgeBatch:
    """check_batch, then one entry per unordered pair: repeated deletions collapse to one,
    repeated insertions (either orientation) to their last occurrence.  Returns b itself when
    nothing repeats.  Self-loops pass through (the apply ignores them)."""
    rep = _scan_batch(b, n, max_insertions)
    if not rep:
        return b
    out = {}
    for bit, kind, (u, v) in ((1, "del", (b.del_u, b.del_v)), (2, "ins", (b.ins_u, b.ins_v))):
        if rep & bit:
            u64, v64 = np.asarray(u, dtype=np.int64), np.asarray(v, dtype=np.int64)
            key = np.minimum(u64, v64) * n + np.maximum(u64, v64)
            out[kind] = np.sort(_last_unique(key))
    dk = out.get("del")
    ik = out.get("ins")
    return EdgeBatch(b.ins_u if ik is None else np.asarray(b.ins_u)[ik],
                     b.ins_v if ik is None else np.asarray(b.ins_v)[ik],
                     b.ins_w if ik is None else np.asarray(b.ins_w)[ik],
                     b.del_u if dk is None else np.asarray(b.del_u)[dk],
                     b.del_v if dk is None else np.asarray(b.del_v)[dk])


class EdgeSet:
    """Host mirror of a graph's edge set as sorted int64 keys u*n+v (u < v), for drawing
    updates (deletions need existing edges, insertions need non-edges)."""

    def __init__(self, g: Graph):
        self.n = g.n
        key = g.u * g.n + g.v
        order = np.argsort(key)
        self.key = key[order]
        self.w = g.w[order]

    def contains(self, key):
        pos = np.searchsorted(self.key, key)
        pos = np.minimum(pos, self.key.shape[0] - 1)
        return (self.key.shape[0] > 0) & (self.key[pos] == key)

    def apply(self, b: EdgeBatch):
        """Deletions, then insertions (upserts; of a pair inserted twice the last weight), as
        merges into the sorted keys: O(m) per batch plus a sort of the batch."""
        n = self.n
        dk = np.minimum(b.del_u, b.del_v).astype(np.int64) * n + np.maximum(b.del_u, b.del_v)
        if dk.shape[0] and self.key.shape[0]:
            pos = np.searchsorted(self.key, dk)
            pos = pos[pos < self.key.shape[0]]
            pos = pos[np.isin(self.key[pos], dk)]
            if pos.shape[0]:
                keep = np.ones(self.key.shape[0], dtype=bool)
                keep[pos] = False
                self.key, self.w = self.key[keep], self.w[keep]
        ik = np.minimum(b.ins_u, b.ins_v).astype(np.int64) * n + np.maximum(b.ins_u, b.ins_v)
        if ik.shape[0]:
            iw = np.asarray(b.ins_w, dtype=self.w.dtype)
            # the last occurrence of each inserted pair
            last = ik.shape[0] - 1 - np.unique(ik[::-1], return_index=True)[1]
            ik, iw = ik[last], iw[last]          # sorted by key
            have = self.contains(ik)
            if have.any():   # upsert: an existing pair takes the new weight
                self.w[np.searchsorted(self.key, ik[have])] = iw[have]
            nk, nw = ik[~have], iw[~have]
            if nk.shape[0]:
                at = np.searchsorted(self.key, nk)
                self.key = np.insert(self.key, at, nk)
                self.w = np.insert(self.w, at, nw)

    def pairs(self):
        return self.key // self.n, self.key % self.n, self.w


def draw_batch(es: EdgeSet, size, ins_frac=0.5, rng=None, wlo=0.5, whi=1.5, unit=False,
               del_pool=None) -> EdgeBatch:
    """`size` updates: round(size*ins_frac) insertions of uniform new non-edge pairs (rejection
    sampled) and the rest deletions of uniform existing edges (without replacement).
    del_pool: optional bool mask over es.key restricting which edges may be deleted; when it
    runs out, deletions fall back to any edge."""
    n = es.n
    n_ins = int(round(size * ins_frac))
    n_del = size - n_ins
    # deletions
    if del_pool is not None and del_pool.sum() >= n_del:
        cand = np.flatnonzero(del_pool)
    else:
        cand = None
    if cand is not None:
        pick = cand[rng.choice(cand.shape[0], size=n_del, replace=False)]
    else:
        pick = rng.choice(es.key.shape[0], size=min(n_del, es.key.shape[0]), replace=False)
    dk = es.key[pick]
    # insertions: rejection-sample pairs that are neither edges nor already drawn
    got = np.zeros(0, dtype=np.int64)
    while got.shape[0] < n_ins:
        need = n_ins - got.shape[0]
        a = rng.integers(0, n, size=2 * need + 8, dtype=np.int64)
        b = rng.integers(0, n, size=2 * need + 8, dtype=np.int64)
        ok = a != b
        k = np.minimum(a, b)[ok] * n + np.maximum(a, b)[ok]
        k = k[~es.contains(k)]
        k = k[~np.isin(k, dk)]
        _, first = np.unique(np.concatenate([got, k]), return_index=True)
        allk = np.concatenate([got, k])
        first.sort()
        got = allk[first][:n_ins] if first.shape[0] >= n_ins else allk[first]
    w = np.ones(n_ins, dtype=np.float32) if unit else rng.uniform(wlo, whi, size=n_ins).astype(np.float32)
    return EdgeBatch(got // n, got % n, w, dk // n, dk % n)


def update_stream(g: Graph, n_batches, batch_size, ins_frac=0.5, seed=0, unit=None,
                  inter_first=False, snapshot_after=None):
    """Pre-drawn batches for graph g (seed S+2).  inter_first: deletions are drawn from
    inter-block edges while any remain, then from all edges (fragmenting stream).  Returns the
    batches and the edge set after the last one; with snapshot_after = k, also a copy of the
    edge set after the first k batches."""
    import copy
    rng = np.random.default_rng(seed + 2)
    es = EdgeSet(g)
    unit = bool(np.all(g.w == 1)) if unit is None else unit
    out, snap = [], None
    for i in range(n_batches):
        pool = None
        if inter_first and g.block is not None:
            uu, vv = es.key // es.n, es.key % es.n
            pool = g.block[uu] != g.block[vv]
        b = draw_batch(es, batch_size, ins_frac, rng, unit=unit, del_pool=pool)
        es.apply(b)
        out.append(b)
        if snapshot_after == i + 1:
            snap = copy.copy(es)
            snap.key, snap.w = es.key.copy(), es.w.copy()
    if snapshot_after is not None:
        return out, es, snap
    return out, es

"""Continuous-batching update service over many device-resident graphs.

No reference counterpart (the reference times one solver per batch, bench.py:93-118); this is
the serving layer the workload asks for:

- Residency and admission.  Every graph lives on the device as an `EdgeStreamLP` (slack CSR,
  FP32 scores and kept residual, FP64 diag).  A graph is admitted when its estimated footprint
  fits the HBM budget; otherwise it waits, and idle graphs are evicted least-recently-used to
  make room.  On the host each graph keeps a checkpoint of its edge set (taken from the device
  store once the batches served since the last one hold about m operations) and those batches;
  an evicted graph is re-admitted by a solve from the checkpoint with the batches applied.
- Continuous batching.  Requests (edge-update batches for one graph) queue per graph.  The
  scheduler takes the requests queued for a graph at that moment (up to max_fuse, and while
  their operations fit 2^21) and repairs them as one group: their deletions and insertions go
  to the device as one ordered sequence (dlp_lp_apply_ops), where every pair ends as its last
  operation leaves it, which is what applying the requests in arrival order gives, with no
  host-side merge.  Requests that arrive while the graph is busy are served by its next repair.
- Fused push and overlap.  One scheduler thread takes, per cycle, the queued requests of every
  idle resident graph and runs the first half of each repair (apply, re-ground, changed-row
  refresh).  The graphs whose change a push can absorb are then repaired by ONE push across
  them (dlp_lp_push_multi: per round one launch per column class covering every such graph,
  replayed from a captured hipGraph), and verified; the others (large changes, or a push that
  spent its edge budget) go to `slots` worker threads for a global solve, each on its own HIP
  stream (default priority).  The C ABI releases the GIL, so global solves of different graphs
  overlap on the device with each other and with the next cycle's push.
- Failures.  A request whose batch is malformed (ragged arrays, ids outside the graph) is
  refused on its own, before fusing; a repair that raises fails the requests fused into it and
  drops the graph's device state, which is rebuilt from the checkpoint and the batches served
  since (the failed ones not among them) when the graph's next request is taken.  Either way the request carries `error`.
- Latency is arrival -> certified result; throughput counts edge updates.
"""
from __future__ import annotations

import ctypes as ct
import itertools
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from .graphs import EdgeBatch, Graph, check_batch
from .stream import EdgeStreamLP

# operations (deletions + insertions) of the requests fused into one repair; a single larger
# request is served alone (the device apply takes it in calls of 2^19 operations)
_MAX_GROUP_OPS = 1 << 21


@dataclass
class Request:
    gid: int
    batch: EdgeBatch
    arrival: float                 # seconds (service clock)
    done: float = float("nan")
    rho: float = float("nan")
    fused_with: int = 1
    error: str | None = None       # why the request was refused or its repair failed

    @property
    def latency_ms(self):
        return (self.done - self.arrival) * 1e3


def fuse(batches):
    """One batch equivalent to applying `batches` in order: per undirected pair the last
    operation wins (deletion, or insertion with its weight; within a batch deletions come
    first).  A single batch goes through the same pass, so a request changes its graph in the
    same way whether it is served alone or fused with others."""
    keys, ops, ws = [], [], []
    n = max(int(max(b.ins_u.max(initial=0), b.ins_v.max(initial=0), b.del_u.max(initial=0),
                    b.del_v.max(initial=0))) for b in batches) + 1
    for t, b in enumerate(batches):
        lo, hi = np.minimum(b.del_u, b.del_v), np.maximum(b.del_u, b.del_v)
        keys.append(lo.astype(np.int64) * n + hi)
        ops.append(np.zeros(lo.shape[0], dtype=np.int8))
        ws.append(np.zeros(lo.shape[0], dtype=np.float32))
        lo, hi = np.minimum(b.ins_u, b.ins_v), np.maximum(b.ins_u, b.ins_v)
        keys.append(lo.astype(np.int64) * n + hi)
        ops.append(np.ones(lo.shape[0], dtype=np.int8))
        ws.append(b.ins_w.astype(np.float32))
    k, o, w = np.concatenate(keys), np.concatenate(ops), np.concatenate(ws)
    # last occurrence per key: reverse, unique first-index, map back
    rk = k[::-1]
    _, first = np.unique(rk, return_index=True)
    last = k.shape[0] - 1 - first
    k, o, w = k[last], o[last], w[last]
    ins, dele = o == 1, o == 0
    return EdgeBatch(k[ins] // n, k[ins] % n, w[ins], k[dele] // n, k[dele] % n)


def footprint_bytes(g: Graph, **lp_kwargs):
    """Device bytes of one resident graph with its global solve reserved (EdgeStreamLP.footprint:
    store with slack, state, direction slots, coarse operator, handle scratch)."""
    return EdgeStreamLP.footprint(g, **lp_kwargs)


@dataclass
class Slot:
    gid: int
    graph: Graph
    lp: EdgeStreamLP | None = None
    reserved: bool = False        # its footprint counts against the budget
    queue: deque = field(default_factory=deque)
    busy: bool = False
    last_used: float = 0.0
    bytes: int = 0
    served: int = 0
    solves: int = 0
    base: Graph | None = None     # checkpoint of the edge set (None: the graph as given)
    history: list = field(default_factory=list)   # batches served since the checkpoint
    hist_ops: int = 0


class Service:
    """Request loop over resident graphs (see the module docstring).

    One scheduler thread takes, in every cycle, the queued requests of each idle resident graph
    (fused per graph), runs the first half of each repair (apply, re-ground, changed-row
    refresh), then repairs every graph whose change a push can absorb with ONE push across
    those graphs (dlp_lp_push_multi: one launch per round and column class, replayed from a
    captured hipGraph), and hands the graphs that need a global solve to `slots` worker
    threads, each on its own HIP stream, so global solves of different graphs overlap on the
    device and with the next cycle.  Admission (and the rebuild of an evicted graph) happens in
    the workers too."""

    def __init__(self, graphs, hbm_budget_bytes, slots=4, eps=1e-6, max_fuse=64, lp_kwargs=None,
                 fused_push=True, push_rounds_per_launch=4, push_max_rounds=200_000, progress=None):
        self.slots_n = max(1, int(slots))
        self.budget = int(hbm_budget_bytes)
        self.eps = eps
        self.max_fuse = max_fuse
        self.lp_kwargs = dict(lp_kwargs or {})
        self.fused_push = bool(fused_push)
        self.progress = progress      # progress(kind, detail): preload steps and every 5 % of requests served
        self.push_rpl, self.push_max_rounds = int(push_rounds_per_launch), int(push_max_rounds)
        self.graphs = {i: Slot(i, g) for i, g in enumerate(graphs)}
        for s in self.graphs.values():
            s.bytes = footprint_bytes(s.graph, **self.lp_kwargs)
            if s.bytes > self.budget:
                raise ValueError(f"graph {s.gid} needs {s.bytes / 1e9:.2f} GB > budget {self.budget / 1e9:.2f} GB")
        self.resident = 0
        self.lock = threading.Lock()
        self.cv = threading.Condition(self.lock)
        self.done = []
        self.stats = {"admissions": 0, "evictions": 0, "fused_requests": 0, "repairs": 0, "refused": 0,
                      "failed_repairs": 0, "cycles": 0, "push_launches": 0, "push_graphs": 0,
                      "global_repairs": 0}
        # seconds spent, summed over threads: host checks of the requests, first halves (apply,
        # ground, refresh), the pushes across graphs, verification after a push, global repairs
        self.times = {"checks": 0.0, "begin": 0.0, "push": 0.0, "verify": 0.0, "global": 0.0, "build": 0.0}
        self._jobs = deque()          # global repairs and builds for the workers
        self._running = 0             # jobs the workers are on
        self._plan_obj, self._plan_key = None, None
        self._n_req = 0
        self._serials = itertools.count()
        self._t0 = None
        self._stop = False

    # -- residency --------------------------------------------------------------------------
    def _new_state(self, s: Slot):
        """A state built from the checkpoint, solved, with the batches served since then applied
        (in order, as one repair): an evicted or failed graph re-enters with every update it
        had served."""
        lp = EdgeStreamLP(s.base if s.base is not None else s.graph, eps=self.eps, **self.lp_kwargs)
        lp.initial_solve()
        if s.history:
            lp.apply_group_and_solve(s.history)
        s.solves += 1
        lp.serial = next(self._serials)
        return lp

    def _served(self, s: Slot, batches):
        """Record batches served on s; past about m operations since the checkpoint, the device
        edge set becomes the new checkpoint (host memory stays O(m) per graph).  Called by the
        thread that repaired s, while s is busy."""
        s.history.extend(batches)
        s.hist_ops += sum(b.size for b in batches)
        if s.hist_ops > max(1 << 20, s.graph.m) and s.lp is not None:
            s.base = s.lp.edge_graph()
            s.history, s.hist_ops = [], 0

    def _admit(self, s: Slot):
        """Called with the lock held: True when s is resident, "build" when it is admitted now
        (memory reserved; the caller builds its state), False when it does not fit yet (idle
        graphs are evicted least-recently-used to make room)."""
        if s.lp is not None:
            return True
        if s.reserved:
            return False            # being built
        while self.resident + s.bytes > self.budget:
            idle = [x for x in self.graphs.values() if x.lp is not None and not x.busy and not x.queue]
            if not idle:
                return False
            victim = min(idle, key=lambda x: x.last_used)
            self._drop_plans()      # a plan keeps its graphs' states alive
            victim.lp = None        # frees its device buffers (handle destroyed with the object)
            victim.reserved = False
            self.resident -= victim.bytes
            self.stats["evictions"] += 1
        s.reserved = True
        self.resident += s.bytes
        self.stats["admissions"] += 1
        return "build"

    def _take(self, s: Slot):
        """Queued requests of s, in arrival order, while their operations fit one group."""
        reqs, ops = [], 0
        while s.queue and len(reqs) < self.max_fuse:
            k = s.queue[0].batch.size
            if reqs and ops + k > _MAX_GROUP_OPS:
                break
            reqs.append(s.queue.popleft())
            ops += k
        return reqs

    def _plan(self):
        """The push plan of the current resident set, rebuilt when the set changes.  The plan
        holds a reference to each member's state, so a state dropped elsewhere stays valid
        until its plan goes (the multi push reads nothing of the graphs outside its call)."""
        lps = {s.gid: s.lp for s in self.graphs.values() if s.lp is not None}
        key = tuple(sorted((g, lp.serial) for g, lp in lps.items()))
        if self._plan_key != key:
            from . import _lib
            self._drop_plans()
            order = [g for g, _ in key]
            hs = (ct.c_void_p * len(order))(*[lps[g].h.value for g in order])
            h = ct.c_void_p()
            _lib.call("dlp_push_plan_create", hs, len(order), ct.byref(h))
            self._plan_obj, self._plan_key = (h, order, [lps[g] for g in order]), key
        return self._plan_obj[:2]

    def _drop_plans(self):
        if self._plan_obj is not None:
            from . import _lib
            _lib.call("dlp_push_plan_destroy", self._plan_obj[0])
        self._plan_obj, self._plan_key = None, None

    # -- completion ---------------------------------------------------------------------------
    def _complete(self, s: Slot, good, bad, st, err):
        held = None
        if err is None and good:
            try:
                self._served(s, [r.batch for r in good])
                held = s.lp.resident_bytes()     # a regrown store holds more than the estimate
            except Exception as e:  # noqa: BLE001  (checkpoint failed: the state is suspect)
                err = f"checkpoint failed: {type(e).__name__}: {e}"
        now = time.perf_counter() - self._t0
        with self.cv:
            if held is not None and s.reserved and held > s.bytes:
                self.resident += held - s.bytes
                s.bytes = held
            for r in bad:
                r.done = now
            if err is not None:
                for r in good:
                    r.done, r.error = now, err
                if s.reserved:
                    # its state is unknown: dropped, rebuilt on next use
                    s.lp = None
                    s.reserved = False
                    self.resident -= s.bytes
                self.stats["failed_repairs"] += 1
            else:
                for r in good:
                    r.done, r.rho, r.fused_with = now, st.rho if st is not None else float("nan"), len(good)
                s.served += len(good)
                self.stats["fused_requests"] += len(good)
                self.stats["repairs"] += int(bool(good))
            before = len(self.done)
            self.done.extend(good + bad)
            if self.progress and self._n_req and (len(self.done) * 20) // self._n_req > (before * 20) // self._n_req:
                self.progress("served", f"{len(self.done)}/{self._n_req} requests at {now:.2f}s, "
                                        f"{self.stats['repairs']} repairs, {self.stats['cycles']} cycles")
            self.stats["refused"] += len(bad)
            s.busy = False
            s.last_used = now
            self.cv.notify_all()

    @staticmethod
    def _split(s: Slot, reqs):
        bad = []
        for r in reqs:
            try:
                check_batch(r.batch, s.graph.n)
            except (IndexError, ValueError) as e:
                r.error = f"refused: {e}"
                bad.append(r)
        return [r for r in reqs if r.error is None], bad

    # -- workers: builds and global repairs -----------------------------------------------------
    def _worker(self, wid):
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            while True:
                with self.cv:
                    while not self._jobs:
                        if self._stop and self._idle_locked():
                            return
                        self.cv.wait(timeout=0.05)
                    job = self._jobs.popleft()
                    self._running += 1
                kind, s, good, bad, req = job
                st, err = None, None
                t_job = time.perf_counter()
                try:
                    if kind == "build":
                        s.lp = self._new_state(s)
                        if good:
                            st = s.lp.apply_group_and_solve([r.batch for r in good])
                    else:
                        st = s.lp.end_request(req, pushed=req.get("pushed"))
                    stream.synchronize()
                except Exception as e:  # noqa: BLE001  (reported on the requests, service goes on)
                    err = f"repair failed: {type(e).__name__}: {e}"
                dt = time.perf_counter() - t_job
                self._complete(s, good, bad, st, err)
                with self.cv:   # workers share these counters
                    self.times["build" if kind == "build" else "global"] += dt
                    if kind != "build":
                        self.stats["global_repairs"] += 1
                    self._running -= 1
                    self.cv.notify_all()

    def _idle_locked(self):
        return (not any(x.queue for x in self.graphs.values()) and not any(x.busy for x in self.graphs.values())
                and not self._jobs and self._running == 0)

    # -- scheduler: one cycle fuses, begins, pushes across graphs, hands off global solves ------
    def _scheduler(self):
        from . import _lib
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            while True:
                picks = []
                with self.cv:
                    while True:
                        if self._stop and self._idle_locked():
                            self.cv.notify_all()
                            return
                        cands = sorted((x for x in self.graphs.values() if x.queue and not x.busy),
                                       key=lambda x: x.queue[0].arrival)
                        for s in cands:
                            a = self._admit(s)
                            if not a:
                                continue
                            s.busy = True
                            reqs = self._take(s)
                            if a == "build":
                                good, bad = self._split(s, reqs)
                                self._jobs.append(("build", s, good, bad, None))
                            else:
                                picks.append((s, reqs))
                        if picks:
                            break
                        self.cv.notify_all()
                        self.cv.wait(timeout=0.05)
                self.stats["cycles"] += 1
                begun = []
                for s, reqs in picks:
                    t = time.perf_counter()
                    good, bad = self._split(s, reqs)
                    if not good:
                        self._complete(s, good, bad, None, None)
                        continue
                    t1 = time.perf_counter()
                    self.times["checks"] += t1 - t
                    try:
                        begun.append((s, good, bad, s.lp.begin_group([r.batch for r in good])))
                    except Exception as e:  # noqa: BLE001
                        self._complete(s, good, bad, None, f"repair failed: {type(e).__name__}: {e}")
                    self.times["begin"] += time.perf_counter() - t1
                stream.synchronize()
                pushers = [x for x in begun if self.fused_push and x[0].lp.push_planned(x[3]["status"])]
                with self.cv:
                    for x in begun:
                        if x not in pushers:
                            self._jobs.append(("repair",) + x)
                    self.cv.notify_all()
                if not pushers:
                    continue
                t = time.perf_counter()
                try:
                    plan, order = self._plan()
                    pos = {g: i for i, g in enumerate(order)}
                    nh = len(order)
                    active = np.zeros(nh, dtype=np.int32)
                    budget = np.zeros(nh, dtype=np.int64)
                    for s, _, _, _ in pushers:
                        active[pos[s.gid]] = 1
                        budget[pos[s.gid]] = s.lp.push_budget_edges()
                    rounds = np.zeros(nh, dtype=np.int64)
                    conv = np.zeros(nh, dtype=np.int32)
                    _lib.call("dlp_lp_push_multi", plan, active.ctypes.data, self.push_max_rounds, self.push_rpl, 1,
                              budget.ctypes.data, rounds.ctypes.data, conv.ctypes.data, stream.cuda_stream)
                    self.stats["push_launches"] += 1
                    self.stats["push_graphs"] += len(pushers)
                    self.times["push"] += time.perf_counter() - t
                except Exception as e:  # noqa: BLE001
                    for s, good, bad, _ in pushers:
                        self._complete(s, good, bad, None, f"repair failed: {type(e).__name__}: {e}")
                    continue
                # converged graphs are verified here; the rest (edge budget spent) go on to a
                # global solve in a worker
                for x in pushers:
                    s, good, bad, req = x
                    i = pos[s.gid]
                    req["pushed"] = (int(rounds[i]), bool(conv[i]))
                    if not conv[i]:
                        with self.cv:
                            self._jobs.append(("repair",) + x)
                            self.cv.notify_all()
                        continue
                    st, err = None, None
                    t = time.perf_counter()
                    try:
                        st = s.lp.end_request(req, pushed=req["pushed"])
                        stream.synchronize()
                    except Exception as e:  # noqa: BLE001
                        err = f"repair failed: {type(e).__name__}: {e}"
                    self.times["verify"] += time.perf_counter() - t
                    self._complete(s, good, bad, st, err)

    # -- driver ----------------------------------------------------------------------------
    def run(self, requests, preload=True):
        """Serve `requests` (list of Request with arrival times in seconds from start, sorted)
        open-loop; returns the completed requests.  preload: admit and solve every graph that
        fits before the clock starts."""
        self._n_req = len(requests)
        if preload:
            for s in sorted(self.graphs.values(), key=lambda x: x.bytes):
                if self.resident + s.bytes <= self.budget:
                    t = time.perf_counter()
                    s.lp = self._new_state(s)
                    s.reserved = True
                    self.resident += s.bytes
                    self.stats["admissions"] += 1
                    if self.progress:
                        self.progress("preload", f"graph {s.gid} n={s.graph.n} k={s.graph.K}: "
                                                 f"{(time.perf_counter() - t) * 1e3:.0f} ms")
        torch.cuda.synchronize()
        self._t0 = time.perf_counter()
        self._stop = False
        threads = [threading.Thread(target=self._scheduler, daemon=True)]
        threads += [threading.Thread(target=self._worker, args=(i,), daemon=True) for i in range(self.slots_n)]
        for t in threads:
            t.start()
        pending = sorted(requests, key=lambda r: r.arrival)
        i = 0
        while i < len(pending):
            now = time.perf_counter() - self._t0
            r = pending[i]
            if r.arrival > now:
                time.sleep(min(r.arrival - now, 0.002))
                continue
            with self.cv:
                while i < len(pending) and pending[i].arrival <= now:
                    self.graphs[pending[i].gid].queue.append(pending[i])
                    i += 1
                self.cv.notify_all()
        with self.cv:
            self._stop = True
            self.cv.notify_all()
        for t in threads:
            t.join()
        self.wall = time.perf_counter() - self._t0
        self._drop_plans()
        return self.done

    def report(self):
        ok = [r for r in self.done if r.error is None]
        lat = np.array([r.latency_ms for r in ok])
        edges = sum(r.batch.size for r in ok)
        pl = self.stats["push_launches"]
        return {"requests": len(ok), "errors": len(self.done) - len(ok), "edge_updates": edges,
                "edge_updates_per_s": edges / self.wall if self.wall > 0 else float("nan"),
                "requests_per_s": len(ok) / self.wall if self.wall > 0 else float("nan"),
                "p50_ms": float(np.percentile(lat, 50)) if lat.size else float("nan"),
                "p99_ms": float(np.percentile(lat, 99)) if lat.size else float("nan"),
                "max_rho": max((r.rho for r in ok), default=float("nan")),
                "mean_fused": float(np.mean([r.fused_with for r in ok])) if ok else 0.0,
                "graphs_per_push_launch": self.stats["push_graphs"] / pl if pl else 0.0,
                **self.stats, "seconds": {k: round(v, 3) for k, v in self.times.items()}}


def poisson_requests(graphs, rate, duration, size_lo=100, size_hi=100_000, seed=0, max_requests=None):
    """Open-loop arrivals at `rate` req/s over `duration` s (seed S+3): graph uniform, batch size
    log-uniform in [size_lo, size_hi], 50/50 insert/delete, drawn from each graph's own edge set."""
    from .graphs import EdgeSet, draw_batch
    rng = np.random.default_rng(seed + 3)
    sets = [EdgeSet(g) for g in graphs]
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration or (max_requests and len(out) >= max_requests):
            break
        gid = int(rng.integers(0, len(graphs)))
        size = int(np.exp(rng.uniform(np.log(size_lo), np.log(size_hi))))
        b = draw_batch(sets[gid], size, 0.5, rng, unit=bool(np.all(graphs[gid].w == 1)))
        sets[gid].apply(b)
        out.append(Request(gid, b, t))
    return out


def stream_requests(graphs, rate, duration, size_lo=100, size_hi=100_000, seed=0, max_requests=None):
    """Open-loop arrivals at `rate` req/s over `duration` s (seed S+4) for workloads too large
    for `poisson_requests`' exact sampling: graph uniform, size log-uniform in [size_lo,
    size_hi], half insertions of uniform pairs (u != v, weights as the graph's: 1 or
    U[0.5, 1.5]) and half deletions.  Each graph keeps its present edges in a FIFO: a deletion
    takes the oldest (the initial edges in a random order first, then the inserted ones in
    insertion order) and an insertion joins the back, so every deletion names an edge present
    at that point of the stream (a uniform insertion that hits a present pair, probability
    about deg/n, makes that pair's second deletion a missing one)."""
    rng = np.random.default_rng(seed + 4)
    arr, gids, sizes = [], [], []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration or (max_requests and len(arr) >= max_requests):
            break
        arr.append(t)
    nr = len(arr)
    gids = rng.integers(0, len(graphs), size=nr)
    sizes = np.exp(rng.uniform(np.log(size_lo), np.log(size_hi), size=nr)).astype(np.int64)
    n_ins = sizes // 2
    n_del = sizes - n_ins
    out = [None] * nr
    for gid, g in enumerate(graphs):
        mine = np.flatnonzero(gids == gid)
        if not mine.size:
            continue
        tot_ins = int(n_ins[mine].sum())
        m = g.u.shape[0]
        perm = rng.permutation(m)
        fu = np.empty(m + tot_ins, dtype=np.int32)
        fv = np.empty(m + tot_ins, dtype=np.int32)
        fu[:m], fv[:m] = g.u[perm], g.v[perm]
        a = rng.integers(0, g.n, size=tot_ins, dtype=np.int64)
        b = rng.integers(0, g.n - 1, size=tot_ins, dtype=np.int64)
        b = b + (b >= a)                              # b != a
        fu[m:], fv[m:] = a, b
        unit = bool(np.all(g.w == 1))
        ws = np.ones(tot_ins, np.float32) if unit else rng.uniform(0.5, 1.5, size=tot_ins).astype(np.float32)
        head, tail = 0, m                             # FIFO [head, tail)
        for r in mine:
            nd = int(min(n_del[r], tail - head))
            ni = int(n_ins[r])
            b_ = EdgeBatch(fu[tail:tail + ni], fv[tail:tail + ni], ws[tail - m:tail - m + ni],
                           fu[head:head + nd], fv[head:head + nd])
            head += nd
            tail += ni
            out[r] = Request(gid, b_, float(arr[r]))
    return out

"""Continuous-batching update service over many device-resident graphs.

No reference counterpart (the reference times one solver per batch, bench.py:93-118); this is
the serving layer the workload asks for:

- Residency and admission.  Every graph lives on the device as an `EdgeStreamLP` (slack CSR,
  FP32 scores and kept residual, FP64 diag).  A graph is admitted when its estimated footprint
  fits the HBM budget; otherwise it waits, and idle graphs are evicted least-recently-used to
  make room (an evicted graph keeps its host-side edge set and is re-admitted with a fresh
  solve).
- Continuous batching.  Requests (edge-update batches for one graph) queue per graph.  A worker
  takes the requests queued for a graph at that moment (up to max_fuse, and while their
  insertions fit one device apply call, 2^19) and fuses them into one batch (per pair, the
  last operation wins, which is what applying them in arrival order gives), so requests that
  arrive while the graph is busy are served by its next repair.
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
  drops the graph's device state, which is rebuilt (initial solve plus the updates already
  served) when the graph's next request is taken.  Either way the request carries `error`.
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

# the device apply takes at most 2^19 insertions per call: requests are fused up to that many,
# and a replayed history is applied in chunks of half of it
_MAX_INSERTIONS = 1 << 19
_REPLAY_CHUNK = 1 << 18


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


def replay_chunks(b: EdgeBatch, chunk=_REPLAY_CHUNK):
    """A fused batch (every pair once, as a deletion or an insertion) split into batches of at
    most `chunk` operations; since no pair repeats, applying them in any order gives b."""
    ops = [(b.del_u[i:i + chunk], b.del_v[i:i + chunk]) for i in range(0, len(b.del_u), chunk)]
    out = [EdgeBatch(b.ins_u[:0], b.ins_v[:0], b.ins_w[:0], du, dv) for du, dv in ops]
    out += [EdgeBatch(b.ins_u[i:i + chunk], b.ins_v[i:i + chunk], b.ins_w[i:i + chunk], b.del_u[:0], b.del_v[:0])
            for i in range(0, len(b.ins_u), chunk)]
    return out


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
    history: list = field(default_factory=list)   # batches served, replayed after an eviction


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
                 fused_push=True, push_rounds_per_launch=4, push_max_rounds=200_000):
        self.slots_n = max(1, int(slots))
        self.budget = int(hbm_budget_bytes)
        self.eps = eps
        self.max_fuse = max_fuse
        self.lp_kwargs = dict(lp_kwargs or {})
        self.fused_push = bool(fused_push)
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
        self._jobs = deque()          # global repairs and builds for the workers
        self._running = 0             # jobs the workers are on
        self._plan_obj, self._plan_key = None, None
        self._serials = itertools.count()
        self._t0 = None
        self._stop = False

    # -- residency --------------------------------------------------------------------------
    def _new_state(self, s: Slot):
        lp = EdgeStreamLP(s.graph, eps=self.eps, **self.lp_kwargs)
        lp.initial_solve()
        s.solves += 1
        lp.serial = next(self._serials)
        return lp

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
        """Queued requests of s fused while their insertions fit one device apply call."""
        reqs, ins = [], 0
        while s.queue and len(reqs) < self.max_fuse:
            k = len(s.queue[0].batch.ins_u)
            if reqs and ins + k > _MAX_INSERTIONS:
                break
            reqs.append(s.queue.popleft())
            ins += k
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
    def _complete(self, s: Slot, good, bad, b, st, err):
        now = time.perf_counter() - self._t0
        with self.cv:
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
                if b is not None:
                    s.history.append(b)
                    if len(s.history) > 32:     # keep the replay log one fused batch
                        s.history = [fuse(s.history)]
                s.served += len(good)
                self.stats["fused_requests"] += len(good)
                self.stats["repairs"] += int(b is not None)
            self.done.extend(good + bad)
            self.stats["refused"] += len(bad)
            s.busy = False
            s.last_used = now
            self.cv.notify_all()

    @staticmethod
    def _split(s: Slot, reqs):
        bad = []
        for r in reqs:
            try:
                check_batch(r.batch, s.graph.n, max_insertions=_MAX_INSERTIONS)
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
                kind, s, good, bad, b, req = job
                st, err = None, None
                try:
                    if kind == "build":
                        lp = self._new_state(s)
                        # an evicted graph re-enters with the updates it had already served
                        for piece in (replay_chunks(fuse(s.history)) if s.history else []):
                            lp.apply_and_solve(piece)
                        s.lp = lp
                        if good:
                            b = fuse([r.batch for r in good])
                            st = s.lp.apply_and_solve(b)
                    else:
                        st = s.lp.end_request(req, pushed=req.get("pushed"))
                    stream.synchronize()
                except Exception as e:  # noqa: BLE001  (reported on the requests, service goes on)
                    err = f"repair failed: {type(e).__name__}: {e}"
                self._complete(s, good, bad, b, st, err)
                with self.cv:
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
                                self._jobs.append(("build", s, good, bad, None, None))
                            else:
                                picks.append((s, reqs))
                        if picks:
                            break
                        self.cv.notify_all()
                        self.cv.wait(timeout=0.05)
                self.stats["cycles"] += 1
                begun = []
                for s, reqs in picks:
                    good, bad = self._split(s, reqs)
                    if not good:
                        self._complete(s, good, bad, None, None, None)
                        continue
                    b = fuse([r.batch for r in good])
                    try:
                        begun.append((s, good, bad, b, s.lp.begin_request(b)))
                    except Exception as e:  # noqa: BLE001
                        self._complete(s, good, bad, b, None, f"repair failed: {type(e).__name__}: {e}")
                stream.synchronize()
                pushers = [x for x in begun if self.fused_push and x[0].lp.push_planned(x[4]["status"])]
                with self.cv:
                    for x in begun:
                        if x not in pushers:
                            self._jobs.append(("repair",) + x)
                    self.cv.notify_all()
                if not pushers:
                    continue
                try:
                    plan, order = self._plan()
                    pos = {g: i for i, g in enumerate(order)}
                    nh = len(order)
                    active = np.zeros(nh, dtype=np.int32)
                    budget = np.zeros(nh, dtype=np.int64)
                    for s, _, _, _, _ in pushers:
                        active[pos[s.gid]] = 1
                        budget[pos[s.gid]] = s.lp.push_budget_edges()
                    rounds = np.zeros(nh, dtype=np.int64)
                    conv = np.zeros(nh, dtype=np.int32)
                    _lib.call("dlp_lp_push_multi", plan, active.ctypes.data, self.push_max_rounds, self.push_rpl, 1,
                              budget.ctypes.data, rounds.ctypes.data, conv.ctypes.data, stream.cuda_stream)
                    self.stats["push_launches"] += 1
                    self.stats["push_graphs"] += len(pushers)
                except Exception as e:  # noqa: BLE001
                    for s, good, bad, b, _ in pushers:
                        self._complete(s, good, bad, b, None, f"repair failed: {type(e).__name__}: {e}")
                    continue
                # converged graphs are verified here; the rest (edge budget spent) go on to a
                # global solve in a worker
                for x in pushers:
                    s, good, bad, b, req = x
                    i = pos[s.gid]
                    req["pushed"] = (int(rounds[i]), bool(conv[i]))
                    if not conv[i]:
                        with self.cv:
                            self._jobs.append(("repair",) + x)
                            self.cv.notify_all()
                        continue
                    st, err = None, None
                    try:
                        st = s.lp.end_request(req, pushed=req["pushed"])
                        stream.synchronize()
                    except Exception as e:  # noqa: BLE001
                        err = f"repair failed: {type(e).__name__}: {e}"
                    self._complete(s, good, bad, b, st, err)

    # -- driver ----------------------------------------------------------------------------
    def run(self, requests, preload=True):
        """Serve `requests` (list of Request with arrival times in seconds from start, sorted)
        open-loop; returns the completed requests.  preload: admit and solve every graph that
        fits before the clock starts."""
        if preload:
            for s in sorted(self.graphs.values(), key=lambda x: x.bytes):
                if self.resident + s.bytes <= self.budget:
                    s.lp = self._new_state(s)
                    s.reserved = True
                    self.resident += s.bytes
                    self.stats["admissions"] += 1
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
                **self.stats}


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

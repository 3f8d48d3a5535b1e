#!/usr/bin/env python3
"""Edge-update throughput and update latency at rho <= 1e-6 on SBM-1M with k = 64, one MI355X.

Workload: SBM n = 1,000,000, 64 equal blocks, mean degree 20 (80% intra-block), weights
U[0.5, 1.5] rounded to FP32, 0.5% labelled with label = block mod 64 (k = 64, 65 solved columns
with h).  After the initial solve, a stream of requests, each a batch of `--batch-edges` edge
updates (50% insertions of new uniform pairs, 50% deletions of existing edges; seed S+2).  A step
is one request: apply on the device-resident slack CSR, re-ground (connected components),
refresh the changed rows' FP64 residual, repair to max |R_uc|/diag_u <= 1e-6 (push, or the
global PCG on FP32 vectors with scaled FP16 search directions and a two-level preconditioner,
replayed from hipGraphs) and verify in FP64.  Latency is submit -> certified.

Printed on stdout: one JSON line.  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "edge-updates/s + p50/p99 update latency at rel. residual 1e-6, SBM-1M k=64"
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK = 78.6     # TFLOP/s, dense FP64 matrix


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def percentile(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs), q)) if xs else float("nan")


def pmc_traffic(kernel):
    """Memory traffic per launch of `kernel` from the committed counter summary (the L2's read
    and write requests to memory, counted by size), or None when it covers another kernel."""
    import re
    path = os.path.join(ROOT, "profiles", "pcg_operator_sbm1m_k64_pmc.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    k = re.search(r"^traffic_kernel: (.+)$", text, re.M)
    b = re.search(r"^traffic_bytes_per_launch: (\d+)$", text, re.M)
    if not (k and b) or k.group(1).strip() != kernel:
        return None
    return int(b.group(1))


# bytes per stored direction value
DIRECTION_BYTES = {"fp32": 4, "bf16": 2, "fp16": 2, "fp8": 1}


def operator_bytes(n, C1, stored, p_bytes=4):
    """Algorithmic bytes of one A*P launch (each distinct element once): row starts and
    lengths, column/weight of every stored entry, the solved mask, P read (p_bytes per value:
    2 for 16-bit directions, 1 for FP8), diag, AP written (FP32)."""
    return 4 * (n + 1) + 4 * n + 8 * stored + n + p_bytes * n * C1 + 4 * n + 4 * n * C1


def verify_bytes(n, n_solved, C1, stored_solved):
    """Algorithmic bytes of the FP64 verification pass (each element once): row offsets and
    lengths and the solved mask of every row, the stored entries (column + FP32 weight) of the
    solved rows, the FP32 solution rows of every row (a solved row's own and its neighbours',
    which cover the graph), the labeled mask and labels, and for the solved rows the FP64
    diagonal and the FP32 residual rows written back."""
    return 4 * (n + 1) + 4 * n + n + 8 * stored_solved + 4 * n * C1 + n + 4 * n + 8 * n_solved + 4 * n_solved * C1


def canonical_components(n, u, v):
    """Connected-component labels of the undirected edge list (u, v) in the reference's
    numbering (a component's label = the rank of its minimum vertex among the component
    minima, cc.py:8-23), by SciPy's csgraph: the independent check of the device labels."""
    import numpy as np
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    G = sp.coo_matrix((np.ones(u.shape[0], dtype=np.int8), (u, v)), shape=(n, n)).tocsr()
    nc, lab = connected_components(G, directed=False)
    first = np.full(nc, n, dtype=np.int64)
    np.minimum.at(first, lab, np.arange(n))
    rank = np.empty(nc, dtype=np.int64)
    rank[np.argsort(first)] = np.arange(nc)
    return rank[lab].astype(np.int32)


def timed(fn, reps):
    """Average ms of fn() over reps launches, HIP events on torch's current stream (the stream
    the library launches on)."""
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def kernel_table(g, C, reps=10):
    """Roofline lines of the other hot kernels at this graph's shape (C columns): the reference
    API's SpMM, frontier Jacobi and push scatter on a frontier of every 10th row, connected
    components, and the FP64 Gram of block PCG on the matrix cores.  Bytes count each distinct
    element once (reads + writes; an atomic read-modify-write twice)."""
    import numpy as np
    import torch
    from dynlp_amd import _lib
    from dynlp_amd import backend as B
    from dynlp_amd.cc import connected_components_csr
    from dynlp_amd.kernels import FrontierOps, spmm_axpy
    ip, ix, d = g.csr()
    n, nnz = g.n, int(ix.shape[0])
    W = B.csr_from_host(ip, ix, d, n, torch.float32)
    gen = torch.Generator(device="cuda").manual_seed(0)
    X = torch.rand((n, C), dtype=torch.float32, device="cuda", generator=gen)
    a = torch.rand(n, dtype=torch.float32, device="cuda", generator=gen) + 20
    rows = []

    def line(name, ms, nbytes=None, flops=None):
        if flops is not None:
            ach = flops / (ms * 1e-3) / 1e12
            rows.append({"kernel": name, "bound": "mfma", "achieved": round(ach, 2), "peak": FP64_MFMA_PEAK,
                         "unit": "TFLOP/s", "frac": round(ach / FP64_MFMA_PEAK, 4), "ms": round(ms, 4),
                         "flops": int(flops)})
        else:
            ach = nbytes / (ms * 1e-3) / 1e9
            rows.append({"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "ms": round(ms, 4),
                         "algorithmic_bytes": int(nbytes)})

    ms = timed(lambda: spmm_axpy(W, X, a=a, b=-1.0), reps)
    line(f"spmm_axpy (A*X, n x {C} FP32, reference API)", ms, 4 * (n + 1) + 8 * nnz + 2 * 4 * n * C + 4 * n)
    f_h = np.arange(0, n, 10, dtype=np.int32)
    f = torch.as_tensor(f_h, device="cuda")
    nf = f_h.size
    deg = np.diff(ip)
    e_f = int(deg[f_h].sum())
    nbrs = np.unique(np.concatenate([ix[ip[r]:ip[r + 1]] for r in f_h[: min(nf, 200_000)]]))
    ops = FrontierOps(W, C)
    rhs = torch.rand((n, C), dtype=torch.float32, device="cuda", generator=gen)
    ms = timed(lambda: ops.jacobi(f, X, rhs, a, 1e-4), reps)
    line(f"frontier_jacobi ({nf} rows)", ms, nf * (4 + 8 + 4 + 3 * C * 4 + 1) + 8 * e_f + 4 * C * nbrs.size)
    D = torch.rand((nf, C), dtype=torch.float32, device="cuda", generator=gen)
    R = torch.zeros((n, C), dtype=torch.float32, device="cuda")
    mark = torch.zeros(n, dtype=torch.uint8, device="cuda")
    ms = timed(lambda: ops.push(f, D, R, mark), reps)
    line(f"push_scatter ({nf} rows, FP32 atomics)", ms, nf * (12 + 4 * C) + 8 * e_f + nbrs.size * (2 * 4 * C + 1))
    ms = timed(lambda: connected_components_csr(W), reps)
    line("connected components (hook + root + scan + label)", ms, 4 * (n + 1) + 4 * nnz + 8 * n + 4 * 4 * n)
    Xg = X if C % 4 == 0 else torch.rand((n, (C + 3) // 4 * 4), dtype=torch.float32, device="cuda", generator=gen)
    G = torch.empty((C, C), dtype=torch.float64, device="cuda")
    ms = timed(lambda: _lib.call("dlp_gram", n, C, B.ptr(Xg), Xg.shape[1], B.ptr(Xg), Xg.shape[1], None, B.ptr(G),
                                 _lib.F32, B.stream()), reps)
    line(f"gram X^T Y (n x {C}, FP64 MFMA)", ms, flops=2.0 * n * C * C)
    # aggregation AMG on this graph's system W_UU: setup, then one V-cycle M(R) on n_u x C FP32
    from dynlp_amd.amg import AMG
    from dynlp_amd.solvers import build_system
    log("kernel table: AMG on the graph's system")
    sysm = build_system(g, "float32")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    amg = AMG(sysm.W, sysm.s)
    torch.cuda.synchronize()
    setup_ms = (time.perf_counter() - t0) * 1e3
    log(f"AMG setup {setup_ms:.0f} ms, levels {amg.sizes}")
    Ru = torch.rand((sysm.n_u, C), dtype=torch.float32, device="cuda", generator=gen)
    ms = timed(lambda: amg(Ru), reps)
    line(f"amg V-cycle M(R) ({len(amg.sizes)} levels {amg.sizes}, n_u x {C} FP32; setup {setup_ms:.0f} ms)", ms,
         amg_cycle_bytes(amg, C))
    return rows


def amg_cycle_bytes(amg, C, s=4):
    """Algorithmic bytes of one V-cycle (nu = 1): its passes depend on each other, so each pass
    counts its distinct elements once.  Per level (n rows, nnz entries, nc aggregates): first
    sweep 3nCs + ns, two residual passes 2(3nCs + 4(n+1) + 8nnz + ns), restriction
    nCs + 2ncCs + 4n, prolongation 2nCs + ncCs + 4n, and on levels whose sweep is not one fused
    pass (Level.pipelined) the separate correction 3nCs + ns; coarsest level: its FP64 inverse
    (8 n^2) applied to an n x C block."""
    import torch
    probe = torch.empty((1, (C + 3) // 4 * 4), dtype=torch.float32)
    total = 0
    for L in amg.levels:
        n, nnz, nc = L.W.shape[0], L.W.nnz, L.nc
        total += 3 * n * C * s + n * s                                   # first sweep
        total += 2 * (3 * n * C * s + 4 * (n + 1) + 8 * nnz + n * s)    # residual passes
        total += n * C * s + 2 * nc * C * s + 4 * n                      # restriction
        total += 2 * n * C * s + nc * C * s + 4 * n                      # prolongation
        if not (s == 4 and L.pipelined(probe)):
            total += 3 * n * C * s + n * s                               # correction
    nc = amg.sizes[-1]
    total += 8 * nc * nc + 2 * nc * C * s
    return total


def request_row(lp, st, b, batch, g, a):
    """The request's row in the reference's CSV schema (dynlp_amd.report.FIELDS): solver
    "dynlp+inc" (the reference's resident incremental solver, whose role this path takes),
    n_u solved rows, nnz stored entries, iters = push rounds + PCG iterations, the request's own
    certificate (cert_self, hmax) and the accuracy of its predictions on the solved rows; tol
    holds the residual-ratio threshold the request is solved to (rho <= eps), delta (the
    Jacobi solvers' change threshold) does not apply; no exact solution is computed, so
    err_max, cert_post, hmax_ref, agree and ref_ms are NaN."""
    from dynlp_amd import report
    C = lp.C
    Z = (lp.Z64 if lp.fp64 else lp.Z32)[:, :C]
    known = lp.solved.bool() & (lp.y >= 0)
    pred = (Z[:, 0] > 0.5).int() if g.K == 2 else Z.argmax(dim=1).int()
    nk = int(known.sum())
    acc = float((pred[known] == lp.y[known]).sum()) / nk if nk else float("nan")
    nnz = lp.stored_entries()
    return report.row(tag=a.tag or g.name, dataset=g.name, K=g.K, batch=batch, solver="dynlp+inc", n_active=g.n,
                      n_u=int(lp.solved.sum()), nnz=nnz, inserted=int(b.ins_u.shape[0]),
                      deleted=int(b.del_u.shape[0]), iters=st.iters + st.extra.get("pcg_iters", 0), edges=st.edges,
                      sweeps=st.edges / max(nnz, 1), ms=st.ms, setup_ms=st.setup_ms, path=st.path,
                      cert_self=st.cert, hmax=st.hmax, acc_true=acc, delta=float("nan"), tol=lp.eps,
                      dtype="float32+float64", group="auto")


# The ladder from the plain FP64 per-batch path to the service: each step changes one thing.
LADDER = [
    ("plain FP64 per-batch path: FP64 state, host-driven push rounds (the reference's IncrementalLP policy), "
     "no graphs", dict(fp64=True)),
    ("FP64 state, global column PCG (Jacobi preconditioner), no graphs", dict(fp64=True, fp64_repair="pcg")),
    ("FP32 state and directions, column PCG (Jacobi), no graphs",
     dict(pcg_directions="fp32", pcg_coarse=0, use_graph=False)),
    ("FP32, column PCG (Jacobi), iterations replayed from hipGraphs", dict(pcg_directions="fp32", pcg_coarse=0)),
    ("FP32 state, FP16 directions, column PCG (Jacobi), hipGraphs", dict(pcg_directions="fp16", pcg_coarse=0)),
    ("FP32 state, FP16 directions, two-level preconditioner, hipGraphs, push for small repairs",
     dict(pcg_directions="fp16")),
]


def run_path(g, batches, Z0, name, kw, steps, batch_edges):
    """Edge-update throughput of one configuration of the resident path on the same requests:
    load the initial solution, one warm-up request, then `steps` timed requests."""
    import torch
    from dynlp_amd.stream import EdgeStreamLP
    lp = EdgeStreamLP(g, eps=1e-6, **kw)
    lp.load_state(Z0)
    lp.apply_and_solve(batches[0])
    lat = []
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for b in batches[1: 1 + steps]:
        ts = time.perf_counter()
        st = lp.apply_and_solve(b)
        lat.append((time.perf_counter() - ts) * 1e3)
    torch.cuda.synchronize()
    T = time.perf_counter() - t1
    v = len(lat) * batch_edges / T
    log(f"{name}: {v:,.0f} edge-updates/s over {len(lat)} requests (p50 {percentile(lat, 50):.1f} ms, "
        f"last path {st.path}, rho {st.rho:.2e})")
    return {"path": name, "value": round(v, 1), "unit": "edge-updates/s", "requests": len(lat),
            "p50_ms": round(percentile(lat, 50), 2), "p99_ms": round(percentile(lat, 99), 2),
            "rho_last": st.rho}


def compare_with_cpu(Z_gpu, Z_cpu, sysm, rho_cpu):
    """The GPU's certified scores after a request against the CPU restatement's FP64 scores
    after the same request (both from the same initial state): max |F_gpu - F_cpu| over the
    solved rows and label columns, with the bound (rho_gpu + rho_cpu) * max h that two
    solutions within rho of F* must satisfy, and the fraction of solved rows whose predicted
    class is the same."""
    import numpy as np
    C, U = sysm.C, sysm.solved
    Fg, Fc = Z_gpu[U, :C], Z_cpu[U, :C]
    diff = float(np.abs(Fg - Fc).max()) if Fg.size else 0.0
    hmax = float(Z_cpu[U, C].max()) if Fg.size else 1.0
    if C == 1:
        agree = float(np.mean((Fg[:, 0] > 0.5) == (Fc[:, 0] > 0.5))) if Fg.size else 1.0
    else:
        agree = float(np.mean(Fg.argmax(axis=1) == Fc.argmax(axis=1))) if Fg.size else 1.0
    return {"max_abs_diff": diff, "bound": (1e-6 + rho_cpu) * hmax, "argmax_agree": agree,
            "rows": int(U.sum())}


def gpu_case_run(g, b):
    """The reference's CPU-runnable case on the GPU: the reference API's Jacobi (ItLP, FP64) from
    the prior to max |dX| <= 1e-6 (so rho <= 1e-6), then the batch b repaired by the resident
    path with the push policy forced."""
    import torch
    from dynlp_amd.solvers import ItLP, build_system
    from dynlp_amd.stream import EdgeStreamLP
    sysd = build_system(g, "float64")
    torch.cuda.synchronize()
    t = time.perf_counter()
    st = ItLP(g.n, g.K, "float64", delta=1e-6).solve(sysd)
    gpu_jacobi_ms = (time.perf_counter() - t) * 1e3
    lp = EdgeStreamLP(g, eps=1e-6, push_frac=1.0, push_max_rows=g.n, push_mass=float("inf"), push_budget=1e9)
    lp.initial_solve()
    torch.cuda.synchronize()
    t = time.perf_counter()
    sb = lp.apply_and_solve(b)
    gpu_batch_ms = (time.perf_counter() - t) * 1e3
    log(f"GPU case: Jacobi {st.iters} sweeps {gpu_jacobi_ms:.1f} ms; batch by {sb.path} {gpu_batch_ms:.1f} ms")
    return {"jacobi_ms": round(gpu_jacobi_ms, 2), "jacobi_sweeps": st.iters, "batch_ms": round(gpu_batch_ms, 2),
            "batch_path": sb.path, "edge_updates_per_s": round(1000 / (gpu_batch_ms / 1e3), 1), "rho": sb.rho}


def full_solves(g):
    """The full solve of this graph from the prior, by the configuration's stated method (the
    reference API's certified DynLP+, solvers.py:244-343, error bound tol = 1e-6, with block PCG
    and the aggregation-AMG preconditioner), its column-PCG and Jacobi forms, and by the
    resident path's initial solve (two-level column PCG; block PCG with the FP64 Gram on the
    matrix cores), each to rho <= 1e-6 at least: wall time, iterations and the certified bound
    on |F - F*|."""
    import torch
    from dynlp_amd.solvers import DynLPPlus, build_system
    from dynlp_amd.stream import EdgeStreamLP
    sysd = build_system(g, "float32")
    rows = []
    for method, block in (("amg", True), ("amg", False), ("pcg", False)):
        solver = DynLPPlus(g.n, g.K, "float32", method=method, tol=1e-6, block=block)
        torch.cuda.synchronize()
        t = time.perf_counter()
        st = solver.solve(sysd)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3
        kind = ("block PCG (FP64 Grams on the matrix cores, FP64 Cholesky)" if block else "column PCG")
        rows.append({"solver": f"DynLP+ ({kind}, {'aggregation-AMG' if method == 'amg' else 'Jacobi'} "
                               "preconditioner), reference API",
                     "ms": round(ms, 1), "setup_ms": round(st.setup_ms, 1), "iters": st.iters,
                     "cert_bound": st.cert, "hmax": st.hmax, "rho": st.cert / st.hmax if st.hmax else None})
        log(f"full solve, {rows[-1]['solver']}: {ms:.0f} ms, {st.iters} iterations, bound {st.cert:.2e}")
    del sysd
    for mode, dirs, name in (("column", "fp16", "resident state: column PCG, two-level preconditioner, FP16 directions"),
                             ("column", "fp8", "resident state: column PCG, two-level preconditioner, FP8 directions"),
                             ("block", "fp16", "resident state: block PCG, FP64 Gram on the matrix cores, Jacobi")):
        lp = EdgeStreamLP(g, eps=1e-6, pcg_mode=mode, pcg_directions=dirs)
        torch.cuda.synchronize()
        t = time.perf_counter()
        st = lp.initial_solve()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3
        rows.append({"solver": name, "ms": round(ms, 1), "iters": lp.pcg_stats()["pcg_iters"],
                     "cert_bound": st.rho * st.hmax, "hmax": st.hmax, "rho": st.rho})
        log(f"full solve, {name}: {ms:.0f} ms, rho {st.rho:.2e}")
        del lp
        torch.cuda.empty_cache()
    return rows


def pmc_probe(g, directions="auto"):
    """Launch sequence for hardware-counter runs (rocprofv3 --pmc ...): the counters of each
    kernel's launches are read against those of one plain copy of known size.  After the initial
    solve: 5 A*P launches (k_pcg_operator), the copy (512 MiB), 3 FP64 verification passes
    (k_verify, refresh mode 3), 3 connected-components passes of the store, 3 reference-API
    SpMMs (n x C1 FP32), 3 push scatters and 3 frontier Jacobi sweeps from every 10th row and 3
    FP64 Gram products on the matrix cores; each kernel's algorithmic bytes (or flops) per launch
    are printed for the summary."""
    import numpy as np
    import torch
    from dynlp_amd import _lib
    from dynlp_amd import backend as B
    from dynlp_amd.kernels import FrontierOps, spmm_axpy
    from dynlp_amd.stream import EdgeStreamLP
    lp = EdgeStreamLP(g, eps=1e-6, pcg_directions=directions)
    lp.initial_solve()
    torch.cuda.synchronize()
    op_ms = lp.time_operator(reps=4)           # 1 + 4 launches
    nbytes = operator_bytes(g.n, lp.C1, lp.stored_entries(), DIRECTION_BYTES[lp.pcg_directions])
    src = torch.ones(128 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"pmc probe: 5 A*P launches ({op_ms:.3f} ms each, {nbytes} algorithmic bytes each); "
        f"copy of {src.numel() * 4} bytes")
    del src, dst
    solved = lp.solved_mask()
    ip_s, rl_s = lp.indptr.cpu().numpy(), lp.rowlen.cpu().numpy()
    vb = verify_bytes(g.n, int(solved.sum()), lp.C1, int(rl_s[solved].sum()))
    for _ in range(3):
        _lib.call("dlp_lp_refresh", lp.h, 3, 0, B.stream())
    torch.cuda.synchronize()
    log(f"pmc probe: 3 verification passes (k_verify), {vb} algorithmic bytes each")
    n, nnz = g.n, lp.stored_entries()
    for _ in range(3):
        lp.components()
    torch.cuda.synchronize()
    log(f"pmc probe: 3 connected-components passes, {4 * (n + 1) + 4 * nnz + 8 * n + 4 * 4 * n} algorithmic bytes each")
    ip, ix, d = g.csr()
    W = B.csr_from_host(ip, ix, d, n, torch.float32)
    C = lp.C1
    gen = torch.Generator(device="cuda").manual_seed(0)
    X = torch.rand((n, C), dtype=torch.float32, device="cuda", generator=gen)
    a = torch.rand(n, dtype=torch.float32, device="cuda", generator=gen) + 20
    for _ in range(3):
        spmm_axpy(W, X, a=a, b=-1.0)
    torch.cuda.synchronize()
    e = int(ix.shape[0])
    log(f"pmc probe: 3 SpMMs (n x {C}), {4 * (n + 1) + 8 * e + 2 * 4 * n * C + 4 * n} algorithmic bytes each")
    f_h = np.arange(0, n, 10, dtype=np.int32)
    f = torch.as_tensor(f_h, device="cuda")
    nf = f_h.size
    e_f = int(np.diff(ip)[f_h].sum())
    nbrs = np.unique(ix[np.concatenate([np.arange(ip[r], ip[r + 1]) for r in f_h])])
    ops = FrontierOps(W, C)
    D = torch.rand((nf, C), dtype=torch.float32, device="cuda", generator=gen)
    R = torch.zeros((n, C), dtype=torch.float32, device="cuda")
    mark = torch.zeros(n, dtype=torch.uint8, device="cuda")
    for _ in range(3):
        ops.push(f, D, R, mark)
    torch.cuda.synchronize()
    log(f"pmc probe: 3 push scatters ({nf} rows), {nf * (12 + 4 * C) + 8 * e_f + nbrs.size * (2 * 4 * C + 1)} "
        f"algorithmic bytes each")
    rhs = torch.rand((n, C), dtype=torch.float32, device="cuda", generator=gen)
    for _ in range(3):
        ops.jacobi(f, X, rhs, a, 1e-4)
    torch.cuda.synchronize()
    log(f"pmc probe: 3 frontier Jacobi sweeps ({nf} rows), "
        f"{nf * (4 + 8 + 4 + 3 * C * 4 + 1) + 8 * e_f + 4 * C * nbrs.size} algorithmic bytes each")
    Xg = torch.rand((n, (C + 3) // 4 * 4), dtype=torch.float32, device="cuda", generator=gen)
    G = torch.empty((C, C), dtype=torch.float64, device="cuda")
    for _ in range(3):
        _lib.call("dlp_gram", n, C, B.ptr(Xg), Xg.shape[1], B.ptr(Xg), Xg.shape[1], None, B.ptr(G), _lib.F32,
                  B.stream())
    torch.cuda.synchronize()
    log(f"pmc probe: 3 Gram products X^T X (n x {C}, FP64 MFMA), {2 * n * C * C} flops each")
    return 0


def trace_summary(path, start=None):
    """Markdown summary of a rocprofv3 kernel trace (`--kernel-trace --output-format csv`):
    from the first launch of a kernel whose name contains `start` (whole trace when None), the
    span, the summed kernel time, the time with 0, 1, 2, 3 and 4 or more kernels running at once
    (launches on different streams overlapping on the device), streams and hardware queues
    seen, and the kernels by total time."""
    import collections
    import csv
    ev = []
    with open(path) as f:
        for r in csv.DictReader(f):
            ev.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"], r["Stream_Id"],
                       r["Queue_Id"]))
    ev.sort()
    if start:
        first = min(e[0] for e in ev if start in e[2])
        ev = [e for e in ev if e[0] >= first]
    t0, t1 = ev[0][0], max(e[1] for e in ev)
    pts = sorted([(e[0], 1) for e in ev] + [(e[1], -1) for e in ev])
    act, last, by = 0, pts[0][0], collections.Counter()
    for t, d in pts:
        by[min(act, 4)] += t - last
        act += d
        last = t
    tot = sum(e[1] - e[0] for e in ev)
    span = t1 - t0
    out = [f"Launches: {len(ev)}; span {span / 1e9:.3f} s; summed kernel time {tot / 1e9:.3f} s "
           f"({tot / span:.2f} x the span); streams {len({e[3] for e in ev})}, "
           f"hardware queues {len({e[4] for e in ev})}.", "",
           "| Kernels running at once | Time (s) | Share of the span |", "|---|---|---|"]
    for k in range(5):
        out.append(f"| {k if k < 4 else '4 or more'} | {by[k] / 1e9:.3f} | {100 * by[k] / span:.1f} % |")
    agg = collections.defaultdict(lambda: [0, 0])
    for s_, e_, name, *_ in ev:
        short = name.split("(")[0].replace("void ", "")
        agg[short][0] += 1
        agg[short][1] += e_ - s_
    out += ["", "| Kernel | Calls | Total ms | Avg us | % of kernel time |", "|---|---|---|---|---|"]
    for name, (c, t) in sorted(agg.items(), key=lambda x: -x[1][1])[:20]:
        out.append(f"| `{name[:90]}` | {c} | {t / 1e6:.1f} | {t / c / 1e3:.1f} | {100 * t / tot:.1f} |")
    return "\n".join(out)


def open_loop(g, batches, rates, seconds, seed, directions="auto"):
    """The same request stream under open-loop Poisson arrivals (seed S+3) at each rate (requests
    per second), served by the continuous-batching service on the one resident graph: requests
    that arrive while a repair runs are fused into the next one.  Each rate starts from a fresh
    state of the graph and serves the stream's first requests; latency is arrival -> certified,
    throughput = edge updates served / wall time from the first arrival to the last result."""
    import numpy as np
    import torch

    from dynlp_amd import serve

    pts = []
    for rate in rates:
        rng = np.random.default_rng(seed + 3)
        arr = np.cumsum(rng.exponential(1.0 / rate, size=len(batches)))
        k = int(np.searchsorted(arr, seconds, side="right"))
        if k < 2:
            continue
        reqs = [serve.Request(0, b, float(t)) for b, t in zip(batches[:k], arr[:k])]
        svc = serve.Service([g], 1 << 62, slots=1, eps=1e-6, lp_kwargs=dict(pcg_directions=directions))
        svc.run(reqs)
        torch.cuda.synchronize()
        rep = svc.report()
        p = {"rate_req_s": rate, "offered_edge_updates_per_s": round(rate * batches[0].size, 1),
             "requests": rep["requests"], "errors": rep["errors"],
             "edge_updates_per_s": round(rep["edge_updates_per_s"], 1),
             "p50_ms": round(rep["p50_ms"], 2), "p99_ms": round(rep["p99_ms"], 2),
             "mean_fused": round(rep["mean_fused"], 2), "repairs": rep["repairs"], "max_rho": rep["max_rho"]}
        log(f"open loop {rate:g} req/s ({p['offered_edge_updates_per_s']:,.0f} offered): "
            f"{p['edge_updates_per_s']:,.0f} edge-updates/s, p50 {p['p50_ms']:.1f} ms p99 {p['p99_ms']:.1f} ms, "
            f"{p['requests']} requests in {p['repairs']} repairs, max rho {p['max_rho']:.2e}")
        pts.append(p)
        del svc
        torch.cuda.empty_cache()
    return pts


def serve_run(a):
    """The serving workload: open-loop Poisson arrivals over resident graphs, each request
    certified at rho <= 1e-6 in FP64 (Service: per-graph fusion, one push across graphs per
    cycle replayed from a hipGraph, global solves on worker streams).  Throughput counts the
    edge updates of every served request over the wall time from the first arrival to the last
    certified result; latency is arrival -> certified."""
    import numpy as np
    import torch

    from dynlp_amd import graphs, serve

    t = time.perf_counter()
    gs = graphs.serving_graphs(a.serve_graphs, seed=a.seed, scale=a.scale)
    log(f"{len(gs)} graphs, n {sum(g.n for g in gs):,} total, m {sum(g.m for g in gs):,} "
        f"({time.perf_counter() - t:.1f}s)")
    t = time.perf_counter()
    reqs = serve.stream_requests(gs, a.serve_rate, a.serve_seconds, seed=a.seed)
    offered = sum(r.batch.size for r in reqs)
    log(f"{len(reqs)} requests, {offered:,} edge updates ({time.perf_counter() - t:.1f}s)")
    budget = int(a.serve_budget_gb * 1e9)
    need = sum(serve.footprint_bytes(g, pcg_directions=a.directions) for g in gs)
    svc = serve.Service(gs, budget, slots=a.serve_slots, eps=1e-6, lp_kwargs=dict(pcg_directions=a.directions),
                        progress=lambda k, d: log(f"{k}: {d}"))
    t = time.perf_counter()
    # preload: every graph that fits is admitted and solved before the clock starts
    done = svc.run(reqs)
    rep = svc.report()
    log(f"served in {svc.wall:.2f}s (with preload {time.perf_counter() - t:.1f}s): {json.dumps(rep)}")
    torch.cuda.synchronize()
    out = {"metric": "serving edge-updates/s + p50/p99 latency at rel. residual 1e-6, Poisson arrivals",
           "value": round(rep["edge_updates_per_s"], 1), "unit": "edge-updates/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "fp32+fp64", "data": "synthetic",
           "p50_ms": round(rep["p50_ms"], 2), "p99_ms": round(rep["p99_ms"], 2),
           "requests_per_s": round(rep["requests_per_s"], 1), "wall_s": round(svc.wall, 3),
           "config": {"workload": f"{len(gs)} resident graphs (ER/SBM, n log-uniform [50k, 2M] x {a.scale:g}, "
                                  f"degree 8-24, k in {{8,16,64}}), Poisson {a.serve_rate:g} req/s over "
                                  f"{a.serve_seconds:g} s, batches log-uniform [100, 100k] (50/50)",
                      "graphs": len(gs), "n_total": int(sum(g.n for g in gs)),
                      "hbm_budget_gb": a.serve_budget_gb, "footprint_gb": round(need / 1e9, 2),
                      "slots": a.serve_slots, "pcg_directions": a.directions},
           "offered_edge_updates": int(offered),
           "service": {k: (float(f"{v:.4g}") if isinstance(v, float) else v) for k, v in rep.items()}}
    print(json.dumps(out), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed requests")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch-edges", type=int, default=10_000)
    ap.add_argument("--scale", type=float, default=1.0, help="multiplies n (1.0 = SBM-1M)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--directions", default="auto", choices=["auto", "fp32", "bf16", "fp16", "fp8"],
                    help="storage of the global solve's search directions on the service path (auto, "
                         "the library's default: FP8 for this workload's 65 columns, 6 %% more edge "
                         "updates per second than FP16)")
    ap.add_argument("--fp64-steps", type=int, default=4,
                    help="requests timed on each path of the ladder from the plain FP64 path to the service")
    ap.add_argument("--cpu-batches", type=int, default=1, help="requests timed on the CPU port")
    ap.add_argument("--no-fp64", action="store_true", help="skip the path ladder (plain FP64 path included)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-case", action="store_true",
                    help="skip the ER-100k (k = 10) CPU case: Jacobi to rho <= 1e-6 and one push batch, CPU and GPU")
    ap.add_argument("--no-kernels", action="store_true", help="skip the per-kernel roofline table")
    ap.add_argument("--open-loop-rates", default="25,50,100,200",
                    help="Poisson arrival rates (requests/s) of the open-loop leg; empty: skip it")
    ap.add_argument("--open-loop-seconds", type=float, default=1.5, help="arrival window of each open-loop rate")
    ap.add_argument("--no-full-solve", action="store_true",
                    help="skip the full solves from the prior (DynLP+ with AMG and Jacobi PCG, resident column and "
                         "block PCG)")
    ap.add_argument("--out", default=None,
                    help="also append one row per timed request to this CSV in the reference's schema "
                         "(bench.py FIELDS), with a .json record of the run next to it")
    ap.add_argument("--tag", default="", help="tag column of the CSV rows")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="counter probe: initial solve, 5 A*P launches, then a device copy of 512 MiB "
                         "(reads 512 MiB, writes 512 MiB) that calibrates FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--serve", action="store_true",
                    help="the serving workload instead: resident graphs (ER or SBM, n log-uniform in "
                         "[50k, 2M] times --scale, degree 8-24, k in {8, 16, 64}) under an HBM budget, Poisson "
                         "arrivals of batches of 100-100k edge updates; prints its own JSON line")
    ap.add_argument("--serve-graphs", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="requests per second")
    ap.add_argument("--serve-seconds", type=float, default=5.0, help="arrival window")
    ap.add_argument("--serve-budget-gb", type=float, default=32.0, help="HBM budget of the resident graphs")
    ap.add_argument("--serve-slots", type=int, default=8, help="worker streams for global solves")
    ap.add_argument("--trace-summary", metavar="CSV", default=None,
                    help="print a markdown summary of a rocprofv3 kernel-trace CSV and exit (no GPU)")
    ap.add_argument("--trace-from", default=None, help="with --trace-summary: start at this kernel's first launch")
    a = ap.parse_args(argv)
    if a.trace_summary:
        print(trace_summary(a.trace_summary, a.trace_from))
        return 0
    if a.gpus != 1:
        print(f"bench.py runs on one GPU: the package has no multi-GPU path (asked for {a.gpus}); "
              "the per-GPU workload is the whole graph", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from dynlp_amd import graphs
    from dynlp_amd.stream import EdgeStreamLP

    t_all = time.perf_counter()
    torch.cuda.init()
    log(f"device {torch.cuda.get_device_name(0)}")
    if a.serve:
        return serve_run(a)
    t = time.perf_counter()
    g = graphs.config_graph(3, scale=a.scale, seed=a.seed)
    log(f"graph n={g.n:,} m={g.m:,} labelled={int(g.labeled.sum()):,} ({time.perf_counter() - t:.1f}s)")
    if a.pmc_probe:
        return pmc_probe(g, a.directions)
    t = time.perf_counter()
    nreq = a.warmup + a.steps
    rates = [float(x) for x in a.open_loop_rates.split(",") if x.strip()]
    # the open-loop leg serves the stream's first requests at each rate: enough for the largest
    n_open = int(max(rates, default=0) * a.open_loop_seconds * 1.3) + 8 if rates else 0
    batches, _, es_final = graphs.update_stream(g, max(nreq, n_open), a.batch_edges, seed=a.seed,
                                                snapshot_after=nreq)
    log(f"update stream: {len(batches)} requests x {a.batch_edges} edges ({time.perf_counter() - t:.1f}s)")

    # ---- mixed-precision service path -----------------------------------------------------
    t = time.perf_counter()
    lp = EdgeStreamLP(g, eps=1e-6, pcg_directions=a.directions)
    st0 = lp.initial_solve()
    Z0 = lp.scores_with_h().astype(np.float32)
    log(f"initial solve: {st0.ms:.0f} ms rho={st0.rho:.2e} (setup {time.perf_counter() - t:.1f}s)")
    Z1 = None
    for i, b in enumerate(batches[: a.warmup]):
        lp.apply_and_solve(b)
        if i == 0 and not a.no_cpu:
            Z1 = lp.scores_with_h()       # the state the CPU leg's first request is compared with
    torch.cuda.synchronize()
    lat, rhos, paths, csv_rows = [], [], {}, []
    pcg0 = lp.pcg_stats()
    paused = 0.0
    t0 = time.perf_counter()
    for t, b in enumerate(batches[a.warmup:nreq]):
        ts = time.perf_counter()
        st = lp.apply_and_solve(b)
        lat.append((time.perf_counter() - ts) * 1e3)
        rhos.append(st.rho)
        paths[st.path] = paths.get(st.path, 0) + 1
        if a.out:
            tp = time.perf_counter()
            csv_rows.append(request_row(lp, st, b, a.warmup + t + 1, g, a))
            paused += time.perf_counter() - tp   # the row's extra device reads are not timed
    torch.cuda.synchronize()
    T = time.perf_counter() - t0 - paused
    pcg1 = lp.pcg_stats()
    edges = a.steps * a.batch_edges
    value = edges / T
    log(f"service: {a.steps} requests in {T:.3f}s -> {value:,.0f} edge-updates/s; p50 {percentile(lat, 50):.1f} ms "
        f"p99 {percentile(lat, 99):.1f} ms; max rho {max(rhos):.2e}; paths {paths}")

    # ---- roofline of the dominant kernel (the global solve's A*P launch) --------------------
    op_ms = lp.time_operator(reps=20)
    stored = lp.stored_entries()
    dirs = lp.pcg_directions
    nbytes = operator_bytes(g.n, lp.C1, stored, DIRECTION_BYTES[dirs])
    achieved = nbytes / (op_ms * 1e-3) / 1e9
    word, wid = {"fp32": ("F32x4", 0), "bf16": ("BF16x4", 2), "fp16": ("F16x4", 3), "fp8": ("F8x4", 5)}[dirs]
    # template arguments: gathers in flight per lane, lane word, waves per SIMD it is built for
    # (the library's defaults: FP8 on wide rows, at most 4 edge lanes, 8 and 6; else 4 and 7)
    wide = 64 // ((lp.C1 + 3) // 4) <= 4
    depth, waves = (8, 6) if dirs == "fp8" and wide else (4, 7)
    kname = (f"k_pcg_operator<{depth}, {word}, {waves}> (A*P on {dirs.upper()} directions with fused p^T A p, "
             "the global solve)")
    traffic = pmc_traffic(f"dlp::k_pcg_operator<{depth}, (dlp::Word){wid}, {waves}>")
    roofline = {"kernel": kname,
                "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": "profiles/pcg_operator_sbm1m_k64_pmc.md (rocprofv3: the L2's memory read and write requests by size)"
                if traffic else None,
                "avg_launch_ms": round(op_ms, 4), "algorithmic_bytes": int(nbytes)}
    log(f"A*P launch: {op_ms:.3f} ms, {nbytes / 1e6:.0f} MB -> {achieved:.0f} GB/s")
    # the FP64 verification pass every global repair ends with (refresh mode 3 over every
    # solved row, on the certified state it leaves unchanged)
    from dynlp_amd import _lib
    from dynlp_amd.backend import stream as cur_stream
    ver_ms = timed(lambda: _lib.call("dlp_lp_refresh", lp.h, 3, 0, cur_stream()), 10)
    ver_kernel = lp.status()["refresh_kernel"]
    solved = lp.solved.bool()
    n_solved = int(solved.sum().item())
    stored_solved = int(lp.rowlen[solved].sum().item())
    vbytes = verify_bytes(g.n, n_solved, lp.C1, stored_solved)
    vname = {3: "k_refresh_pipe<4> (listed rows: the pipelined verification did not apply)",
             2: "k_verify<4> (FP64 residual of every solved row, column maxima, h maximum)",
             1: "k_refresh_v4 (FP64 refresh of every row: the pipelined verification did not apply)",
             0: "k_refresh (FP64 refresh of every row: the pipelined verification did not apply)"}[ver_kernel]
    verify = {"kernel": vname, "bound": "hbm", "ms": round(ver_ms, 4), "algorithmic_bytes": int(vbytes),
              "solved_rows": n_solved,
              "achieved": round(vbytes / (ver_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
              "frac": round(vbytes / (ver_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"verification pass: {ver_ms:.3f} ms, {vbytes / 1e6:.0f} MB -> {verify['achieved']:.0f} GB/s")
    # CC labels of the final graph, bit for bit against SciPy's components in the reference's numbering
    t = time.perf_counter()
    u_f, v_f, _ = es_final.pairs()
    cc_exact = bool(np.array_equal(lp.components(), canonical_components(g.n, np.r_[u_f, v_f], np.r_[v_f, u_f])))
    log(f"CC labels of the final graph bit-exact against SciPy: {cc_exact} ({time.perf_counter() - t:.1f}s)")
    pcg_dirs, pcg_coarse = lp.pcg_directions, lp.pcg_coarse
    pcg_iters = pcg1["pcg_iters"] - pcg0["pcg_iters"]
    err = lp.error_report()
    log(f"final state: rel l2 residual {err['rel_l2_residual']:.2e}, certified |F - F*| <= {err['cert_bound']:.2e}, "
        f"argmax certified on {100 * err['argmax_certified_frac']:.3f}% of solved rows")
    pcg_ms = (pcg1["pcg_us"] - pcg0["pcg_us"]) / 1e3
    del lp
    torch.cuda.empty_cache()

    # ---- the same requests under open-loop Poisson arrivals (continuous batching) --------------
    openl = None
    if rates:
        openl = {"workload": f"the same {a.batch_edges}-edge requests, Poisson arrivals (seed S+3) over "
                             f"{a.open_loop_seconds:g} s per rate, served by the continuous-batching service on "
                             "the one resident graph (requests arriving during a repair are fused into the "
                             "next); latency = arrival -> certified",
                 "points": open_loop(g, batches, rates, a.open_loop_seconds, a.seed, a.directions)}

    # ---- the same requests on simpler paths: the plain FP64 per-batch path (the denominator of
    # the 5x target) and the ladder from it to the service, one change at a time ------------------
    ladder = []
    if not a.no_fp64 and a.fp64_steps > 0:
        for name, kw in LADDER:
            ladder.append(run_path(g, batches, Z0, name, kw, a.fp64_steps, a.batch_edges))
            torch.cuda.empty_cache()
        ladder.append({"path": f"the service (above): FP32 state, {pcg_dirs.upper()} directions, two-level "
                               "preconditioner, hipGraph replay, push for small repairs", "value": round(value, 1),
                       "unit": "edge-updates/s", "requests": a.steps, "p50_ms": round(percentile(lat, 50), 2),
                       "p99_ms": round(percentile(lat, 99), 2)})
    fp64 = ladder[0] if ladder else None

    # ---- the configuration's full solve by its stated method ------------------------------------
    full = None
    if not a.no_full_solve:
        full = full_solves(g)
        torch.cuda.empty_cache()

    # ---- the other hot kernels at this graph's shape ------------------------------------------
    ktab = None
    if not a.no_kernels:
        ktab = kernel_table(g, 65)
        for r in ktab:
            log(f"{r['kernel']}: {r['ms']:.3f} ms -> {r['achieved']} {r['unit']} ({100 * r['frac']:.1f}% of peak)")
        torch.cuda.empty_cache()

    # ---- CPU baseline: the FP64 NumPy/SciPy restatement on the host --------------------------
    cpu, agreement = None, None
    if not a.no_cpu and a.cpu_batches > 0:
        from oracle.incremental import CPUIncremental
        t = time.perf_counter()
        port = CPUIncremental(g.n, g.u, g.v, g.w, g.labeled, g.y, g.K, eps=1e-6, Z=Z0)
        log(f"CPU port setup {time.perf_counter() - t:.1f}s")
        cms = []
        for i, b in enumerate(batches[: a.cpu_batches]):
            out = port.solve_batch(b.ins_u, b.ins_v, b.ins_w, b.del_u, b.del_v)
            cms.append(out["ms"])
            log(f"CPU request: {out['ms']:.0f} ms rho={out['rho']:.2e} path={out['path']}")
            if i == 0 and Z1 is not None:
                agreement = compare_with_cpu(Z1, port.Z, port.sys, out["rho"])
                log(f"GPU vs CPU after request 1: max |F_gpu - F_cpu| = {agreement['max_abs_diff']:.2e} "
                    f"(bound {agreement['bound']:.2e}), argmax agreement {agreement['argmax_agree']:.6f}")
        cpu = {"value": round(len(cms) * a.batch_edges / (sum(cms) / 1e3), 1), "unit": "edge-updates/s",
               "cores": 1, "kind": "port",
               "sample": f"first {len(cms)} request(s) of {a.batch_edges} edge updates on the same graph from the same "
                         "initial state: FP64 NumPy/SciPy restatement that updates the edge set, rebuilds the "
                         "grounding and W_UU from it, refreshes the changed rows' residual and repairs by host push "
                         "(few violators) or column PCG (single-threaded sparse kernels)",
               "host_cores_available": len(os.sched_getaffinity(0))}
    case = None
    if not a.no_cpu_case:
        # the reference's CPU-runnable case, ER n = 100,000, degree 10, k = 10, 1 % labelled:
        # Jacobi to rho <= 1e-6, then one batch of 1,000 edge updates by residual push
        from oracle.incremental import cpu_case
        gc = graphs.config_graph(1, seed=a.seed)
        bc = graphs.update_stream(gc, 1, 1000, seed=a.seed)[0][0]
        hc = cpu_case(gc, bc)
        log(f"CPU case (ER-100k, k=10): Jacobi {hc['jacobi_sweeps']} sweeps {hc['jacobi_ms']:.0f} ms; batch of "
            f"1,000 by {hc['batch_path']} {hc['batch_ms']:.0f} ms (rho {hc['rho']:.2e})")
        case = {"workload": "ER n=100,000, degree 10, k=10, 1% labelled: Jacobi to rho <= 1e-6, then one batch "
                            "of 1,000 edge updates (50/50) by residual push",
                "cpu": dict(hc, cores=1, kind="port (FP64 NumPy/SciPy, single-threaded sparse kernels)"),
                "gpu": gpu_case_run(gc, bc)}

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "edge-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(T / a.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "sbm-1m-k64 edge-update stream" if a.scale == 1.0 else f"sbm-{g.n}-k64 edge-update stream",
                   "n": g.n, "blocks": 64, "deg": 20, "p_in": 0.8, "k": 64, "label_frac": 0.005,
                   "batch_edges": a.batch_edges, "ins_frac": 0.5, "eps": 1e-6, "seed": a.seed,
                   "pcg_directions": pcg_dirs, "pcg_coarse_aggregates": pcg_coarse},
        "p50_ms": round(percentile(lat, 50), 3), "p99_ms": round(percentile(lat, 99), 3),
        "rho_max": max(rhos), "paths": paths,
        "pcg_iters_per_request": round(pcg_iters / a.steps, 2), "pcg_ms_per_request": round(pcg_ms / a.steps, 3),
        "error": err, "cc_bit_exact": cc_exact, "open_loop": openl,
        "fp64_path": fp64,
        "speedup_vs_fp64_path": round(value / fp64["value"], 2) if fp64 else None,
        "path_ladder": ladder or None,
        "roofline": roofline, "verify_pass": verify, "cpu_baseline": cpu, "cpu_agreement": agreement,
        "cpu_case": case, "full_solve": full, "kernels": ktab,
        "wall_s": round(time.perf_counter() - t_all, 1),
    }
    if a.out:
        from dynlp_amd import report
        for r in csv_rows:
            r["run_s"] = out["wall_s"]
        report.write(csv_rows, a.out, args=vars(a), env={"gpu": torch.cuda.get_device_name(0),
                                                          "torch": torch.__version__}, run_s=out["wall_s"])
        log(f"{len(csv_rows)} rows -> {a.out}\n" + report.summarize(csv_rows))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

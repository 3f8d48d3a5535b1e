"""Writes the golden fixtures of this directory from the FP64 oracle (test infrastructure).

Each fixture holds inputs (graph pairs and weights, labels, update batches) and the expected
outputs the oracle computes for them.  Tests read the committed .npz files: the CPU suite checks
that the oracle still reproduces them from the stored inputs, the GPU suite runs the HIP path on
the same stored inputs.  Inputs are stored as arrays, so the fixtures do not depend on the graph
generators staying unchanged.  Run from the repository root:

    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from dynlp_amd import graphs  # noqa: E402
from oracle import lp as O  # noqa: E402


def graph_arrays(g):
    return dict(n=np.int64(g.n), u=g.u.astype(np.int32), v=g.v.astype(np.int32), w=g.w.astype(np.float32),
                labeled=g.labeled.astype(np.uint8), y=g.y.astype(np.int32), K=np.int64(g.K))


def system_fixture(g):
    """[F*, h] of the harmonic system with its solved set, diag and column residual ratios."""
    S = O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, g.K)
    Z = O.solve_system(S)
    rho = O.rho_cols(S.diag, O.residual(S, Z))
    return dict(graph_arrays(g), solved=S.solved.astype(np.uint8), diag=S.diag, Z=Z, rho=rho)


def stream_fixture(g, n_batches, size, seed, ins_frac=0.5, inter_first=False, with_scores=True):
    """Edge-update batches and, after each batch, CC labels of the whole graph, the solved set
    and (with_scores) [F*, h] of the rebuilt system."""
    batches, _ = graphs.update_stream(g, n_batches, size, ins_frac=ins_frac, seed=seed, inter_first=inter_first)
    es = {(int(a), int(b)): float(x) for a, b, x in zip(g.u, g.v, g.w)}
    out = graph_arrays(g)
    out["n_batches"] = np.int64(n_batches)
    for t, b in enumerate(batches):
        out[f"b{t}_ins_u"], out[f"b{t}_ins_v"] = b.ins_u.astype(np.int32), b.ins_v.astype(np.int32)
        out[f"b{t}_ins_w"] = b.ins_w.astype(np.float32)
        out[f"b{t}_del_u"], out[f"b{t}_del_v"] = b.del_u.astype(np.int32), b.del_v.astype(np.int32)
        O.apply_edge_batch(es, zip(b.ins_u.tolist(), b.ins_v.tolist(), b.ins_w.tolist()),
                           zip(b.del_u.tolist(), b.del_v.tolist()))
        u = np.array([k[0] for k in es], dtype=np.int64)
        v = np.array([k[1] for k in es], dtype=np.int64)
        w = np.array(list(es.values()))
        out[f"b{t}_cc"], _ = O.connected_components(g.n, np.r_[u, v], np.r_[v, u])
        S = O.build_system(g.n, u, v, w, g.labeled, g.y, g.K)
        out[f"b{t}_solved"] = S.solved.astype(np.uint8)
        if with_scores:
            out[f"b{t}_Z"] = O.solve_system(S)
    return out


def spmm_fixture(seed=11):
    g = graphs.sbm(500, 5, 8, 0.8, 2, 0.02, seed=seed)
    ip, ix, d = g.csr()
    W = O.csr(ip, ix, d, g.n)
    rs = np.random.default_rng(seed)
    out = dict(indptr=ip.astype(np.int32), indices=ix.astype(np.int32), data=d.astype(np.float32))
    a = rs.random(g.n).astype(np.float32)
    out["a"] = a
    for C in (1, 17, 65):
        X = rs.standard_normal((g.n, C)).astype(np.float32)
        out[f"X{C}"] = X
        out[f"Y{C}"] = O.spmm_axpy(W, X.astype(np.float64), a.astype(np.float64), -1.0)
    return out


def amg_fixture(seed=7):
    """The V-cycle operator of a two-level-or-deeper hierarchy applied to a fixed block."""
    g = graphs.sbm(300, 6, 6, 0.85, 2, 0.03, seed=seed)
    S = O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, g.K)
    Wl, dl, rl = S.local()
    amg = O.AMG(Wl, S.s[S.U], max_coarse=20)
    X = np.random.default_rng(seed).standard_normal((Wl.shape[0], 8))
    return dict(graph_arrays(g), X=X, MX=amg(X), sizes=np.array(amg.sizes, dtype=np.int64),
                agg0=amg.levels[0]["agg"].astype(np.int64))


def main():
    fixtures = {
        "sys_er_k2": lambda: system_fixture(graphs.erdos_renyi(1000, 8, 2, label_frac=0.02, seed=21)),
        "sys_er_k16": lambda: system_fixture(graphs.erdos_renyi(1000, 8, 16, label_frac=0.03, seed=22)),
        "sys_sbm_k64": lambda: system_fixture(graphs.sbm(640, 64, 10, 0.8, 64, 0.1, seed=23)),
        "stream_sbm_k8": lambda: stream_fixture(graphs.sbm(1500, 8, 10, 0.8, 8, 0.02, seed=24), 5, 150, seed=26),
        "cc_fragmenting": lambda: stream_fixture(graphs.sbm(2000, 16, 3.0, 0.9, 16, 0.01, seed=27, unit=True), 5,
                                                 300, seed=29, ins_frac=0.3, inter_first=True, with_scores=False),
        "spmm": spmm_fixture,
        "amg": amg_fixture,
    }
    for name, make in fixtures.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **make())
        print(f"{path}: {os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    main()

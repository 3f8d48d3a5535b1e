"""Connected components on the device (cc.connected_components, reference cc.py:8-23).

Labels are int32 in [0, n_comp): a component's label is the rank of its minimum vertex id
among all component minima, bit-identical to the reference's hook/jump + unique.
"""
from __future__ import annotations

import torch

from . import _lib
from .backend import CSR, device, ptr, stream


def _workspace(n):
    nbytes = int(_lib.load().dlp_cc_workspace_bytes(max(int(n), 1)))
    return torch.empty(nbytes, dtype=torch.uint8, device=device()), nbytes


def connected_components(n, rows, cols):
    """(labels int32[n], n_comp) for an undirected edge list (cc.py:8-23)."""
    n = int(n)
    if n == 0:
        return torch.zeros(0, dtype=torch.int32, device=device()), 0
    rows = torch.as_tensor(rows, device=device()).to(torch.int32).contiguous()
    cols = torch.as_tensor(cols, device=device()).to(torch.int32).contiguous()
    if rows.shape != cols.shape:
        raise ValueError("rows and cols differ in length")
    labels = torch.empty(n, dtype=torch.int32, device=device())
    nc = torch.zeros(1, dtype=torch.int32, device=device())
    ws, nb = _workspace(n)
    _lib.call("dlp_cc_edges", n, int(rows.shape[0]), ptr(rows), ptr(cols), ptr(labels), ptr(nc), ptr(ws), nb,
              stream())
    return labels, int(nc.item())


def connected_components_csr(W: CSR, mask=None, out=None, count=None, ws=None):
    """Components of a resident (slack) CSR graph, optionally restricted to mask != 0.
    With `out`/`count`/`ws` given nothing is allocated and the count stays on the device."""
    n = W.shape[0]
    labels = torch.empty(n, dtype=torch.int32, device=device()) if out is None else out
    nc = torch.zeros(1, dtype=torch.int32, device=device()) if count is None else count
    if ws is None:
        ws, nb = _workspace(n)
    else:
        nb = ws.numel()
    _lib.call("dlp_cc_csr", n, ptr(W.indptr), ptr(W.rowlen), ptr(W.indices), ptr(mask), ptr(labels), ptr(nc),
              ptr(ws), nb, stream())
    if count is None:
        return labels, int(nc.item())
    return labels, nc


def _lightest(nc, cu, cv, de):
    """For every component with an outgoing edge (cc.py:26-35): the component, the one across
    its lightest outgoing edge (ties: the smallest such component) and that edge's weight."""
    from .backend import scatter_min
    dev = de.device
    best = torch.full((nc,), float("inf"), dtype=torch.float32, device=dev)
    scatter_min(best, cu, de)
    across = torch.full((nc,), nc, dtype=torch.int64, device=dev)
    sel = de <= best[cu.long()]
    scatter_min(across, cu[sel], cv[sel].long())
    c = torch.nonzero(across < nc).flatten()
    return c, across[c], best[c]


def fh_components(m, rows, cols, d, k, ground=None):
    """Felzenszwalb-Huttenlocher merging in parallel Boruvka rounds (cc.py:38-74): every
    component takes its lightest outgoing edge e and merges across it when
        d(e) <= min(Int(C1) + k/|C1|, Int(C2) + k/|C2|),
    Int(C) being C's largest internal dissimilarity; the merges of a round are joined by
    connected components.  When no edge qualifies and ground (per vertex) is given, components
    with zero ground then join across their lightest outgoing edge until none can.  FP32
    dissimilarities, as the reference keeps them.  Returns (labels int32[m], count)."""
    from .backend import scatter_add, scatter_max
    dev = device()
    lab = torch.arange(m, dtype=torch.int32, device=dev)
    Int = torch.zeros(m, dtype=torch.float32, device=dev)
    size = torch.ones(m, dtype=torch.float32, device=dev)
    rows = torch.as_tensor(rows, device=dev).long()
    cols = torch.as_tensor(cols, device=dev).long()
    d = torch.as_tensor(d, device=dev).to(torch.float32)
    gvert = None if ground is None else torch.as_tensor(ground, device=dev).to(torch.float32)
    nc, fh = m, True
    while rows.shape[0]:
        cu, cv = lab[rows].long(), lab[cols].long()
        ext = cu != cv
        if not fh:   # grounding pass: only components without ground pick an edge
            g = torch.zeros(nc, dtype=torch.float32, device=dev)
            scatter_add(g, lab, gvert)
            ext &= g[cu] == 0
        if not bool(ext.any()):
            break
        c, t, dc = _lightest(nc, cu[ext], cv[ext], d[ext])
        if fh:
            ok = dc <= torch.minimum(Int[c] + k / size[c], Int[t] + k / size[t])
            if not bool(ok.any()):
                if gvert is None:
                    break
                fh = False
                continue
            c, t, dc = c[ok], t[ok], dc[ok]
        root, nc = connected_components(nc, c, t)
        root = root.long()
        Int2 = torch.zeros(nc, dtype=torch.float32, device=dev)
        size2 = torch.zeros(nc, dtype=torch.float32, device=dev)
        scatter_max(Int2, root, Int)
        scatter_max(Int2, root[c], dc)
        scatter_add(size2, root, size)
        lab, Int, size = root[lab.long()].to(torch.int32), Int2, size2
    return lab, nc


def adaptive_components(m, rows, cols, w, s_r, s_c, wmax, ground=None):
    """Component-adaptive merging (cc.py:77-83): dissimilarity 1 - w/wmax scaled by the two
    endpoints' own mean dissimilarities, k = the mean scaled dissimilarity."""
    if m == 0:
        return torch.zeros(0, dtype=torch.int32, device=device()), 0
    w = torch.as_tensor(w, device=device())
    s_rc = torch.as_tensor(s_r, device=device()) * torch.as_tensor(s_c, device=device())
    d = (1 - w / wmax) / torch.sqrt(s_rc).to(w.dtype)   # in w's precision; FP32 from fh_components on
    return fh_components(m, rows, cols, d, float(d.mean()) if d.shape[0] else 0.0, ground)

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

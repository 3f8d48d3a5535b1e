"""Connected components on the device (cc.connected_components, reference cc.py:8-23).

Labels are int32 in [0, n_comp): a component's label is the rank of its minimum vertex id
among all component minima, bit-identical to the reference's hook/jump + unique.
"""
from __future__ import annotations

import ctypes as ct

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


def fh_components(m, rows, cols, d, k, ground=None):
    """Felzenszwalb-Huttenlocher merging of m vertices over the edge list (rows, cols) with
    dissimilarities d (FP32), in parallel Boruvka rounds on the device (dlp_fh_components,
    cc.py:38-74): a component merges across its lightest outgoing edge e when
        d(e) <= min(Int(C1) + k/|C1|, Int(C2) + k/|C2|),
    Int(C) being C's largest internal dissimilarity; when no edge qualifies and ground (per
    vertex, >= 0) is given, components with no ground then join across their lightest outgoing
    edge until none can.  Returns (labels int32[m], count)."""
    dev = device()
    m = int(m)
    labels = torch.empty(max(m, 1), dtype=torch.int32, device=dev)[:m]
    rows = torch.as_tensor(rows, device=dev).to(torch.int32).contiguous()
    cols = torch.as_tensor(cols, device=dev).to(torch.int32).contiguous()
    d = torch.as_tensor(d, device=dev).to(torch.float32).contiguous()
    if not (rows.shape == cols.shape == d.shape):
        raise ValueError("rows, cols and d differ in length")
    g = None if ground is None else torch.as_tensor(ground, device=dev).to(torch.float32).contiguous()
    nb = int(_lib.load().dlp_fh_workspace_bytes(max(m, 1)))
    ws = torch.empty(nb, dtype=torch.uint8, device=dev)
    nc = ct.c_int32(0)
    _lib.call("dlp_fh_components", m, int(rows.shape[0]), ptr(rows), ptr(cols), ptr(d), float(k), ptr(g),
              ptr(labels), ct.byref(nc), ptr(ws), nb, stream())
    return labels, int(nc.value)


def _dtype_of(w):
    return _lib.F64 if w.dtype == torch.float64 else _lib.F32


def adaptive_components(m, rows, cols, w, s_r, s_c, wmax, ground=None, scale_index=None):
    """Component-adaptive merging (cc.py:77-83): the dissimilarity 1 - w/wmax scaled by the two
    endpoints' own mean dissimilarities (s_r, s_c per edge), k = the mean scaled
    dissimilarity; both on the device (dlp_fh_dissimilarity), then fh_components.  With
    scale_index=(ir, ic), s_r and s_c are per-vertex scales read at ir[e], ic[e] instead."""
    dev = device()
    if int(m) == 0:
        return torch.zeros(0, dtype=torch.int32, device=dev), 0
    w = torch.as_tensor(w, device=dev)
    if w.dtype != torch.float64:
        w = w.to(torch.float32)
    w = w.contiguous()
    rows = torch.as_tensor(rows, device=dev).to(torch.int32).contiguous()
    cols = torch.as_tensor(cols, device=dev).to(torch.int32).contiguous()
    s_r = torch.as_tensor(s_r, device=dev).to(torch.float32).contiguous()
    s_c = torch.as_tensor(s_c, device=dev).to(torch.float32).contiguous()
    ir, ic = (None, None) if scale_index is None else (
        torch.as_tensor(x, device=dev).to(torch.int32).contiguous() for x in scale_index)
    ne = int(w.shape[0])
    d = torch.empty(max(ne, 1), dtype=torch.float32, device=dev)[:ne]
    nb = int(_lib.load().dlp_fh_dissimilarity_workspace_bytes())
    ws = torch.empty(nb, dtype=torch.uint8, device=dev)
    total = ct.c_double(0.0)
    _lib.call("dlp_fh_dissimilarity", ne, ptr(w), _dtype_of(w), ptr(s_r), ptr(s_c), ptr(ir), ptr(ic), float(wmax),
              ptr(d), ct.byref(total), ptr(ws), nb, stream())
    return fh_components(m, rows, cols, d, total.value / ne if ne else 0.0, ground)


def vertex_scales(m, lr, w, wmax):
    """Local scale of each of m vertices over its stored entries (lr[e] = row, w[e] > 0 live):
    the mean of 1 - w/wmax, plus 1e-6 (FP32, stream.py:72-74), on the device (dlp_fh_scale)."""
    dev = device()
    w = torch.as_tensor(w, device=dev)
    if w.dtype != torch.float64:
        w = w.to(torch.float32)
    w = w.contiguous()
    lr = torch.as_tensor(lr, device=dev).to(torch.int32).contiguous()
    s = torch.empty(max(int(m), 1), dtype=torch.float32, device=dev)
    cnt = torch.empty_like(s)
    _lib.call("dlp_fh_scale", int(m), int(w.shape[0]), ptr(lr), ptr(w), _dtype_of(w), float(wmax), ptr(s), ptr(cnt),
              stream())
    return s[: int(m)]

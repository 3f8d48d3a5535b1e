"""Device backend: torch-ROCm tensors as device buffers, HIP kernels through the C ABI.

Mirrors dynlp/backend.py of the reference (backend.py:1-56): the array namespace, CSR
container, unbuffered scatters, device sync, memory use and a synchronising Timer.  The
reference bound `xp` to CuPy unconditionally; here DYNLP_BACKEND selects the backend and the
only product backend is "hip" (one MI355X, gfx950).  Any other value raises: the FP64 CPU
restatement lives outside this package (it is test infrastructure), so there is no CPU
fallback to fall into.
"""
from __future__ import annotations

import os
import time

import torch

from . import _lib

BACKEND = os.environ.get("DYNLP_BACKEND", "hip")
if BACKEND != "hip":
    raise ImportError(f"DYNLP_BACKEND={BACKEND!r}: this package runs on the HIP backend only")

xp = torch
float32, float64, int32, int64, uint8 = torch.float32, torch.float64, torch.int32, torch.int64, torch.uint8


def device():
    return torch.device("cuda", torch.cuda.current_device())


def stream():
    """Current torch stream as a raw hipStream_t (what the C ABI takes)."""
    return torch.cuda.current_stream().cuda_stream


def ptr(t):
    return None if t is None else t.data_ptr()


def dtype_code(dt):
    if dt == torch.float32:
        return _lib.F32
    if dt == torch.float64:
        return _lib.F64
    raise TypeError(f"unsupported dtype {dt}")


def asarray(x, dtype=None):
    """Device tensor from a host array / tensor (copies host data to the device)."""
    if isinstance(x, torch.Tensor):
        t = x.to(device())
    else:
        t = torch.as_tensor(x).to(device())
    if dtype is not None:
        t = t.to(dtype)
    return t.contiguous()


def zeros(*shape, dtype=torch.float32):
    return torch.zeros(*shape, dtype=dtype, device=device())


def empty(*shape, dtype=torch.float32):
    return torch.empty(*shape, dtype=dtype, device=device())


_SCATTER_DTYPES = {torch.float32: 0, torch.float64: 1, torch.int32: 2, torch.int64: 3}


def _scatter(op, a, idx, v):
    """a[idx[i]] op= v[i] on the device (dlp_scatter), unbuffered: repeated indices all apply.
    v: one row per index (shape (m, *a.shape[1:])), or one value for every element; indices
    outside a's first dimension are skipped."""
    if a.dtype not in _SCATTER_DTYPES or not a.is_contiguous():
        raise TypeError(f"scatter: contiguous float32/float64/int32/int64 target expected, got {a.dtype}")
    idx = torch.as_tensor(idx, device=a.device)
    if idx.dtype not in (torch.int32, torch.int64):
        idx = idx.long()
    idx = idx.contiguous()
    m = int(idx.shape[0])
    cols = 1
    for d in a.shape[1:]:
        cols *= int(d)
    v = torch.as_tensor(v, dtype=a.dtype, device=a.device)
    if v.numel() == 1:
        v, stride = v.reshape(1).contiguous(), 0
    else:
        v, stride = v.expand(m, *a.shape[1:]).contiguous(), cols
    _lib.call("dlp_scatter", op, _SCATTER_DTYPES[a.dtype], int(a.shape[0]), cols, a.data_ptr(), m, idx.data_ptr(),
              int(idx.dtype == torch.int64), v.data_ptr(), stride, None, stream())


def scatter_add(a, idx, v):
    """a[idx] += v, unbuffered (repeated indices accumulate) (backend.py:19-20)."""
    _scatter(0, a, idx, v)


def scatter_max(a, idx, v):
    """a[idx] = max(a[idx], v), unbuffered (backend.py:23-24)."""
    _scatter(1, a, idx, v)


def scatter_min(a, idx, v):
    """a[idx] = min(a[idx], v), unbuffered (backend.py:27-28)."""
    _scatter(2, a, idx, v)


def sync():
    torch.cuda.synchronize()


def mem_used_gb() -> float:
    return torch.cuda.memory_allocated() / 1e9


class Timer:
    """Wall-clock timer that synchronizes the device on enter and exit (backend.py:43-56)."""

    def __init__(self):
        self.ms = 0.0

    def __enter__(self):
        sync()
        self._t = time.perf_counter()
        return self

    def __exit__(self, *exc):
        sync()
        self.ms += (time.perf_counter() - self._t) * 1e3


class CSR:
    """Device CSR (int32 indptr/indices, FP32 or FP64 data).  `rowlen` (optional) makes it a
    slack CSR: row r holds rowlen[r] live entries from indptr[r]."""

    def __init__(self, data, indices, indptr, shape, rowlen=None):
        self.data = data
        self.indices = indices
        self.indptr = indptr
        self.shape = tuple(int(s) for s in shape)
        self.rowlen = rowlen
        self._nnz = None

    @property
    def nnz(self):
        if self._nnz is None:
            self._nnz = int(self.rowlen.sum()) if self.rowlen is not None else int(self.indptr[-1])
        return self._nnz

    @property
    def avg_deg(self):
        return self.nnz / max(self.shape[0], 1)

    @property
    def dtype(self):
        return self.data.dtype

    def astype(self, dt):
        if dt == self.data.dtype:
            return self
        return CSR(self.data.to(dt), self.indices, self.indptr, self.shape, self.rowlen)

    def __matmul__(self, X):
        from .kernels import spmm_axpy
        if X.dim() == 1:
            return spmm_axpy(self, X[:, None].contiguous())[:, 0]
        return spmm_axpy(self, X)

    def row_ids(self):
        """Row index of every stored entry (graphs.csr_rows)."""
        counts = (self.indptr[1:] - self.indptr[:-1]).long()
        return torch.repeat_interleave(torch.arange(self.shape[0], device=self.indptr.device, dtype=torch.int32),
                                       counts)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.data.cpu().numpy(), self.indices.cpu().numpy(), self.indptr.cpu().numpy()),
                             shape=self.shape)


def csr(data, indices, indptr, shape):
    """(backend.py:35-36)."""
    return CSR(data, indices, indptr, shape)


def csr_from_host(indptr, indices, data, n, dtype=torch.float32):
    return CSR(asarray(data, dtype), asarray(indices, torch.int32), asarray(indptr, torch.int32), (n, n))


class _SparseModule:
    """`sp` of the reference's backend (backend.py:6-7, its array module's scipy.sparse): here
    only the CSR constructor, over device tensors."""

    @staticmethod
    def csr_matrix(arg, shape=None, dtype=None):
        """csr_matrix((data, indices, indptr), shape=...) or csr_matrix(<SciPy sparse matrix>):
        a CSR of device tensors (int32 indices, FP32 or FP64 values)."""
        if isinstance(arg, tuple):
            data, indices, indptr = arg
            if shape is None:
                raise ValueError("csr_matrix((data, indices, indptr)) needs shape")
        else:
            m = arg.tocsr()
            data, indices, indptr, shape = m.data, m.indices, m.indptr, m.shape
        src = dtype if dtype is not None else getattr(data, "dtype", "float32")
        dt = torch.float64 if str(src) in ("float64", "torch.float64", "<class 'numpy.float64'>") else torch.float32
        return CSR(asarray(data, dt), asarray(indices, torch.int32), asarray(indptr, torch.int32),
                   (int(shape[0]), int(shape[1])))


sp = _SparseModule()

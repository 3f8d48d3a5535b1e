"""The no-allocation connected-components entry point (dlp_cc_csr_static) against the
workspace entry point it is a variant of (dlp_cc_csr), through the C ABI: labels and component
counts bit-exact on two calls in a row, the second on a smaller graph (the scratch keeps the
first call's contents), and sizes beyond the scratch refused before any device work."""
import ctypes as ct

import numpy as np
import pytest

from dynlp_amd import graphs

pytestmark = pytest.mark.gpu


def _csr_on_device(g, mask_frac, seed):
    import torch
    from dynlp_amd.backend import device
    ip, ix, _ = g.csr()
    dev = device()
    rs = np.random.RandomState(seed)
    mask = (rs.rand(g.n) >= mask_frac).astype(np.uint8)
    t = lambda a, dt: torch.as_tensor(a).to(dev, dt).contiguous()  # noqa: E731
    return t(ip, torch.int32), t(ix, torch.int32), t(mask, torch.uint8)


def _run_both(n, ip, ix, mask):
    import torch
    from dynlp_amd import _lib
    from dynlp_amd.backend import device, ptr, stream
    dev = device()
    lab_ws = torch.full((n,), -7, dtype=torch.int32, device=dev)
    lab_st = torch.full((n,), -9, dtype=torch.int32, device=dev)
    nc = torch.zeros(2, dtype=torch.int32, device=dev)
    wsb = _lib.load().dlp_cc_workspace_bytes(n)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    _lib.call("dlp_cc_csr", n, ptr(ip), None, ptr(ix), ptr(mask), ptr(lab_ws), ptr(nc), ptr(ws), wsb, stream())
    _lib.call("dlp_cc_csr_static", n, ptr(ip), None, ptr(ix), ptr(mask), ptr(lab_st), ptr(nc) + 4, stream())
    torch.cuda.synchronize()
    return lab_ws.cpu().numpy(), lab_st.cpu().numpy(), nc.cpu().numpy()


def test_static_cc_matches_workspace_cc_on_two_calls():
    import torch
    from oracle import lp as O
    big = graphs.sbm(200_000, 32, 6, 0.9, 4, 0.01, seed=3)
    small = graphs.erdos_renyi(5_000, 2, 4, 0.01, 4, unit=True)
    for g, frac, seed in ((big, 0.3, 1), (small, 0.1, 2)):
        ip, ix, mask = _csr_on_device(g, frac, seed)
        free0 = torch.cuda.mem_get_info()[0]
        a, b, nc = _run_both(g.n, ip, ix, mask)
        assert np.array_equal(a, b), g.name
        assert nc[0] == nc[1] and nc[0] == int(a.max()) + 1
        # against the oracle: vertices outside the mask are singletons
        m = mask.cpu().numpy().astype(bool)
        u, v = g.u[m[g.u] & m[g.v]], g.v[m[g.u] & m[g.v]]
        ref, nref = O.connected_components(g.n, np.r_[u, v], np.r_[v, u])
        assert np.array_equal(b, ref) and nc[1] == nref
        # the static call allocated nothing (its array exists since the first call)
        _lib_static_again(g.n, ip, ix, mask)
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] >= free0 - (64 << 20)


def _lib_static_again(n, ip, ix, mask):
    import torch
    from dynlp_amd import _lib
    from dynlp_amd.backend import device, ptr, stream
    lab = torch.empty(n, dtype=torch.int32, device=device())
    nc = torch.empty(1, dtype=torch.int32, device=device())
    free0 = torch.cuda.mem_get_info()[0]
    _lib.call("dlp_cc_csr_static", n, ptr(ip), None, ptr(ix), ptr(mask), ptr(lab), ptr(nc), stream())
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0


def test_static_cc_refuses_sizes_beyond_its_scratch():
    from dynlp_amd import _lib
    lib = _lib.load()
    nmax = lib.dlp_cc_static_max_n()
    assert nmax >= 2_000_000
    rc = lib.dlp_cc_csr_static(nmax + 1, None, None, None, None, None, None, None)
    assert rc == -1 and b"scratch" in lib.dlp_last_error()
    # n = 0 writes a zero count
    import torch
    from dynlp_amd.backend import device, ptr, stream
    nc = torch.full((1,), 5, dtype=torch.int32, device=device())
    _lib.call("dlp_cc_csr_static", 0, None, None, None, None, None, ptr(nc), stream())
    assert int(nc.item()) == 0

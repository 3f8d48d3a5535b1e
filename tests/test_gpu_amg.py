"""The AMG's coarsest level in FP64 on the device against NumPy and the oracle."""
import numpy as np
import pytest

from dynlp_amd import graphs
from oracle import lp as O

pytestmark = pytest.mark.gpu


def test_spd_inverse_and_apply_match_numpy():
    """Gauss-Jordan inverse of an SPD operator diag - W and its FP64 application, against
    numpy.linalg.inv (relative 1e-11; a condition number near 1e4)."""
    import torch
    from dynlp_amd import _lib
    from dynlp_amd.backend import csr_from_host, ptr, stream
    g = graphs.sbm(700, 7, 6, 0.8, 2, 0.02, seed=11)
    ip, ix, d = g.csr()
    W = csr_from_host(ip, ix, d, g.n, torch.float64)
    rs = np.random.RandomState(0)
    diag = np.bincount(np.repeat(np.arange(g.n), np.diff(ip)), weights=d.astype(np.float64), minlength=g.n)
    diag = diag + rs.uniform(0.01, 0.1, g.n)
    A_ref = np.diag(diag) - O.csr(ip, ix, d, g.n).toarray()
    inv_ref = np.linalg.inv(A_ref)
    A = torch.empty((g.n, g.n), dtype=torch.float64, device="cuda")
    scratch = torch.empty_like(A)
    dg = torch.as_tensor(diag, device="cuda")
    _lib.call("dlp_dense_operator_f64", g.n, ptr(W.indptr), ptr(W.indices), ptr(W.data), ptr(dg), _lib.F64, ptr(A),
              stream())
    np.testing.assert_array_equal(A.cpu().numpy(), A_ref)
    _lib.call("dlp_spd_inverse_f64", g.n, ptr(A), ptr(scratch), stream())
    inv = A.cpu().numpy()
    assert np.abs(inv - inv_ref).max() <= 1e-11 * np.abs(inv_ref).max()
    for dt, tol in ((torch.float64, 1e-11), (torch.float32, 1e-5)):
        R = torch.as_tensor(rs.standard_normal((g.n, 37)), device="cuda").to(dt)
        X = torch.empty_like(R)
        _lib.call("dlp_dense_apply_f64", g.n, 37, ptr(A), ptr(R), 37, ptr(X), 37, 1 if dt == torch.float64 else 0,
                  stream())
        want = inv_ref @ R.double().cpu().numpy()
        np.testing.assert_allclose(X.double().cpu().numpy(), want, rtol=tol, atol=tol * np.abs(want).max())


def test_amg_coarse_inverse_is_fp64_and_matches_oracle():
    """The AMG's coarse inverse is kept in FP64 and equals the inverse of the oracle's coarsest
    operator on the same graph; M(I) as a whole matches the oracle's V-cycle (FP64 system)."""
    import torch
    from dynlp_amd.amg import AMG
    from dynlp_amd.backend import csr_from_host
    g = graphs.sbm(1500, 12, 6, 0.85, 2, 0.02, seed=7)
    sysm = O.build_system(g.n, g.u, g.v, g.w, g.labeled, g.y, 2)
    Wl, dl, rl = sysm.local()
    s = sysm.s[sysm.U]
    ref = O.AMG(Wl, s, max_coarse=50)
    W = csr_from_host(Wl.indptr, Wl.indices, Wl.data, Wl.shape[0], torch.float64)
    amg = AMG(W, torch.as_tensor(s, device="cuda"), max_coarse=50)
    assert amg.sizes == ref.sizes and amg.inv.dtype == torch.float64
    np.testing.assert_allclose(amg.inv.cpu().numpy(), ref.inv, rtol=1e-9, atol=1e-12 * np.abs(ref.inv).max())
    Xr = np.random.RandomState(1).standard_normal((Wl.shape[0], 5))
    got = amg(torch.as_tensor(Xr, device="cuda")).cpu().numpy()
    np.testing.assert_allclose(got, ref(Xr), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("mode", [0, 1])
def test_level_pass_matches_dense(mode):
    """dlp_amg_level against the dense FP64 operator: mode 0 R - A X, mode 1 one weighted-Jacobi
    sweep X + omega (R - A X) / diag (FP32 blocks, relative 1e-5), rows of 1 to 96 entries."""
    import torch
    from dynlp_amd import _lib
    from dynlp_amd.backend import csr_from_host, ptr, stream
    g = graphs.sbm(2500, 5, 12, 0.8, 2, 0.02, seed=13)
    ip, ix, d = g.csr()
    W = csr_from_host(ip, ix, d, g.n, torch.float32)
    rs = np.random.RandomState(1)
    C = 68
    diag = (np.bincount(np.repeat(np.arange(g.n), np.diff(ip)), weights=d.astype(np.float64), minlength=g.n)
            + rs.uniform(0.01, 0.1, g.n)).astype(np.float32)
    X = rs.standard_normal((g.n, C)).astype(np.float32)
    R = rs.standard_normal((g.n, C)).astype(np.float32)
    A = np.diag(diag.astype(np.float64)) - O.csr(ip, ix, d, g.n).toarray()
    res = R.astype(np.float64) - A @ X.astype(np.float64)
    ref = res if mode == 0 else X + 0.7 * res / diag[:, None].astype(np.float64)
    dev = lambda a: torch.as_tensor(a, device="cuda").contiguous()  # noqa: E731
    Xd, Rd, dd = dev(X), dev(R), dev(diag)
    out = torch.empty_like(Xd)
    _lib.call("dlp_amg_level", g.n, C, ptr(W.indptr), ptr(W.indices), ptr(W.data), ptr(dd), ptr(Xd), ptr(Rd), C,
              0.7, mode, ptr(out), stream())
    scale = np.abs(ref).max()
    assert np.abs(out.cpu().numpy() - ref).max() <= 1e-5 * scale


def test_fp32_vcycle_pipelined_levels_match_row_operator_levels():
    """The V-cycle with its sparse levels on the pipelined passes equals the one on the row
    operator (FP32 rounding: relative 1e-5), and stays symmetric positive definite."""
    import torch
    from dynlp_amd import amg as amgmod
    from dynlp_amd.amg import AMG
    from dynlp_amd.solvers import build_system
    g = graphs.sbm(3000, 12, 10, 0.85, 4, 0.02, seed=14)
    s = build_system(g, "float32")
    a = AMG(s.W, s.s, max_coarse=60)
    assert len(a.levels) >= 2 and a.levels[0].W.avg_deg <= amgmod.PIPE_MAX_DEG
    R = torch.rand((s.n_u, 7), dtype=torch.float32, device="cuda", generator=torch.Generator(device="cuda").manual_seed(3))
    keep = amgmod.PIPE_MAX_DEG
    try:
        amgmod.PIPE_MAX_DEG = 0
        Y0 = a(R).double().cpu().numpy()
        amgmod.PIPE_MAX_DEG = 1 << 30
        Y1 = a(R).double().cpu().numpy()
    finally:
        amgmod.PIPE_MAX_DEG = keep
    assert np.abs(Y1 - Y0).max() <= 1e-5 * np.abs(Y0).max()
    Rn = R.double().cpu().numpy()
    assert np.all(np.einsum("ij,ij->j", Rn, Y1) > 0)      # r^T M r > 0 per column

"""The HIP path on the committed golden fixtures (inputs and expected outputs written by the
FP64 oracle, tests/golden/make_fixtures.py)."""
import numpy as np
import pytest

from tests.golden import fixtures as F

pytestmark = pytest.mark.gpu


def test_spmm_fixture():
    import torch
    from dynlp_amd import backend as B
    from dynlp_amd.kernels import spmm_axpy
    fx = F.load("spmm")
    n = fx["indptr"].shape[0] - 1
    W = B.csr_from_host(fx["indptr"], fx["indices"], fx["data"], n, torch.float32)
    a = torch.as_tensor(fx["a"], device="cuda")
    for C in (1, 17, 65):
        X = torch.as_tensor(fx[f"X{C}"], device="cuda")
        Y = spmm_axpy(W, X, a=a, b=-1.0).double().cpu().numpy()
        np.testing.assert_allclose(Y, fx[f"Y{C}"], rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("name", ["sys_er_k2", "sys_er_k16", "sys_sbm_k64"])
@pytest.mark.parametrize("dirs", ["fp32", "bf16", "fp16", "fp8"])
def test_system_fixture_initial_solve(name, dirs):
    """Initial solve of the stored graph: solved set exact, diag to 1e-12, every label column
    within the certificate rho * hmax of the stored F*."""
    from dynlp_amd.stream import EdgeStreamLP
    fx = F.load(name)
    g = F.graph(fx)
    lp = EdgeStreamLP(g, eps=1e-6, pcg_directions=dirs)
    st = lp.initial_solve()
    assert np.array_equal(lp.solved_mask(), fx["solved"].astype(bool))
    np.testing.assert_allclose(lp.diag64.cpu().numpy(), fx["diag"], rtol=1e-12)
    Z, Zs = lp.scores_with_h(), fx["Z"]
    C = Zs.shape[1] - 1
    hmax = Zs[:, C].max()
    assert st.rho <= 1e-6 * 1.0001
    assert np.abs(Z[:, :C] - Zs[:, :C]).max() <= 1e-6 * hmax * 1.05
    assert np.abs(Z[:, C] - Zs[:, C]).max() <= 0.05 * hmax   # h to its own looser threshold


@pytest.mark.parametrize("fp64", [False, True])
def test_stream_fixture(fp64):
    """Five batches replayed on the resident path: per batch the solved set and CC labels match
    the stored ones exactly and the label columns are within the certificate of the stored F*."""
    from dynlp_amd.stream import EdgeStreamLP
    fx = F.load("stream_sbm_k8")
    g = F.graph(fx)
    lp = EdgeStreamLP(g, eps=1e-6, fp64=fp64)
    lp.initial_solve()
    for t, b in enumerate(F.batches(fx)):
        st = lp.apply_and_solve(b)
        assert np.array_equal(lp.solved_mask(), fx[f"b{t}_solved"].astype(bool)), t
        assert np.array_equal(lp.components(), fx[f"b{t}_cc"]), t
        Z, Zs = lp.scores_with_h(), fx[f"b{t}_Z"]
        C = Zs.shape[1] - 1
        assert st.rho <= 1e-6 * 1.0001, t
        assert np.abs(Z[:, :C] - Zs[:, :C]).max() <= 1e-6 * Zs[:, C].max() * 1.05, t
        # the request's own certificate (solvers.py:295-301): hmax bounds the exact h from
        # above, and rho * hmax bounds the error of every label score
        assert np.isfinite(st.cert) and st.hmax >= Zs[:, C].max() * (1 - 1e-9), t
        assert np.abs(Z[:, :C] - Zs[:, :C]).max() <= st.cert, t
        assert st.cert == pytest.approx(st.rho * st.hmax, rel=1e-12), t


def test_cc_fixture_bit_exact():
    """CC labels of the fragmenting stream, bit-exact: on the edge list and on the resident
    slack CSR after each batch (its grounding too)."""
    from dynlp_amd.cc import connected_components
    from dynlp_amd.stream import EdgeStreamLP
    fx = F.load("cc_fragmenting")
    g = F.graph(fx)
    lp = EdgeStreamLP(g, eps=1e-6)
    lp.initial_solve()
    es = {(int(a), int(c)) for a, c in zip(g.u, g.v)}
    for t, b in enumerate(F.batches(fx)):
        lp.apply_and_solve(b)
        es -= {(int(a), int(c)) for a, c in zip(b.del_u, b.del_v)}
        es |= {(int(a), int(c)) for a, c in zip(b.ins_u, b.ins_v)}
        u = np.array([p[0] for p in es], dtype=np.int32)
        v = np.array([p[1] for p in es], dtype=np.int32)
        lab, nc = connected_components(g.n, np.r_[u, v], np.r_[v, u])
        assert np.array_equal(lab.cpu().numpy(), fx[f"b{t}_cc"]), t
        assert np.array_equal(lp.components(), fx[f"b{t}_cc"]), t
        assert np.array_equal(lp.solved_mask(), fx[f"b{t}_solved"].astype(bool)), t


def test_amg_fixture():
    """Device AMG on the stored system: same hierarchy sizes and first-level aggregates, and
    M X equal to the stored V-cycle output."""
    import torch
    from dynlp_amd.amg import AMG
    from dynlp_amd.solvers import build_system
    fx = F.load("amg")
    s = build_system(F.graph(fx), "float64")
    amg = AMG(s.W, s.s, max_coarse=20)
    assert amg.sizes == fx["sizes"].tolist()
    assert np.array_equal(amg.levels[0].agg.cpu().numpy(), fx["agg0"])
    MX = amg(torch.as_tensor(fx["X"], device="cuda")).cpu().numpy()
    np.testing.assert_allclose(MX, fx["MX"], rtol=1e-10, atol=1e-10)

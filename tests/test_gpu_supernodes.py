"""Device scatters and supernode components against the oracle (through the C ABI)."""
import numpy as np
import pytest

from oracle import lp as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt", ["float32", "float64", "int32", "int64"])
@pytest.mark.parametrize("op", ["add", "max", "min"])
def test_scatter_matches_numpy(op, dt):
    """Unbuffered scatters with repeated indices, one or several columns and a broadcast
    scalar (backend.py:11-28): integer results exact; floating add to rounding (atomic order),
    max/min exact."""
    import torch
    from dynlp_amd import backend as B
    rs = np.random.RandomState(1)
    n, m = 500, 20000
    idx = rs.randint(0, n, m)
    for cols in (1, 3):
        shape = (n,) if cols == 1 else (n, cols)
        a0 = (rs.standard_normal(shape) * 100).astype(dt)
        v = (rs.standard_normal((m,) + shape[1:]) * 100).astype(dt)
        ref = a0.copy()
        getattr(np, {"add": "add", "max": "maximum", "min": "minimum"}[op]).at(ref, idx, v)
        a = torch.as_tensor(a0).cuda()
        getattr(B, f"scatter_{op}")(a, torch.as_tensor(idx).cuda(), torch.as_tensor(v).cuda())
        got = a.cpu().numpy()
        if dt.startswith("float") and op == "add":
            np.testing.assert_allclose(got, ref, rtol=1e-4 if dt == "float32" else 1e-12, atol=1e-2)
        else:
            assert np.array_equal(got, ref), (op, dt, cols)
    # one value for every index (int32 indices)
    a0 = np.zeros(n, dtype=dt)
    ref = a0.copy()
    getattr(np, {"add": "add", "max": "maximum", "min": "minimum"}[op]).at(ref, idx, np.array(3, dtype=dt))
    a = torch.as_tensor(a0).cuda()
    getattr(B, f"scatter_{op}")(a, torch.as_tensor(idx.astype(np.int32)).cuda(), 3)
    assert np.array_equal(a.cpu().numpy(), ref)


def _random_fh_case(rs, m, ne, integer_d):
    a = rs.randint(0, m, ne)
    b = rs.randint(0, m, ne)
    if integer_d:   # many exact ties
        d = rs.randint(0, 6, ne).astype(np.float32) / 8
    else:
        d = rs.uniform(0.0, 1.0, ne).astype(np.float32)
    rows, cols = np.r_[a, b], np.r_[b, a]
    return rows, cols, np.r_[d, d]


@pytest.mark.parametrize("integer_d", [False, True])
def test_fh_components_match_oracle(integer_d):
    """Labels and counts bit-exact against the oracle's Boruvka FH rounds, with and without
    the grounding pass, over a range of k (from nothing merging to everything)."""
    from dynlp_amd.cc import fh_components
    rs = np.random.RandomState(7 + integer_d)
    for trial in range(12):
        m = int(rs.choice([5, 60, 3000, 40000]))
        ne = int(m * rs.uniform(0.5, 4))
        rows, cols, d = _random_fh_case(rs, m, ne, integer_d)
        k = float(rs.choice([0.0, 0.05, 0.3, 1.0, 5.0]))
        ground = None
        if trial % 2:
            ground = (rs.rand(m) < 0.05).astype(np.float32) * rs.uniform(0.5, 2, m).astype(np.float32)
        want, nw = O.fh_components(m, rows, cols, d, k, ground)
        got, ng = fh_components(m, rows, cols, d, k, ground)
        assert ng == nw, (trial, m, k)
        assert np.array_equal(got.cpu().numpy(), want), (trial, m, k)


def test_adaptive_components_and_scales_match_oracle():
    """The local scales, the dissimilarities and the adaptive components of new vertices
    (stream.py:67-81, cc.py:77-83) against the oracle, FP32 and FP64 weights."""
    import torch
    from dynlp_amd.cc import adaptive_components, vertex_scales
    rs = np.random.RandomState(3)
    for dt in (np.float32, np.float64):
        m, ne = 5000, 30000
        lr = rs.randint(0, m, ne)
        lc = rs.randint(0, m, ne)
        w = rs.uniform(0.0, 1.0, ne).astype(dt)
        w[rs.rand(ne) < 0.1] = 0
        wmax = 1.0
        s_want = O.vertex_scales(m, lr, w, wmax)
        s = vertex_scales(m, torch.as_tensor(lr).cuda(), torch.as_tensor(w).cuda(), wmax).cpu().numpy()
        np.testing.assert_allclose(s, s_want, rtol=2e-6)          # FP32 atomic sums
        on = w > 0
        r, c, we = lr[on], lc[on], w[on]
        d, k = O.adaptive_dissimilarity(we, s[r], s[c], wmax)      # the device's scales, exactly
        want, nw = O.fh_components(m, r, c, d, k)
        got, ng = adaptive_components(m, r, c, we, s[r], s[c], wmax)
        assert ng == nw and np.array_equal(got.cpu().numpy(), want)
        got2, ng2 = adaptive_components(m, r, c, we, s, s, wmax, scale_index=(r, c))
        assert ng2 == nw and np.array_equal(got2.cpu().numpy(), want)

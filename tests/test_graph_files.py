"""Edge-list files (the reference's graphs.py:176-214 `from_npz` / `build_cached` / `load`):
round trips, duplicate and loop handling against scipy's COO coalescing, defaults, errors."""
import numpy as np
import pytest
import scipy.sparse as sp

from dynlp_amd import graphs


def test_round_trip_keeps_graph_and_order(tmp_path):
    g = graphs.sbm(3000, 6, 8, 0.8, 6, 0.02, seed=3)
    g.meta["order"] = np.random.default_rng(5).permutation(g.n)
    path = tmp_path / "g.npz"
    graphs.save_npz(g, path)
    with np.load(path) as z:
        assert set(z.files) == {"src", "dst", "w", "y", "order"}
        assert np.all(z["src"] < z["dst"])                      # upper triangle, as build_cached writes
    h = graphs.from_npz(path, label_frac=0.02, seed=3)
    assert (h.n, h.K, h.m) == (g.n, g.K, g.m)
    assert np.array_equal(h.u, g.u) and np.array_equal(h.v, g.v)
    assert np.array_equal(h.w, g.w) and np.array_equal(h.y, g.y)
    assert np.array_equal(h.meta["order"], g.meta["order"])
    ip, ix, d = h.csr()
    jp, jx, e = g.csr()
    assert np.array_equal(ip, jp) and np.array_equal(ix, jx) and np.array_equal(d, e)


def test_duplicates_summed_loops_dropped_like_scipy_coalesce(tmp_path):
    """Against scipy's own COO duplicate summation of both directions (what the reference's
    `coalesce` calls), on a multigraph with repeated pairs in both directions and self-loops;
    every pair occurs at most twice so the FP32 sums do not depend on the order."""
    rng = np.random.default_rng(11)
    n, m = 400, 3000
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    key = np.minimum(a, b) * n + np.maximum(a, b)
    _, idx, cnt = np.unique(key, return_index=True, return_counts=True)
    keep = np.isin(key, key[idx[cnt <= 2]])
    a, b = a[keep], b[keep]
    w = rng.uniform(0.5, 1.5, a.shape[0]).astype(np.float32)
    assert np.any(a == b) and np.any(cnt == 2)
    y = rng.integers(0, 5, n).astype(np.int32)
    path = tmp_path / "multi.npz"
    np.savez(path, src=a, dst=b, w=w, y=y)
    g = graphs.from_npz(path)
    ip, ix, d = g.csr()
    r, c, v = np.r_[a, b], np.r_[b, a], np.r_[w, w]
    off = r != c
    A = sp.coo_matrix((v[off], (r[off], c[off])), shape=(n, n), dtype=np.float32).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    assert np.array_equal(ip, A.indptr) and np.array_equal(ix, A.indices)
    assert np.array_equal(d, A.data)
    assert np.all(g.u < g.v)


def test_small_case_by_hand(tmp_path):
    path = tmp_path / "tiny.npz"
    np.savez(path, src=np.array([0, 1, 2, 2, 3, 0]), dst=np.array([1, 0, 2, 3, 2, 3]),
             w=np.array([1, 2, 5, 0.5, 0.25, 4], dtype=np.float32), y=np.array([0, 1, -1, 2]))
    g = graphs.from_npz(path, label_frac=1.0)
    assert (g.n, g.K) == (4, 3)
    assert g.u.tolist() == [0, 2, 0] and g.v.tolist() == [1, 3, 3]
    assert g.w.tolist() == [3.0, 0.75, 4.0]
    assert g.labeled.tolist() == [True, True, False, True]  # y < 0 is never labelled


def test_defaults_unit_weights_and_order(tmp_path):
    path = tmp_path / "now.npz"
    np.savez(path, src=np.array([0, 1, 2]), dst=np.array([1, 2, 3]), y=np.array([0, 0, 1, 1]))
    g = graphs.from_npz(path, seed=7)
    assert g.w.dtype == np.float32 and np.all(g.w == 1)
    assert np.array_equal(g.meta["order"], np.random.RandomState(8).permutation(4))


@pytest.mark.parametrize("arrays,msg", [
    (dict(src=[0], dst=[1]), "missing"),
    (dict(src=[0], dst=[4], y=[0, 1, 0, 1]), "vertex ids"),
    (dict(src=[0], dst=[1], y=[0, 1], w=[0.0]), "positive"),
    (dict(src=[0], dst=[1], y=[0, 1], order=[0, 0]), "permutation"),
    (dict(src=[0], dst=[1], y=[-1, -1]), "no vertex"),
])
def test_bad_files_rejected(tmp_path, arrays, msg):
    path = tmp_path / "bad.npz"
    np.savez(path, **{k: np.asarray(v) for k, v in arrays.items()})
    with pytest.raises(ValueError, match=msg):
        graphs.from_npz(path)


def test_load_switch(tmp_path):
    g = graphs.load("sbm", n=2000, K=4, deg=6, p_in=0.9, seed=1)
    assert g.n == 2000 and g.K == 4 and g.meta["kind"] == "sbm"
    e = graphs.load("er", n=1000, K=3, deg=5, seed=2)
    assert e.n == 1000 and e.K == 3
    path = tmp_path / "e.npz"
    graphs.save_npz(e, path)
    f = graphs.load(str(path), seed=2)
    assert np.array_equal(f.u, e.u) and np.array_equal(f.w, e.w)
    with pytest.raises(ValueError):
        graphs.load("imdb")

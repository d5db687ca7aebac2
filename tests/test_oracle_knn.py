"""Pins the oracle's kNN: grid search == brute force bit for bit, brute force ==
scipy's cKDTree where the cut is untied, the group masks == the reference's
Inf-masked searches (src/calculateneighbors.jl:23-41), and the reference
fixture's cuts are untied (so its sparsity pattern is geometry-determined)."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from rbffd_amd.nodes import jittered_grid


def _sets():
    rng = np.random.default_rng(7)
    return {
        "jittered": jittered_grid(40, 31, 0.25, seed=3),
        "regular": jittered_grid(30, 30, 0.0, seed=0),        # many exact ties
        "uniform": rng.uniform(0, 1, size=(1500, 2)),
        "clustered": np.concatenate([rng.normal(0.2, 0.01, size=(500, 2)),
                                     rng.normal(0.7, 0.2, size=(700, 2))]),
        "line": np.stack([np.linspace(0, 1, 300), np.zeros(300)], axis=1),
        "duplicates": np.repeat(rng.uniform(0, 1, size=(200, 2)), 2, axis=0),
    }


@pytest.mark.parametrize("name", list(_sets().keys()))
@pytest.mark.parametrize("k", [1, 13, 31, 57])
def test_grid_equals_brute(O, name, k):
    X = _sets()[name]
    rng = np.random.default_rng(k)
    Q = np.concatenate([X, rng.uniform(-0.2, 1.2, size=(200, 2))])
    ib, db = O.knn(X, Q, k, method="brute")
    ig, dg = O.knn(X, Q, k, method="grid")
    assert np.array_equal(ib, ig)
    assert np.array_equal(db, dg)
    # ascending (d2, index)
    assert np.all((np.diff(db, axis=1) > 0) | ((np.diff(db, axis=1) == 0) & (np.diff(ib, axis=1) > 0)))


def test_brute_matches_kdtree_where_untied(O):
    X = _sets()["uniform"]
    k = 20
    ib, db = O.knn(X, X, k, method="brute")
    dt, it = cKDTree(X).query(X, k=k + 1)
    untied = dt[:, k - 1] < dt[:, k]
    assert untied.mean() > 0.99
    for q in np.nonzero(untied)[0]:
        assert set(ib[q]) == set(it[q, :k])
    assert np.allclose(np.sqrt(db), dt[:, :k], rtol=1e-14, atol=0)
    assert np.array_equal(ib[:, 0], np.arange(len(X)))  # self first


def test_groups_equal_inf_masked_search(O):
    # a small domain with two boundary markers and their ghosts, searched the
    # way the reference does it: invisible nodes moved to (Inf, Inf)
    rng = np.random.default_rng(5)
    inner = rng.uniform(0.05, 0.95, size=(300, 2))
    t = np.linspace(0, 1, 25)
    left = np.stack([np.zeros_like(t), t], 1)
    bottom = np.stack([t[1:], np.zeros_like(t[1:])], 1)
    gl = left - [0.04, 0.0]
    gb = bottom - [0.0, 0.04]
    X = np.concatenate([inner, left, bottom, gl, gb])
    i0 = 0
    idx_in = np.arange(i0, i0 + 300)
    bc = [np.arange(300, 325), np.arange(325, 349)]
    gh = [np.arange(349, 374), np.arange(374, 398)]
    xg, qg = O.boundary_groups(len(X), idx_in, bc, gh)
    n = 13
    got, _ = O.knn(X, X, n, xg, qg, method="brute")
    got_grid, _ = O.knn(X, X, n, xg, qg, method="grid")
    assert np.array_equal(got, got_grid)
    for b in range(2):
        Xm = np.full_like(X, np.inf)
        vis = np.concatenate([idx_in, bc[0], bc[1], gh[b]])
        Xm[vis] = X[vis]
        for q in np.concatenate([bc[b], gh[b]]):
            ref, _ = O.knn(Xm, X[q:q + 1], n, method="brute")
            assert np.array_equal(got[q], ref[0])
    full, _ = O.knn(X, X[idx_in], n, method="brute")
    assert np.array_equal(got[idx_in], full)


@pytest.mark.parametrize("n", [20, 42])
def test_reference_fixture_cut_untied(O, fitted_nodes, n):
    X, _ = fitted_nodes
    _, d2 = O.knn(X, X, n + 1, method="brute")
    assert np.all(d2[:, n - 1] < d2[:, n])

"""K1 (kNN) on the GPU against the oracle: indices and squared distances bit
for bit, including exact ties, masks and far-away queries."""
import numpy as np
import pytest

from rbffd_amd.nodes import jittered_grid

pytestmark = pytest.mark.gpu


def _run(ctx, X, Q, k, xg=None, qg=None):
    import rbffd_amd

    eng = rbffd_amd.Engine(ctx)
    Xd = ctx.to_device(X)
    Qd = Xd if Q is X else ctx.to_device(Q)
    xgd = ctx.to_device(xg.astype(np.int32)) if xg is not None else None
    qgd = ctx.to_device(qg.astype(np.int32)) if qg is not None else None
    idx, d2 = eng.knn(Xd, Qd, k, xgd, qgd, want_d2=True)
    return idx.get(), d2.get()


def _sets():
    rng = np.random.default_rng(17)
    return {
        "poisson2500": jittered_grid(51, 51, 0.25, seed=1),
        "regular_ties": jittered_grid(40, 40, 0.0, seed=0),
        "uniform": rng.uniform(0, 1, size=(3000, 2)),
        "clustered": np.concatenate([rng.normal(0.3, 0.01, size=(800, 2)), rng.normal(0.6, 0.2, size=(900, 2))]),
        "duplicates": np.repeat(rng.uniform(0, 1, size=(600, 2)), 2, axis=0),
        "line": np.stack([np.linspace(0, 1, 500), np.full(500, 0.25)], axis=1),
        # one cell per grid row, and a few: the row runs of the cell walk clamp at both ends
        "vertical_line": np.stack([np.full(400, 0.5), np.linspace(0, 2, 400)], axis=1),
        "strip": np.stack([rng.uniform(0, 0.02, size=1500), rng.uniform(0, 1, size=1500)], axis=1),
    }


@pytest.mark.parametrize("name", list(_sets().keys()))
@pytest.mark.parametrize("k", [1, 13, 20, 31, 42, 57])
def test_knn_matches_oracle(ctx, O, name, k):
    X = _sets()[name]
    print(f"knn {name} k={k} N={len(X)}", flush=True)
    idx, d2 = _run(ctx, X, X, k)
    ri, rd = O.knn(X, X, k, method="brute")
    assert np.array_equal(idx, ri)
    assert np.array_equal(d2, rd)


def test_knn_queries_off_the_nodes(ctx, O, fitted_nodes):
    X, Y = fitted_nodes
    rng = np.random.default_rng(3)
    Q = np.concatenate([Y, rng.uniform(-3, 3, size=(500, 2))])
    for k in (1, 20):
        print(f"knn off-node queries k={k}", flush=True)
        idx, d2 = _run(ctx, X, Q, k)
        ri, rd = O.knn(X, Q, k, method="brute")
        assert np.array_equal(idx, ri)
        assert np.array_equal(d2, rd)


def test_knn_boundary_masks(ctx, O):
    rng = np.random.default_rng(5)
    inner = rng.uniform(0.05, 0.95, size=(600, 2))
    t = np.linspace(0, 1, 40)
    left = np.stack([np.zeros_like(t), t], 1)
    bottom = np.stack([t[1:], np.zeros_like(t[1:])], 1)
    X = np.concatenate([inner, left, bottom, left - [0.03, 0.0], bottom - [0.0, 0.03]])
    idx_in = np.arange(600)
    bc = [np.arange(600, 640), np.arange(640, 679)]
    gh = [np.arange(679, 719), np.arange(719, 758)]
    xg, qg = O.boundary_groups(len(X), idx_in, bc, gh)
    for k in (13, 31):
        print(f"knn masks k={k}", flush=True)
        idx, d2 = _run(ctx, X, X, k, xg, qg)
        ri, rd = O.knn(X, X, k, xg, qg, method="brute")
        assert np.array_equal(idx, ri)
        assert np.array_equal(d2, rd)


def test_knn_too_few_nodes_is_an_error(ctx):
    X = jittered_grid(3, 3, 0.0, seed=0)
    with pytest.raises(ValueError):
        _run(ctx, X, X, 10)


def test_knn_full_size_1m(ctx, O):
    X = jittered_grid(1024, 1024, 0.25, seed=2)
    print("knn 1M nodes k=31", flush=True)
    idx, d2 = _run(ctx, X, X, 31)
    ri, rd = O.knn(X, X, 31, method="grid", nthreads=16)
    assert np.array_equal(idx, ri)
    assert np.array_equal(d2, rd)


@pytest.mark.parametrize("k", [5, 8, 20])
def test_knn_masked_query_with_too_few_visible_nodes(ctx, O, k):
    # the 6 nodes of group 2 see only each other (no group-0 nodes), so with
    # k > 6 their searches come up short and the call is an error; with k <= 6
    # every query is filled and equals brute force
    rng = np.random.default_rng(9)
    X = rng.uniform(0, 1, size=(300, 2))
    g = np.ones(300, np.int32)
    g[:6] = 2
    print(f"knn short masked queries k={k}", flush=True)
    if k > 6:
        with pytest.raises(ValueError):
            _run(ctx, X, X, k, g, g)
    else:
        idx, d2 = _run(ctx, X, X, k, g, g)
        ri, rd = O.knn(X, X, k, g, g, method="brute")
        assert np.array_equal(idx, ri)
        assert np.array_equal(d2, rd)

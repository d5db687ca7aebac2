"""Stencil shapes without a specialised kernel, and output rows that share a
stencil (Y distinct from X), on the GPU against the oracle.

The reference accepts any stencil size n and polynomial degree
(src/generate_operator.jl:29, src/polynomialbasis.jl:6).  The library solves
them with a generic kernel that pads the system to a compiled size (16, 32,
48, 64, 96 or 128 equations) and, above 128 equations, with a kernel that keeps
the system in a device workspace.  With Y distinct from X the rows mapped to
one X node are solved against one factorisation of its system
(src/generate_operator.jl:55-65, :158).

Error measure and bar as in test_gpu_weights: per-row relative weight error
max_j |w_j - w_ref_j| / max_j |w_ref_j| against the high-precision
(__float128) oracle, at most 1e-10 or 4x the error of a backward-stable FP64
LU solve of the same systems (the oracle's solve="lu"), whichever is larger;
the CSR pattern bit for bit.
"""

import numpy as np
import pytest

from rbffd_amd.nodes import jittered_grid

pytestmark = pytest.mark.gpu

TOL = 1e-10
OPS = ["E", "Dx", "Dy", "Dxx", "Dyy", "Dxy", "Dxk", "Dyk", "Lap", "Lapk"]

# (PHS power, polynomial degree, stencil size, padded size): none of these has a
# specialised kernel.  An even power r^2k is a polynomial of degree 2k, so its
# matrix has rank at most (2k+1)(2k+2)/2 and most even-power systems are
# singular in exact arithmetic (an FP64 LU solve misses by O(1)); r^8 with
# degree 1 and n = 21 is a well-posed one (FP64 LU within 1e-10 of exact).
SHAPES = [
    (3, 1, 7, 16),      # n + P = 10
    (8, 1, 21, 32),     # even power, 24
    (5, 3, 25, 48),     # 35
    (7, 4, 35, 64),     # 50
    (3, 4, 40, 64),     # 55
    (5, 5, 50, 96),     # 71
    (9, 6, 64, 96),     # 92
    (3, 7, 90, 128),    # 126
    (5, 8, 100, 0),     # 145: workspace-resident kernel
]


def rowrel(got, ref):
    return np.abs(got - ref).max(axis=-1) / np.abs(ref).max(axis=-1)


def _opl(O, ops, k):
    return [(O.OP_NAMES[o], k if o in ("Dxk", "Dyk", "Lapk") else 0) for o in ops]


def _ref_csr(st_rows, W):
    order = np.argsort(st_rows, axis=1, kind="stable")
    return np.take_along_axis(st_rows, order, 1), np.take_along_axis(W, order, 1)


def _gpu_weights(ctx, X, Y, st, y2x, p, polydeg, ops, k, method):
    import rbffd_amd

    eng = rbffd_amd.Engine(ctx)
    Xd = ctx.to_device(X)
    Yd = Xd if Y is X else ctx.to_device(Y)
    std = ctx.to_device(st.astype(np.int32))
    y2xd = ctx.to_device(y2x.astype(np.int32)) if y2x is not None else None
    col, vals, status = eng.weights(Xd, Yd, std, y2xd, p, polydeg, ops, k=k, method=method)
    return col.get(), [v.get() for v in vals], status.get()


def _check(O, X, Y, st, y2x, p, polydeg, ops, k, method, col, vals, rows):
    """Compare the GPU rows `rows` with the quad oracle under the bar."""
    Wq, _, sq = O.weights(X, Y, st, y2x, p, polydeg, _opl(O, ops, k), method=method, quad=True, rows=rows)
    Wl, _, _ = O.weights(X, Y, st, y2x, p, polydeg, _opl(O, ops, k), method=method, solve="lu", rows=rows)
    assert not sq.any()
    st_rows = st[rows] if y2x is None else st[y2x[rows]]
    bad = []
    for o, name in enumerate(ops):
        rc, rq = _ref_csr(st_rows, Wq[:, o, :])
        _, rl = _ref_csr(st_rows, Wl[:, o, :])
        assert np.array_equal(col[rows], rc), name
        e = rowrel(vals[o][rows], rq).max()
        e_lu = rowrel(rl, rq).max()
        print(f"  {name}: vs quad {e:.2e} (fp64 LU {e_lu:.2e})", flush=True)
        if e > max(TOL, 4 * e_lu):
            bad.append((name, e, e_lu))
    assert not bad, bad


def _sample(M, count, seed):
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(M, size=min(M, count), replace=False))


@pytest.mark.parametrize("p,polydeg,n,nsb", SHAPES)
@pytest.mark.parametrize("method", [1, 3])
def test_shapes_collocated(ctx, O, p, polydeg, n, nsb, method):
    side = 20 if n <= 40 else 26
    X = jittered_grid(side, side + 1, 0.3, seed=n + method)
    k = 4 if polydeg >= 4 else 2
    ops = OPS if method == 1 else OPS[:8]
    print(f"shape p={p} deg={polydeg} n={n} (padded {nsb or 'workspace'}) method={method}", flush=True)
    st, _ = O.knn(X, X, n)
    col, vals, status = _gpu_weights(ctx, X, X, st, None, p, polydeg, ops, k, method)
    assert not status.any()
    # the large systems are slow in quad precision: a row sample (edges included)
    rows = np.arange(len(X)) if n <= 25 else _sample(len(X), 48, n)
    _check(O, X, X, st, None, p, polydeg, ops, k, method, col, vals, rows)


@pytest.mark.parametrize("p,polydeg,n", [(5, 3, 25), (8, 1, 21), (9, 6, 64), (5, 8, 100)])
def test_shapes_y_distinct_from_x(ctx, O, p, polydeg, n, fitted_nodes):
    X, Y = fitted_nodes
    print(f"shape X != Y p={p} deg={polydeg} n={n}", flush=True)
    st, _ = O.knn(X, X, n)
    y2x = O.knn(X, Y, 1)[0][:, 0]
    ops = ["E", "Dx", "Dy", "Dxx", "Dyy", "Dxy"]
    col, vals, status = _gpu_weights(ctx, X, Y, st, y2x, p, polydeg, ops, 0, 1)
    assert not status.any()
    rows = _sample(len(Y), 400 if n <= 25 else 48, n)
    _check(O, X, Y, st, y2x, p, polydeg, ops, 0, 1, col, vals, rows)


def test_grouped_rows_equal_single_rows(ctx, O, fitted_nodes, monkeypatch):
    """Rows that share a factorisation get the same bits as rows solved one per
    task (the generic kernel at n = 25, degree 3): every right-hand-side column
    sees the same operations wherever it sits in the task."""
    X, Y = fitted_nodes
    st, _ = O.knn(X, X, 25)
    y2x = O.knn(X, Y, 1)[0][:, 0]
    ops = ["Dx", "Dy", "Lap"]
    print("grouped vs one row per task, n = 25", flush=True)
    c1, v1, s1 = _gpu_weights(ctx, X, Y, st, y2x, 5, 3, ops, 0, 1)
    monkeypatch.setenv("RBFFD_GROUP_ROWS", "0")
    c2, v2, s2 = _gpu_weights(ctx, X, Y, st, y2x, 5, 3, ops, 0, 1)
    assert np.array_equal(c1, c2) and np.array_equal(s1, s2)
    for a, b in zip(v1, v2):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("p,polydeg,n", [(3, 3, 20), (3, 5, 42)])
def test_one_factorisation_per_x_node(ctx, O, p, polydeg, n, fitted_nodes):
    """generate_operator(X, Y, ...) on the reference's Poisson fixture (2000 X,
    6000 Y; test/poisson_test.jl:56) and with the mesh test's shape (n = 42,
    degree 5; test/mesh_import_test.jl:78) factors each X node's system once, as
    the reference's loop over X does (src/generate_operator.jl:55-65), and the
    weights stay within the bar."""
    import rbffd_amd

    X, Y = fitted_nodes
    eng = rbffd_amd.Engine(ctx)
    Xd, Yd = ctx.to_device(X), ctx.to_device(Y)
    ops = ["E", "Dx", "Dy", "Dxx", "Dyy", "Dxy"]
    y2x_d = ctx.empty((len(Y),), np.int32)
    st_d = ctx.empty((len(X), n), np.int32)
    col, vals, stats = eng.generate(Xd, Yd, p, n, polydeg, ops, stencils_out=st_d, y2x_out=y2x_d)
    y2x = y2x_d.get()
    used = len(np.unique(y2x))
    print(f"  n={n}: {len(Y)} rows, {used} X nodes used, {stats['factorizations']} factorisations, "
          f"solve {stats['ms_solve']:.3f} ms", flush=True)
    assert stats["factorizations"] == used
    st = st_d.get()
    rows = _sample(len(Y), 300, n)
    _check(O, X, Y, st, y2x, p, polydeg, ops, 0, 1, col.get(), [v.get() for v in vals], rows)


@pytest.mark.parametrize("p,polydeg,n", [(3, 3, 20), (7, 4, 31), (3, 5, 42)])
def test_grouped_specialised_equal_single_rows(ctx, O, p, polydeg, n, fitted_nodes, monkeypatch):
    """A shape with a specialised kernel solves Y rows grouped by X node in
    that kernel's one-slot layout (16, 32 or 64 lanes for n <= 16, 32, 64).
    Each right-hand-side column sees the operations it would see in a task
    of its own in that layout, so the grouped weights equal one-row-per-task
    weights in the same layout bit for bit; against the quad oracle they meet
    the bar, and the generic kernel's grouped weights agree with them to
    rounding."""
    X, Y = fitted_nodes
    st, _ = O.knn(X, X, n)
    y2x = O.knn(X, Y, 1)[0][:, 0]
    ops = ["Dx", "Dy", "Lap"]
    g = 16 if n <= 16 else (32 if n <= 32 else 64)
    print(f"grouped specialised vs one row per task, n = {n}, {g} lanes", flush=True)
    monkeypatch.setenv("RBFFD_GROUP_ROWS", "1")
    c1, v1, s1 = _gpu_weights(ctx, X, Y, st, y2x, p, polydeg, ops, 0, 1)
    assert not s1.any()
    monkeypatch.setenv("RBFFD_GROUP_ROWS", "0")
    monkeypatch.setenv("RBFFD_LU_GROUP", str(g))
    c2, v2, s2 = _gpu_weights(ctx, X, Y, st, y2x, p, polydeg, ops, 0, 1)
    assert np.array_equal(c1, c2) and np.array_equal(s1, s2)
    for name, a, b in zip(ops, v1, v2):
        nd = int((a != b).any(axis=-1).sum()) if a.ndim > 1 else int((a != b).sum())
        print(f"  {name}: rows differing {nd}, max rel {rowrel(a.reshape(len(Y), -1), b.reshape(len(Y), -1)).max():.3e}", flush=True)
    for a, b in zip(v1, v2):
        assert np.array_equal(a, b)
    monkeypatch.delenv("RBFFD_LU_GROUP")
    monkeypatch.setenv("RBFFD_GROUP_ROWS", "1")
    monkeypatch.setenv("RBFFD_TUNED_GROUPS", "0")
    c3, v3, s3 = _gpu_weights(ctx, X, Y, st, y2x, p, polydeg, ops, 0, 1)
    assert np.array_equal(c1, c3) and not s3.any()
    for a, b in zip(v1, v3):
        assert rowrel(a, b).max() < 1e-9
    rows = _sample(len(Y), 300, n)
    _check(O, X, Y, st, y2x, p, polydeg, ops, 0, 1, c1, v1, rows)


def test_grouped_specialised_operator_chunks(ctx, O, fitted_nodes, monkeypatch):
    """Nine operators run in chunks of six right-hand sides; with at most four
    Y rows per X node every chunk fits the specialised kernel's 24-column
    block (one factorisation per X node per chunk), and the grouped weights
    equal one-row-per-task weights in the same layout bit for bit (n = 31,
    degree 4, Dxk/Dyk at k = 3)."""
    X, Y = fitted_nodes
    n, p, polydeg = 31, 7, 4
    st, _ = O.knn(X, X, n)
    y2x_all = O.knn(X, Y, 1)[0][:, 0]
    # at most four rows per X node (the fixture has up to five)
    keep = np.ones(len(Y), bool)
    seen = {}
    for i, x in enumerate(y2x_all):
        seen[x] = seen.get(x, 0) + 1
        keep[i] = seen[x] <= 4
    Y = np.ascontiguousarray(Y[keep])
    y2x = y2x_all[keep]
    ops = ["E", "Dx", "Dy", "Dxx", "Dyy", "Dxy", "Lap", "Dxk", "Dyk"]
    print(f"grouped specialised, nine operators, {len(Y)} rows", flush=True)
    monkeypatch.setenv("RBFFD_GROUP_ROWS", "1")
    c1, v1, s1 = _gpu_weights(ctx, X, Y, st, y2x, p, polydeg, ops, 3, 1)
    assert not s1.any()
    monkeypatch.setenv("RBFFD_GROUP_ROWS", "0")
    monkeypatch.setenv("RBFFD_LU_GROUP", "32")
    c2, v2, s2 = _gpu_weights(ctx, X, Y, st, y2x, p, polydeg, ops, 3, 1)
    assert np.array_equal(c1, c2) and np.array_equal(s1, s2)
    for name, a, b in zip(ops, v1, v2):
        assert np.array_equal(a, b), name
    rows = _sample(len(Y), 200, 9)
    _check(O, X, Y, st, y2x, p, polydeg, ops, 3, 1, c1, v1, rows)

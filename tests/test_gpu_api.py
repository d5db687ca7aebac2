"""The Python mirror of the reference API on the GPU: generate_operator (three
methods), hyperviscosity_operator, the reference's integration tests run on
GPU-built operators, and the reference's error behaviour."""
import numpy as np
import pytest
import scipy.sparse as sp

from rbffd_amd.nodes import jittered_grid

pytestmark = pytest.mark.gpu
NAMES = ("E", "Dx", "Dy", "Dxx", "Dyy", "Dxy")


def _csc(t):
    cp, rv, nz, sh = t
    return sp.csc_matrix((nz, rv, cp), shape=sh)


def _lu_reference(O, X, p, n, polydeg, ops, k):
    st, _ = O.knn(X, X, n)
    opl = [(O.OP_NAMES[o], k) for o in ops]
    W, _, _ = O.weights(X, X, st, None, p, polydeg, opl, solve="lu")
    rows = np.repeat(np.arange(len(X)), n).reshape(len(X), n)
    return {name: O.sparse_csc(rows, st, W[:, j, :]) for j, name in enumerate(ops)}


def _rowerr(A, B):
    D = (A - B).tocsr()
    R = B.tocsr()
    num = np.zeros(A.shape[0])
    den = np.zeros(A.shape[0])
    np.maximum.at(num, np.repeat(np.arange(A.shape[0]), np.diff(D.indptr)), np.abs(D.data))
    np.maximum.at(den, np.repeat(np.arange(A.shape[0]), np.diff(R.indptr)), np.abs(R.data))
    return float((num / np.maximum(den, 1e-300)).max())


def _same(A, ref_tuple, tol=1e-10):
    B = _csc(ref_tuple)
    assert A.shape == B.shape
    assert np.array_equal(A.indptr, B.indptr)
    assert np.array_equal(A.indices, B.indices)
    # per-row relative error
    D = (A - B).tocsr()
    R = B.tocsr()
    num = np.zeros(A.shape[0])
    den = np.zeros(A.shape[0])
    np.maximum.at(num, np.repeat(np.arange(A.shape[0]), np.diff(D.indptr)), np.abs(D.data))
    np.maximum.at(den, np.repeat(np.arange(A.shape[0]), np.diff(R.indptr)), np.abs(R.data))
    err = (num / np.maximum(den, 1e-300)).max()
    assert err <= tol, err
    return err


def test_generate_operator_method1_and_poisson(ctx, O, fitted_nodes):
    import rbffd_amd
    from test_oracle_weights import load_poisson_case, poisson_error

    X, Y, yin, yneu, ydir, xn, yn = load_poisson_case(O, fitted_nodes)
    print("generate_operator(X, Y, 3, 20, 3) on the reference fixture", flush=True)
    E, Dx, Dy, Dxx, Dyy, Dxy = rbffd_amd.generate_operator(X, Y, 3, 20, 3, ctx=ctx)
    ref = O.generate_operator(X, Y, 3, 20, 3, quad=True)
    for name, A in zip(NAMES, (E, Dx, Dy, Dxx, Dyy, Dxy)):
        print(f"  {name} row error {_same(A, ref[name]):.2e}", flush=True)
    err = poisson_error([E, Dx, Dy, Dxx, Dyy], X, Y, yin, yneu, ydir, xn, yn, O)
    print(f"  Poisson relative L2 error {err:.6f}", flush=True)
    assert err < 0.0027   # test/poisson_test.jl:132


def test_generate_operator_method3(ctx, O):
    import rbffd_amd

    X = jittered_grid(30, 30, 0.25, seed=12)
    print("generate_operator(X, 5, 13, 2)", flush=True)
    mats = rbffd_amd.generate_operator(X, 5, 13, 2, ctx=ctx)
    ref = O.generate_operator(X, None, 5, 13, 2, quad=True)
    for name, A in zip(NAMES, mats):
        _same(A, ref[name])


def test_hyperviscosity_self_consistency(ctx, fitted_nodes):
    # test/hyperviscosity_test.jl:26-33
    import rbffd_amd

    X, _ = fitted_nodes
    print("hyperviscosity_operator(2, X, X, 5, 42, 5)", flush=True)
    E, Dx, Dy, Dxx, Dyy, Dxy = rbffd_amd.generate_operator(X, X, 5, 42, 5, ctx=ctx)
    Dxk, Dyk = rbffd_amd.hyperviscosity_operator(2, X, X, 5, 42, 5, ctx=ctx)
    for a, b in ((Dxk, Dxx), (Dyk, Dyy)):
        assert abs(sp.linalg.norm(a - b)) <= np.sqrt(2.2e-16) * max(sp.linalg.norm(a), sp.linalg.norm(b))


def test_hyperviscosity_order6(ctx, O):
    import rbffd_amd

    X = jittered_grid(40, 40, 0.3, seed=3)
    print("hyperviscosity_operator(6, X, X, 9, 57, 6)", flush=True)
    Dxk, Dyk = rbffd_amd.hyperviscosity_operator(6, X, X, 9, 57, 6, ctx=ctx)
    ref = O.generate_operator(X, X, 9, 57, 6, ops=("Dxk", "Dyk"), k=6, quad=True)
    lu = _lu_reference(O, X, 9, 57, 6, ("Dxk", "Dyk"), 6)
    # cond(A) of r^9 / degree-6 stencils reaches 1e8 at the domain edge: the bar
    # is 1e-10 or 4x the error of a backward-stable FP64 LU solve
    for A, name in ((Dxk, "Dxk"), (Dyk, "Dyk")):
        e_lu = _rowerr(_csc(lu[name]), _csc(ref[name]))
        e = _same(A, ref[name], tol=max(1e-10, 4 * e_lu))
        print(f"  {name}: {e:.2e} (fp64 LU {e_lu:.2e})", flush=True)


def _boundary_domain():
    rng = np.random.default_rng(2)
    inner = jittered_grid(30, 30, 0.25, seed=4)[:, :] * 0.9 + 0.05
    t = np.linspace(0, 1, 31)
    left = np.stack([np.zeros_like(t), t], 1)
    right = np.stack([np.ones_like(t), t], 1)
    X = np.concatenate([inner, left, right, left - [1.0 / 30, 0], right + [1.0 / 30, 0]])
    ni = len(inner)
    idx_in = np.arange(ni)
    bc = [np.arange(ni, ni + 31), np.arange(ni + 31, ni + 62)]
    gh = [np.arange(ni + 62, ni + 93), np.arange(ni + 93, ni + 124)]
    del rng
    return X, idx_in, bc, gh


def test_generate_operator_method2(ctx, O):
    import rbffd_amd

    X, idx_in, bc, gh = _boundary_domain()
    Y = X.copy()
    print("generate_operator with boundary index sets", flush=True)
    mats = rbffd_amd.generate_operator(X, Y, 5, 21, 3, idx_in, bc, gh, None, None, None, ctx=ctx)
    ref = O.generate_operator(X, Y, 5, 21, 3, boundary=(idx_in, bc, gh), quad=True)
    for name, A in zip(NAMES, mats):
        _same(A, ref[name])


def test_error_behaviour(ctx):
    import rbffd_amd

    X = jittered_grid(12, 12, 0.25, seed=1)
    print("error paths", flush=True)
    with pytest.raises(ValueError):
        rbffd_amd.generate_operator(X[:10], X[:10], 5, 13, 2, ctx=ctx)       # n > N
    with pytest.raises(NotImplementedError):
        rbffd_amd.generate_operator(X, X, 5, 129, 2, ctx=ctx)                # n above the supported 128
    Xd = np.concatenate([X, X[50:51]])                                     # duplicated node
    with pytest.raises(rbffd_amd.SingularException):
        rbffd_amd.generate_operator(Xd, Xd, 5, 13, 2, ctx=ctx)
    Xn = X.copy()
    Xn[7] = np.nan
    with pytest.raises(rbffd_amd.NonFiniteError):
        rbffd_amd.generate_operator(Xn, Xn, 5, 13, 2, ctx=ctx)
    Xb, idx_in, bc, gh = _boundary_domain()
    with pytest.raises(rbffd_amd.UndefRefError):
        rbffd_amd.generate_operator(Xb, Xb, 5, 21, 3, idx_in[1:], bc, gh, None, None, None, ctx=ctx)
    # the context still works after errors
    mats = rbffd_amd.generate_operator(X, X, 5, 13, 2, ctx=ctx)
    assert mats[0].shape == (144, 144)

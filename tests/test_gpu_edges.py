"""Edge cases of the drop-in path: no Y nodes (the reference's loops run zero
times and sparse() of empty vectors is 0 x 0), stencils that take every node
(n = N), a single query, ragged CSR rows with empty rows in the product, and
CSC output with Julia's 1-based indices through the host entry point."""
import numpy as np
import pytest
import scipy.sparse as sp

from rbffd_amd.nodes import jittered_grid

pytestmark = pytest.mark.gpu


def test_no_y_nodes(ctx):
    import rbffd_amd

    X = jittered_grid(10, 10, 0.25, seed=3)
    print("generate_operator with an empty Y", flush=True)
    mats = rbffd_amd.generate_operator(X, np.zeros((0, 2)), 5, 13, 2, ctx=ctx)
    assert len(mats) == 6
    for A in mats:
        assert A.shape == (0, 0) and A.nnz == 0


def test_stencil_of_every_node(ctx, O):
    import rbffd_amd

    rng = np.random.default_rng(5)
    X = rng.uniform(0, 1, size=(13, 2))
    print("n = N = 13", flush=True)
    mats = rbffd_amd.generate_operator(X, X, 5, 13, 2, ctx=ctx)
    ref = O.generate_operator(X, X, 5, 13, 2, quad=True)
    for name, A in zip(("E", "Dx", "Dy", "Dxx", "Dyy", "Dxy"), mats):
        cp, rv, nz, shape = ref[name]
        B = sp.csc_matrix((nz, rv, cp), shape=shape)
        assert A.shape == (13, 13) and A.nnz == 169
        assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
        d = np.abs((A - B).toarray()).max(axis=1) / np.abs(B.toarray()).max(axis=1)
        assert d.max() <= 1e-10, (name, d.max())


def test_single_query(ctx, O):
    import rbffd_amd

    X = jittered_grid(20, 20, 0.25, seed=8)
    Y = np.array([[0.37, 0.61]])
    print("one Y node off the grid", flush=True)
    mats = rbffd_amd.generate_operator(X, Y, 5, 13, 2, ctx=ctx)
    ref = O.generate_operator(X, Y, 5, 13, 2, quad=True)
    for name, A in zip(("E", "Dx", "Dy", "Dxx", "Dyy", "Dxy"), mats):
        cp, rv, nz, shape = ref[name]
        B = sp.csc_matrix((nz, rv, cp), shape=shape)
        assert A.shape == B.shape
        assert np.array_equal(A.indices, B.indices)
        assert np.abs((A - B).toarray()).max() <= 1e-10 * np.abs(B.toarray()).max()


def test_ragged_rows_spmv(ctx, O):
    import rbffd_amd

    eng = rbffd_amd.Engine(ctx)
    rng = np.random.default_rng(9)
    M, N = 2000, 1500
    lens = rng.integers(0, 40, size=M)
    lens[::11] = 0
    rowptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = rng.integers(0, N, size=rowptr[-1]).astype(np.int32)
    val = rng.standard_normal(rowptr[-1])
    print("ragged CSR, empty rows, 1 and 8 fields", flush=True)
    for F in (1, 8):
        x = rng.standard_normal((N, F)) if F > 1 else rng.standard_normal(N)
        y = ctx.empty((M, F) if F > 1 else (M,), np.float64)
        eng.spmv(ctx.to_device(col), ctx.to_device(val), ctx.to_device(x), y, rowptr=ctx.to_device(rowptr))
        ref = sp.csr_matrix((val, col, rowptr), shape=(M, N)) @ x
        assert np.allclose(y.get(), ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
        assert not y.get()[lens == 0].any()


def test_host_entry_point_one_based(ctx, O):
    import ctypes

    from rbffd_amd import _lib

    X = jittered_grid(15, 15, 0.25, seed=2)
    N, n = len(X), 13
    L = _lib.load()
    ops = np.array([_lib.OP["Dx"], _lib.OP["Lap"]], np.int32)
    ks = np.zeros(2, np.int32)
    colptr = np.empty(N + 1, np.int64)
    rowval = np.empty(N * n, np.int64)
    nz = [np.empty(N * n) for _ in range(2)]
    nzp = (ctypes.c_void_p * 2)(*[a.ctypes.data for a in nz])
    ncols = ctypes.c_int64()
    print("host entry point, index base 1", flush=True)
    code = L.rbffd_generate_operator(ctx.handle, X.ctypes.data, N, X.ctypes.data, N, 5, n, 2, None, 2,
                                     ops.ctypes.data, ks.ctypes.data, 0, 1, colptr.ctypes.data,
                                     rowval.ctypes.data, ctypes.cast(nzp, ctypes.c_void_p), ctypes.byref(ncols),
                                     None)
    _lib.check(code)
    assert ncols.value == N
    ref = O.generate_operator(X, X, 5, n, 2, ops=("Dx", "Lap"), quad=True)
    cp, rv, _, _ = ref["Dx"]
    assert colptr[0] == 1 and np.array_equal(colptr - 1, cp) and np.array_equal(rowval - 1, rv)
    # the values through the host entry point, per-row relative error vs the quad oracle
    for o, name in enumerate(("Dx", "Lap")):
        _, _, zref, _ = ref[name]
        num = np.zeros(N)
        den = np.zeros(N)
        np.maximum.at(num, rv, np.abs(nz[o] - zref))
        np.maximum.at(den, rv, np.abs(zref))
        err = (num / den).max()
        print(f"  {name}: nzval vs quad {err:.2e}", flush=True)
        assert err <= 1e-10


def test_host_entry_point_large_rows_widened_on_the_host(ctx):
    # 600,625 rows x 31 entries: past two staging buffers of row indices, so
    # the int32 rows cross in chunks and the host widens them (index base 1);
    # the result equals the device-side CSC of the same rows
    import ctypes

    import rbffd_amd
    from rbffd_amd import _lib

    L = _lib.load()
    X = jittered_grid(775, 775, 0.25, seed=12)
    N, n = len(X), 31
    print(f"host entry point at {N} rows", flush=True)
    ops = np.array([_lib.OP["Lap"]], np.int32)
    ks = np.zeros(1, np.int32)
    colptr = np.empty(N + 1, np.int64)
    rowval = np.empty(N * n, np.int64)
    nz = np.empty(N * n)
    nzp = (ctypes.c_void_p * 1)(nz.ctypes.data)
    ncols = ctypes.c_int64()
    code = L.rbffd_generate_operator(ctx.handle, X.ctypes.data, N, X.ctypes.data, N, 7, n, 4, None, 1,
                                     ops.ctypes.data, ks.ctypes.data, 0, 1, colptr.ctypes.data,
                                     rowval.ctypes.data, ctypes.cast(nzp, ctypes.c_void_p), ctypes.byref(ncols),
                                     None)
    _lib.check(code)
    eng = rbffd_amd.Engine(ctx)
    Xd = ctx.to_device(X)
    col, vals, _ = eng.generate(Xd, Xd, 7, n, 4, ["Lap"])
    cp, rv, cv, used = eng.csr_to_csc(col, vals, N, index_base=1)
    assert ncols.value == used
    assert np.array_equal(colptr, cp.get())
    assert np.array_equal(rowval, rv.get())
    assert np.array_equal(nz, cv[0].get())

"""The drop-in boundary as a Julia caller sees it.

* Every matrix generate_operator returns owns its index arrays, as the
  reference's six separate sparse() calls give independent matrices
  (src/generate_operator.jl:176-181): a structural edit of one leaves the
  others intact.
* Operator application on the device: the call sequence of the Julia
  binding's DeviceOperator / mul! (julia/RBFFDAmd.jl), replayed through ctypes
  -- host CSC in, its CSR uploaded once, rbffd_spmv_device, the product back --
  against scipy's A * x, the reference's SparseMatrixCSC * Vector
  (test/poisson_test.jl:121, examples/adv_diff_test.jl:151-152).
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from rbffd_amd.nodes import jittered_grid

pytestmark = pytest.mark.gpu


def test_returned_matrices_are_independent(ctx):
    import rbffd_amd

    X = jittered_grid(30, 31, 0.25, seed=4)
    print("independent index arrays of the six operators", flush=True)
    mats = rbffd_amd.generate_operator(X, X, 5, 13, 2, ctx=ctx)
    for i in range(len(mats)):
        for j in range(i + 1, len(mats)):
            assert not np.shares_memory(mats[i].indices, mats[j].indices)
            assert not np.shares_memory(mats[i].indptr, mats[j].indptr)
    saved = [(A.indptr.copy(), A.indices.copy(), A.data.copy()) for A in mats]
    Dx = mats[1]
    Dx.data[::3] = 0.0
    Dx.eliminate_zeros()          # compacts Dx's index arrays in place
    Dx[0, len(X) - 1] = 1.0       # a new structural entry
    for A, (ip, ix, d) in zip(mats[:1] + mats[2:], saved[:1] + saved[2:]):
        assert np.array_equal(A.indptr, ip) and np.array_equal(A.indices, ix) and np.array_equal(A.data, d)


def _julia_mul(ctx, A, x):
    """What RBFFDAmd.mul!(y, DeviceOperator(A), x) does, call for call."""
    from rbffd_amd import _lib

    L = _lib.load()
    h = ctx.handle
    M, N = A.shape
    At = sp.csc_matrix(A.T)                       # sparse(transpose(A)): colptr/rowval = CSR of A
    At.sort_indices()
    F = 1 if x.ndim == 1 else x.shape[1]
    xr = np.ascontiguousarray(x.reshape(N, F))    # node-major, transpose(reshape(x, N, F)) in Julia
    yr = np.empty((M, F))
    bufs = []

    def upload(a):
        p = ctypes.c_void_p()
        _lib.check(L.rbffd_device_malloc(h, max(a.nbytes, 16), ctypes.byref(p)))
        _lib.check(L.rbffd_memcpy_h2d(h, p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))
        bufs.append(p)
        return p

    try:
        rp = upload(np.ascontiguousarray(At.indptr, dtype=np.int64))
        cl = upload(np.ascontiguousarray(At.indices, dtype=np.int32))
        vl = upload(np.ascontiguousarray(At.data, dtype=np.float64))
        xd = upload(xr)
        yd = ctypes.c_void_p()
        _lib.check(L.rbffd_device_malloc(h, yr.nbytes, ctypes.byref(yd)))
        bufs.append(yd)
        _lib.check(L.rbffd_spmv_device(h, M, rp, cl, vl, 0, F, xd, yd, 1.0, 0.0, 0, None))
        _lib.check(L.rbffd_memcpy_d2h(h, yr.ctypes.data_as(ctypes.c_void_p), yd, yr.nbytes))
    finally:
        for p in bufs:
            L.rbffd_device_free(h, p)
    return yr[:, 0] if x.ndim == 1 else yr


@pytest.mark.parametrize("F", [1, 4])
def test_device_operator_product(ctx, fitted_nodes, F):
    import rbffd_amd

    X, Y = fitted_nodes
    print(f"device product of generated operators, {F} field(s)", flush=True)
    E, Dx, Dy, Dxx, Dyy, Dxy = rbffd_amd.generate_operator(X, Y, 3, 20, 3, ctx=ctx)
    rng = np.random.default_rng(F)
    x = rng.standard_normal((X.shape[0], F)) if F > 1 else rng.standard_normal(X.shape[0])
    # an operator sum as the harness builds it (test/poisson_test.jl:75-113)
    for A in (Dx, Dxx + Dyy, E):
        ref = A @ x
        got = _julia_mul(ctx, A, x)
        scale = np.abs(ref).max()
        assert np.allclose(got.reshape(ref.shape), ref, rtol=1e-13, atol=1e-13 * scale)
        op = rbffd_amd.DeviceOperator(A, ctx=ctx)
        got2 = op @ x
        assert np.array_equal(got2.reshape(ref.shape), got.reshape(ref.shape))
        # the same operator through an application plan, rows in a random order
        order = np.random.default_rng(7).permutation(A.shape[0])
        opp = rbffd_amd.DeviceOperator(A, ctx=ctx, plan=True, order=order)
        assert np.array_equal((opp @ x).reshape(ref.shape), got.reshape(ref.shape))

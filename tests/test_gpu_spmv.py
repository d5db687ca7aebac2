"""K4 (SpMV / multi-field SpMM) and K3 (CSR -> CSC) on the GPU against the
oracle and scipy."""
import numpy as np
import pytest
import scipy.sparse as sp

from rbffd_amd.nodes import jittered_grid

pytestmark = pytest.mark.gpu


def _operator(ctx, X, n=31, p=7, polydeg=4):
    import rbffd_amd

    eng = rbffd_amd.Engine(ctx)
    Xd = ctx.to_device(X)
    col, vals, _ = eng.generate(Xd, Xd, p, n, polydeg, ["Lap", "Dx"])
    return eng, Xd, col, vals


@pytest.mark.parametrize("F", [1, 2, 4, 8])
@pytest.mark.parametrize("vtype", ["f64", "f32"])
def test_spmv_matches_oracle(ctx, O, F, vtype):
    X = jittered_grid(80, 70, 0.25, seed=6)
    N, n = len(X), 31
    print(f"spmv F={F} {vtype}", flush=True)
    eng, Xd, col, vals = _operator(ctx, X)
    rng = np.random.default_rng(F)
    x = rng.standard_normal((N, F)) if F > 1 else rng.standard_normal(N)
    xd = ctx.to_device(x)
    yd = ctx.empty(x.shape, np.float64)
    v = vals[0].get()
    if vtype == "f32":
        v = v.astype(np.float32)
    vd = ctx.to_device(v)
    eng.spmv(col, vd, xd, yd, uniform_n=n)
    y = yd.get()
    rowptr = np.arange(N + 1, dtype=np.int64) * n
    ref = O.spmv_csr(rowptr, col.get().ravel(), v.astype(np.float64).ravel(), x)
    scale = np.abs(v).max() * np.abs(x).max() * n
    assert np.abs(y - ref).max() <= 1e-13 * scale
    # explicit rowptr path, alpha/beta
    rpd = ctx.to_device(rowptr)
    y2 = ctx.to_device(np.ones_like(x))
    eng.spmv(col, vd, xd, y2, rowptr=rpd, alpha=2.0, beta=-1.0)
    assert np.abs(y2.get() - (2 * ref - 1)).max() <= 2e-13 * scale
    # a locality order changes which rows run together, not the results
    order = eng.locality_order(Xd)
    oh = order.get()
    assert np.array_equal(np.sort(oh), np.arange(N))
    y3 = ctx.empty(x.shape, np.float64)
    eng.spmv(col, vd, xd, y3, uniform_n=n, order=order)
    assert np.array_equal(y3.get(), y)


def test_csr_to_csc_matches_sparse(ctx, O):
    X = jittered_grid(60, 50, 0.25, seed=7)
    N, n = len(X), 31
    print("csr -> csc", flush=True)
    eng, Xd, col, vals = _operator(ctx, X)
    cp, rv, cv, used = eng.csr_to_csc(col, vals, N, index_base=0)
    c = col.get()
    rows = np.repeat(np.arange(N), n)
    for o in range(2):
        rcp, rrv, rnz, shape = O.sparse_csc(rows, c.ravel(), vals[o].get().ravel())
        assert used == shape[1]
        assert np.array_equal(cp.get()[: used + 1], rcp)
        assert np.array_equal(rv.get(), rrv)
        assert np.array_equal(cv[o].get(), rnz)
    cp1, rv1, _, _ = eng.csr_to_csc(col, vals[:1], N, index_base=1)
    assert np.array_equal(cp1.get(), cp.get() + 1)
    assert np.array_equal(rv1.get(), rv.get() + 1)

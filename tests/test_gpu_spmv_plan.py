"""The application plan (rbffd_spmv_plan_*): products and SSP-RK3 steps through
a plan equal the plain kernels (rbffd_spmv_device, rbffd_ssprk3_device, which
test_gpu_spmv and test_gpu_timestep check against the oracle) bit for bit, in
each kind of row block: halo staged in LDS, halo listed but larger than the
LDS share at 8 fields, and more than 2048 distinct columns (read through the
plan's copy of the columns)."""
import numpy as np
import pytest
import scipy.sparse as sp

from rbffd_amd.nodes import jittered_grid

pytestmark = pytest.mark.gpu


def _lap(ctx, X, n=31, p=7, polydeg=4):
    import rbffd_amd

    eng = rbffd_amd.Engine(ctx)
    Xd = ctx.to_device(X)
    col, vals, _ = eng.generate(Xd, Xd, p, n, polydeg, ["Lap"])
    return eng, Xd, col, vals[0]


def _fields(rng, N, F):
    return rng.standard_normal((N, F)) if F > 1 else rng.standard_normal(N)


@pytest.mark.parametrize("F", [1, 2, 4, 8])
@pytest.mark.parametrize("vtype", ["f64", "f32"])
@pytest.mark.parametrize("ordered", [True, False])
def test_plan_equals_spmv(ctx, F, vtype, ordered):
    X = jittered_grid(80, 70, 0.25, seed=6)
    N, n = len(X), 31
    print(f"plan F={F} {vtype} ordered={ordered}", flush=True)
    eng, Xd, col, v = _lap(ctx, X)
    vd = v if vtype == "f64" else ctx.to_device(v.get().astype(np.float32))
    order = eng.locality_order(Xd) if ordered else None
    plan = eng.spmv_plan(col, vd, N, N, uniform_n=n, order=order)
    assert plan.info["blocks"] == -(-N // 256) and plan.info["nnz"] == N * n
    assert plan.info["overflow_blocks"] == 0 and 0 < plan.info["max_halo"] <= 1024
    x = ctx.to_device(_fields(np.random.default_rng(F), N, F))
    y0, y1 = ctx.empty(x.shape, np.float64), ctx.empty(x.shape, np.float64)
    eng.spmv(col, vd, x, y0, uniform_n=n)
    plan.apply(x, y1)
    assert np.array_equal(y0.get(), y1.get())
    ones = np.ones(x.shape)
    y0, y1 = ctx.to_device(ones), ctx.to_device(ones)
    eng.spmv(col, vd, x, y0, uniform_n=n, alpha=2.0, beta=-1.0)
    plan.apply(x, y1, alpha=2.0, beta=-1.0)
    assert np.array_equal(y0.get(), y1.get())
    plan.close()


@pytest.mark.parametrize("F", [1, 8])
def test_plan_listed_halo_past_lds(ctx, F):
    # rows in node order on a wide grid: a block of 256 rows spans most of a
    # grid row, so its halo is over 1024 nodes (the LDS share at 8 fields)
    X = jittered_grid(300, 24, 0.25, seed=8, lx=299 / 23)
    N, n = len(X), 31
    print(f"plan wide grid F={F}", flush=True)
    eng, Xd, col, v = _lap(ctx, X)
    plan = eng.spmv_plan(col, v, N, N, uniform_n=n)
    assert plan.info["overflow_blocks"] == 0 and 1024 < plan.info["max_halo"] <= 2048
    x = ctx.to_device(_fields(np.random.default_rng(3), N, F))
    y0, y1 = ctx.empty(x.shape, np.float64), ctx.empty(x.shape, np.float64)
    eng.spmv(col, v, x, y0, uniform_n=n)
    plan.apply(x, y1)
    assert np.array_equal(y0.get(), y1.get())


@pytest.mark.parametrize("F", [1, 4, 8])
def test_plan_overflow_blocks_and_ragged_rows(ctx, F):
    import rbffd_amd

    # rows 0..299 draw random columns from all N (over 2048 distinct per block);
    # the rest are banded; row lengths 0..40, some rows empty
    rng = np.random.default_rng(11)
    M, N = 1500, 40000
    lens = rng.integers(0, 41, size=M)
    lens[[5, 700, 1499]] = 0
    rows, cols = [], []
    for i, k in enumerate(lens):
        if i < 300:
            c = rng.choice(N, size=k, replace=False)
        else:
            c = (i + rng.choice(60, size=k, replace=False)) % N
        rows.append(np.full(k, i))
        cols.append(c)
    A = sp.csr_matrix((rng.standard_normal(lens.sum()), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(M, N))
    A.sort_indices()
    print(f"plan overflow F={F} nnz={A.nnz}", flush=True)
    eng = rbffd_amd.Engine(ctx)
    rp = ctx.to_device(A.indptr.astype(np.int64))
    cl = ctx.to_device(A.indices.astype(np.int32))
    vl = ctx.to_device(A.data.astype(np.float64))
    order = ctx.to_device(rng.permutation(M).astype(np.int32))
    for od in (None, order):
        plan = eng.spmv_plan(cl, vl, M, N, rowptr=rp, order=od)
        assert plan.info["overflow_blocks"] >= 1 and plan.info["nnz"] == A.nnz
        x = ctx.to_device(_fields(np.random.default_rng(F), N, F))
        shape = (M, F) if F > 1 else (M,)
        y0, y1 = ctx.to_device(np.full(shape, 3.0)), ctx.to_device(np.full(shape, 3.0))
        eng.spmv(cl, vl, x, y0, rowptr=rp, alpha=0.5, beta=2.0)
        plan.apply(x, y1, alpha=0.5, beta=2.0)
        got = y1.get()
        assert np.array_equal(y0.get(), got)
        xh = x.get()
        ref = 0.5 * (A @ xh) + 2.0 * 3.0
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("F", [1, 8])
@pytest.mark.parametrize("vtype", ["f64", "f32"])
def test_plan_ssprk3_equals_plain(ctx, F, vtype):
    X = jittered_grid(64, 60, 0.25, seed=9)
    N, n = len(X), 31
    print(f"plan ssprk3 F={F} {vtype}", flush=True)
    eng, Xd, col, v = _lap(ctx, X)
    vd = v if vtype == "f64" else ctx.to_device(v.get().astype(np.float32))
    h = 1.0 / 63
    fixed_h = np.zeros(N, np.uint8)
    fixed_h[::17] = 1
    fixed = ctx.to_device(fixed_h)
    order = eng.locality_order(Xd)
    u0 = _fields(np.random.default_rng(5), N, F)
    ua, ub = ctx.to_device(u0), ctx.to_device(u0)
    dt = 0.05 * h * h
    eng.ssprk3(col, vd, ua, dt, 5, uniform_n=n, fixed=fixed, order=order)
    plan = eng.spmv_plan(col, vd, N, N, uniform_n=n, order=order)
    plan.ssprk3(ub, dt, 5, fixed=fixed)
    a, b = ua.get(), ub.get()
    assert np.isfinite(a).all() and not np.array_equal(a, u0)
    assert np.array_equal(a, b)


def test_plan_errors_and_empty(ctx):
    import rbffd_amd

    X = jittered_grid(20, 20, 0.25, seed=2)
    eng, Xd, col, v = _lap(ctx, X, n=13, p=5, polydeg=2)
    N = len(X)
    plan = eng.spmv_plan(col, v, N, N, uniform_n=13)
    x = ctx.to_device(np.ones((N, 3)))
    y = ctx.empty((N, 3), np.float64)
    with pytest.raises(NotImplementedError):
        plan.apply(x, y)
    with pytest.raises(ValueError):
        plan.apply(ctx.to_device(np.ones(N + 1)), ctx.empty((N,), np.float64))
    empty = eng.spmv_plan(col, v, 0, N, uniform_n=13)
    assert empty.info["blocks"] == 0 and empty.info["nnz"] == 0
    empty.close()
    plan.close()
    plan.close()

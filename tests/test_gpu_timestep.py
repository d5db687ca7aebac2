"""Operator algebra and SSP-RK3 time stepping on the GPU against a numpy
restatement of the same Shu-Osher stages over the oracle's CSR product:
FP64 values, FP32 values (the reference uses the same rounded values), F = 1
and 8 fields, Dirichlet rows held fixed."""
import numpy as np
import pytest

from rbffd_amd.nodes import jittered_grid

pytestmark = pytest.mark.gpu


def _cpu_ssprk3(O, rowptr, col, val, u, dt, steps, fixed):
    u = u.copy()

    def stage(w, a, b, c):
        r = a * u + b * w + c * O.spmv_csr(rowptr, col, val, w)
        r[fixed] = u[fixed]
        return r

    for _ in range(steps):
        u1 = stage(u, 0.0, 1.0, dt)
        u2 = stage(u1, 0.75, 0.25, 0.25 * dt)
        u = stage(u2, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0 * dt)
    return u


@pytest.mark.parametrize("F", [1, 8])
@pytest.mark.parametrize("vtype", ["f64", "f32"])
def test_ssprk3_matches_numpy(ctx, O, F, vtype):
    import rbffd_amd

    eng = rbffd_amd.Engine(ctx)
    nx = 48
    X = jittered_grid(nx, nx, 0.25, seed=7)
    N, n, h = len(X), 31, 1.0 / (nx - 1)
    Xd = ctx.to_device(X)
    print(f"ssprk3 F={F} {vtype}", flush=True)
    col, vals, _ = eng.generate(Xd, Xd, 7, n, 4, ["Dx", "Dy", "Lap", "Lapk"], k=3)
    coef = [-1.0, -0.5, 1e-3, -1e-8 * h**6]
    L64 = eng.combine(vals, coef)
    L32 = ctx.empty((N, n), np.float32)
    eng.combine(vals, coef, out32=L32)
    hv = [v.get() for v in vals]
    ref = coef[0] * hv[0]
    for c, v in zip(coef[1:], hv[1:]):
        ref = c * v + ref
    assert np.allclose(L64.get(), ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())
    assert np.array_equal(L32.get(), L64.get().astype(np.float32))
    Lv = L64 if vtype == "f64" else L32
    Lh = Lv.get().astype(np.float64).ravel()
    colh = col.get().ravel()
    rowptr = np.arange(N + 1, dtype=np.int64) * n
    rng = np.random.default_rng(F)
    c = rng.uniform(0.3, 0.7, size=(F, 2))
    u0 = np.stack([np.exp(-((X[:, 0] - cx) ** 2 + (X[:, 1] - cy) ** 2) / 0.1**2) for cx, cy in c], 1)
    u0 = u0[:, 0] if F == 1 else np.ascontiguousarray(u0)
    edge = (X[:, 0] <= 0) | (X[:, 0] >= 1) | (X[:, 1] <= 0) | (X[:, 1] >= 1)
    fixed = ctx.to_device(edge.astype(np.uint8))
    dt, steps = 0.05 * h, 20
    u = ctx.to_device(u0)
    eng.ssprk3(col, Lv, u, dt, steps, uniform_n=n, fixed=fixed)
    got = u.get()
    # the same steps with the rows in locality order give the same bits
    u_o = ctx.to_device(u0)
    eng.ssprk3(col, Lv, u_o, dt, steps, uniform_n=n, fixed=fixed, order=eng.locality_order(Xd))
    assert np.array_equal(u_o.get(), got)
    want = _cpu_ssprk3(O, rowptr, colh, Lh, u0, dt, steps, edge)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"  max relative difference {err:.2e}", flush=True)
    assert err <= 1e-12
    assert np.array_equal(got[edge], u0[edge])


def _closure_setup(ctx, nodes=1500, seed=3):
    import scipy.sparse as sp

    import rbffd_amd
    from rbffd_amd.nodes import ghosted_rectangle
    from rbffd_amd.operators import GhostClosure

    X, inner, bcs, ghosts = ghosted_rectangle(nodes, 1.5, 0.25, seed)
    E, Dx, Dy, Dxx, Dyy, _ = rbffd_amd.generate_operator(X, X, 7, 31, 4, inner, bcs, ghosts, inner, bcs, ghosts,
                                                         ctx=ctx)
    h = 1.5 / (int(round(np.sqrt(nodes / 1.5))) - 1)
    L = (-1.0 * Dx - 0.5 * Dy + 1e-3 * (Dxx + Dyy)).tocsr()
    L.sort_indices()
    # left/right markers close with Dx, bottom/top with Dy (the normal derivative)
    closures = [GhostClosure(ctx, D, b, g) for D, b, g in zip((Dx, Dx, Dy, Dy), bcs, ghosts)]
    return X, L, closures, h


def _cpu_closure_ssprk3(L, closures, u, dt, steps):
    """The same stages over scipy products, the closure after every stage
    with the device's inverse blocks."""
    u = u.copy()
    blocks = []
    for g in closures:
        inner, _, _, ghost = g.host
        ginv = g.ginv.get().T          # column-major on the device
        blocks.append((inner, ginv, ghost))

    def close(v):
        for inner, ginv, ghost in blocks:
            v[ghost] = -(ginv @ (inner @ v))
        return v

    def stage(w, a, b, c):
        return close(a * u + b * w + c * (L @ w))

    u = close(u)
    for _ in range(steps):
        u1 = stage(u, 0.0, 1.0, dt)
        u2 = stage(u1, 0.75, 0.25, 0.25 * dt)
        u = stage(u2, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0 * dt)
    return u


@pytest.mark.parametrize("F", [1, 4])
def test_ssprk3_ghost_closure(ctx, F):
    """Ghost-node closure (examples/adv_diff_test.jl:107-146,168-182) on a
    ghosted rectangle, against a scipy restatement; the Neumann rows of the
    last-closed marker vanish; two runs give the same bits."""
    import rbffd_amd

    eng = rbffd_amd.Engine(ctx)
    X, L, closures, h = _closure_setup(ctx)
    N = len(X)
    print(f"ghost closure: {N} nodes, markers {[g.nbc for g in closures]}, F={F}", flush=True)
    rng = np.random.default_rng(F)
    c = rng.uniform(0.3, 0.7, size=(F, 2)) * [1.5, 1.0]
    u0 = np.stack([np.exp(-((X[:, 0] - cx) ** 2 + (X[:, 1] - cy) ** 2) / 0.1**2) for cx, cy in c], 1)
    dt, steps = 0.05 * h, 20
    rowptr = ctx.to_device(L.indptr.astype(np.int64))
    col = ctx.to_device(L.indices.astype(np.int32))
    vals = ctx.to_device(L.data)
    runs = []
    for _ in range(2):
        u = ctx.to_device(np.ascontiguousarray(u0[:, 0] if F == 1 else u0))
        eng.ssprk3_closure(col, vals, u, dt, steps, closures, rowptr=rowptr)
        runs.append(u.get().reshape(N, F))
    assert np.array_equal(runs[0], runs[1])
    got = runs[0]
    for f in range(F):
        want = _cpu_closure_ssprk3(L, closures, u0[:, f], dt, steps)
        err = np.abs(got[:, f] - want).max() / np.abs(want).max()
        print(f"  field {f}: max relative difference {err:.2e}", flush=True)
        assert err <= 1e-12
    inner, Wg, bc, ghost = closures[-1].host
    resid = inner @ got + Wg @ got[ghost]
    assert np.abs(resid).max() <= 1e-10 * np.abs(inner).max() * np.abs(got).max()


def test_dense_inverse(ctx):
    import rbffd_amd

    eng = rbffd_amd.Engine(ctx)
    rng = np.random.default_rng(11)
    A = rng.standard_normal((50, 50)) + 8 * np.eye(50)
    inv = eng.dense_inverse(ctx.to_device(np.ascontiguousarray(A.T))).get().T
    ref = np.linalg.inv(A)
    assert np.abs(inv - ref).max() <= 1e-13 * np.abs(ref).max()
    S = A.copy()
    S[:, 7] = 0.0
    with pytest.raises(rbffd_amd.SingularException):
        eng.dense_inverse(ctx.to_device(np.ascontiguousarray(S.T)))

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

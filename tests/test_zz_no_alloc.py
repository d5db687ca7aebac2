"""rbffd_weights_device_noalloc, the entry point that allocates nothing, against
rbffd_weights_device through the C ABI: the same kernels, so the same bits.
Two calls in a row, the second with a smaller input, for a specialised shape
with Y = X (no workspace), a shape for the generic kernel with Y distinct from X
(row grouping takes its buffers from the static array) and a system above 128
equations (the workspace-resident kernel); a call that needs more than the
array holds is refused with an error."""
import ctypes

import numpy as np
import pytest

from rbffd_amd.nodes import jittered_grid

pytestmark = pytest.mark.gpu

def _fn(L):
    return L.rbffd_weights_device_noalloc


def _call(ctx, fn, X, Y, st, y2x, p, polydeg, ops, k=0):
    from rbffd_amd import _lib

    Xd = ctx.to_device(X)
    Yd = Xd if Y is X else ctx.to_device(Y)
    std = ctx.to_device(st.astype(np.int32))
    y2xd = ctx.to_device(y2x.astype(np.int32)) if y2x is not None else None
    M, n = len(Y), st.shape[1]
    kinds = np.array([_lib.OP[o] for o in ops], np.int32)
    ks = np.array([k if o in ("Dxk", "Dyk", "Lapk") else 0 for o in ops], np.int32)
    col = ctx.empty((M, n), np.int32)
    vals = [ctx.empty((M, n), np.float64) for _ in ops]
    status = ctx.empty((M,), np.int32)
    vp = (ctypes.c_void_p * len(ops))(*[v.ptr for v in vals])
    code = fn(ctx.handle, Xd.ptr, len(X), Yd.ptr, M, std.ptr, n, y2xd.ptr if y2xd is not None else None, int(p),
              int(polydeg), 1, len(ops), kinds.ctypes.data_as(ctypes.c_void_p), ks.ctypes.data_as(ctypes.c_void_p),
              0, col.ptr, ctypes.cast(vp, ctypes.c_void_p), status.ptr)
    _lib.check(code)
    ctx.synchronize()
    return col.get(), [v.get() for v in vals], status.get()


CASES = [
    # (p, polydeg, n, collocated, (nx, ny) first call, (nx, ny) second call)
    (7, 4, 31, True, (60, 61), (33, 30)),
    (5, 3, 25, False, (50, 51), (28, 29)),
    (5, 8, 100, True, (24, 25), (20, 21)),
]


@pytest.mark.parametrize("p,polydeg,n,colloc,size1,size2", CASES)
def test_noalloc_matches_weights_device(ctx, O, p, polydeg, n, colloc, size1, size2):
    from rbffd_amd import _lib

    L = _lib.load()
    na = _fn(L)
    for (nx, ny) in (size1, size2):
        X = jittered_grid(nx, ny, 0.3, seed=nx)
        if colloc:
            Y, y2x = X, None
        else:
            Y = jittered_grid(2 * nx - 1, 2 * ny - 1, 0.2, seed=ny)
            y2x = O.knn(X, Y, 1)[0][:, 0]
        st, _ = O.knn(X, X, n)
        print(f"no-alloc p={p} deg={polydeg} n={n} X={len(X)} Y={len(Y)}", flush=True)
        ops = ["Dx", "Dy", "Lap"]
        c1, v1, s1 = _call(ctx, L.rbffd_weights_device, X, Y, st, y2x, p, polydeg, ops)
        c2, v2, s2 = _call(ctx, na, X, Y, st, y2x, p, polydeg, ops)
        assert not s1.any()
        assert np.array_equal(c1, c2) and np.array_equal(s1, s2)
        for a, b in zip(v1, v2):
            assert np.array_equal(a, b)


def test_noalloc_refuses_more_than_the_array(ctx, O):
    from rbffd_amd import _lib

    L = _lib.load()
    na = _fn(L)
    # grouping needs about 20 bytes per X node and 16 per Y row: 4M Y rows do not fit in 64 MiB
    X = jittered_grid(40, 40, 0.25, seed=3)
    st, _ = O.knn(X, X, 13)
    M = 4_200_000
    rng = np.random.default_rng(0)
    Y = rng.uniform(0.0, 1.0, size=(M, 2))
    y2x = rng.integers(0, len(X), size=M).astype(np.int32)
    print("no-alloc refusal", flush=True)
    with pytest.raises(ValueError, match="static array"):
        _call(ctx, na, X, Y, st, y2x, 5, 2, ["Dx"])

"""Every compiled lane-group layout of the FP64 local-solve kernel, forced one
at a time (RBFFD_LU_GROUP), against the oracle.

The autotuner picks a layout per (shape, right-hand-side count) by timing, and
inputs below 4096 stencils use a fixed default, so without forcing, a layout
that only wins at full size would be checked by the benchmark alone.  Each
layout here runs through rbffd_generate_device (kNN + solves) on a small
jittered grid; the stats report the layout that ran.

Bar (as in test_gpu_weights): the CSR pattern bit for bit; weights within
1e-10 of the high-precision oracle per row, or within 4x the error of a
backward-stable FP64 LU solve of the same systems where conditioning puts
1e-10 out of reach.
"""
import os

import numpy as np
import pytest

from rbffd_amd.nodes import jittered_grid

pytestmark = pytest.mark.gpu

TOL = 1e-10


def layouts(n, polydeg):
    """Lanes per stencil G compiled for a shape: a lane holds ceil(n / G) row
    slots of n + P + 3 doubles (at three right-hand sides), a layout is
    compiled where those take at most 200 registers and at least half of a
    group's lanes hold a node, and 64 lanes when no layout qualifies.  Here:
    n = 13 -> 16; 21, 31 -> 16, 32; 43, 57 -> 64."""
    ns = n + (polydeg + 1) * (polydeg + 2) // 2
    used = [g for g in (16, 32, 64) if -(-n // g) * (ns + 3) * 2 <= 200 and 2 * n > g]
    return used or [64]


CASES = [(p, d, n, g) for (p, d, n) in [(5, 2, 13), (3, 3, 21), (7, 4, 31), (7, 5, 43), (9, 6, 57)]
         for g in layouts(n, d)]


def rowrel(got, ref):
    return (np.abs(got - ref).max(axis=-1) / np.abs(ref).max(axis=-1)).max()


@pytest.mark.parametrize("ops", [["Dx", "Dy", "Lap"], ["Lap"], ["E", "Dx", "Dy", "Dxx", "Dyy", "Dxy"]])
@pytest.mark.parametrize("p,polydeg,n,g", CASES)
def test_forced_layout(ctx, O, monkeypatch, p, polydeg, n, g, ops):
    import rbffd_amd

    monkeypatch.setenv("RBFFD_LU_GROUP", str(g))
    X = jittered_grid(30, 28, 0.3, seed=200 + n)
    print(f"layout G={g} p={p} deg={polydeg} n={n} ops={ops}", flush=True)
    eng = rbffd_amd.Engine(ctx)
    Xd = ctx.to_device(X)
    st_d = ctx.empty((len(X), n), np.int32)
    col, vals, stats = eng.generate(Xd, Xd, p, n, polydeg, ops, stencils_out=st_d)
    assert stats["lanes_per_stencil"] == g
    assert stats["singular"] == 0 and stats["nonfinite"] == 0
    st = st_d.get()
    st_ref, _ = O.knn(X, X, n)
    assert np.array_equal(st, st_ref)
    opl = [(O.OP_NAMES[o], 0) for o in ops]
    Wq, _, sq = O.weights(X, X, st, None, p, polydeg, opl, quad=True)
    Wl, _, _ = O.weights(X, X, st, None, p, polydeg, opl, solve="lu")
    assert not sq.any()
    order = np.argsort(st, axis=1, kind="stable")
    assert np.array_equal(col.get(), np.take_along_axis(st, order, 1))
    for o, name in enumerate(ops):
        rq = np.take_along_axis(Wq[:, o, :], order, 1)
        rl = np.take_along_axis(Wl[:, o, :], order, 1)
        e, e_lu = rowrel(vals[o].get(), rq), rowrel(rl, rq)
        print(f"  {name}: vs quad {e:.2e} (fp64 LU {e_lu:.2e})", flush=True)
        assert e <= max(TOL, 4 * e_lu), name

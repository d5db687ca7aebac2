"""The headline workload at its full size on the GPU against the oracle:
1,048,576 jittered nodes (1024 x 1024, +-0.25h, seed 2), r^7, degree 4,
n = 31, operators Dx, Dy, Lap (the reference's generate_operator(X, p, n,
polydeg) at src/generate_operator.jl:29 with Y = X).

Bar, per row relative weight error max_j |w - w_ref| / max_j |w_ref|:
  * the CSR pattern equals the oracle's kNN stencils, sorted, bit for bit on
    every row;
  * against the high-precision (__float128) oracle on every 256th row:
    <= 1e-10;
  * against the FP64 oracle that restates the reference's inv(A) * RHS on
    every row: <= 1e-10, except where that oracle's own error explains the
    gap -- on the 64 rows where the GPU and that oracle differ most, the
    high-precision oracle shows the GPU within 1e-10 of exact and the gap
    within 1.5 x (that oracle's error + ours) on each row (inv(A) * RHS
    carries cond(A) eps of its own).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rowrel(got, ref):
    return np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)


def test_gen1m_against_oracle(ctx, O):
    import rbffd_amd
    from rbffd_amd.nodes import config_nodes

    X, prm = config_nodes("gen1m")
    p, polydeg, n = prm["p"], prm["polydeg"], prm["n"]
    N = len(X)
    ops = ["Dx", "Dy", "Lap"]
    print(f"full-size workload: {N} nodes, n = {n}", flush=True)
    eng = rbffd_amd.Engine(ctx)
    Xd = ctx.to_device(X)
    col, vals, stats = eng.generate(Xd, Xd, p, n, polydeg, ops)
    print(f"  GPU solve {stats['ms_solve']:.2f} ms, kNN {stats['ms_knn']:.2f} ms", flush=True)
    colh = col.get()
    wh = [v.get() for v in vals]
    st, _ = O.knn(X, X, n)
    assert np.array_equal(colh, np.sort(st, axis=1)), "CSR pattern differs from the oracle's stencils"
    order = np.argsort(st, axis=1, kind="stable")
    opl = [(O.OP_DX, 0), (O.OP_DY, 0), (O.OP_LAP, 0)]
    rows = np.arange(0, N, 256, dtype=np.int64)
    Wq, _, sq = O.weights(X, X, st, None, p, polydeg, opl, quad=True, rows=rows)
    assert not sq.any()
    W, _, sd = O.weights(X, X, st, None, p, polydeg, opl)
    assert not sd.any()
    for o, name in enumerate(ops):
        rq = np.take_along_axis(Wq[:, o, :], order[rows], 1)
        rd = np.take_along_axis(W[:, o, :], order, 1)
        eq = rowrel(wh[o][rows], rq).max()
        ed = rowrel(wh[o], rd)
        worst = np.sort(np.argsort(ed)[-64:]).astype(np.int64)
        Ww, _, _ = O.weights(X, X, st, None, p, polydeg, [opl[o]], quad=True, rows=worst)
        rw = np.take_along_axis(Ww[:, 0, :], order[worst], 1)
        e_gpu_w = rowrel(wh[o][worst], rw)
        e_inv_w = rowrel(rd[worst], rw)
        print(f"  {name}: vs quad {eq:.2e} (every 256th row), {e_gpu_w.max():.2e} (64 worst rows); vs inv "
              f"oracle max {ed.max():.2e} p99 {np.percentile(ed, 99):.2e}; inv oracle's own error on the worst "
              f"rows up to {e_inv_w.max():.2e}", flush=True)
        assert eq <= 1e-10 and e_gpu_w.max() <= 1e-10, name
        assert (ed[worst] <= np.maximum(1e-10, 1.5 * (e_inv_w + e_gpu_w))).all(), name

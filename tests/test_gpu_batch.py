"""The uneven batch (many node sets of different stencil size and degree in
one call, recorded as a hipGraph and replayed): every replay equals the eager
run bit for bit, the eager run equals one single-set call per set, one set
with method-2 groups matches the oracle, and a failing set is reported
without disturbing the others."""
import numpy as np
import pytest

from rbffd_amd.nodes import ghosted_rectangle

pytestmark = pytest.mark.gpu
OPS = ["Dx", "Dy", "Lap"]


def _sets(ctx, rbffd_amd, specs):
    sets, host = [], []
    for i, (nodes, deg, grouped) in enumerate(specs):
        X, inner, bcs, ghosts = ghosted_rectangle(nodes, 1.0 + 0.5 * i, 0.25, 100 + i)
        n = 2 * (deg + 1) * (deg + 2) // 2 + 1
        S = dict(X=ctx.to_device(X), p=2 * (deg // 2) + 3, n=n, polydeg=deg)
        xg = qg = None
        if grouped:
            xg, qg = rbffd_amd.boundary_groups(len(X), inner, bcs, ghosts)
            S["xgroup"], S["qgroup"] = ctx.to_device(xg), ctx.to_device(qg)
        sets.append(S)
        host.append((X, inner, bcs, ghosts, xg, qg))
    return sets, host


SPECS = [(3000, 2, True), (1500, 3, False), (4200, 4, True), (2500, 5, True), (900, 2, False), (3600, 4, False)]


def test_batch_graph_replays_equal_eager_and_single_calls(ctx):
    import rbffd_amd

    eng = rbffd_amd.Engine(ctx)
    sets, host = _sets(ctx, rbffd_amd, SPECS)
    print("batch of 6 sets, eager run + capture", flush=True)
    out, st, stats = eng.generate_batch(sets, OPS, use_graph=True)
    assert not st.any() and stats["singular"] == 0 and stats["nonfinite"] == 0
    eager = [(c.get(), [v.get() for v in vs]) for c, vs, _ in out]
    for S, (c, vs, _) in zip(sets, out):
        S["colidx"], S["vals"] = c, vs
    for rep in range(3):
        for c, vs, _ in out:
            c.zero_()
            for v in vs:
                v.zero_()
        print(f"graph replay {rep}", flush=True)
        out2, st2, _ = eng.generate_batch(sets, OPS, use_graph=True)
        assert not st2.any()
        for (c, vs, _), (ce, vse) in zip(out2, eager):
            assert np.array_equal(c.get(), ce)
            for v, ve in zip(vs, vse):
                assert np.array_equal(v.get(), ve)
        ctx.synchronize()
    print("single-set calls", flush=True)
    for S, (X, inner, bcs, ghosts, xg, _), (ce, vse) in zip(sets, host, eager):
        bnd = (inner, bcs, ghosts) if xg is not None else None
        col, vals, _ = eng.generate(S["X"], S["X"], S["p"], S["n"], S["polydeg"], OPS, boundary=bnd)
        assert np.array_equal(col.get(), ce)
        for v, ve in zip(vals, vse):
            assert np.array_equal(v.get(), ve)


def test_batch_grouped_set_matches_oracle(ctx, O):
    import rbffd_amd

    eng = rbffd_amd.Engine(ctx)
    sets, host = _sets(ctx, rbffd_amd, [(2600, 3, True), (1800, 4, True)])
    print("batch of 2 grouped sets vs oracle", flush=True)
    out, st, _ = eng.generate_batch(sets, OPS, use_graph=False)
    assert not st.any()
    for S, (X, inner, bcs, ghosts, xg, qg), (c, vs, _) in zip(sets, host, out):
        stn, _ = O.knn(X, X, S["n"], xgroup=xg, qgroup=qg)
        assert np.array_equal(c.get(), np.sort(stn, axis=1))
        Wq, _, sq = O.weights(X, X, stn, None, S["p"], S["polydeg"],
                              [(O.OP_DX, 0), (O.OP_DY, 0), (O.OP_LAP, 0)], quad=True)
        assert not sq.any()
        order = np.argsort(stn, axis=1, kind="stable")
        for o in range(3):
            ref = np.take_along_axis(Wq[:, o, :], order, 1)
            got = vs[o].get()
            err = (np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)).max()
            print(f"  n={S['n']} op {OPS[o]}: {err:.2e}", flush=True)
            assert err <= 1e-10


def test_batch_reports_the_failing_set(ctx):
    import rbffd_amd

    eng = rbffd_amd.Engine(ctx)
    sets, host = _sets(ctx, rbffd_amd, [(1200, 2, False), (1500, 3, False), (1300, 2, False)])
    X = host[1][0].copy()
    X[17] = np.nan
    sets[1]["X"] = ctx.to_device(X)
    print("batch with a non-finite set", flush=True)
    with pytest.raises(rbffd_amd.NonFiniteError):
        eng.generate_batch(sets, OPS, use_graph=False)
    assert list(eng.last_set_status) == [0, 2, 0]
    with pytest.raises(ValueError):
        bad = dict(sets[0])
        bad["n"] = 5000
        eng.generate_batch([bad], OPS)
    # the context still works
    out, st, _ = eng.generate_batch([sets[0], sets[2]], OPS, use_graph=True)
    assert not st.any()


def test_batch_graph_after_an_eager_call_with_bigger_sets(ctx):
    # graph A, an eager batch B whose sets are larger (its workspace grows and
    # the buffers A's graph recorded are reallocated), then graph A again:
    # the second A must not replay the stale graph; its outputs equal A's first
    import rbffd_amd

    eng = rbffd_amd.Engine(ctx)
    small = [(1200, 2, False), (900, 3, True), (1500, 4, False)]
    big = [(9000, 4, False), (7000, 5, True), (12000, 2, False)]
    sets_a, _ = _sets(ctx, rbffd_amd, small)
    sets_b, _ = _sets(ctx, rbffd_amd, big)
    print("graph A", flush=True)
    out, st, _ = eng.generate_batch(sets_a, OPS, use_graph=True)
    assert not st.any()
    first = [(c.get(), [v.get() for v in vs]) for c, vs, _ in out]
    for S, (c, vs, _) in zip(sets_a, out):
        S["colidx"], S["vals"] = c, vs
    print("eager B (bigger sets)", flush=True)
    _, stb, _ = eng.generate_batch(sets_b, OPS, use_graph=False)
    assert not stb.any()
    for c, vs, _ in out:
        c.zero_()
        for v in vs:
            v.zero_()
    print("graph A again", flush=True)
    out2, st2, _ = eng.generate_batch(sets_a, OPS, use_graph=True)
    assert not st2.any()
    for (c, vs, _), (cf, vsf) in zip(out2, first):
        assert np.array_equal(c.get(), cf)
        for v, vf in zip(vs, vsf):
            assert np.array_equal(v.get(), vf)

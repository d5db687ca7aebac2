"""Benchmark of the RBF-FD operator engine on one MI355X.

A step is one pass of the hot path over the whole node set with the inputs
resident in HBM: exact kNN stencils (n = 31) of 1,048,576 jittered nodes, the
nearest-node map, and the local solves of r^7 / degree-4 systems for the
operators d/dx, d/dy and the Laplacian, written as FP64 CSR rows.

The JSON line reports:
  value                stencils/s of the whole step, FP64 Gauss-Jordan solver (weights
                       within 1e-10 of the high-precision oracle)
  roofline             the local-solve kernel: algorithmic flops per launch
                       ((2/3) Nsys^3 + 2 Nsys^2 r per stencil, Nsys = 46, r = 3)
                       over its HIP-event duration, against the FP64 peak;
                       traffic = HBM bytes per launch from the rocprofv3 PMC
                       summary committed with the sources
  cpu_baseline         the oracle (a C restatement of the reference's
                       per-stencil algorithm, inv(A) * RHS) on the host cores
                       over all rows of the workload; the same pass gives the
                       GPU weights' max and p99 row error per operator
  end_to_end           the host-level call (host arrays in, CSC out): H2D,
                       kNN, solves, CSR->CSC, D2H
  spmv                 y = L x for the generated Laplacian (1 and 8 fields),
                       algorithmic bytes over HIP-event time, beside a
                       device-to-device copy of 1 GiB and a 1 GiB read; the
                       same products through an application plan (F1_plan,
                       F8_plan: Morton-ordered row blocks, x staged in LDS)
  mixed                the FP32 Gauss-Jordan + FP64 refinement solver, same
                       workload, and its ratio to the FP64 solver
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "radialbasisfinitedifferences.jl_amd")]

METRIC = "stencils/s (1M nodes, deg-4 PHS); SpMV GB/s; mixed vs FP64 at ≤1e-10"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
OPS = ["Dx", "Dy", "Lap"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_share():
    """Host threads this job may use: OMP_NUM_THREADS when the host sets it
    (the job's CPU share), else the CPUs in the affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def read_traffic(name):
    """HBM bytes per launch recorded by a rocprofv3 --pmc summary."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def row_errors(got, ref):
    """max and p99 of the per-row relative error max_j |got - ref| / max_j |ref|."""
    e = np.abs(got - ref).max(axis=1) / np.maximum(np.abs(ref).max(axis=1), 1e-300)
    return float(e.max()), float(np.percentile(e, 99))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-e2e", action="store_true", help="skip the host-level end-to-end leg")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the other workloads (hyperviscosity 4M, uneven batch, time stepping)")
    ap.add_argument("--extras", default="all",
                    help="comma-separated subset of the extra workloads: hyperviscosity_4m, uneven_batch_64, "
                         "advdiff_ssprk3_4m, poisson_solve, y_distinct_4m (default all)")
    ap.add_argument("--spmv-only", action="store_true",
                    help="profiling workload: generate the Laplacian of the 1M-node set, then for 1 and 8 "
                         "fields 10 plain products and 10 through an application plan; prints their mean times")
    args = ap.parse_args()
    if args.spmv_only:
        spmv_only()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import rbffd_amd
    from rbffd_amd.nodes import config_nodes

    ctx = rbffd_amd.Context(0)
    info = ctx.info()
    log("device", info)
    eng = rbffd_amd.Engine(ctx)
    X, prm = config_nodes("gen1m")
    p, polydeg, n = prm["p"], prm["polydeg"], prm["n"]
    N = len(X)
    Xd = ctx.to_device(X)
    P = (polydeg + 1) * (polydeg + 2) // 2
    nsys = n + P
    flops_per_stencil = (2.0 / 3.0) * nsys**3 + 2.0 * nsys**2 * len(OPS)

    def run(precision, steps, warmup, tag, col, vals, st_d):
        stats = None
        for _ in range(warmup):
            eng.generate(Xd, Xd, p, n, polydeg, OPS, precision=precision, colidx=col, vals=vals,
                         stencils_out=st_d)
        ctx.synchronize()
        ms_solve, ms_knn, ms_fb = [], [], []
        t0 = time.perf_counter()
        ctx.timer_start()
        for _ in range(steps):
            _, _, stats = eng.generate(Xd, Xd, p, n, polydeg, OPS, precision=precision, colidx=col,
                                       vals=vals, stencils_out=st_d)
            ms_solve.append(stats["ms_solve"])
            ms_knn.append(stats["ms_knn"])
            ms_fb.append(stats["ms_fallback"])
        ms = ctx.timer_stop()
        ctx.synchronize()
        wall = (time.perf_counter() - t0) * 1e3
        log(f"[{tag}] {steps} steps: {ms / steps:.3f} ms/step (wall {wall / steps:.3f}), "
            f"knn {np.mean(ms_knn):.3f} ms, solve {np.mean(ms_solve):.3f} ms, "
            f"fallback rows {stats['refined_fallback']}, lanes/stencil {stats['lanes_per_stencil']}")
        return {"ms_step": ms / steps, "ms_solve": float(np.mean(ms_solve)), "ms_knn": float(np.mean(ms_knn)),
                "ms_fallback": float(np.mean(ms_fb)), "fallback_rows": int(stats["refined_fallback"]),
                "lanes": int(stats["lanes_per_stencil"])}

    col = ctx.empty((N, n), np.int32)
    vals = [ctx.empty((N, n), np.float64) for _ in OPS]
    st_d = ctx.empty((N, n), np.int32)
    log("timing FP64 solver")
    r64 = run("fp64", args.steps, args.warmup, "fp64", col, vals, st_d)
    value = N / (r64["ms_step"] * 1e-3)
    achieved = N * flops_per_stencil / (r64["ms_solve"] * 1e-3) / 1e12
    colh = col.get()
    wh = [v.get() for v in vals]

    mixed = None
    if not args.no_mixed:
        log("timing mixed solver")
        mcol = ctx.empty((N, n), np.int32)
        mvals = [ctx.empty((N, n), np.float64) for _ in OPS]
        rm = run("mixed", args.steps, args.warmup, "mixed", mcol, mvals, st_d)
        mh = [v.get() for v in mvals]
        dev = max(row_errors(a, b)[0] for a, b in zip(mh, wh))
        mixed = {"experimental": True,
                 "note": "FP32 Gauss-Jordan + FP64 refinement; scalar FP32 issues at the FP64 rate here "
                         "(packed FP32 is not used), so it is slower and FP64 is the default",
                 "stencils_per_s": N / (rm["ms_step"] * 1e-3), "ms_per_step": rm["ms_step"],
                 "ms_solve": rm["ms_solve"], "ms_fallback": rm["ms_fallback"], "fallback_rows": rm["fallback_rows"],
                 "max_row_dev_vs_fp64": dev, "mixed_vs_fp64": (N / (rm["ms_step"] * 1e-3)) / value}
        del mcol, mvals

    e2e = None
    if not args.no_e2e:
        log("timing the host-level call (H2D + kNN + solves + CSR->CSC + D2H)")
        rbffd_amd.operators(X[:4096], X[:4096], p, n, polydeg, OPS, ctx=ctx)  # warm the small paths
        calls, walls = [], []
        for _ in range(3):  # the first call at this size also grows the device and pinned buffers
            t0 = time.perf_counter()
            mats, est = rbffd_amd.operators(X, X, p, n, polydeg, OPS, ctx=ctx, return_stats=True)
            walls.append(time.perf_counter() - t0)
            calls.append(est["host_call_s"])
            del mats
        tc, dt = float(np.median(calls[1:])), float(np.median(walls[1:]))
        e2e = {"stencils_per_s": N / tc, "seconds": tc, "first_call_seconds": calls[0],
               "python_wrapper_seconds": dt,
               "note": "the C host entry point: host arrays in, SparseMatrixCSC-layout arrays out, PCIe "
                       "included; seconds is the median of the calls after the first; python_wrapper_seconds "
                       "adds the scipy matrices of the Python mirror"}
        log(f"  end to end {tc * 1e3:.1f} ms in the C call ({N / tc:.3e} stencils/s; first call "
            f"{calls[0] * 1e3:.1f} ms), {dt * 1e3:.1f} ms with the Python wrapper")

    # operator application: the Laplacian on 1 and 8 fields, and a plain copy
    log("timing SpMV / SpMM")
    nnz = N * n
    spmv = {}
    morton = eng.locality_order(Xd)
    for F in (1, 8):
        x = ctx.to_device(np.random.default_rng(F).standard_normal((N, F)) if F > 1 else
                          np.random.default_rng(1).standard_normal(N))
        y = ctx.empty((N, F) if F > 1 else (N,), np.float64)
        for tag, od in (("", None),):
            for _ in range(3):
                eng.spmv(col, vals[2], x, y, uniform_n=n, order=od)
            ctx.synchronize()
            reps = 50
            ctx.timer_start()
            for _ in range(reps):
                eng.spmv(col, vals[2], x, y, uniform_n=n, order=od)
            t = ctx.timer_stop() / reps
            bytes_alg = nnz * 12 + 8 * F * (N + N) + (4 * N if od is not None else 0)
            spmv[f"F{F}{tag}"] = {"ms": t, "GB_s": bytes_alg / (t * 1e-3) / 1e9, "bytes": bytes_alg,
                                  "frac_hbm_peak": bytes_alg / (t * 1e-3) / 1e9 / PEAK_HBM_GBS}
            log(f"  spmv F={F}{tag}: {t * 1e3:.1f} us, {spmv[f'F{F}{tag}']['GB_s']:.0f} GB/s")
        # the same product through an application plan (Morton-ordered blocks of
        # 256 rows, x staged in LDS); GB/s counted on the plain CSR's bytes, so
        # the two compare at equal work
        plan = eng.spmv_plan(col, vals[2], N, N, uniform_n=n, order=morton)
        y2 = ctx.empty(y.shape, np.float64)
        for _ in range(3):
            plan.apply(x, y2)
        ctx.synchronize()
        ctx.timer_start()
        for _ in range(reps):
            plan.apply(x, y2)
        t = ctx.timer_stop() / reps
        bytes_alg = nnz * 12 + 8 * F * (N + N)
        spmv[f"F{F}_plan"] = {"ms": t, "GB_s": bytes_alg / (t * 1e-3) / 1e9, "bytes": bytes_alg,
                              "frac_hbm_peak": bytes_alg / (t * 1e-3) / 1e9 / PEAK_HBM_GBS,
                              "speedup": spmv[f"F{F}"]["ms"] / t,
                              "bitwise_equal_to_plain": bool(np.array_equal(y.get(), y2.get())), "plan": plan.info}
        log(f"  spmv F={F} through a plan: {t * 1e3:.1f} us ({spmv[f'F{F}_plan']['speedup']:.2f}x)")
        plan.close()
    a = ctx.empty((1 << 27,), np.float64)
    b = ctx.empty((1 << 27,), np.float64)
    b.copy_from_device(a)
    reps = 10
    ctx.timer_start()
    for _ in range(reps):
        b.copy_from_device(a)
    t = ctx.timer_stop() / reps
    copy_gbs = 2 * a.nbytes / (t * 1e-3) / 1e9
    spmv["copy_1GiB_GB_s"] = copy_gbs
    log(f"  device copy of 1 GiB: {t:.3f} ms, {copy_gbs:.0f} GB/s (read + write)")
    # read-only stream: the ceiling for SpMV, whose traffic is almost all reads
    out = ctx.empty((1,), np.float64)
    eng.read_sum(a, out)
    ctx.timer_start()
    for _ in range(reps):
        eng.read_sum(a, out)
    t = ctx.timer_stop() / reps
    read_gbs = a.nbytes / (t * 1e-3) / 1e9
    spmv["read_1GiB_GB_s"] = read_gbs
    for F in (1, 8):
        spmv[f"F{F}"]["frac_of_read_ceiling"] = spmv[f"F{F}"]["GB_s"] / read_gbs
    log(f"  device read of 1 GiB: {t:.3f} ms, {read_gbs:.0f} GB/s")
    del a, b, out

    cpu = None
    accuracy = None
    if not args.no_cpu:
        log("CPU baseline (oracle restatement, all rows)")
        from oracle import oracle as O

        threads = cpu_share()
        t0 = time.perf_counter()
        st, _ = O.knn(X, X, n, nthreads=threads)
        W, _, status = O.weights(X, X, st, None, p, polydeg, [(O.OP_DX, 0), (O.OP_DY, 0), (O.OP_LAP, 0)],
                                 nthreads=threads)
        dt = time.perf_counter() - t0
        same_pattern = bool(np.array_equal(colh, np.sort(st, axis=1)))
        order = np.argsort(st, axis=1, kind="stable")
        accuracy = {"vs_fp64_oracle_inv": {}}
        for o, name in enumerate(OPS):
            ref = np.take_along_axis(W[:, o, :], order, 1)
            mx, p99 = row_errors(wh[o], ref)
            accuracy["vs_fp64_oracle_inv"][name] = {"max": mx, "p99": p99}
        del W
        # the high-precision oracle on a sample of rows
        rows = np.arange(0, N, 256, dtype=np.int64)
        Wq, _, _ = O.weights(X, X, st, None, p, polydeg, [(O.OP_DX, 0), (O.OP_DY, 0), (O.OP_LAP, 0)],
                             quad=True, rows=rows, nthreads=threads)
        accuracy["vs_quad_oracle"] = {"rows": len(rows)}
        for o, name in enumerate(OPS):
            ref = np.take_along_axis(Wq[:, o, :], order[rows], 1)
            mx, p99 = row_errors(wh[o][rows], ref)
            accuracy["vs_quad_oracle"][name] = {"max": mx, "p99": p99}
        vq = max(accuracy["vs_quad_oracle"][name]["max"] for name in OPS)
        accuracy["bar"] = ("relative weight error per row <= 1e-10 against the high-precision oracle; against "
                           "the reference-semantics FP64 oracle (inv(A) * RHS) the difference is that oracle's "
                           "own FP64 error")
        accuracy["max_vs_quad"] = vq
        accuracy["pass"] = bool(vq <= 1e-10 and same_pattern)
        cpu = {"value": N / dt, "unit": "stencils/s", "cores": threads, "kind": "port",
               "sample": f"all {N} rows of the workload, kNN + weights (inv(A) * RHS), {dt:.1f} s wall",
               "host_cpus_visible": os.cpu_count(),
               "cores_note": "threads = this job's CPU share on the host (OMP_NUM_THREADS, else the affinity "
                             "mask), all of it used",
               "gpu_pattern_equal": same_pattern}
        log(f"  CPU {cpu['value']:.0f} stencils/s on {threads} threads; pattern equal {same_pattern}")
        log(f"  accuracy {json.dumps(accuracy)}")
        # CPU SpMV of the same Laplacian, one thread and all threads
        rowptr = np.arange(0, nnz + 1, n, dtype=np.int64)
        xv = np.random.default_rng(1).standard_normal(N)
        cpu_spmv = {}
        for th in (1, threads):
            O.spmv_csr(rowptr, colh, wh[2], xv, nthreads=th)
            t0 = time.perf_counter()
            reps = 5
            for _ in range(reps):
                O.spmv_csr(rowptr, colh, wh[2], xv, nthreads=th)
            ts = (time.perf_counter() - t0) / reps
            cpu_spmv[f"threads_{th}"] = {"ms": ts * 1e3, "GB_s": (nnz * 12 + 8 * (N + 1) + 16 * N) / ts / 1e9}
        cpu["spmv"] = cpu_spmv
        log(f"  CPU SpMV {json.dumps(cpu_spmv)}")

    extras = None
    if not args.no_extras:
        del col, vals, st_d, Xd
        extras = {}
        wanted = None if args.extras == "all" else set(args.extras.split(","))
        for name, fn in (("hyperviscosity_4m", extra_hyper4m), ("uneven_batch_64", extra_batch),
                         ("advdiff_ssprk3_4m", extra_advdiff), ("poisson_solve", extra_poisson_solve),
                         ("y_distinct_4m", extra_ydistinct)):
            if wanted is not None and name not in wanted:
                continue
            try:
                extras[name] = fn(rbffd_amd, ctx, eng, args.no_cpu)
            except Exception as e:  # a failed extra is reported, the headline stands
                extras[name] = {"error": f"{type(e).__name__}: {e}"}
            log(f"  {name}: {json.dumps(extras[name])}")

    G = r64["lanes"]
    out = {
        "metric": METRIC, "value": value, "unit": "stencils/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r64["ms_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "gen1m: 1,048,576 jittered nodes (1024x1024, +-0.25h, seed 2), PHS r^7, "
                               "degree-4 monomials, n=31, operators Dx, Dy, Lap as FP64 CSR",
                   "nodes": N, "stencil": n, "polydeg": polydeg, "phs_power": p, "ops": OPS},
        "roofline": {"bound": "valu-fp64", "kernel": f"solve_lu_kernel<31,15,3,{G}>",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": read_traffic("solve_lu_gen1m_pmc.txt")},
        "kernel_only_stencils_per_s": N / (r64["ms_solve"] * 1e-3),
        "ms_knn": r64["ms_knn"], "ms_solve": r64["ms_solve"], "lanes_per_stencil": G,
        "end_to_end": e2e, "accuracy": accuracy,
        "spmv": spmv, "mixed": mixed, "cpu_baseline": cpu, "extras": extras, "device": info,
    }
    print(json.dumps(out))


def spmv_only():
    import rbffd_amd
    from rbffd_amd.nodes import config_nodes

    ctx = rbffd_amd.Context(0)
    eng = rbffd_amd.Engine(ctx)
    X, prm = config_nodes("gen1m")
    N, n = len(X), prm["n"]
    Xd = ctx.to_device(X)
    col, vals, _ = eng.generate(Xd, Xd, prm["p"], n, prm["polydeg"], ["Lap"])
    morton = eng.locality_order(Xd)
    out = {}
    for F in (1, 8):
        x = ctx.to_device(np.random.default_rng(F).standard_normal((N, F)))
        y = ctx.empty((N, F), np.float64)
        ctx.synchronize()
        ctx.timer_start()
        for _ in range(10):
            eng.spmv(col, vals[0], x, y, uniform_n=n)
        out[f"F{F}_ms"] = ctx.timer_stop() / 10
        # the same products through an application plan
        plan = eng.spmv_plan(col, vals[0], N, N, uniform_n=n, order=morton)
        ctx.synchronize()
        ctx.timer_start()
        for _ in range(10):
            plan.apply(x, y)
        out[f"F{F}_plan_ms"] = ctx.timer_stop() / 10
        plan.close()
    print(json.dumps(out))


def extra_hyper4m(rbffd_amd, ctx, eng, no_cpu):
    """4,194,304 nodes (2048^2, +-0.3h), r^9, degree 6, n = 57: Lap^3, and the
    reference-semantics pair d^6/dx^6, d^6/dy^6, FP64 LU."""
    from rbffd_amd.nodes import config_nodes

    log("extra: hyperviscosity, 4M nodes, n = 57")
    X, prm = config_nodes("hyper4m")
    N, n = len(X), prm["n"]
    Xd = ctx.to_device(X)
    st = ctx.empty((N, n), np.int32)
    res = {"nodes": N, "stencil": n, "polydeg": prm["polydeg"], "phs_power": prm["p"]}
    outs = {}
    for tag, ops, k in (("lap3", ["Lapk"], 3), ("dxk_dyk6", ["Dxk", "Dyk"], 6)):
        col = ctx.empty((N, n), np.int32)
        vals = [ctx.empty((N, n), np.float64) for _ in ops]
        eng.generate(Xd, Xd, prm["p"], n, prm["polydeg"], ops, k=k, colidx=col, vals=vals, stencils_out=st)
        ctx.synchronize()
        ctx.timer_start()
        _, _, stats = eng.generate(Xd, Xd, prm["p"], n, prm["polydeg"], ops, k=k, colidx=col, vals=vals,
                                   stencils_out=st)
        ms = ctx.timer_stop()
        nsys = n + 28
        flops = (2.0 / 3.0) * nsys**3 + 2.0 * nsys**2 * len(ops)
        res[tag] = {"ms": ms, "stencils_per_s": N / (ms * 1e-3), "ms_solve": stats["ms_solve"],
                    "ms_knn": stats["ms_knn"], "lanes_per_stencil": stats["lanes_per_stencil"],
                    "solve_tflops": N * flops / (stats["ms_solve"] * 1e-3) / 1e12}
        outs[tag] = (ops, k, col, [v.get() for v in vals])
        del col, vals
    if not no_cpu:
        from oracle import oracle as O

        rows = np.arange(0, N, 16384, dtype=np.int64)
        stn, _ = O.knn(X, X[rows], n, nthreads=cpu_share())
        full = np.zeros((N, n), np.int32)
        full[rows] = stn
        order = np.argsort(stn, axis=1, kind="stable")
        for tag, (ops, k, _, hv) in outs.items():
            opl = [(O.OP_LAPK if o == "Lapk" else (O.OP_XK if o == "Dxk" else O.OP_YK), k) for o in ops]
            Wq, _, _ = O.weights(X, X, full, None, prm["p"], prm["polydeg"], opl, quad=True, rows=rows, nthreads=cpu_share())
            Wl, _, _ = O.weights(X, X, full, None, prm["p"], prm["polydeg"], opl, solve="lu", rows=rows,
                                 nthreads=cpu_share())
            acc = {}
            for o, name in enumerate(ops):
                ref = np.take_along_axis(Wq[:, o, :], order, 1)
                lu = np.take_along_axis(Wl[:, o, :], order, 1)
                acc[name] = {"gpu_vs_quad_max": row_errors(hv[o][rows], ref)[0],
                             "fp64_lu_oracle_vs_quad_max": row_errors(lu, ref)[0]}
            res[tag]["accuracy_rows"] = len(rows)
            res[tag]["accuracy"] = acc
    res["mixed_precision"] = "not compiled for n + P = 85 (the FP32 path covers n + P <= 64)"
    return res


def extra_batch(rbffd_amd, ctx, eng, no_cpu):
    """64 node sets, 10k-500k nodes each, degree 2-5 (n = 13, 21, 31, 43), ghost
    nodes with method-2 groups, one captured graph per batch."""
    from rbffd_amd.nodes import batch_sets

    log("extra: uneven batch of 64 node sets")
    specs = batch_sets()
    sets, total = [], 0
    for S in specs:
        X = S["X"]
        xg, qg = rbffd_amd.boundary_groups(len(X), S["X_idx_in"], S["X_idx_bc"], S["X_idx_bc_g"])
        sets.append(dict(X=ctx.to_device(X), p=S["p"], n=S["n"], polydeg=S["polydeg"], xgroup=ctx.to_device(xg),
                         qgroup=ctx.to_device(qg)))
        total += len(X)
    ops = ["Dx", "Dy", "Lap"]
    out, st, stats = eng.generate_batch(sets, ops, use_graph=True)  # eager run, tuning, capture
    for S, (c, vs, s) in zip(sets, out):
        S["colidx"], S["vals"], S["status"] = c, vs, s
    eng.generate_batch(sets, ops, use_graph=True)
    times = []
    for _ in range(3):
        _, st, stats = eng.generate_batch(sets, ops, use_graph=True)
        times.append(stats["ms_solve"])
    t_graph = float(np.median(times))
    _, _, s2 = eng.generate_batch(sets, ops, use_graph=False)
    t_eager = s2["ms_solve"]
    ctx.synchronize()
    t0 = time.perf_counter()
    for S in sets:  # one generation call per set, as a caller of the single-set API would
        eng.generate(S["X"], S["X"], S["p"], S["n"], S["polydeg"], ops, colidx=S["colidx"], vals=S["vals"])
    ctx.synchronize()
    t_loop = (time.perf_counter() - t0) * 1e3
    return {"sets": len(sets), "nodes": total, "failed_sets": int((st != 0).sum()),
            "graph_ms": t_graph, "stencils_per_s_graph": total / (t_graph * 1e-3),
            "eager_one_stream_ms": t_eager, "per_set_calls_ms": t_loop,
            "stencil_sizes": sorted({S["n"] for S in specs})}


def extra_advdiff(rbffd_amd, ctx, eng, no_cpu):
    """4,194,304 nodes, n = 31 Dx, Dy, Lap and Lap^3 fused into one operator,
    8 Gaussian fields, 1000 SSP-RK3 steps with FP64 and with FP32 values."""
    from rbffd_amd.nodes import config_nodes

    log("extra: advection-diffusion, 4M nodes, 8 fields, 1000 SSP-RK3 steps")
    X, prm = config_nodes("advdiff4m")
    N, n = len(X), prm["n"]
    h = 1.0 / 2047
    Xd = ctx.to_device(X)
    col, vals, _ = eng.generate(Xd, Xd, prm["p"], n, prm["polydeg"], ["Dx", "Dy", "Lap", "Lapk"], k=3)
    gamma, kappa = 1e-8, 1e-3
    coef = [-1.0, -0.5, kappa, -gamma * h**6]
    L64 = eng.combine(vals, coef)
    L32 = ctx.empty((N, n), np.float32)
    eng.combine(vals, coef, out32=L32)
    del vals
    rng = np.random.default_rng(5)
    cen = rng.uniform(0.2, 0.8, size=(8, 2))
    u0 = np.ascontiguousarray(np.stack([np.exp(-((X[:, 0] - cx) ** 2 + (X[:, 1] - cy) ** 2) / 0.05**2)
                                        for cx, cy in cen], 1))
    edge = (X[:, 0] <= 0) | (X[:, 0] >= 1) | (X[:, 1] <= 0) | (X[:, 1] >= 1)
    fixed = ctx.to_device(edge.astype(np.uint8))
    dt, steps = 0.05 * h, 1000
    work = ctx.empty((2 * N * 8,), np.float64)
    res = {"nodes": N, "fields": 8, "steps": steps, "dt": dt, "kappa": kappa, "gamma_h6": gamma * h**6,
           "velocity": [1.0, 0.5], "bc": "boundary nodes held at their initial values",
           "dt_note": "dt = 0.05h: kappa * dt * rho(Lap) ~ 2.05 (rho ~ 20/h^2) keeps SSP-RK3 stable; 0.2h gives ~8"}
    finals = {}
    nnz = N * n
    order = None  # the grid's row-major node order already keeps stencils local
    for tag, Lv, vb in (("fp64", L64, 8), ("fp32_values", L32, 4)):
        u = ctx.to_device(u0)
        eng.ssprk3(col, Lv, u, dt, 2, uniform_n=n, fixed=fixed, work=work, order=order)  # warm-up
        u.copy_from_host(u0)
        ctx.synchronize()
        ctx.timer_start()
        eng.ssprk3(col, Lv, u, dt, steps, uniform_n=n, fixed=fixed, work=work, order=order)
        ms = ctx.timer_stop()
        stage_bytes = nnz * (4 + vb) + 3 * 8 * 8 * N
        res[tag] = {"ms": ms, "ms_per_step": ms / steps, "stage_GB_s": 3 * steps * stage_bytes / (ms * 1e-3) / 1e9}
        finals[tag] = u.get()
    # the same runs through an application plan (rows in blocks of 256 of the
    # Morton order, each block's x rows staged in LDS); bitwise equal
    morton = eng.locality_order(Xd)
    for tag, Lv, vb in (("fp64_plan", L64, 8), ("fp32_values_plan", L32, 4)):
        plan = eng.spmv_plan(col, Lv, N, N, uniform_n=n, order=morton)
        u = ctx.to_device(u0)
        plan.ssprk3(u, dt, 2, fixed=fixed, work=work)  # warm-up
        u.copy_from_host(u0)
        ctx.synchronize()
        ctx.timer_start()
        plan.ssprk3(u, dt, steps, fixed=fixed, work=work)
        ms = ctx.timer_stop()
        base = tag[:-len("_plan")]
        res[tag] = {"ms": ms, "ms_per_step": ms / steps, "speedup": res[base]["ms"] / ms,
                    "bitwise_equal_to_plain": bool(np.array_equal(u.get(), finals[base])), "plan": plan.info}
        plan.close()
    ref = finals["fp64"]
    res["fp32_values_final_rel_err"] = float(np.abs(finals["fp32_values"] - ref).max() / np.abs(ref).max())
    res["final_max_abs"] = float(np.abs(ref).max())
    return res


def extra_ydistinct(rbffd_amd, ctx, eng, no_cpu):
    """Y distinct from X: X the 1024^2 workload nodes (r^7, degree 4, n = 31),
    Y a 2047^2 jittered grid (4,190,209 rows, about 4 per X node), operators
    Dx, Dy, Lap: the library's own choice (timed on the first large call),
    rows grouped by X node (each task's rows solved against one factorisation
    of its system; an X node with more rows than a task holds takes several),
    and one factorisation per row."""
    from rbffd_amd.nodes import config_nodes, jittered_grid

    log("extra: Y distinct from X, 1M X nodes, 4.2M Y rows")
    X, prm = config_nodes("gen1m")
    Y = jittered_grid(2047, 2047, 0.25, seed=7)
    Xd, Yd = ctx.to_device(X), ctx.to_device(Y)
    ops = ["Dx", "Dy", "Lap"]
    M, n = len(Y), prm["n"]
    col = ctx.empty((M, n), np.int32)
    vals = [ctx.empty((M, n), np.float64) for _ in ops]
    res = {"x_nodes": len(X), "y_rows": M, "stencil": n}
    old = os.environ.get("RBFFD_GROUP_ROWS")
    try:
        outs = {}
        for tag, flag in (("default_choice", None), ("grouped_by_x_node", "1"),
                          ("one_factorisation_per_row", "0")):
            if flag is None:
                os.environ.pop("RBFFD_GROUP_ROWS", None)
            else:
                os.environ["RBFFD_GROUP_ROWS"] = flag
            eng.generate(Xd, Yd, prm["p"], n, prm["polydeg"], ops, colidx=col, vals=vals)
            ctx.synchronize()
            _, _, stats = eng.generate(Xd, Yd, prm["p"], n, prm["polydeg"], ops, colidx=col, vals=vals)
            res[tag] = {"ms_solve": stats["ms_solve"], "ms_knn": stats["ms_knn"],
                        "factorizations": stats["factorizations"],
                        "rows_per_s_solve": M / (stats["ms_solve"] * 1e-3)}
            outs[tag] = [v.get() for v in vals]
        a, b = outs["grouped_by_x_node"], outs["one_factorisation_per_row"]
        res["max_row_dev_between_the_two"] = max(row_errors(x, y)[0] for x, y in zip(a, b))
    finally:
        if old is None:
            os.environ.pop("RBFFD_GROUP_ROWS", None)
        else:
            os.environ["RBFFD_GROUP_ROWS"] = old
    return res


def extra_poisson_solve(rbffd_amd, ctx, eng, no_cpu):
    """`D \\ f` of the Poisson harness on the GPU (rbffd_lsq_solve_device).
    Square: the 51 x 51 Dirichlet system (r^5, degree 2, n = 13) by QR.
    Oversampled: X the same 51 x 51 nodes, Y a 101 x 101 jittered grid,
    Laplacian rows inside and E rows on the boundary, scaled as the reference
    scales its blocks; QR and CGLS (tol 1e-12).  Times are device time of the
    whole call (checks, densify or transpose, solve)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    from rbffd_amd.nodes import config_nodes, jittered_grid

    log("extra: Poisson harness solves (square 2601, oversampled 10201 x 2601)")
    X, prm = config_nodes("poisson2500")
    u = lambda P: np.sin(np.pi * P[:, 0]) * np.sin(np.pi * P[:, 1])
    edge = lambda P: (P[:, 0] <= 0) | (P[:, 0] >= 1) | (P[:, 1] <= 0) | (P[:, 1] >= 1)

    def timed(A, b, method):
        A = sp.csr_matrix(A)
        A.sort_indices()
        dev = [ctx.to_device(a) for a in (A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data, b)]
        x = ctx.zeros((A.shape[1],), np.float64)
        eng.lsq_solve(*dev, x, method=method, tol=1e-12, max_iters=100000)   # warm-up (handles, workspace)
        x.zero_()
        ctx.synchronize()
        ctx.timer_start()
        its, rel = eng.lsq_solve(*dev, x, method=method, tol=1e-12, max_iters=100000)
        ms = ctx.timer_stop()
        return x.get(), ms, its, rel

    res = {}
    E, Lap = rbffd_amd.operators(X, X, prm["p"], prm["n"], prm["polydeg"], ["E", "Lap"], ctx=ctx)
    ed = np.where(edge(X))[0]
    A = sp.lil_matrix(Lap.tocsr())
    A[ed, :] = E.tocsr()[ed, :]
    A = A.tocsr()
    b = -2 * np.pi**2 * u(X)
    b[ed] = u(X)[ed]
    x, ms, _, _ = timed(A, b, "qr")
    t0 = time.perf_counter()
    ref = sla.spsolve(A.tocsc(), b)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    res["square_2601_qr"] = {"ms": ms, "rel_l2_error": float(np.linalg.norm(x - u(X)) / np.linalg.norm(u(X))),
                             "vs_splu": float(np.linalg.norm(x - ref) / np.linalg.norm(ref)),
                             "cpu_splu_ms": cpu_ms}

    Y = jittered_grid(101, 101, 0.25, seed=11)
    E, Lap = rbffd_amd.operators(X, Y, prm["p"], prm["n"], prm["polydeg"], ["E", "Lap"], ctx=ctx)
    yb = edge(Y)
    yi = np.where(~yb)[0]
    yd = np.where(yb)[0]
    d2 = np.sort(((X[:, None, :] - X[None, :, :]) ** 2).sum(-1), axis=1)[:, 1]
    h = float(np.mean(np.sqrt(d2)))
    A = sp.vstack([Lap.tocsr()[yi] / np.sqrt(len(yi)), E.tocsr()[yd] / (h * np.sqrt(len(yd)))]).tocsr()
    b = np.concatenate([-2 * np.pi**2 * u(Y[yi]) / np.sqrt(len(yi)), u(Y[yd]) / (h * np.sqrt(len(yd)))])
    t0 = time.perf_counter()
    ref = np.linalg.lstsq(A.toarray(), b, rcond=None)[0]
    cpu_ms = (time.perf_counter() - t0) * 1e3
    for method in ("qr", "cgls"):
        x, ms, its, rel = timed(A, b, method)
        res[f"oversampled_10201x2601_{method}"] = {
            "ms": ms, "iterations": its, "relres": rel,
            "rel_l2_error": float(np.linalg.norm(x - u(X)) / np.linalg.norm(u(X))),
            "vs_lstsq": float(np.linalg.norm(x - ref) / np.linalg.norm(ref))}
    res["oversampled_cpu_dense_lstsq_ms"] = cpu_ms
    return res


if __name__ == "__main__":
    main()

"""Benchmark of the RBF-FD operator engine on one MI355X.

A step is one pass of the hot path over the whole node set with the inputs
resident in HBM: exact kNN stencils (n = 31) of 1,048,576 jittered nodes, the
nearest-node map, and the local solves of r^7 / degree-4 systems for the
operators d/dx, d/dy and the Laplacian, written as FP64 CSR rows.

The JSON line reports:
  value                stencils/s of the whole step, FP64 solver (the bar:
                       weights within 1e-10 of the oracle)
  roofline             the local-solve kernel: algorithmic flops per launch
                       ((2/3) Nsys^3 + 2 Nsys^2 r per stencil, Nsys = 46, r = 3)
                       over its HIP-event duration, against the FP64 peak
  cpu_baseline         the oracle (a C restatement of the reference's
                       per-stencil algorithm, inv(A) * RHS) on the host cores,
                       over a bounded sample of the same rows
  spmv / spmm          y = L x for the generated Laplacian (1 and 8 fields),
                       algorithmic bytes over HIP-event time
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(name):
    """HBM bytes per launch recorded by the rocprofv3 --pmc summaries."""
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true")
    args = ap.parse_args()
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
    ops = ["Dx", "Dy", "Lap"]
    Xd = ctx.to_device(X)
    col = ctx.empty((N, n), np.int32)
    vals = [ctx.empty((N, n), np.float64) for _ in ops]
    st_d = ctx.empty((N, n), np.int32)
    P = (polydeg + 1) * (polydeg + 2) // 2
    nsys = n + P
    flops_per_stencil = (2.0 / 3.0) * nsys**3 + 2.0 * nsys**2 * len(ops)

    def run(precision, steps, warmup, tag):
        for _ in range(warmup):
            eng.generate(Xd, Xd, p, n, polydeg, ops, precision=precision, colidx=col, vals=vals,
                         stencils_out=st_d)
        ctx.synchronize()
        ms_solve, ms_knn, ms_fb, fb = [], [], [], 0
        t0 = time.perf_counter()
        ctx.timer_start()
        for s in range(steps):
            _, _, stats = eng.generate(Xd, Xd, p, n, polydeg, ops, precision=precision, colidx=col,
                                       vals=vals, stencils_out=st_d)
            ms_solve.append(stats["ms_solve"])
            ms_knn.append(stats["ms_knn"])
            ms_fb.append(stats["ms_fallback"])
            fb = stats["refined_fallback"]
        ms = ctx.timer_stop()
        ctx.synchronize()
        wall = (time.perf_counter() - t0) * 1e3
        log(f"[{tag}] {steps} steps: {ms / steps:.3f} ms/step (wall {wall / steps:.3f}), "
            f"knn {np.mean(ms_knn):.3f} ms, solve {np.mean(ms_solve):.3f} ms, fallback rows {fb}")
        return ms / steps, float(np.mean(ms_solve)), float(np.mean(ms_knn)), float(np.mean(ms_fb)), fb

    log("timing FP64 solver")
    ms_step, ms_solve, ms_knn, _, _ = run("fp64", args.steps, args.warmup, "fp64")
    value = N / (ms_step * 1e-3)
    achieved = N * flops_per_stencil / (ms_solve * 1e-3) / 1e12

    colh = col.get()
    wh = [v.get() for v in vals]

    mixed = None
    if not args.no_mixed:
        log("timing mixed solver")
        mcol = ctx.empty((N, n), np.int32)
        mvals = [ctx.empty((N, n), np.float64) for _ in ops]
        col_fp64, vals_fp64 = col, vals
        col, vals = mcol, mvals
        msm, mss, _, msfb, fb = run("mixed", args.steps, args.warmup, "mixed")
        col, vals = col_fp64, vals_fp64
        mh = [v.get() for v in mvals]
        dev = max(float((np.abs(a - b).max(axis=1) / np.abs(b).max(axis=1)).max()) for a, b in zip(mh, wh))
        mixed = {"stencils_per_s": N / (msm * 1e-3), "ms_per_step": msm, "ms_solve": mss, "ms_fallback": msfb,
                 "fallback_rows": int(fb), "max_row_dev_vs_fp64": dev,
                 "mixed_vs_fp64": (N / (msm * 1e-3)) / value}

    # operator application: Laplacian, 1 and 8 fields
    log("timing SpMV / SpMM")
    nnz = N * n
    spmv = {}
    for F in (1, 8):
        x = ctx.to_device(np.random.default_rng(F).standard_normal((N, F)) if F > 1 else
                          np.random.default_rng(1).standard_normal(N))
        y = ctx.empty((N, F) if F > 1 else (N,), np.float64)
        for _ in range(3):
            eng.spmv(col, vals[2], x, y, uniform_n=n)
        ctx.synchronize()
        reps = 50
        ctx.timer_start()
        for _ in range(reps):
            eng.spmv(col, vals[2], x, y, uniform_n=n)
        t = ctx.timer_stop() / reps
        bytes_alg = nnz * 12 + 8 * F * (N + N)
        spmv[f"F{F}"] = {"ms": t, "GB_s": bytes_alg / (t * 1e-3) / 1e9, "bytes": bytes_alg,
                         "frac_hbm_peak": bytes_alg / (t * 1e-3) / 1e9 / PEAK_HBM_GBS}
        log(f"  spmv F={F}: {t * 1e3:.1f} us, {spmv[f'F{F}']['GB_s']:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (oracle restatement)")
        from oracle import oracle as O

        threads = min(os.cpu_count() or 1, 16)
        rows = np.arange(0, N, 8, dtype=np.int64)  # every 8th row: 131,072 stencils
        t0 = time.perf_counter()
        st, _ = O.knn(X, X[rows], n, nthreads=threads)
        W, _, status = O.weights(X, X, _scatter(st, rows, N, n), None, p, polydeg,
                                 [(O.OP_DX, 0), (O.OP_DY, 0), (O.OP_LAP, 0)], rows=rows, nthreads=threads)
        dt = time.perf_counter() - t0
        # the GPU rows of the same sample, against the oracle (relative row error)
        gst = np.sort(st, axis=1)
        same_pattern = bool(np.array_equal(colh[rows], gst))
        errs = []
        for o in range(3):
            ref = np.take_along_axis(W[:, o, :], np.argsort(st, axis=1, kind="stable"), 1)
            errs.append(float((np.abs(wh[o][rows] - ref).max(axis=1) / np.abs(ref).max(axis=1)).max()))
        cpu = {"value": len(rows) / dt, "unit": "stencils/s", "cores": threads, "kind": "port",
               "sample": f"{len(rows)} of {N} rows (every 8th), kNN + weights, {dt:.1f} s",
               "gpu_pattern_equal": same_pattern, "gpu_max_row_err_vs_oracle": max(errs)}
        log(f"  CPU {cpu['value']:.0f} stencils/s on {threads} threads; GPU err vs oracle {max(errs):.2e}")

    out = {
        "metric": METRIC, "value": value, "unit": "stencils/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "gen1m: 1,048,576 jittered nodes (1024x1024, +-0.25h, seed 2), PHS r^7, "
                               "degree-4 monomials, n=31, operators Dx, Dy, Lap as FP64 CSR",
                   "nodes": N, "stencil": n, "polydeg": polydeg, "phs_power": p, "ops": ops},
        "roofline": {"bound": "mfma", "pipe": "valu-fp64", "kernel": "solve_kernel<double,31,15,3>",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": read_traffic("solve_gen1m_pmc.txt")},
        "kernel_only_stencils_per_s": N / (ms_solve * 1e-3),
        "ms_knn": ms_knn, "ms_solve": ms_solve,
        "spmv": spmv, "mixed": mixed, "cpu_baseline": cpu, "device": info,
    }
    print(json.dumps(out))


def _scatter(st_rows, rows, N, n):
    full = np.zeros((N, n), np.int32)
    full[rows] = st_rows
    return full


if __name__ == "__main__":
    main()

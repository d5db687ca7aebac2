"""Host-side mirror of the operator API of RadialBasisFiniteDifferences.jl,
running on librbffd (gfx950 HIP kernels behind a C ABI).

    generate_operator(X, Y, p, n, polydeg)                       src/generate_operator.jl:29
    generate_operator(X, Y, p, n, polydeg, X_idx_in, X_idx_bc,
                      X_idx_bc_g, Y_idx_in, Y_idx_bc, Y_idx_bc_g)   src/generate_operator.jl:192
    generate_operator(X, p, n, polydeg)                           src/generate_operator.jl:354
        -> (E, Dx, Dy, Dxx, Dyy, Dxy)
    hyperviscosity_operator(k, X, Y, p, n, polydeg[, index sets])  src/hyperviscosity_operator.jl:26,177
    hyperviscosity_operator(k, X, p, n, polydeg)                   src/hyperviscosity_operator.jl:314
        -> (Dxk, Dyk)

Arguments keep the reference's meaning.  Points are (N, 2) float64 arrays
(the memory of Vector{SVector{2,Float64}}); index sets are 0-based here (the
Julia binding passes 1-based sets with index_base = 1).  Results are
scipy.sparse.csc_matrix with the layout sparse(I, J, V) gives: rows sorted in
each column, explicit zeros kept, shape (number of Y nodes, largest referenced
column + 1).  Errors follow the reference: SingularException for an exactly
singular local system, UndefRefError for a node in no index set, ValueError
where Julia raises ArgumentError (e.g. n larger than the node count).
"""
from __future__ import annotations

import ctypes
import os
import re
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import OP, PREC, check, rbffd_batch_set, rbffd_boundary, rbffd_ghost_closure, rbffd_stats
from .device import Context, DeviceArray

_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _pts(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != 2:
        raise ValueError("points must be an (N, 2) array")
    return a


def _vp(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def host_threads() -> int:
    """Threads for large host copies: RBFFD_HOST_THREADS, else OMP_NUM_THREADS
    (a positive leading integer), else the CPUs this process may run on; at
    most 32.  The same rule as the C library's device-to-host copy pool."""
    for name in ("RBFFD_HOST_THREADS", "OMP_NUM_THREADS"):
        m = re.match(r"\s*([+-]?\d+)", os.environ.get(name, ""))
        if m and int(m.group(1)) > 0:
            return min(32, int(m.group(1)))
    try:
        return max(1, min(32, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(32, os.cpu_count() or 1))


_copy_pool = (None, 0)


def _host_copy(a: np.ndarray, dtype=None) -> np.ndarray:
    """A copy of the contiguous 1-D array a, converted to dtype when given.
    Large copies are split over host_threads() threads (numpy releases the GIL
    while it copies): touching the new pages for the first time costs as much
    as the copy itself."""
    global _copy_pool
    out = np.empty(a.shape, a.dtype if dtype is None else dtype)
    T = host_threads()
    if T == 1 or a.nbytes < (16 << 20):
        np.copyto(out, a)
        return out
    pool, size = _copy_pool
    if pool is None or size != T:
        if pool is not None:
            pool.shutdown(wait=True)
        pool = ThreadPoolExecutor(max_workers=T)
        _copy_pool = (pool, T)
    edges = [len(a) * i // T for i in range(T + 1)]

    def part(i):
        np.copyto(out[edges[i]:edges[i + 1]], a[edges[i]:edges[i + 1]])

    list(pool.map(part, range(T)))
    return out


class _Boundary:
    """Keeps the flattened index sets alive for the duration of a call."""

    def __init__(self, X_idx_in, X_idx_bc, X_idx_bc_g, index_base=0):
        self.in_idx = np.ascontiguousarray(np.asarray(X_idx_in, dtype=np.int64).ravel())
        bcs = [np.asarray(b, dtype=np.int64).ravel() for b in X_idx_bc]
        ghs = [np.asarray(g, dtype=np.int64).ravel() for g in X_idx_bc_g]
        if len(bcs) != len(ghs):
            raise ValueError("X_idx_bc and X_idx_bc_g must have one entry per boundary marker")
        self.bc_ptr = np.zeros(len(bcs) + 1, np.int64)
        self.bc_ptr[1:] = np.cumsum([len(b) for b in bcs])
        self.gh_ptr = np.zeros(len(ghs) + 1, np.int64)
        self.gh_ptr[1:] = np.cumsum([len(g) for g in ghs])
        self.bc_idx = np.ascontiguousarray(np.concatenate(bcs) if bcs else np.zeros(0, np.int64))
        self.gh_idx = np.ascontiguousarray(np.concatenate(ghs) if ghs else np.zeros(0, np.int64))
        self.s = rbffd_boundary(index_base, len(self.in_idx), _vp(self.in_idx), len(bcs), _vp(self.bc_ptr),
                                _vp(self.bc_idx), _vp(self.gh_ptr), _vp(self.gh_idx))


def boundary_groups(N, X_idx_in, X_idx_bc, X_idx_bc_g, index_base=0):
    """Method-2 group ids (xgroup, qgroup) of N nodes: the visibility rule of
    calculateneighbors (src/calculateneighbors.jl:16-42, 83-94)."""
    L = _lib.load()
    bnd = _Boundary(X_idx_in, X_idx_bc, X_idx_bc_g, index_base)
    xg = np.empty(N, np.int32)
    qg = np.empty(N, np.int32)
    check(L.rbffd_boundary_groups(ctypes.byref(bnd.s), int(N), _vp(xg), _vp(qg)))
    return xg, qg


def _op_arrays(ops, k):
    kinds = np.array([OP[o] for o in ops], np.int32)
    ks = np.array([k if o in ("Dxk", "Dyk", "Lapk") else 0 for o in ops], np.int32)
    return kinds, ks


def operators(X, Y, p, n, polydeg, ops, k=0, boundary=None, precision="fp64", ctx=None,
              return_stats=False):
    """Operators `ops` (names from E, Dx, Dy, Dxx, Dyy, Dxy, Dxk, Dyk, Lap,
    Lapk) of the reference pipeline; Y None selects method 3, boundary =
    (X_idx_in, X_idx_bc, X_idx_bc_g) selects method 2.  Returns a tuple of
    csc_matrix in the order of `ops`."""
    ctx = ctx or default_context()
    L = _lib.load()
    X = _pts(X)
    N = len(X)
    Yp = None if Y is None else _pts(Y)
    M = N if Yp is None else len(Yp)
    n = int(n)
    kinds, ks = _op_arrays(ops, int(k))
    nops = len(ops)
    nnz = M * n
    colptr = np.empty(N + 1, np.int64)
    rowval = np.empty(nnz, np.int64)
    nz = [np.empty(nnz, np.float64) for _ in range(nops)]
    nzp = (ctypes.c_void_p * nops)(*[_vp(a) for a in nz])
    ncols = ctypes.c_int64()
    stats = rbffd_stats()
    bnd = None
    if boundary is not None:
        bnd = _Boundary(*boundary)
    Y_ptr = None
    if Yp is not None:
        Y_ptr = _vp(X) if Y is X else _vp(Yp)
    t0 = time.perf_counter()
    code = L.rbffd_generate_operator(ctx.handle, _vp(X), N, Y_ptr, M, int(p), n, int(polydeg),
                                     ctypes.byref(bnd.s) if bnd else None, nops, _vp(kinds), _vp(ks),
                                     PREC[precision], 0, _vp(colptr), _vp(rowval),
                                     ctypes.cast(nzp, ctypes.c_void_p), ctypes.byref(ncols), ctypes.byref(stats))
    t_call = time.perf_counter() - t0
    check(code)
    nc = int(ncols.value)
    # every matrix owns its index arrays, as the reference's separate sparse()
    # calls return independent matrices (src/generate_operator.jl:176-181): a
    # structural edit of one leaves the others intact.  scipy keeps 32-bit
    # indices whenever they fit, and converting int64 ones itself costs a scan
    # and a single-threaded copy per matrix, so the indices are converted here,
    # once, and copied per matrix with the host threads.
    idx_t = np.int32 if max(nnz, M, nc) <= np.iinfo(np.int32).max else np.int64
    rv = _host_copy(rowval, idx_t)
    cp = colptr[: nc + 1].astype(idx_t)
    mats = tuple(sp.csc_matrix((nz[o], rv if o == 0 else _host_copy(rv), cp if o == 0 else cp.copy()),
                               shape=(M, nc), copy=False) for o in range(nops))
    if return_stats:
        st = stats.as_dict()
        st["host_call_s"] = t_call  # the C entry point alone: H2D, kernels, D2H
        return mats, st
    return mats


_SIX = ("E", "Dx", "Dy", "Dxx", "Dyy", "Dxy")


def generate_operator(*args, precision="fp64", ctx=None):
    """Mirror of generate_operator (src/generate_operator.jl:29,192,354):
    returns (E, Dx, Dy, Dxx, Dyy, Dxy)."""
    if len(args) == 4:
        X, p, n, polydeg = args
        return operators(X, None, p, n, polydeg, _SIX, precision=precision, ctx=ctx)
    if len(args) == 5:
        X, Y, p, n, polydeg = args
        return operators(X, Y, p, n, polydeg, _SIX, precision=precision, ctx=ctx)
    if len(args) == 11:
        X, Y, p, n, polydeg, X_idx_in, X_idx_bc, X_idx_bc_g, _Yin, _Ybc, _Ybcg = args
        # the Y index sets are accepted and unused, as in the reference
        # (src/calculateneighbors.jl:1-2)
        return operators(X, Y, p, n, polydeg, _SIX, boundary=(X_idx_in, X_idx_bc, X_idx_bc_g),
                         precision=precision, ctx=ctx)
    raise TypeError("generate_operator takes (X, p, n, polydeg), (X, Y, p, n, polydeg) "
                    "or (X, Y, p, n, polydeg, X_idx_in, X_idx_bc, X_idx_bc_g, Y_idx_in, Y_idx_bc, Y_idx_bc_g)")


def hyperviscosity_operator(k_deriv, *args, precision="fp64", ctx=None):
    """Mirror of hyperviscosity_operator (src/hyperviscosity_operator.jl:26,177,314):
    returns (Dxk, Dyk), the k-th pure derivatives scaled by s^k."""
    ops = ("Dxk", "Dyk")
    if len(args) == 4:
        X, p, n, polydeg = args
        return operators(X, None, p, n, polydeg, ops, k=k_deriv, precision=precision, ctx=ctx)
    if len(args) == 5:
        X, Y, p, n, polydeg = args
        return operators(X, Y, p, n, polydeg, ops, k=k_deriv, precision=precision, ctx=ctx)
    if len(args) == 11:
        X, Y, p, n, polydeg, X_idx_in, X_idx_bc, X_idx_bc_g, _Yin, _Ybc, _Ybcg = args
        return operators(X, Y, p, n, polydeg, ops, k=k_deriv, boundary=(X_idx_in, X_idx_bc, X_idx_bc_g),
                         precision=precision, ctx=ctx)
    raise TypeError("hyperviscosity_operator takes (k, X, p, n, polydeg), (k, X, Y, p, n, polydeg) "
                    "or (k, X, Y, p, n, polydeg, 6 index sets)")


class Engine:
    """Device-resident entry points (inputs and outputs stay in HBM)."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.L = _lib.load()

    def knn(self, X: DeviceArray, Q: DeviceArray, k: int, xgroup=None, qgroup=None, want_d2=False):
        N, M = X.shape[0], Q.shape[0]
        idx = self.ctx.empty((M, k), np.int32)
        d2 = self.ctx.empty((M, k), np.float64) if want_d2 else None
        check(self.L.rbffd_knn_device(self.ctx.handle, X.ptr, N, Q.ptr, M, int(k),
                                      xgroup.ptr if xgroup is not None else None,
                                      qgroup.ptr if qgroup is not None else None,
                                      idx.ptr, d2.ptr if d2 is not None else None))
        return (idx, d2) if want_d2 else idx

    def weights(self, X, Y, stencils, y2x, p, polydeg, ops, k=0, method=1, precision="fp64",
                colidx=None, vals=None, status=None):
        n = stencils.shape[1]
        M = Y.shape[0]
        kinds, ks = _op_arrays(ops, k)
        if colidx is None:
            colidx = self.ctx.empty((M, n), np.int32)
        if vals is None:
            vals = [self.ctx.empty((M, n), np.float64) for _ in ops]
        if status is None:
            status = self.ctx.empty((M,), np.int32)
        vp = (ctypes.c_void_p * len(ops))(*[v.ptr for v in vals])
        check(self.L.rbffd_weights_device(self.ctx.handle, X.ptr, X.shape[0], Y.ptr, M, stencils.ptr, n,
                                          y2x.ptr if y2x is not None else None, int(p), int(polydeg), int(method),
                                          len(ops), _vp(kinds), _vp(ks), PREC[precision], colidx.ptr,
                                          ctypes.cast(vp, ctypes.c_void_p), status.ptr))
        return colidx, vals, status

    def generate(self, X, Y, p, n, polydeg, ops, k=0, boundary=None, precision="fp64",
                 colidx=None, vals=None, stencils_out=None, y2x_out=None):
        """kNN + weights on device.  Y None: method 3; Y is X: collocated."""
        N = X.shape[0]
        M = N if Y is None else Y.shape[0]
        kinds, ks = _op_arrays(ops, k)
        if colidx is None:
            colidx = self.ctx.empty((M, n), np.int32)
        if vals is None:
            vals = [self.ctx.empty((M, n), np.float64) for _ in ops]
        vp = (ctypes.c_void_p * len(ops))(*[v.ptr for v in vals])
        stats = rbffd_stats()
        bnd = _Boundary(*boundary) if boundary is not None else None
        Y_ptr = None if Y is None else Y.ptr
        code = self.L.rbffd_generate_device(self.ctx.handle, X.ptr, N, Y_ptr, M, int(p), int(n), int(polydeg),
                                            ctypes.byref(bnd.s) if bnd else None, len(ops), _vp(kinds), _vp(ks),
                                            PREC[precision], colidx.ptr, ctypes.cast(vp, ctypes.c_void_p),
                                            stencils_out.ptr if stencils_out is not None else None,
                                            y2x_out.ptr if y2x_out is not None else None, ctypes.byref(stats))
        check(code)
        return colidx, vals, stats.as_dict()

    def generate_batch(self, sets, ops, k=0, use_graph=True):
        """Uneven batch: sets is a list of dicts with X (DeviceArray, N x 2), p,
        n, polydeg and optionally xgroup/qgroup (DeviceArray int32, method-2
        groups from boundary_groups) and colidx/vals/status outputs (allocated
        here when absent).  Returns (per-set [(colidx, vals, status)], set
        status array, stats); raises on the first failing set like the
        single-set calls."""
        kinds, ks = _op_arrays(ops, k)
        arr = (rbffd_batch_set * len(sets))()
        out = []
        for i, S in enumerate(sets):
            X = S["X"]
            N, n = X.shape[0], int(S["n"])
            col = S["colidx"] if S.get("colidx") is not None else self.ctx.empty((N, n), np.int32)
            vals = S["vals"] if S.get("vals") is not None else [self.ctx.empty((N, n), np.float64) for _ in ops]
            st = S["status"] if S.get("status") is not None else self.ctx.empty((N,), np.int32)
            b = arr[i]
            b.X, b.N, b.p, b.n, b.polydeg = X.ptr, N, int(S["p"]), n, int(S["polydeg"])
            b.xgroup = S["xgroup"].ptr if S.get("xgroup") is not None else None
            b.qgroup = S["qgroup"].ptr if S.get("qgroup") is not None else None
            b.colidx = col.ptr
            for o, v in enumerate(vals):
                b.vals[o] = v.ptr
            b.status = st.ptr
            out.append((col, vals, st))
        set_status = np.zeros(len(sets), np.int32)
        stats = rbffd_stats()
        self._batch_keepalive = (arr, out)
        code = self.L.rbffd_generate_batch_device(self.ctx.handle, len(sets), arr, len(ops), _vp(kinds), _vp(ks),
                                                  1 if use_graph else 0, _vp(set_status), ctypes.byref(stats))
        if code != 0:
            self.last_set_status = set_status
        check(code)
        return out, set_status, stats.as_dict()

    def combine(self, vals, coef, out=None, out32=None):
        """out = sum_o coef[o] vals[o] for value arrays sharing one pattern;
        out32 receives the FP32 rounding."""
        nnz = vals[0].size
        if out is None and out32 is None:
            out = self.ctx.empty(vals[0].shape, np.float64)
        vp = (ctypes.c_void_p * len(vals))(*[v.ptr for v in vals])
        cf = np.ascontiguousarray(coef, np.float64)
        check(self.L.rbffd_combine_device(self.ctx.handle, nnz, len(vals), ctypes.cast(vp, ctypes.c_void_p),
                                          _vp(cf), out.ptr if out is not None else None,
                                          out32.ptr if out32 is not None else None))
        return out if out is not None else out32

    def ssprk3(self, col, vals, u, dt, steps, uniform_n=0, rowptr=None, fixed=None, work=None, order=None):
        """Advance u (N or N x F, device) by `steps` SSP-RK3 steps of du/dt = L u."""
        N = u.shape[0]
        F = 1 if len(u.shape) == 1 else u.shape[1]
        if work is None:
            work = self.ctx.empty((2 * N * F,), np.float64)
        vt = 1 if vals.dtype == np.float32 else 0
        check(self.L.rbffd_ssprk3_device(self.ctx.handle, N, rowptr.ptr if rowptr is not None else None, col.ptr,
                                         vals.ptr, vt, int(uniform_n), order.ptr if order is not None else None, F,
                                         u.ptr, work.ptr,
                                         fixed.ptr if fixed is not None else None, float(dt), int(steps)))
        return u

    def ssprk3_closure(self, col, vals, u, dt, steps, closures, uniform_n=0, rowptr=None, fixed=None, work=None,
                       order=None):
        """ssprk3 with ghost-node closures: `closures` is a list of GhostClosure
        (device arrays), applied in order before the first step and after every
        stage."""
        N = u.shape[0]
        F = 1 if len(u.shape) == 1 else u.shape[1]
        if work is None:
            work = self.ctx.empty((2 * N * F,), np.float64)
        vt = 1 if vals.dtype == np.float32 else 0
        arr = (rbffd_ghost_closure * max(1, len(closures)))()
        for i, g in enumerate(closures):
            arr[i] = rbffd_ghost_closure(g.nbc, g.rowptr.ptr, g.col.ptr, g.vals.ptr, g.ginv.ptr, g.ghost.ptr)
        check(self.L.rbffd_ssprk3_closure_device(self.ctx.handle, N, rowptr.ptr if rowptr is not None else None,
                                                 col.ptr, vals.ptr, vt, int(uniform_n),
                                                 order.ptr if order is not None else None, F, u.ptr, work.ptr,
                                                 fixed.ptr if fixed is not None else None, len(closures), arr,
                                                 float(dt), int(steps)))
        return u

    def dense_inverse(self, A):
        """inv(A) of a dense n x n device matrix (column-major, i.e. the
        transpose of a C-ordered array) as a new device array."""
        n = A.shape[0]
        out = self.ctx.empty((n, n), np.float64)
        check(self.L.rbffd_dense_inverse_device(self.ctx.handle, n, A.ptr, out.ptr))
        return out

    def csr_to_csc(self, colidx, vals, ncols, index_base=0):
        M, n = colidx.shape
        colptr = self.ctx.empty((ncols + 1,), np.int64)
        rowval = self.ctx.empty((M * n,), np.int64)
        cvals = [self.ctx.empty((M * n,), np.float64) for _ in vals]
        vp = (ctypes.c_void_p * max(1, len(vals)))(*[v.ptr for v in vals])
        cp = (ctypes.c_void_p * max(1, len(vals)))(*[v.ptr for v in cvals])
        used = ctypes.c_int64()
        check(self.L.rbffd_csr_to_csc_device(self.ctx.handle, M, n, ncols, colidx.ptr, len(vals),
                                             ctypes.cast(vp, ctypes.c_void_p), colptr.ptr, rowval.ptr,
                                             ctypes.cast(cp, ctypes.c_void_p), int(index_base), ctypes.byref(used)))
        return colptr, rowval, cvals, int(used.value)

    def spmv(self, col, vals, x, y, uniform_n=0, rowptr=None, alpha=1.0, beta=0.0, order=None):
        M = y.shape[0]
        F = 1 if len(x.shape) == 1 else x.shape[1]
        vt = 1 if vals.dtype == np.float32 else 0
        check(self.L.rbffd_spmv_device(self.ctx.handle, M, rowptr.ptr if rowptr is not None else None, col.ptr,
                                       vals.ptr, vt, F, x.ptr, y.ptr, float(alpha), float(beta), int(uniform_n),
                                       order.ptr if order is not None else None))
        return y

    def lsq_solve(self, rowptr, col, vals, b, x, method="auto", tol=1e-12, max_iters=100000):
        """x = argmin ||D x - b|| for D (M x N, M >= N) in device CSR (int64 rowptr,
        int32 columns, float64 values); x (device, N) holds the start (CGLS) and
        gets the solution.  method: "auto", "qr" (direct) or "cgls".  Returns
        (CGLS iterations, relative normal-equation residual)."""
        M = b.shape[0]
        N = x.shape[0]
        it = _lib._I64(0)
        rel = _lib._D(0.0)
        m = {"auto": 0, "cgls": 1, "qr": 2}[method]
        check(self.L.rbffd_lsq_solve_device(self.ctx.handle, M, N, rowptr.ptr, col.ptr, vals.ptr, b.ptr, x.ptr, m,
                                            float(tol), int(max_iters), ctypes.byref(it), ctypes.byref(rel)))
        return int(it.value), float(rel.value)

    def read_sum(self, x, out=None):
        """Device sum of x (float64) by a read-only stream; returns the 1-element
        device array (asynchronous)."""
        out = out if out is not None else self.ctx.empty((1,), np.float64)
        check(self.L.rbffd_read_sum_device(self.ctx.handle, x.ptr, x.size, out.ptr))
        return out

    def spmv_plan(self, col, vals, M, N, rowptr=None, uniform_n=0, order=None):
        """An application plan of the CSR operator (M x N) for repeated
        products (rbffd_spmv_plan_create); see SpmvPlan."""
        return SpmvPlan(self, col, vals, M, N, rowptr=rowptr, uniform_n=uniform_n, order=order)

    def locality_order(self, X):
        """Morton order of the nodes of X (device, N x 2), as an int32 device array."""
        order = self.ctx.empty((X.shape[0],), np.int32)
        check(self.L.rbffd_locality_order_device(self.ctx.handle, X.ptr, X.shape[0], order.ptr))
        return order


class SpmvPlan:
    """An application plan of one CSR operator (rbffd_spmv_plan_create): the
    rows in blocks of 256 of `order` (a locality order such as
    Engine.locality_order of the rows' nodes), each block's distinct columns
    listed, so a product stages a block's x rows in LDS.  Products equal
    Engine.spmv bit for bit.  `info`: row blocks, largest listed halo, blocks
    over 2048 distinct columns, nonzeros.  The plan keeps no reference to the
    arrays it was made from."""

    def __init__(self, eng, col, vals, M, N, rowptr=None, uniform_n=0, order=None):
        self.eng, self.ctx, self.L = eng, eng.ctx, eng.L
        self.shape = (int(M), int(N))
        vt = 1 if vals.dtype == np.float32 else 0
        h = ctypes.c_void_p()
        info = (ctypes.c_int64 * 4)()
        check(self.L.rbffd_spmv_plan_create(self.ctx.handle, int(M), int(N),
                                            rowptr.ptr if rowptr is not None else None, col.ptr, vals.ptr, vt,
                                            int(uniform_n), order.ptr if order is not None else None,
                                            ctypes.byref(h), ctypes.cast(info, ctypes.c_void_p)))
        self.handle = h
        self.info = {"blocks": info[0], "max_halo": info[1], "overflow_blocks": info[2], "nnz": info[3]}

    def apply(self, x, y, alpha=1.0, beta=0.0):
        """y = alpha A x + beta y (device arrays, 1, 2, 4 or 8 fields node-major)."""
        if x.shape[0] != self.shape[1] or y.shape[0] != self.shape[0] or x.shape[1:] != y.shape[1:]:
            raise ValueError("SpmvPlan: shape mismatch")
        F = 1 if len(x.shape) == 1 else x.shape[1]
        check(self.L.rbffd_spmv_plan_apply(self.ctx.handle, self.handle, F, x.ptr, y.ptr, float(alpha),
                                           float(beta)))
        return y

    def ssprk3(self, u, dt, steps, fixed=None, work=None):
        """Engine.ssprk3 with this plan's operator and row order."""
        N = u.shape[0]
        F = 1 if len(u.shape) == 1 else u.shape[1]
        if work is None:
            work = self.ctx.empty((2 * N * F,), np.float64)
        check(self.L.rbffd_ssprk3_plan_device(self.ctx.handle, self.handle, F, u.ptr, work.ptr,
                                              fixed.ptr if fixed is not None else None, float(dt), int(steps)))
        return u

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            check(self.L.rbffd_spmv_plan_destroy(self.handle))
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceOperator:
    """A sparse operator resident in device memory, applied by librbffd's SpMV
    (rbffd_spmv_device): the `A * x` / `mul!(y, A, x)` of the reference's
    harnesses (test/poisson_test.jl:121, examples/adv_diff_test.jl:151-152) on
    a matrix that generate_operator returned (or any scipy sparse matrix).
    The CSR of A (int64 row pointers, int32 columns, float64 values) is
    uploaded once; products take host or device vectors, with 1 to 8 fields
    stored node-major (x[node, field]).  plan=True builds an application plan
    (SpmvPlan; `order`, a permutation of the rows such as a Morton order of
    their nodes, keeps its row blocks compact) for an operator applied many
    times; products are bitwise the same."""

    def __init__(self, A, ctx=None, plan=False, order=None):
        self.ctx = ctx or default_context()
        self.eng = Engine(self.ctx)
        R = sp.csr_matrix(A)
        R.sort_indices()
        if R.nnz > np.iinfo(np.int32).max or R.shape[1] > np.iinfo(np.int32).max:
            raise ValueError("DeviceOperator: more than 2^31 - 1 entries or columns")
        self.shape = R.shape
        self.nnz = R.nnz
        self.rowptr = self.ctx.to_device(R.indptr.astype(np.int64))
        self.col = self.ctx.to_device(R.indices.astype(np.int32))
        self.vals = self.ctx.to_device(R.data.astype(np.float64))
        self.plan = None
        if plan:
            od = None
            if order is not None:
                od = order if isinstance(order, DeviceArray) else self.ctx.to_device(
                    np.ascontiguousarray(order, dtype=np.int32))
            self.plan = self.eng.spmv_plan(self.col, self.vals, self.shape[0], self.shape[1], rowptr=self.rowptr,
                                           order=od)

    def apply(self, x: DeviceArray, y: DeviceArray, alpha=1.0, beta=0.0):
        """y = alpha A x + beta y on device arrays (asynchronous on the context stream)."""
        if x.shape[0] != self.shape[1] or y.shape[0] != self.shape[0] or x.shape[1:] != y.shape[1:]:
            raise ValueError("DeviceOperator: shape mismatch")
        if self.plan is not None:
            return self.plan.apply(x, y, alpha=alpha, beta=beta)
        return self.eng.spmv(self.col, self.vals, x, y, rowptr=self.rowptr, alpha=alpha, beta=beta)

    def __matmul__(self, x):
        """A @ x for a host vector or (N, F) block; returns a host array."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        xd = self.ctx.to_device(x)
        yd = self.ctx.empty((self.shape[0],) + x.shape[1:], np.float64)
        self.apply(xd, yd)
        return yd.get()


def solve(D, f, method="auto", tol=1e-12, max_iters=100000, ctx=None, info=False):
    """`D \\ f` of the reference's PDE harness (test/poisson_test.jl:121,
    test/mesh_import_test.jl:150) for a scipy sparse D with at least as many
    rows as columns: the least-squares solution (the solution when square), on
    the GPU by a dense QR ("qr") or CGLS on the column-equilibrated normal
    equations ("cgls"); "auto" takes QR while the dense copy has at most 2^28
    entries.  With info=True also returns (CGLS iterations, relative
    normal-equation residual).  A rank-deficient D raises SingularException
    (QR), as Julia's `\\` does for a singular square D."""
    ctx = ctx or default_context()
    A = sp.csr_matrix(D)
    A.sort_indices()
    M, N = A.shape
    f = np.ascontiguousarray(np.asarray(f, dtype=np.float64).ravel())
    if f.shape[0] != M:
        raise ValueError("f must have one entry per row of D")
    eng = Engine(ctx)
    rowptr = ctx.to_device(A.indptr.astype(np.int64))
    col = ctx.to_device(A.indices.astype(np.int32))
    vals = ctx.to_device(A.data.astype(np.float64))
    b = ctx.to_device(f)
    x = ctx.zeros((N,), np.float64)
    its, rel = eng.lsq_solve(rowptr, col, vals, b, x, method=method, tol=tol, max_iters=max_iters)
    out = x.get()
    return (out, (its, rel)) if info else out


class GhostClosure:
    """Device data of one marker's ghost-node closure
    (examples/adv_diff_test.jl:107-146): the CSR of D[bc, :] without the
    marker's ghost columns, inv(D[bc, ghost]) and the ghost node indices."""

    def __init__(self, ctx, D, bc, ghost, eng=None):
        """D: scipy sparse N x N normal-derivative operator of the marker (Dx or
        Dy); bc, ghost: its boundary nodes and their ghost nodes, paired in
        order.  inv(D[bc, ghost]) is computed on the device."""
        D = sp.csr_matrix(D)
        bc = np.asarray(bc, dtype=np.int64)
        ghost = np.asarray(ghost, dtype=np.int64)
        rows = D[bc]
        keep = np.ones(D.shape[1], dtype=bool)
        keep[ghost] = False
        inner = (rows @ sp.diags(keep.astype(np.float64))).tocsr()
        inner.eliminate_zeros()
        inner.sort_indices()
        Wg = rows[:, ghost].toarray()
        eng = eng or Engine(ctx)
        self.nbc = len(bc)
        self.rowptr = ctx.to_device(inner.indptr.astype(np.int64))
        self.col = ctx.to_device(inner.indices.astype(np.int32))
        self.vals = ctx.to_device(inner.data.astype(np.float64))
        self.ginv = eng.dense_inverse(ctx.to_device(np.ascontiguousarray(Wg.T)))  # column-major
        self.ghost = ctx.to_device(ghost.astype(np.int32))
        self.host = (inner, Wg, bc, ghost)

"""TEST INFRASTRUCTURE ONLY -- the oracle.

Python glue over ``liboracle.so`` (the C restatement in ``rbffd_oracle.c``)
plus a numpy restatement of the sparse assembly of the reference.  Only
``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module; the product path never does.

Reference behaviour restated here (paths relative to the reference root):

* ``generate_operator(X, Y, p, n, polydeg)``      src/generate_operator.jl:29-190
* ``generate_operator(X, Y, ..., index sets)``    src/generate_operator.jl:192-352
* ``generate_operator(X, p, n, polydeg)``          src/generate_operator.jl:354-491
* ``hyperviscosity_operator(k, ...)``             src/hyperviscosity_operator.jl:26-440
* ``calculateneighbors``                          src/calculateneighbors.jl:1-97
* ``sparse(I, J, V)`` assembly                    src/generate_operator.jl:171-182
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

OP_E, OP_DX, OP_DY, OP_DXX, OP_DYY, OP_DXY, OP_XK, OP_YK, OP_LAP, OP_LAPK = range(10)
OP_NAMES = {"E": OP_E, "Dx": OP_DX, "Dy": OP_DY, "Dxx": OP_DXX, "Dyy": OP_DYY,
            "Dxy": OP_DXY, "Dxk": OP_XK, "Dyk": OP_YK, "Lap": OP_LAP, "Lapk": OP_LAPK}

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-B" if force else "liboracle.so"], cwd=_HERE, check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        L.orc_monomials.argtypes = [ctypes.c_int, P, P]
        L.orc_monomials.restype = ctypes.c_int
        L.orc_phs_deriv.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_phs_deriv.restype = ctypes.c_double
        L.orc_op_phs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double]
        L.orc_op_phs.restype = ctypes.c_double
        L.orc_phs_deriv_q.argtypes = L.orc_phs_deriv.argtypes
        L.orc_phs_deriv_q.restype = ctypes.c_double
        knn_args = [P, ctypes.c_long, P, ctypes.c_long, ctypes.c_int, P, P, P, P, ctypes.c_int]
        L.orc_knn_brute.argtypes = knn_args
        L.orc_knn_grid.argtypes = knn_args
        L.orc_weights.argtypes = [P, ctypes.c_long, P, ctypes.c_long, P, ctypes.c_int, P,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int,
                                  P, ctypes.c_long, P, P, P, ctypes.c_int]
        L.orc_weights.restype = ctypes.c_long
        L.orc_spmv_csr.argtypes = [ctypes.c_long, P, P, P, ctypes.c_int, P, P, ctypes.c_int]
        L.orc_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _pts(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert a.ndim == 2 and a.shape[1] == 2
    return a


def monomials(polydeg: int):
    """Exponents (a, b) of x^a y^b in basis order (src/polynomialbasis.jl:8-9)."""
    P = lib().orc_monomials(polydeg, None, None)
    ea = np.zeros(P, np.int32)
    eb = np.zeros(P, np.int32)
    lib().orc_monomials(polydeg, _p(ea), _p(eb))
    return ea, eb


def phs_deriv(m, a, b, vx, vy, quad=False):
    f = lib().orc_phs_deriv_q if quad else lib().orc_phs_deriv
    return f(int(m), int(a), int(b), float(vx), float(vy))


def op_phs(op, k, m, vx, vy, sx=1.0, sy=1.0):
    return lib().orc_op_phs(int(op), int(k), int(m), float(vx), float(vy), float(sx), float(sy))


def _groups(g, n):
    if g is None:
        return None
    g = np.ascontiguousarray(g, dtype=np.int32)
    assert g.shape == (n,)
    return g


def knn(X, Q, k, xgroup=None, qgroup=None, method="grid", nthreads=0):
    """k nearest X nodes of every query, ascending (d^2, index); 0-based."""
    X = _pts(X)
    Q = _pts(Q)
    N, M = len(X), len(Q)
    idx = np.empty((M, k), np.int32)
    d2 = np.empty((M, k), np.float64)
    xg = _groups(xgroup, N)
    qg = _groups(qgroup, M)
    if (xg is None) != (qg is None):
        raise ValueError("xgroup and qgroup go together")
    f = lib().orc_knn_grid if method == "grid" else lib().orc_knn_brute
    f(_p(X), N, _p(Q), M, int(k), _p(xg), _p(qg), _p(idx), _p(d2), int(nthreads))
    return idx, d2


def boundary_groups(N, X_idx_in, X_idx_bc, X_idx_bc_g):
    """Group ids that restate the masked searches of calculateneighbors
    (src/calculateneighbors.jl:16-42 and :83-87), 0-based index sets.

    Returns (xgroup, qgroup): marker b (1-based) queries with group b and sees
    nodes of group 0 (interior and all boundaries) and of group b (its ghosts);
    interior queries see everything (qgroup -1); nodes in no set carry -2.
    """
    xg = np.full(N, -2, np.int32)
    qg = np.full(N, -1, np.int32)
    xg[np.asarray(X_idx_in, dtype=np.int64)] = 0
    for bc in X_idx_bc:
        xg[np.asarray(bc, dtype=np.int64)] = 0
    for b, (bc, gh) in enumerate(zip(X_idx_bc, X_idx_bc_g), start=1):
        gh = np.asarray(gh, dtype=np.int64)
        xg[gh] = b
        qg[np.asarray(bc, dtype=np.int64)] = b
        qg[gh] = b
    qg[np.asarray(X_idx_in, dtype=np.int64)] = -1
    return xg, qg


def weights(X, Y, stencils, y2x, m, polydeg, ops, method=1, quad=False, solve="inv",
            rows=None, nthreads=0):
    """Per-row stencil weights in stencil order.

    ops: list of (kind, k) pairs (kind from OP_*).  Returns W (nrows, nops, n),
    scale (nrows, 2), status (nrows,).
    """
    X = _pts(X)
    Y = _pts(Y)
    stencils = np.ascontiguousarray(stencils, dtype=np.int32)
    n = stencils.shape[1]
    y2x_a = None if y2x is None else np.ascontiguousarray(y2x, dtype=np.int32)
    kinds = np.array([o[0] for o in ops], np.int32)
    ks = np.array([o[1] for o in ops], np.int32)
    if rows is None:
        nrows = len(Y)
        rows_a = None
    else:
        rows_a = np.ascontiguousarray(rows, dtype=np.int64)
        nrows = len(rows_a)
    W = np.empty((nrows, len(ops), n), np.float64)
    scale = np.empty((nrows, 2), np.float64)
    status = np.empty(nrows, np.int32)
    lib().orc_weights(_p(X), len(X), _p(Y), len(Y), _p(stencils), n, _p(y2x_a),
                      int(m), int(polydeg), int(method), len(ops), _p(kinds), _p(ks),
                      1 if quad else 0, 0 if solve == "inv" else 1,
                      _p(rows_a), nrows, _p(W), _p(scale), _p(status), int(nthreads))
    return W, scale, status


def sparse_csc(rows, cols, vals, shape=None):
    """Restatement of Julia's sparse(I, J, V) (SparseArrays, Julia 1.9): sums
    duplicates, keeps explicit zeros, rows sorted within each column, size
    (max(I), max(J)) unless given.  0-based inputs; returns (colptr, rowval,
    nzval, shape) with 0-based colptr/rowval."""
    rows = np.asarray(rows, np.int64).ravel()
    cols = np.asarray(cols, np.int64).ravel()
    vals = np.asarray(vals, np.float64).ravel()
    if shape is None:
        shape = (int(rows.max()) + 1, int(cols.max()) + 1)
    order = np.lexsort((rows, cols))
    r, c, v = rows[order], cols[order], vals[order]
    key = c * shape[0] + r
    uniq, first = np.unique(key, return_index=True)
    sums = np.add.reduceat(v, first) if len(v) else v
    ur, uc = uniq % shape[0], uniq // shape[0]
    colptr = np.zeros(shape[1] + 1, np.int64)
    np.add.at(colptr, uc + 1, 1)
    colptr = np.cumsum(colptr)
    return colptr, ur.astype(np.int64), sums, shape


def csr_from_rows(stencil_rows, W_rows):
    """CSR with columns sorted inside each row, from per-row stencil order.
    stencil_rows (M, n) int, W_rows (M, n) float -> (rowptr, col, val)."""
    M, n = stencil_rows.shape
    order = np.argsort(stencil_rows, axis=1, kind="stable")
    col = np.take_along_axis(stencil_rows, order, axis=1)
    val = np.take_along_axis(W_rows, order, axis=1)
    rowptr = np.arange(M + 1, dtype=np.int64) * n
    return rowptr, col.reshape(-1).astype(np.int32), val.reshape(-1)


def spmv_csr(rowptr, col, val, x, nthreads=0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    F = 1 if x.ndim == 1 else x.shape[1]
    M = len(rowptr) - 1
    y = np.empty((M, F) if x.ndim == 2 else M, np.float64)
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    lib().orc_spmv_csr(M, _p(rowptr), _p(col), _p(val), F, _p(x), _p(y), int(nthreads))
    return y


def generate_operator(X, Y=None, p=3, n=20, polydeg=3, ops=("E", "Dx", "Dy", "Dxx", "Dyy", "Dxy"),
                      k=0, boundary=None, method=None, quad=False):
    """Full CPU pipeline as the reference runs it; returns dict name -> CSC tuple
    (colptr, rowval, nzval, shape), all 0-based.

    Y None selects method 3 (src/generate_operator.jl:354).  boundary =
    (X_idx_in, X_idx_bc, X_idx_bc_g) selects method 2 (0-based sets)."""
    X = _pts(X)
    if Y is None:
        meth = 3
        Yp = X
    else:
        meth = 1
        Yp = _pts(Y)
    if method is not None:
        meth = method
    if boundary is not None:
        xg, qg = boundary_groups(len(X), *boundary)
        st, _ = knn(X, X, n, xg, qg)
    else:
        st, _ = knn(X, X, n)
    if meth == 3:
        y2x = None
    else:
        y2x, _ = knn(X, Yp, 1)
        y2x = y2x[:, 0]
    op_list = [(OP_NAMES[o], k if o in ("Dxk", "Dyk", "Lapk") else 0) for o in ops]
    W, _, status = weights(X, Yp, st, y2x, p, polydeg, op_list, method=1 if meth == 2 else meth,
                           quad=quad)
    if status.any():
        raise FloatingPointError(f"{int((status != 0).sum())} stencils failed (singular or non-finite)")
    cols = st if y2x is None else st[y2x]
    rows = np.repeat(np.arange(len(Yp)), n).reshape(len(Yp), n)
    out = {}
    for j, name in enumerate(ops):
        out[name] = sparse_csc(rows, cols, W[:, j, :])
    return out

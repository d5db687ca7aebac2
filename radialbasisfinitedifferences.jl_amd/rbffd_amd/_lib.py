"""ctypes binding of librbffd.so (the C ABI declared in include/rbffd.h).

The library is required: importing the operators without a built
``librbffd.so`` raises immediately; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "librbffd.so")

RBFFD_OK = 0
RBFFD_ERR_ARG = 1
RBFFD_ERR_UNSUPPORTED = 2
RBFFD_ERR_RUNTIME = 3
RBFFD_ERR_SINGULAR = 4
RBFFD_ERR_NONFINITE = 5
RBFFD_ERR_UNDEFINED = 6

OP = {"E": 0, "Dx": 1, "Dy": 2, "Dxx": 3, "Dyy": 4, "Dxy": 5, "Dxk": 6, "Dyk": 7, "Lap": 8, "Lapk": 9}
PREC = {"fp64": 0, "mixed": 1, "auto": 2}


class SingularException(ArithmeticError):
    """A local system is exactly singular (LinearAlgebra.SingularException in
    the reference, raised by inv at src/interpolationmatrix.jl:8)."""


class UndefRefError(RuntimeError):
    """A node belongs to no index set, so the reference never fills its
    neighbour list (src/calculateneighbors.jl:8-41) and reads an undefined
    reference."""


class NonFiniteError(FloatingPointError):
    """Non-finite coordinates or weights."""


class RbffdError(RuntimeError):
    pass


class rbffd_boundary(ctypes.Structure):
    _fields_ = [
        ("index_base", ctypes.c_int32),
        ("n_in", ctypes.c_int64),
        ("in_idx", ctypes.c_void_p),
        ("n_markers", ctypes.c_int32),
        ("bc_ptr", ctypes.c_void_p),
        ("bc_idx", ctypes.c_void_p),
        ("ghost_ptr", ctypes.c_void_p),
        ("ghost_idx", ctypes.c_void_p),
    ]


class rbffd_stats(ctypes.Structure):
    _fields_ = [
        ("rows", ctypes.c_int64),
        ("singular", ctypes.c_int64),
        ("nonfinite", ctypes.c_int64),
        ("refined_fallback", ctypes.c_int64),
        ("ms_knn", ctypes.c_double),
        ("ms_solve", ctypes.c_double),
        ("ms_fallback", ctypes.c_double),
        ("lanes_per_stencil", ctypes.c_int64),
        ("factorizations", ctypes.c_int64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class rbffd_ghost_closure(ctypes.Structure):
    _fields_ = [("nbc", ctypes.c_int64), ("rowptr", ctypes.c_void_p), ("col", ctypes.c_void_p),
                ("vals", ctypes.c_void_p), ("ginv", ctypes.c_void_p), ("ghost", ctypes.c_void_p)]


class rbffd_batch_set(ctypes.Structure):
    _fields_ = [
        ("X", ctypes.c_void_p),
        ("N", ctypes.c_int64),
        ("p", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("polydeg", ctypes.c_int32),
        ("xgroup", ctypes.c_void_p),
        ("qgroup", ctypes.c_void_p),
        ("colidx", ctypes.c_void_p),
        ("vals", ctypes.c_void_p * 10),
        ("status", ctypes.c_void_p),
    ]


_V = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double

# name -> (restype, argtypes); every function declared in include/rbffd.h
SIGNATURES = {
    "rbffd_version": (ctypes.c_int, []),
    "rbffd_last_error": (ctypes.c_char_p, []),
    "rbffd_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_V)]),
    "rbffd_destroy": (ctypes.c_int, [_V]),
    "rbffd_set_stream": (ctypes.c_int, [_V, _V]),
    "rbffd_synchronize": (ctypes.c_int, [_V]),
    "rbffd_device_malloc": (ctypes.c_int, [_V, _I64, ctypes.POINTER(_V)]),
    "rbffd_device_free": (ctypes.c_int, [_V, _V]),
    "rbffd_memcpy_h2d": (ctypes.c_int, [_V, _V, _V, _I64]),
    "rbffd_memcpy_d2h": (ctypes.c_int, [_V, _V, _V, _I64]),
    "rbffd_memcpy_d2d": (ctypes.c_int, [_V, _V, _V, _I64]),
    "rbffd_memset_device": (ctypes.c_int, [_V, _V, _I32, _I64]),
    "rbffd_timer_start": (ctypes.c_int, [_V]),
    "rbffd_timer_stop": (ctypes.c_int, [_V, ctypes.POINTER(ctypes.c_float)]),
    "rbffd_device_info": (ctypes.c_int, [_V, ctypes.c_char_p, _I32, ctypes.POINTER(_I32), ctypes.c_char_p, _I32]),
    "rbffd_knn_device": (ctypes.c_int, [_V, _V, _I64, _V, _I64, _I32, _V, _V, _V, _V]),
    "rbffd_weights_device": (ctypes.c_int, [_V, _V, _I64, _V, _I64, _V, _I32, _V, _I32, _I32, _I32, _I32,
                                            _V, _V, _I32, _V, _V, _V]),
    "rbffd_weights_device_noalloc": (ctypes.c_int, [_V, _V, _I64, _V, _I64, _V, _I32, _V, _I32, _I32, _I32, _I32,
                                                    _V, _V, _I32, _V, _V, _V]),
    "rbffd_generate_device": (ctypes.c_int, [_V, _V, _I64, _V, _I64, _I32, _I32, _I32,
                                             ctypes.POINTER(rbffd_boundary), _I32, _V, _V, _I32, _V, _V, _V, _V,
                                             ctypes.POINTER(rbffd_stats)]),
    "rbffd_csr_to_csc_device": (ctypes.c_int, [_V, _I64, _I32, _I64, _V, _I32, _V, _V, _V, _V, _I32,
                                               ctypes.POINTER(_I64)]),
    "rbffd_spmv_device": (ctypes.c_int, [_V, _I64, _V, _V, _V, _I32, _I32, _V, _V, _D, _D, _I32, _V]),
    "rbffd_locality_order_device": (ctypes.c_int, [_V, _V, _I64, _V]),
    "rbffd_spmv_plan_create": (ctypes.c_int, [_V, _I64, _I64, _V, _V, _V, _I32, _I32, _V, ctypes.POINTER(_V), _V]),
    "rbffd_spmv_plan_apply": (ctypes.c_int, [_V, _V, _I32, _V, _V, _D, _D]),
    "rbffd_spmv_plan_destroy": (ctypes.c_int, [_V]),
    "rbffd_ssprk3_plan_device": (ctypes.c_int, [_V, _V, _I32, _V, _V, _V, _D, _I64]),
    "rbffd_read_sum_device": (ctypes.c_int, [_V, _V, _I64, _V]),
    "rbffd_lsq_solve_device": (ctypes.c_int, [_V, _I64, _I64, _V, _V, _V, _V, _V, _I32, _D, _I64,
                                              ctypes.POINTER(_I64), ctypes.POINTER(_D)]),
    "rbffd_combine_device": (ctypes.c_int, [_V, _I64, _I32, _V, _V, _V, _V]),
    "rbffd_ssprk3_device": (ctypes.c_int, [_V, _I64, _V, _V, _V, _I32, _I32, _V, _I32, _V, _V, _V, _D, _I64]),
    "rbffd_ssprk3_closure_device": (ctypes.c_int, [_V, _I64, _V, _V, _V, _I32, _I32, _V, _I32, _V, _V, _V, _I32,
                                                   ctypes.POINTER(rbffd_ghost_closure), _D, _I64]),
    "rbffd_dense_inverse_device": (ctypes.c_int, [_V, _I64, _V, _V]),
    "rbffd_boundary_groups": (ctypes.c_int, [ctypes.POINTER(rbffd_boundary), _I64, _V, _V]),
    "rbffd_generate_batch_device": (ctypes.c_int, [_V, _I32, ctypes.POINTER(rbffd_batch_set), _I32, _V, _V, _I32,
                                                   _V, ctypes.POINTER(rbffd_stats)]),
    "rbffd_generate_operator": (ctypes.c_int, [_V, _V, _I64, _V, _I64, _I32, _I32, _I32,
                                               ctypes.POINTER(rbffd_boundary), _I32, _V, _V, _I32, _I32, _V, _V,
                                               _V, ctypes.POINTER(_I64), ctypes.POINTER(rbffd_stats)]),
}

_lib = None


def load():
    """Load librbffd.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it first (python "
                          "radialbasisfinitedifferences.jl_amd/build_lib.py); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(code: int):
    if code == RBFFD_OK:
        return
    msg = load().rbffd_last_error().decode(errors="replace")
    if code == RBFFD_ERR_ARG:
        raise ValueError(msg)
    if code == RBFFD_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if code == RBFFD_ERR_SINGULAR:
        raise SingularException(msg)
    if code == RBFFD_ERR_NONFINITE:
        raise NonFiniteError(msg)
    if code == RBFFD_ERR_UNDEFINED:
        raise UndefRefError(msg)
    raise RbffdError(f"librbffd error {code}: {msg}")

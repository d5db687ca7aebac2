"""The C-ABI library loads and exports every entry point include/rbffd.h
declares (no GPU needed: nothing here launches a kernel)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def declared():
    text = open(os.path.join(ROOT, "include", "rbffd.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(rbffd_\w+)\(", text, flags=re.M)))


def test_header_declares_the_expected_surface():
    names = declared()
    for required in ("rbffd_knn_device", "rbffd_weights_device", "rbffd_generate_device",
                     "rbffd_csr_to_csc_device", "rbffd_spmv_device", "rbffd_generate_operator",
                     "rbffd_create", "rbffd_destroy", "rbffd_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from rbffd_amd import _lib

    L = _lib.load()
    for name in declared():
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert set(_lib.SIGNATURES) == set(declared())
    assert L.rbffd_version() == 1


def test_library_links_only_rocm_runtime():
    # the product library must not pull in the oracle
    from rbffd_amd import _lib

    data = open(_lib.LIB_PATH, "rb").read()
    assert b"liboracle" not in data
    assert b"orc_weights" not in data


def test_no_device_reports_an_error_not_a_crash():
    from rbffd_amd import _lib

    L = _lib.load()
    n = ctypes.c_int()
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be present")
    h = ctypes.c_void_p()
    code = L.rbffd_create(0, ctypes.byref(h))
    assert code != 0
    assert L.rbffd_last_error()

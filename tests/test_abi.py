"""The C ABI libraries load, export every symbol their headers in include/
declare, and the host-side argument checks behave -- without touching a GPU."""
import ctypes as C
import os
import re

import numpy as np

from benchmark_spmv_using_csr_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header="bhsparse_spmv.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(bhs_\w+)\s*\(", txt)))


def test_every_header_is_covered():
    assert sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")) == \
        ["bhs_mmio.h", "bhs_vendor.h", "bhsparse_spmv.h"]


def test_host_library_exports_the_reader():
    L = _lib.host_lib()
    names = _declared("bhs_mmio.h")
    assert names == ["bhs_csr_free", "bhs_mm_read"]
    for name in names:
        assert hasattr(L, name), name


def test_header_and_binding_agree():
    assert _declared() == sorted(_lib.exported_names())


def test_library_exports_every_declared_symbol():
    L = _lib.spmv_lib()
    for name in _declared():
        assert hasattr(L, name), name


def test_vendor_library_exports_every_declared_symbol():
    from benchmark_spmv_using_csr_amd import vendor
    L = vendor.vendor_lib()
    names = _declared("bhs_vendor.h")
    assert names == ["bhs_vendor_d_csrmv_bench", "bhs_vendor_s_csrmv_bench", "bhs_vendor_version"]
    for name in names:
        assert hasattr(L, name), name
    assert vendor.version().startswith("4.")


def test_error_strings():
    L = _lib.spmv_lib()
    assert L.bhs_error_string(0) == b"success"
    assert L.bhs_error_string(-4) == b"invalid matrix"
    assert L.bhs_error_string(-99) == b"unknown error"


def test_host_side_argument_checks():
    L = _lib.spmv_lib()
    h = C.c_void_p()
    assert L.bhs_d_create(C.byref(h)) == 0
    # wrong precision for the handle
    assert L.bhs_s_spmv(h, C.c_float(1.0)) == _lib.BHSPARSE_UNSUPPORTED_VALUE_TYPE
    # calls before init_platform / init_data
    rp = np.array([0, 1], np.int32)
    ci = np.array([0], np.int32)
    v = np.ones(1)
    assert L.bhs_d_init_data(h, 1, 1, 1, rp, ci, v, v, None) == _lib.BHSPARSE_NOT_INITIALIZED
    assert L.bhs_d_spmv(h, 1.0) == _lib.BHSPARSE_NOT_INITIALIZED
    bad = _lib.Config(7, 8, 4, 3, 0)
    assert L.bhs_d_set_config(h, C.byref(bad)) == _lib.BHSPARSE_INVALID_ARGUMENT
    bad = _lib.Config(0, 6, 4, 3, 0)
    assert L.bhs_d_set_config(h, C.byref(bad)) == _lib.BHSPARSE_INVALID_ARGUMENT
    bad = _lib.Config(2, 8, 4, 3, 3)        # column panels must be 2, 4 or 8
    assert L.bhs_d_set_config(h, C.byref(bad)) == _lib.BHSPARSE_INVALID_ARGUMENT
    assert L.bhs_d_set_config(h, C.byref(_lib.Config(2, 8, 4, 3, 4))) == 0
    good = _lib.Config(0, 16, 2, 5, 0)
    assert L.bhs_d_set_config(h, C.byref(good)) == 0
    out = _lib.Config()
    assert L.bhs_d_get_config(h, C.byref(out)) == 0 and out.as_tuple() == (0, 16, 2, 5, 0)
    # memory mode (USE_SVM_ALWAYS, spmv_cuda/Makefile:6): 0 / 1 / 2, device by default
    md = C.c_int(-1)
    assert L.bhs_d_get_memory_mode(h, C.byref(md)) == 0 and md.value == 0
    assert L.bhs_d_set_memory_mode(h, 3) == _lib.BHSPARSE_INVALID_ARGUMENT
    assert L.bhs_d_set_memory_mode(h, -1) == _lib.BHSPARSE_INVALID_ARGUMENT
    assert L.bhs_d_set_memory_mode(h, 2) == 0
    assert L.bhs_d_get_memory_mode(h, C.byref(md)) == 0 and md.value == 2
    assert L.bhs_d_set_memory_mode(h, 1) == 0
    assert L.bhs_d_get_memory_mode(h, C.byref(md)) == 0 and md.value == 1
    assert L.bhs_d_get_memory_mode(h, None) == _lib.BHSPARSE_INVALID_ARGUMENT
    assert L.bhs_s_set_memory_mode(h, 1) == _lib.BHSPARSE_UNSUPPORTED_VALUE_TYPE
    assert L.bhs_d_set_memory_mode(h, 0) == 0
    # the stream exists only after init_platform
    st = C.c_void_p()
    assert L.bhs_d_get_stream(h, C.byref(st)) == _lib.BHSPARSE_NOT_INITIALIZED
    assert L.bhs_d_get_stream(h, None) == _lib.BHSPARSE_INVALID_ARGUMENT
    assert L.bhs_d_free_mem(h) == 0 and L.bhs_d_free_mem(h) == 0     # twice is fine
    assert L.bhs_d_destroy(h) == 0
    assert L.bhs_d_spmv(None, 1.0) == _lib.BHSPARSE_INVALID_ARGUMENT

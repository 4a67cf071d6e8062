"""rocSPARSE CSR SpMV on the same matrix (libbhs_vendor.so, include/bhs_vendor.h):
the comparison column the reference fills with cuSPARSE (call_cusparse_ref,
spmv_cuda/main.cu:49-183).  Not on the bhsparse_spmv path."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import PKG, check

VENDOR_LIB = os.path.join(PKG, "libbhs_vendor.so")
ALGS = {"adaptive": 0, "rowsplit": 1, "lrb": 2, "nnzsplit": 3}
_vendor = None


def vendor_lib():
    global _vendor
    if _vendor is None:
        if not os.path.exists(VENDOR_LIB):
            raise FileNotFoundError(f"{VENDOR_LIB} is missing: run __graft_entry__.build() first")
        L = C.CDLL(VENDOR_LIB)
        for p, vt in (("d", C.c_double), ("s", C.c_float)):
            f = getattr(L, f"bhs_vendor_{p}_csrmv_bench")
            f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                          C.c_void_p, vt, C.c_int, C.c_int, C.c_int, C.c_longlong, C.c_void_p,
                          C.c_void_p, C.c_void_p]
            f.restype = C.c_int
        L.bhs_vendor_version.restype = C.c_int
        _vendor = L
    return _vendor


def version() -> str:
    v = vendor_lib().bhs_vendor_version()
    return f"{v // 100000}.{(v // 100) % 1000}.{v % 100}"


def csrmv(A, x, dtype=np.float64, alg="adaptive", alpha=1.0, warmup=3, reps=10,
          flush_bytes=0, device=0):
    """(per-call ms, y, analysis ms) of rocSPARSE y = alpha*A*x on the GPU."""
    dtype = np.dtype(dtype)
    p = "d" if dtype == np.float64 else "s"
    rp = np.ascontiguousarray(A.row_ptr, np.int32)
    col = np.ascontiguousarray(A.col, np.int32)
    val = np.ascontiguousarray(A.val, dtype)
    xx = np.ascontiguousarray(x, dtype)
    y = np.zeros(A.m, dtype)
    ms = np.zeros(max(reps, 1), np.float64)
    an = C.c_double(0)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    vt = C.c_double if p == "d" else C.c_float
    check(getattr(vendor_lib(), f"bhs_vendor_{p}_csrmv_bench")(
        device, A.m, A.n, int(A.nnz), vp(rp), vp(col), vp(val), vp(xx), vt(alpha), ALGS[alg],
        warmup, reps, flush_bytes, vp(ms), vp(y), C.byref(an)), f"rocsparse csrmv ({alg})")
    return ms[:reps], y, an.value

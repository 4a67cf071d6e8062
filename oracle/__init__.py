"""CPU oracle for the speculative segmented-sum CSR SpMV.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or the timed CPU baseline).  The shipped library never uses it.

The numerical routines live in ``oracle.c`` (built by ``oracle/Makefile`` into
``liboracle.so``); each cites the reference file:line it restates
(weifengliu-ssslab/Benchmark_SpMV_using_CSR).  This module is a thin ctypes
wrapper over numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


class TileStats(C.Structure):
    _fields_ = [(n, C.c_longlong) for n in
                ("n_tiles", "n_waves", "n_fast", "n_dirty", "n_sync", "n_lead_empty")]


class RefStats(C.Structure):
    _fields_ = [(n, C.c_longlong) for n in ("n_tiles", "n_bunches", "n_fast", "n_dirty")]


def build() -> None:
    """Compile liboracle.so from oracle.c (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.oracle_spmv_serial_f64.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.oracle_spmv_serial_f32.argtypes = [C.c_int, _i32p, _i32p, _f32p, _f32p, _f32p]
        L.oracle_spmv_omp_f64.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int]
        L.oracle_spmv_omp_f64.restype = C.c_int
        L.oracle_spmv_omp_f32.argtypes = [C.c_int, _i32p, _i32p, _f32p, _f32p, _f32p, C.c_int]
        L.oracle_spmv_omp_f32.restype = C.c_int
        L.oracle_time_spmv_f64.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p,
                                           C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.oracle_time_spmv_f64.restype = C.c_double
        L.oracle_time_spmv_f32.argtypes = [C.c_int, _i32p, _i32p, _f32p, _f32p, _f32p,
                                           C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.oracle_time_spmv_f32.restype = C.c_double
        L.oracle_upper_bound.argtypes = [_i32p, C.c_int, C.c_int]
        L.oracle_upper_bound.restype = C.c_int
        L.oracle_partition.argtypes = [C.c_int, C.c_int, _i32p, C.c_int, _i32p]
        L.oracle_tile_stats_compute.argtypes = [C.c_int, C.c_int, _i32p, C.c_int, C.c_int,
                                                C.POINTER(TileStats)]
        L.oracle_segsum_ref.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i32p,
                                        _i32p, _f64p, _f64p, _f64p, C.POINTER(RefStats)]
        L.oracle_segsum_ref.restype = C.c_int
        L.oracle_rel_error_f64.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                           C.c_void_p]
        L.oracle_rel_error_f64.restype = C.c_double
        L.oracle_mlp_loss_grad.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, _f64p, _f64p, _f64p,
                                           _f64p, C.c_int, C.c_void_p, C.c_void_p]
        L.oracle_mlp_loss_grad.restype = C.c_double
        L.oracle_set_mutation.argtypes = [C.c_int]
        L.oracle_set_mutation.restype = None
        _lib = L
    return _lib


# Planted mistakes (oracle.h ORACLE_MUT_*), for tests only.
MUTATIONS = {
    "serial_drop_last": 1, "serial_sign": 2, "serial_f32_acc": 3, "serial_transpose": 4,
    "lower_bound": 5, "no_carry": 6, "no_span": 7, "no_calibration": 8, "no_fixup": 9,
}
# planted mistakes of the tuner's MLP backprop (caught by the finite-difference pin)
MLP_MUTATIONS = {"mlp_no_mask": 10, "mlp_relu_grad": 11, "mlp_transpose": 12}


class mutated:
    """Context manager: run the oracle with one planted mistake, then restore it."""

    def __init__(self, name):
        self.kind = MUTATIONS[name] if name in MUTATIONS else MLP_MUTATIONS[name]

    def __enter__(self):
        lib().oracle_set_mutation(self.kind)
        return self

    def __exit__(self, *exc):
        lib().oracle_set_mutation(0)
        return False


def _csr(row_ptr, col, val, dtype=np.float64):
    return (np.ascontiguousarray(row_ptr, np.int32), np.ascontiguousarray(col, np.int32),
            np.ascontiguousarray(val, dtype))


def spmv_serial(row_ptr, col, val, x):
    """FP64 serial CSR loop (spmv_cuda/main.cu:755-761) -- the parity oracle."""
    rp, c, v = _csr(row_ptr, col, val)
    x = np.ascontiguousarray(x, np.float64)
    m = len(rp) - 1
    y = np.empty(m, np.float64)
    lib().oracle_spmv_serial_f64(m, rp, c, v, x, y)
    return y


def spmv_serial_f32(row_ptr, col, val, x):
    """The same loop with value_type = float (USE_DOUBLE=0 build)."""
    rp, c, v = _csr(row_ptr, col, val, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    m = len(rp) - 1
    y = np.empty(m, np.float32)
    lib().oracle_spmv_serial_f32(m, rp, c, v, x, y)
    return y


def spmv_omp(row_ptr, col, val, x, nthreads=0):
    """OpenMP row-parallel loop (spmv_cuda/main.cu:377-393); returns (y, threads)."""
    dt = np.float32 if np.asarray(val).dtype == np.float32 else np.float64
    rp, c, v = _csr(row_ptr, col, val, dt)
    x = np.ascontiguousarray(x, dt)
    m = len(rp) - 1
    y = np.empty(m, dt)
    f = lib().oracle_spmv_omp_f32 if dt == np.float32 else lib().oracle_spmv_omp_f64
    t = f(m, rp, c, v, x, y, int(nthreads))
    return y, t


def time_spmv(row_ptr, col, val, x, nthreads, reps):
    """Median seconds per call of the serial (nthreads=1) or OpenMP loop, in
    the arrays' own precision; returns (seconds, threads_used)."""
    dt = np.float32 if np.asarray(val).dtype == np.float32 else np.float64
    rp, c, v = _csr(row_ptr, col, val, dt)
    x = np.ascontiguousarray(x, dt)
    m = len(rp) - 1
    y = np.empty(m, dt)
    used = C.c_int(0)
    f = lib().oracle_time_spmv_f32 if dt == np.float32 else lib().oracle_time_spmv_f64
    s = f(m, rp, c, v, x, y, int(nthreads), int(reps), C.byref(used))
    return s, used.value


def upper_bound(row_ptr, key):
    rp = np.ascontiguousarray(row_ptr, np.int32)
    return lib().oracle_upper_bound(rp, len(rp), int(key))


def partition(row_ptr, nnz, P):
    """tile_row[p] = rowOf(p*P), p in [0, ceil(nnz/P)] (kernels_cuda.h:39-49)."""
    rp = np.ascontiguousarray(row_ptr, np.int32)
    NP = (int(nnz) + P - 1) // P
    out = np.empty(NP + 1, np.int32)
    lib().oracle_partition(len(rp) - 1, int(nnz), rp, int(P), out)
    return out


def tile_stats(row_ptr, nnz, P, T):
    rp = np.ascontiguousarray(row_ptr, np.int32)
    st = TileStats()
    lib().oracle_tile_stats_compute(len(rp) - 1, int(nnz), rp, int(P), int(T), C.byref(st))
    return {n: getattr(st, n) for n, _ in TileStats._fields_}


def segsum_ref(row_ptr, col, val, x, W=64, S=8, T=5):
    """The reference's speculative segmented sum restated on the CPU (FP64).
    Returns (y, stats)."""
    rp, c, v = _csr(row_ptr, col, val)
    x = np.ascontiguousarray(x, np.float64)
    m = len(rp) - 1
    y = np.empty(m, np.float64)
    st = RefStats()
    rc = lib().oracle_segsum_ref(int(W), int(S), int(T), m, int(rp[-1]) if m >= 0 else 0,
                                 rp, c, v, x, y, C.byref(st))
    if rc != 0:
        raise ValueError("oracle_segsum_ref: bad arguments")
    return y, {n: getattr(st, n) for n, _ in RefStats._fields_}


def rel_error(row_ptr, col, val, x, y, y_ref, per_row=False):
    """max_i |y_i - y*_i| / sum_j |a_ij x_j| in FP64 (0/0 -> 0)."""
    rp, c, v = _csr(row_ptr, col, np.asarray(val, np.float64))
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    yr = np.ascontiguousarray(y_ref, np.float64)
    m = len(rp) - 1
    err = np.empty(max(m, 1), np.float64) if per_row else None
    e = lib().oracle_rel_error_f64(m, rp, c, v, x, y, yr,
                                   err.ctypes.data_as(C.c_void_p) if per_row else None)
    return (e, err[:m]) if per_row else e


def mlp_loss_grad(n_in, h, n_out, params, norm, X, T, M):
    """FP64 MLP forward + masked MSE + backprop; returns (loss, outputs, grad)."""
    params = np.ascontiguousarray(params, np.float64)
    norm = np.ascontiguousarray(norm, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    T = np.ascontiguousarray(T, np.float64)
    M = np.ascontiguousarray(M, np.float64)
    N = X.shape[0]
    out = np.empty((N, n_out), np.float64)
    grad = np.empty(len(params), np.float64)
    loss = lib().oracle_mlp_loss_grad(n_in, h, n_out, params, norm, X.ravel(), T.ravel(),
                                      M.ravel(), N, out.ctypes.data_as(C.c_void_p),
                                      grad.ctypes.data_as(C.c_void_p))
    return loss, out, grad

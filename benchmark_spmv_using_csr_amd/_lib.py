"""ctypes bindings to the in-tree native libraries.

``libbhsparse_spmv.so`` (HIP, gfx950) holds every kernel and the C ABI of
``include/bhsparse_spmv.h``; ``libbhs_host.so`` holds the seeded generators.
Loading fails loudly when a library is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
SPMV_LIB = os.path.join(PKG, "libbhsparse_spmv.so")
HOST_LIB = os.path.join(PKG, "libbhs_host.so")

BHSPARSE_SUCCESS = 0
BHSPARSE_UNSUPPORTED_VALUE_TYPE = -1
BHSPARSE_UNSUPPORTED_DEVICE = -2
BHSPARSE_INVALID_ARGUMENT = -3
BHSPARSE_INVALID_MATRIX = -4
BHSPARSE_OUT_OF_MEMORY = -5
BHSPARSE_HIP_ERROR = -6
BHSPARSE_NOT_INITIALIZED = -7

BHS_VARIANT_SEGSUM = 0
BHS_VARIANT_CSRVEC = 1


class BhsError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        super().__init__(f"{where} failed with {code}" + (f": {detail}" if detail else ""))


class Config(C.Structure):
    _fields_ = [("variant", C.c_int), ("items_per_lane", C.c_int),
                ("waves_per_block", C.c_int), ("tiles_per_wave", C.c_int), ("tpv", C.c_int)]

    def as_tuple(self):
        return (self.variant, self.items_per_lane, self.waves_per_block, self.tiles_per_wave,
                self.tpv)


class SpmvStats(C.Structure):
    _fields_ = [(n, C.c_longlong) for n in
                ("n_tiles", "n_waves", "n_fast", "n_dirty", "n_sync", "n_lead_empty",
                 "chunk_rows")]


class BatchSchedule(C.Structure):
    _fields_ = [("items_per_lane", C.c_int), ("tiles_per_wave", C.c_int),
                ("waves_per_block", C.c_int), ("grid_blocks", C.c_int), ("order", C.c_int),
                ("x_stage", C.c_int)]

    def as_tuple(self):
        return (self.items_per_lane, self.tiles_per_wave, self.waves_per_block,
                self.grid_blocks, self.order, self.x_stage)


class Csr(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("nnz", C.c_int),
                ("row_ptr", C.POINTER(C.c_int)), ("col", C.POINTER(C.c_int)),
                ("val", C.POINTER(C.c_double))]


_spmv = None
_host = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_H = C.c_void_p

# (name suffix, argtypes after the handle) for both precisions; T marks the
# value pointer type, V the value scalar.
_PER_TYPE = {
    "create": [C.POINTER(C.c_void_p)],
    "init_platform": [_H, C.c_int],
    "init_data": [_H, C.c_int, C.c_int, C.c_int, _i32p, _i32p, "T", "T", C.c_void_p],
    "set_x": [_H, "T"],
    "set_memory_mode": [_H, C.c_int],
    "get_memory_mode": [_H, C.POINTER(C.c_int)],
    "spmv": [_H, "V"],
    "spmv_async": [_H, "V"],
    "get_stream": [_H, C.POINTER(C.c_void_p)],
    "get_y": [_H],
    "get_y_to": [_H, "T"],
    "free_mem": [_H],
    "free_platform": [_H],
    "destroy": [_H],
    "set_config": [_H, C.POINTER(Config)],
    "get_config": [_H, C.POINTER(Config)],
    "get_partition": [_H, C.c_void_p, C.c_int, C.POINTER(C.c_int)],
    "get_stats": [_H, C.POINTER(SpmvStats)],
    "bench": [_H, "V", C.c_int, C.c_int, C.c_longlong, C.c_void_p, C.c_void_p],
    "row_features": [_H, C.c_void_p],
    "bench_floor": [_H, C.c_int, C.c_longlong, C.c_void_p],
    "bench_event_overhead": [_H, C.c_int, C.c_longlong, C.c_void_p],
    "autotune": [_H, C.POINTER(Config), C.c_int, C.c_int, C.c_longlong, C.c_void_p,
                 C.POINTER(C.c_int)],
    "apply_tuner": [_H, C.c_void_p, C.POINTER(Config), C.c_int, C.POINTER(C.c_int)],
    "apply_tuner_topk": [_H, C.c_void_p, C.POINTER(Config), C.c_int, C.c_int, C.c_int,
                         C.c_longlong, C.POINTER(C.c_int)],
}

_BATCH = {
    "batch_create": ([C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], "int"),
    "batch_set_x": ([C.c_void_p, C.c_void_p], "int"),
    "batch_set_schedule": ([C.c_void_p, C.POINTER(BatchSchedule)], "int"),
    "batch_get_schedule": ([C.c_void_p, C.POINTER(BatchSchedule)], "int"),
    "batch_run": ([C.c_void_p, "V"], "int"),
    "batch_run_plain": ([C.c_void_p, "V"], "int"),
    "batch_get_y": ([C.c_void_p, C.c_void_p, C.c_int], "int"),
    "batch_bench": ([C.c_void_p, "V", C.c_int, C.c_int, C.c_int, C.c_void_p], "int"),
    "batch_autotune": ([C.c_void_p, C.POINTER(BatchSchedule), C.c_int, C.c_int, C.c_void_p,
                        C.POINTER(C.c_int)], "int"),
    "batch_destroy": ([C.c_void_p], "int"),
}

_MLP = {
    "bhs_tuner_transform": ([C.c_void_p, C.c_void_p], C.c_int),
    "bhs_mlp_param_count": ([C.c_int, C.c_int, C.c_int], C.c_longlong),
    "bhs_mlp_create": ([C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    "bhs_mlp_destroy": ([C.c_void_p], C.c_int),
    "bhs_mlp_set_params": ([C.c_void_p, C.c_void_p, C.c_longlong, C.c_void_p, C.c_void_p], C.c_int),
    "bhs_mlp_get_params": ([C.c_void_p, C.c_void_p, C.c_longlong], C.c_int),
    "bhs_mlp_forward": ([C.c_void_p, C.c_void_p, C.c_int, C.c_void_p], C.c_int),
    "bhs_mlp_loss_grad": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                           C.c_void_p], C.c_int),
    "bhs_mlp_train": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_float,
                       C.c_float, C.c_void_p], C.c_int),
    "bhs_mlp_loss_grad_f64": ([C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                               C.c_void_p], C.c_int),
}


def exported_names():
    """Every symbol include/bhsparse_spmv.h declares."""
    names = ["bhs_error_string", "bhs_last_error", "bhs_device_synchronize", "bhs_device_count",
             "bhs_stream_copy"] + list(_MLP)
    for p in ("s", "d"):
        names += [f"bhs_{p}_{k}" for k in _PER_TYPE]
        names += [f"bhs_{p}_{k}" for k in _BATCH]
        names.append(f"bhs_{p}_spmv_noalloc")
    return names


def spmv_lib():
    """The HIP library; raises if it has not been built."""
    global _spmv
    if _spmv is not None:
        return _spmv
    if not os.path.exists(SPMV_LIB):
        raise FileNotFoundError(
            f"{SPMV_LIB} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
    L = C.CDLL(SPMV_LIB)
    for p, tp, vt in (("s", _f32p, C.c_float), ("d", _f64p, C.c_double)):
        for k, args in _PER_TYPE.items():
            f = getattr(L, f"bhs_{p}_{k}")
            f.argtypes = [tp if a == "T" else vt if a == "V" else a for a in args]
            f.restype = C.c_int
    L.bhs_error_string.argtypes = [C.c_int]
    L.bhs_error_string.restype = C.c_char_p
    L.bhs_last_error.argtypes = []
    L.bhs_last_error.restype = C.c_char_p
    L.bhs_device_synchronize.restype = C.c_int
    L.bhs_device_count.restype = C.c_int
    L.bhs_stream_copy.argtypes = [C.c_int, C.c_longlong, C.c_int, C.c_void_p]
    L.bhs_stream_copy.restype = C.c_int
    for p, vt in (("s", C.c_float), ("d", C.c_double)):
        for k, (args, _) in _BATCH.items():
            f = getattr(L, f"bhs_{p}_{k}")
            f.argtypes = [vt if a == "V" else a for a in args]
            f.restype = C.c_int
    for name, (args, res) in _MLP.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    for p, vt in (("s", C.c_float), ("d", C.c_double)):
        f = getattr(L, f"bhs_{p}_spmv_noalloc")   # device pointers, alpha, stream
        f.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 5 + [vt, C.c_void_p]
        f.restype = C.c_int
    _spmv = L
    return L


def host_lib():
    global _host
    if _host is not None:
        return _host
    if not os.path.exists(HOST_LIB):
        raise FileNotFoundError(f"{HOST_LIB} is missing: run __graft_entry__.build() first")
    L = C.CDLL(HOST_LIB)
    P = C.POINTER(Csr)
    L.bhs_csr_free.argtypes = [P]
    L.bhs_gen_vector.argtypes = [C.c_ulonglong, C.c_int, _f64p]
    L.bhs_gen_uniform_rows.argtypes = [C.c_ulonglong, C.c_int, C.c_int, C.c_int, C.c_int]
    L.bhs_gen_uniform_rows.restype = P
    L.bhs_gen_power_law.argtypes = [C.c_ulonglong, C.c_int, C.c_int, C.c_double, C.c_int,
                                    C.c_double, C.c_longlong]
    L.bhs_gen_power_law.restype = P
    L.bhs_gen_laplacian7.argtypes = [C.c_int, C.c_int, C.c_int]
    L.bhs_gen_laplacian7.restype = P
    L.bhs_gen_block3.argtypes = [C.c_ulonglong, C.c_int, C.c_int, C.c_int]
    L.bhs_gen_block3.restype = P
    L.bhs_gen_banded.argtypes = [C.c_ulonglong, C.c_int, C.c_int, C.c_int]
    L.bhs_gen_banded.restype = P
    L.bhs_mm_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(P)]
    L.bhs_mm_read.restype = C.c_int
    _host = L
    return L


def check(rc: int, where: str) -> None:
    if rc != BHSPARSE_SUCCESS:
        L = spmv_lib()
        raise BhsError(rc, where, (L.bhs_last_error() or b"").decode(errors="replace"))

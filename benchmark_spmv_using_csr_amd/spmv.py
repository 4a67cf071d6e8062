"""Python mirror of the reference's operator class.

``bhsparse_spmv`` keeps the method names and meaning of ``bhsparse_spmv_cuda``
(spmv_cuda/bhsparse_spmv_cuda.h:7-55): ``init_platform`` / ``prepare_mem`` /
``run_benchmark`` / ``get_y`` / ``free_platform`` / ``free_mem``, plus the
camel-case names ``initPlatform`` / ``initData`` / ``spmv``.  Every call goes
through the C ABI of ``libbhsparse_spmv.so`` (gfx950 HIP kernels).  Integer
return codes are kept (0 = BHSPARSE_SUCCESS) as in the reference; ``check``
turns them into exceptions where a caller prefers that.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Config, SpmvStats, check, spmv_lib


MEMORY_DEVICE = 0        # BHS_MEMORY_DEVICE
MEMORY_MANAGED = 1       # BHS_MEMORY_MANAGED (the reference's USE_SVM_ALWAYS=1 build)
MEMORY_HOST_MAPPED = 2   # BHS_MEMORY_HOST_MAPPED (pinned host memory mapped into the device)


class bhsparse_spmv:  # noqa: N801  (reference class name)
    """y = alpha * A * x on one GPU.  dtype selects the fp32 (bhs_s_*) or fp64
    (bhs_d_*) build -- the reference's USE_DOUBLE switch (common.h:27-31)."""

    def __init__(self, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        if self.dtype == np.float64:
            self._p = "d"
        elif self.dtype == np.float32:
            self._p = "s"
        else:
            raise TypeError("value type must be float32 or float64")
        self._L = spmv_lib()
        self._h = C.c_void_p()
        check(self._fn("create")(C.byref(self._h)), "create")
        self._keep = None      # host arrays the library borrows (y)
        self.m = self.n = self.nnz = 0

    def _fn(self, name):
        return getattr(self._L, f"bhs_{self._p}_{name}")

    # -- reference surface ----------------------------------------------
    def init_platform(self, device: int = 0) -> int:
        return self._fn("init_platform")(self._h, int(device))

    def prepare_mem(self, m, n, nnz, row_ptr, col, val, x, y) -> int:
        rp = np.ascontiguousarray(row_ptr, np.int32)
        ci = np.ascontiguousarray(col, np.int32)
        v = np.ascontiguousarray(val, self.dtype)
        xx = np.ascontiguousarray(x, self.dtype)
        # the library reads row_ptr[0..m], col/val[0..nnz) and x[0..n) from
        # these buffers: shorter ones would be read past their end
        m, n, nnz = int(m), int(n), int(nnz)
        if m < 0 or n < 0 or nnz < 0:
            raise ValueError("m, n and nnz must be non-negative")
        if len(rp) < m + 1 or len(ci) < nnz or len(v) < nnz or len(xx) < n:
            raise ValueError(f"arrays shorter than the sizes: len(row_ptr)={len(rp)} < m+1="
                             f"{m + 1}, or len(col)={len(ci)} / len(val)={len(v)} < nnz={nnz}, "
                             f"or len(x)={len(xx)} < n={n}")
        if len(ci) == 0:
            ci = np.zeros(1, np.int32)
            v = np.zeros(1, self.dtype)
        if len(xx) == 0:
            xx = np.zeros(1, self.dtype)
        if not (isinstance(y, np.ndarray) and y.dtype == self.dtype and y.flags.c_contiguous
                and y.size >= m):
            raise TypeError("y must be a C-contiguous array of the handle's dtype with m entries")
        self._keep = y
        self.m, self.n, self.nnz = int(m), int(n), int(nnz)
        return self._fn("init_data")(self._h, int(m), int(n), int(nnz), rp, ci, v, xx,
                                     y.ctypes.data_as(C.c_void_p))

    def run_benchmark(self) -> int:
        return self.spmv(1.0)

    def get_y(self) -> int:
        return self._fn("get_y")(self._h)

    def free_platform(self) -> int:
        return self._fn("free_platform")(self._h)

    def free_mem(self) -> int:
        rc = self._fn("free_mem")(self._h)
        self._keep = None
        return rc

    # -- camel-case names ---------------------------------------------------
    def initPlatform(self, device: int = 0) -> int:  # noqa: N802
        return self.init_platform(device)

    def initData(self, m, n, nnz, row_ptr, col, val, x, y) -> int:  # noqa: N802
        return self.prepare_mem(m, n, nnz, row_ptr, col, val, x, y)

    def spmv(self, alpha=1.0) -> int:
        return self._fn("spmv")(self._h, float(alpha))

    # -- extras -----------------------------------------------------------
    def spmv_async(self, alpha=1.0) -> int:
        return self._fn("spmv_async")(self._h, float(alpha))

    def set_x(self, x) -> int:
        xx = np.ascontiguousarray(x, self.dtype)
        if len(xx) < self.n:
            raise ValueError(f"x has {len(xx)} entries, the matrix has n={self.n} columns")
        if len(xx) == 0:
            xx = np.zeros(1, self.dtype)
        return self._fn("set_x")(self._h, xx)

    def set_memory_mode(self, mode) -> int:
        """MEMORY_DEVICE, MEMORY_MANAGED or MEMORY_HOST_MAPPED for the next
        prepare_mem (spmv_cuda/Makefile:6, bhsparse_spmv_cuda.h:250-263)."""
        return self._fn("set_memory_mode")(self._h, int(mode))

    def memory_mode(self) -> int:
        md = C.c_int(-1)
        check(self._fn("get_memory_mode")(self._h, C.byref(md)), "get_memory_mode")
        return md.value

    def y(self) -> np.ndarray:
        out = np.empty(max(self.m, 1), self.dtype)
        check(self._fn("get_y_to")(self._h, out), "get_y_to")
        return out[: self.m]

    def set_config(self, variant=0, items_per_lane=8, waves_per_block=4, tiles_per_wave=3,
                   tpv=0) -> int:
        c = Config(int(variant), int(items_per_lane), int(waves_per_block), int(tiles_per_wave),
                   int(tpv))
        return self._fn("set_config")(self._h, C.byref(c))

    def get_config(self) -> Config:
        c = Config()
        check(self._fn("get_config")(self._h, C.byref(c)), "get_config")
        return c

    def partition(self) -> np.ndarray:
        np_ = C.c_int(0)
        check(self._fn("get_partition")(self._h, None, 0, C.byref(np_)), "get_partition")
        out = np.empty(np_.value + 1, np.int32)
        check(self._fn("get_partition")(self._h, out.ctypes.data_as(C.c_void_p), len(out),
                                        C.byref(np_)), "get_partition")
        return out

    def stats(self) -> dict:
        s = SpmvStats()
        check(self._fn("get_stats")(self._h, C.byref(s)), "get_stats")
        return {n: getattr(s, n) for n, _ in SpmvStats._fields_}

    def bench(self, alpha=1.0, warmup=10, reps=50, flush_bytes=0):
        """Per-rep (total_ms, main_kernel_ms) arrays, HIP-event timed on the
        handle's stream."""
        t = np.zeros(max(reps, 1), np.float64)
        k = np.zeros(max(reps, 1), np.float64)
        check(self._fn("bench")(self._h, float(alpha), int(warmup), int(reps), int(flush_bytes),
                                t.ctypes.data_as(C.c_void_p), k.ctypes.data_as(C.c_void_p)),
              "bench")
        return t[:reps], k[:reps]

    def bench_floor(self, reps=20, flush_bytes=0) -> np.ndarray:
        """Per-rep ms of the data-movement floor kernel for this matrix."""
        ms = np.zeros(max(reps, 1), np.float64)
        check(self._fn("bench_floor")(self._h, int(reps), int(flush_bytes),
                                      ms.ctypes.data_as(C.c_void_p)), "bench_floor")
        return ms[:reps]

    def bench_event_overhead(self, reps=20, flush_bytes=0) -> np.ndarray:
        """Per-rep ms of an empty interval between the main-kernel timing events."""
        ms = np.zeros(max(reps, 1), np.float64)
        check(self._fn("bench_event_overhead")(self._h, int(reps), int(flush_bytes),
                                               ms.ctypes.data_as(C.c_void_p)),
              "bench_event_overhead")
        return ms[:reps]

    def autotune(self, classes, reps=7, flush_bytes=1 << 30):
        """Exhaustive autotuner: (us per class, -1 where a class cannot run;
        index of the fastest, which is installed)."""
        arr = (Config * len(classes))(*[Config(*c) for c in classes])
        us = np.zeros(len(classes), np.float64)
        best = C.c_int(-1)
        check(self._fn("autotune")(self._h, arr, len(classes), int(reps), int(flush_bytes),
                                   us.ctypes.data_as(C.c_void_p), C.byref(best)), "autotune")
        return us, best.value

    def apply_tuner(self, net, classes, topk=1, reps=7, flush_bytes=1 << 30) -> int:
        """Learned choice on the GPU; installs and returns the class index.
        With topk > 1 the network's topk best-predicted classes are timed and
        the fastest is installed (a learned shortlist for the autotuner)."""
        arr = (Config * len(classes))(*[Config(*c) for c in classes])
        chosen = C.c_int(-1)
        if topk == 1:
            check(self._fn("apply_tuner")(self._h, net.handle, arr, len(classes), C.byref(chosen)),
                  "apply_tuner")
        else:
            check(self._fn("apply_tuner_topk")(self._h, net.handle, arr, len(classes), int(topk),
                                               int(reps), int(flush_bytes), C.byref(chosen)),
                  "apply_tuner_topk")
        return chosen.value

    def row_features(self) -> np.ndarray:
        f = np.zeros(8, np.float64)
        check(self._fn("row_features")(self._h, f.ctypes.data_as(C.c_void_p)), "row_features")
        return f

    def close(self):
        if self._h:
            self._fn("destroy")(self._h)
            self._h = C.c_void_p()
            self._keep = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count() -> int:
    return spmv_lib().bhs_device_count()


def stream_copy(nbytes: int, reps: int = 10, device: int = 0) -> np.ndarray:
    """Per-rep ms of a device-to-device copy of nbytes (bandwidth probe)."""
    ms = np.zeros(reps, np.float64)
    check(spmv_lib().bhs_stream_copy(int(device), int(nbytes), int(reps),
                                     ms.ctypes.data_as(C.c_void_p)), "stream_copy")
    return ms


def device_synchronize() -> None:
    check(spmv_lib().bhs_device_synchronize(), "device_synchronize")

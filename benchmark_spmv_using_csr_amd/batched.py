"""Batched solves: many independent CSR matrices, each with k <= 16
right-hand sides, in one persistent-kernel launch replayed as a hipGraph
(bhs_*_batch_*).  The baseline runs one plain SpMV per matrix per right-hand
side, back to back.  No reference counterpart (the reference runs one matrix
and one right-hand side per process, spmv_cuda/main.cu:813-816)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import matgen
from ._lib import BatchSchedule, check, spmv_lib

KMAX = 16


def _ptrs(arrays, ctype):
    return (C.c_void_p * len(arrays))(*[a.ctypes.data_as(C.c_void_p) for a in arrays])


class SpmvBatch:
    def __init__(self, mats, ks, dtype=np.float64, device=0):
        self.dtype = np.dtype(dtype)
        self._p = "d" if self.dtype == np.float64 else "s"
        self.L = spmv_lib()
        self.mats = mats
        self.ks = [int(k) for k in ks]
        n = len(mats)
        m = np.array([A.m for A in mats], np.int32)
        nn = np.array([A.n for A in mats], np.int32)
        nnz = np.array([A.nnz for A in mats], np.int32)
        kk = np.array(self.ks, np.int32)
        self._rp = [np.ascontiguousarray(A.row_ptr, np.int32) for A in mats]
        self._col = [np.ascontiguousarray(A.col if A.nnz else np.zeros(1), np.int32) for A in mats]
        self._val = [np.ascontiguousarray(A.val if A.nnz else np.zeros(1), self.dtype)
                     for A in mats]
        for i, A in enumerate(mats):          # the library reads these many entries
            if (len(self._rp[i]) < A.m + 1 or len(self._col[i]) < A.nnz
                    or len(self._val[i]) < A.nnz or int(self._rp[i][A.m]) != A.nnz):
                raise ValueError(f"matrix {i}: row_ptr/col/val shorter than m+1 / nnz")
        self.h = C.c_void_p()
        check(self._fn("batch_create")(C.byref(self.h), device, n, m.ctypes.data_as(C.c_void_p),
                                        nn.ctypes.data_as(C.c_void_p),
                                        nnz.ctypes.data_as(C.c_void_p),
                                        _ptrs(self._rp, C.c_int), _ptrs(self._col, C.c_int),
                                        _ptrs(self._val, None), kk.ctypes.data_as(C.c_void_p)),
              "batch_create")

    def _fn(self, name):
        return getattr(self.L, f"bhs_{self._p}_{name}")

    def set_x(self, Xs):
        self._x = [np.ascontiguousarray(X, self.dtype).reshape(A.n, k) if A.n * k else
                   np.zeros((1, 1), self.dtype) for X, A, k in zip(Xs, self.mats, self.ks)]
        check(self._fn("batch_set_x")(self.h, _ptrs(self._x, None)), "batch_set_x")

    def set_schedule(self, items_per_lane=8, tiles_per_wave=4, waves_per_block=4, grid_blocks=0,
                     order=1, x_stage=0):
        s = BatchSchedule(items_per_lane, tiles_per_wave, waves_per_block, grid_blocks, order,
                          x_stage)
        check(self._fn("batch_set_schedule")(self.h, C.byref(s)), "batch_set_schedule")

    def schedule(self):
        s = BatchSchedule()
        check(self._fn("batch_get_schedule")(self.h, C.byref(s)), "batch_get_schedule")
        return s.as_tuple()

    def run(self, alpha=1.0):
        check(self._fn("batch_run")(self.h, alpha), "batch_run")

    def run_plain(self, alpha=1.0):
        check(self._fn("batch_run_plain")(self.h, alpha), "batch_run_plain")

    def y(self, plain=False):
        Ys = [np.empty((max(A.m, 1), k), self.dtype) for A, k in zip(self.mats, self.ks)]
        check(self._fn("batch_get_y")(self.h, _ptrs(Ys, None), int(plain)), "batch_get_y")
        return [Y[:A.m] for Y, A in zip(Ys, self.mats)]

    def bench(self, plain=False, warmup=2, reps=5, alpha=1.0):
        ms = np.zeros(reps, np.float64)
        check(self._fn("batch_bench")(self.h, alpha, int(plain), warmup, reps,
                                      ms.ctypes.data_as(C.c_void_p)), "batch_bench")
        return ms

    def autotune(self, cands, reps=5):
        arr = (BatchSchedule * len(cands))(*[BatchSchedule(*c) for c in cands])
        ms = np.zeros(len(cands), np.float64)
        best = C.c_int(-1)
        check(self._fn("batch_autotune")(self.h, arr, len(cands), reps,
                                         ms.ctypes.data_as(C.c_void_p), C.byref(best)),
              "batch_autotune")
        return ms, best.value

    def close(self):
        if self.h:
            self._fn("batch_destroy")(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Candidate schedules for the batch autotuner: (S, T, waves/block, grid, order,
# x_stage).  The x-staged ones run on fp64 batches only (the autotuner skips
# them elsewhere).
SCHEDULES = [(S, T, wpb, 0, 1, 0) for S in (4, 8) for T in (1, 4, 16) for wpb in (1, 4)] + \
            [(4, T, wpb, 0, 1, 0) for T in (32, 64) for wpb in (1, 4)] + [(8, 4, 4, 0, 0, 0)] + \
            [(4, T, 1, 0, 1, xs) for xs in (2, 4) for T in (8, 16, 32)]


def batch_set(seed=5, n_mats=256, m_lo=1e3, m_hi=2e5, dtype=np.float64):
    """The batched workload: n_mats matrices, m log-uniform in [m_lo, m_hi], the
    family drawn uniformly from the four tuner families, k in {1, 4, 16}
    right-hand sides per matrix drawn uniformly, X uniform in [-1, 1]."""
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0xBA7C]))
    mats, ks, Xs = [], [], []
    for i in range(n_mats):
        fam = int(rng.integers(4))
        k = int(rng.choice([1, 4, 16]))
        s = seed * 100000 + i
        A = matgen.family_matrix(s, m_lo, m_hi, fam)
        X = matgen.vector(s + 17, A.n * k).reshape(A.n, k)
        mats.append(A)
        ks.append(k)
        Xs.append(X.astype(dtype))
    return mats, ks, Xs

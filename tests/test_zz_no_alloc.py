"""The allocation-free entry point bhs_{s,d}_spmv_noalloc through the C ABI
(GPU).  It runs the kernels of a handle's default tile over device arrays the
caller owns, with the tile partition and the cross-wave partials in the
library's static device scratch; its y must equal, bit for bit, the handle
path with the same tile (same kernels, same summation order), on two calls in
a row, the second with a smaller matrix, and stay within the FP64 oracle's
bound.  The caller's device arrays come from hipMalloc of the HIP runtime the
library itself links (ctypes); the entry point under test allocates nothing."""
import ctypes as C

import numpy as np
import pytest

import cases
import oracle
from benchmark_spmv_using_csr_amd import _lib, matgen
from benchmark_spmv_using_csr_amd.spmv import bhsparse_spmv

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-13, np.float32: 1e-5}
DEFAULT_TILE = (0, 8, 4, 4, 0)     # the tile the allocation-free entry point runs


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def handle_y(c, dtype, alpha):
    op = bhsparse_spmv(dtype)
    _lib.check(op.init_platform(0), "init_platform")
    _lib.check(op.set_config(*DEFAULT_TILE), "set_config")
    y = np.zeros(max(c.m, 1), dtype)
    _lib.check(op.prepare_mem(c.m, c.n, c.nnz, c.row_ptr, c.col, c.val.astype(dtype),
                              c.x.astype(dtype), y), "prepare_mem")
    _lib.check(op.spmv(alpha), "spmv")
    _lib.check(op.get_y(), "get_y")
    op.close()
    return y[:c.m]


_H2D, _D2H = 1, 2     # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip():
    """The HIP runtime libbhsparse_spmv.so is linked against (loaded with it)."""
    _lib.spmv_lib()
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipFree.argtypes = [C.c_void_p]
    for f in (H.hipMalloc, H.hipMemcpy, H.hipFree, H.hipDeviceSynchronize, H.hipSetDevice):
        f.restype = C.c_int
    assert H.hipSetDevice(0) == 0
    return H


class DeviceArrays:
    """hipMalloc'd copies of host arrays (freed on exit)."""

    def __init__(self, *arrays):
        self.H = hip()
        self.ptrs = []
        for a in arrays:
            a = np.ascontiguousarray(a)
            p = C.c_void_p()
            assert self.H.hipMalloc(C.byref(p), max(a.nbytes, 16)) == 0
            self.ptrs.append(p)
            if a.nbytes:
                assert self.H.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, _H2D) == 0

    def get(self, i, like):
        out = np.empty_like(like)
        assert self.H.hipDeviceSynchronize() == 0
        assert self.H.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptrs[i], out.nbytes,
                                _D2H) == 0
        return out

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.H.hipDeviceSynchronize()
        for p in self.ptrs:
            self.H.hipFree(p)
        return False


def call_noalloc(dtype, m, n, nnz, ptrs, alpha, stream=None):
    L = _lib.spmv_lib()
    fn = L.bhs_d_spmv_noalloc if dtype == np.float64 else L.bhs_s_spmv_noalloc
    return fn(m, n, nnz, *ptrs, alpha, stream)


def noalloc_y(c, dtype, alpha, y_init=7.0):
    y0 = np.full(max(c.m, 1), y_init, dtype)        # prior contents must not matter
    with DeviceArrays(np.asarray(c.row_ptr, np.int32), np.asarray(c.col, np.int32),
                      c.val.astype(dtype), c.x.astype(dtype), y0) as d:
        _lib.check(call_noalloc(dtype, c.m, c.n, c.nnz, d.ptrs, alpha), "spmv_noalloc")
        return d.get(4, y0)[:c.m]


def power_law_case(seed, m, n, target):
    A = matgen.power_law(seed, m, n, 2.0, min(n, 50000), 0.1, target)
    x = matgen.vector(seed + 1, n)
    return cases.Case(f"power-law m={m} nnz={A.nnz}", A.m, A.n, A.row_ptr, A.col, A.val, x)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_equals_the_handle_path_twice(dtype):
    # a larger matrix, then a smaller one: the second call must not see the
    # first one's partition or partials (same scratch, rewritten)
    big = power_law_case(11, 200_000, 180_000, 2_000_003)
    small = power_law_case(12, 9_000, 12_000, 60_001)
    for c, alpha in ((big, 1.0), (small, -0.75)):
        say(f"noalloc {c.name} {np.dtype(dtype).name} alpha={alpha}")
        assert c.nnz % 512 != 0                        # the last tile is partial
        yh = handle_y(c, dtype, alpha)
        yn = noalloc_y(c, dtype, alpha)
        assert np.array_equal(yn, yh), c.name
        v = c.val.astype(dtype).astype(np.float64)
        xx = c.x.astype(dtype).astype(np.float64)
        yr = alpha * oracle.spmv_serial(c.row_ptr, c.col, v, xx)
        assert oracle.rel_error(c.row_ptr, c.col, alpha * v, xx, yn, yr) <= TOL[dtype]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_adversarial_shapes(dtype):
    # leading/trailing empty rows, dirty tiles, no nonzeros, fewer than a tile
    for c in cases.adversarial(512, 4, seed=3, integer=False):
        say(f"noalloc adversarial {c.name!r} {np.dtype(dtype).name}")
        yh = handle_y(c, dtype, 1.0)
        yn = noalloc_y(c, dtype, 1.0)
        assert np.array_equal(yn, yh), c.name


def test_noalloc_rejects_what_it_cannot_hold():
    L = _lib.spmv_lib()
    with DeviceArrays(np.zeros(64, np.int32), np.zeros(64, np.float64)) as d:
        ip, p = d.ptrs
        # scratch too small for 2^30 nonzeros: refused before anything is launched
        rc = L.bhs_d_spmv_noalloc(1, 1, 1 << 30, ip, ip, p, p, p, 1.0, None)
        assert rc == -3 and b"scratch" in L.bhs_last_error()
        # misaligned val
        rc = L.bhs_d_spmv_noalloc(1, 1, 1, ip, ip, C.c_void_p(p.value + 8), p, p, 1.0, None)
        assert rc == -3
        assert L.bhs_d_spmv_noalloc(-1, 1, 1, ip, ip, p, p, p, 1.0, None) == -3


def test_noalloc_under_graph_capture():
    """The callers it is for: captured into a hipGraph on the caller's stream and
    replayed, with x changed and eager work done between replays; every replay
    gives the eager call's y for the current x, bit for bit."""
    H = hip()
    for f, args in (("hipStreamCreate", [C.POINTER(C.c_void_p)]),
                    ("hipStreamBeginCapture", [C.c_void_p, C.c_int]),
                    ("hipStreamEndCapture", [C.c_void_p, C.POINTER(C.c_void_p)]),
                    ("hipGraphInstantiate", [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_size_t]),
                    ("hipGraphLaunch", [C.c_void_p, C.c_void_p]),
                    ("hipStreamSynchronize", [C.c_void_p]),
                    ("hipGraphExecDestroy", [C.c_void_p]), ("hipGraphDestroy", [C.c_void_p]),
                    ("hipStreamDestroy", [C.c_void_p])):
        getattr(H, f).argtypes = args
        getattr(H, f).restype = C.c_int
    c = power_law_case(31, 50_000, 40_000, 400_001)
    rng = np.random.default_rng(4)
    xs = [c.x] + [rng.uniform(-1, 1, c.n) for _ in range(3)]
    say(f"noalloc graph capture {c.name}")
    y0 = np.zeros(c.m)
    with DeviceArrays(np.asarray(c.row_ptr, np.int32), np.asarray(c.col, np.int32), c.val,
                      c.x, y0) as d:
        st = C.c_void_p()
        assert H.hipStreamCreate(C.byref(st)) == 0
        # eager results for every x, on the same stream
        eager = []
        for xv in xs:
            assert H.hipMemcpy(d.ptrs[3], np.ascontiguousarray(xv).ctypes.data_as(C.c_void_p),
                               xv.nbytes, _H2D) == 0
            _lib.check(call_noalloc(np.float64, c.m, c.n, c.nnz, d.ptrs, 1.0, st), "noalloc")
            assert H.hipStreamSynchronize(st) == 0
            eager.append(d.get(4, y0).copy())
        graph, gexec = C.c_void_p(), C.c_void_p()
        assert H.hipStreamBeginCapture(st, 0) == 0             # hipStreamCaptureModeGlobal
        rc = call_noalloc(np.float64, c.m, c.n, c.nnz, d.ptrs, 1.0, st)
        assert H.hipStreamEndCapture(st, C.byref(graph)) == 0
        _lib.check(rc, "noalloc under capture")
        assert H.hipGraphInstantiate(C.byref(gexec), graph, None, None, 0) == 0
        for k in (1, 2, 3, 0):
            xv = xs[k]
            assert H.hipMemcpy(d.ptrs[3], np.ascontiguousarray(xv).ctypes.data_as(C.c_void_p),
                               xv.nbytes, _H2D) == 0
            # eager work between replays, on the device's other arrays
            assert H.hipMemcpy(d.ptrs[4], y0.ctypes.data_as(C.c_void_p), y0.nbytes, _H2D) == 0
            assert H.hipGraphLaunch(gexec, st) == 0
            assert H.hipStreamSynchronize(st) == 0
            assert np.array_equal(d.get(4, y0), eager[k]), k
        assert H.hipGraphExecDestroy(gexec) == 0 and H.hipGraphDestroy(graph) == 0
        assert H.hipStreamDestroy(st) == 0

"""Unified-memory modes through the C ABI: the reference's USE_SVM_ALWAYS=1
build (spmv_cuda/Makefile:6, bhsparse_spmv_cuda.h:250-263) keeps A, x and y in
memory that host and device share.  BHS_MEMORY_MANAGED does it with
hipMallocManaged where the device has managed memory; BHS_MEMORY_HOST_MAPPED
with pinned host memory mapped into the device, which needs no managed memory.
The same kernels run on those copies, so y must equal the device-memory result
bit for bit, and meet the oracle's bound (exact on integer data).  Also: set_x
and get_y in a unified mode, the mode switch refused while data is loaded, and
a device without managed memory refusing MANAGED with
BHSPARSE_UNSUPPORTED_DEVICE while the handle stays usable."""
import numpy as np
import pytest

import cases
import oracle
from benchmark_spmv_using_csr_amd import _lib, matgen
from benchmark_spmv_using_csr_amd.spmv import (MEMORY_DEVICE, MEMORY_HOST_MAPPED,
                                               MEMORY_MANAGED, bhsparse_spmv)

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-13, np.float32: 1e-5}
# (variant, items per lane, waves per block, tiles per wave, tpv)
CONFIGS = [(0, 8, 4, 2, 0), (2, 8, 4, 1, 4), (1, 8, 4, 1, 0)]
SMALL = cases.Case("2x2", 2, 2, np.array([0, 1, 2], np.int32), np.array([1, 0], np.int32),
                   np.array([2.0, 3.0]), np.array([1.0, 1.0]))


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def load(c, dtype, cfg, mode):
    """A handle holding case c in `mode`: (handle, y, init_data's code)."""
    op = bhsparse_spmv(dtype)
    _lib.check(op.set_memory_mode(mode), "set_memory_mode")
    _lib.check(op.init_platform(0), "init_platform")
    _lib.check(op.set_config(*cfg), "set_config")
    y = np.full(max(c.m, 1), np.nan, dtype)
    rc = op.prepare_mem(c.m, c.n, c.nnz, c.row_ptr, c.col, c.val.astype(dtype),
                        c.x.astype(dtype), y)
    return op, y, rc


def run(c, dtype, cfg, mode, alpha=1.0):
    op, y, rc = load(c, dtype, cfg, mode)
    _lib.check(rc, "prepare_mem")
    assert op.memory_mode() == mode
    _lib.check(op.spmv(alpha), "spmv")
    _lib.check(op.get_y(), "get_y")
    return y[:c.m], op


def has_managed_memory():
    op, _, rc = load(SMALL, np.float64, CONFIGS[0], MEMORY_MANAGED)
    op.close()
    assert rc in (0, _lib.BHSPARSE_UNSUPPORTED_DEVICE), rc
    return rc == 0


def check_oracle(c, dtype, y, integer, alpha=1.0):
    v = c.val.astype(dtype).astype(np.float64)
    x = c.x.astype(dtype).astype(np.float64)
    yr = alpha * oracle.spmv_serial(c.row_ptr, c.col, v, x)
    if integer:
        assert np.array_equal(y.astype(np.float64), yr), c.name
    assert oracle.rel_error(c.row_ptr, c.col, v, x, y, yr) <= TOL[dtype], c.name


@pytest.mark.parametrize("mode", [MEMORY_HOST_MAPPED, MEMORY_MANAGED])
@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_unified_equals_device_and_oracle(mode, cfg, dtype):
    if mode == MEMORY_MANAGED and not has_managed_memory():
        pytest.skip("device reports no managed memory (hipDeviceAttributeManagedMemory = 0)")
    for integer in (True, False):
        for c in cases.adversarial(64 * cfg[1], cfg[3], seed=3, integer=integer)[:12]:
            say(f"memory mode {mode} {c.name!r} cfg={cfg} {np.dtype(dtype).name} int={integer}")
            yu, op = run(c, dtype, cfg, mode)
            op.close()
            yd, op = run(c, dtype, cfg, MEMORY_DEVICE)
            op.close()
            assert not np.isnan(yu).any(), c.name
            assert np.array_equal(yu, yd), c.name
            check_oracle(c, dtype, yu, integer)


def test_no_managed_memory_is_an_error_not_a_crash():
    say("managed memory capability")
    op, y, rc = load(SMALL, np.float64, CONFIGS[0], MEMORY_MANAGED)
    if rc == 0:
        op.close()
        pytest.skip("device has managed memory; the refusal path does not arise")
    assert rc == _lib.BHSPARSE_UNSUPPORTED_DEVICE
    assert b"managed memory" in _lib.spmv_lib().bhs_last_error()
    assert op.spmv(1.0) == _lib.BHSPARSE_NOT_INITIALIZED       # nothing half-loaded
    _lib.check(op.set_memory_mode(MEMORY_DEVICE), "set_memory_mode")
    _lib.check(op.prepare_mem(SMALL.m, SMALL.n, SMALL.nnz, SMALL.row_ptr, SMALL.col,
                              SMALL.val, SMALL.x, y), "prepare_mem")
    _lib.check(op.spmv(1.0), "spmv")
    _lib.check(op.get_y(), "get_y")
    assert y.tolist() == [2.0, 3.0]
    op.close()


def test_host_mapped_set_x_and_mode_switch():
    A = matgen.power_law(5, 30000, 30000, 1.8, 3000, 0.1, 0)
    c = cases.Case("pl", A.m, A.n, A.row_ptr, A.col, A.val, matgen.vector(2, A.n))
    say("host-mapped set_x / mode switch")
    y, op = run(c, np.float64, (0, 8, 4, 2, 0), MEMORY_HOST_MAPPED, alpha=-0.5)
    check_oracle(c, np.float64, y, False, alpha=-0.5)
    # loaded in host-mapped memory: switching needs free_mem first
    assert op.set_memory_mode(MEMORY_DEVICE) == _lib.BHSPARSE_INVALID_ARGUMENT
    assert op.set_memory_mode(MEMORY_HOST_MAPPED) == 0
    x2 = matgen.vector(9, A.n)
    _lib.check(op.set_x(x2), "set_x")
    _lib.check(op.spmv(1.0), "spmv")
    c2 = cases.Case("pl-x2", A.m, A.n, A.row_ptr, A.col, A.val, x2)
    check_oracle(c2, np.float64, op.y(), False)
    _lib.check(op.free_mem(), "free_mem")
    assert op.set_memory_mode(MEMORY_DEVICE) == 0 and op.memory_mode() == MEMORY_DEVICE
    op.close()

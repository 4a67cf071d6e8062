"""Golden fixtures (tests/golden/*.json, written by make_golden.py without the
oracle): the oracle must reproduce them exactly, every planted mistake in the
oracle must break at least one of its pins, and the HIP library (through the C
ABI) must reproduce them exactly too -- the data are small integers, so every
summation order gives the exact y."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import cases
import oracle
from benchmark_spmv_using_csr_amd import _lib, matgen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["six_by_six", "laplace_ramp", "int_skewed", "tall_dirty"]
REF_SHAPES = [(64, 4, 1), (32, 4, 7), (64, 8, 5), (8, 16, 6)]


def load(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        d = json.load(f)
    for k, t in (("row_ptr", np.int32), ("col", np.int32), ("val", np.float64),
                 ("x", np.float64), ("y", np.float64)):
        d[k] = np.asarray(d[k], t)
    return d


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    d = load(name)
    rp, col, val, x, y = d["row_ptr"], d["col"], d["val"], d["x"], d["y"]
    assert np.array_equal(oracle.spmv_serial(rp, col, val, x), y)
    assert np.array_equal(oracle.spmv_serial_f32(rp, col, val, x), y.astype(np.float32))
    assert np.array_equal(oracle.spmv_omp(rp, col, val, x, 3)[0], y)
    for W, S, T in REF_SHAPES:
        assert np.array_equal(oracle.segsum_ref(rp, col, val, x, W, S, T)[0], y), (W, S, T)
    for P, tr in d.get("tile_row", {}).items():
        assert np.array_equal(oracle.partition(rp, int(rp[-1]), int(P)), tr), P


def pins_hold():
    """The oracle's checks that do not rest on the oracle: golden fixtures,
    scipy, the Laplacian closed form, bisection.  True if all pass."""
    import bisect
    try:
        for name in NAMES:
            d = load(name)
            rp, col, val, x, y = d["row_ptr"], d["col"], d["val"], d["x"], d["y"]
            if not np.array_equal(oracle.spmv_serial(rp, col, val, x), y):
                return False
            for W, S, T in REF_SHAPES:
                if not np.array_equal(oracle.segsum_ref(rp, col, val, x, W, S, T)[0], y):
                    return False
            for P, tr in d.get("tile_row", {}).items():
                if not np.array_equal(oracle.partition(rp, int(rp[-1]), int(P)), tr):
                    return False
        c = cases.from_lengths("r", np.random.default_rng(0).integers(0, 40, 300), n=500, seed=0)
        A = sp.csr_matrix((c.val, c.col, c.row_ptr), shape=(c.m, c.n))
        if np.max(np.abs(oracle.spmv_serial(c.row_ptr, c.col, c.val, c.x) - A @ c.x)) > 1e-13:
            return False
        rp = c.row_ptr
        for key in range(0, int(rp[-1]) + 1, 7):
            if oracle.upper_bound(rp, key) != bisect.bisect_right(list(rp), key):
                return False
        return True
    except AssertionError:
        return False


def test_pins_hold_for_the_oracle_as_written():
    assert pins_hold()


@pytest.mark.parametrize("mistake", sorted(oracle.MUTATIONS))
def test_planted_oracle_mistake_breaks_a_pin(mistake):
    with oracle.mutated(mistake):
        assert not pins_hold(), mistake
    assert pins_hold()


# ---- the HIP library against the fixtures (GPU) ----------------------------
SEGSUM = [(0, 4, 1, 1, 0), (0, 8, 4, 3, 0), (0, 16, 2, 2, 0), (0, 32, 2, 1, 0), (0, 4, 4, 16, 0),
          (2, 4, 1, 3, 2), (2, 8, 4, 2, 4), (2, 16, 2, 1, 8)]
CSRVEC = [(1, 8, 4, 1, 0), (1, 8, 4, 1, 2), (1, 8, 4, 1, 64)]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name", NAMES)
def test_library_reproduces_golden(name, dtype):
    from benchmark_spmv_using_csr_amd.spmv import bhsparse_spmv
    d = load(name)
    print(f"[gpu] golden {name} {np.dtype(dtype).name}", flush=True)
    op = bhsparse_spmv(dtype)
    _lib.check(op.init_platform(0), "init_platform")
    y = np.full(d["m"], 7, dtype)                 # prior contents must not matter
    _lib.check(op.prepare_mem(d["m"], d["n"], int(d["row_ptr"][-1]), d["row_ptr"], d["col"],
                              d["val"].astype(dtype), d["x"].astype(dtype), y), "prepare_mem")
    for cfg in SEGSUM + CSRVEC:
        _lib.check(op.set_config(*cfg), "set_config")
        if cfg[0] == 0:               # the plain layout's partition is the fixture's
            P = 64 * cfg[1]
            tr = op.partition()
            if str(P) in d.get("tile_row", {}):
                assert np.array_equal(tr, d["tile_row"][str(P)]), cfg
        _lib.check(op.spmv(1.0), "spmv")
        _lib.check(op.get_y(), "get_y")
        assert np.array_equal(y, d["y"].astype(dtype)), cfg
    op.close()

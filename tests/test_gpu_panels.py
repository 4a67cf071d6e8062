"""The column-panel variant (BHS_VARIANT_SEGSUM_PANELS) through the C ABI:
partition bit-exact on the stacked matrix, y bit-exact against the lane-exact
model run on the stacked matrix (its wave-boundary partials added in the
combine, chunk by chunk) and summed over panels in order, y within the error
bound of the FP64 oracle (exact on integer data), consistent when the handle
switches between layouts, and no slower than a few times the plain layout on
banded matrices (every stacked row is nonempty)."""
import numpy as np
import pytest

import cases
import kernel_model as km
import oracle
from benchmark_spmv_using_csr_amd import _lib, matgen
from benchmark_spmv_using_csr_amd.spmv import bhsparse_spmv

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-13, np.float32: 1e-5}
PANELS = 2
# (items per lane, waves per block, tiles per wave, panels)
CONFIGS = [(4, 1, 3, 2), (8, 4, 2, 4), (16, 2, 1, 8)]


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def run(c, dtype, cfg, alpha=1.0):
    S, wpb, tpw, npan = cfg
    op = bhsparse_spmv(dtype)
    _lib.check(op.init_platform(0), "init_platform")
    _lib.check(op.set_config(PANELS, S, wpb, tpw, npan), "set_config")
    y = np.full(max(c.m, 1), np.nan, dtype)
    _lib.check(op.prepare_mem(c.m, c.n, c.nnz, c.row_ptr, c.col, c.val.astype(dtype),
                              c.x.astype(dtype), y), "prepare_mem")
    _lib.check(op.spmv(alpha), "spmv")
    _lib.check(op.get_y(), "get_y")
    return y[:c.m], op


def model(c, dtype, cfg):
    S, _, tpw, npan = cfg
    st, idx = cases.stack_panels(c, npan)
    return km.segsum_panels(st.row_ptr, st.col, st.val, idx, c.x, S=S, TPW=tpw, dtype=dtype)


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_adversarial_against_model_and_oracle(cfg, dtype):
    S, _, tpw, _ = cfg
    for integer in (True, False):
        for c in cases.adversarial(64 * S, tpw, seed=S + tpw, integer=integer):
            say(f"panels {c.name!r} cfg={cfg} {np.dtype(dtype).name} int={integer}")
            y, op = run(c, dtype, cfg)
            op.close()
            assert not np.isnan(y).any(), c.name
            assert np.array_equal(y, model(c, dtype, cfg)), c.name
            v = c.val.astype(dtype).astype(np.float64)
            x = c.x.astype(dtype).astype(np.float64)
            yr = oracle.spmv_serial(c.row_ptr, c.col, v, x)
            if integer:
                assert np.array_equal(y.astype(np.float64), yr), c.name
            assert oracle.rel_error(c.row_ptr, c.col, v, x, y, yr) <= TOL[dtype], c.name


def long_chain_case(seed, wave_nnz, npan):
    """Short random rows around two long ones: in every panel the first long
    row spans about 150 waves (three 64-entry chunks of wave partials, the last
    one partial) and the second exactly 64, next to rows spanning 2-3 waves."""
    rng = np.random.default_rng(seed)
    n = 40000
    lens = list(rng.integers(0, 12, 3000))
    lens[100] = 150 * wave_nnz * npan + 37
    lens[2000] = 64 * wave_nnz * npan
    lens[2500] = 2 * wave_nnz * npan + 5
    lens[2900] = 3 * wave_nnz * npan
    rp = np.zeros(len(lens) + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    # long rows: every column in order, repeated in blocks so that each panel
    # (a contiguous range of n / npan columns) gets the same share
    col = np.empty(int(rp[-1]), np.int32)
    for i, L in enumerate(lens):
        if L > n:
            col[rp[i]:rp[i + 1]] = np.sort(np.arange(L, dtype=np.int64) * n // L)
        else:
            col[rp[i]:rp[i + 1]] = np.sort(rng.choice(n, L, replace=False))
    val = rng.integers(-8, 9, int(rp[-1])).astype(np.float64)
    x = rng.integers(-8, 9, n).astype(np.float64)
    return cases.Case("long chains", len(lens), n, rp.astype(np.int32), col, val, x)


@pytest.mark.parametrize("cfg", [(4, 1, 1, 2), (8, 2, 1, 4)])
def test_wave_chains_longer_than_a_chunk(cfg):
    """Rows whose partials span 64 and more waves: the combine sums a chain in
    chunks of 64 wave partials; bit-exact against the model and, on integer
    data, exactly the oracle's y."""
    S, _, tpw, npan = cfg
    c = long_chain_case(S, 64 * S * tpw, npan)
    say(f"panels long chains cfg={cfg} nnz={c.nnz}")
    y, op = run(c, np.float64, cfg)
    op.close()
    assert np.array_equal(y, model(c, np.float64, cfg))
    yr = oracle.spmv_serial(c.row_ptr, c.col, c.val, c.x)
    assert np.array_equal(y, yr)


def many_chains_case(seed, wave_nnz, npan):
    """Ten neighbouring rows, each about 2.5 waves long in every panel, among
    short random rows: one combine wave receives 10 x npan chains, more than
    the slots it loads up front, so most of them come from the linked list."""
    rng = np.random.default_rng(seed)
    n = 40000
    lens = list(rng.integers(0, 12, 600))
    for i in range(10, 20):
        lens[i] = int(2.5 * wave_nnz * npan) + i
    rp = np.zeros(len(lens) + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(n, L, replace=False)) for L in lens]).astype(np.int32)
    val = rng.integers(-8, 9, int(rp[-1])).astype(np.float64)
    x = rng.integers(-8, 9, n).astype(np.float64)
    return cases.Case("many chains", len(lens), n, rp.astype(np.int32), col, val, x)


@pytest.mark.parametrize("cfg", [(4, 1, 1, 8), (8, 2, 1, 4), (4, 1, 2, 2)])
def test_more_chains_than_slots_in_one_combine_wave(cfg):
    S, _, tpw, npan = cfg
    c = many_chains_case(S + npan, 64 * S * tpw, npan)
    say(f"panels many chains cfg={cfg} nnz={c.nnz}")
    y, op = run(c, np.float64, cfg)
    op.close()
    assert np.array_equal(y, model(c, np.float64, cfg))
    assert np.array_equal(y, oracle.spmv_serial(c.row_ptr, c.col, c.val, c.x))


@pytest.mark.parametrize("npan", [2, 4, 8])
def test_partition_of_the_stacked_matrix(npan):
    A = matgen.power_law(11, 20000, 20000, 1.8, 5000, 0.1, 0)
    c = cases.Case("pl", A.m, A.n, A.row_ptr, A.col, A.val, matgen.vector(1, A.n))
    say(f"panel partition npan={npan}")
    y, op = run(c, np.float64, (8, 4, 3, npan))
    st, _ = cases.stack_panels(c, npan)
    assert np.array_equal(op.partition(), oracle.partition(st.row_ptr, st.nnz, 64 * 8))
    op.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_power_law_full_size_and_layout_switches(dtype):
    A = matgen.power_law_1m()
    x = matgen.vector(102, A.n)
    c = cases.Case("power-law", A.m, A.n, A.row_ptr, A.col, A.val, x)
    op = bhsparse_spmv(dtype)
    _lib.check(op.init_platform(0), "init_platform")
    y = np.zeros(A.m, dtype)
    _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val.astype(dtype),
                              x.astype(dtype), y), "prepare_mem")
    v = A.val.astype(dtype).astype(np.float64)
    x64 = x.astype(dtype).astype(np.float64)
    yr = oracle.spmv_serial(A.row_ptr, A.col, v, x64)
    outs = {}
    for cfg in ((PANELS, 4, 1, 8, 2), (0, 4, 1, 8, 0), (PANELS, 8, 4, 4, 4), (PANELS, 4, 1, 8, 2)):
        say(f"full power-law {np.dtype(dtype).name} cfg={cfg}")
        _lib.check(op.set_config(*cfg), "set_config")
        _lib.check(op.spmv(1.0), "spmv")
        _lib.check(op.get_y(), "get_y")
        e = oracle.rel_error(A.row_ptr, A.col, v, x64, y, yr)
        assert e <= TOL[dtype], (cfg, e)
        if cfg in outs:
            assert np.array_equal(outs[cfg], y), cfg      # same layout again: same bits
        outs[cfg] = y.copy()
    op.close()


def test_banded_matrix_no_pathological_slowdown():
    """On a banded matrix every row's entries fall in one or two panels; the
    stacked matrix keeps only nonempty (panel, row) pairs, so no tile spans a
    long run of empty rows and the panel layout stays within a small factor of
    the plain one."""
    A = matgen.laplacian7(64, 64, 64)
    x = matgen.vector(3, A.n)
    op = bhsparse_spmv(np.float64)
    _lib.check(op.init_platform(0), "init_platform")
    y = np.zeros(A.m)
    _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val, x, y), "prepare_mem")
    yr = oracle.spmv_serial(A.row_ptr, A.col, A.val, x)
    t = {}
    for cfg in ((0, 4, 4, 8, 0), (PANELS, 4, 4, 8, 2), (PANELS, 4, 4, 8, 8)):
        say(f"banded laplace 64^3 cfg={cfg}")
        _lib.check(op.set_config(*cfg), "set_config")
        ms, _ = op.bench(1.0, warmup=2, reps=7, flush_bytes=0)
        t[cfg] = float(np.median(ms))
        _lib.check(op.get_y(), "get_y")
        assert oracle.rel_error(A.row_ptr, A.col, A.val, x, y, yr) <= 1e-13, cfg
    base = t[(0, 4, 4, 8, 0)]
    for cfg, v in t.items():
        assert v <= 4 * base + 0.05, (cfg, v, base)
    op.close()

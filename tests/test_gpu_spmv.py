"""GPU parity of the segmented-sum SpMV through the C ABI.

Bars: partitions and counters bit-exact against the oracle; y bit-exact
against the lane-exact kernel model; y within |y - y*| / sum|a x| <= 1e-13
(fp64) / 1e-5 (fp32) of the FP64 serial oracle; exact on integer data."""
import numpy as np
import pytest

import cases
import kernel_model as km
import oracle
from benchmark_spmv_using_csr_amd import _lib, matgen
from benchmark_spmv_using_csr_amd.spmv import bhsparse_spmv

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-13, np.float32: 1e-5}
# (items per lane S, waves per block, tiles per wave)
CONFIGS = [(8, 4, 3), (4, 1, 1), (16, 2, 2), (32, 2, 1), (8, 1, 7), (16, 4, 5)]


def say(msg):
    print(f"[gpu] {msg}", flush=True)


def make_op(dtype, S=None, wpb=4, tpw=3, variant=0, tpv=0):
    op = bhsparse_spmv(dtype)
    _lib.check(op.init_platform(0), "init_platform")
    if S is not None:
        _lib.check(op.set_config(variant, S, wpb, tpw, tpv), "set_config")
    return op


def gpu_spmv(c, dtype, cfg=(8, 4, 3), alpha=1.0, variant=0, tpv=0):
    S, wpb, tpw = cfg
    op = make_op(dtype, S, wpb, tpw, variant, tpv)
    y = np.full(max(c.m, 1), np.nan, dtype)
    _lib.check(op.prepare_mem(c.m, c.n, c.nnz, c.row_ptr, c.col, c.val.astype(dtype),
                              c.x.astype(dtype), y), "prepare_mem")
    _lib.check(op.spmv(alpha), "spmv")
    _lib.check(op.get_y(), "get_y")
    return y[:c.m], op


def rel_err(c, dtype, y):
    v = c.val.astype(dtype).astype(np.float64)
    x = c.x.astype(dtype).astype(np.float64)
    yr = oracle.spmv_serial(c.row_ptr, c.col, v, x)
    return oracle.rel_error(c.row_ptr, c.col, v, x, y, yr), yr


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_six_by_six_reference_flow(dtype):
    # init_platform -> prepare_mem -> run_benchmark -> get_y (main.cu:490-501),
    # repeated: the second call must give the same answer (it does not in the
    # reference, whose calibration adds to stale y)
    c = cases.six_by_six()
    op = bhsparse_spmv(dtype)
    assert op.init_platform() == 0
    y = np.zeros(6, dtype)
    assert op.prepare_mem(6, 6, 15, c.row_ptr, c.col, c.val.astype(dtype), c.x.astype(dtype),
                          y) == 0
    for _ in range(3):
        assert op.run_benchmark() == 0
        assert op.get_y() == 0
        assert np.array_equal(y, cases.SIX_BY_SIX_Y.astype(dtype))
    assert op.free_platform() == 0 and op.free_mem() == 0


@pytest.mark.parametrize("cfg", CONFIGS)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_adversarial_against_model_and_oracle(cfg, dtype):
    S, wpb, tpw = cfg
    for integer in (True, False):
        for c in cases.adversarial(64 * S, tpw, seed=S + tpw, integer=integer):
            say(f"adversarial {c.name!r} cfg={cfg} {np.dtype(dtype).name} int={integer}")
            y, _ = gpu_spmv(c, dtype, cfg)
            assert not np.isnan(y).any(), c.name
            ym = km.segsum(c.row_ptr, c.col, c.val.astype(dtype), c.x.astype(dtype), S=S,
                           TPW=tpw, dtype=dtype)
            assert np.array_equal(y, ym), (c.name, np.flatnonzero(y != ym)[:8])
            e, yr = rel_err(c, dtype, y)
            if integer:
                assert np.array_equal(y.astype(np.float64), yr), c.name
            assert e <= TOL[dtype], (c.name, e)


@pytest.mark.parametrize("S", [4, 8, 16, 32])
def test_partition_bit_exact(S):
    A = matgen.power_law(11, 20000, 20000, 1.8, 5000, 0.1, 0)
    c = cases.Case("pl", A.m, A.n, A.row_ptr, A.col, A.val, matgen.vector(1, A.n))
    say(f"partition S={S} nnz={c.nnz}")
    op = make_op(np.float64, S, 4, 3)
    y = np.empty(c.m)
    _lib.check(op.prepare_mem(c.m, c.n, c.nnz, c.row_ptr, c.col, c.val, c.x, y), "prepare_mem")
    assert np.array_equal(op.partition(), oracle.partition(c.row_ptr, c.nnz, 64 * S))


@pytest.mark.parametrize("cfg", CONFIGS[:4])
def test_counters_match_oracle(cfg):
    S, wpb, tpw = cfg
    A = matgen.power_law(12, 30000, 30000, 1.7, 20000, 0.1, 0)
    c = cases.Case("pl", A.m, A.n, A.row_ptr, A.col, A.val, matgen.vector(2, A.n))
    say(f"counters cfg={cfg} nnz={c.nnz}")
    _, op = gpu_spmv(c, np.float64, cfg)
    st = op.stats()
    ts = oracle.tile_stats(c.row_ptr, c.nnz, 64 * S, tpw)
    for k in ("n_tiles", "n_waves", "n_fast", "n_dirty", "n_sync", "n_lead_empty"):
        assert st[k] == ts[k], k
    assert st["n_fast"] > 0 and st["n_dirty"] > 0 and st["n_sync"] > 0
    assert st["chunk_rows"] == expected_chunk_rows(c, S)


def expected_chunk_rows(c, S):
    """256 when at least 5% of the tiles span 128-255 rows, else 128."""
    tr = oracle.partition(c.row_ptr, c.nnz, 64 * S)
    span = np.diff(np.asarray(tr, np.int64))
    return 256 if 20 * int(((span >= 128) & (span < 256)).sum()) >= len(span) else 128


@pytest.mark.parametrize("S", [4, 8, 16])
def test_row_chunk_choice(S):
    """A matrix of about 190 rows per tile takes the 256-row chunk, one of
    about 30 the 128-row chunk; both give the model's bits."""
    rng = np.random.default_rng(S)
    for per_tile in (190, 30):
        Lm = 64 * S / per_tile
        lens = (np.floor(Lm) + (rng.random(20000) < Lm - np.floor(Lm))).astype(np.int64)
        lens[rng.random(20000) < 0.05] = 0
        c = cases.from_lengths(f"{per_tile} rows per tile", lens, 3000, seed=S)
        say(f"row chunk S={S} rows/tile~{per_tile}")
        y, op = gpu_spmv(c, np.float64, (S, 4, 3))
        want = expected_chunk_rows(c, S)
        assert want == (256 if per_tile == 190 else 128)
        assert op.stats()["chunk_rows"] == want
        assert np.array_equal(y, km.segsum(c.row_ptr, c.col, c.val, c.x, S=S, TPW=3))


@pytest.mark.parametrize("S", [4, 8])
def test_long_tiles_under_the_256_row_chunk(S):
    """Tiles longer than the 256-row chunk (walked in the long-tile loop, empty
    runs skipped by galloping) on a matrix that takes the 256-row chunk: runs
    of 300-5,000 empty rows and of length-1 rows among ~190-row tiles."""
    rng = np.random.default_rng(100 + S)
    Lm = 64 * S / 190
    parts = []
    for blk in range(12):
        lens = (np.floor(Lm) + (rng.random(1500) < Lm - np.floor(Lm))).astype(np.int64)
        parts.append(lens)
        parts.append(np.zeros(int(rng.choice([300, 700, 5000])), np.int64))
        parts.append(np.ones(int(rng.choice([257, 400])), np.int64))
    lens = np.concatenate(parts)
    c = cases.from_lengths("long tiles, 256-row chunk", lens, 3000, seed=S)
    say(f"long tiles under the 256-row chunk S={S} m={c.m}")
    y, op = gpu_spmv(c, np.float64, (S, 4, 3))
    assert op.stats()["chunk_rows"] == expected_chunk_rows(c, S) == 256
    span = np.diff(np.asarray(oracle.partition(c.row_ptr, c.nnz, 64 * S), np.int64))
    assert (span >= 256).any()
    assert np.array_equal(y, km.segsum(c.row_ptr, c.col, c.val, c.x, S=S, TPW=3))
    v, x = c.val, c.x
    yr = oracle.spmv_serial(c.row_ptr, c.col, v, x)
    assert oracle.rel_error(c.row_ptr, c.col, v, x, y, yr) <= 1e-13


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_repeatable_and_independent_of_previous_y(dtype):
    A = matgen.power_law(13, 50000, 50000, 2.0, 20000, 0.1, 0)
    x1 = matgen.vector(3, A.n).astype(dtype)
    x2 = matgen.vector(4, A.n).astype(dtype)
    say(f"repeatable nnz={A.nnz}")
    op = make_op(dtype, 8, 4, 3)
    y = np.zeros(A.m, dtype)
    _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val.astype(dtype), x1, y),
               "prepare_mem")
    _lib.check(op.spmv(1.0), "spmv")
    ya = op.y()
    _lib.check(op.spmv(1.0), "spmv")
    assert np.array_equal(op.y(), ya)                 # bitwise repeatable
    _lib.check(op.set_x(x2), "set_x")
    _lib.check(op.spmv(1.0), "spmv")
    yb = op.y()
    ym = km.segsum(A.row_ptr, A.col, A.val.astype(dtype), x2, S=8, TPW=3, dtype=dtype)
    assert np.array_equal(yb, ym)                     # nothing left over from x1


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_alpha(dtype):
    c = cases.from_lengths("r", np.random.default_rng(9).integers(0, 40, 3000), n=4000, seed=9)
    y1, _ = gpu_spmv(c, dtype, alpha=1.0)
    y2, _ = gpu_spmv(c, dtype, alpha=-0.5)
    assert np.array_equal(y2, (-0.5 * y1).astype(dtype))
    y3, _ = gpu_spmv(c, dtype, alpha=0.3)
    ym = km.segsum(c.row_ptr, c.col, c.val.astype(dtype), c.x.astype(dtype), S=8, TPW=3,
                   alpha=0.3, dtype=dtype)
    assert np.array_equal(y3, ym)


@pytest.mark.parametrize("tpv", [0, 2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_csr_vector_variant(tpv, dtype):
    for c in cases.adversarial(512, 3, seed=tpv)[::3]:
        say(f"csr-vector tpv={tpv} {c.name!r}")
        y, _ = gpu_spmv(c, dtype, variant=1, tpv=tpv)
        e, _ = rel_err(c, dtype, y)
        assert e <= TOL[dtype], (c.name, e)


def _full(name):
    if name == "uniform":
        A = matgen.uniform_65536()
        x = matgen.vector(101, A.n)
    elif name == "power-law":
        A = matgen.power_law_1m()
        x = matgen.vector(102, A.n)
    elif name == "laplace":
        A = matgen.laplacian_160()
        x = matgen.vector(103, A.n)
    else:
        A = matgen.block3_2m()
        x = matgen.vector(104, A.n)
    return A, x


@pytest.mark.parametrize("n", [33, 131072, 131073, 262144])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_x_read_ahead_sizes(n, dtype):
    """x at and around the largest size the kernels read ahead into L2 (1 MB:
    131,072 doubles or 262,144 floats), and a tiny one; the first and the last
    column are in use.  y within the oracle's bound for csr-vector and the
    segmented sum."""
    rng = np.random.default_rng(n)
    m = 5000
    lens = rng.integers(1, 17, m)
    rp = np.zeros(m + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = rng.integers(0, n, int(rp[-1])).astype(np.int32)
    for i in range(m):
        col[rp[i]:rp[i + 1]] = np.sort(col[rp[i]:rp[i + 1]])
    col[0] = 0
    col[-1] = n - 1
    c = cases.Case(f"read-ahead n={n}", m, n, rp.astype(np.int32), col,
                   rng.standard_normal(int(rp[-1])), rng.standard_normal(n))
    for variant, tpv, cfg in ((1, 8, (8, 4, 3)), (0, 0, (8, 1, 1)), (0, 0, (4, 4, 2))):
        say(f"read-ahead n={n} {np.dtype(dtype).name} variant={variant} cfg={cfg}")
        y, _ = gpu_spmv(c, dtype, cfg=cfg, variant=variant, tpv=tpv)
        e, _ = rel_err(c, dtype, y)
        assert e <= TOL[dtype], (n, variant, cfg, e)


@pytest.mark.parametrize("name,dtype", [("uniform", np.float64), ("uniform", np.float32),
                                        ("power-law", np.float64), ("power-law", np.float32),
                                        ("laplace", np.float64), ("block3", np.float64)])
def test_benchmark_matrices(name, dtype):
    A, x = _full(name)
    c = cases.Case(name, A.m, A.n, A.row_ptr, A.col, A.val, x)
    x0 = np.ones(A.n, dtype)
    for cfg in ((8, 4, 3), (16, 4, 2)):
        say(f"full {name} {np.dtype(dtype).name} cfg={cfg} nnz={A.nnz}")
        op = make_op(dtype, *cfg)
        y = np.zeros(A.m, dtype)
        _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val.astype(dtype), x0, y),
                   "prepare_mem")
        _lib.check(op.spmv(1.0), "spmv")          # leaves A*1 in y ...
        _lib.check(op.set_x(x.astype(dtype)), "set_x")
        _lib.check(op.spmv(1.0), "spmv")          # ... which every row must overwrite
        _lib.check(op.get_y(), "get_y")
        e, _ = rel_err(c, dtype, y)
        say(f"  max rel err {e:.3e}")
        assert e <= TOL[dtype], e
        op.close()


def test_invalid_matrices_rejected():
    op = make_op(np.float64)
    y = np.zeros(3)
    ci = np.array([0, 1, 2], np.int32)
    v = np.ones(3)
    x = np.ones(3)
    assert op.prepare_mem(3, 3, 3, np.array([0, 2, 1, 3], np.int32), ci, v, x, y) == \
        _lib.BHSPARSE_INVALID_MATRIX
    assert op.prepare_mem(3, 3, 3, np.array([0, 1, 2, 3], np.int32),
                          np.array([0, 1, 3], np.int32), v, x, y) == _lib.BHSPARSE_INVALID_MATRIX
    assert op.prepare_mem(3, 3, 3, np.array([0, 1, 2, 2], np.int32), ci, v, x, y) == \
        _lib.BHSPARSE_INVALID_MATRIX
    assert op.prepare_mem(3, 3, 3, np.array([0, 1, 2, 3], np.int32), ci, v, x, y) == 0


def test_row_features():
    A = matgen.power_law(21, 40000, 40000, 2.2, 10000, 0.2, 0)
    op = make_op(np.float32)
    y = np.zeros(A.m, np.float32)
    _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val.astype(np.float32),
                              np.ones(A.n, np.float32), y), "prepare_mem")
    f = op.row_features()
    L = np.diff(A.row_ptr).astype(np.float64)
    ne = L > 0
    span = (A.col[A.row_ptr[1:][ne] - 1] - A.col[A.row_ptr[:-1][ne]] + 1).astype(np.float64)
    want = [A.m, A.nnz, L.mean(), L.var(), L.max(), 1 - ne.mean(), span.mean(), 0.0]
    np.testing.assert_allclose(f, want, rtol=1e-12)


def test_bench_entry_point():
    A = matgen.uniform_rows(5, 20000, 20000, 0, 32)
    op = make_op(np.float64)
    y = np.zeros(A.m)
    _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val, matgen.vector(5, A.n), y),
               "prepare_mem")
    t, k = op.bench(1.0, warmup=2, reps=5, flush_bytes=64 << 20)
    # whole calls and main kernels come from separate passes of cold calls
    assert len(t) == 5 and np.all(t > 0) and np.all(k > 0)
    assert np.median(k) <= np.median(t) * 1.05 + 0.002
    ov = op.bench_event_overhead(reps=5, flush_bytes=64 << 20)
    assert len(ov) == 5 and np.all(ov >= 0) and np.median(ov) < np.median(k)


@pytest.mark.parametrize("x0", [np.inf, np.nan])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_non_finite_x0_stays_in_its_rows(x0, dtype):
    # items past nnz in the last tile read col 0 / val 0; their products are
    # masked, so an infinite or NaN x[0] reaches only the rows that hold column 0
    # (fast-track tiles: the long rows; general and dirty tiles: the rest)
    rng = np.random.default_rng(5)
    lens = np.concatenate([rng.integers(0, 9, 3000), [4000, 2500], rng.integers(0, 5, 777)])
    c = cases.from_lengths("x0 non-finite", lens, n=6000, seed=9)
    assert c.nnz % 512 != 0 and c.nnz % 256 != 0
    c.x = c.x.copy()
    c.x[0] = x0
    has0 = np.array([0 in c.col[c.row_ptr[i]:c.row_ptr[i + 1]] for i in range(c.m)])
    assert has0.any() and (~has0).any()
    v = c.val.astype(dtype).astype(np.float64)
    xx = c.x.astype(dtype).astype(np.float64)
    yr = oracle.spmv_serial(c.row_ptr, c.col, v, xx)
    for variant, cfg, tpv in [(0, (8, 4, 3), 0), (0, (4, 1, 1), 0), (0, (16, 2, 5), 0),
                              (2, (8, 1, 4), 4), (1, (8, 4, 1), 8)]:
        say(f"x[0]={x0} variant={variant} cfg={cfg} {np.dtype(dtype).name}")
        y, op = gpu_spmv(c, dtype, cfg, variant=variant, tpv=tpv)
        op.close()
        assert np.all(np.isfinite(y[~has0])), (variant, cfg)
        assert not np.any(np.isfinite(y[has0])), (variant, cfg)
        _, err = oracle.rel_error(c.row_ptr, c.col, v, xx, y, yr, per_row=True)
        assert np.max(err[~has0]) <= TOL[dtype], (variant, cfg)

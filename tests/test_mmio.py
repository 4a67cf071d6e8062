"""Matrix Market reader (include/bhs_mmio.h) against fixtures whose expected
dense matrices were built from the entry lists (tests/golden/make_golden.py),
its reference-layout mode against the reference's counting sort
(spmv_cuda/main.cu:619-710, restated below for the layout only), its error
codes, and a write/read round trip."""
import json
import os

import numpy as np
import pytest

from benchmark_spmv_using_csr_amd import matgen
from benchmark_spmv_using_csr_amd._lib import BhsError

MTX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mtx")
EXPECTED = json.load(open(os.path.join(MTX, "expected.json")))
REJECT = json.load(open(os.path.join(MTX, "reject.json")))


def csr_ok(A):
    assert A.row_ptr[0] == 0 and np.all(np.diff(A.row_ptr) >= 0)
    assert A.nnz == len(A.col) == len(A.val)
    assert np.all((A.col >= 0) & (A.col < A.n))


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_sorted_merged_read_matches_expected(name):
    A = matgen.read_mtx(os.path.join(MTX, name))
    e = EXPECTED[name]
    assert (A.m, A.n) == (e["m"], e["n"])
    csr_ok(A)
    for i in range(A.m):
        c = A.col[A.row_ptr[i]:A.row_ptr[i + 1]]
        assert np.all(np.diff(c) > 0), (name, i)          # sorted, no duplicates
    np.testing.assert_array_equal(A.dense(), np.array(e["dense"]))


def reference_layout(path):
    """Entries of every row in the order the reference's counting sort puts
    them (main.cu:653-710): file order, each mirrored entry right after its
    source entry's turn in the other row; duplicates kept."""
    lines = [l.split() for l in open(path) if l.strip() and not l.startswith("%")]
    banner = open(path).readline().lower().split()
    field, sym = banner[3], banner[4]
    m, n, _ = (int(v) for v in lines[0])
    rows = [[] for _ in range(m)]
    for t in lines[1:]:
        i, j = int(t[0]) - 1, int(t[1]) - 1
        v = 1.0 if field == "pattern" else float(t[2])
        rows[i].append((j, v))
        if sym in ("symmetric", "hermitian", "skew-symmetric") and i != j:
            rows[j].append((i, -v if sym == "skew-symmetric" else v))
    return rows


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_reference_layout_mode(name):
    path = os.path.join(MTX, name)
    A = matgen.read_mtx(path, sort=False, merge=False)
    csr_ok(A)
    want = reference_layout(path)
    for i in range(A.m):
        got = list(zip(A.col[A.row_ptr[i]:A.row_ptr[i + 1]].tolist(),
                       A.val[A.row_ptr[i]:A.row_ptr[i + 1]].tolist()))
        assert got == want[i], (name, i)
    np.testing.assert_array_equal(A.dense(), np.array(EXPECTED[name]["dense"]))


@pytest.mark.parametrize("name", sorted(REJECT))
def test_rejected_files(name):
    with pytest.raises(BhsError) as ei:
        matgen.read_mtx(os.path.join(MTX, name))
    assert ei.value.code == REJECT[name], name


def test_missing_file_and_bad_flags():
    with pytest.raises(BhsError) as ei:
        matgen.read_mtx(os.path.join(MTX, "does_not_exist.mtx"))
    assert ei.value.code == -3
    with pytest.raises(BhsError):
        matgen.read_mtx(os.path.join(MTX, "general_real.mtx"), sort=False, merge=True)


def test_write_read_round_trip(tmp_path):
    A = matgen.power_law(5, 3000, 2500, 2.0, 400, 0.1, 0)
    p = tmp_path / "pl.mtx"
    matgen.write_mtx(A, p)
    B = matgen.read_mtx(p)
    assert (B.m, B.n) == (A.m, A.n)
    np.testing.assert_array_equal(B.row_ptr, A.row_ptr)
    np.testing.assert_array_equal(B.col, A.col)
    np.testing.assert_array_equal(B.val, A.val)


@pytest.mark.gpu
def test_read_file_through_the_library():
    import oracle
    from benchmark_spmv_using_csr_amd import _lib
    from benchmark_spmv_using_csr_amd.spmv import bhsparse_spmv
    for name in sorted(EXPECTED):
        A = matgen.read_mtx(os.path.join(MTX, name))
        x = matgen.vector(4, A.n)
        op = bhsparse_spmv(np.float64)
        _lib.check(op.init_platform(0), "init_platform")
        y = np.zeros(A.m)
        _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val, x, y), "prepare_mem")
        _lib.check(op.spmv(1.0), "spmv")
        _lib.check(op.get_y(), "get_y")
        yr = oracle.spmv_serial(A.row_ptr, A.col, A.val, x)
        assert oracle.rel_error(A.row_ptr, A.col, A.val, x, y, yr) <= 1e-13, name
        np.testing.assert_allclose(y, np.array(EXPECTED[name]["dense"]) @ x, rtol=0, atol=1e-12)
        op.close()

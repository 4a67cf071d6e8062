"""Seeded workload generators: determinism and the summary statistics of
every benchmark configuration.  CPU only."""
import numpy as np
import pytest

from benchmark_spmv_using_csr_amd import matgen


def _check_csr(A, sample=4000):
    rp = A.row_ptr
    assert rp[0] == 0 and rp[-1] == A.nnz and np.all(np.diff(rp) >= 0)
    assert A.col.min(initial=0) >= 0 and A.col.max(initial=0) < max(A.n, 1)
    step = max(1, A.m // sample)
    for i in range(0, A.m, step):
        c = A.col[rp[i]:rp[i + 1]]
        assert np.all(np.diff(c) > 0)          # distinct and sorted
    if A.nnz:
        assert np.abs(A.val).max() <= 1.0


def test_determinism():
    a = matgen.uniform_rows(7, 3000, 3000, 0, 32)
    b = matgen.uniform_rows(7, 3000, 3000, 0, 32)
    c = matgen.uniform_rows(8, 3000, 3000, 0, 32)
    assert np.array_equal(a.row_ptr, b.row_ptr) and np.array_equal(a.col, b.col)
    assert np.array_equal(a.val, b.val)
    assert not np.array_equal(a.row_ptr, c.row_ptr)
    assert np.array_equal(matgen.vector(3, 9999), matgen.vector(3, 9999))


def test_uniform_65536():
    A = matgen.uniform_65536()
    _check_csr(A)
    L = A.row_lengths()
    assert A.m == A.n == 65536 and L.min() == 0 and L.max() == 32
    assert abs(A.nnz - 65536 * 16) < 0.01 * 65536 * 16
    assert 0.02 < (L == 0).mean() < 0.045


def test_power_law_1m():
    A = matgen.power_law_1m()
    _check_csr(A, 2000)
    L = A.row_lengths()
    assert A.m == A.n == 1 << 20
    assert abs(A.nnz - 10_000_000) < 0.005 * 10_000_000
    assert (L == 0).sum() == int(0.1 * A.m)
    assert L.max() > 50_000                     # long rows span many tiles and waves


def test_laplacian_160():
    A = matgen.laplacian_160()
    assert A.m == 4_096_000 and A.nnz == 28_518_400
    assert np.array_equal(np.unique(A.val), [-1.0, 6.0])


def test_block3_small_and_shape():
    A = matgen.block3(3, 2000, 9, 5000)
    _check_csr(A)
    assert A.m == 2000 and A.row_lengths().max() <= 27
    # the full-size recipe: 666,667 block rows, the last one truncated to 2 rows
    assert (2_000_000 + 2) // 3 == 666_667 and 3 * 666_666 + 2 == 2_000_000


@pytest.mark.parametrize("fam", range(4))
def test_families(fam):
    A = matgen.family_matrix(100 + fam, 1e3, 2e4, fam)
    _check_csr(A)
    assert A.m >= 1000


def test_named_workloads_are_the_documented_matrices():
    """The benchmark matrices by name (bench.py --workload, the profile driver):
    sizes as the workload descriptions state, and x of length n."""
    import pytest
    from benchmark_spmv_using_csr_amd import matgen
    A, x = matgen.workload("uniform")
    assert (A.m, A.n) == (65536, 65536) and len(x) == A.n
    A, x = matgen.workload("laplace")
    assert A.m == 160 ** 3 and A.nnz == 7 * 160 ** 3 - 6 * 160 ** 2 and len(x) == A.n
    with pytest.raises(ValueError):
        matgen.workload("nope")

"""The rocSPARSE comparison column computes the same y: every algorithm it is
timed with matches the FP64 oracle within the bound the library itself is held
to, so the bench compares like with like."""
import numpy as np
import pytest

import oracle
from benchmark_spmv_using_csr_amd import matgen, vendor

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-13, np.float32: 1e-5}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("alg", sorted(vendor.ALGS))
def test_rocsparse_matches_oracle(alg, dtype):
    print(f"[gpu] rocsparse {alg} {np.dtype(dtype).name}", flush=True)
    for A in (matgen.power_law(7, 20000, 20000, 2.0, 5000, 0.1, 0),
              matgen.uniform_rows(8, 3000, 5000, 0, 40)):
        x = matgen.vector(9, A.n)
        ms, y, an = vendor.csrmv(A, x.astype(dtype), dtype, alg=alg, warmup=1, reps=3)
        assert np.all(ms > 0) and an >= 0
        v64 = A.val.astype(dtype).astype(np.float64)
        x64 = x.astype(dtype).astype(np.float64)
        yr = oracle.spmv_serial(A.row_ptr, A.col, v64, x64)
        assert oracle.rel_error(A.row_ptr, A.col, v64, x64, y, yr) <= TOL[dtype], (alg, A.name)

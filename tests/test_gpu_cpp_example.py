"""The C++ drop-in (include/bhsparse_spmv.hpp) end to end: the reference
driver's own call sequence (new bhsparse_spmv_cuda, init_platform, prepare_mem,
run_benchmark, get_y, free_platform, free_mem; main.cu:490-542) on its 6x6
example, compiled against the library by the build."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "benchmark_spmv_using_csr_amd", "examples", "spmv_small")


def test_reference_call_sequence_in_cpp():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing: run __graft_entry__.build() first")
    print("[gpu] C++ example", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "y = [6, 39, 21, 0, 12, 42] PASS" in r.stdout, r.stdout

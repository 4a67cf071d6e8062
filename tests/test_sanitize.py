"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

Builds tests/sanitize/host_checks.cpp with the Matrix Market reader, the
matrix generators and the CPU oracle, all instrumented, and runs it: every
fixture file (malformed ones included), the generators, the oracle's entry
points, and allocation failure under an injected cap (the reader must return
BHSPARSE_OUT_OF_MEMORY, the generators nullptr).  No GPU code is involved."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1", "-fopenmp"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                    reason="needs gcc/g++")
def test_host_code_under_asan_ubsan(tmp_path):
    csrc = os.path.join(ROOT, "benchmark_spmv_using_csr_amd", "csrc")
    oracle_o = str(tmp_path / "oracle.o")
    subprocess.run(["gcc", *SAN, "-std=c11", "-ffp-contract=off", "-c",
                    os.path.join(ROOT, "oracle", "oracle.c"), "-o", oracle_o], check=True)
    exe = str(tmp_path / "host_checks")
    subprocess.run(["g++", *SAN, "-std=c++17",
                    os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp"),
                    os.path.join(csrc, "mmio.cpp"), os.path.join(csrc, "matgen.cpp"),
                    oracle_o, "-lm", "-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "mtx"), str(tmp_path)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "host checks: ok" in r.stdout

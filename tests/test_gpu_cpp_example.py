"""The C++ drop-in (include/bhsparse_spmv.hpp) end to end: the reference
driver's own call sequence (new bhsparse_spmv_cuda, init_platform, prepare_mem,
run_benchmark, get_y, free_platform, free_mem; main.cu:490-542) on its 6x6
example, compiled against the library by the build, once as the default
build and once with -DUSE_SVM_ALWAYS=1 (the reference's unified-memory build,
spmv_cuda/Makefile:6), which puts A, x and y in unified memory."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXAMPLES = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "benchmark_spmv_using_csr_amd", "examples")


# the unified-memory build uses managed memory where the device has it and
# host-mapped memory where it does not
@pytest.mark.parametrize("name,memory", [("spmv_small", ("device",)),
                                         ("spmv_small_svm", ("managed", "host-mapped"))])
def test_reference_call_sequence_in_cpp(name, memory):
    exe = os.path.join(EXAMPLES, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: run __graft_entry__.build() first")
    print(f"[gpu] C++ example {name}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert any(f"memory: {md}\n" in r.stdout for md in memory), r.stdout
    assert "y = [6, 39, 21, 0, 12, 42] PASS" in r.stdout, r.stdout
    # and the OpenCL driver's sequence (init_kernels, get_y checked, sync_device)
    assert "opencl sequence: y = [6, 39, 21, 0, 12, 42] PASS" in r.stdout, r.stdout


def test_spmv_async_replays_from_a_captured_hipgraph():
    """bhs_d_spmv_async captured into a hipGraph (examples/graph_capture.cpp):
    four replays per variant, with eager calls on the same handle between them,
    each bitwise equal to the eager result for the x it reads."""
    exe = os.path.join(EXAMPLES, "graph_capture")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: run __graft_entry__.build() first")
    print("[gpu] C++ graph capture", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("PASS"), r.stdout
    assert r.stdout.count("match the eager result") == 4, r.stdout

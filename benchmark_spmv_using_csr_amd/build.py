"""Build recipes for the native libraries (run by ``__graft_entry__.build()``).

* ``libbhsparse_spmv.so`` -- HIP kernels + the C ABI of include/bhsparse_spmv.h,
  compiled for gfx950 with hipcc.
* ``libbhs_host.so``      -- host-side C++ (seeded matrix generators).

Both land next to this file (in-tree, so they travel with the repository).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
ROOT = os.path.dirname(PKG)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["bhs_api.hip", "bhs_mlp.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",            # products are rounded before they are summed
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} ... -o {cmd[-1] if cmd else ''}")
    return r.stdout


def build_hip(jobs: int = 4) -> str:
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(PKG, "libbhsparse_spmv.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_vendor() -> str:
    """libbhs_vendor.so: the rocSPARSE comparison column (its own library, so the
    drop-in library does not depend on rocSPARSE)."""
    src = os.path.join(CSRC, "bhs_vendor.hip")
    obj = os.path.join(CSRC, "bhs_vendor.o")
    _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
    out = os.path.join(PKG, "libbhs_vendor.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", obj, "-lrocsparse", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_host() -> str:
    out = os.path.join(PKG, "libbhs_host.so")
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall",
          os.path.join(CSRC, "matgen.cpp"), os.path.join(CSRC, "mmio.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_example() -> str:
    """The C++ drop-in example (reference's `./spmv 0`) against the built
    library, as the default build and as the unified-memory build
    (-DUSE_SVM_ALWAYS=1, the reference's spmv_cuda/Makefile:6 switch)."""
    for name, defs in (("spmv_small", []), ("spmv_small_svm", ["-DUSE_SVM_ALWAYS=1"])):
        out = os.path.join(PKG, "examples", name)
        tmp = out + ".tmp"
        _run(["g++", "-O2", "-std=c++17", *defs, f"-I{os.path.join(ROOT, 'include')}",
              os.path.join(PKG, "examples", "spmv_small.cpp"), f"-L{PKG}", "-lbhsparse_spmv",
              f"-Wl,-rpath,{PKG}", "-o", tmp])
        os.replace(tmp, out)
    # hipGraph capture of spmv_async (host code; needs the HIP runtime headers)
    out = os.path.join(PKG, "examples", "graph_capture")
    tmp = out + ".tmp"
    _run([HIPCC, "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}",
          os.path.join(PKG, "examples", "graph_capture.cpp"), f"-L{PKG}", "-lbhsparse_spmv",
          f"-Wl,-rpath,{PKG}", "-o", tmp])
    os.replace(tmp, out)
    return os.path.join(PKG, "examples", "spmv_small")


def build_all() -> None:
    with ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(build_hip), ex.submit(build_host), ex.submit(build_vendor)]
        for f in futs:
            f.result()
    build_example()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(PKG, "libbhsparse_spmv.so"), os.path.join(PKG, "libbhs_host.so"))

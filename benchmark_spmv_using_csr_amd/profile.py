"""Profiling driver: runs a fixed number of SpMV calls on one workload (and a
copy of known size for counter calibration) so that rocprofv3 can be pointed
at it, e.g.

  rocprofv3 --kernel-trace --stats -f csv -d OUT -- \\
      python3 -m benchmark_spmv_using_csr_amd.profile --workload power-law --dtype f64
  rocprofv3 --pmc FETCH_SIZE -f csv -d OUT -- python3 -m benchmark_spmv_using_csr_amd.profile ...

Each call is preceded by the 1 GiB flush write (cold caches) unless --flush-mb 0.
"""
from __future__ import annotations

import argparse
import sys

import numpy as np

from . import _lib, matgen
from .spmv import bhsparse_spmv, stream_copy


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="power-law", choices=list(matgen.WORKLOADS))
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32"])
    ap.add_argument("--config", default="", help="S,waves_per_block,tiles_per_wave")
    ap.add_argument("--variant", type=int, default=0,
                    help="0 segmented sum, 1 csr-vector, 2 segmented sum over column panels")
    ap.add_argument("--tpv", type=int, default=0,
                    help="csr-vector lanes per row (variant 1) or column panels (variant 2)")
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--flush-mb", type=int, default=1024)
    ap.add_argument("--copy-mb", type=int, default=256)
    ap.add_argument("--floor", action="store_true",
                    help="also time the matrix's data-movement floor (floor_kernel) --reps times")
    ap.add_argument("--no-spmv", action="store_true", help="skip the SpMV calls (with --floor)")
    ap.add_argument("--batched", default="",
                    help="S,T,waves_per_block,grid,order[,x_stage]: run the 256-matrix batch "
                         "(fp64) with this schedule --reps times instead of the single-matrix "
                         "workload")
    a = ap.parse_args(argv)
    if a.batched:
        from . import batched
        mats, ks, Xs = batched.batch_set()
        b = batched.SpmvBatch(mats, ks, np.float64)
        b.set_schedule(*(int(v) for v in a.batched.split(",")))
        b.set_x(Xs)
        ms = b.bench(plain=False, warmup=1, reps=a.reps)
        print(f"batched {len(mats)} matrices, {sum(ks)} right-hand sides, schedule {a.batched}: "
              f"median {np.median(ms):.4f} ms over {a.reps} graph replays", flush=True)
        b.close()
        if a.copy_mb > 0:                  # the counters' calibration, in the same process
            cm = stream_copy(a.copy_mb << 20, reps=a.reps)
            print(f"copy {a.copy_mb} MiB: median {np.median(cm) * 1e3:.2f} us", flush=True)
        return 0
    dtype = np.float64 if a.dtype == "f64" else np.float32
    A, x = matgen.workload(a.workload)
    op = bhsparse_spmv(dtype)
    _lib.check(op.init_platform(0), "init_platform")
    if a.config:
        S, wpb, tpw = (int(v) for v in a.config.split(","))
        _lib.check(op.set_config(a.variant, S, wpb, tpw, a.tpv), "set_config")
    y = np.zeros(A.m, dtype)
    _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val.astype(dtype),
                              x.astype(dtype), y), "prepare_mem")
    if not a.no_spmv:
        t, k = op.bench(1.0, warmup=1, reps=a.reps, flush_bytes=a.flush_mb << 20)
        print(f"spmv {A.name} {a.dtype}: median {np.median(t) * 1e3:.2f} us, main kernel mean "
              f"{np.mean(k) * 1e3:.2f} us over {a.reps} calls", flush=True)
    if a.floor:
        f = op.bench_floor(reps=a.reps, flush_bytes=a.flush_mb << 20)
        print(f"floor {A.name} {a.dtype}: median {np.median(f) * 1e3:.2f} us over {a.reps} calls",
              flush=True)
    op.close()
    if a.copy_mb > 0:
        ms = stream_copy(a.copy_mb << 20, reps=a.reps)
        gbps = 2 * (a.copy_mb << 20) / (np.median(ms) * 1e-3) / 1e9
        print(f"copy {a.copy_mb} MiB: median {np.median(ms) * 1e3:.2f} us = {gbps:.1f} GB/s "
              f"(read+write)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

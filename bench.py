#!/usr/bin/env python3
"""Benchmark: effective GB/s of the speculative segmented-sum CSR SpMV on one
MI355X, on the power-law workload (m = n = 2^20, P(k) ~ k^-2 on [1, 1e5],
10% empty rows, ~10M nonzeros, values and x uniform in [-1, 1], FP64).

One step = one y = A*x (main kernel + calibration kernel) with A and x already
resident in HBM.  L2 and the MALL are flushed before every step (a 1 GiB
write outside the timed events), so every step starts cold.  Steps are timed
with HIP events on the handle's stream inside libbhsparse_spmv.so; the whole
timed loop is also bracketed by device synchronisation and host wall clock.

Effective bytes = nnz*(sizeof(val)+4) + 4*(m+1) + n*sizeof(val) + m*sizeof(val).
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from benchmark_spmv_using_csr_amd import _lib, matgen  # noqa: E402
from benchmark_spmv_using_csr_amd.spmv import (bhsparse_spmv, device_synchronize,  # noqa: E402
                                               stream_copy)

HBM_PEAK_GBPS = 8000.0
# the tile comes from the learned tuner's shortlist: its TUNER_TOPK best-predicted
# classes are timed and the fastest is kept
TUNER_TOPK = 5
METRIC = "effective GB/s (% HBM peak) CSR SpMV; tuned+batched vs plain speedup at equal err"


def eff_bytes(m, n, nnz, vsize, k=1):
    return nnz * (vsize + 4) + 4 * (m + 1) + k * (n + m) * vsize


def load_traffic(name, cfg, key="config"):
    """HBM bytes per launch of the main kernel from the committed rocprofv3
    --pmc summary for this workload and tile (FETCH_SIZE + WRITE_SIZE, scaled by
    the copy-kernel calibration recorded with it); None if absent.  key: the
    entry field that names the tile ("config", or "schedule" for the batch)."""
    p = os.path.join(ROOT, "profiles", name)
    if not name or not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            d = json.load(f)
        for e in d.get("entries", [d]):
            if tuple(e.get(key, ())) == tuple(cfg):
                return float(e["hbm_bytes_per_launch"])
        return None
    except Exception:
        return None


def median(a):
    return float(np.median(a))


def run_gpu(A, x, dtype, cfg, warmup, steps, flush_bytes, memory=0):
    """Times the workload with the learned tuner's tile (or `cfg` when given),
    then the fixed default tile and the exhaustive autotuner's best.  memory:
    0 = device memory, 1 = managed memory, 2 = host-mapped memory (the last two
    are USE_SVM_ALWAYS=1, spmv_cuda/Makefile:6)."""
    from benchmark_spmv_using_csr_amd import tuner

    op = bhsparse_spmv(dtype)
    _lib.check(op.set_memory_mode(memory), "set_memory_mode")
    _lib.check(op.init_platform(0), "init_platform")
    y = np.zeros(A.m, dtype)
    t0 = time.perf_counter()
    _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val.astype(dtype),
                              x.astype(dtype), y), "prepare_mem")
    t_init = time.perf_counter() - t0
    top1 = None
    if cfg is None:
        net = tuner.make_net(tuner.load_net())
        top1 = op.apply_tuner(net, tuner.CLASSES)          # the network's own pick
        t0 = time.perf_counter()
        chosen = op.apply_tuner(net, tuner.CLASSES, topk=TUNER_TOPK, reps=7,
                                flush_bytes=flush_bytes)
        t_tune = time.perf_counter() - t0
        cfg = tuner.CLASSES[chosen]
        source = (f"learned tuner (data/tuner_net.json): its {TUNER_TOPK} best-predicted "
                  "classes timed, fastest kept")
    else:
        _lib.check(op.set_config(*cfg), "set_config")
        t_tune = 0.0
        chosen = tuner.CLASSES.index(cfg) if cfg in tuner.CLASSES else None
        source = "command line"
    device_synchronize()
    w0 = time.perf_counter()
    ms_total, ms_main = op.bench(1.0, warmup=warmup, reps=steps, flush_bytes=flush_bytes)
    device_synchronize()
    wall = time.perf_counter() - w0
    _lib.check(op.get_y(), "get_y")
    reps2 = max(10, steps // 2)
    wt, wm = op.bench(1.0, warmup=3, reps=reps2, flush_bytes=0)
    fl = op.bench_floor(reps=reps2, flush_bytes=flush_bytes)
    ov = op.bench_event_overhead(reps=reps2, flush_bytes=flush_bytes)
    # the method without the tuner, and the exhaustive autotuner
    _lib.check(op.set_config(*tuner.CLASSES[tuner.DEFAULT_CLASS]), "set_config")
    dt, _ = op.bench(1.0, warmup=3, reps=reps2, flush_bytes=flush_bytes)
    us, best = op.autotune(tuner.CLASSES, reps=9, flush_bytes=flush_bytes)
    # the plain (speculative) layout at its fastest tile of the sweep: the
    # reference's own layout, where empty rows make dirty tiles and the
    # calibration adds cross-wave partials
    plain = [i for i, c in enumerate(tuner.CLASSES) if c[0] == 0 and us[i] > 0]
    pcfg = tuner.CLASSES[min(plain, key=lambda i: us[i])]
    _lib.check(op.set_config(*pcfg), "set_config")
    p_total, p_main = op.bench(1.0, warmup=3, reps=reps2, flush_bytes=flush_bytes)
    p_stats = op.stats()
    op.close()
    # one-time cost of the chosen layout: a fresh handle, the upload, then the
    # set_config that builds it (column panels: the stacked copy of A)
    layout_ms = None
    if cfg[0] == 2:
        op2 = bhsparse_spmv(dtype)
        _lib.check(op2.set_memory_mode(memory), "set_memory_mode")
        _lib.check(op2.init_platform(0), "init_platform")
        y2 = np.zeros(A.m, dtype)
        _lib.check(op2.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val.astype(dtype),
                                   x.astype(dtype), y2), "prepare_mem")
        device_synchronize()
        t0 = time.perf_counter()
        _lib.check(op2.set_config(*cfg), "set_config")
        device_synchronize()
        layout_ms = (time.perf_counter() - t0) * 1e3
        op2.close()
    return {
        "plain_cfg": pcfg, "plain_total": p_total, "plain_main": p_main, "plain_stats": p_stats,
        "layout_ms": layout_ms,
        "cfg": cfg, "cfg_source": source, "tune_s": t_tune, "chosen": chosen, "top1": top1,
        "ms_total": ms_total, "ms_main": ms_main, "wall_s": wall, "init_s": t_init,
        "warm_total": wt, "warm_main": wm, "floor": fl, "event_overhead": ov,
        "default_ms": median(dt),
        "sweep_us": us, "sweep_best": best, "y": y,
    }


def run_batched():
    """256 uneven matrices x {1, 4, 16} right-hand sides: per-matrix plain
    launches vs one persistent launch replayed as a hipGraph, schedule from the
    batch autotuner."""
    from benchmark_spmv_using_csr_amd import batched

    t0 = time.perf_counter()
    mats, ks, Xs = batched.batch_set()
    b = batched.SpmvBatch(mats, ks, np.float64)
    b.set_x(Xs)
    setup = time.perf_counter() - t0
    plain = b.bench(plain=True, warmup=1, reps=5)
    ms, best = b.autotune(batched.SCHEDULES, reps=5)
    sched = batched.SCHEDULES[best]
    bt = b.bench(plain=False, warmup=2, reps=10)
    # the same plain launches replayed from one hipGraph: what is left of the
    # speedup without the per-launch overhead
    pg = b.bench(plain=2, warmup=1, reps=5)
    b.run(1.0)
    yb = b.y()
    b.run_plain(1.0)
    yp = b.y(plain=True)
    identical = all(np.array_equal(u, v) for u, v in zip(yb, yp))
    nnz = sum(A.nnz for A in mats)
    byts = sum(eff_bytes(A.m, A.n, A.nnz, 8, k) for A, k in zip(mats, ks))
    b.close()
    ach = byts / (median(bt) * 1e-3) / 1e9
    return {
        "workload": "256 matrices, m log-uniform [1e3, 2e5], 4 families, k in {1,4,16}, fp64",
        "n_matrices": len(mats), "nnz": int(nnz), "rhs_total": int(sum(ks)),
        "plain_ms": round(median(plain), 4), "batched_ms": round(median(bt), 4),
        "speedup": round(median(plain) / median(bt), 3),
        "plain_graph_ms": round(median(pg), 4),
        "speedup_vs_plain_graph": round(median(pg) / median(bt), 3),
        "batched_effective_gbps": round(ach, 1),
        "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBPS, 4),
                     "traffic": load_traffic("batched_256mat_f64_traffic.json", sched,
                                             key="schedule"),
                     "definition": "effective bytes of the batch (A once, x and y per "
                                   "right-hand side) / median graph replay (queue reset, "
                                   "batched_kernel, calibration)"},
        "schedule": {"items_per_lane": sched[0], "tiles_per_wave": sched[1],
                     "waves_per_block": sched[2], "grid_blocks": sched[3],
                     "order": "matrix" if sched[4] == 0 else "most work first",
                     "x_stage": sched[5]},
        "batched_equals_plain_bitwise": bool(identical), "setup_s": round(setup, 2),
    }


OTHER_WORKLOADS = (("power-law", "f32"), ("laplace", "f64"), ("block3", "f64"),
                   ("uniform", "f64"))


def run_others(flush_bytes, vendor_cols, reps=30):
    """The other benchmark matrices (and fp32 on the power-law one), each with
    the learned tuner's shortlisted tile, timed cold like the headline, beside
    rocSPARSE's best algorithm timed the same way."""
    from benchmark_spmv_using_csr_amd import tuner

    net = tuner.make_net(tuner.load_net())
    rows = []
    for wl, dt in OTHER_WORKLOADS:
        dtype = np.float64 if dt == "f64" else np.float32
        A, x = matgen.workload(wl)
        op = bhsparse_spmv(dtype)
        _lib.check(op.init_platform(0), "init_platform")
        y = np.zeros(A.m, dtype)
        _lib.check(op.prepare_mem(A.m, A.n, A.nnz, A.row_ptr, A.col, A.val.astype(dtype),
                                  x.astype(dtype), y), "prepare_mem")
        top1 = op.apply_tuner(net, tuner.CLASSES)          # the network's own pick
        chosen = op.apply_tuner(net, tuner.CLASSES, topk=TUNER_TOPK, reps=7,
                                flush_bytes=flush_bytes)
        cfg = tuner.CLASSES[chosen]
        tot, main_ms = op.bench(1.0, warmup=5, reps=reps, flush_bytes=flush_bytes)
        _lib.check(op.get_y(), "get_y")
        # the exhaustive autotuner as the yardstick of both picks
        us, best = op.autotune(tuner.CLASSES, reps=7, flush_bytes=flush_bytes)
        op.close()
        b = eff_bytes(A.m, A.n, A.nnz, np.dtype(dtype).itemsize)
        med, km = median(tot), float(np.mean(main_ms))
        row = {"workload": A.name, "dtype": dt, "m": A.m, "nnz": A.nnz,
               "tile": tuner.class_name(cfg), "ms_per_step": round(med, 5),
               "gbps": round(b / (med * 1e-3) / 1e9, 1),
               "kernel_mean_us": round(km * 1e3, 3),
               "roofline_frac": round(b / (km * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
               "whole_call_frac": round(b / (med * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
               "tuner": {"network_top1": tuner.class_name(tuner.CLASSES[top1]),
                         "network_top1_regret_vs_exhaustive":
                             round(float(us[top1]) / float(us[best]) - 1, 4),
                         "exhaustive_best": tuner.class_name(tuner.CLASSES[best]),
                         "regret_vs_exhaustive":
                             round(float(us[chosen]) / float(us[best]) - 1, 4)}}
        if vendor_cols:
            v = run_vendor(A, x, dtype, flush_bytes, y, reps=10)
            row["vendor_best"] = {"alg": v["best_alg"], "cold_us": v["best_cold_us"],
                                  "max_abs_diff_vs_bhs_over_max_y":
                                      v["algs"][v["best_alg"]]["max_abs_diff_vs_bhs_over_max_y"]}
            row["bhs_speedup_vs_vendor_best"] = round(v["best_cold_us"] / (med * 1e3), 3)
        rows.append(row)
    return rows


def run_vendor(A, x, dtype, flush_bytes, y_bhs, reps=20):
    """rocSPARSE's CSR algorithms on the same matrix, timed the same way (cold,
    HIP events per call): the vendor column of the reference's comparison
    (cuSPARSE, main.cu:49-183)."""
    from benchmark_spmv_using_csr_amd import vendor

    b = eff_bytes(A.m, A.n, A.nnz, np.dtype(dtype).itemsize)
    out = {"library": f"rocSPARSE {vendor.version()} (rocsparse_v2_spmv, CSR, i32 indices)"}
    algs = {}
    scale = max(float(np.max(np.abs(y_bhs))), 1e-300)
    for alg in vendor.ALGS:
        ms, y, an = vendor.csrmv(A, x.astype(dtype), dtype, alg=alg, warmup=3, reps=reps,
                                 flush_bytes=flush_bytes)
        algs[alg] = {"cold_us": round(median(ms) * 1e3, 3),
                     "gbps": round(b / (median(ms) * 1e-3) / 1e9, 1),
                     "analysis_ms": round(an, 2),
                     "max_abs_diff_vs_bhs_over_max_y": float(np.max(np.abs(y - y_bhs)) / scale)}
    best = min(algs, key=lambda k: algs[k]["cold_us"])
    out.update({"algs": algs, "best_alg": best, "best_cold_us": algs[best]["cold_us"],
                "best_gbps": algs[best]["gbps"]})
    return out


def cpu_baseline(A, x, dtype, y_gpu, budget_s=12.0):
    """The reference's CPU paths timed on this host: the serial loop
    (main.cu:755-761) and the OpenMP row-parallel loop (main.cu:377-393) on all
    host cores, in the same precision and byte formula.  Also checks the GPU
    result against the FP64 serial oracle."""
    import oracle  # the CPU baseline leg is the only place bench.py loads the oracle

    v = A.val.astype(dtype)
    xx = x.astype(dtype)
    s1, _ = oracle.time_spmv(A.row_ptr, A.col, v, xx, 1, 3)
    ncores = os.cpu_count() or 1
    try:
        ncores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    so, used = oracle.time_spmv(A.row_ptr, A.col, v, xx, ncores, 5)
    reps = int(max(5, min(5000, budget_s / max(so, 1e-4))))
    so, used = oracle.time_spmv(A.row_ptr, A.col, v, xx, ncores, reps)
    b = eff_bytes(A.m, A.n, A.nnz, np.dtype(dtype).itemsize)
    v64, x64 = v.astype(np.float64), xx.astype(np.float64)
    yr = oracle.spmv_serial(A.row_ptr, A.col, v64, x64)
    err = oracle.rel_error(A.row_ptr, A.col, v64, x64, y_gpu, yr)
    return {
        "value": round(b / so / 1e9, 3), "unit": "GB/s", "cores": used, "kind": "port",
        "sample": f"{A.name}, {np.dtype(dtype).name}: OpenMP row loop (main.cu:377-393), "
                  f"median of {reps} calls on {used} threads; serial loop "
                  f"{b / s1 / 1e9:.3f} GB/s (median of 3 calls, 1 core)",
        "serial_gbps": round(b / s1 / 1e9, 3),
        "gpu_max_rel_err_vs_fp64_oracle": err,
        "host": {"nproc": os.cpu_count(), "affinity_cores": ncores, "cpu_model": cpu_model()},
    }


def cpu_model():
    """The host CPU's model name (first 'model name' of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    ap.add_argument("--config", default="",
                    help="variant,items_per_lane,waves_per_block,tiles_per_wave,tpv "
                         "(default: the learned tuner's choice)")
    ap.add_argument("--flush-mb", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-vendor", action="store_true")
    ap.add_argument("--no-others", action="store_true",
                    help="skip the other matrices (fp32 power-law, Laplacian, block3, uniform)")
    ap.add_argument("--workload", choices=list(matgen.WORKLOADS), default="power-law")
    ap.add_argument("--mtx", default="",
                    help="Matrix Market file to use instead of a synthetic workload "
                         "(sorted, duplicates merged; x uniform in [-1, 1])")
    ap.add_argument("--memory", choices=["device", "managed", "host-mapped"], default="device",
                    help="where A, x and y live: hipMalloc (default), hipMallocManaged, or "
                         "pinned host memory mapped into the device (the last two are the "
                         "reference's USE_SVM_ALWAYS=1 build)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported "
              "(the multi-GPU nnz sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    from benchmark_spmv_using_csr_amd import tuner

    dtype = np.float64 if a.dtype == "f64" else np.float32
    cfg = tuple(int(v) for v in a.config.split(",")) if a.config else None
    if a.mtx:
        A = matgen.read_mtx(a.mtx)
        A.name = f"Matrix Market file {os.path.basename(a.mtx)} m={A.m} n={A.n}"
        x = matgen.vector(105, A.n)
    else:
        A, x = matgen.workload(a.workload)

    managed = a.memory != "device"
    r = run_gpu(A, x, dtype, cfg, a.warmup, a.steps, a.flush_mb << 20,
                {"device": 0, "managed": 1, "host-mapped": 2}[a.memory])
    cfg = r["cfg"]
    vs = np.dtype(dtype).itemsize
    b = eff_bytes(A.m, A.n, A.nnz, vs)
    med_total = median(r["ms_total"])
    # the main kernel's duration: the mean interval between the start and end
    # events attached to its dispatch (hipExtLaunchKernelGGL), on the stream it
    # runs on.  It agrees with rocprofv3's kernel-trace average for the same
    # kernel (profiles/); the fixed cost of an empty interval between the same
    # events is reported beside it, not subtracted
    ev_ms = median(r["event_overhead"])
    mean_main = float(np.mean(r["ms_main"]))
    value = b / (med_total * 1e-3) / 1e9
    achieved = b / (mean_main * 1e-3) / 1e9
    S = cfg[1]
    prof = "" if (a.mtx or managed) else f"segsum_{a.workload.replace('-', '')}_{a.dtype}_traffic.json"
    traffic = load_traffic(prof, cfg)
    floor_ms = median(r["floor"])
    copy_ms = stream_copy(512 << 20, reps=10)
    copy_gbps = 2 * (512 << 20) / (median(copy_ms) * 1e-3) / 1e9
    us, best = r["sweep_us"], r["sweep_best"]
    tname = "double" if vs == 8 else "float"
    kname = (f"csrvec_kernel<{tname},{cfg[4]}>" if cfg[0] == 1 else
             f"segsum_kernel<{tname},{S}>" + (f" over {cfg[4]} column panels" if cfg[0] == 2
                                              else ""))
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(med_total, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": a.dtype,
        "data": ("Matrix Market file" if a.mtx else "synthetic (seeded generator)") +
                (f", A, x and y in {a.memory} memory" if managed else
                 ", inputs resident in HBM") + ", L2+MALL flushed per step",
        "config": {
            "workload": A.name, "m": A.m, "n": A.n, "nnz": A.nnz,
            "effective_bytes": b,
            "kernel": {0: "speculative segmented sum", 1: "csr-vector",
                       2: "speculative segmented sum over column panels"}[cfg[0]],
            "tile": tuner.class_name(cfg), "tile_source": r["cfg_source"],
            "parallelism": "single GPU",
            "memory": a.memory,
        },
        "pct_hbm_peak": round(100 * value / HBM_PEAK_GBPS, 2),
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "definition": "achieved = the call's algorithmic bytes / the main kernel's mean "
                          "duration (dispatch-attached HIP events); whole_call_frac uses the "
                          "median whole call (every kernel of it) instead",
            "whole_call_frac": round(value / HBM_PEAK_GBPS, 4),
            "kernel": kname,
            "kernel_mean_us": round(mean_main * 1e3, 3),
            "event_overhead_us": round(ev_ms * 1e3, 3),
            "measured_ceilings": {
                "stream_copy_gbps": round(copy_gbps, 1),
                "matrix_floor_us": round(floor_ms * 1e3, 3),
                "matrix_floor_gbps": round(b / (floor_ms * 1e-3) / 1e9, 1),
                "kernel_frac_of_floor": round(floor_ms / mean_main, 4),
                "note": "floor = same val/col stream and x gathers as the main kernel with no "
                        "segmentation and no y (bhs_*_bench_floor); on random-column matrices "
                        "the x gathers, not HBM, bound every CSR SpMV",
            },
        },
        "tuner": {
            "chosen": tuner.class_name(cfg), "tuning_s": round(r["tune_s"], 5),
            "shortlist": TUNER_TOPK,
            "network_top1": (tuner.class_name(tuner.CLASSES[r["top1"]])
                             if r["top1"] is not None else None),
            "network_top1_regret_vs_exhaustive": (
                round(float(us[r["top1"]]) / float(us[best]) - 1, 4)
                if r["top1"] is not None else None),
            "default_tile": tuner.class_name(tuner.CLASSES[tuner.DEFAULT_CLASS]),
            "default_ms": round(r["default_ms"], 5),
            "tuned_speedup_vs_default": (
                round(float(us[tuner.DEFAULT_CLASS]) / float(us[r["chosen"]]), 3)
                if r["chosen"] is not None else round(r["default_ms"] / med_total, 3)),
            "exhaustive_best": tuner.class_name(tuner.CLASSES[best]),
            "exhaustive_best_us": round(float(us[best]), 3),
            "chosen_us_in_sweep": (round(float(us[r["chosen"]]), 3)
                                   if r["chosen"] is not None else None),
            # both times from the same exhaustive sweep (same flush, same reps)
            "regret_vs_exhaustive": (round(float(us[r["chosen"]]) / float(us[best]) - 1, 4)
                                     if r["chosen"] is not None else None),
        },
        "timing": {
            "median_step_ms": med_total, "p10_step_ms": float(np.percentile(r["ms_total"], 10)),
            "p90_step_ms": float(np.percentile(r["ms_total"], 90)),
            "mean_main_kernel_ms": mean_main,
            "wall_s_timed_loop_incl_flush": round(r["wall_s"], 4),
            "warm_median_step_ms": median(r["warm_total"]),
            "warm_gbps": round(b / (median(r["warm_total"]) * 1e-3) / 1e9, 2),
            "init_data_s": round(r["init_s"], 4),
        },
    }
    # the plain layout (the reference's layout, with its speculative fix-up and
    # calibration) at its fastest tile, on the same matrix and timing
    pc = r["plain_cfg"]
    p_med = median(r["plain_total"])
    p_main = float(np.mean(r["plain_main"]))
    p_traffic = load_traffic(prof, pc)
    p_ach = b / (p_main * 1e-3) / 1e9
    out["plain_layout"] = {
        "tile": tuner.class_name(pc), "ms_per_step": round(p_med, 5),
        "gbps": round(b / (p_med * 1e-3) / 1e9, 2),
        "dirty_tiles": int(r["plain_stats"]["n_dirty"]), "tiles": int(r["plain_stats"]["n_tiles"]),
        "waves_calibrated": int(r["plain_stats"]["n_sync"]),
        "roofline": {"bound": "hbm", "achieved": round(p_ach, 2), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(p_ach / HBM_PEAK_GBPS, 4),
                     "traffic": p_traffic, "kernel_mean_us": round(p_main * 1e3, 3)},
    }
    if r["layout_ms"] is not None:
        saved = p_med - med_total
        out["layout_build"] = {
            "layout": out["config"]["tile"], "build_ms": round(r["layout_ms"], 2),
            "saved_per_call_ms": round(saved, 5),
            "break_even_calls": (int(np.ceil(r["layout_ms"] / saved)) if saved > 0 else None),
            "note": "one-time set_config that builds the layout from the device copy of A, "
                    "against the plain layout's fastest tile",
        }
    if not a.no_vendor:
        v = run_vendor(A, x, dtype, a.flush_mb << 20, r["y"])
        v["bhs_speedup_vs_best"] = round(v["best_cold_us"] / (med_total * 1e3), 3)
        out["vendor"] = v
    if not a.no_batched:
        out["batched"] = run_batched()
    if not a.no_others and not a.mtx and a.memory == "device":
        out["other_workloads"] = run_others(a.flush_mb << 20, not a.no_vendor)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(A, x, dtype, r["y"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""The committed rocprofv3 traffic summaries that bench.py reports as
roofline.traffic (CPU only): each entry's bytes follow from its own counters
and copy-kernel calibration, its algorithmic bytes from the byte formula on
the generated matrix, its tile is a tuner class, and the summary it cites
exists."""
import json
import os

import numpy as np
import pytest

from benchmark_spmv_using_csr_amd import matgen, tuner

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def eff_bytes(m, n, nnz, v):
    return nnz * (v + 4) + 4 * (m + 1) + (n + m) * v


@pytest.fixture(scope="module")
def power_law():
    return matgen.power_law_1m()


@pytest.mark.parametrize("dt,v", [("f64", 8), ("f32", 4)])
def test_traffic_entries_are_consistent(dt, v, power_law):
    A = power_law
    with open(os.path.join(ROOT, "profiles", f"segsum_powerlaw_{dt}_traffic.json")) as f:
        d = json.load(f)
    assert d["entries"]
    seen = set()
    for e in d["entries"]:
        cfg = tuple(e["config"])
        assert cfg in tuner.CLASSES, cfg
        assert cfg not in seen
        seen.add(cfg)
        cal = e["copy_calibration"]
        rf = cal["bytes_read_kib"] / cal["fetch_size_kib"]
        wf = cal["bytes_written_kib"] / cal["write_size_kib"]
        assert abs(rf - cal["read_factor"]) < 1e-3 and abs(wf - cal["write_factor"]) < 1e-3
        hb = (e["fetch_size_kib"] * rf + e["write_size_kib"] * wf) * 1024
        assert abs(hb - e["hbm_bytes_per_launch"]) <= 2e-6 * hb
        assert e["algorithmic_bytes_per_launch"] == eff_bytes(A.m, A.n, A.nnz, v)
        # the main kernel moves at least its algorithmic bytes, less x refetches
        assert e["hbm_bytes_per_launch"] > 0.9 * e["algorithmic_bytes_per_launch"]
        assert os.path.exists(os.path.join(ROOT, e["source"])), e["source"]


def test_bench_reads_the_entries():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    with open(os.path.join(ROOT, "profiles", "segsum_powerlaw_f64_traffic.json")) as f:
        d = json.load(f)
    for e in d["entries"]:
        got = bench.load_traffic("segsum_powerlaw_f64_traffic.json", tuple(e["config"]))
        assert got == float(e["hbm_bytes_per_launch"])
    assert bench.load_traffic("segsum_powerlaw_f64_traffic.json", (0, 16, 8, 99, 0)) is None
    assert bench.load_traffic("", (2, 8, 1, 8, 4)) is None
    assert np.isfinite(bench.HBM_PEAK_GBPS)


def test_batched_traffic_entries_are_consistent():
    from benchmark_spmv_using_csr_amd import batched
    with open(os.path.join(ROOT, "profiles", "batched_256mat_f64_traffic.json")) as f:
        d = json.load(f)
    assert d["entries"]
    for e in d["entries"]:
        assert tuple(e["schedule"]) in batched.SCHEDULES, e["schedule"]
        cal = e["copy_calibration"]
        rf = cal["bytes_read_kib"] / cal["fetch_size_kib"]
        wf = cal["bytes_written_kib"] / cal["write_size_kib"]
        hb = (e["fetch_size_kib"] * rf + e["write_size_kib"] * wf) * 1024
        assert abs(hb - e["hbm_bytes_per_launch"]) <= 2e-6 * hb
        assert e["hbm_bytes_per_launch"] > e["algorithmic_bytes_per_launch"]
        assert os.path.exists(os.path.join(ROOT, e["source"])), e["source"]

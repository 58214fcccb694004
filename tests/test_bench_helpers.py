"""bench.py's measurement arithmetic and its one-GPU contract (no GPU needed)."""

from __future__ import annotations

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_roofline_is_algorithmic_bytes_over_the_average_launch():
    # 2+1: per visit 3 rows x 12 B of cell heads, per compared point a 16 B FP32 record
    st = {"launches": 4, "items": 4 << 20, "ms": 0.4, "visits": 10**9, "points": 5 * 10**8}
    r = bench.roofline_of(st, 2, 64, True, traffic=123.0)
    per_launch = (10**9 * 3 * 12 + 5 * 10**8 * 16) / 4
    assert r["algorithmic_bytes_per_launch"] == pytest.approx(per_launch)
    assert r["achieved"] == pytest.approx(per_launch / 1e-4 / 1e9, rel=1e-3)
    assert r["frac"] == pytest.approx(r["achieved"] / bench.HBM_PEAK_GBS, rel=1e-3)
    assert (r["bound"], r["unit"], r["peak"], r["traffic"]) == ("hbm", "GB/s", 8000.0, 123.0)
    # 3+1 FP64 form: 9 rows, and 3 doubles plus a link per point
    r3 = bench.roofline_of(st, 3, 128, False)
    assert r3["algorithmic_bytes_per_launch"] == pytest.approx((10**9 * 9 * 12 + 5 * 10**8 * 28) / 4)


def test_mfma_figure_counts_two_flops_per_coefficient_and_column():
    st = {"launches": 2, "items": 10**9, "ms": 2.0, "visits": 0, "points": 0}
    m = bench.mfma_roofline(st, 2, 64)
    flops = 10**9 * 2 * 2 * 2 * 63 * 32 / 2  # per launch: items x 2 d K 32, K = 2(F-1)
    assert m["flops_per_launch"] == pytest.approx(flops)
    assert m["achieved"] == pytest.approx(flops / 1e-3 / 1e12, rel=1e-3)
    assert (m["bound"], m["unit"], m["peak"]) == ("mfma", "TFLOP/s", 2500.0)
    # models shorter than one column block count their T columns only
    assert bench.mfma_roofline(st, 2, 16)["flops_per_launch"] == pytest.approx(10**9 * 2 * 2 * 2 * 15 * 16 / 2)


def test_committed_profiles_feed_the_bench_line():
    for w in ("cfg2", "cfg3", "cfg5"):
        assert bench.measured_traffic(w, True) > 0
        assert 0 < bench.measured_valu_busy(w, True) <= 100
    assert bench.measured_traffic("cfg2", False) is None  # the FP64 form has no summary


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""

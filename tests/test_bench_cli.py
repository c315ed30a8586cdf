"""bench.py's contract pieces that need no GPU: the one-GPU refusal and the
roofline / VALU arithmetic on a synthetic kernel record."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_roofline_fields():
    # 4 launches of 2 ms moving 8e9 algorithmic bytes each: 4 TB/s, half the peak
    stat = {"name": "expand", "launches": 4, "ms": 8.0, "bytes": 32e9}
    r = bench.roofline(stat, traffic=9e9)
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == bench.HBM_PEAK_GBS
    assert r["achieved"] == pytest.approx(4000.0)
    assert r["frac"] == pytest.approx(0.5)
    assert r["avg_launch_ms"] == pytest.approx(2.0)
    assert r["algorithmic_bytes_per_launch"] == 8e9
    assert r["traffic"] == 9e9


def test_valu_issue_arithmetic():
    insts, lanes = bench.PMC_VALU["leaf2"]
    # one launch lasting exactly insts / (SIMDs x clock) seconds issues one
    # wave instruction per SIMD and cycle
    ms = insts / (bench.SIMDS * bench.CLOCK_HZ) * 1e3
    v = bench.valu_issue({"name": "leaf2", "launches": 1, "ms": ms}, counts=int(insts * 64 / 700))
    assert v["wave_insts_per_simd_cycle"] == pytest.approx(1.0, rel=1e-3)
    assert v["lane_insts_per_count"] == pytest.approx(700, abs=1)
    assert v["lanes_active"] == lanes
    assert bench.valu_issue({"name": "leaf1", "launches": 1, "ms": 1.0}, counts=1) is None

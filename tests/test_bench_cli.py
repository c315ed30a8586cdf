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
    insts, lanes, _ = bench.PMC_VALU["leaf2"]
    # one launch lasting exactly insts / (SIMDs x clock) seconds issues one
    # wave instruction per SIMD and cycle
    ms = insts / (bench.SIMDS * bench.CLOCK_HZ) * 1e3
    v = bench.valu_issue({"name": "leaf2", "launches": 1, "ms": ms}, counts=int(insts * 64 / 700))
    assert v["wave_insts_per_simd_cycle"] == pytest.approx(1.0, rel=1e-3)
    assert v["lane_insts_per_count"] == pytest.approx(700, abs=1)
    assert v["lanes_active"] == lanes
    assert bench.valu_issue({"name": "leaf1", "launches": 1, "ms": 1.0}, counts=1) is None


def test_valu_roofline_arithmetic():
    # a leaf launch whose single-issued slots fill every SIMD issue slot of the
    # launch time sits at frac 1; the calibrated ceiling is the best
    # single-issue probe's share of the same slots
    insts, lanes, dual = bench.PMC_VALU["leaf2"]
    slots_per_s = bench.SIMDS * bench.CLOCK_HZ / 4
    ms = insts * (1 - dual) / slots_per_s * 1e3
    stat = {"name": "leaf2", "launches": 2, "ms": 2 * ms, "bytes": 2 * 1e8}
    hbm = bench.roofline(stat, traffic=1.4e8)
    calib = {f"{n}_w8": {"wave_insts_per_s": 0.9 * slots_per_s} for n in bench.PROBE_KINDS}
    calib["v_bcnt_u32_b32_w8"] = {"wave_insts_per_s": 0.95 * slots_per_s}
    ceiling = bench.calibrated_issue_ceiling(calib)
    assert ceiling == pytest.approx(0.95)
    r = bench.roofline_valu(stat, hbm, ceiling)
    assert r["bound"] == "valu" and r["frac"] == pytest.approx(1.0, rel=1e-3)
    assert r["peak"] == pytest.approx(614.4) and r["achieved"] == pytest.approx(614.4, rel=1e-3)
    assert r["traffic"] == 1.4e8 and r["hbm"]["frac"] == hbm["frac"]
    assert r["lanes_active"] == lanes and r["calibrated_ceiling_frac"] == pytest.approx(0.95)

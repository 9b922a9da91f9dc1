"""Host logic of bench.py: the one-GPU guard, the roofline object of the dominant kernel and the
committed traffic summary it reads. No GPU is used."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_roofline_of_k_lite_move_is_bytes_over_event_time():
    tm = {"steps": 200, "lite_move_ms": 200 * 0.025, "contact_ms": 200 * 0.06}
    info = {"contact_spheres": 1000}
    r = bench.roofline(tm, 1_000_000, 0.004, True, info, 1.5e8)
    units = 1_000_000 - 1000
    assert r["kernel_ms"] == pytest.approx(0.025)
    assert r["units"] == units
    assert r["achieved"] == pytest.approx(168.0 * units / 25e-6 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    assert r["traffic"] == 1.5e8 and r["bound"] == "hbm" and r["unit"] == "GB/s"
    # split mode: the contact pass over all spheres
    s = bench.roofline(tm, 1_000_000, 0.004, False, info, None)
    assert s["kernel_ms"] == pytest.approx(0.06)
    assert s["achieved"] == pytest.approx(bench.contact_kernel_bytes(0.004, False) * 1e6 / 60e-6 / 1e9)


def test_committed_traffic_summary_matches_its_workload_only():
    with open(bench.TRAFFIC_FILE) as f:
        t = json.load(f)
    assert t["np"] == 1_000_000
    # the copy calibration: FETCH_SIZE counts half of the bytes a copy reads, WRITE_SIZE all it writes
    assert t["copy_scale_fetch"] == pytest.approx(2.0, rel=0.01)
    assert t["copy_scale_write"] == pytest.approx(1.0, rel=0.01)
    v = bench.measured_traffic(1_000_000, True)
    k = t["k_lite_move"]
    assert v == pytest.approx((k["fetch_kib"] * t["copy_scale_fetch"] + k["write_kib"] * t["copy_scale_write"]) * 1024)
    # the kernel moves at most its algorithmic bytes per sphere (skipped reads and writes of
    # settled spheres make it less), and at least its state reads and writes of position and velocity
    assert 1e6 * 112 < v <= 1e6 * (bench.LITE_MOVE_BYTES + 16)
    assert bench.measured_traffic(999_999, True) is None
    assert bench.measured_traffic(1_000_000, False) is None

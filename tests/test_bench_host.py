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


def test_kernel_rooflines_are_bytes_over_event_time():
    tm = {"steps": 200, "lite_move_ms": 200 * 0.025, "contact_ms": 200 * 0.09, "pair_pass_ms": 200 * 0.04,
          "collide_ms": 200 * 0.03}
    info = {"contact_spheres": 100000}   # a dense full pass: k_pairs walks the pairs
    kr = bench.kernel_rooflines(tm, 1_000_000, info, 6000, 200, True)
    units = 1_000_000 - 100000
    assert kr["k_lite_move"]["ms"] == pytest.approx(0.025)
    assert kr["k_lite_move"]["units"] == units
    assert kr["k_lite_move"]["gbps"] == pytest.approx(168.0 * units / 25e-6 / 1e9)
    # k_pairs: 228 B per full-pass sphere and 44 B per passing pair; k_collide: 480 B per
    # sphere and 40 B per wall contact
    assert kr["k_pairs"]["bytes"] == 228.0 * 100000 + 44.0 * 6000
    assert kr["k_pairs"]["gbps"] == pytest.approx((228.0 * 100000 + 44.0 * 6000) / 40e-6 / 1e9)
    assert kr["k_collide"]["bytes"] == 480.0 * 100000 + 40.0 * 200
    for k in kr.values():
        assert k["frac"] == pytest.approx(k["gbps"] / 8000.0)
    r = bench.dominant_roofline(kr, {"k_lite_move": 1.5e8})
    assert r["kernel"] == "k_pairs" and r["kernel_ms"] == pytest.approx(0.04)
    assert r["traffic"] is None and r["bound"] == "hbm" and r["unit"] == "GB/s"
    tm["lite_move_ms"] = 200 * 0.05
    r = bench.dominant_roofline(bench.kernel_rooflines(tm, 1_000_000, info, 6000, 200, True), {"k_lite_move": 1.5e8})
    assert r["kernel"] == "k_lite_move" and r["traffic"] == 1.5e8
    # split mode: no fused kernels
    assert set(bench.kernel_rooflines(tm, 1_000_000, info, 6000, 200, False)) == {"k_pairs"}
    # a small full pass: k_collide walks the pairs, k_pairs only reads the counters
    small = bench.kernel_rooflines(tm, 1_000_000, {"contact_spheres": 3000}, 900, 200, True)
    assert "k_pairs" not in small and small["k_collide"]["walks_pairs"]
    assert small["k_collide"]["bytes"] == 416.0 * 3000 + 40.0 * 200 + 44.0 * 900


def test_host_cpu_reports_model_and_cores():
    model, usable, total = bench.host_cpu()
    assert isinstance(model, str) and 1 <= usable <= total


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

"""bench.py pieces that run without a GPU.

* `--gpus N` with N > 1 exits with status 2 and says why (one GPU only);
* the instruction-issue bound in the roofline block is the closed form
  N^2 / 64 / 1024 SIMDs x (4 cycles per full-rate and 16 per quarter-rate
  FP64 wave-instruction) / clock, and is absent for the FP32 mixed kernel;
* the committed HBM-traffic summary read into `roofline.traffic` holds an
  entry per RHS variant, scaled by its copy-kernel calibration.
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_instruction_bound_closed_form():
    import bench
    b = bench.instruction_bound("exact", 16384, 2000.0, 0.5)
    # 38.38 VALU instructions per 64 interactions (the counters of the power-of-two
    # mass path): 36.38 full-rate x 4 + 2 quarter-rate x 16 cycles
    assert b["cycles_per_64_interactions"] == pytest.approx(36.38 * 4 + 32)
    want_ms = 16384 ** 2 / 64 / 1024 * (36.38 * 4 + 32) / 2.0e6
    assert b["bound_ms"] == pytest.approx(want_ms)
    assert b["frac"] == pytest.approx(want_ms / 0.5)
    # twice the clock halves the bound; four times the particles quadruples it twice
    assert bench.instruction_bound("exact", 16384, 4000.0, 0.5)["bound_ms"] == \
        pytest.approx(want_ms / 2)
    assert bench.instruction_bound("exact", 32768, 2000.0, 0.5)["bound_ms"] == \
        pytest.approx(want_ms * 4)
    assert bench.instruction_bound("fast", 16384, 2000.0, 0.5)["quarter_rate"] == 1
    assert bench.instruction_bound("mixed", 16384, 2000.0, 0.5) is None


def test_traffic_summary_is_consistent():
    d = json.load(open(os.path.join(ROOT, "profiles", "nbody_rhs_n16384_traffic.json")))
    cal = d["calibration"]
    # the calibration copy moves 64 MiB each way
    assert cal["write_kib"] * cal["write_scale"] == pytest.approx(65536.0)
    assert cal["fetch_kib"] * cal["fetch_scale"] == pytest.approx(65536.0)
    for var in ("exact", "fast", "mixed"):
        k = d["kernels"][var]
        want = k["fetch_kib"] * 1024 * cal["fetch_scale"] + k["write_kib"] * 1024 * cal["write_scale"]
        assert k["hbm_bytes_per_launch"] == pytest.approx(want)
        # at least the algorithmic bytes (7N doubles read, 6N written at N=16384)
        assert k["hbm_bytes_per_launch"] >= 13 * 16384 * 8

"""Host-side pieces of bench.py: the committed counter summary it reads for
the dominant kernel's traffic and vector-ALU floor, and its one-GPU limit."""
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

import bench  # noqa: E402

PROFILE = os.path.join(REPO, "profiles", "traffic_channel_128x64x64.json")


def test_profile_record_prefers_the_tile_run():
    with open(PROFILE) as f:
        d = json.load(f)
    base = bench.profile_record("fluid")
    assert base == d["fluid"]
    assert bench.profile_record("fluid", (8, 8, 8)) == d["fluid"]  # no 8x8x8-keyed entry: the base one
    if "fluid@16x8x4" in d:
        rec = bench.profile_record("fluid", (16, 8, 4))
        assert rec == d["fluid@16x8x4"]
        assert rec["bytes_per_launch"] != base["bytes_per_launch"]
    assert bench.profile_record("no_such_stage") is None
    assert bench.profile_record("no_such_stage", (16, 8, 4)) is None


def test_traffic_and_valu_floor_from_the_summary():
    rec = bench.profile_record("flux_update")
    assert bench.measured_traffic("flux_update") == rec["bytes_per_launch"]
    # copy-calibrated: FETCH_SIZE KiB x its scale + WRITE_SIZE KiB x its scale
    with open(PROFILE) as f:
        cal = json.load(f)["_calibration"]
    b = rec["fetch_kib_per_dispatch"] * 1024 * cal["fetch_scale"] + rec["write_kib_per_dispatch"] * 1024 * cal["write_scale"]
    assert abs(b - rec["bytes_per_launch"]) <= 1.0
    v = bench.valu_roofline("flux_update", 0.040)
    floor_ms = rec["valu_insts_per_launch"] / (256 * 4 * 2.4e9 / 4) * 1e3
    assert v["floor_ms"] == pytest.approx(floor_ms)
    assert v["frac"] == pytest.approx(floor_ms / 0.040)
    assert bench.valu_roofline("no_such_stage", 0.040) is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr

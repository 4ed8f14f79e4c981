"""bench.py pieces that run without a GPU: the one-GPU guard, the figures read from the
committed rocprofv3 summaries, the VALU-issue and HBM roofline arithmetic and the serving
mix."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    for argv, env in ((["--gpus", "2"], {}), (["--gpus", "1"], {"WORLD_SIZE": "4"})):
        e = dict(os.environ, **env)
        p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *argv],
                           capture_output=True, text=True, env=e, timeout=120)
        assert p.returncode == 2
        assert "one GPU only" in p.stderr
        assert p.stdout.strip() == ""  # no JSON line for a run that did not happen


def test_profile_figures_present_and_consistent():
    """The committed PMC summary carries the figures the bench line quotes."""
    lane = bench.profile_value("valu_lane_instr_per_ray_step")
    traffic = bench.traffic_from_profiles()
    assert lane is not None and 150 < lane < 500
    assert traffic is not None and traffic > 0
    assert bench.profile_value("no_such_key") is None


def test_valu_roofline_arithmetic():
    lane = bench.profile_value("valu_lane_instr_per_ray_step")
    rates = {"fma_sgpr": 3.3e13, "add_f32": 8.0e13, "fma_f64": 1.7e13}   # lane-instr/s
    # the step rate at which the kernel issues exactly the fastest form's rate
    full = 8.0e13 / lane
    r = bench.valu_roofline(full, rates, {"fma_sgpr": 50.0, "add_f32": 20.0, "fma_f64": 99.0})
    assert r["bound"] == "valu_issue" and r["unit"] == "TOP/s"
    assert r["frac"] == pytest.approx(1.0, abs=1e-4)
    assert r["peak"] == pytest.approx(80.0) and "add_f32" in r["peak_source"]
    assert r["calibration_tops"]["fma_sgpr"] == pytest.approx(33.0)
    r = bench.valu_roofline(0.25 * full, rates)
    assert r["frac"] == pytest.approx(0.25, abs=1e-4)
    json.dumps(r)  # goes into the JSON line as is
    assert bench.valu_roofline(full, {}) is None


def test_hbm_roofline_arithmetic():
    traffic = bench.traffic_from_profiles()
    n_vox, n_rays = 512 ** 3, 10_000_000
    r = bench.hbm_roofline(traffic / 8.0e12, n_vox, n_rays)   # duration at exactly 8 TB/s
    assert r["frac"] == pytest.approx(1.0, abs=1e-4) and r["unit"] == "GB/s"
    # compulsory: 16 B per voxel + 68 B of per-ray I/O
    assert r["compulsory_bytes"] == 16 * n_vox + 68 * n_rays
    assert r["traffic_over_compulsory"] == pytest.approx(traffic / r["compulsory_bytes"],
                                                         rel=1e-3)
    json.dumps(r)


def test_algorithmic_bytes_per_ray_step():
    # 4 RHS evaluations x 8 corners x 4 fp32 channels (gx, gy, gz, kappa)
    assert bench.BYTES_PER_STEP == 512
    assert bench.HBM_PEAK_GBS == 8000.0


def test_serving_mix_is_seeded_and_in_range():
    a = bench.serving_mix(64, seed=7)
    b = bench.serving_mix(64, seed=7)
    assert [r.field for r in a] == [r.field for r in b]
    assert [r.beam for r in a] == [r.beam for r in b]
    sides = {r.field.side for r in a}
    assert sides <= {64, 128, 192, 256, 384} and len(sides) >= 3
    rays = np.array([r.beam.n_rays for r in a])
    assert rays.min() >= 10_000 and rays.max() <= 4_000_000
    assert {r.lwl for r in a} <= {532e-9, 1064e-9}
    for r in a:
        # absorption exactly for the Kolmogorov fields
        assert r.inv_brems == (r.field.kind == "kolmogorov")
    assert [r.rid for r in a] == list(range(64))
    c = bench.serving_mix(64, seed=8)
    assert [r.field for r in a] != [r.field for r in c]

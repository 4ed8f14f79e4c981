"""bench.py pieces that run without a GPU: the one-GPU guard, the figures read from the
committed rocprofv3 summaries, the VALU-issue roofline arithmetic and the serving mix."""
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
    ghz = bench.profile_value("sclk_ghz_under_load")
    traffic = bench.traffic_from_profiles()
    assert lane is not None and 150 < lane < 500
    assert ghz is not None and 1.0 < ghz < 2.5
    assert traffic is not None and traffic > 0
    assert bench.profile_value("no_such_key") is None


def test_valu_roofline_arithmetic():
    lane = bench.profile_value("valu_lane_instr_per_ray_step")
    ghz = bench.profile_value("sclk_ghz_under_load")
    n_cu = 256
    # the step rate at which the kernel would issue exactly 0.5 wave64 VALU per SIMD-cycle
    full = 0.5 * 4 * n_cu * ghz * 1e9 * 64.0 / lane
    r = bench.valu_roofline(full, n_cu)
    assert r["frac"] == pytest.approx(1.0, abs=1e-4)
    assert r["simds"] == 1024 and r["peak"] == 0.5
    r = bench.valu_roofline(0.25 * full, n_cu)
    assert r["frac"] == pytest.approx(0.25, abs=1e-4)
    json.dumps(r)  # goes into the JSON line as is
    assert bench.valu_roofline(full, 0) is None


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

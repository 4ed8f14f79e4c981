"""The three tracing configurations at their full sizes.

C1 (128^3 sine, 1e5 rays, 1064 nm) is checked ray for ray against the FP64
oracle, with its 256 x 256 detector image.  C2 (256^3 quadratic trough, 1e6
rays, 532 nm) is checked on every ray against the closed-form oscillation and
on a seeded 20000-ray subset against the oracle.  C3 (512^3 turbulent box with
absorption, 1e7 rays, 532 nm) is checked on a seeded 1e6-ray subset against
the oracle and on every ray through properties that hold whatever the field.
"""
import numpy as np
import pytest

from _parity import C, assert_parity, gpu_trace, oracle_trace

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _subset(gpu, idx):
    out = dict(gpu)
    out["states"] = gpu["states"][:, :, idx]
    out["tau"] = gpu["tau"][:, idx] if gpu["tau"] is not None else None
    out["path"] = gpu["path"][idx]
    out["steps"] = gpu["steps"][idx]
    return out


def _edge_distance(v, edges):
    i = np.clip(np.searchsorted(edges, v), 1, len(edges) - 1)
    return np.minimum(np.abs(v - edges[i - 1]), np.abs(v - edges[i]))


def test_c1_sine_128_full_parity_and_detector(orc):
    import torch

    from sloth_amd.core.beam import Beam
    from sloth_amd.fields import sine_field
    from sloth_amd.processing.diagnostics import histogram2d_device
    from sloth_amd.shared.propagation import jones_device

    L = 10e-3
    ne, coords = sine_field(128, L)
    s0 = Beam(100000, 2e-3, 0.0, L / 2, beam_type="circular", seeded=True).s0
    t = np.array([0.0, np.sqrt(8.0) * L / C])
    gpu = gpu_trace(ne, coords, s0, t, lwl=1064e-9)
    ref = oracle_trace(orc, ne, coords, s0, gpu["nsub"], gpu["h"], lwl=1064e-9)
    print("C1 parity", assert_parity(gpu, ref, coords))
    assert np.median(gpu["steps"]) in (255, 256)
    st = torch.from_numpy(gpu["states"][-1].astype(np.float32)).cuda()
    rf, _ = jones_device(st, "z")
    rf_ref = orc.jones(ref["saves"][-1][:6], "z")
    edges = np.linspace(-5e-3, 5e-3, 257)
    counts, _ = histogram2d_device(rf[0], rf[2], edges, edges)
    H_ref, _, _ = np.histogram2d(rf_ref[0], rf_ref[2], bins=[edges, edges])
    near = (_edge_distance(rf_ref[0], edges) < 1e-5 * 5e-3) | \
           (_edge_distance(rf_ref[2], edges) < 1e-5 * 5e-3)
    diff = np.abs(counts.cpu().numpy() - H_ref.T).sum()
    assert diff <= 2 * near.sum(), (diff, near.sum())
    assert counts.sum().item() == 100000


def test_c2_trough_256_closed_form_all_rays_and_oracle_subset(orc):
    from sloth_amd.core.beam import Beam
    from sloth_amd.fields import quadratic_trough

    L = 12e-3
    ne, coords = quadratic_trough(256, L)
    s0 = Beam(1000000, 4e-3, 0.0, L / 2, beam_type="square", seeded=True).s0
    depth = 11e-3
    t = np.linspace(0, depth, 3) / C
    gpu = gpu_trace(ne, coords, s0, t, lwl=532e-9)
    omega = 2 * np.pi * C / 532e-9
    w = C * np.sqrt(5e24 / (3.14207787e-4 * omega ** 2)) / 3e-3
    x0, y0 = s0[0].astype(np.float64), s0[1].astype(np.float64)
    fin = gpu["states"][-1]
    # the trilinear interpolant of the exactly linear gradient is exact: every ray follows
    # x0 cos(w t) to the RK4 truncation error
    np.testing.assert_allclose(fin[0], x0 * np.cos(w * t[-1]), rtol=0, atol=2e-9)
    np.testing.assert_allclose(fin[1], y0 * np.cos(w * t[-1]), rtol=0, atol=2e-9)
    idx = np.sort(np.random.default_rng(20).choice(s0.shape[1], 20000, replace=False))
    ref = oracle_trace(orc, ne, coords, s0[:, idx], gpu["nsub"], gpu["h"], lwl=532e-9)
    print("C2 subset parity", assert_parity(_subset(gpu, idx), ref, coords))


def test_c3_turbulent_512_oracle_subset_and_invariants(orc):
    import torch

    from sloth_amd.core.beam import Beam
    from sloth_amd.fields import kolmogorov_field

    L = 20e-3
    dev = torch.device("cuda:0")
    ne, coords = kolmogorov_field(512, L, seed=1, device=dev)
    s0 = Beam(10_000_000, 8e-3, 0.0, L / 2, beam_type="square", seeded=True).s0
    t = np.array([0.0, np.sqrt(8.0) * L / C])
    gpu = gpu_trace(ne, coords, s0, t, lwl=532e-9, inv_brems=True, Te=20.0, Z=1.0,
                    coulomb_log=10.0)
    steps = gpu["steps"]
    assert gpu["total"] == int(steps.sum())
    assert steps.min() >= 1021 and steps.max() <= 1025          # every ray crosses 20 mm
    tau = gpu["tau"][-1]
    assert np.all(tau > 0) and np.all(np.isfinite(gpu["states"][-1]))
    # path length: at least the box depth, and the straight run to the end of the schedule
    assert np.all(gpu["path"] >= L * (1 - 1e-6))
    # a seeded 1e6-ray subset (1e9 oracle ray-steps) against the FP64 oracle; the count
    # of rays over the energy bar, each shown to pass within 1e-5 of the extent of a face
    # or to differ by one step, is printed
    idx = np.sort(np.random.default_rng(30).choice(s0.shape[1], 1_000_000, replace=False))
    print("C3: oracle on 1e6 rays ...", flush=True)
    ref = oracle_trace(orc, ne, coords, s0[:, idx], gpu["nsub"], gpu["h"], lwl=532e-9,
                       inv_brems=True, Te=20.0, Z=1.0, coulomb_log=10.0)
    print("C3 subset parity", assert_parity(_subset(gpu, idx), ref, coords), flush=True)


def test_trough_646_packed_grid_above_4gib_closed_form():
    """A packed grid above 4 GiB (646^3 x 16 B) takes 64-bit corner addresses; the
    quadratic trough's closed form pins every ray there as on the smaller grids."""
    import torch

    from sloth_amd.core.beam import Beam
    from sloth_amd.fields import quadratic_trough

    L = 12e-3
    n = 646
    assert n ** 3 * 16 > 0xFFFFFFFF
    ne, coords = quadratic_trough(n, L)
    s0 = Beam(20000, 4e-3, 0.0, L / 2, beam_type="square", seeded=True).s0
    t = np.linspace(0, 11e-3, 3) / C
    gpu = gpu_trace(ne, coords, s0, t, lwl=532e-9)
    assert gpu["grid"].nbytes > 0xFFFFFFFF
    del ne
    omega = 2 * np.pi * C / 532e-9
    w = C * np.sqrt(5e24 / (3.14207787e-4 * omega ** 2)) / 3e-3
    x0, y0 = s0[0].astype(np.float64), s0[1].astype(np.float64)
    for k in (1, 2):
        st = gpu["states"][k]
        np.testing.assert_allclose(st[0], x0 * np.cos(w * t[k]), rtol=0, atol=2e-9)
        np.testing.assert_allclose(st[1], y0 * np.cos(w * t[k]), rtol=0, atol=2e-9)
    del gpu
    torch.cuda.empty_cache()

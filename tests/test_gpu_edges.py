"""Edge cases on the GPU: empty and degenerate inputs through each entry point."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def test_histogram_empty_all_nan_and_all_outside():
    import torch

    from sloth_amd.processing.diagnostics import histogram2d_device

    e = np.linspace(-1.0, 1.0, 9)
    for case, x in (("empty", np.zeros(0)), ("nan", np.full(50, np.nan)),
                    ("outside", np.full(50, 3.0))):
        t = torch.from_numpy(x).cuda()
        counts, wsum = histogram2d_device(t, t, e, e, weights=torch.ones(x.size, dtype=torch.float64))
        assert counts.sum().item() == 0 and wsum.abs().sum().item() == 0.0, case
    # the last edge belongs to the last bin, the first to the first (numpy's rule)
    x = torch.tensor([-1.0, 1.0, 1.0 + 1e-12], dtype=torch.float64).cuda()
    counts, _ = histogram2d_device(x, x, e, e)
    c = counts.cpu().numpy()
    assert c[0, 0] == 1 and c[-1, -1] == 1 and c.sum() == 2


def test_optics_empty_and_all_rejected():
    from sloth_amd.processing import diagnostics as D

    d = D.Shadowgraphy(np.zeros((4, 0), np.float32))
    d.two_lens_solve()
    assert d.rf.shape == (4, 0)
    d.histogram(pix_x=10, pix_y=10)
    assert d.H.sum() == 0
    far = np.zeros((4, 5), np.float32)
    far[0] = 1.0                                  # 1 m off axis: outside every aperture
    d = D.Schlieren(far)
    d.DF_solve()
    assert np.isnan(d.rf).all()
    d.histogram(pix_x=10, pix_y=10)
    assert d.H.sum() == 0


def test_serving_single_ray_and_empty_requests():
    from sloth_amd.serving import BeamSpec, FieldSpec, RequestSpec, ServingEngine, run_offline

    reqs = [RequestSpec(rid=0, field=FieldSpec(kind="sine", side=32, length=6e-3),
                        beam=BeamSpec(n_rays=1, half_width=1e-3, seed=1), bins=16),
            RequestSpec(rid=1, field=FieldSpec(kind="kolmogorov", side=32, length=6e-3, seed=2),
                        beam=BeamSpec(n_rays=0, half_width=1e-3, seed=2), bins=16),
            RequestSpec(rid=2, field=FieldSpec(kind="kolmogorov", side=40, length=6e-3, seed=3),
                        beam=BeamSpec(n_rays=3000, half_width=2e-3, seed=3), bins=16)]
    off, _ = run_offline(reqs)
    resp, stats = ServingEngine(max_slots=4, blocks=64, batch_work=1e5).run(reqs, realtime=False)
    assert stats["ok"] == 3
    assert resp[1].n_rays == 0 and resp[1].ray_steps == 0 and resp[1].image.sum().item() == 0
    assert resp[0].n_rays == 1 and resp[0].ray_steps == off[0]["steps"] > 0
    np.testing.assert_array_equal(resp[2].rf.cpu().numpy(), off[2]["rf"].cpu().numpy())
    np.testing.assert_array_equal(resp[0].rf.cpu().numpy(), off[0]["rf"].cpu().numpy())


def test_trace_rays_that_miss_the_grid():
    """Rays that never touch the box move in straight lines and count no steps."""
    from sloth_amd.core import engine
    from sloth_amd.fields import sine_field

    L = 6e-3
    ne, coords = sine_field(24, L)
    grid = engine.PackedGrid(ne, *coords, 2 * np.pi * 299792458.0 / 1064e-9)
    s0 = np.zeros((6, 3), np.float32)
    s0[0] = [5e-3, -4e-3, 0.0]        # two beside the box, one heading away from it
    s0[1] = [0.0, 0.0, 0.0]
    s0[2] = [-L / 2, -L / 2, -L]
    s0[5] = [299792458.0, 299792458.0, -299792458.0]
    t = np.array([0.0, 2 * L / 299792458.0])
    r = engine.trace(grid, s0, t, save_all=False, want_path=True)
    assert r.steps.cpu().numpy().tolist() == [0, 0, 0]
    end = r.states[0].cpu().numpy()
    np.testing.assert_allclose(end[2], s0[2] + s0[5] * t[-1], rtol=1e-6)
    np.testing.assert_allclose(r.path.cpu().numpy(), 2 * L, rtol=1e-6)


def test_trace_non_finite_rays_leave_the_others_alone():
    """NaN or infinite start states come out non-finite and do not change any other ray's
    result (bit for bit).  A non-finite position is outside the box from the start and counts
    no steps; a non-finite velocity leaves the box after the first step, which counts once
    when the ray starts on the box face.  As in the reference, where NaN propagates through
    the interpolation, the detector histogram then drops such rays."""
    from sloth_amd.core import engine
    from sloth_amd.core.beam import Beam
    from sloth_amd.fields import kolmogorov_field

    L = 6e-3
    ne, coords = kolmogorov_field(40, L, seed=9)
    grid = engine.PackedGrid(ne, *coords, 2 * np.pi * 299792458.0 / 532e-9, inv_brems=True,
                             Te=20.0, Z=1.0)
    good = Beam(3000, 2.5e-3, 0.0, L / 2, beam_type="square", seeded=True).s0
    bad = good[:, :6].copy()
    bad[0, 0] = np.nan                    # position
    bad[4, 1] = np.nan                    # velocity
    bad[2, 2] = np.inf                    # position
    bad[5, 3] = np.inf                    # velocity
    bad[1, 4] = -np.inf
    bad[0, 5] = 1e38                      # finite but far outside every coordinate range
    mixed = np.concatenate([good[:, :1500], bad, good[:, 1500:]], axis=1)
    t = np.array([0.0, 0.6 * L / 299792458.0, 1.5 * L / 299792458.0])
    a = engine.trace(grid, good, t, save_all=True, want_path=True)
    b = engine.trace(grid, mixed, t, save_all=True, want_path=True)
    keep = np.r_[0:1500, 1506:mixed.shape[1]]
    st_a, st_b = a.states.cpu().numpy(), b.states.cpu().numpy()
    np.testing.assert_array_equal(st_b[:, :, keep], st_a)
    np.testing.assert_array_equal(b.tau.cpu().numpy()[:, keep], a.tau.cpu().numpy())
    np.testing.assert_array_equal(b.path.cpu().numpy()[keep], a.path.cpu().numpy())
    np.testing.assert_array_equal(b.steps.cpu().numpy()[keep], a.steps.cpu().numpy())
    bad_steps = b.steps.cpu().numpy()[1500:1506].tolist()
    assert bad_steps[0] == bad_steps[2] == bad_steps[4] == bad_steps[5] == 0, bad_steps
    assert bad_steps[1] <= 1 and bad_steps[3] <= 1, bad_steps
    assert int(b.total_steps.item()) == int(a.total_steps.item()) + sum(bad_steps)
    fin = st_b[-1][:, 1500:1506]
    assert not np.isfinite(fin[:, :5]).all(axis=0).any()
    assert np.isfinite(fin[:, 5]).all()   # the far-away ray stays finite and moves straight
    np.testing.assert_allclose(fin[0, 5], 1e38, rtol=1e-6)


def test_const_device_shares_reads_across_streams_and_evicts():
    """Small read-only uploads are shared by content; a hit on another stream waits for the
    upload; entries past the cap are evicted and new ones still hold the right values."""
    import torch

    from sloth_amd.core import engine

    dev = torch.device("cuda", 0)
    a = np.linspace(-1.0, 1.0, 257).astype(np.float32)
    side = torch.cuda.Stream(dev)
    t1 = engine.const_device(a, dev, side)
    t2 = engine.const_device(a.copy(), dev)          # same bytes: the same tensor
    assert t1 is t2
    t3 = engine.const_device(a.astype(np.float64), dev)
    assert t3 is not t1 and t3.dtype == torch.float64
    with torch.cuda.stream(side):
        out = (t2 * 2.0).cpu().numpy()               # read on the side stream
    np.testing.assert_array_equal(out, 2.0 * a)
    torch.cuda.synchronize()
    # fill past the cap: eviction waits for the device, later uploads are exact
    for i in range(engine._CONST_MAX + 8):
        v = np.full(3, float(i), np.float64)
        np.testing.assert_array_equal(engine.const_device(v, dev).cpu().numpy(), v)
    assert len(engine._CONST) <= engine._CONST_MAX
    big = np.zeros((1 << 20) // 4 + 1, np.float32)   # above 1 MiB: an ordinary copy
    assert engine.const_device(big, dev) is not engine.const_device(big, dev)

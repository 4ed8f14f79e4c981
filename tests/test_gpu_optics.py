"""k5 optics on the GPU against the NumPy oracle: every diagnostic chain, the element
functions on host arrays, and detector images after a solve."""
import numpy as np
import pytest

from oracle import optics as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _jvec(n, seed, dtype=np.float32, half=12e-3, ang=2e-3):
    """A solve-like (4, N) Jones vector in metres (float32 as the solver returns it)."""
    rng = np.random.default_rng(seed)
    r = np.empty((4, n))
    r[0] = rng.uniform(-half, half, n)
    r[2] = rng.uniform(-half, half, n)
    r[1] = rng.normal(0, ang, n)
    r[3] = rng.normal(0, ang, n)
    return r.astype(dtype)


def _same(got, ref, rtol=1e-13):
    """NaN where the oracle has NaN; finite values equal to fp64 rounding."""
    got = np.asarray(got)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    f = ~np.isnan(ref)
    scale = np.maximum(np.abs(ref[f]), 1e-30)
    err = np.max(np.abs(got[f] - ref[f]) / scale) if f.any() else 0.0
    assert err <= rtol, err
    return err


def _edge_distance(v, edges):
    i = np.clip(np.searchsorted(edges, v), 1, len(edges) - 1)
    return np.minimum(np.abs(v - edges[i - 1]), np.abs(v - edges[i]))


CHAINS = [("single_lens_custom", {}), ("single_lens", {}), ("two_lens", {}),
          ("DF", {"R_stop": 1}), ("LF", {"R_stop": 1}), ("incoherent_custom", {}),
          ("incoherent", {})]


@pytest.mark.parametrize("name,kw", CHAINS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_chain_matches_oracle(name, kw, dtype):
    from sloth_amd.processing import diagnostics as D

    rf = _jvec(50000, 11, dtype, half=20e-3)
    L, R, fp = 400, 25, 3.5
    ref = O.apply(O.m_to_mm(rf), O.chain(name, L=L, R=R, focal_plane=fp, **kw))
    cls = {"single_lens_custom": D.Shadowgraphy, "single_lens": D.Shadowgraphy,
           "two_lens": D.Shadowgraphy, "DF": D.Schlieren, "LF": D.Schlieren,
           "incoherent_custom": D.Refractometry, "incoherent": D.Refractometry}[name]
    d = cls(rf, L=L, R=R, focal_plane=fp)
    method = {"single_lens_custom": "single_lens_custom_solve", "single_lens": "single_lens_solve",
              "two_lens": "two_lens_solve", "DF": "DF_solve", "LF": "LF_solve",
              "incoherent_custom": "incoherent_custom_solve",
              "incoherent": "incoherent_solve"}[name]
    getattr(d, method)(**({"R": kw["R_stop"]} if "R_stop" in kw else {}))
    kept = (~np.isnan(ref[0])).mean()
    assert 0.02 < kept < 0.999, kept          # the chain both keeps and rejects rays
    _same(d.rf, ref)
    assert d.rf.dtype == np.float64
    # r0 is m_to_mm of the input, in the input's precision
    np.testing.assert_array_equal(d.r0, O.m_to_mm(rf))


def test_element_functions_on_host_arrays_match_oracle():
    from sloth_amd.processing import diagnostics as D

    r = O.m_to_mm(_jvec(4000, 12, np.float64))
    _same(D.travel(r, 321.5), O.travel(r, 321.5))
    _same(D.lens(r, 200, 150.0), O.lens(r, 200, 150.0))
    _same(D.sym_lens(r, 77.0), O.lens(r, 77.0, 77.0))
    for fn, args in ((D.circular_aperture, (6.0,)), (D.circular_stop, (4.0,)),
                     (D.rect_aperture, (5, 7)), (D.knife_edge, (0.5, "y", -1))):
        mine, theirs = r.copy(), r.copy()
        out = fn(mine, *args)
        assert out is mine                                   # in place, as the reference
        ref = getattr(O, fn.__name__)(theirs, *args)
        _same(mine, ref, rtol=0)
    mask = D.annular_stop(r, 3.0, 8.0)
    s = r[0] ** 2 + r[2] ** 2
    np.testing.assert_array_equal(mask, (s > 9.0) & (s < 64.0))


def test_solved_detector_image_matches_oracle_histogram():
    from sloth_amd.processing import diagnostics as D

    rf = _jvec(200000, 13, np.float32, half=8e-3, ang=5e-3)
    w = np.random.default_rng(14).uniform(0.2, 1.0, rf.shape[1])
    d = D.Shadowgraphy(rf, weights=w, L=400, R=25, Lx=18, Ly=13.5)
    d.single_lens_solve()
    d.histogram(bin_scale=8, pix_x=3448, pix_y=2574)
    ref_rf = O.apply(O.m_to_mm(rf), O.chain("single_lens", L=400, R=25))
    H, xe, ye = O.histogram(ref_rf, Lx=18, Ly=13.5, pix_x=3448, pix_y=2574, bin_scale=8,
                            weights=w)
    np.testing.assert_array_equal(d.xedges, xe)
    np.testing.assert_array_equal(d.yedges, ye)
    # positions agree to fp64 rounding: only a ray within that of a bin edge could move
    fin = ~np.isnan(ref_rf[0])
    edge = int(((_edge_distance(ref_rf[0, fin], xe) < 1e-9) |
                (_edge_distance(ref_rf[2, fin], ye) < 1e-9)).sum())
    assert np.abs(d.H - H).sum() <= 2 * edge * w.max() + 1e-9 * H.sum()
    # counts (no weights) and the CCD mask, through the same solve
    c = D.Shadowgraphy(rf, L=400, R=25, Lx=18, Ly=13.5, ccd_shape_m=(13.5e-3, 9e-3))
    c.single_lens_solve()
    c.histogram(bin_scale=8)
    Hc, _, _ = O.histogram(ref_rf, Lx=18, Ly=13.5, pix_x=3448, pix_y=2574, bin_scale=8,
                           ccd_half_mm=(6.75, 4.5))
    assert np.abs(c.H - Hc).sum() <= 2 * edge
    assert 0 < c.H.sum() < fin.sum()


def test_device_tensor_input_stays_on_device_until_read():
    import torch

    from sloth_amd.processing import diagnostics as D

    rf = _jvec(3000, 15, np.float32)
    t = torch.from_numpy(rf).cuda()
    d = D.Schlieren(t, L=400, R=25)
    d.LF_solve(R=2)
    assert isinstance(d._rf, torch.Tensor) and d._rf.is_cuda
    ref = O.apply(O.m_to_mm(rf), O.chain("LF", L=400, R=25, R_stop=2))
    _same(d.rf, ref)
    d.histogram(pix_x=100, pix_y=100)
    H, _, _ = O.histogram(ref, pix_x=100, pix_y=100)
    np.testing.assert_array_equal(d.H, H)
    d.plot_rays(pix_x=100, pix_y=100)                       # plain: bins r0 (mm, unsolved)
    Hp, _, _ = O.histogram(O.m_to_mm(rf).astype(np.float64), pix_x=100, pix_y=100)
    np.testing.assert_array_equal(d.H, Hp)


def test_helpers_lens_cutoff_ccd_mask_transmission():
    from sloth_amd.processing import diagnostics as D

    rf = _jvec(5000, 16, np.float64, half=20.0, ang=0.05)
    got = D.lens_cutoff(rf, L=400, R=25)
    keep = np.pow(np.pow(400 * np.tan(rf[1]) + rf[0], 2) + np.pow(400 * np.tan(rf[3]) + rf[2],
                                                                 2), 0.5) <= 25
    assert got.shape[1] == keep.sum() or abs(got.shape[1] - keep.sum()) <= 1
    w = np.linspace(0.1, 1.0, rf.shape[1])
    out, wo, inside = D.apply_ccd_mask(rf * 1e-3, w, ccd_shape_m=(13.5e-3, 18e-3))
    ref_inside = (np.abs(rf[0] * 1e-3) <= 6.75e-3) & (np.abs(rf[2] * 1e-3) <= 9e-3)
    np.testing.assert_array_equal(inside, ref_inside)
    assert np.isnan(out[:, ~inside]).all() and (wo[~inside] == 0).all()
    Hw = np.array([[0.0, 2.0], [3.0, 0.5]])
    Hp = np.array([[0.0, 4.0], [3.0, 1.0]])
    T = D.transmission_map(Hw, Hp)
    assert np.isnan(T[0, 0]) and T[0, 1] == 0.5 and T[1, 0] == 1.0 and T[1, 1] == 0.5

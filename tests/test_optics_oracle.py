"""Pins of the optics oracle (oracle/optics.py) and host-side checks of the optics API.

The oracle restates the reference's element functions; these tests check it
against what geometric optics fixes independently: imaging conditions of the
diagnostics' lens systems, the focus of a thin lens, and the boundaries of
apertures, stops, the rectangular aperture and the knife edge.
"""
import ctypes

import numpy as np
import pytest

from oracle import optics as O


def _rays(n, seed=0, half=10.0, ang=2e-3):
    rng = np.random.default_rng(seed)
    r = np.empty((4, n))
    r[0] = rng.uniform(-half, half, n)
    r[2] = rng.uniform(-half, half, n)
    r[1] = rng.normal(0, ang, n)
    r[3] = rng.normal(0, ang, n)
    return r


def test_two_lens_telescope_images_with_magnification_plus_one():
    # travel L, lens L/2, travel 2L, lens L/2, travel L: system matrix [[1, 0], [4/L, 1]]
    L = 400.0
    r0 = _rays(1000, 1, half=5.0, ang=1e-3)
    out = O.apply(r0, O.chain("two_lens", L=L, R=1e9))
    np.testing.assert_allclose(out[0], r0[0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(out[2], r0[2], rtol=0, atol=1e-12)
    np.testing.assert_allclose(out[1], r0[1] + 4 * r0[0] / L, rtol=0, atol=1e-15)


def test_single_lens_images_with_magnification_minus_two():
    # object 3L/4 before a lens of L/2, image 3L/2 behind it: 1/f = 1/do + 1/di, M = -di/do
    L = 400.0
    r0 = _rays(1000, 2, half=5.0, ang=1e-3)
    out = O.apply(r0, O.chain("single_lens", L=L, R=1e9))
    np.testing.assert_allclose(out[0], -2 * r0[0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(out[2], -2 * r0[2], rtol=0, atol=1e-12)
    np.testing.assert_allclose(out[1], -2 * r0[0] / L - r0[1] / 2, rtol=0, atol=1e-15)


def test_lens_focuses_parallel_rays_and_travel_is_a_shear():
    f = 250.0
    r = _rays(500, 3)
    r[1] = r[3] = 0.0
    focal = O.travel(O.lens(r, f, 2 * f), f)
    assert np.max(np.abs(focal[0])) < 1e-12 * np.max(np.abs(r[0]))
    np.testing.assert_allclose(focal[2], r[2] / 2, rtol=1e-14)      # y lens has 2f
    t = O.travel(_rays(10, 4), 123.0)
    base = _rays(10, 4)
    np.testing.assert_allclose(t[0], base[0] + 123.0 * base[1], rtol=1e-15)
    np.testing.assert_array_equal(t[1], base[1])


def test_aperture_stop_rectangle_and_knife_edge_boundaries():
    R = 25.0
    r = np.zeros((4, 6))
    r[0] = [25.0, 25.0 + 1e-9, 0.0, 17.0, 18.0, -3.0]
    r[2] = [0.0, 0.0, 25.0, 17.0, 18.0, 4.0]
    a = O.circular_aperture(r.copy(), R)
    # kept on the rim (strict >), dropped just outside; 17^2 * 2 = 578 < 625 kept; 648 > 625
    assert list(np.isnan(a[0])) == [False, True, False, False, True, False]
    assert np.isnan(a[:, 1]).all()                        # all four rows become NaN
    s = O.circular_stop(r.copy(), 5.0)                    # strict <: only |(-3, 4)| = 5 stays
    assert list(np.isnan(s[0])) == [False, False, False, False, False, False]
    s = O.circular_stop(r.copy(), 5.0 + 1e-9)
    assert list(np.isnan(s[0])) == [False, False, False, False, False, True]
    q = np.zeros((4, 4))
    q[0] = [20.0, 20.0, 0.0, -16.0]
    q[2] = [0.0, 40.0, 40.0, -31.0]
    rect = O.rect_aperture(q.copy(), 15, 30)              # rejects only when both are outside
    assert list(np.isnan(rect[0])) == [False, True, False, True]
    k = O.knife_edge(q.copy(), 1.0, "x", +1)
    assert list(np.isnan(k[0])) == [True, True, False, False]
    k = O.knife_edge(q.copy(), 1.0, "y", -1)
    assert list(np.isnan(k[0])) == [True, False, False, True]


def test_nan_rays_stay_nan_through_linear_elements():
    r = _rays(4, 5)
    r[:, 2] = np.nan
    out = O.lens(O.travel(r, 10.0), 3.0, 4.0)
    assert np.isnan(out[:, 2]).all() and np.isfinite(out[:, [0, 1, 3]]).all()


def test_m_to_mm_keeps_float32_and_rounds_in_it():
    r = np.float32(_rays(100, 6) * 1e-3)
    mm = O.m_to_mm(r)
    assert mm.dtype == np.float32
    np.testing.assert_array_equal(mm[0], r[0] * np.float32(1000.0))
    np.testing.assert_array_equal(mm[1], r[1])


def test_oracle_histogram_is_numpy_with_masks():
    r = _rays(2000, 7, half=12.0)
    r[:, :5] = np.nan
    H, xe, ye = O.histogram(r, Lx=18, Ly=13.5, pix_x=36, pix_y=27, ccd_half_mm=(6.75, 9.0))
    keep = ~np.isnan(r[0]) & (np.abs(r[0]) <= 6.75) & (np.abs(r[2]) <= 9.0)
    H2, _, _ = np.histogram2d(r[0, keep], r[2, keep], bins=[36, 27], range=[[-9, 9],
                                                                          [-6.75, 6.75]])
    np.testing.assert_array_equal(H, H2.T)


def test_rtm_abi_rejects_bad_elements_without_a_device():
    from sloth_amd import _native as N

    L = N.lib()

    def call(ops, n_ops=None):
        arr = (N.Optic * max(len(ops), 1))(*ops)
        return L.slt_rtm(None, 0, 0, 0, 1.0, arr, len(ops) if n_ops is None else n_ops, None,
                         0, None)

    ok = N.Optic(N.SLT_OPTIC_TRAVEL, 0, 1.0, 0.0)
    assert call([ok]) == N.SLT_OK                              # n = 0: nothing to do
    assert call([N.Optic(9, 0, 0.0, 0.0)]) == N.SLT_ERR_INVALID
    assert "unknown kind" in N.last_error()
    assert call([N.Optic(N.SLT_OPTIC_KNIFE_EDGE, 1, 0.0, 1.0)]) == N.SLT_ERR_INVALID
    assert call([N.Optic(N.SLT_OPTIC_KNIFE_EDGE, 0, 0.0, 0.0)]) == N.SLT_ERR_INVALID
    assert call([ok], n_ops=N.SLT_MAX_OPTICS + 1) == N.SLT_ERR_INVALID


def test_knife_edge_errors_like_the_reference(capsys):
    from sloth_amd.processing.diagnostics import knife_edge

    with pytest.raises(UnboundLocalError):
        knife_edge(np.zeros((4, 3)), 0.0, "x", 0)
    assert "Direction must be <0 or >0" in capsys.readouterr().out
    with pytest.raises(UnboundLocalError):
        knife_edge(np.zeros((4, 3)), 0.0, "z", 1)


def test_optic_struct_matches_header():
    from sloth_amd import _native as N

    assert ctypes.sizeof(N.Optic) == 24 and N.Optic.a.offset == 8

"""The reference's own acceptance tests (tests/test_quadratic_trough.py), run through the
drop-in API of sloth_amd with the same domains, parameters and tolerances.

Test classes mirrored: TestQuadraticTroughNoBrems (:195-294), TestQuadraticTroughWithBrems
(:301-372), TestBenchmark (:379-467), TestAPIContracts (:530-592), TestAbsorptionFormula
(:599-663), TestDiagnosticWeighting (:670-753), the histogram parts of
TestPrefilterAndAmplitudeHelpers (:760-879) and TestCCDMask's Diagnostic part (:1066-1115).
Plotting helpers, transmission_map, apply_ccd_mask and absorption_sanity_check are analysis
utilities outside the traced path and are not provided.
"""
import os
import pickle
import tempfile
import time
import warnings

import numpy as np
import pytest

from _parity import C

pytestmark = pytest.mark.gpu

_ALPHA = 3.14207787e-4
_LWL = 1064e-9
_OMEGA = 2 * np.pi * C / _LWL
_A = 5e-3
_HALF_Z = 11e-3
_DEPTH_MAX = 20e-3
_ANGLE = np.pi / 4.0
_T_TRACE = _DEPTH_MAX / C
_OMEGA_OSC = _ANGLE / _T_TRACE
_NE_0 = (_OMEGA_OSC * _A * np.sqrt(_ALPHA) * _OMEGA / C) ** 2
_X0_VALS = np.array([0.5e-3, 1.0e-3, 1.5e-3, 2.0e-3])
_TOL_X_M = 0.2e-3
_TOL_VX_REL = 0.10


@pytest.fixture(scope="module", autouse=True)
def _dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


class _QuadraticTroughDomain:
    def __init__(self, ne_0, a, half_z, n_xy=24, n_z=24, inv_brems=False, Te=None, Z=None):
        x = np.linspace(-a, a, n_xy)
        y = np.linspace(-a, a, n_xy)
        z = np.linspace(-half_z, half_z, n_z)
        XX, _, _ = np.meshgrid(x, y, z, indexing="ij")
        self.ne = (ne_0 * (XX / a) ** 2).astype(np.float32)
        self.x = np.asarray(x, dtype=np.float32)
        self.y = np.asarray(y, dtype=np.float32)
        self.z = np.asarray(z, dtype=np.float32)
        self.probing_direction = "z"
        self.lengths = np.array([x[-1] - x[0], y[-1] - y[0], z[-1] - z[0]], dtype=np.float32)
        self.dims = np.array([len(x), len(y), len(z)], dtype=np.int32)
        self.Np_total = None
        self.inv_brems = inv_brems
        self.Te = np.float32(Te) if Te is not None else None
        self.Z = np.float32(Z) if Z is not None else None


def _collimated_rays(x0_vals, z_start):
    s0 = np.zeros((6, len(x0_vals)))
    s0[0] = x0_vals
    s0[2] = z_start
    s0[5] = C
    return s0.astype(np.float32)


def _tsd(*a, **k):
    from sloth_amd.core.propagator import trace_and_save_depths

    return trace_and_save_depths(*a, **k)


# -- TestQuadraticTroughNoBrems ---------------------------------------------

def _trough_run(comps=(0, 1), n_xy=24, n_z=24):
    dom = _QuadraticTroughDomain(_NE_0, _A, _HALF_Z, n_xy=n_xy, n_z=n_z)
    s0 = _collimated_rays(_X0_VALS, z_start=-_HALF_Z)
    return _tsd(s0, dom, step=_DEPTH_MAX, depth_max=_DEPTH_MAX, output_path=None, lwl=_LWL,
                jones_components=list(comps), jitted=True, verbose=False)


def test_x_position_matches_analytic():
    x_final = np.asarray(_trough_run()["jvec"][-1][0])
    np.testing.assert_allclose(x_final, _X0_VALS * np.cos(_ANGLE), atol=_TOL_X_M)


def test_x_velocity_matches_analytic():
    vx_final = np.asarray(_trough_run()["jvec"][-1][1]) * C
    vx_expect = -_X0_VALS * _OMEGA_OSC * np.sin(_ANGLE)
    np.testing.assert_allclose(vx_final, vx_expect, rtol=_TOL_VX_REL)


def test_y_position_unchanged():
    y_final = np.asarray(_trough_run(comps=(2, 3), n_xy=20, n_z=20)["jvec"][-1][0])
    np.testing.assert_allclose(y_final, 0.0, atol=0.1e-3)


def test_vacuum_straight_rays():
    n, half = 16, 2e-3
    coords = np.float32(np.linspace(-half, half, n))

    class _Vac:
        ne = np.zeros((n, n, n), dtype=np.float32)
        x = y = z = coords
        probing_direction = "z"
        lengths = np.array([2 * half] * 3, dtype=np.float32)
        dims = np.array([n] * 3, dtype=np.int32)
        Np_total = None
        inv_brems = False

    x0 = np.array([0.5e-3, 1.0e-3])
    res = _tsd(_collimated_rays(x0, -half), _Vac(), step=2 * half, depth_max=2 * half,
               output_path=None, lwl=_LWL, jones_components=[0], jitted=True, verbose=False)
    np.testing.assert_allclose(np.asarray(res["jvec"][-1][0]), x0, atol=1e-6)


# -- TestQuadraticTroughWithBrems -------------------------------------------

def _brems_run():
    dom = _QuadraticTroughDomain(_NE_0, _A, _HALF_Z, n_xy=20, n_z=20, inv_brems=True, Te=100.0,
                                 Z=1.0)
    return _tsd(_collimated_rays(_X0_VALS, -_HALF_Z), dom, step=_DEPTH_MAX / 4,
                depth_max=_DEPTH_MAX, output_path=None, lwl=_LWL, jones_components=[0, 1],
                jitted=True, verbose=False)


def test_brems_x_position_and_amplitude():
    res = _brems_run()
    np.testing.assert_allclose(np.asarray(res["jvec"][-1][0]), _X0_VALS * np.cos(_ANGLE),
                               atol=_TOL_X_M)
    amps = [float(np.asarray(a).mean()) for a in res["amplitude"]]
    assert amps[-1] < 1.0
    for i in range(1, len(amps)):
        assert amps[i] <= amps[i - 1] + 1e-6
    jv = np.asarray(res["jvec"][-1])
    amp = np.asarray(res["amplitude"][-1])
    assert np.sqrt(np.nanmean((jv * amp) ** 2)) <= np.sqrt(np.nanmean(jv ** 2)) + 1e-10


# -- TestBenchmark -----------------------------------------------------------

def test_inv_brems_overhead_and_geometry_identical():
    rng = np.random.RandomState(0)
    s0 = _collimated_rays(_A * 0.3 * (rng.rand(30) - 0.5), -_HALF_Z)
    plain = _QuadraticTroughDomain(_NE_0, _A, _HALF_Z, n_xy=16, n_z=16)
    ib = _QuadraticTroughDomain(_NE_0, _A, _HALF_Z, n_xy=16, n_z=16, inv_brems=True, Te=100.0,
                                Z=1.0)
    kw = dict(step=_DEPTH_MAX, depth_max=_DEPTH_MAX, output_path=None, lwl=_LWL,
              jones_components=[0, 1], jitted=True, verbose=False)
    _tsd(s0, plain, **kw), _tsd(s0, ib, **kw)
    t0 = time.perf_counter()
    r_plain = _tsd(s0, plain, **kw)
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    r_ib = _tsd(s0, ib, **kw)
    t_ib = time.perf_counter() - t0
    assert t_ib / max(t_plain, 1e-9) < 10.0
    np.testing.assert_allclose(np.asarray(r_ib["jvec"][-1][0]),
                               np.asarray(r_plain["jvec"][-1][0]), atol=1e-5)


# -- TestAPIContracts --------------------------------------------------------

def _vacuum_domain(half=2e-3, n=12):
    class D:
        pass

    d = D()
    c = np.float32(np.linspace(-half, half, n))
    d.ne = np.zeros((n, n, n), np.float32)
    d.x = d.y = d.z = c
    d.probing_direction = "z"
    d.lengths = np.array([2 * half] * 3, np.float32)
    d.dims = np.array([n] * 3, np.int32)
    d.Np_total = None
    d.inv_brems = False
    return d


def _small_rays(Np=8, z_start=-2e-3):
    s0 = np.zeros((6, Np))
    s0[0] = np.linspace(-5e-4, 5e-4, Np)
    s0[2] = z_start
    s0[5] = C
    return s0.astype(np.float32)


def _api(domain, **kw):
    d = dict(step=500e-6, depth_max=1e-3, output_path=None, lwl=1064e-9, jitted=True,
             verbose=False)
    d.update(kw)
    return _tsd(_small_rays(), domain, **d)


def test_api_contracts():
    res = _api(_vacuum_domain())
    assert {"depth_saves", "jvec", "jones_components"} <= set(res)
    assert "amplitude" not in res and "jvec_unweighted" not in res
    res = _api(_vacuum_domain(), step=200e-6, depth_max=1e-3)
    np.testing.assert_allclose(np.diff(res["depth_saves"]), 200e-6, rtol=1e-6)
    with pytest.raises(ValueError, match="smaller than the requested depth_max"):
        _api(_vacuum_domain(half=2e-3), depth_max=10e-3)
    full = _api(_vacuum_domain())
    pos = _api(_vacuum_domain(), jones_components="position")
    for j in range(len(full["jvec"])):
        np.testing.assert_array_equal(pos["jvec"][j], full["jvec"][j][[0, 2], :])
    with tempfile.NamedTemporaryFile(suffix=".pkl", delete=False) as f:
        path = f.name
    try:
        r = _api(_vacuum_domain(), output_path=path)
        with open(path, "rb") as fh:
            loaded = pickle.load(fh)
        np.testing.assert_array_equal(loaded["depth_saves"], r["depth_saves"])
        np.testing.assert_array_equal(loaded["jvec"][0], r["jvec"][0])
    finally:
        os.remove(path)


# -- TestAbsorptionFormula ---------------------------------------------------

def test_analytic_amplitude_match():
    from sloth_amd.core.propagator import kappa_inv_brems

    half, depth = 3e-3, 2e-3
    omega = 2 * np.pi * C / 1064e-9
    kappa = float(kappa_inv_brems(np.float32(1e25), np.float32(100.0), 1.0, omega))
    n = 16

    class D:
        pass

    d = D()
    c = np.float32(np.linspace(-half, half, n))
    d.ne = np.full((n, n, n), 1e25, np.float32)
    d.x = d.y = d.z = c
    d.probing_direction = "z"
    d.lengths = np.array([2 * half] * 3, np.float32)
    d.dims = np.array([n] * 3, np.int32)
    d.Np_total = None
    d.inv_brems = True
    d.Te = np.float32(100.0)
    d.Z = np.float32(1.0)
    res = _tsd(_small_rays(10, -half), d, step=500e-6, depth_max=depth, output_path=None,
               lwl=1064e-9, jitted=True, verbose=False)
    amp = float(np.asarray(res["amplitude"][-1]).mean())
    np.testing.assert_allclose(amp, np.exp(-kappa * depth / C), rtol=0.05)


@pytest.mark.parametrize("Te_eV, Z_val", [(100.0, 1.0), (10.0, 4.0)])
def test_kappa_coulomb_log_formula(Te_eV, Z_val):
    from scipy.constants import e as e_charge, epsilon_0 as eps0

    from sloth_amd.core.propagator import kappa_inv_brems

    omega = 2 * np.pi * C / 1064e-9
    ne_cc = 1e25 * 1e-6
    v_the = 4.19e5 * np.sqrt(Te_eV)
    o_pe = 5.64e4 * np.sqrt(ne_cc)
    b_min = Z_val * e_charge / (4.0 * np.pi * eps0 * Te_eV)
    CL = max(2.0, np.log(v_the / (max(o_pe, omega) * b_min)))
    kappa_ref = 3.1e-5 * Z_val * C * (ne_cc / omega) ** 2 * CL * Te_eV ** (-1.5)
    got = float(kappa_inv_brems(np.float32(1e25), np.float32(Te_eV), float(Z_val), omega))
    np.testing.assert_allclose(got, kappa_ref, rtol=1e-4)


def test_negative_te_raises_and_overcritical_warns():
    from sloth_amd.core.domain import ScalarDomain
    from sloth_amd.core.propagator import precompute_gradients

    with pytest.raises(AssertionError, match="eV"):
        ScalarDomain([4e-3, 4e-3, 4e-3], [8, 8, 8], inv_brems=True, Te=-10.0, Z=1.0)
    omega = 2 * np.pi * C / 1064e-9
    ncr = _ALPHA * omega ** 2
    c = np.float32(np.linspace(-3e-3, 3e-3, 8))
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        precompute_gradients(np.full((8, 8, 8), 1.5 * ncr, np.float32), c, c, c, omega)
    assert len(caught) == 1 and issubclass(caught[0].category, UserWarning)


# -- TestDiagnosticWeighting / TestPrefilterAndAmplitudeHelpers -----------------

def _weighting_runs(seed, scale, np_):
    rng = np.random.RandomState(seed)
    s0 = _collimated_rays(_A * scale * (rng.rand(np_) - 0.5), -_HALF_Z)
    kw = dict(step=_DEPTH_MAX, depth_max=_DEPTH_MAX, output_path=None, lwl=_LWL,
              jones_components=[0, 1, 2, 3], jitted=True, verbose=False)
    plain = _tsd(s0, _QuadraticTroughDomain(_NE_0, _A, _HALF_Z), **kw)
    ib = _tsd(s0, _QuadraticTroughDomain(_NE_0, _A, _HALF_Z, inv_brems=True, Te=100.0, Z=1.0),
              **kw)
    return plain, ib


def test_diagnostic_weighting_contract():
    from sloth_amd.processing.diagnostics import Diagnostic

    plain, ib = _weighting_runs(7, 0.4, 30)
    jvec = plain["jvec"][-1]
    amp = np.asarray(ib["amplitude"][-1])
    d_plain = Diagnostic(jvec, L=100, R=50, Lx=50, Ly=50)
    d_wt = Diagnostic(jvec, weights=amp, L=100, R=50, Lx=50, Ly=50)
    d_plain.histogram(pix_x=50, pix_y=50)
    d_wt.histogram(pix_x=50, pix_y=50)
    assert d_plain.H.sum() == pytest.approx(30, abs=2)
    assert d_wt.H.sum() < d_plain.H.sum()
    np.testing.assert_array_equal(d_plain.xedges, d_wt.xedges)
    np.testing.assert_array_equal(d_plain.yedges, d_wt.yedges)


def test_prefilter_and_unit_and_low_weights():
    from sloth_amd.processing.diagnostics import Diagnostic

    plain, ib = _weighting_runs(99, 0.5, 40)
    jvec = plain["jvec"][-1]
    d = Diagnostic(jvec, L=100, R=1, prefilter_input=True)
    d.histogram(pix_x=20, pix_y=20)
    assert d.H.sum() <= 40
    amp = np.asarray(ib["amplitude"][-1])
    d = Diagnostic(ib["jvec"][-1], weights=amp, L=100, R=1, prefilter_input=True)
    assert d.rf.shape[-1] == len(d.weights)
    d.histogram(pix_x=20, pix_y=20)
    n = np.asarray(jvec).shape[-1]
    a = Diagnostic(jvec, L=100, R=50, Lx=50, Ly=50)
    b = Diagnostic(jvec, weights=np.ones(n), L=100, R=50, Lx=50, Ly=50)
    lo = Diagnostic(jvec, weights=np.full(n, 0.1), L=100, R=50, Lx=50, Ly=50)
    for x in (a, b, lo):
        x.histogram(pix_x=30, pix_y=30)
    np.testing.assert_array_equal(a.H, b.H)
    assert abs(lo.H.sum() / a.H.sum() - 0.1) < 0.01


def test_diagnostic_ccd_mask_contract():
    from sloth_amd.processing.diagnostics import Diagnostic

    rng = np.random.RandomState(42)
    rf_m = np.zeros((4, 500))
    rf_m[0] = (rng.rand(500) - 0.5) * 40e-3
    rf_m[2] = (rng.rand(500) - 0.5) * 40e-3
    rf_m[1] = rng.randn(500) * 1e-4
    rf_m[3] = rng.randn(500) * 1e-4
    rng = np.random.RandomState(42)
    rf_mm = np.zeros((4, 500))
    rf_mm[0] = (rng.rand(500) - 0.5) * 40
    rf_mm[2] = (rng.rand(500) - 0.5) * 40
    rf_mm[1] = rng.randn(500) * 1e-4
    rf_mm[3] = rng.randn(500) * 1e-4
    kw = dict(L=100, R=500, Lx=50, Ly=50)
    no_ccd = Diagnostic(rf_m, **kw)
    ccd = Diagnostic(rf_m, ccd_shape_m=(13.5e-3, 18e-3), **kw)
    no_ccd.rf = rf_mm.copy()
    ccd.rf = rf_mm.copy()
    no_ccd.histogram(pix_x=20, pix_y=20)
    ccd.histogram(pix_x=20, pix_y=20)
    assert ccd.H.sum() < no_ccd.H.sum()
    d0 = Diagnostic(rf_m, **kw)
    d1 = Diagnostic(rf_m, ccd_shape_m=None, **kw)
    d0.histogram(pix_x=20, pix_y=20)
    d1.histogram(pix_x=20, pix_y=20)
    np.testing.assert_array_equal(d0.H, d1.H)

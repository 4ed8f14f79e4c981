"""Host-side logic and the C ABI surface (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C = 299792458.0


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "sloth_hip.h")).read()
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(slt_\w+)\s*\(", text, re.M)))


def test_constants_equal_scipy():
    import scipy.constants as sc

    from sloth_amd import constants as K

    assert K.C_LIGHT == sc.c and K.E_CHARGE == sc.e and K.EPSILON_0 == sc.epsilon_0


def test_library_loads_and_exports_every_declared_symbol():
    from sloth_amd import _native as N

    lib = ctypes.CDLL(N.LIB_PATH)
    declared = _header_symbols()
    assert len(declared) >= 10
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert set(N.EXPORTED) == set(declared)
    L = N.lib()
    assert L.slt_version() == N.SLT_ABI_VERSION == 3
    assert L.slt_last_error() in (b"", None) or isinstance(L.slt_last_error(), bytes)


def _desc(x, y, z, axis=2, kappa=0):
    from sloth_amd import _native as N

    L = N.lib()
    d = N.GridDesc()
    xs, ys, zs = (np.ascontiguousarray(c, dtype=np.float32) for c in (x, y, z))
    rc = L.slt_grid_desc_init(xs.ctypes.data, ys.ctypes.data, zs.ctypes.data, xs.size, ys.size,
                              zs.size, axis, kappa, ctypes.byref(d))
    return rc, d


def test_grid_desc_layout_probing_axis_fastest():
    x = np.float32(np.linspace(-1e-3, 1e-3, 5))
    y = np.float32(np.linspace(-2e-3, 2e-3, 6))
    z = np.float32(np.linspace(0, 7e-3, 7))
    rc, d = _desc(x, y, z, 2)
    assert rc == 0 and list(d.stride) == [42, 7, 1] and d.nchan == 3
    rc, d = _desc(x, y, z, 0, 1)
    assert rc == 0 and list(d.stride) == [1, 35, 5] and d.nchan == 4
    rc, d = _desc(x, y, z, 1)
    assert rc == 0 and list(d.stride) == [42, 1, 6]
    assert d.lo[0] == x[0] and d.hi[2] == z[-1]
    assert d.inv_d[1] == pytest.approx(5 / (float(y[-1]) - float(y[0])), rel=1e-6)
    from sloth_amd import _native as N

    assert N.lib().slt_grid_bytes(ctypes.byref(d)) == 5 * 6 * 7 * 16


def test_grid_desc_classifies_and_rejects_coordinates():
    """A linspace (up to fp32 rounding) is uniform; any other strictly ascending axis
    makes the grid non-uniform, with knots in grid units; an axis that is not strictly
    ascending is rejected, as the reference interpolator rejects it."""
    from sloth_amd import _native as N

    c = np.float32(np.linspace(0, 1, 8))
    rc, d = _desc(c, c, c)
    assert rc == 0 and d.uniform == 1 and not d.knots
    wobble = c.copy()
    wobble[3] += np.float32(1e-7)                    # within fp32 rounding of the linspace
    assert _desc(wobble, c, c)[1].uniform == 1
    bad = c.copy()
    bad[3] += 0.05
    rc, d = _desc(bad, c, c)
    assert rc == 0 and d.uniform == 0 and not d.knots
    kn = np.empty(24, dtype=np.float32)
    assert N.lib().slt_grid_knots(bad.ctypes.data, c.ctypes.data, c.ctypes.data,
                                  ctypes.byref(d), kn.ctypes.data) == 0
    ref = np.concatenate([((bad.astype(np.float64) - bad[0]) * 7 / (float(bad[-1]) - bad[0])),
                          np.arange(8.0), np.arange(8.0)]).astype(np.float32)
    np.testing.assert_array_equal(kn, ref)
    assert kn[0] == 0 and kn[7] == 7 and kn[3] == pytest.approx(3 + 0.05 * 7, rel=1e-6)
    flat = c.copy()
    flat[4] = flat[3]                                # not strictly ascending
    assert _desc(flat, c, c)[0] == N.SLT_ERR_INVALID and "strictly" in N.last_error()
    assert _desc(c[:1], c, c)[0] == N.SLT_ERR_INVALID
    assert _desc(c[::-1].copy(), c, c)[0] == N.SLT_ERR_INVALID
    assert _desc(c, c, c, axis=3)[0] == N.SLT_ERR_INVALID


def test_schedule_lands_on_every_save():
    from sloth_amd.core.engine import schedule

    t = np.array([0.0, 1.0, 1.0, 2.5, 2.75])
    nsub, h = schedule(t, 0.3)
    assert list(nsub) == [4, 0, 5, 1]
    np.testing.assert_allclose(np.cumsum(nsub * h), t[1:] - t[0], rtol=1e-15)
    assert np.all(h <= 0.3)
    with pytest.raises(ValueError):
        schedule([0.0, -1.0], 0.1)


def test_beam_draw_order_matches_randomstate():
    from sloth_amd.core.beam import Beam

    n = 1000
    b = Beam(n, 2e-3, 1e-3, 5e-3, probing_direction="z", beam_type="circular", seeded=True)
    rng = np.random.RandomState(42)
    t = 2 * np.pi * rng.rand(n)
    u = rng.power(2, n)
    phi = np.pi * rng.rand(n)
    chi = 1e-3 * rng.randn(n)
    np.testing.assert_allclose(b.s0[0], 2e-3 * u * np.cos(t), rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(b.s0[1], 2e-3 * u * np.sin(t), rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(b.s0[5], C * np.cos(chi), rtol=1e-6)
    np.testing.assert_allclose(b.s0[3], C * np.sin(chi) * np.cos(phi), rtol=1e-5, atol=1e-3)
    assert np.all(b.s0[2] == np.float32(-5e-3))
    assert b.s0.dtype == np.float32 and b.s0.shape == (6, n)
    # unseeded beams differ between calls, seeded ones repeat
    b2 = Beam(n, 2e-3, 1e-3, 5e-3, beam_type="circular", seeded=True)
    np.testing.assert_array_equal(b.s0, b2.s0)


@pytest.mark.parametrize("pd", ["x", "y", "z"])
def test_beam_square_and_rectangular(pd):
    from sloth_amd.core.beam import Beam

    b = Beam(500, 4e-3, 0.0, 6e-3, probing_direction=pd, beam_type="square", seeded=True)
    ax = "xyz".index(pd)
    assert np.all(b.s0[ax] == np.float32(-6e-3))
    others = [a for a in range(3) if a != ax]
    assert np.abs(b.s0[others]).max() <= 4e-3
    assert np.all(b.s0[3 + ax] == np.float32(C))
    r = Beam(500, [1e-3, 2e-3], 0.0, 6e-3, probing_direction=pd, beam_type="rectangular",
             seeded=True)
    assert np.abs(r.s0[others[0]]).max() <= 1e-3 and np.abs(r.s0[others[1]]).max() <= 2e-3
    with pytest.raises(NotImplementedError):
        Beam(10, 1e-3, 0.0, 1e-3, beam_type="even")


def test_scalar_domain_protocol():
    from sloth_amd.core.domain import ScalarDomain

    d = ScalarDomain([4e-3, 4e-3, 8e-3], [8, 9, 10], inv_brems=True, Te=20.0, Z=1.0, Np=100)
    assert d.x.dtype == np.float32 and d.x[0] == np.float32(-2e-3) and d.z[-1] == np.float32(4e-3)
    assert d.ne.shape == (8, 9, 10) and d.region_count == 1 and d.Np_total == 100
    with pytest.raises(AssertionError, match="eV"):
        ScalarDomain([4e-3, 4e-3, 4e-3], [8, 8, 8], inv_brems=True, Te=-10.0, Z=1.0)
    with pytest.raises(AssertionError):
        ScalarDomain([4e-3] * 3, [8] * 3, inv_brems=True, Te=None, Z=1.0)


def test_scalar_domain_accepts_only_the_reference_keywords():
    """The reference constructor's keyword list (src/core/domain.py:57-59): a misspelt
    keyword (e.g. inverse_brems for inv_brems) is a TypeError, not a silent no-op; an
    unknown ne_type and a malformed lengths/dims raise as there (:232-258, :522-566)."""
    from sloth_amd.core.domain import ScalarDomain

    ScalarDomain(1e-3, 8, auto_batching=False, iteration=1, region_count=1, leeway_factor=1.2,
                 extra_info=False, memory_reporting=False, memory_limit=None, Np=10, s=1.0,
                 s1=0.1, s2=0.1, Ly=1e-3, ne_0=1e24, coulomb_log=10.0)
    for bad in ("inverse_brems", "Te_eV", "ne0", "S1"):
        with pytest.raises(TypeError):
            ScalarDomain(1e-3, 8, **{bad: 1.0})
    with pytest.raises(AssertionError, match="No valid profile"):
        ScalarDomain(1e-3, 8, ne_type="test_B")
    with pytest.raises(Exception, match="lengths must have len = 3"):
        ScalarDomain([1e-3, 1e-3], 8)
    with pytest.raises(Exception, match="n must have len = 3"):
        ScalarDomain(1e-3, [8, 8])
    d = ScalarDomain(2e-3, [6, 7, 8])
    assert d.ne.shape == (6, 7, 8) and not d.ne.any()          # no ne_type: empty domain
    for kind in ("test_slab", "test_linear_cos", "test_exponential_cos"):
        ne = ScalarDomain(2e-3, [6, 7, 8], ne_type=kind).ne
        assert ne.flags.writeable and ne.flags.c_contiguous and ne.dtype == np.float32


def test_scalar_domain_generated_profiles():
    """The generated profiles (src/core/domain.py:566-683), pinned by what their formulas
    fix: test_slab is linear in x through ne_0 at x = 0; test_linear_cos is linear in x at
    every y, periodic in y with period Ly, ne_0 (1 + s2) at x = 0 on a crest, and
    independent of z; test_exponential_cos grows by 10 per distance s in x, vanishes on
    troughs and is 2 ne_0 10^(x/s) on crests.  All in float32, as the reference computes
    them."""
    from sloth_amd.core.domain import ScalarDomain

    d = ScalarDomain([4e-3, 4e-3, 2e-3], [9, 5, 3], ne_type="test_slab", s=0.5, ne_0=1e24)
    assert d.ne.dtype == np.float32
    x = d.x.astype(np.float64)
    np.testing.assert_allclose(d.ne[:, 2, 1], 1e24 * (1 + 0.5 * x / 4e-3), rtol=1e-6)
    assert d.ne[4, 0, 0] == np.float32(1e24)                   # x = 0

    Ly = 1e-3
    d = ScalarDomain([4e-3, 4e-3, 2e-3], [9, 17, 4], ne_type="test_linear_cos", s1=0.3, s2=0.2,
                     ne_0=2e23, Ly=Ly)
    ne = d.ne.astype(np.float64)
    assert np.array_equal(d.ne[..., 0], d.ne[..., 3])          # no z dependence
    second = ne[2:] - 2 * ne[1:-1] + ne[:-2]                   # linear in x
    assert np.abs(second).max() < 1e-5 * ne.max()
    # y spacing is 0.25 mm, so 4 cells = one period of 1 mm
    np.testing.assert_allclose(ne[:, 4:], ne[:, :-4], rtol=1e-5)
    yi = int(np.argmin(np.abs(np.cos(2 * np.pi * d.y.astype(np.float64) / Ly) - 1.0)))
    assert ne[4, yi, 0] == pytest.approx(2e23 * 1.2, rel=1e-6)

    s = -2e-3
    d = ScalarDomain([10e-3, 2e-3, 1e-3], [11, 9, 2], ne_type="test_exponential_cos")
    ne = d.ne.astype(np.float64)
    x = d.x.astype(np.float64)
    crest = 4                                                  # y = 0: cos = 1
    assert d.y[crest] == 0
    np.testing.assert_allclose(ne[:, crest, 0], 2e24 * 10 ** (x / s), rtol=2e-6)
    np.testing.assert_allclose(ne[1:, crest, 0] / ne[:-1, crest, 0],
                               10 ** ((x[1] - x[0]) / s), rtol=2e-6)
    trough = 0                                                 # y = -1 mm = -Ly: cos = 1;
    assert ne[:, trough, 0] == pytest.approx(ne[:, crest, 0], rel=1e-5)
    half = 2                                                   # y = -0.5 mm: cos = -1
    assert np.abs(ne[:, half, 0]).max() < 1e-6 * ne.max()


def _vac(n=12, half=2e-3):
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


def test_trace_and_save_depths_validation_before_any_device_work():
    # the reference raises these ValueErrors before tracing (propagator.py:1083-1111)
    from sloth_amd.core.propagator import trace_and_save_depths

    s0 = np.zeros((6, 4), np.float32)
    s0[5] = C
    with pytest.raises(ValueError, match="positive"):
        trace_and_save_depths(s0, _vac(), 0.0, 1e-3, None, verbose=False)
    with pytest.raises(ValueError, match="out of range"):
        trace_and_save_depths(s0, _vac(), 1e-4, 1e-3, None, jones_components=[0, 5],
                              verbose=False)
    with pytest.raises(ValueError, match="smaller than the requested depth_max"):
        trace_and_save_depths(s0, _vac(half=2e-3), 1e-4, 10e-3, None, verbose=False)


def test_solve_rejects_beam_object():
    from sloth_amd.core.beam import Beam
    from sloth_amd.core.propagator import solve

    b = Beam(10, 1e-3, 0.0, 2e-3, seeded=True)
    with pytest.raises(AssertionError, match="direct output of the Beam"):
        solve(b, _vac(), 4e-3, verbose=False)


def test_product_fails_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from sloth_amd.core.propagator import precompute_gradients

    c = np.float32(np.linspace(0, 1, 4))
    with pytest.raises(RuntimeError, match="HIP device"):
        precompute_gradients(np.zeros((4, 4, 4), np.float32), c, c, c, 1e15)


def test_oracle_not_imported_by_product():
    # the product package must never reach the checker
    import subprocess
    import sys

    for root, _, files in os.walk(os.path.join(ROOT, "sloth_amd")):
        for f in files:
            if f.endswith(".py"):
                text = open(os.path.join(root, f)).read()
                assert "oracle" not in re.sub(r"#.*", "", text).replace('"""', "").split(
                    "import")[0] or "from oracle" not in text, f
                assert "from oracle" not in text and "import oracle" not in text, f
    code = ("import sys; import sloth_amd.core.propagator, sloth_amd.processing.diagnostics; "
            "print(any(m.startswith('oracle') for m in sys.modules))")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True)
    assert out.stdout.strip() == "False", out.stderr

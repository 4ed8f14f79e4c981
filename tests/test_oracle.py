"""Pins the FP64 oracle (CPU only).

Every oracle function is checked against something that does not share its
code: a library routine (numpy.gradient, scipy RegularGridInterpolator), a
closed form (harmonic trough, vacuum line, uniform absorbing slab), an
independent restatement in the reference's own test file (the NRL kappa
formula, tests/test_quadratic_trough.py:630-646), the convergence order of
RK4, or the second (NumPy) restatement of the oracle.
"""
import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

C = 299792458.0
ALPHA = 3.14207787e-4


def _f32lin(a, b, n):
    return np.float32(np.linspace(a, b, n)).astype(np.float64)


# ---------------------------------------------------------------------------
# precompute_gradients
# ---------------------------------------------------------------------------

def test_gradient_matches_numpy_gradient(orc):
    rng = np.random.default_rng(1)
    ne = rng.uniform(1e23, 1e24, (9, 7, 11)).astype(np.float32)
    x, y, z = _f32lin(-1e-3, 1e-3, 9), _f32lin(-2e-3, 1e-3, 7), _f32lin(0, 4e-3, 11)
    omega = orc.omega_of(532e-9)
    got = orc.gradients(ne, x, y, z, omega)
    f = ne.astype(np.float64) * orc.gradient_coeff(omega)
    for ax, (g, c) in enumerate(zip(got, (x, y, z))):
        ref = np.gradient(f, c, axis=ax, edge_order=1)
        np.testing.assert_allclose(g, ref, rtol=1e-13, atol=0)


def test_gradient_nonuniform_coordinates(orc):
    rng = np.random.default_rng(2)
    ne = rng.uniform(0, 1e24, (6, 5, 8))
    x = np.cumsum(rng.uniform(0.5, 1.5, 6)) * 1e-4
    y = np.cumsum(rng.uniform(0.5, 1.5, 5)) * 1e-4
    z = np.cumsum(rng.uniform(0.5, 1.5, 8)) * 1e-4
    got = orc.gradients(ne, x, y, z, 1e15)
    f = ne * orc.gradient_coeff(1e15)
    for ax, (g, c) in enumerate(zip(got, (x, y, z))):
        np.testing.assert_allclose(g, np.gradient(f, c, axis=ax), rtol=1e-12)


def test_gradient_exact_for_quadratic_interior(orc):
    # second-order central differences are exact for quadratics (any spacing)
    x = np.cumsum(np.r_[0.0, np.random.default_rng(3).uniform(0.5, 1.5, 9)]) * 1e-3
    y = z = np.linspace(0, 1e-3, 4)
    X = x[:, None, None] * np.ones((1, 4, 4))
    omega = 2e15
    coeff = orc.gradient_coeff(omega)
    ne = 1e24 * (X / 1e-3) ** 2
    gx, gy, gz = orc.gradients(ne, x, y, z, omega)
    exact = coeff * 1e24 * 2 * X / 1e-6
    np.testing.assert_allclose(gx[1:-1], exact[1:-1], rtol=1e-10)
    # constant along y and z: zero up to the rounding of the spacing differences
    assert np.abs(gy).max() <= 1e-12 * np.abs(gx).max()
    assert np.abs(gz).max() <= 1e-12 * np.abs(gx).max()


# ---------------------------------------------------------------------------
# RegularGridInterpolator (fill 0 outside)
# ---------------------------------------------------------------------------

def test_interp_matches_scipy_including_outside_and_faces(orc):
    rng = np.random.default_rng(4)
    cx, cy, cz = _f32lin(-1, 1, 7), _f32lin(-2, 2, 5), _f32lin(0, 1, 9)
    V = rng.standard_normal((7, 5, 9))
    pts = np.stack([rng.uniform(-1.3, 1.3, 4000), rng.uniform(-2.3, 2.3, 4000),
                    rng.uniform(-0.2, 1.2, 4000)], 1)
    # points exactly on grid nodes and on the box faces count as inside
    pts[:50, 0] = cx[-1]
    pts[50:100, 1] = cy[0]
    pts[100:150] = np.stack([cx[rng.integers(0, 7, 50)], cy[rng.integers(0, 5, 50)],
                             cz[rng.integers(0, 9, 50)]], 1)
    ref = RegularGridInterpolator((cx, cy, cz), V, bounds_error=False, fill_value=0.0)(pts)
    np.testing.assert_allclose(orc.interp_points((cx, cy, cz), V, pts), ref, rtol=1e-13,
                               atol=1e-15)
    got_np, _ = orc.np_interp((cx, cy, cz), [V], pts)
    np.testing.assert_allclose(got_np[0], ref, rtol=1e-13, atol=1e-15)


def test_interp_exact_for_trilinear_field(orc):
    cx, cy, cz = _f32lin(-1, 1, 6), _f32lin(0, 3, 4), _f32lin(-2, 0, 5)
    X, Y, Z = np.meshgrid(cx, cy, cz, indexing="ij")
    f = lambda x, y, z: 1 + 2 * x - 3 * y + 0.5 * z + x * y - 2 * y * z + 0.25 * x * y * z
    rng = np.random.default_rng(5)
    pts = np.stack([rng.uniform(-1, 1, 500), rng.uniform(0, 3, 500), rng.uniform(-2, 0, 500)], 1)
    got = orc.interp_points((cx, cy, cz), f(X, Y, Z), pts)
    np.testing.assert_allclose(got, f(*pts.T), rtol=1e-12, atol=1e-12)


def test_interp_brute_force_tiny_grid(orc):
    # 2^3 grid: the interpolant is sum over corners of value * prod(weights)
    cx, cy, cz = np.array([0.0, 1.0]), np.array([0.0, 2.0]), np.array([-1.0, 1.0])
    V = np.arange(8, dtype=np.float64).reshape(2, 2, 2) ** 1.5
    p = np.array([[0.25, 0.5, 0.0]])
    tx, ty, tz = 0.25, 0.25, 0.5
    want = 0.0
    for i in (0, 1):
        for j in (0, 1):
            for k in (0, 1):
                want += V[i, j, k] * (tx if i else 1 - tx) * (ty if j else 1 - ty) * (tz if k else 1 - tz)
    assert orc.interp_points((cx, cy, cz), V, p)[0] == pytest.approx(want, rel=1e-15)


# ---------------------------------------------------------------------------
# kappa_inv_brems
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("Te, Z", [(100.0, 1.0), (10.0, 4.0), (20.0, 1.0), (5.0, 1.0)])
def test_kappa_matches_reference_test_formula(orc, Te, Z):
    # independent restatement copied in spirit from tests/test_quadratic_trough.py:633-642
    from scipy.constants import e as e_charge, epsilon_0 as eps0

    ne = 1e25
    omega = 2 * np.pi * C / 1064e-9
    ne_cc = ne * 1e-6
    v_the = 4.19e5 * np.sqrt(Te)
    o_pe = 5.64e4 * np.sqrt(ne_cc)
    b_min = Z * e_charge / (4.0 * np.pi * eps0 * Te)
    CL = max(2.0, np.log(v_the / (max(o_pe, omega) * b_min)))
    want = 3.1e-5 * Z * C * (ne_cc / omega) ** 2 * CL * Te ** (-1.5)
    got, cl = orc.kappa(np.array([ne]), Te, Z, omega, return_cl=True)
    assert got[0] == pytest.approx(want, rel=1e-12)
    assert cl[0] == pytest.approx(CL, rel=1e-12)


def test_kappa_coulomb_log_clamp_and_override(orc):
    omega = 2 * np.pi * C / 1064e-9
    # low Te clamps CL at 2 (reference tests/test_quadratic_trough.py:1165-1170)
    _, cl = orc.kappa(np.array([1e25]), 5.0, 1.0, omega, return_cl=True)
    assert cl[0] == 2.0
    k2 = orc.kappa(np.array([1e25]), 5.0, 1.0, omega)
    k10 = orc.kappa(np.array([1e25]), 5.0, 1.0, omega, coulomb_log=10.0)
    assert k10[0] == pytest.approx(5 * k2[0], rel=1e-14)
    # turbulent-box operating point: 2e24 m^-3, 20 eV, Z=1, 532 nm -> clamped to 2
    _, cl = orc.kappa(np.array([2e24]), 20.0, 1.0, 2 * np.pi * C / 532e-9, return_cl=True)
    assert cl[0] == 2.0


def test_kappa_grid_te_z(orc):
    rng = np.random.default_rng(6)
    ne = rng.uniform(1e23, 1e25, (3, 4, 5))
    te = rng.uniform(5, 200, (3, 4, 5))
    z = rng.uniform(1, 5, (3, 4, 5))
    got = orc.kappa(ne, te, z, 3e15)
    for idx in [(0, 0, 0), (2, 3, 4), (1, 2, 3)]:
        assert got[idx] == pytest.approx(orc.kappa(np.array([ne[idx]]), te[idx], z[idx], 3e15)[0],
                                         rel=1e-15)


# ---------------------------------------------------------------------------
# RK4 trace
# ---------------------------------------------------------------------------

def _trough(orc, n=33, half=5e-3, half_z=11e-3, lwl=1064e-9, ne0=3.8e25):
    x = _f32lin(-half, half, n)
    z = _f32lin(-half_z, half_z, n)
    X = x[:, None, None] * np.ones((1, n, n))
    Y = x[None, :, None] * np.ones((n, 1, n))
    ne = ne0 * ((X / half) ** 2 + (Y / half) ** 2)
    omega = orc.omega_of(lwl)
    grads = orc.gradients(ne, x, x, z, omega)
    w_osc = C * np.sqrt(ne0 / (ALPHA * omega ** 2)) / half
    return grads, (x, x, z), w_osc


def _rays(x0, y0, z0, vx=0.0, vy=0.0):
    n = len(x0)
    s0 = np.zeros((6, n))
    s0[0], s0[1], s0[2] = x0, y0, z0
    s0[3], s0[4] = vx, vy
    s0[5] = np.sqrt(C ** 2 - np.asarray(vx) ** 2 - np.asarray(vy) ** 2)
    return s0


def test_rk4_harmonic_trough_closed_form_and_order(orc):
    grads, coords, w = _trough(orc)
    x0 = np.array([0.5e-3, 1e-3, -1.5e-3, 2e-3])
    y0 = np.array([0.0, 0.7e-3, 1e-3, -0.3e-3])
    # start exactly on the entry face (fp32 coordinate, as the reference's fp32 rays do);
    # a start a hair outside would lose 1/6 of the first step's force
    s0 = _rays(x0, y0, coords[2][0])
    T = 20e-3 / C
    errs = []
    for nsub in (40, 80, 160):
        saves, steps = orc.trace(grads, None, coords, s0, [nsub], [T / nsub], naive=True)
        fin = saves[-1]
        np.testing.assert_allclose(fin[0], x0 * np.cos(w * T), rtol=0, atol=5e-9)
        np.testing.assert_allclose(fin[1], y0 * np.cos(w * T), rtol=0, atol=5e-9)
        np.testing.assert_allclose(fin[3], -x0 * w * np.sin(w * T), rtol=1e-6, atol=1e-3)
        assert np.all(fin[5] == C)  # no z force: vz stays c exactly
        assert np.all(steps == nsub)
        errs.append(np.abs(fin[0] - x0 * np.cos(w * T)).max())
    # fourth order: halving h divides the error by ~16
    assert 10 < errs[0] / errs[1] < 24 and 10 < errs[1] / errs[2] < 24


def test_rk4_vacuum_straight_lines_and_ballistic_skip(orc):
    n = 12
    c = _f32lin(-2e-3, 2e-3, n)
    zero = np.zeros((n, n, n))
    rng = np.random.default_rng(7)
    s0 = _rays(rng.uniform(-3e-3, 3e-3, 50), rng.uniform(-3e-3, 3e-3, 50), -5e-3,
               vx=rng.uniform(-3e6, 3e6, 50), vy=rng.uniform(-3e6, 3e6, 50))
    T = 12e-3 / C
    saves, steps = orc.trace((zero, zero, zero), zero, (c, c, c), s0, [100], [T / 100])
    fin = saves[-1]
    np.testing.assert_allclose(fin[:3], s0[:3] + s0[3:6] * T, rtol=1e-13, atol=1e-17)
    np.testing.assert_array_equal(fin[3:6], s0[3:6])
    np.testing.assert_allclose(fin[7], C * T, rtol=1e-13)  # path length = c T
    assert np.all(fin[6] == 0)


def test_rk4_uniform_slab_absorption_closed_form(orc):
    n = 16
    c = _f32lin(-3e-3, 3e-3, n)
    ne = np.full((n, n, n), 1e25)
    omega = orc.omega_of(1064e-9)
    kap = orc.kappa(ne, 100.0, 1.0, omega)
    grads = orc.gradients(ne, c, c, c, omega)
    s0 = _rays(np.linspace(-5e-4, 5e-4, 10), np.zeros(10), -3e-3)
    depth = 2e-3
    t = np.linspace(0, depth, 5) / C
    nsub = np.full(4, 7)
    saves, _ = orc.trace(grads, kap, (c, c, c), s0, nsub, np.diff(t) / 7)
    for k in range(5):
        np.testing.assert_allclose(saves[k, 6], kap[0, 0, 0] * t[k], rtol=1e-12)
    # geometry unaffected by absorption (tests/test_quadratic_trough.py:449-467)
    saves0, _ = orc.trace(grads, None, (c, c, c), s0, nsub, np.diff(t) / 7)
    np.testing.assert_array_equal(saves0[:, :6], saves[:, :6])


def test_rk4_exits_box_then_straight(orc):
    # after leaving the box the ray is ballistic: position = exit state + v t
    grads, coords, w = _trough(orc, n=17, half_z=4e-3)
    s0 = _rays(np.array([1e-3]), np.array([0.5e-3]), coords[2][0])
    h = 0.5 * (coords[2][1] - coords[2][0]) / C
    nsub = 400
    saves, steps = orc.trace(grads, None, coords, s0, [nsub], [h], naive=True)
    fast, steps_f = orc.trace(grads, None, coords, s0, [nsub], [h], naive=False)
    np.testing.assert_allclose(fast[-1], saves[-1], rtol=1e-13, atol=1e-18)
    assert steps[0] == steps_f[0]
    assert 30 <= steps[0] <= 34  # 8 mm at 0.25 mm per step


def test_c_oracle_matches_numpy_oracle_random_field(orc):
    rng = np.random.default_rng(8)
    n = (9, 8, 10)
    cx, cy, cz = _f32lin(-1e-3, 1e-3, n[0]), _f32lin(-1e-3, 1e-3, n[1]), _f32lin(-1e-3, 1e-3, n[2])
    ne = rng.uniform(1e24, 5e25, n)
    omega = orc.omega_of(532e-9)
    grads = orc.gradients(ne, cx, cy, cz, omega)
    kap = orc.kappa(ne, 20.0, 1.0, omega)
    m = 64
    s0 = _rays(rng.uniform(-1.2e-3, 1.2e-3, m), rng.uniform(-1.2e-3, 1.2e-3, m),
               rng.uniform(-1.5e-3, -0.9e-3, m), vx=rng.uniform(-2e7, 2e7, m),
               vy=rng.uniform(-2e7, 2e7, m))
    nsub = np.array([5, 9, 13], dtype=np.int32)
    h = np.array([1e-13, 7e-14, 5e-14])
    a, sa = orc.trace(grads, kap, (cx, cy, cz), s0, nsub, h, naive=True)
    b, sb = orc.np_trace(grads, kap, (cx, cy, cz), s0, nsub, h)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-20)
    np.testing.assert_array_equal(sa, sb)
    c, sc = orc.trace(grads, kap, (cx, cy, cz), s0, nsub, h, naive=False)
    np.testing.assert_allclose(c, a, rtol=1e-12, atol=1e-20)
    np.testing.assert_array_equal(sc, sa)


def test_jones_closed_form(orc):
    st = np.array([[1.0], [2.0], [3.0], [0.1], [-0.2], [1.0]])
    np.testing.assert_allclose(orc.jones(st, "z")[:, 0], [1.0, np.arctan(0.1), 2.0, np.arctan(-0.2)])
    np.testing.assert_allclose(orc.jones(st, "x")[:, 0], [2.0, np.arctan(-2.0), 3.0, np.arctan(10.0)])
    np.testing.assert_allclose(orc.jones(st, "y")[:, 0], [1.0, np.arctan(-0.5), 3.0, np.arctan(-5.0)])


def test_face_gap_closed_forms(orc):
    """The oracle's face gap (smallest distance of a stage point, the start point
    excepted, to the surface of the closed box) on vacuum rays along z, where every
    stage point is known: z = lo + j h c / 2 while the step touches the box."""
    n = 9
    c = _f32lin(-2e-3, 2e-3, n)
    lo, hi = float(c[0]), float(c[-1])
    zero = np.zeros((n, n, n))
    hc = (hi - lo) / 7.3            # no stage point on a face
    nsub = 12
    # rays: centred; 0.4 mm inside the x = hi face; on the x = hi face; 0.1 mm outside it
    xs = np.array([0.0, hi - 0.4e-3, hi, hi + 0.1e-3])
    s0 = _rays(xs, np.zeros(4), lo)
    _, steps, gap = orc.trace((zero, zero, zero), zero, (c, c, c), s0, [nsub], [hc / C],
                              face_gap=True)
    # z distances of the stage points of the steps that start inside (start point excepted)
    zpts = [lo + j * hc / 2 for j in range(1, 2 * 8 + 1)]   # steps 0..7 start at z <= hi
    zgap = min(min(z - lo, hi - z) if z <= hi else z - hi for z in zpts)
    assert zgap > 1e-5
    np.testing.assert_allclose(gap[0], zgap, rtol=1e-9)
    np.testing.assert_allclose(gap[1], min(zgap, 0.4e-3), rtol=1e-9)
    assert gap[2] == 0.0                 # on the face the whole way: inside, distance 0
    assert np.isinf(gap[3]) and steps[3] == 0   # misses the box: no RK4 step at all
    # the start point itself is excepted (it sits on the z = lo face for every ray)
    assert gap[0] > 0

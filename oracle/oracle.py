"""FP64 CPU oracle for SLOTH's ray-tracing hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product package ``sloth_amd`` never does, and nothing here is
ever the thing measured on the GPU.

Two restatements of the same algorithm live here:

* the C library ``sloth_oracle.c`` (built by ``build()`` into
  ``oracle/_build/libsloth_oracle.so``), OpenMP over rays -- the fast oracle and
  the timed CPU baseline;
* NumPy restatements (``np_interp``, ``np_trace``) vectorised over rays -- an
  independent second implementation used to cross-check the C one.

Reference functions restated (MAGPIE-ICL/SLOTH, file:line):

* ``gradients``      src/core/propagator.py:39-93  (precompute_gradients)
* ``kappa``          src/core/propagator.py:116-164 (kappa_inv_brems)
* ``interp_points``  src/core/interpolator.py:15-106 (RegularGridInterpolator)
* ``trace``          src/core/propagator.py:168-243 (dsdt) under fixed-step RK4
* ``jones``          src/shared/propagation.py:4-90 (ray_to_Jonesvector,
  keep_current_plane=True)

Pinning (see tests/test_oracle.py): gradients against ``numpy.gradient`` and the
exact derivative of quadratics; interpolation against
``scipy.interpolate.RegularGridInterpolator`` and exactness on trilinear
fields; kappa against the reference test's independent restatement of the NRL
formula (tests/test_quadratic_trough.py:630-646) and the Coulomb-log clamp;
RK4 against the closed-form harmonic trough, vacuum straight lines, the
uniform-slab ``exp(-kappa d / c)`` law and the fourth-order convergence rate;
the C and NumPy restatements against each other.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SRC = os.path.join(_HERE, "sloth_oracle.c")
_LIB = os.path.join(_HERE, "_build", "libsloth_oracle.so")

C_LIGHT = 299792458.0          # scipy.constants.c
E_CHARGE = 1.602176634e-19     # scipy.constants.e
EPSILON_0 = 8.8541878188e-12   # scipy.constants.epsilon_0 (CODATA 2022)
ALPHA_CRIT = 3.14207787e-4     # m_e eps0 / e^2, src/core/propagator.py:48-50,73,88

_lib = None


def build(verbose: bool = False) -> str:
    """Compile the C oracle with gcc (-O3, OpenMP) into oracle/_build."""
    os.makedirs(os.path.dirname(_LIB), exist_ok=True)
    cmd = ["gcc", "-O3", "-std=c11", "-fopenmp", "-fPIC", "-shared", "-fno-fast-math",
           "-ffp-contract=off", _SRC, "-o", _LIB + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(_LIB + ".tmp", _LIB)
    return _LIB


# The timed CPU baseline: the same C source built for the host it runs on.  It is
# compiled where it is used (the GPU box's host at bench time), into a temporary
# directory, because -march=native code may not run on another CPU.
NATIVE_FLAGS = ["-O3", "-march=native", "-std=gnu11", "-fopenmp", "-fPIC", "-shared"]
_native = None


def load_native() -> str:
    """Build (once per process) and load the -O3 -march=native oracle; returns its flags.

    GNU C mode lets gcc contract multiply-adds into FMAs, so this compile is faster and
    not bit-identical to the one the tests check against; it is timed, never used as the
    checker."""
    global _native
    if _native is None:
        import tempfile

        d = tempfile.mkdtemp(prefix="sloth_oracle_native_")
        out = os.path.join(d, "libsloth_oracle_native.so")
        subprocess.run(["gcc", *NATIVE_FLAGS, _SRC, "-o", out, "-lm"], check=True)
        _native = _bind(ctypes.CDLL(out))
    return " ".join(NATIVE_FLAGS)


def load():
    """Load (building first if absent) the C oracle."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        build()
    _lib = _bind(ctypes.CDLL(_LIB))
    return _lib


def _bind(lib):
    P = ctypes.POINTER
    d = ctypes.c_double
    i32, i64 = ctypes.c_int, ctypes.c_int64
    lib.orc_gradient_axis.argtypes = [P(d), P(d), i32, i32, i32, i32, P(d)]
    lib.orc_kappa.argtypes = [P(d), i64, P(d), i64, P(d), i64, d, d, d, d, P(d), P(d)]
    lib.orc_interp_points.argtypes = [i32, i32, i32, P(d), P(d), P(d), P(d), i64, P(d), P(d)]
    lib.orc_trace.argtypes = [i32, i32, i32, P(d), P(d), P(d), P(d), P(d), P(d), P(d),
                              i64, P(d), i32, P(ctypes.c_int32), P(d), i32, P(d),
                              P(ctypes.c_int64), P(d), i32]
    lib.orc_max_threads.restype = ctypes.c_int
    return lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def omega_of(lwl: float) -> float:
    """Laser angular frequency, src/core/propagator.py:467,1058."""
    return 2 * np.pi * (C_LIGHT / lwl)


def gradient_coeff(omega: float) -> float:
    """-0.5 c^2 / (alpha omega^2), evaluated in float64 (propagator.py:88)."""
    return np.float64(-0.5) * float(C_LIGHT) ** 2 / (ALPHA_CRIT * float(omega) ** 2)


def gradients(ne, x, y, z, omega):
    """FP64 (dndx, dndy, dndz) of coeff*ne with numpy.gradient semantics."""
    lib = load()
    ne = _f64(ne)
    nx, ny, nz = ne.shape
    f = ne * gradient_coeff(omega)
    outs = []
    for axis, coord in enumerate((x, y, z)):
        coord = _f64(coord)
        out = np.empty_like(f)
        lib.orc_gradient_axis(_dp(f), _dp(coord), nx, ny, nz, axis, _dp(out))
        outs.append(out)
    return tuple(outs)


def kappa(ne, Te, Z, omega, coulomb_log=None, return_cl=False):
    """FP64 inverse-bremsstrahlung rate [1/s] (propagator.py:116-164)."""
    lib = load()
    ne = _f64(ne)
    shape = ne.shape
    ne1 = ne.reshape(-1)
    te = _f64(Te).reshape(-1)
    zz = _f64(Z).reshape(-1)
    if te.size not in (1, ne1.size) or zz.size not in (1, ne1.size):
        raise ValueError("Te and Z must be scalars or match ne")
    out = np.empty_like(ne1)
    cl = np.empty_like(ne1)
    lib.orc_kappa(_dp(ne1), ne1.size, _dp(te), te.size, _dp(zz), zz.size, float(omega),
                  E_CHARGE, EPSILON_0, float(coulomb_log) if coulomb_log else 0.0,
                  _dp(out), _dp(cl))
    if return_cl:
        return out.reshape(shape), cl.reshape(shape)
    return out.reshape(shape)


def interp_points(coords, values, pts):
    """RegularGridInterpolator(fill_value=0) of one grid at (N,3) points."""
    lib = load()
    cx, cy, cz = (_f64(c) for c in coords)
    values = _f64(values)
    pts = _f64(pts).reshape(-1, 3)
    out = np.empty(pts.shape[0])
    lib.orc_interp_points(cx.size, cy.size, cz.size, _dp(cx), _dp(cy), _dp(cz), _dp(values),
                          pts.shape[0], _dp(pts), _dp(out))
    return out


def trace(grads, kap, coords, s0, nsub, h, naive=False, nthreads=0, face_gap=False,
          native=False):
    """Fixed-step RK4 trace.  Returns (saves[(S+1),8,N], steps[N]), and with
    ``face_gap=True`` also gap[N]: per ray, the smallest distance [m] from any stage
    point of its RK4 steps (the start point excepted) to the surface of the closed
    grid box.

    Rows of each save: x, y, z, vx, vy, vz, tau, path (length of the polyline through
    the step end points).  ``native=True`` runs the -march=native build (``load_native``,
    the timed CPU baseline) instead of the checker build.
    """
    if native:
        load_native()
    lib = _native if native else load()
    dndx, dndy, dndz = (_f64(g) for g in grads)
    kp = _f64(kap) if kap is not None else None
    cx, cy, cz = (_f64(c) for c in coords)
    s0 = _f64(s0)
    n = s0.shape[1]
    nsub = np.ascontiguousarray(np.asarray(nsub, dtype=np.int32))
    h = _f64(h)
    nseg = nsub.size
    saves = np.empty((nseg + 1, 8, n))
    steps = np.empty(n, dtype=np.int64)
    gap = np.empty(n) if face_gap else None
    lib.orc_trace(cx.size, cy.size, cz.size, _dp(cx), _dp(cy), _dp(cz),
                  _dp(dndx), _dp(dndy), _dp(dndz), _dp(kp) if kp is not None else None,
                  n, _dp(s0), nseg, nsub.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                  _dp(h), 1 if naive else 0, _dp(saves),
                  steps.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                  _dp(gap) if gap is not None else None, int(nthreads))
    if face_gap:
        return saves, steps, gap
    return saves, steps


def max_threads() -> int:
    return int(load().orc_max_threads())


def jones(state6, probing_direction="z"):
    """ray_to_Jonesvector(keep_current_plane=True), propagation.py:4-90."""
    x, y, z, vx, vy, vz = (np.asarray(state6[i], dtype=np.float64) for i in range(6))
    if probing_direction == "x":
        return np.stack([y, np.arctan(vy / vx), z, np.arctan(vz / vx)])
    if probing_direction == "y":
        return np.stack([x, np.arctan(vx / vy), z, np.arctan(vz / vy)])
    if probing_direction == "z":
        return np.stack([x, np.arctan(vx / vz), y, np.arctan(vy / vz)])
    raise ValueError(probing_direction)


# ---------------------------------------------------------------------------
# Independent NumPy restatement (vectorised over rays)
# ---------------------------------------------------------------------------

def np_interp(coords, channels, pts):
    """Vectorised RegularGridInterpolator over several channels sharing a grid.

    Returns (values[C, N], inside[N]).  interpolator.py:79-104.
    """
    pts = np.asarray(pts, dtype=np.float64)
    idx, t = [], []
    oob = np.zeros(pts.shape[0], dtype=bool)
    for a, g in enumerate(coords):
        g = np.asarray(g, dtype=np.float64)
        x = pts[:, a]
        i = np.searchsorted(g, x) - 1
        i = np.where(i < 0, 0, i)
        i = np.where(i > g.size - 2, g.size - 2, i)
        idx.append(i)
        t.append((x - g[i]) / (g[i + 1] - g[i]))
        oob |= x < g[0]
        oob |= x > g[-1]
    out = np.zeros((len(channels), pts.shape[0]))
    for ex in (0, 1):
        for ey in (0, 1):
            for ez in (0, 1):
                w = np.ones(pts.shape[0])
                w = w * (t[0] if ex else 1 - t[0])
                w = w * (t[1] if ey else 1 - t[1])
                w = w * (t[2] if ez else 1 - t[2])
                for c, ch in enumerate(channels):
                    if ch is None:
                        continue
                    out[c] += ch[idx[0] + ex, idx[1] + ey, idx[2] + ez] * w
    out[:, oob] = 0.0
    return out, ~oob


def np_trace(grads, kap, coords, s0, nsub, h):
    """Textbook RK4 (every step a full RK4 step), vectorised over rays.

    Returns (saves[(S+1),8,N], steps[N]) exactly like ``trace(..., naive=True)``.
    """
    chans = [np.asarray(g, dtype=np.float64) for g in grads]
    chans.append(np.asarray(kap, dtype=np.float64) if kap is not None else None)
    s = np.asarray(s0, dtype=np.float64)
    r = s[0:3].T.copy()
    v = s[3:6].T.copy()
    n = r.shape[0]
    tau = np.zeros(n)
    path = np.zeros(n)
    steps = np.zeros(n, dtype=np.int64)
    saves = [np.vstack([r.T, v.T, tau, path])]

    def rhs(p):
        vals, inside = np_interp(coords, chans, p)
        return vals[0:3].T, vals[3], inside

    for nk, hh in zip(np.asarray(nsub), np.asarray(h, dtype=np.float64)):
        for _ in range(int(nk)):
            a1, k1, i1 = rhs(r)
            r2 = r + 0.5 * hh * v
            v2 = v + 0.5 * hh * a1
            a2, k2, i2 = rhs(r2)
            r3 = r + 0.5 * hh * v2
            v3 = v + 0.5 * hh * a2
            a3, k3, i3 = rhs(r3)
            r4 = r + hh * v3
            v4 = v + hh * a3
            a4, k4, i4 = rhs(r4)
            dr = hh / 6.0 * (v + 2.0 * v2 + 2.0 * v3 + v4)
            r = r + dr
            v = v + hh / 6.0 * (a1 + 2.0 * a2 + 2.0 * a3 + a4)
            tau = tau + hh / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
            path = path + np.linalg.norm(dr, axis=1)
            steps += (i1 | i2 | i3 | i4)
        saves.append(np.vstack([r.T, v.T, tau, path]))
    return np.stack(saves), steps

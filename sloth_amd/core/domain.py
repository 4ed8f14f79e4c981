"""Scalar electron-density domain (the attribute protocol the tracer reads).

Mirrors ``core.domain.ScalarDomain`` (src/core/domain.py:27-800): the keyword
arguments of its constructor (:57-59), the attributes ``solve`` /
``trace_and_save_depths`` read (``ne, x, y, z, lengths, dims,
probing_direction, inv_brems, Te, Z, Np_total, region_count,
ray_batch_count``) and the generated test profiles ``test_null``,
``test_slab``, ``test_linear_cos`` and ``test_exponential_cos`` (:566-683),
evaluated in float32 in the reference's operation order.  Coordinates are
float32 ``linspace(-L/2, L/2, n)`` (domain.py:380-382); Te and Z are cast to
float32 and Te must be > 0 (domain.py:192-207).  The memory-probing
auto-batching arguments (``auto_batching``, ``iteration``, ``region_count``,
``leeway_factor``, ``coord_backup``, ``future_dims``, ``memory_reporting``,
``memory_limit``, ``extra_info``) are accepted and have no effect: on a 288 GB
device there is one region and one ray batch, always.  ``coulomb_log`` is an
addition (a fixed Coulomb logarithm instead of the computed one).  Duck-typed
objects with the same attributes are accepted everywhere a ScalarDomain is.
"""
from __future__ import annotations

import numpy as np

# ne_type values and the profile each generates (domain.py:517-565); None with no
# ne gives an empty domain, as the reference's constructor does (:453-466)
PROFILES = ("test_null", "test_slab", "test_linear_cos", "test_exponential_cos", "import")


def _three(v, what, label):
    """A scalar for all three axes or a length-3 sequence (domain.py:232-258)."""
    a = np.asarray(v)
    if a.ndim == 0:
        return np.array([a.item()] * 3)
    if a.shape != (3,):
        raise Exception(f"{what} must have len = 3: {label}")
    return a


class ScalarDomain:
    def __init__(self, lengths, dims, *, ne_type=None, probing_direction="z", auto_batching=True,
                 iteration=1, region_count=1, leeway_factor=None, coord_backup=None,
                 future_dims=None, extra_info=False, memory_reporting=False, memory_limit=None,
                 Np=None, s=None, s1=None, s2=None, Ly=None, ne_0=None, ne=None,
                 inv_brems=False, Te=None, Z=None, coulomb_log=None):
        lengths = _three(lengths, "lengths", "(x,y,z)").astype(np.float64)
        dims = _three(dims, "n", "(x_n, y_n, z_n)").astype(np.int64)
        if probing_direction not in ("x", "y", "z"):
            raise AssertionError("Invalid entry for probing_direction!")
        if ne_type is not None and ne_type not in PROFILES:
            raise AssertionError(
                "\nNo valid profile detected! Ensure passed name is correct or call yourself.")
        self.lengths = lengths.astype(np.float32)
        self.dims = dims.astype(np.int32)
        self.x_length, self.y_length, self.z_length = (float(v) for v in lengths)
        self.x_n, self.y_n, self.z_n = (int(v) for v in dims)
        self.probing_direction = probing_direction
        self.ne_type = ne_type
        self.s, self.s1, self.s2, self.Ly, self.ne_0 = s, s1, s2, Ly, ne_0
        self.Np_total = np.int64(Np) if Np is not None else None
        self.region_count = 1
        self.ray_batch_count = 1
        self.coord_backup = None
        self.future_dims = None
        self.extra_info = extra_info
        self.coulomb_log = coulomb_log

        assert not inv_brems or (Te is not None and Z is not None), \
            "\nTe and Z must both be provided when inv_brems=True."
        self.inv_brems = inv_brems
        if Te is not None:
            te = np.asarray(Te, dtype=np.float32)
            assert float(np.min(te)) > 0.0, (
                "\nTe must be a strictly positive electron temperature in eV "
                "(received a zero or negative value).  Do not pass Te in Kelvin or Joules.")
            self.Te = te
        else:
            self.Te = None
        self.Z = np.asarray(Z, dtype=np.float32) if Z is not None else None

        self.x = np.float32(np.linspace(-self.x_length / 2, self.x_length / 2, self.x_n))
        self.y = np.float32(np.linspace(-self.y_length / 2, self.y_length / 2, self.y_n))
        self.z = np.float32(np.linspace(-self.z_length / 2, self.z_length / 2, self.z_n))

        if ne is not None:
            self.ne = ne
        elif ne_type in (None, "test_null"):
            self.ne = np.zeros((self.x_n, self.y_n, self.z_n), dtype=np.float32)
        elif ne_type == "import":
            raise AssertionError("ne_type='import' needs ne=...")
        else:
            self.ne = getattr(self, ne_type)()

    # -- generated profiles, float32 over the (x, y, z) grid ----------------------
    def _grid(self, axis):
        """The coordinate of `axis` broadcast over the grid (float32, no copy)."""
        c = (self.x, self.y, self.z)[axis]
        shape = [1, 1, 1]
        shape[axis] = -1
        return np.broadcast_to(c.reshape(shape), (self.x_n, self.y_n, self.z_n))

    def _full(self, a):
        # a fresh, writable C-order float32 grid (a broadcast view is read-only)
        return np.array(np.broadcast_to(a, (self.x_n, self.y_n, self.z_n)), dtype=np.float32,
                        order="C", copy=True)

    def test_slab(self):
        """ne_0 (1 + s x / L_x): a linear gradient along x (domain.py:580-603; s = 1,
        ne_0 = 2e23 m^-3 unless given)."""
        s = 1 if self.s is None else self.s
        n0 = 2e23 if self.ne_0 is None else self.ne_0
        X = self._grid(0)
        return self._full(n0 * (1.0 + s * X / np.float32(self.x_length)))

    def test_linear_cos(self):
        """ne_0 (1 + s1 x / L_x) (1 + s2 cos(2 pi y / Ly)): a sinusoidal perturbation in y
        growing linearly in x (domain.py:605-637; s1 = s2 = 0.1, ne_0 = 2e23 m^-3, Ly = 1
        unless given)."""
        s1 = 0.1 if self.s1 is None else self.s1
        s2 = 0.1 if self.s2 is None else self.s2
        n0 = 2e23 if self.ne_0 is None else self.ne_0
        ly = 1 if self.Ly is None else self.Ly
        X, Y = self._grid(0), self._grid(1)
        grow = 1.0 + s1 * X / np.float32(self.x_length)
        wave = 1 + s2 * np.cos(2 * np.float32(np.pi) * Y / np.float32(ly))
        return self._full(n0 * grow * wave)

    def test_exponential_cos(self):
        """ne_0 10^(x / s) (1 + cos(2 pi y / Ly)): a sinusoidal perturbation in y with an
        exponential envelope in x, evaluated step by step as the reference does
        (domain.py:640-683; ne_0 = 1e24 m^-3, Ly = 1e-3 m, s = -2e-3 m unless given)."""
        n0 = 1e24 if self.ne_0 is None else self.ne_0
        ly = 1e-3 if self.Ly is None else self.Ly
        s = -2e-3 if self.s is None else self.s
        envelope = np.float32(10) ** (self._grid(0) / np.float32(s))
        wave = np.float32(2) * (np.float32(np.pi) * (self._grid(1) / np.float32(ly)))
        wave = 1 + np.cos(wave)
        return self._full(np.float32(n0) * (envelope * wave))

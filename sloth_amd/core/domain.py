"""Scalar electron-density domain (the attribute protocol the tracer reads).

Mirrors the parts of ``core.domain.ScalarDomain`` (src/core/domain.py:27-800)
that ``solve`` / ``trace_and_save_depths`` read: ``ne, x, y, z, lengths, dims,
probing_direction, inv_brems, Te, Z, Np_total, region_count,
ray_batch_count``.  Coordinates are float32 ``linspace(-L/2, L/2, n)``
(domain.py:380-382); Te and Z are cast to float32 and Te must be > 0
(domain.py:192-207).  Memory-probing auto-batching is not needed on a
288 GB device: one region and one ray batch, always.  Duck-typed objects with
the same attributes are accepted everywhere a ScalarDomain is.
"""
from __future__ import annotations

import numpy as np


class ScalarDomain:
    def __init__(self, lengths, dims, *, ne_type=None, probing_direction="z", Np=None, ne=None,
                 inv_brems=False, Te=None, Z=None, coulomb_log=None, ne_0=None, s=None,
                 **_ignored):
        lengths = np.broadcast_to(np.asarray(lengths, dtype=np.float64).reshape(-1), (3,))
        dims = np.broadcast_to(np.asarray(dims, dtype=np.int64).reshape(-1), (3,))
        if probing_direction not in ("x", "y", "z"):
            raise AssertionError("Invalid entry for probing_direction!")
        self.lengths = lengths.astype(np.float32)
        self.dims = dims.astype(np.int32)
        self.x_length, self.y_length, self.z_length = (float(v) for v in lengths)
        self.x_n, self.y_n, self.z_n = (int(v) for v in dims)
        self.probing_direction = probing_direction
        self.ne_type = ne_type
        self.Np_total = Np
        self.region_count = 1
        self.ray_batch_count = 1
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
        elif ne_type == "test_slab":
            # ne_0 (1 + s x / L_x), src/core/domain.py:580-603
            n0 = 2e23 if ne_0 is None else ne_0
            sc = 1.0 if s is None else s
            xx = self.x[:, None, None].astype(np.float64)
            self.ne = np.ascontiguousarray(np.broadcast_to(
                n0 * (1.0 + sc * xx / self.x_length), (self.x_n, self.y_n, self.z_n)
            ).astype(np.float32))
        elif ne_type == "import":
            raise AssertionError("ne_type='import' needs ne=...")
        else:
            raise NotImplementedError(f"ne_type={ne_type!r} is not provided by sloth_amd")

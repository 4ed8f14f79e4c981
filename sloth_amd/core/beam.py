"""Probe-beam ray initialiser (host side).

Mirrors ``core.beam.Beam`` (src/core/beam.py:8-346): the same constructor
signature, the same ``RandomState(42 if seeded else None)`` stream and the
same draw order per beam type, so a seeded beam reproduces the reference's
ray set (up to the reference computing the trigonometry in float32).  Ray
synthesis stays on the host in NumPy to keep that RNG stream; it is input
preparation, not part of the traced path.

``device_rng=True`` (an extension) makes the rays on the GPU instead, with the
k6 kernel's counter-based generator: same distributions, a different (but
seeded and launch-independent) stream, and ``s0`` a device tensor.
"""
from __future__ import annotations

import os

import numpy as np

from ..constants import C_LIGHT

_BEAM_TYPES = {"circular": 0, "square": 1, "rectangular": 2, "rect_trackers": 2, "linear": 3}


def device_beam(n, beam_size, divergence, ne_extent, *, probing_direction="z",
                beam_type="circular", seed=42, device=None, stream=None):
    """(6, n) float32 rays made on the GPU by ``slt_beam`` (Philox4x32-10 per ray)."""
    import torch

    from .. import _native as N

    if beam_type not in _BEAM_TYPES:
        raise ValueError("beam_type unrecognised! Accepted: circular, square, rectangular, "
                         "linear, rect_trackers.")
    if beam_type in ("rectangular", "rect_trackers"):
        if len(beam_size) != 2:
            raise AssertionError("Must pass a list of length 2 to initialise a rectangular beam.")
        a, b = float(beam_size[0]), float(beam_size[1])
    else:
        if not np.isscalar(beam_size):
            raise AssertionError("Received a sized beam_size, expected a float.")
        a = b = float(beam_size)
    axis = {"x": 0, "y": 1, "z": 2}.get(probing_direction, 1)
    dev = device if device is not None else N.require_device()
    s0 = torch.empty((6, int(n)), dtype=torch.float32, device=dev)
    N.check(N.lib().slt_beam(int(n), _BEAM_TYPES[beam_type], axis, a, b, float(divergence),
                             float(ne_extent), int(seed) & 0xFFFFFFFFFFFFFFFF, s0.data_ptr(),
                             N.stream_handle(stream)))
    return s0


class Beam:
    def __init__(self, Np, beam_size, divergence, ne_extent, *, probing_direction="z",
                 beam_type="circular", seeded=False, device_rng=False):
        self.Np = np.int64(Np)
        self.beam_size = beam_size
        self.divergence = divergence
        self.ne_extent = ne_extent
        self.probing_direction = probing_direction
        self.beam_type = beam_type
        self.seed = 42 if seeded else None
        if device_rng:
            if beam_type == "even":
                raise NotImplementedError("beam_type='even' is not functional in the reference")
            seed = self.seed if seeded else int.from_bytes(os.urandom(8), "little")
            self.s0 = device_beam(self.Np, beam_size, divergence, ne_extent,
                                  probing_direction=probing_direction, beam_type=beam_type,
                                  seed=seed)
            return
        self.init_beam(-self.ne_extent, self.seed)

    def init_beam(self, ne_extent, seed):
        """Fill ``self.s0`` (6, Np) float32: x, y, z [m], vx, vy, vz [m/s]."""
        c = C_LIGHT
        n = int(self.Np)
        rng = np.random.RandomState(seed)
        s0 = np.zeros((6, n), dtype=np.float64)
        pd = self.probing_direction
        bt = self.beam_type

        def place(a, b, chi, phi):
            # velocity about the probing axis and transverse positions a, b
            # (src/core/beam.py:115-145, 158-188, 204-234)
            if pd == "x":
                s0[3] = c * np.cos(chi)
                s0[4] = c * np.sin(chi) * np.cos(phi)
                s0[5] = c * np.sin(chi) * np.sin(phi)
                s0[0] = ne_extent
                s0[1] = a
                s0[2] = b
            elif pd == "z":
                s0[3] = c * np.sin(chi) * np.cos(phi)
                s0[4] = c * np.sin(chi) * np.sin(phi)
                s0[5] = c * np.cos(chi)
                s0[0] = a
                s0[1] = b
                s0[2] = ne_extent
            else:  # the reference defaults every other value to y
                s0[4] = c * np.cos(chi)
                s0[3] = c * np.sin(chi) * np.cos(phi)
                s0[5] = c * np.sin(chi) * np.sin(phi)
                s0[0] = a
                s0[1] = ne_extent
                s0[2] = b

        if bt == "circular":
            if not np.isscalar(self.beam_size):
                raise AssertionError("Received a sized beam_size, expected a float.")
            t = 2 * np.pi * rng.rand(n)
            u = rng.power(2, n)
            phi = np.pi * rng.rand(n)
            chi = self.divergence * rng.randn(n)
            r = self.beam_size * u
            place(r * np.cos(t), r * np.sin(t), chi, phi)
        elif bt == "square":
            if not np.isscalar(self.beam_size):
                raise AssertionError("Received a sized beam_size, expected a float.")
            t = 2 * rng.rand(n) - 1.0
            u = 2 * rng.rand(n) - 1.0
            phi = np.pi * rng.rand(n)
            chi = self.divergence * rng.randn(n)
            place(self.beam_size * u, self.beam_size * t, chi, phi)
        elif bt in ("rectangular", "rect_trackers"):
            if len(self.beam_size) != 2:
                raise AssertionError(
                    "Must pass a list of length 2 to initialise a rectangular beam.")
            t = 2 * rng.rand(n) - 1.0
            u = 2 * rng.rand(n) - 1.0
            phi = np.pi * rng.rand(n)
            chi = self.divergence * rng.randn(n)
            place(self.beam_size[0] * u, self.beam_size[1] * t, chi, phi)
        elif bt == "linear":
            if not np.isscalar(self.beam_size):
                raise AssertionError("Received a sized beam_size, expected a float.")
            t = 2 * rng.rand(n) - 1.0
            chi = self.divergence * rng.randn(n)
            # always along z in the x-z plane (src/core/beam.py:238-254)
            s0[3] = c * np.sin(chi)
            s0[4] = 0.0
            s0[5] = c * np.cos(chi)
            s0[0] = self.beam_size * t
            s0[1] = 0.0
            s0[2] = ne_extent
        elif bt == "even":
            # the reference's concentric-disc beam never produces rays
            # (src/core/beam.py:255-272); refuse it explicitly
            raise NotImplementedError("beam_type='even' is not functional in the reference")
        else:
            raise ValueError(
                "beam_type unrecognised! Accepted: circular, square, rectangular, linear, "
                "rect_trackers.")
        self.s0 = s0.astype(np.float32)

    def save_rays_pos(self, fn=None):
        """Save ``s0`` as .npy (src/core/beam.py:348-370)."""
        from datetime import datetime

        if fn is None:
            fn = "{} rays.npy".format(datetime.now().strftime("%Y-%m-%d_%H-%M-%S"))
        else:
            fn = "{}.npy".format(fn)
        with open(fn, "wb") as f:
            np.save(f, self.s0)

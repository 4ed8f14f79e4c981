"""Synthetic electron-density fields for the benchmark workloads and tests.

Input preparation, not the traced path.  Coordinates follow
``ScalarDomain``: float32 ``linspace(-L/2, L/2, n)`` per axis.

* ``sine_field``       n_e = n0 (1 + a sin(2 pi x / lam))
* ``quadratic_trough`` n_e = n0 (1 + (x^2 + y^2) / R^2)
* ``kolmogorov_field`` Gaussian random field with power spectrum P(k) ~ k^p
  (p = -11/3 by default) on k_min <= |k| <= k_max, zero mean and unit
  variance, from a seeded NumPy RandomState.  The reference's
  ``gaussian3D.domain_fft`` (src/field_generator/gaussian3D.py:215-273) uses a
  k^-5/3 default and normalises by max|field|; both are options here.
"""
from __future__ import annotations

import numpy as np


def coords(n: int, length: float) -> np.ndarray:
    return np.float32(np.linspace(-length / 2, length / 2, n))


def sine_field(n, length, n0=1e24, amp=0.1, wavelength=2e-3):
    x = coords(n, length).astype(np.float64)
    prof = n0 * (1.0 + amp * np.sin(2 * np.pi * x / wavelength))
    ne = np.broadcast_to(prof[:, None, None], (n, n, n)).astype(np.float32)
    c = coords(n, length)
    return np.ascontiguousarray(ne), (c, c.copy(), c.copy())


def quadratic_trough(n, length, n0=5e24, R=3e-3):
    x = coords(n, length).astype(np.float64)
    prof = n0 * (1.0 + (x[:, None] ** 2 + x[None, :] ** 2) / R ** 2)
    ne = np.broadcast_to(prof[:, :, None], (n, n, n)).astype(np.float32)
    c = coords(n, length)
    return np.ascontiguousarray(ne), (c, c.copy(), c.copy())


def gaussian_random_field(n, length, *, l_max, l_min, exponent=-11.0 / 3.0, seed=0,
                          normalise="std", device=None):
    """Real GRF on an n^3 grid of edge ``length`` with P(k) ~ k^exponent on the band.

    White noise is drawn from ``np.random.RandomState(seed)`` (float32) so the
    field is reproducible; the FFTs run with torch on ``device`` when given
    (rocFFT on the GPU), else with numpy.  Returns float32 (n, n, n).
    """
    rng = np.random.RandomState(seed)
    noise = rng.standard_normal((n, n, n)).astype(np.float32)
    dx = length / (n - 1)
    k1 = 2 * np.pi * np.fft.fftfreq(n, d=dx)
    k1r = 2 * np.pi * np.fft.rfftfreq(n, d=dx)
    kmin, kmax = 2 * np.pi / l_max, 2 * np.pi / l_min
    if device is not None:
        import torch

        t = torch.from_numpy(noise).to(device)
        F = torch.fft.rfftn(t)
        del t
        kx = torch.from_numpy(k1.astype(np.float32)).to(device)
        kz = torch.from_numpy(k1r.astype(np.float32)).to(device)
        k = torch.sqrt(kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2)
        band = (k >= kmin) & (k <= kmax)
        amp = torch.where(band, torch.clamp(k, min=kmin) ** (exponent / 2.0),
                          torch.zeros_like(k))
        del k, band
        F *= amp
        del amp
        g = torch.fft.irfftn(F, s=(n, n, n))
        del F
        g = g - g.mean()
        if normalise == "std":
            g = g / g.std()
        else:
            g = g / g.abs().max()
        return g.to("cpu", torch.float32).numpy()
    F = np.fft.rfftn(noise)
    k = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + k1r[None, None, :] ** 2)
    band = (k >= kmin) & (k <= kmax)
    amp = np.zeros_like(k)
    amp[band] = k[band] ** (exponent / 2.0)
    g = np.fft.irfftn(F * amp, s=(n, n, n), axes=(0, 1, 2))
    g = g - g.mean()
    g = g / (g.std() if normalise == "std" else np.abs(g).max())
    return g.astype(np.float32)


def kolmogorov_field(n, length, *, n0=2e24, frac=0.3, l_max=None, l_min=0.2e-3, seed=0,
                     exponent=-11.0 / 3.0, device=None):
    """n_e = max(0, n0 (1 + frac g)), g a unit-variance k^exponent GRF."""
    g = gaussian_random_field(n, length, l_max=length if l_max is None else l_max, l_min=l_min,
                              exponent=exponent, seed=seed, device=device)
    ne = np.maximum(0.0, n0 * (1.0 + frac * g.astype(np.float64))).astype(np.float32)
    c = coords(n, length)
    return ne, (c, c.copy(), c.copy())

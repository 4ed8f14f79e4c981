"""NumPy oracle for the Gaussian-random-field generators.

TEST INFRASTRUCTURE ONLY: ``tests/`` import this module to check
``sloth_amd.field_generator.gaussian3D``; the product never imports it.

Restated from MAGPIE-ICL/SLOTH ``src/field_generator/gaussian3D.py``: ``cos``
(:19-151, Fourier summation over random modes), ``fft`` (:153-213, Timmer &
Koenig synthesis) and ``domain_fft`` (:215-273, band-limited field, NumPy's
global generator seeded with 42).  All arithmetic in fp64 NumPy, FFTs by
numpy.fft.

Pinned in ``tests/test_fields_oracle.py`` by properties the constructions fix:
``domain_fft``'s spectrum vanishes outside the band and the field's maximum
magnitude is 1; ``fft``'s Hermitian noise makes the field real and its zeroed
k = 0 term makes its mean 0; every ``cos`` mode is a plane wave of wavenumber
wn, so a one-mode field satisfies the Helmholtz equation lap r = -wn^2 r.
"""
from __future__ import annotations

import numpy as np


def cos(k_func, lx, ly, lz, nx, ny, nz, nmodes, wn1):
    dx, dy, dz = lx / nx, ly / ny, lz / nz
    wnn = max(np.pi / dx, np.pi / dy, np.pi / dz)
    dk = (wnn - wn1) / nmodes
    wn = wn1 + 0.5 * dk + np.arange(0, nmodes) * dk
    A = np.sqrt(2.0 * np.clip(k_func(wn), 0.0, None) * np.full(nmodes, dk) ** 3)
    psi = [2.0 * np.pi * np.random.uniform(0.0, 1.0, nmodes) for _ in range(4)]
    theta = 2.0 * np.pi * np.random.uniform(0.0, 1.0, nmodes)
    phi = 2.0 * np.pi * np.random.uniform(0.0, 1.0, nmodes)
    kx = np.sin(theta) * np.cos(phi) * wn
    ky = np.sin(theta) * np.sin(phi) * wn
    kz = np.cos(theta) * wn
    xc = dx / 2.0 + np.arange(nx) * dx
    yc = dy / 2.0 + np.arange(ny) * dy
    zc = dz / 2.0 + np.arange(nz) * dz
    X, Y, Z = xc[:, None, None, None], yc[None, :, None, None], zc[None, None, :, None]
    ax, ay, az = kx * X, ky * Y, kz * Z
    s = (np.cos(ax + ay + az + psi[0]) + np.cos(ax + ay - az + psi[1]) +
         np.cos(ax - ay + az + psi[2]) + np.cos(ax - ay - az + psi[3]))
    return (A * np.sqrt(2.0) * s).sum(axis=-1), wn


def fft(k_func, N):
    M = 2 * N + 1
    k = np.fft.fftfreq(M)
    KX, KY, KZ = np.meshgrid(k, k, k)
    K = np.fft.fftshift(np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2))
    Wr = np.random.randn(M, M, M)
    Wi = np.random.randn(M, M, M)
    W = (Wr + np.flip(Wr)) + 1j * (Wi - np.flip(Wi))
    with np.errstate(divide="ignore", invalid="ignore"):
        F = np.fft.ifftshift(W * np.sqrt(k_func(K)))
    F[0, 0, 0] = 0
    return np.fft.ifftn(F)


def domain_fft(k_func, l_max, l_min, extent, res, factor):
    dx = extent / res
    kx = 2 * np.pi * np.fft.fftfreq(2 * res, d=dx)
    kz = 2 * np.pi * np.fft.fftfreq(int(2 * res * factor), d=dx)
    kxx, kyy, kzz = np.meshgrid(kx, kx, kz)
    k = np.sqrt(kxx ** 2 + kyy ** 2 + kzz ** 2, dtype=np.float32)
    S = np.zeros_like(k)
    band = (k >= 2 * np.pi / l_max) & (k <= 2 * np.pi / l_min)
    S[band] = k_func(k[band])
    np.random.seed(seed=42)
    noise = np.random.normal(0, 1, k.shape) + 1j * np.random.normal(0, 1, k.shape)
    field = np.fft.ifftn(noise * np.sqrt(S)).real
    return field / np.abs(field).max(), k, band

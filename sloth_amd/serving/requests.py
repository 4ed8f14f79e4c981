"""Ray-trace request specifications and their on-device synthesis.

A request names its density field (kind, grid side, box length, seed), its
beam (shape, half-width, ray count, seed), the wavelength and the absorption
settings.  ``materialize`` builds the packed grid and the rays directly in HBM
(rocFFT through torch for the Kolmogorov field, the k6 counter-based beam
kernel for the rays), so the host never touches grid-sized data.  The same spec always
produces the same inputs, which is what lets a request be compared alone and
inside a serving mix.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from ..constants import C_LIGHT
from ..core import engine
from ..core.beam import device_beam


@dataclass(frozen=True)
class FieldSpec:
    kind: str = "kolmogorov"      # "kolmogorov" | "sine"
    side: int = 128               # grid points per axis
    length: float = 10e-3         # box edge [m]
    seed: int = 0
    n0: float = 2e24              # mean density [m^-3]
    frac: float = 0.3             # relative fluctuation (kolmogorov) / sine amplitude
    l_min: float = 0.2e-3         # smallest scale of the kolmogorov band [m]
    wavelength: float = 2e-3      # sine period [m]


@dataclass(frozen=True)
class BeamSpec:
    n_rays: int = 10000
    half_width: float = 3e-3      # square beam half-width [m]
    seed: int = 0


@dataclass
class RequestSpec:
    rid: int
    field: FieldSpec
    beam: BeamSpec
    lwl: float = 532e-9
    inv_brems: bool = True
    Te: float = 20.0
    Z: float = 1.0
    coulomb_log: float | None = 10.0
    bins: int = 256
    arrival: float = 0.0          # seconds after the start of a run
    extra: dict = field(default_factory=dict)

    def horizon(self) -> float:
        """The reference solve horizon sqrt(8) L / c (src/core/propagator.py:616)."""
        return math.sqrt(8.0) * self.field.length / C_LIGHT

    def footprint_bytes(self) -> int:
        """HBM needed while the request is resident: the packed grid and the field's
        transients (16 + 16 B per voxel); per ray the rays, order, resumable state and
        outputs (~160 B) and, at set-up and finish, the permutation workspace and the
        caller-order copies (~130 B); the detector image."""
        n3 = self.field.side ** 3
        return int(16 * n3 + 16 * n3 + 288 * self.beam.n_rays + 8 * self.bins ** 2 + (1 << 20))


def _coords(spec: FieldSpec):
    c = np.float32(np.linspace(-spec.length / 2, spec.length / 2, spec.side))
    return c, c.copy(), c.copy()


_AMP: dict = {}  # (n, length, l_min, device) -> (spectrum, event recorded after building it)
_AMP_MAX = 8


def _spectrum_amp(n: int, length: float, l_min: float, device) -> torch.Tensor:
    """|k|^(-11/6) on the band 2 pi / length <= |k| <= 2 pi / l_min of the (n, n, n / 2 + 1)
    half spectrum, zero elsewhere.  It depends only on the grid, so it is built once per
    (n, length, l_min) and shared (read-only) by every request of that shape."""
    key = (n, float(length), float(l_min), str(device))
    use = torch.cuda.current_stream(device)
    hit = _AMP.get(key)
    if hit is not None:
        amp, built = hit
        use.wait_event(built)   # built on another stream: wait for it on the GPU
        return amp
    dx = length / (n - 1)
    kx = torch.fft.fftfreq(n, d=dx, device=device, dtype=torch.float32) * (2 * math.pi)
    kz = torch.fft.rfftfreq(n, d=dx, device=device, dtype=torch.float32) * (2 * math.pi)
    k = torch.sqrt(kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2)
    kmin, kmax = 2 * math.pi / length, 2 * math.pi / l_min
    amp = torch.where((k >= kmin) & (k <= kmax), torch.clamp(k, min=kmin) ** (-11.0 / 6.0),
                      torch.zeros_like(k))
    if len(_AMP) >= _AMP_MAX:
        torch.cuda.synchronize(device)   # no queued kernel still reads the evicted spectrum
        _AMP.pop(next(iter(_AMP)))
    built = torch.cuda.Event()
    built.record(use)
    _AMP[key] = (amp, built)
    return amp


def field_on_device(spec: FieldSpec, device) -> torch.Tensor:
    """(n, n, n) float32 electron density [m^-3] built on the device."""
    n = spec.side
    if spec.kind == "sine":
        x = engine.const_device(_coords(spec)[0].astype(np.float64), device)
        prof = spec.n0 * (1.0 + spec.frac * torch.sin(2 * math.pi * x / spec.wavelength))
        return prof.to(torch.float32)[:, None, None].expand(n, n, n).contiguous()
    if spec.kind != "kolmogorov":
        raise ValueError(f"unknown field kind {spec.kind!r}")
    gen = torch.Generator(device=device)
    gen.manual_seed(int(spec.seed))
    F = _half_spectrum_noise(n, gen, device)
    F *= _spectrum_amp(n, spec.length, spec.l_min, device)
    g = torch.fft.irfftn(F, s=(n, n, n))
    del F
    # n0 (1 + frac (g - mean) / std), clipped at 0: one std/mean pass, then
    # g * a + b in place with a = frac n0 / std, b = n0 - mean a
    std, mean = torch.std_mean(g)
    a = (spec.frac * spec.n0) / std
    return g.mul_(a).add_(spec.n0 - mean * a).clamp_(min=0.0)


def _half_spectrum_noise(n: int, gen: torch.Generator, device) -> torch.Tensor:
    """White complex Gaussian noise on the (n, n, n / 2 + 1) half spectrum of a real n^3
    field, drawn directly in k-space (the spectrum of real white noise, without the noise
    field and its forward transform).  In the planes kz = 0 and, for even n, kz = n / 2 a
    mode's conjugate partner (-kx, -ky) lies in the same plane, so those planes are made
    Hermitian, P <- (P + conj(P(-kx, -ky))) / sqrt(2), which keeps every mode's expected
    power and makes the self-conjugate modes real; the real inverse transform is then the
    exact one."""
    F = torch.view_as_complex(
        torch.randn((n, n, n // 2 + 1, 2), generator=gen, device=device, dtype=torch.float32))
    neg = engine.const_device((-np.arange(n, dtype=np.int64)) % n, device)  # -k mod n
    for kz in ((0, n // 2) if n % 2 == 0 else (0,)):
        P = F[:, :, kz]
        Q = P.index_select(0, neg).index_select(1, neg).conj()
        P.add_(Q).mul_(1.0 / math.sqrt(2.0))
    return F


def beam_on_device(spec: BeamSpec, length: float, device) -> torch.Tensor:
    """(6, N) collimated square beam entering the z = -L/2 face, made by the k6 kernel."""
    return device_beam(spec.n_rays, spec.half_width, 0.0, float(np.float32(length / 2)),
                       probing_direction="z", beam_type="square", seed=int(spec.seed) + 7919,
                       device=device)


@dataclass
class Materialized:
    """Device-resident inputs of one request."""
    grid: engine.PackedGrid
    rays: torch.Tensor
    save_times: np.ndarray
    edges: np.ndarray


def materialize(req: RequestSpec, device) -> Materialized:
    ne = field_on_device(req.field, device)
    omega = 2 * math.pi * C_LIGHT / req.lwl
    grid = engine.PackedGrid(ne, *_coords(req.field), omega, inv_brems=req.inv_brems,
                             Te=req.Te if req.inv_brems else None,
                             Z=req.Z if req.inv_brems else None,
                             coulomb_log=req.coulomb_log, device=device)
    del ne
    rays = beam_on_device(req.beam, req.field.length, device)
    L = req.field.length
    return Materialized(grid, rays, np.array([0.0, req.horizon()]),
                        np.linspace(-L / 2, L / 2, req.bins + 1))

"""Analytic and turbulent electron-density domains, built on the GPU.

Mirrors ``shared.domains`` (src/shared/domains.py:1-62): ``grid`` ('ij'
meshgrid of linspace coordinates), ``distance``, ``spherical`` (nested shells),
``density_step`` (rho1 for y > 0, rho2 below) and ``turbulent_box`` (ne_0 plus a
``gaussian3D.domain_fft`` field, spectrum k^power).  The grids are torch
tensors on the device (the reference builds jax arrays); pass ``device`` to
choose it.
"""
from __future__ import annotations

import torch

from .. import _native as _N
from ..field_generator.gaussian3D import gaussian3D


def distance(x, y, z):
    return torch.sqrt(x ** 2 + y ** 2 + z ** 2)


def grid(lengths, n_cells, *, device=None, dtype=torch.float32):
    """(XX, YY, ZZ) of linspace(-len/2, len/2, n) per axis, 'ij' indexing."""
    dev = device if device is not None else _N.require_device()
    if isinstance(n_cells, (float, int)):
        n_cells = 3 * [n_cells]
    axes = [torch.linspace(-lengths[a] / 2, lengths[a] / 2, int(n_cells[a]), dtype=dtype,
                           device=dev) for a in range(3)]
    return torch.meshgrid(*axes, indexing="ij")


def spherical(radii, ne, rho, lengths, n_cells, *, device=None):
    """Shell i (radii[i-1] < r <= radii[i], radii[-1] = 0 implied first) holds ne[i]
    and, in the i-th returned grid, rho[i]."""
    radii = (0,) + tuple(radii)
    XX, YY, ZZ = grid(lengths, n_cells, device=device)
    ne_grid = torch.zeros_like(XX)
    rho_grid_list = [torch.zeros_like(XX) for _ in range(len(rho))]
    d = distance(XX, YY, ZZ)
    for i in range(len(radii) - 1):
        shell = (radii[i] < d) & (d <= radii[i + 1])
        ne_grid[shell] = ne[i]
        rho_grid_list[i][shell] = rho[i]
    return ne_grid, rho_grid_list


def density_step(rho1, rho2, lengths, n_cells, *, device=None):
    """rho1 where y > 0, rho2 where y <= 0."""
    _, YY, _ = grid(lengths, n_cells, device=device)
    density = torch.full_like(YY, rho1)
    density[YY <= 0] = rho2
    return density


def turbulent_box(ne_0, max_pert, l_max, l_min, extent, n_cells, power=-5 / 3):
    """ne_0 + max_pert * (band-limited k^power field normalised to max |.| = 1);
    n_cells must be even (the field has n_cells points per axis)."""
    def k_func(k):
        return k ** (power)

    field = gaussian3D(k_func)
    noise3D = field.domain_fft(l_max, l_min, extent, n_cells // 2, factor=1.0)
    return ne_0 + noise3D * max_pert

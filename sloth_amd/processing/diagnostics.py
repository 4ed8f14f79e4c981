"""Detector histogram of traced rays, on the GPU.

Mirrors ``processing.diagnostics.Diagnostic`` (src/processing/diagnostics.py:171-282)
up to and including ``histogram``: the same constructor arguments, the NaN
and CCD masks, optional per-ray weights, and numpy.histogram2d's binning and
``H.T`` layout.  Binning runs in the k3 kernel (``slt_histogram2d``): integer
counts are bit-identical to numpy's; weighted sums equal numpy's up to fp64
summation order.  ``H`` is returned as float64 like numpy's.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _native as N
from ..core.engine import to_device


def m_to_mm(r):
    """Scale the position rows (0 and 2) of a Jones vector from m to mm."""
    rr = np.copy(r)
    rr[0::2, :] *= 1e3
    return rr


def mm_to_m(r):
    rr = np.copy(r)
    rr[0::2, :] *= 1e-3
    return rr


def histogram2d_device(x: torch.Tensor, y: torch.Tensor, xedges: np.ndarray, yedges: np.ndarray,
                       weights: torch.Tensor | None = None, mask: torch.Tensor | None = None):
    """numpy.histogram2d(x, y, bins=(xedges, yedges), weights).T on the GPU.

    x, y: float32 or float64 device vectors of one dtype (any stride).
    Returns (nyb, nxb) tensors: counts as int64 and, when weights are given,
    fp64 weighted sums.
    """
    if x.dtype != y.dtype or x.dtype not in (torch.float32, torch.float64):
        raise TypeError("x and y must both be float32 or both float64")
    dev = x.device
    nxb, nyb = len(xedges) - 1, len(yedges) - 1
    xe = to_device(np.asarray(xedges, dtype=np.float64), dev)
    ye = to_device(np.asarray(yedges, dtype=np.float64), dev)
    counts = torch.zeros((nyb, nxb), dtype=torch.int32, device=dev)
    wsum = None
    w = None
    if weights is not None:
        w = weights.to(device=dev, dtype=torch.float64).contiguous()
        wsum = torch.zeros((nyb, nxb), dtype=torch.float64, device=dev)
    m = None
    if mask is not None:
        m = mask.to(device=dev, dtype=torch.uint8).contiguous()
    n = x.numel()
    N.check(N.lib().slt_histogram2d(x.data_ptr(), x.stride(0), y.data_ptr(), y.stride(0),
                                    1 if x.dtype == torch.float64 else 0, n, N.ptr(m),
                                    xe.data_ptr(), nxb, ye.data_ptr(), nyb, N.ptr(w),
                                    counts.data_ptr(), N.ptr(wsum), N.stream_handle()))
    return counts.to(torch.int64) & 0xFFFFFFFF, wsum


def _outer_edges(lo, hi):
    """numpy's _get_outer_edges for an explicit range."""
    lo, hi = float(lo), float(hi)
    if lo > hi:
        raise ValueError("max must be larger than min in range parameter.")
    if not (np.isfinite(lo) and np.isfinite(hi)):
        raise ValueError(f"supplied range of [{lo}, {hi}] is not finite")
    if lo == hi:
        lo, hi = lo - 0.5, hi + 0.5
    return lo, hi


class Diagnostic:
    """Ray diagnostic base: holds a (4, N) Jones vector and bins it on a detector."""

    def __init__(self, rf, *, weights=None, focal_plane=0, L=400, R=25, Lx=18, Ly=13.5,
                 prefilter_input=False, ccd_shape_m=None):
        self.focal_plane, self.L, self.R, self.Lx, self.Ly = focal_plane, L, R, Lx, Ly
        if ccd_shape_m is not None:
            self._ccd_half_x_mm = ccd_shape_m[0] * 1e3 / 2.0
            self._ccd_half_y_mm = ccd_shape_m[1] * 1e3 / 2.0
        else:
            self._ccd_half_x_mm = None
            self._ccd_half_y_mm = None
        assert rf is not None, "rf should not be None"
        assert rf.shape[0] == 4 or (len(rf.shape) == 3 and rf.shape[1] == 4), \
            "\nIncorrect format for rf, are you sure you passed the right variable?"
        if len(rf.shape) == 3:
            rf = rf[-1, :, :]
        self.Np = rf.shape[-1]
        if isinstance(rf, torch.Tensor):
            rf = rf.detach().cpu().numpy()
        self.rf = np.asarray(rf)
        self.Np_inc = self.Np
        if prefilter_input:
            xp = self.L * np.tan(self.rf[1]) + self.rf[0]
            yp = self.L * np.tan(self.rf[3]) + self.rf[2]
            keep = np.sqrt(xp ** 2 + yp ** 2) <= self.R
            self.rf = self.rf[:, keep]
            if weights is not None:
                weights = np.asarray(weights)[keep]
            self.Np_inc = self.rf.shape[-1]
        self.r0 = m_to_mm(self.rf)
        self.weights = np.asarray(weights, dtype=np.float64) if weights is not None else None

    def histogram(self, *, bin_scale=1, pix_x=3448, pix_y=2574, clear_mem=False,
                  plain_plot=False, extra_info=True):
        """Bin rows 0 and 2 into ``H`` (pix_y/bin_scale, pix_x/bin_scale) over ±Lx/2, ±Ly/2.

        src/processing/diagnostics.py:243-276 (NaN mask, optional CCD mask,
        numpy.histogram2d, transpose).
        """
        dev = N.require_device()
        matrix = self.r0 if plain_plot else self.rf
        nxb = int(np.floor(pix_x / bin_scale))
        nyb = int(np.floor(pix_y / bin_scale))
        if nxb < 1 or nyb < 1:
            raise ValueError("`bins[0]` must be positive, when an integer")
        xlo, xhi = _outer_edges(-self.Lx / 2, self.Lx / 2)
        ylo, yhi = _outer_edges(-self.Ly / 2, self.Ly / 2)
        self.xedges = np.linspace(xlo, xhi, nxb + 1)
        self.yedges = np.linspace(ylo, yhi, nyb + 1)
        m = np.asarray(matrix)
        keep = ~np.isnan(m[0]) & ~np.isnan(m[2])
        if self._ccd_half_x_mm is not None:
            keep &= (np.abs(m[0]) <= self._ccd_half_x_mm) & (np.abs(m[2]) <= self._ccd_half_y_mm)
        dt = np.float32 if m.dtype == np.float32 else np.float64
        xs = torch.from_numpy(np.ascontiguousarray(m[0], dtype=dt)).to(dev)
        ys = torch.from_numpy(np.ascontiguousarray(m[2], dtype=dt)).to(dev)
        counts, wsum = histogram2d_device(
            xs, ys, self.xedges, self.yedges,
            weights=None if self.weights is None else torch.from_numpy(self.weights),
            mask=torch.from_numpy(keep.astype(np.uint8)))
        H = counts.to(torch.float64) if wsum is None else wsum
        self.H = H.cpu().numpy()
        if clear_mem:
            self.r0 = None
            self.rf = None

    def plot(self, ax, clim=None, cmap=None):
        ax.imshow(self.H, interpolation="nearest", origin="lower", clim=clim, cmap=cmap,
                  extent=[self.xedges[0], self.xedges[-1], self.yedges[0], self.yedges[-1]])

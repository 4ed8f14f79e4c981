"""Ray diagnostics on the GPU: ray-transfer-matrix optics and the detector histogram.

Mirrors ``processing.diagnostics`` (src/processing/diagnostics.py): the optic
elements ``lens``/``sym_lens``/``travel``/``circular_aperture``/
``circular_stop``/``annular_stop``/``rect_aperture``/``knife_edge`` (:28-128),
the helpers ``ray``/``d2r``/``lens_cutoff`` (:140-169), ``Diagnostic`` with its
NaN and CCD masks and ``histogram`` (:171-282), the ``Shadowgraphy``,
``Schlieren`` and ``Refractometry`` solve chains (:284-406), and
``transmission_map``/``apply_ccd_mask`` (:541-603).

The optics run in the k5 kernel (``slt_rtm``): a solve method hands its whole
element chain to one launch, which walks it per ray in fp64 registers.
Binning runs in the k3 kernel (``slt_histogram2d``): integer counts are
bit-identical to numpy.histogram2d's; weighted sums equal numpy's up to fp64
summation order.  The element functions accept the reference's host (4, N)
arrays (and return host arrays, modifying them in place where the reference
does) as well as device tensors.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _native as N
from ..core.engine import const_device, to_device


def m_to_mm(r):
    """Scale the position rows (0 and 2) of a Jones vector from m to mm (:16-20)."""
    rr = np.copy(r)
    rr[0::2, :] *= 1e3
    return rr


def mm_to_m(r):
    """Scale the position rows (0 and 2) from mm to m (:22-26)."""
    rr = np.copy(r)
    rr[0::2, :] *= 1e-3
    return rr


# ---------------------------------------------------------------------------
# k5: optic chains
# ---------------------------------------------------------------------------

def _optic(kind, a=0.0, b=0.0, axis=0) -> N.Optic:
    o = N.Optic()
    o.kind = kind
    o.axis = axis
    o.a = float(a)
    o.b = float(b)
    return o


def _device_rays(r, device=None) -> torch.Tensor:
    """(4, N) float32/float64 device tensor with unit column stride (no copy if it is one)."""
    if isinstance(r, torch.Tensor):
        t = r.detach()
        if not t.is_cuda:
            t = t.to(device if device is not None else N.require_device())
    else:
        a = np.asarray(r)
        if a.dtype not in (np.float32, np.float64):
            a = a.astype(np.float64)
        t = to_device(a, device if device is not None else N.require_device())
    if t.dim() != 2 or t.shape[0] != 4:
        raise AssertionError("Expected a (4, N) ray array [x, theta, y, phi].")
    if t.dtype not in (torch.float32, torch.float64):
        t = t.to(torch.float64)
    if t.shape[1] > 0 and t.stride(1) != 1:
        t = t.contiguous()
    return t


def rtm_device(rays, ops, *, pos_scale: float = 1.0) -> torch.Tensor:
    """Apply an element chain to (4, N) rays on the GPU; returns a new (4, N) fp64 tensor.

    ``ops`` is a list of ``N.Optic`` (see ``include/sloth_hip.h``); rows 0 and 2
    are first multiplied by ``pos_scale`` in the input's precision.
    """
    t = _device_rays(rays)
    if len(ops) > N.SLT_MAX_OPTICS:
        raise ValueError(f"at most {N.SLT_MAX_OPTICS} optic elements per chain")
    n = t.shape[1]
    out = torch.empty((4, n), dtype=torch.float64, device=t.device)
    arr = (N.Optic * max(len(ops), 1))(*ops)
    N.check(N.lib().slt_rtm(t.data_ptr(), 1 if t.dtype == torch.float64 else 0,
                            t.stride(0) if n else n, n, float(pos_scale), arr, len(ops),
                            out.data_ptr(), n, N.stream_handle()))
    return out


def _apply(r, op, in_place: bool):
    """One element on host or device rays, returned in the caller's kind of array."""
    out = rtm_device(r, [op])
    if isinstance(r, torch.Tensor):
        if in_place:
            r.copy_(out.to(device=r.device, dtype=r.dtype))
            return r
        return out
    host = out.cpu().numpy()
    if in_place:
        r[...] = host.astype(r.dtype, copy=False)
        return r
    return host


def lens(r, f1, f2):
    """Thin lens, focal lengths f1 (x) and f2 (y): [[1, 0], [-1/f, 1]] per axis (:28-43)."""
    return _apply(r, _optic(N.SLT_OPTIC_LENS, f1, f2), in_place=False)


def sym_lens(r, f):
    """Axisymmetric lens (:45-50)."""
    return lens(r, f, f)


def travel(r, d):
    """Free travel over d: [[1, d], [0, 1]] per axis (:52-65)."""
    return _apply(r, _optic(N.SLT_OPTIC_TRAVEL, d), in_place=False)


def circular_aperture(r, R, E=None):
    """Rays with x^2 + y^2 > R^2 become NaN, in place (:67-75)."""
    return _apply(r, _optic(N.SLT_OPTIC_CIRCULAR_APERTURE, R), in_place=True)


def circular_stop(r, R):
    """Rays with x^2 + y^2 < R^2 become NaN, in place (:77-85)."""
    return _apply(r, _optic(N.SLT_OPTIC_CIRCULAR_STOP, R), in_place=True)


def annular_stop(r, R1, R2):
    """The filter R1^2 < x^2 + y^2 < R2^2 (returned, r untouched, as in :87-96)."""
    t = _device_rays(r)
    x2 = t[0].to(torch.float64) * t[0].to(torch.float64)
    y2 = t[2].to(torch.float64) * t[2].to(torch.float64)
    s = x2 + y2
    filt = (s > float(R1) ** 2) & (s < float(R2) ** 2)
    return filt if isinstance(r, torch.Tensor) else filt.cpu().numpy()


def rect_aperture(r, Lx, Ly):
    """Rays with x^2 > Lx^2 and y^2 > Ly^2 become NaN, in place (:98-109)."""
    return _apply(r, _optic(N.SLT_OPTIC_RECT_APERTURE, Lx, Ly), in_place=True)


def knife_edge(r, offset, axis, direction):
    """Knife edge along 'x' or 'y': direction > 0 rejects r[axis] > offset, < 0 rejects
    r[axis] < offset, in place (:111-128)."""
    if axis not in ("x", "y") or direction == 0:
        if direction == 0:
            print("Direction must be <0 or >0")
        # the reference then indexes with an unassigned filter
        raise UnboundLocalError("knife_edge: axis must be 'x' or 'y' and direction <0 or >0")
    a = 0 if axis == "x" else 2
    return _apply(r, _optic(N.SLT_OPTIC_KNIFE_EDGE, offset, 1.0 if direction > 0 else -1.0, a),
                  in_place=True)


def clear_rays(self):
    """Drop a diagnostic's ray arrays (:130-138)."""
    self.r0 = None
    self.rf = None


def ray(x, θ, y, ϕ):
    """A 4x1 sympy Matrix ray (:140-145)."""
    from sympy import Matrix

    return Matrix([x, θ, y, ϕ])


def d2r(d):
    """Degrees to radians (:147-149)."""
    return d * np.pi / 180


def lens_cutoff(rf, *, L=400, R=25):
    """Keep rays whose extrapolation over L lands within R of the axis (:151-169)."""
    t = _device_rays(rf)
    xp = L * torch.tan(t[1]) + t[0]
    yp = L * torch.tan(t[3]) + t[2]
    keep = torch.pow(torch.pow(xp, 2) + torch.pow(yp, 2), 0.5) <= R
    return t[:, keep].cpu().numpy()


def transmission_map(H_wt, H_plain):
    """Per-pixel mean transmission H_wt / H_plain, NaN where H_plain == 0 (:541-565)."""
    wt = torch.as_tensor(np.asarray(H_wt, dtype=np.float64))
    pl = torch.as_tensor(np.asarray(H_plain, dtype=np.float64))
    dev = N.require_device()
    wt, pl = wt.to(dev), pl.to(dev)
    T = torch.where(pl > 0, wt / pl, torch.full_like(wt, float("nan")))
    return T.cpu().numpy()


def apply_ccd_mask(rf, weights=None, ccd_shape_m=(13.5e-3, 18e-3)):
    """NaN (geometry) and 0 (weights) outside the CCD footprint (:567-603).

    Returns (rf_out (4, N) float64, weights_out or None, inside mask).
    """
    t = _device_rays(rf).to(torch.float64).clone()
    half_x = ccd_shape_m[0] / 2.0
    half_y = ccd_shape_m[1] / 2.0
    inside = (t[0].abs() <= half_x) & (t[2].abs() <= half_y)
    t[:, ~inside] = float("nan")
    w = None
    if weights is not None:
        w = torch.as_tensor(np.array(weights, dtype=np.float64, copy=True)).to(t.device)
        w[~inside] = 0.0
        w = w.cpu().numpy()
    return t.cpu().numpy(), w, inside.cpu().numpy()


# ---------------------------------------------------------------------------
# k3: detector histogram
# ---------------------------------------------------------------------------

def histogram2d_device(x: torch.Tensor, y: torch.Tensor, xedges: np.ndarray, yedges: np.ndarray,
                       weights: torch.Tensor | None = None, mask: torch.Tensor | None = None,
                       order: torch.Tensor | None = None):
    """numpy.histogram2d(x, y, bins=(xedges, yedges), weights).T on the GPU.

    x, y: float32 or float64 device vectors of one dtype (any stride).
    order: optional int32 device permutation of the rays (the Morton order they
    were traced in, ``TraceResult.order``): visiting rays in that order lets a
    wavefront combine its atomics; counts are the same either way.
    Returns (nyb, nxb) tensors: counts as int64 and, when weights are given,
    fp64 weighted sums.  NaN positions are dropped, as values outside the edges.
    """
    if x.dtype != y.dtype or x.dtype not in (torch.float32, torch.float64):
        raise TypeError("x and y must both be float32 or both float64")
    dev = x.device
    nxb, nyb = len(xedges) - 1, len(yedges) - 1
    xe = const_device(np.asarray(xedges, dtype=np.float64), dev)
    ye = const_device(np.asarray(yedges, dtype=np.float64), dev)
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
    o = None
    if order is not None:
        if order.numel() != n:
            raise ValueError("order must hold one entry per ray")
        o = order.to(device=dev, dtype=torch.int32).contiguous()
    N.check(N.lib().slt_histogram2d(x.data_ptr(), x.stride(0), y.data_ptr(), y.stride(0),
                                    1 if x.dtype == torch.float64 else 0, n, N.ptr(o), N.ptr(m),
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


def _host(a):
    return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a


class Diagnostic:
    """Ray diagnostic base: holds a (4, N) Jones vector, images it through optics, bins it.

    ``rf`` may be a host array or a device tensor; after a solve the rays stay
    on the device until ``rf`` is read.  ``r0`` is the input in mm.
    """

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
            rf = rf.detach()
        else:
            rf = np.asarray(rf)
        self.Np_inc = self.Np
        if prefilter_input:
            if isinstance(rf, torch.Tensor):
                xp = self.L * torch.tan(rf[1]) + rf[0]
                yp = self.L * torch.tan(rf[3]) + rf[2]
                keep = torch.sqrt(xp ** 2 + yp ** 2) <= self.R
                rf = rf[:, keep]
                if weights is not None:
                    weights = torch.as_tensor(weights).to(keep.device)[keep]
            else:
                xp = self.L * np.tan(rf[1]) + rf[0]
                yp = self.L * np.tan(rf[3]) + rf[2]
                keep = np.sqrt(xp ** 2 + yp ** 2) <= self.R
                rf = rf[:, keep]
                if weights is not None:
                    weights = np.asarray(_host(weights))[keep]
            self.Np_inc = rf.shape[-1]
        if self.Np == self.Np_inc:
            print("\nAll rays retained at diagnostic input.")
        else:
            print("\n{} rays received, {} retained after optional prefilter.".format(
                self.Np, self.Np_inc))
            print(" --> {} % filtered before diagnostic solve.".format(
                round((1 - self.Np_inc / self.Np) * 100, 3)))
        self._in = rf              # input rays [m], host array or device tensor
        self._rf = rf              # current rays: the input until a solve replaces them
        self._r0 = None            # input in mm, made on first use
        if weights is None:
            self.weights = None
        elif isinstance(weights, torch.Tensor):
            self.weights = weights.detach().to(torch.float64)
        else:
            self.weights = np.asarray(weights, dtype=np.float64)

    # -- ray arrays: host views made on demand --------------------------------
    @property
    def rf(self):
        if isinstance(self._rf, torch.Tensor):
            self._rf = self._rf.cpu().numpy()
        return self._rf

    @rf.setter
    def rf(self, value):
        self._rf = value

    @property
    def r0(self):
        if self._r0 is None and self._in is not None:
            self._r0 = m_to_mm(_host(self._in))
        return self._r0

    @r0.setter
    def r0(self, value):
        self._r0 = value
        if value is None:
            self._in = None

    # -- optics ---------------------------------------------------------------
    def _solve(self, ops):
        """Run an element chain on r0 (the input in mm) in one k5 launch."""
        if self._in is None:
            raise ValueError("the diagnostic's rays were cleared")
        if self._r0 is not None:   # r0 was read or set: solve from it, as the reference does
            src, scale = self._r0, 1.0
        else:                      # r0 = m_to_mm(input), scaled inside the kernel
            src, scale = self._in, 1e3
        self._rf = rtm_device(src, ops, pos_scale=scale)

    @staticmethod
    def _travel(d):
        return _optic(N.SLT_OPTIC_TRAVEL, d)

    @staticmethod
    def _lens(f1, f2):
        return _optic(N.SLT_OPTIC_LENS, f1, f2)

    @staticmethod
    def _aperture(R):
        return _optic(N.SLT_OPTIC_CIRCULAR_APERTURE, R)

    @staticmethod
    def _stop(R):
        return _optic(N.SLT_OPTIC_CIRCULAR_STOP, R)

    @staticmethod
    def _rect(Lx, Ly):
        return _optic(N.SLT_OPTIC_RECT_APERTURE, Lx, Ly)

    # -- detector -------------------------------------------------------------
    def _matrix_device(self, plain_plot: bool) -> torch.Tensor:
        if plain_plot:
            if self._r0 is not None:
                return _device_rays(self._r0)
            if self._in is None:
                raise ValueError("the diagnostic's rays were cleared")
            return rtm_device(self._in, [], pos_scale=1e3)
        if self._rf is None:
            raise ValueError("the diagnostic's rays were cleared")
        return _device_rays(self._rf)

    def histogram(self, *, bin_scale=1, pix_x=3448, pix_y=2574, clear_mem=False,
                  plain_plot=False, extra_info=True):
        """Bin rows 0 and 2 into ``H`` (pix_y/bin_scale, pix_x/bin_scale) over ±Lx/2, ±Ly/2.

        src/processing/diagnostics.py:243-276 (NaN mask, optional CCD mask,
        numpy.histogram2d, transpose).
        """
        N.require_device()
        nxb = int(np.floor(pix_x / bin_scale))
        nyb = int(np.floor(pix_y / bin_scale))
        if nxb < 1 or nyb < 1:
            raise ValueError("`bins[0]` must be positive, when an integer")
        xlo, xhi = _outer_edges(-self.Lx / 2, self.Lx / 2)
        ylo, yhi = _outer_edges(-self.Ly / 2, self.Ly / 2)
        self.xedges = np.linspace(xlo, xhi, nxb + 1)
        self.yedges = np.linspace(ylo, yhi, nyb + 1)
        m = self._matrix_device(plain_plot)
        keep = None
        if self._ccd_half_x_mm is not None:
            keep = (m[0].abs() <= self._ccd_half_x_mm) & (m[2].abs() <= self._ccd_half_y_mm)
        w = None
        if self.weights is not None:
            w = self.weights if isinstance(self.weights, torch.Tensor) \
                else torch.from_numpy(self.weights)
        counts, wsum = histogram2d_device(m[0], m[2], self.xedges, self.yedges, weights=w,
                                          mask=keep)
        H = counts.to(torch.float64) if wsum is None else wsum
        self.H = H.cpu().numpy()
        if clear_mem:
            clear_rays(self)

    def plot(self, ax, clim=None, cmap=None):
        ax.imshow(self.H, interpolation="nearest", origin="lower", clim=clim, cmap=cmap,
                  extent=[self.xedges[0], self.xedges[-1], self.yedges[0], self.yedges[-1]])

    def plot_rays(self, *, bin_scale=1, pix_x=3448, pix_y=2574, clear_mem=False):
        self.histogram(bin_scale=bin_scale, pix_x=pix_x, pix_y=pix_y, clear_mem=clear_mem,
                       plain_plot=True)


class Shadowgraphy(Diagnostic):
    """Shadowgraphy: one lens (M ~ 2) or a two-lens M = 1 telescope (:284-315)."""

    def single_lens_custom_solve(self, f=200, obj_dist=400, img_dist=400):
        self._solve([self._travel(obj_dist - self.focal_plane), self._aperture(self.R),
                     self._lens(f, f), self._travel(img_dist)])

    def single_lens_solve(self):
        self._solve([self._travel(3 * self.L / 4 - self.focal_plane), self._aperture(self.R),
                     self._lens(self.L / 2, self.L / 2), self._travel(3 * self.L / 2)])

    def two_lens_solve(self):
        self._solve([self._travel(self.L - self.focal_plane), self._aperture(self.R),
                     self._lens(self.L / 2, self.L / 2), self._travel(self.L * 2),
                     self._aperture(self.R), self._lens(self.L / 2, self.L / 2),
                     self._travel(self.L)])


class Schlieren(Diagnostic):
    """Dark- and light-field schlieren: M = 1 telescope with a focal-plane stop (:317-368)."""

    def DF_solve(self, R=1):
        self._solve([self._travel(self.L - self.focal_plane), self._aperture(self.R),
                     self._lens(self.L, self.L), self._travel(self.L), self._stop(R),
                     self._travel(self.L), self._aperture(self.R), self._lens(self.L, self.L),
                     self._travel(self.L)])

    def LF_solve(self, R=1):
        self._solve([self._travel(self.L - self.focal_plane), self._aperture(self.R),
                     self._lens(self.L, self.L), self._travel(self.L), self._aperture(R),
                     self._travel(self.L), self._aperture(self.R), self._lens(self.L, self.L),
                     self._travel(self.L)])


class Refractometry(Diagnostic):
    """Imaging refractometer: spherical lens then a hybrid lens (:370-406)."""

    def incoherent_custom_solve(self, f1=200, f3=200, img_f1_dist=600, img_dist=400):
        self._solve([self._travel(2 * f1 - self.focal_plane), self._aperture(self.R),
                     self._lens(f1, f1), self._travel(img_f1_dist), self._rect(15, 30),
                     self._aperture(self.R), self._lens((2 * f3) / 3, f1),
                     self._travel(img_dist)])

    def incoherent_solve(self):
        self._solve([self._travel(3 * self.L / 4 - self.focal_plane), self._aperture(self.R),
                     self._lens(self.L / 2, self.L / 2), self._travel(3 * self.L / 2),
                     self._rect(15, 30), self._aperture(self.R),
                     self._lens(self.L / 3, self.L / 2), self._travel(self.L)])

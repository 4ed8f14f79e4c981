"""NumPy oracle for the ray-transfer-matrix optics ahead of the detector.

TEST INFRASTRUCTURE ONLY.  ``tests/`` import this module as the checker of
the k5 kernel (``slt_rtm``) and of ``sloth_amd.processing.diagnostics``; the
product package never imports it.

Restated from MAGPIE-ICL/SLOTH ``src/processing/diagnostics.py``:

* ``m_to_mm``                         :16-20
* ``lens`` / ``sym_lens`` / ``travel`` :28-65  (4x4 block matrices, ``np.matmul``)
* ``circular_aperture`` / ``circular_stop`` / ``rect_aperture`` /
  ``knife_edge``                       :67-128 (boolean filters, rejected rays set to NaN)
* the solve chains of ``Shadowgraphy`` (:290-315), ``Schlieren`` (:325-368) and
  ``Refractometry`` (:376-406), applied to ``r0 = m_to_mm(rf)``
* ``histogram``                        :243-276 (NaN and CCD masks, numpy.histogram2d, H.T)

Pinned in ``tests/test_optics_oracle.py`` by what the optics fix
independently of this restatement: a two-lens telescope at f = L/2 with
spacings L, 2L, L images the object plane with magnification +1; a single lens
at f = L/2 with object distance 3L/4 images it at 3L/2 with magnification -2;
a lens maps every ray parallel to the axis through its focus; aperture, stop,
rectangle and knife-edge boundaries on rays placed just inside and outside; and
``numpy.histogram2d`` itself for the detector.
"""
from __future__ import annotations

import numpy as np


def m_to_mm(r):
    rr = np.copy(r)
    rr[0::2, :] *= 1e3
    return rr


def lens(r, f1, f2):
    L = np.zeros((4, 4))
    L[:2, :2] = np.asarray([[1, 0], [-1 / f1, 1]])
    L[2:, 2:] = np.asarray([[1, 0], [-1 / f2, 1]])
    return np.matmul(L, r)


def travel(r, d):
    m = np.asarray([[1, d], [0, 1]])
    L = np.zeros((4, 4))
    L[:2, :2] = m
    L[2:, 2:] = m
    return np.matmul(L, r)


def circular_aperture(r, R):
    filt = r[0, :] ** 2 + r[2, :] ** 2 > R ** 2
    r[:, filt] = None
    return r


def circular_stop(r, R):
    filt = r[0, :] ** 2 + r[2, :] ** 2 < R ** 2
    r[:, filt] = None
    return r


def rect_aperture(r, Lx, Ly):
    filt = (r[0, :] ** 2 > Lx ** 2) * (r[2, :] ** 2 > Ly ** 2)
    r[:, filt] = None
    return r


def knife_edge(r, offset, axis, direction):
    a = {"x": 0, "y": 2}[axis]
    filt = r[a, :] > offset if direction > 0 else r[a, :] < offset
    r[:, filt] = None
    return r


# The diagnostics' chains as element lists: (name, args); ``solve`` runs one on r0.
def chain(name, *, L=400, R=25, focal_plane=0, **kw):
    fp = focal_plane
    if name == "single_lens_custom":
        f, od, idist = kw.get("f", 200), kw.get("obj_dist", 400), kw.get("img_dist", 400)
        return [("travel", od - fp), ("aperture", R), ("lens", f, f), ("travel", idist)]
    if name == "single_lens":
        return [("travel", 3 * L / 4 - fp), ("aperture", R), ("lens", L / 2, L / 2),
                ("travel", 3 * L / 2)]
    if name == "two_lens":
        return [("travel", L - fp), ("aperture", R), ("lens", L / 2, L / 2), ("travel", L * 2),
                ("aperture", R), ("lens", L / 2, L / 2), ("travel", L)]
    if name in ("DF", "LF"):
        Rs = kw.get("R_stop", 1)
        mid = ("stop", Rs) if name == "DF" else ("aperture", Rs)
        return [("travel", L - fp), ("aperture", R), ("lens", L, L), ("travel", L), mid,
                ("travel", L), ("aperture", R), ("lens", L, L), ("travel", L)]
    if name == "incoherent_custom":
        f1, f3 = kw.get("f1", 200), kw.get("f3", 200)
        i1, idist = kw.get("img_f1_dist", 600), kw.get("img_dist", 400)
        return [("travel", 2 * f1 - fp), ("aperture", R), ("lens", f1, f1), ("travel", i1),
                ("rect", 15, 30), ("aperture", R), ("lens", (2 * f3) / 3, f1),
                ("travel", idist)]
    if name == "incoherent":
        return [("travel", 3 * L / 4 - fp), ("aperture", R), ("lens", L / 2, L / 2),
                ("travel", 3 * L / 2), ("rect", 15, 30), ("aperture", R),
                ("lens", L / 3, L / 2), ("travel", L)]
    raise ValueError(name)


def apply(r0, elements):
    """Run an element list on a copy of r0 (mm), as the reference's solve methods do."""
    r = np.array(r0, copy=True)
    for e in elements:
        kind, args = e[0], e[1:]
        if kind == "travel":
            r = travel(r, *args)
        elif kind == "lens":
            r = lens(r, *args)
        elif kind == "aperture":
            r = circular_aperture(r, *args)
        elif kind == "stop":
            r = circular_stop(r, *args)
        elif kind == "rect":
            r = rect_aperture(r, *args)
        elif kind == "knife":
            r = knife_edge(r, *args)
        else:
            raise ValueError(kind)
    return r


def histogram(rf, *, Lx=18, Ly=13.5, pix_x=3448, pix_y=2574, bin_scale=1, weights=None,
              ccd_half_mm=None):
    """Diagnostic.histogram on a solved (mm) ray array: (H.T, xedges, yedges)."""
    mask = ~np.isnan(rf[0]) & ~np.isnan(rf[2])
    if ccd_half_mm is not None:
        mask &= (np.abs(rf[0]) <= ccd_half_mm[0]) & (np.abs(rf[2]) <= ccd_half_mm[1])
    w = None if weights is None else np.asarray(weights, dtype=np.float64)[mask]
    H, xe, ye = np.histogram2d(rf[0, mask], rf[2, mask],
                               bins=[np.floor(pix_x / bin_scale).astype(np.int64),
                                     np.floor(pix_y / bin_scale).astype(np.int64)],
                               range=[[-Lx / 2, Lx / 2], [-Ly / 2, Ly / 2]], weights=w)
    return H.T, xe, ye

"""Write tests/golden/fields_vectors.npz from the reference's own random-field generators.

Run in the build container, where MAGPIE-ICL/SLOTH is checked out at /root/reference:

    python tests/golden/make_fields_golden.py

``src/field_generator/gaussian3D.py`` of the reference needs only NumPy, so it is imported
from its file and run: ``cos`` (:19-151), ``fft`` (:153-212) and ``domain_fft`` (:215-273)
on small grids with the power spectrum k^(-11/3), NumPy's global generator seeded as
recorded (``domain_fft`` seeds it with 42 itself).  The vectors pin the spectral amplitude
of each construction (sqrt(k_func) in ``fft`` and ``domain_fft``, A_m in ``cos``), which
the oracle's property tests leave open.  Only inputs and outputs are stored (plain arrays,
no pickles); the reference's code is not copied.
"""
from __future__ import annotations

import contextlib
import importlib.util
import io
import os

import numpy as np

REF = "/root/reference/src/field_generator/gaussian3D.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fields_vectors.npz")

# the cases: every parameter the generators take, and the global seed set before the call
COS = dict(lx=1.0, ly=0.8, lz=1.2, nx=7, ny=6, nz=5, nmodes=24, wn1=2 * np.pi / 1.2, seed=11)
FFT = dict(N=4, seed=12)
DOMAIN = [dict(l_max=2.0, l_min=0.3, extent=1.0, res=8, factor=1.0),
          dict(l_max=2.0, l_min=0.2, extent=1.0, res=6, factor=1.5)]


def kolmogorov(k):
    return k ** (-11.0 / 3.0)


def main():
    spec = importlib.util.spec_from_file_location("ref_gaussian3D", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    out = {}
    quiet = contextlib.redirect_stdout(io.StringIO())
    with quiet, np.errstate(divide="ignore", invalid="ignore"):
        g = mod.gaussian3D(kolmogorov)
        np.random.seed(COS["seed"])
        out["cos_field"] = np.asarray(g.cos(*(COS[k] for k in ("lx", "ly", "lz", "nx", "ny",
                                                              "nz", "nmodes", "wn1"))))
        np.random.seed(FFT["seed"])
        out["fft_field"] = np.asarray(g.fft(FFT["N"]))
        for i, d in enumerate(DOMAIN):
            out[f"domain{i}_field"] = np.asarray(
                g.domain_fft(d["l_max"], d["l_min"], d["extent"], d["res"], d["factor"]))
            out[f"domain{i}_x"] = np.asarray(g.xc)
            out[f"domain{i}_z"] = np.asarray(g.zc)
    for k, v in COS.items():
        out[f"cos_{k}"] = np.asarray(v)
    for k, v in FFT.items():
        out[f"fft_{k}"] = np.asarray(v)
    for i, d in enumerate(DOMAIN):
        for k, v in d.items():
            out[f"domain{i}_{k}"] = np.asarray(v)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items() if v.ndim})


if __name__ == "__main__":
    main()

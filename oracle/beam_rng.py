"""NumPy oracle for the device beam (k6, ``slt_beam``).

TEST INFRASTRUCTURE ONLY: ``tests/`` import this module to check the kernel;
the product never imports it.

Philox4x32-10 is restated from Salmon, Moraes, Dror and Shaw, "Parallel random
numbers: as easy as 1, 2, 3" (SC 2011): ten rounds of two 32x32->64-bit
multiplies (0xD2511F53, 0xCD9E8D57) with the key bumped by the Weyl constants
(0x9E3779B9, 0xBB67AE85) between rounds.  Pinned in tests/test_beam_oracle.py
against the known-answer vectors published with the paper's Random123 library,
and the beam formulas against the reference's distributions
(src/core/beam.py:57-346) by their moments.
"""
from __future__ import annotations

import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    c = [np.asarray(v, dtype=np.uint64) & MASK for v in (c0, c1, c2, c3)]
    k0, k1 = int(k0) & 0xFFFFFFFF, int(k1) & 0xFFFFFFFF
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [hi1 ^ c[1] ^ np.uint64(k0), lo1, hi0 ^ c[3] ^ np.uint64(k1), lo0]
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    return [v.astype(np.uint32) for v in c]


def unit(x):
    return ((np.asarray(x, dtype=np.uint32) >> 8).astype(np.float64) + 0.5) / 16777216.0


TYPES = {"circular": 0, "square": 1, "rectangular": 2, "linear": 3}


def beam(n, beam_type, probing_axis, size_a, size_b, divergence, ne_extent, seed):
    """(6, n) float32 rays exactly as the kernel computes them (fp64 math)."""
    i = np.arange(n, dtype=np.uint64)
    lo, hi = i & MASK, i >> np.uint64(32)
    k0, k1 = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
    w0 = philox4x32_10(lo, hi, 0, 0, k0, k1)
    w1 = philox4x32_10(lo, hi, 1, 0, k0, k1)
    ty = TYPES[beam_type]
    g = np.sqrt(-2.0 * np.log(unit(w1[0]))) * np.cos(2.0 * np.pi * unit(w1[1]))
    chi = divergence * g
    b = size_a if ty == 1 else size_b
    if ty == 0:
        t = 2.0 * np.pi * unit(w0[0])
        r = size_a * np.sqrt(unit(w0[1]))
        phi = np.pi * unit(w0[2])
        p, q = r * np.cos(t), r * np.sin(t)
    elif ty in (1, 2):
        t = 2.0 * unit(w0[0]) - 1.0
        u = 2.0 * unit(w0[1]) - 1.0
        phi = np.pi * unit(w0[2])
        p, q = size_a * u, b * t
    else:
        p, q, phi = size_a * (2.0 * unit(w0[0]) - 1.0), np.zeros(n), np.zeros(n)
    c = 299792458.0
    vl, vt1, vt2 = c * np.cos(chi), c * np.sin(chi) * np.cos(phi), c * np.sin(chi) * np.sin(phi)
    z0 = np.full(n, -ne_extent)
    if ty == 3:
        s = [p, np.zeros(n), z0, c * np.sin(chi), np.zeros(n), c * np.cos(chi)]
    elif probing_axis == 0:
        s = [z0, p, q, vl, vt1, vt2]
    elif probing_axis == 2:
        s = [p, q, z0, vt1, vt2, vl]
    else:
        s = [p, z0, q, vt1, vl, vt2]
    return np.array(s, dtype=np.float64).astype(np.float32)

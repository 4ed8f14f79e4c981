"""Write tests/golden/trace_vectors.npz: small seeded cases of the traced path.

Inputs (density, fp32 coordinates, rays, save schedule, wavelength, Te/Z/CL) and the FP64
oracle's outputs (saved states, tau, path, in-box steps).  The reference's JAX path cannot
run offline, so these vectors come from the oracle; they pin it against drift (the oracle
itself is pinned by tests/test_oracle.py) and give the GPU path fixed cases to match.

Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

C = 299792458.0


def lin(n, L):
    return np.float32(np.linspace(-L / 2, L / 2, n))


def schedule(t, h_nom):
    dt = np.diff(t)
    nsub = np.maximum(np.ceil(dt / h_nom - 1e-9), 0).astype(np.int32)
    h = np.where(nsub > 0, dt / np.maximum(nsub, 1), 0.0)
    return nsub, h.astype(np.float32).astype(np.float64)


def case_sine():
    n, L = 16, 4e-3
    x = lin(n, L).astype(np.float64)
    ne = np.broadcast_to((1e24 * (1 + 0.1 * np.sin(2 * np.pi * x / 1e-3)))[:, None, None],
                         (n, n, n)).astype(np.float32)
    rng = np.random.RandomState(1)
    m = 64
    s0 = np.zeros((6, m), np.float32)
    s0[0] = rng.uniform(-1.5e-3, 1.5e-3, m)
    s0[1] = rng.uniform(-1.5e-3, 1.5e-3, m)
    s0[2] = np.float32(-L / 2)
    s0[5] = C
    t = np.array([0.0, 0.5, 1.0, np.sqrt(8.0)]) * L / C
    return dict(ne=ne, x=lin(n, L), y=lin(n, L), z=lin(n, L), s0=s0, t=t, lwl=1064e-9,
                inv_brems=0, Te=0.0, Z=0.0, cl=0.0)


def case_turbulent():
    n, L = 20, 5e-3
    rng = np.random.RandomState(2)
    g = rng.standard_normal((n, n, n))
    # smooth the noise a little with a 3-point average per axis
    for a in range(3):
        g = (np.roll(g, 1, a) + g + np.roll(g, -1, a)) / 3.0
    g /= g.std()
    ne = np.maximum(0.0, 2e24 * (1 + 0.3 * g)).astype(np.float32)
    m = 64
    s0 = np.zeros((6, m), np.float32)
    s0[0] = rng.uniform(-3e-3, 3e-3, m)
    s0[1] = rng.uniform(-3e-3, 3e-3, m)
    s0[2] = rng.uniform(-4e-3, -2.5e-3, m)   # some start outside the box
    s0[3] = rng.uniform(-3e6, 3e6, m)
    s0[4] = rng.uniform(-3e6, 3e6, m)
    s0[5] = np.sqrt(C ** 2 - s0[3].astype(np.float64) ** 2 - s0[4].astype(np.float64) ** 2)
    t = np.array([0.0, 0.7, 1.6]) * L / C
    return dict(ne=ne, x=lin(n, L), y=lin(n, L), z=lin(n, L), s0=s0, t=t, lwl=532e-9,
                inv_brems=1, Te=20.0, Z=1.0, cl=10.0)


def run(c):
    omega = 2 * np.pi * C / c["lwl"]
    coords = [c[k].astype(np.float64) for k in "xyz"]
    ne64 = c["ne"].astype(np.float64)
    grads = O.gradients(ne64, *coords, omega)
    kap = O.kappa(ne64, c["Te"], c["Z"], omega, coulomb_log=c["cl"] or None) if c["inv_brems"] else None
    h_nom = 0.5 * min(float(co[1] - co[0]) for co in coords) / C
    nsub, h = schedule(c["t"], h_nom)
    saves, steps = O.trace(grads, kap, coords, c["s0"].astype(np.float64), nsub, h)
    return nsub, h, saves, steps


def main():
    out = {}
    for name, c in (("sine", case_sine()), ("turb", case_turbulent())):
        nsub, h, saves, steps = run(c)
        for k, v in c.items():
            out[f"{name}_{k}"] = np.asarray(v)
        out[f"{name}_nsub"] = nsub
        out[f"{name}_h"] = h
        out[f"{name}_saves"] = saves
        out[f"{name}_steps"] = steps
    path = os.path.join(HERE, "trace_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()

"""The NumPy CPU-baseline runner (oracle/np_baseline.py) as bench.py starts it: a
child interpreter with a forked process pool.  On a small grid and ray sample it
must count exactly the in-box steps the oracle's own NumPy trace counts."""
import json
import os
import subprocess
import sys

import numpy as np

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_np_baseline_child_counts_the_oracle_steps(tmp_path):
    n, L = 24, 4e-3
    x = np.linspace(-L / 2, L / 2, n).astype(np.float32)
    rng = np.random.default_rng(5)
    ne = (1e24 * (1 + 0.1 * rng.standard_normal((n, n, n)))).astype(np.float32)
    omega = 2 * np.pi * O.C_LIGHT / 532e-9
    m = 40
    s0 = np.zeros((6, m))
    s0[0] = rng.uniform(-1.5e-3, 1.5e-3, m)
    s0[1] = rng.uniform(-1.5e-3, 1.5e-3, m)
    s0[2] = -L / 2
    s0[5] = O.C_LIGHT
    nsub = np.array([30, 50], dtype=np.int32)
    h = np.array([0.5 * (L / (n - 1)) / O.C_LIGHT] * 2, dtype=np.float32)
    path = str(tmp_path / "in.npz")
    np.savez(path, ne=ne, x=x, y=x, z=x, s0=s0, nsub=nsub, h=h, omega=np.float64(omega))
    r = subprocess.run([sys.executable, "-m", "oracle.np_baseline", path, "2", "60"], cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["workers"] == 2 and res["rays"] == m and res["value"] > 0
    # the same sample through the oracle directly: the steps the rate counts
    c64 = [x.astype(np.float64)] * 3
    g = O.gradients(ne.astype(np.float64), *c64, omega)
    k = O.kappa(ne.astype(np.float64), 20.0, 1.0, omega, coulomb_log=10.0)
    _, st = O.np_trace(g, k, c64, s0, nsub, h.astype(np.float64))
    assert st.sum() > 0
    assert abs(res["value"] * res["seconds"] - st.sum()) <= 1e-9 * st.sum()

"""Timing of the oracle's NumPy restatement (``np_trace``: RK4 vectorised over
rays, the closest analogue of the reference's ``jax.vmap`` over rays on a CPU),
split across the host's cores with a process pool.

Test and measurement infrastructure only: bench.py's CPU-baseline leg runs it
as a child interpreter (``python -m oracle.np_baseline IN.npz CORES SECONDS``),
so that the pool's workers are forked from a process that never opened the
GPU.  IN.npz holds ``ne`` (float32 grid), ``x``, ``y``, ``z`` (axes), ``s0``
(6 x N start states of a seeded ray sample), ``nsub``/``h`` (the step
schedule) and ``omega``.  Prints one JSON object: in-box ray-steps per second
over the timed sample, the worker count, the rays timed, the rays per worker
and the seconds.  Every ray runs every schedule step (no straight-line tail
outside the grid, as a vmapped solver steps every ray), and only in-box steps
count.

Reference functions the timed code restates (MAGPIE-ICL/SLOTH, file:line):
``dsdt`` src/core/propagator.py:168-243 under fixed-step RK4, with the
``RegularGridInterpolator`` of src/core/interpolator.py:15-106.
"""
from __future__ import annotations

import json
import multiprocessing as mp
import sys
import time

import numpy as np

from oracle import oracle as O

_D: dict = {}


def _work(job) -> int:
    idx, nsub, h = job
    d = _D
    _, steps = O.np_trace(d["grads"], d["kap"], d["c64"], d["s0"][:, idx], nsub, h)
    return int(steps.sum())


def _timed(pool, n: int, workers: int, nsub, h) -> tuple[int, float]:
    jobs = [(idx, nsub, h) for idx in np.array_split(np.arange(n), workers)]
    t0 = time.perf_counter()
    steps = sum(pool.map(_work, jobs, chunksize=1))
    return steps, time.perf_counter() - t0


def main(path: str, workers: int, target_s: float, width: int = 1024) -> None:
    z = np.load(path)
    c64 = [z[k].astype(np.float64) for k in ("x", "y", "z")]
    ne64 = z["ne"].astype(np.float64)
    omega = float(z["omega"])
    nsub = z["nsub"].astype(np.int64)
    h = z["h"].astype(np.float32).astype(np.float64)
    _D.update(grads=O.gradients(ne64, *c64, omega),
              kap=O.kappa(ne64, 20.0, 1.0, omega, coulomb_log=10.0),
              c64=c64, s0=z["s0"].astype(np.float64))
    del ne64
    total = _D["s0"].shape[1]
    with mp.get_context("fork").Pool(workers) as pool:
        # a vector of `width` rays per worker (NumPy's per-call overhead dominates much
        # narrower ones); the first 16 schedule steps estimate the full run's time, and
        # the width shrinks only if that exceeds target_s
        n = min(total, width * workers)
        k = int(min(16, nsub[0]))
        _, dt = _timed(pool, n, workers, np.array([k]), h[:1])
        est = dt / k * float(nsub.sum())
        if est > target_s:
            n = min(n, max(128 * workers, int(n * target_s / est) // workers * workers))
        steps, dt = _timed(pool, n, workers, nsub, h)
    print(json.dumps({"value": steps / dt, "rays": n, "seconds": dt, "workers": workers,
                      "width": n // workers}), flush=True)


if __name__ == "__main__":
    main(sys.argv[1], int(sys.argv[2]), float(sys.argv[3]))

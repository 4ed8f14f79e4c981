"""Golden vectors (tests/golden/trace_vectors.npz, written by tests/golden/make_golden.py):
the oracle must reproduce them (CPU); the GPU path must match them at 1e-5 (gpu)."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "trace_vectors.npz")
C = 299792458.0


def _case(name):
    z = np.load(GOLD)  # allow_pickle stays False: plain arrays only
    return {k[len(name) + 1:]: z[k] for k in z.files if k.startswith(name + "_")}


@pytest.mark.parametrize("name", ["sine", "turb"])
def test_oracle_reproduces_golden(orc, name):
    c = _case(name)
    omega = 2 * np.pi * C / float(c["lwl"])
    coords = [c[k].astype(np.float64) for k in "xyz"]
    ne64 = c["ne"].astype(np.float64)
    grads = orc.gradients(ne64, *coords, omega)
    kap = None
    if int(c["inv_brems"]):
        kap = orc.kappa(ne64, float(c["Te"]), float(c["Z"]), omega,
                        coulomb_log=float(c["cl"]) or None)
    saves, steps = orc.trace(grads, kap, coords, c["s0"].astype(np.float64), c["nsub"], c["h"])
    np.testing.assert_allclose(saves, c["saves"], rtol=1e-12, atol=1e-25)
    np.testing.assert_array_equal(steps, c["steps"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["sine", "turb"])
def test_gpu_matches_golden(orc, name):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from _parity import assert_parity, gpu_trace

    c = _case(name)
    inv = bool(int(c["inv_brems"]))
    gpu = gpu_trace(c["ne"], (c["x"], c["y"], c["z"]), c["s0"], c["t"], lwl=float(c["lwl"]),
                    inv_brems=inv, Te=float(c["Te"]) if inv else None,
                    Z=float(c["Z"]) if inv else None, coulomb_log=float(c["cl"]) or None)
    np.testing.assert_array_equal(gpu["nsub"], c["nsub"])
    # the face gaps of the golden trajectories, for the energy exception
    omega = 2 * np.pi * C / float(c["lwl"])
    c64 = [c[k].astype(np.float64) for k in "xyz"]
    ne64 = c["ne"].astype(np.float64)
    kap = (orc.kappa(ne64, float(c["Te"]), float(c["Z"]), omega,
                     coulomb_log=float(c["cl"]) or None) if inv else None)
    _, _, gap = orc.trace(orc.gradients(ne64, *c64, omega), kap, c64,
                          c["s0"].astype(np.float64), c["nsub"], c["h"], face_gap=True)
    ref = {"saves": c["saves"], "steps": c["steps"], "face_gap": gap}
    print(name, assert_parity(gpu, ref, (c["x"], c["y"], c["z"]), step_mismatch_frac=0.02))

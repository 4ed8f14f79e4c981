"""k6 device beam against the NumPy Philox oracle."""
import numpy as np
import pytest

from oracle import beam_rng as B

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


@pytest.mark.parametrize("bt,axis,sa,sb", [("square", 2, 3e-3, 3e-3), ("rectangular", 0, 1e-3, 4e-3),
                                           ("linear", 2, 2e-3, 0.0), ("circular", 1, 2e-3, 0.0)])
def test_device_beam_matches_oracle(bt, axis, sa, sb):
    from sloth_amd.core.beam import device_beam

    n, seed, div = 100003, 1234567890123, 2e-3
    pd = "xyz"[axis]
    size = (sa, sb) if bt == "rectangular" else sa
    got = device_beam(n, size, div, 5e-3, probing_direction=pd, beam_type=bt, seed=seed)
    got = got.cpu().numpy()
    ref = B.beam(n, bt, axis, sa, sb, div, 5e-3, seed)
    if bt == "circular":
        # positions go through cos/sin: one float32 ulp at most
        np.testing.assert_allclose(got, ref, rtol=2e-7, atol=1e-12)
    else:
        # positions: integer draw -> exact fp64 arithmetic -> one rounding to fp32
        pos = [0, 1, 2]
        np.testing.assert_array_equal(got[pos], ref[pos])
        np.testing.assert_allclose(got[3:], ref[3:], rtol=2e-7, atol=1e-3)


def test_beam_class_device_rng_and_prefix_stability():
    import torch

    from sloth_amd.core.beam import Beam, device_beam

    b = Beam(5000, 1e-3, 1e-4, 4e-3, beam_type="circular", seeded=True, device_rng=True)
    assert isinstance(b.s0, torch.Tensor) and b.s0.is_cuda and tuple(b.s0.shape) == (6, 5000)
    a = device_beam(2000, 1e-3, 1e-4, 4e-3, seed=42)
    np.testing.assert_array_equal(b.s0[:, :2000].cpu().numpy(), a.cpu().numpy())

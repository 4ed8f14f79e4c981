"""Pins of the device-beam oracle (oracle/beam_rng.py) and host checks of the beam API."""
import numpy as np
import pytest

from oracle import beam_rng as B


def test_philox4x32_10_known_answers():
    """The known-answer vectors published with Random123 (philox4x32_10)."""
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for c, k, want in kat:
        got = B.philox4x32_10(*c, *k)
        assert [int(v) for v in got] == list(want)


def test_beam_distributions():
    n, c = 200000, 299792458.0
    s = B.beam(n, "circular", 2, 2e-3, 0.0, 1e-3, 5e-3, 7).astype(np.float64)
    r2 = s[0] ** 2 + s[1] ** 2
    assert r2.max() <= (2e-3) ** 2 * (1 + 1e-6)
    assert abs(r2.mean() / (2e-3) ** 2 - 0.5) < 0.01          # uniform on the disc
    assert abs(np.arctan2(s[1], s[0]).mean()) < 0.02
    chi = np.arccos(np.clip(s[5] / c, -1, 1))
    assert abs(np.sqrt(np.mean(chi ** 2)) / 1e-3 - 1) < 0.02      # |chi| of N(0, div)
    np.testing.assert_allclose(np.sqrt((s[3:] ** 2).sum(0)), c, rtol=1e-6)
    assert np.all(s[2] == np.float32(-5e-3))
    q = B.beam(n, "square", 0, 3e-3, 0.0, 0.0, 4e-3, 9).astype(np.float64)
    assert np.abs(q[1:3]).max() <= 3e-3 and abs(np.mean(q[1] ** 2) / (3e-3) ** 2 - 1 / 3) < 0.01
    assert np.all(q[0] == np.float32(-4e-3)) and np.all(q[3] == np.float32(c))
    rect = B.beam(n, "rectangular", 1, 1e-3, 4e-3, 0.0, 4e-3, 9).astype(np.float64)
    assert np.abs(rect[0]).max() <= 1e-3 and np.abs(rect[2]).max() <= 4e-3
    assert np.abs(rect[2]).max() > 3.9e-3 and np.all(rect[1] == np.float32(-4e-3))
    lin = B.beam(n, "linear", 2, 1e-3, 0.0, 0.0, 4e-3, 9)
    assert np.all(lin[1] == 0) and np.all(lin[4] == 0)
    # independent streams for different seeds, prefix-stable in n
    assert not np.array_equal(B.beam(100, "square", 2, 1.0, 0.0, 0.0, 1.0, 1),
                              B.beam(100, "square", 2, 1.0, 0.0, 0.0, 1.0, 2))
    np.testing.assert_array_equal(B.beam(1000, "circular", 2, 1.0, 0.0, 0.1, 1.0, 3)[:, :100],
                                  B.beam(100, "circular", 2, 1.0, 0.0, 0.1, 1.0, 3))


def test_device_beam_argument_errors_and_abi():
    from sloth_amd import _native as N
    from sloth_amd.core.beam import device_beam

    with pytest.raises(TypeError):           # len() of a float, as the reference's check
        device_beam(10, 1e-3, 0.0, 1e-3, beam_type="rectangular")
    with pytest.raises(AssertionError):
        device_beam(10, (1e-3, 2e-3, 3e-3), 0.0, 1e-3, beam_type="rectangular")
    with pytest.raises(AssertionError):
        device_beam(10, (1e-3, 2e-3), 0.0, 1e-3, beam_type="square")
    with pytest.raises(ValueError):
        device_beam(10, 1e-3, 0.0, 1e-3, beam_type="hexagonal")
    L = N.lib()
    assert L.slt_beam(-1, 0, 2, 1.0, 1.0, 0.0, 1.0, 0, None, None) == N.SLT_ERR_INVALID
    assert L.slt_beam(10, 4, 2, 1.0, 1.0, 0.0, 1.0, 0, None, None) == N.SLT_ERR_INVALID
    assert L.slt_beam(0, 0, 2, 1.0, 1.0, 0.0, 1.0, 0, None, None) == N.SLT_OK

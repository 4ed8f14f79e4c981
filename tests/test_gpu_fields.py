"""Random-field generators on the GPU against the NumPy oracle (same seeds, same draws)."""
import numpy as np
import pytest

from oracle import fields as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _k53(k):
    return k ** (-5 / 3)


def _close(a, b, rtol=1e-12):
    scale = np.abs(b).max()
    assert np.abs(a - b).max() <= rtol * scale, np.abs(a - b).max() / scale


@pytest.mark.parametrize("factor", [1.0, 1.5])
def test_domain_fft_matches_oracle(factor):
    from sloth_amd.field_generator.gaussian3D import gaussian3D

    g = gaussian3D(_k53)
    f = g.domain_fft(2.0, 0.25, 1.0, 16, factor)
    ref, _, _ = F.domain_fft(_k53, 2.0, 0.25, 1.0, 16, factor)
    assert f.shape == ref.shape == (32, 32, int(32 * factor))
    _close(f, ref)
    assert np.abs(f).max() == 1.0
    np.testing.assert_array_equal(g.xc, np.linspace(-1.0, 1.0, 32, endpoint=False,
                                                    dtype=np.float32))


def test_fft_and_cos_match_oracle():
    from sloth_amd.field_generator.gaussian3D import gaussian3D

    g = gaussian3D(_k53)
    np.random.seed(5)
    a = g.fft(6)
    np.random.seed(5)
    b = F.fft(_k53, 6).real
    _close(a, b)
    np.random.seed(9)
    c = g.cos(1.0, 2.0, 1.5, 12, 10, 8, 40, 2 * np.pi)
    np.random.seed(9)
    d, _ = F.cos(_k53, 1.0, 2.0, 1.5, 12, 10, 8, 40, 2 * np.pi)
    assert c.shape == (12, 10, 8)
    _close(c, d, rtol=1e-11)


def test_generators_match_the_reference_fields():
    """The GPU generators against the reference's own outputs (same arguments, same seeds),
    at fp64 FFT and cosine rounding."""
    from sloth_amd.field_generator.gaussian3D import gaussian3D

    from test_fields_oracle import _k113, golden_fields

    z, cos_args, dom = golden_fields()
    g = gaussian3D(_k113)
    np.random.seed(int(z["cos_seed"]))
    _close(g.cos(*cos_args), z["cos_field"], rtol=1e-11)
    np.random.seed(int(z["fft_seed"]))
    with np.errstate(divide="ignore", invalid="ignore"):
        _close(g.fft(int(z["fft_N"])), z["fft_field"])
    for i, args in enumerate(dom):
        f = g.domain_fft(*args)
        assert f.shape == z[f"domain{i}_field"].shape
        _close(f, z[f"domain{i}_field"])
        np.testing.assert_array_equal(g.xc, z[f"domain{i}_x"])
        np.testing.assert_array_equal(g.zc, z[f"domain{i}_z"])


def test_turbulent_box_spherical_and_step_domains():
    import torch

    from sloth_amd.shared import domains as D

    ne = D.turbulent_box(1e24, 0.3e24, 2e-3, 0.2e-3, 1e-3, 32)
    ref, _, _ = F.domain_fft(lambda k: k ** (-5 / 3), 2e-3, 0.2e-3, 1e-3, 16, 1.0)
    _close(ne, 1e24 + ref * 0.3e24)
    L = [2.0, 2.0, 2.0]
    ne_s, rho = D.spherical((0.5, 1.0), (3.0, 2.0), (7.0, 9.0), L, 21)
    XX, YY, ZZ = D.grid(L, 21)
    r = torch.sqrt(XX ** 2 + YY ** 2 + ZZ ** 2)
    assert torch.all(ne_s[(r > 0) & (r <= 0.5)] == 3.0)
    assert torch.all(ne_s[(r > 0.5) & (r <= 1.0)] == 2.0)
    assert torch.all(ne_s[r > 1.0] == 0.0) and torch.all(rho[1][(r > 0.5) & (r <= 1.0)] == 9.0)
    step = D.density_step(1.0, 4.0, L, [5, 6, 7])
    assert step.shape == (5, 6, 7)
    assert torch.all(step[:, YY.new_tensor(np.linspace(-1, 1, 6)) > 0, :] == 1.0)
    assert torch.all(step[:, :3, :] == 4.0)

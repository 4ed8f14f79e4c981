"""Pins of the random-field oracle (oracle/fields.py): the reference generators' own outputs
(tests/golden/fields_vectors.npz) and properties the constructions fix; host tests of the
VTK image files."""
import base64
import os
import zlib

import numpy as np
import pytest

from oracle import fields as F


def _k53(k):
    return k ** (-5 / 3)


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fields_vectors.npz")


def _k113(k):
    return k ** (-11.0 / 3.0)


def golden_fields():
    """The reference's own fields (tests/golden/make_fields_golden.py ran
    src/field_generator/gaussian3D.py) with the arguments and seeds that made them."""
    z = np.load(GOLD)  # plain arrays; allow_pickle stays False
    cos_args = ([float(z[f"cos_{a}"]) for a in ("lx", "ly", "lz")]
                + [int(z[f"cos_{a}"]) for a in ("nx", "ny", "nz", "nmodes")]
                + [float(z["cos_wn1"])])
    dom = [([float(z[f"domain{i}_{a}"]) for a in ("l_max", "l_min", "extent")]
            + [int(z[f"domain{i}_res"]), float(z[f"domain{i}_factor"])]) for i in (0, 1)]
    return z, cos_args, dom


def test_oracle_reproduces_the_reference_fields():
    """The oracle's cos / fft / domain_fft equal the reference generators' outputs on the
    same arguments and seeds: the spectral amplitudes (A_m of the mode sum, sqrt of the
    spectrum in both FFT constructions), the draw order and the normalisation are the
    reference's."""
    z, cos_args, dom = golden_fields()
    np.random.seed(int(z["cos_seed"]))
    r, _ = F.cos(_k113, *cos_args)
    np.testing.assert_allclose(r, z["cos_field"], rtol=1e-12, atol=1e-12 * np.abs(r).max())
    np.random.seed(int(z["fft_seed"]))
    s = F.fft(_k113, int(z["fft_N"]))
    np.testing.assert_allclose(s.real, z["fft_field"], rtol=1e-12,
                               atol=1e-12 * np.abs(s.real).max())
    for i, args in enumerate(dom):
        f, _, _ = F.domain_fft(_k113, *args)
        np.testing.assert_allclose(f, z[f"domain{i}_field"], rtol=1e-12, atol=1e-13)
    # the vectors do constrain the amplitude: the field of the spectrum itself instead of
    # its square root (a plausible slip) is far from the reference's
    k, band = F.domain_fft(_k113, *dom[0])[1:]
    S = np.zeros_like(k)
    S[band] = _k113(k[band])
    np.random.seed(42)
    noise = np.random.normal(0, 1, k.shape) + 1j * np.random.normal(0, 1, k.shape)
    wrong = np.fft.ifftn(noise * S).real
    wrong /= np.abs(wrong).max()
    assert np.abs(wrong - z["domain0_field"]).max() > 0.1


def test_domain_fft_is_band_limited_normalised_and_seeded():
    l_max, l_min, extent, res = 2.0, 0.25, 1.0, 16
    f1, k, band = F.domain_fft(_k53, l_max, l_min, extent, res, 1.0)
    f2, _, _ = F.domain_fft(_k53, l_max, l_min, extent, res, 1.0)
    np.testing.assert_array_equal(f1, f2)                     # seed 42 every call
    assert f1.shape == (32, 32, 32) and f1.dtype == np.float64
    assert np.abs(f1).max() == 1.0
    # the real part of a band-limited complex field stays in the band's symmetric hull
    spec = np.abs(np.fft.fftn(f1))
    outside = ~(band | np.flip(np.roll(band, -1, axis=(0, 1, 2)), axis=(0, 1, 2)))
    assert spec[outside].max() < 1e-12 * spec.max()
    assert spec[band].max() > 0


def test_fft_field_is_real_with_zero_mean():
    np.random.seed(7)
    s = F.fft(_k53, 6)
    assert s.shape == (13, 13, 13)
    assert np.abs(s.imag).max() < 1e-12 * np.abs(s.real).max()
    assert abs(s.real.mean()) < 1e-12 * np.abs(s.real).max()


def test_cos_modes_are_plane_waves_of_wavenumber_wn():
    """Only the first of 64 modes has power: four cosines of wavevectors of norm wn[0],
    so lap r = -wn[0]^2 r."""
    np.random.seed(3)
    n, L = 48, 1.0
    r, wn = F.cos(lambda k: np.where(k == k.min(), 1.0, 0.0), L, L, L, n, n, n, 64,
                  2 * np.pi / L)
    h = L / n
    lap = (r[2:, 1:-1, 1:-1] + r[:-2, 1:-1, 1:-1] + r[1:-1, 2:, 1:-1] + r[1:-1, :-2, 1:-1] +
           r[1:-1, 1:-1, 2:] + r[1:-1, 1:-1, :-2] - 6 * r[1:-1, 1:-1, 1:-1]) / h ** 2
    inner = r[1:-1, 1:-1, 1:-1]
    # second-order finite differences: relative error ~ (wn h)^2 / 12
    err = np.abs(lap + wn[0] ** 2 * inner).max() / (wn[0] ** 2 * np.abs(inner).max())
    assert err < (wn[0] * h) ** 2 / 6


# ---------------------------------------------------------------------------
# .vti / .pvti
# ---------------------------------------------------------------------------

def _pvti(path, shape, spacing, source):
    nx, ny, nz = shape
    with open(path, "w") as fh:
        fh.write(f'<?xml version="1.0"?><VTKFile type="PImageData" version="0.1" '
                 f'byte_order="LittleEndian"><PImageData WholeExtent="0 {nx} 0 {ny} 0 {nz}" '
                 f'Origin="0 0 0" Spacing="{spacing[0]} {spacing[1]} {spacing[2]}">'
                 f'<PCellData><PDataArray type="Float64" Name="rnec"/></PCellData>'
                 f'<Piece Extent="0 {nx} 0 {ny} 0 {nz}" Source="{source}"/></PImageData></VTKFile>')


def _vti(path, shape, body, fmt, header_type=None, compressor=None, appended=None):
    nx, ny, nz = shape
    attrs = f' header_type="{header_type}"' if header_type else ""
    attrs += f' compressor="{compressor}"' if compressor else ""
    off = ' offset="0"' if fmt == "appended" else ""
    xml = (f'<?xml version="1.0"?><VTKFile type="ImageData" version="0.1" byte_order="LittleEndian"'
           f'{attrs}><ImageData WholeExtent="0 {nx} 0 {ny} 0 {nz}" Origin="0 0 0" Spacing="1 1 1">'
           f'<Piece Extent="0 {nx} 0 {ny} 0 {nz}"><CellData Scalars="rnec">'
           f'<DataArray type="Float64" Name="rnec" format="{fmt}"{off}>{body}</DataArray>'
           f'</CellData></Piece></ImageData>').encode()
    if appended is not None:
        xml += b'<AppendedData encoding="raw">_' + appended + b'</AppendedData>'
    with open(path, "wb") as fh:
        fh.write(xml + b"</VTKFile>")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_export_pvti_round_trip(tmp_path, dtype, capsys):
    from sloth_amd.utils.handle_filetypes import export_pvti, pvti_readin

    a = np.random.default_rng(0).normal(size=(5, 4, 3)).astype(dtype)
    export_pvti(a, str(tmp_path / "ne"), extent_x=2.5, extent_y=2.0, extent_z=1.5)
    img, shape, spacing = pvti_readin(str(tmp_path / "ne.pvti"))
    assert shape == (5, 4, 3) and img.dtype == dtype
    np.testing.assert_array_equal(img, a)
    np.testing.assert_allclose(spacing, [2.5 / 2, 2.0 / 2, 1.5 / 1])
    assert "wrote " in capsys.readouterr().out


def test_pvti_readin_documented_encodings(tmp_path):
    """ascii, inline base64 (UInt32 header, raw and zlib blocks) and appended raw arrays."""
    from sloth_amd.utils.handle_filetypes import pvti_readin

    shape = (3, 2, 2)
    a = np.arange(12, dtype=np.float64).reshape(shape, order="F") * 0.5 - 1.0
    flat = a.flatten(order="F")
    raw = flat.tobytes()
    cases = {
        "ascii": dict(fmt="ascii", body=" ".join(repr(float(v)) for v in flat)),
        "b64": dict(fmt="binary", header_type="UInt32",
                    body=(base64.b64encode(np.uint32(len(raw)).tobytes()) +
                          base64.b64encode(raw)).decode()),
    }
    comp = zlib.compress(raw)
    hdr = np.array([1, len(raw), len(raw), len(comp)], dtype=np.uint32).tobytes()
    cases["zlib"] = dict(fmt="binary", header_type="UInt32", compressor="vtkZLibDataCompressor",
                         body=(base64.b64encode(hdr) + base64.b64encode(comp)).decode())
    cases["appended"] = dict(fmt="appended", header_type="UInt64", body="",
                             appended=np.uint64(len(raw)).tobytes() + raw)
    for name, kw in cases.items():
        _vti(tmp_path / f"{name}.vti", shape, **kw)
        _pvti(tmp_path / f"{name}.pvti", shape, (0.1, 0.2, 0.3), f"{name}.vti")
        img, sh, sp = pvti_readin(str(tmp_path / f"{name}.pvti"))
        np.testing.assert_array_equal(img, a, err_msg=name)
        np.testing.assert_allclose(sp, [0.1, 0.2, 0.3])


def test_gaussian3d_export_scalar_field_round_trip(tmp_path):
    from sloth_amd.field_generator.gaussian3D import gaussian3D
    from sloth_amd.utils.handle_filetypes import pvti_readin

    g = gaussian3D(_k53)
    g.ne = np.random.default_rng(1).normal(size=(8, 8, 6))
    g.export_scalar_field(fname=str(tmp_path / "field"))
    img, shape, spacing = pvti_readin(str(tmp_path / "field.pvti"))
    np.testing.assert_array_equal(img, g.ne)
    # no coordinates: xc = arange(-4, 4); spacing = max(xc) / ((n - 1) // 2)
    np.testing.assert_allclose(spacing, [3 / 3, 3 / 3, 3 / 2])
    assert os.path.exists(tmp_path / "field.vti")

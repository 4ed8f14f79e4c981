"""Jones epilogue and detector histogram (k3) against numpy / the oracle."""
import numpy as np
import pytest

from _parity import C, gpu_trace, oracle_trace

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _edge_heavy_sample(n, lo, hi, nb, rng, dtype):
    e = np.linspace(lo, hi, nb + 1)
    v = rng.uniform(lo - 0.1 * (hi - lo), hi + 0.1 * (hi - lo), n)
    k = n // 5
    v[:k] = e[rng.integers(0, nb + 1, k)]          # exactly on edges, incl. both ends
    v[k:k + 10] = hi
    v[k + 10:k + 20] = np.nan
    return v.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_histogram_bit_exact_vs_numpy(dtype):
    import torch

    from sloth_amd.processing.diagnostics import histogram2d_device

    rng = np.random.default_rng(20)
    n = 200_000
    x = _edge_heavy_sample(n, -9.0, 9.0, 256, rng, dtype)
    y = _edge_heavy_sample(n, -6.75, 6.75, 192, rng, dtype)
    xe, ye = np.linspace(-9.0, 9.0, 257), np.linspace(-6.75, 6.75, 193)
    w = rng.uniform(0, 1, n)
    counts, wsum = histogram2d_device(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), xe,
                                      ye, weights=torch.from_numpy(w))
    keep = ~np.isnan(x) & ~np.isnan(y)
    H, _, _ = np.histogram2d(x[keep], y[keep], bins=[256, 192], range=[[-9, 9], [-6.75, 6.75]])
    Hw, _, _ = np.histogram2d(x[keep], y[keep], bins=[256, 192], range=[[-9, 9], [-6.75, 6.75]],
                              weights=w[keep])
    np.testing.assert_array_equal(counts.cpu().numpy(), H.T.astype(np.int64))
    np.testing.assert_allclose(wsum.cpu().numpy(), Hw.T, rtol=1e-12, atol=1e-12)


def test_histogram_spatially_sorted_input_combines_and_stays_exact():
    """Rays in a spatially coherent order (sorted by cell, as after a Morton-ordered
    trace) and no order array: the wavefront combine path, against numpy, also with a
    mask and with more distinct cells per wavefront than are combined."""
    import torch

    from sloth_amd.processing.diagnostics import histogram2d_device

    rng = np.random.default_rng(31)
    n = 300_000
    xe, ye = np.linspace(-1.0, 1.0, 65), np.linspace(-1.0, 1.0, 49)
    x = rng.uniform(-1.05, 1.05, n)
    y = rng.uniform(-1.05, 1.05, n)
    w = rng.uniform(0, 1, n)
    for label, key in (("sorted", np.lexsort((x, np.floor((y + 1) * 24)))),
                       ("runs of 8", np.argsort(np.arange(n) % (n // 8), kind="stable"))):
        xs, ys, ws = x[key], y[key], w[key]
        mask = (np.arange(n) % 7) != 3
        counts, wsum = histogram2d_device(torch.from_numpy(xs).cuda(),
                                          torch.from_numpy(ys).cuda(), xe, ye,
                                          weights=torch.from_numpy(ws),
                                          mask=torch.from_numpy(mask))
        H, _, _ = np.histogram2d(xs[mask], ys[mask], bins=[xe, ye])
        Hw, _, _ = np.histogram2d(xs[mask], ys[mask], bins=[xe, ye], weights=ws[mask])
        np.testing.assert_array_equal(counts.cpu().numpy(), H.T.astype(np.int64), err_msg=label)
        np.testing.assert_allclose(wsum.cpu().numpy(), Hw.T, rtol=1e-12, atol=1e-12,
                                   err_msg=label)


def test_diagnostic_histogram_matches_numpy_and_weights_contract():
    from sloth_amd.processing.diagnostics import Diagnostic

    rng = np.random.RandomState(42)
    rf = np.zeros((4, 500))
    rf[0] = (rng.rand(500) - 0.5) * 40
    rf[2] = (rng.rand(500) - 0.5) * 40
    rf[1] = rng.randn(500) * 1e-4
    rf[3] = rng.randn(500) * 1e-4
    d = Diagnostic(rf, L=100, R=500, Lx=50, Ly=50)
    d.histogram(pix_x=20, pix_y=30)
    H, xe, ye = np.histogram2d(rf[0], rf[2], bins=[20, 30], range=[[-25, 25], [-25, 25]])
    np.testing.assert_array_equal(d.H, H.T)
    np.testing.assert_array_equal(d.xedges, xe)
    ones = Diagnostic(rf, weights=np.ones(500), L=100, R=500, Lx=50, Ly=50)
    ones.histogram(pix_x=20, pix_y=30)
    np.testing.assert_array_equal(ones.H, d.H)            # unit weights == counts
    low = Diagnostic(rf, weights=np.full(500, 0.1), L=100, R=500, Lx=50, Ly=50)
    low.histogram(pix_x=20, pix_y=30)
    assert abs(low.H.sum() / d.H.sum() - 0.1) < 1e-12
    # CCD mask on post-RTM (mm) coordinates reduces the count
    ccd = Diagnostic(rf, ccd_shape_m=(13.5e-3, 18e-3), L=100, R=500, Lx=50, Ly=50)
    ccd.rf = rf.copy()
    ccd.histogram(pix_x=20, pix_y=20)
    inside = (np.abs(rf[0]) <= 6.75) & (np.abs(rf[2]) <= 9.0)
    H2, _, _ = np.histogram2d(rf[0][inside], rf[2][inside], bins=[20, 20], range=[[-25, 25]] * 2)
    np.testing.assert_array_equal(ccd.H, H2.T)


def test_traced_detector_image_matches_oracle_except_edge_rays(orc):
    """256x256 detector image of traced rays: counts equal the oracle's except for rays whose
    oracle position lies within the parity tolerance of a bin edge."""
    import torch

    from sloth_amd.core.beam import Beam
    from sloth_amd.fields import sine_field
    from sloth_amd.processing.diagnostics import histogram2d_device
    from sloth_amd.shared.propagation import jones_device

    L = 10e-3
    ne, coords = sine_field(64, L)
    s0 = Beam(20000, 2e-3, 0.0, 5e-3, beam_type="circular", seeded=True).s0
    # solve() horizon; a horizon of exactly L/c would put every collimated ray's last
    # stage point on the exit face, where rounding decides inside/outside
    t = np.array([0.0, np.sqrt(8.0) * L / C])
    gpu = gpu_trace(ne, coords, s0, t, lwl=1064e-9)
    ref = oracle_trace(orc, ne, coords, s0, gpu["nsub"], gpu["h"], lwl=1064e-9)
    st = torch.from_numpy(gpu["states"][-1].astype(np.float32)).cuda()
    rf, _ = jones_device(st, "z")
    rf_ref = orc.jones(ref["saves"][-1][:6], "z")
    # angles [rad]: the 1e-5 direction bar
    np.testing.assert_allclose(rf.cpu().numpy()[[1, 3]], rf_ref[[1, 3]], rtol=0, atol=1e-5)
    xe = np.linspace(-4e-3, 4e-3, 257)
    ye = np.linspace(-4e-3, 4e-3, 257)
    counts, _ = histogram2d_device(rf[0], rf[2], xe, ye)
    H_gpu = counts.cpu().numpy()
    # numpy on the GPU's own positions: bit-exact
    H_np, _, _ = np.histogram2d(rf[0].cpu().numpy(), rf[2].cpu().numpy(), bins=[xe, ye])
    np.testing.assert_array_equal(H_gpu, H_np.T)
    # oracle positions: only rays within tol of an edge may land elsewhere
    H_ref, _, _ = np.histogram2d(rf_ref[0], rf_ref[2], bins=[xe, ye])
    tol = 1e-5 * np.sqrt(8.0) * L
    near = np.zeros(rf_ref.shape[1], bool)
    for v, e in ((rf_ref[0], xe), (rf_ref[2], ye)):
        near |= np.min(np.abs(v[:, None] - e[None, :]), axis=1) <= tol
    diff = np.abs(H_gpu - H_ref.T).sum()
    assert diff <= 2 * near.sum(), (diff, near.sum())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ordered_histogram_same_counts_as_numpy(dtype):
    """Visiting rays through an order (clustered, so wavefronts share cells, and random)
    gives numpy's counts exactly and its weighted sums to summation order; a mask
    still applies to the ray, not to the position in the order."""
    import torch

    from sloth_amd.processing.diagnostics import histogram2d_device

    rng = np.random.default_rng(21)
    n = 300_001
    x = _edge_heavy_sample(n, -9.0, 9.0, 64, rng, dtype)
    y = _edge_heavy_sample(n, -6.75, 6.75, 48, rng, dtype)
    xe, ye = np.linspace(-9.0, 9.0, 65), np.linspace(-6.75, 6.75, 49)
    w = rng.uniform(0, 1, n)
    mask = rng.uniform(size=n) < 0.9
    keep = ~np.isnan(x) & ~np.isnan(y) & mask
    H, _, _ = np.histogram2d(x[keep], y[keep], bins=[xe, ye])
    Hw, _, _ = np.histogram2d(x[keep], y[keep], bins=[xe, ye], weights=w[keep])
    clustered = np.lexsort((np.nan_to_num(y), np.nan_to_num(x)))  # runs of equal cells
    for order in (clustered, rng.permutation(n)):
        counts, wsum = histogram2d_device(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(),
                                          xe, ye, weights=torch.from_numpy(w),
                                          mask=torch.from_numpy(mask),
                                          order=torch.from_numpy(order.astype(np.int32)))
        np.testing.assert_array_equal(counts.cpu().numpy(), H.T.astype(np.int64))
        np.testing.assert_allclose(wsum.cpu().numpy(), Hw.T, rtol=1e-12, atol=1e-12)
        counts_only, none = histogram2d_device(torch.from_numpy(x).cuda(),
                                               torch.from_numpy(y).cuda(), xe, ye,
                                               mask=torch.from_numpy(mask),
                                               order=torch.from_numpy(order.astype(np.int32)))
        assert none is None
        np.testing.assert_array_equal(counts_only.cpu().numpy(), H.T.astype(np.int64))


def test_ordered_histogram_skips_entries_outside_the_rays():
    """Entries of order outside [0, n) are skipped; a repeated entry counts twice."""
    import torch

    from sloth_amd.processing.diagnostics import histogram2d_device

    rng = np.random.default_rng(22)
    n = 1000
    x = rng.uniform(-1, 1, n).astype(np.float32)
    order = np.arange(n, dtype=np.int32)
    order[10], order[20], order[30] = -1, n, 5
    e = np.linspace(-1, 1, 9)
    counts, _ = histogram2d_device(torch.from_numpy(x).cuda(), torch.from_numpy(x).cuda(), e, e,
                                   order=torch.from_numpy(order))
    vis = order[(order >= 0) & (order < n)]
    H, _, _ = np.histogram2d(x[vis], x[vis], bins=[e, e])
    np.testing.assert_array_equal(counts.cpu().numpy(), H.T.astype(np.int64))
    with pytest.raises(ValueError):
        histogram2d_device(torch.from_numpy(x).cuda(), torch.from_numpy(x).cuda(), e, e,
                           order=torch.from_numpy(order[:-1]))


def _morton_keys_np(p, q, lo_a, hi_a, lo_b, hi_b):
    """15-bit quantisation in float32 and bit interleaving (a on even bits)."""
    def quant(v, lo, hi):
        with np.errstate(invalid="ignore"):
            u = (v - np.float32(lo)) / (np.float32(hi) - np.float32(lo))
        u = np.where(np.isnan(u), np.float32(0), np.clip(u, np.float32(0), np.float32(1)))
        return (u.astype(np.float32) * np.float32(32767.0)).astype(np.uint32)

    qa, qb = quant(p, lo_a, hi_a), quant(q, lo_b, hi_b)
    key = np.zeros(p.shape, dtype=np.int64)
    for bit in range(15):
        key |= ((qa.astype(np.int64) >> bit) & 1) << (2 * bit)
        key |= ((qb.astype(np.int64) >> bit) & 1) << (2 * bit + 1)
    return key


def test_morton_keys_and_ray_order():
    """Keys are the 30-bit interleave of the 15-bit transverse positions (clamped, NaN
    to 0); ray_order is the stable sort of them."""
    import torch

    from sloth_amd import _native as N
    from sloth_amd.core import engine
    from sloth_amd.fields import sine_field

    L = 4e-3
    ne, coords = sine_field(16, L)
    grid = engine.PackedGrid(ne, *coords, 2 * np.pi * C / 1064e-9)
    rng = np.random.default_rng(23)
    n = 50_000
    s0 = np.zeros((6, n), dtype=np.float32)
    s0[0] = rng.uniform(-0.6 * L, 0.6 * L, n)     # some clamp at either side
    s0[1] = rng.uniform(-0.6 * L, 0.6 * L, n)
    s0[1, :7] = np.nan
    s0[0, 7:20] = s0[0, 20:33]                     # equal keys keep their input order
    s0[1, 7:20] = s0[1, 20:33]
    s0[2] = -L / 2
    s0[5] = C
    t = torch.from_numpy(s0).cuda()
    keys = torch.empty(n, dtype=torch.int32, device="cuda")
    import ctypes
    N.check(N.lib().slt_morton_keys(t.data_ptr(), n, ctypes.byref(grid.desc), keys.data_ptr(),
                                    None))
    d = grid.desc
    ref = _morton_keys_np(s0[0], s0[1], d.lo[0], d.hi[0], d.lo[1], d.hi[1])
    got = keys.cpu().numpy().astype(np.int64)
    np.testing.assert_array_equal(got, ref)
    assert got.min() >= 0 and got.max() < (1 << 30)
    order = engine.ray_order(grid, t).cpu().numpy()
    np.testing.assert_array_equal(order, np.argsort(ref, kind="stable"))

"""Adjoint (k4) checks: transposed stencil identity, d kappa / d n_e, dL/dn_e against
central finite differences of the FP64 oracle's loss on 64 voxels drawn from the beam's
own columns (not from the adjoint's output), the unsorted many-cells-per-flush path
against the sorted one, and the full config-A problem (256^3, 2e6 rays, K = 16) through
directional derivatives against central differences of the GPU forward.

The reference has no gradient path; jax.grad through its diffrax solve is not
runnable here, so the oracle-FD comparison is the pin (the adjoint is that of
the discrete RK4 map both compute).
"""
import numpy as np
import pytest

from _parity import C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _setup(n=24, L=4e-3, seed=31, rays=3000, cl=10.0):
    from sloth_amd.core.beam import Beam
    from sloth_amd.fields import kolmogorov_field

    ne, coords = kolmogorov_field(n, L, seed=seed, n0=5e24, frac=0.3, l_min=0.5e-3)
    s0 = Beam(rays, 1.5e-3, 0.0, L / 2, beam_type="square", seeded=True).s0
    return ne, coords, s0, cl


def test_unpack_is_the_transpose_of_pack():
    """<P ne, y> == <ne, unpack(y)> for the gradient channels, P = the packed map
    (coeff times the gradient stencil times 1 / cell size per axis)."""
    import torch

    from sloth_amd.adjoint import unpack
    from sloth_amd.core import engine

    rng = np.random.default_rng(1)
    shape = (9, 11, 7)
    coords = [np.float32(np.linspace(-a, a, k)) for a, k in zip((1e-3, 2e-3, 1.5e-3), shape)]
    ne = rng.uniform(1e23, 1e24, shape).astype(np.float32)
    omega = 2 * np.pi * C / 532e-9
    for pd in ("z", "x"):
        g = engine.PackedGrid(ne, *coords, omega, probing_direction=pd)
        vox = g.voxel_view()
        y = torch.from_numpy(rng.standard_normal((*shape, 3)).astype(np.float32)).cuda()
        adj = torch.zeros(g.nbytes // 4, dtype=torch.float32, device="cuda")
        view = torch.as_strided(adj, (*shape, 4), (*[int(g.desc.stride[a]) * 4 for a in range(3)], 1))
        view[..., :3] = y
        ne_d = torch.from_numpy(ne).cuda()
        cfg = type("cfg", (), {"Te": None, "Z": None, "coulomb_log": None})
        dne = unpack(adj, g, ne_d, cfg).double().cpu().numpy()
        lhs = sum(float((vox[..., a].double().cpu() * y[..., a].double().cpu()).sum())
                  for a in range(3))
        rhs = float((ne.astype(np.float64) * dne).sum())
        assert rhs == pytest.approx(lhs, rel=2e-5)


@pytest.mark.parametrize("ne_val, te, cl", [(1e25, 100.0, None), (5e24, 5.0, None),
                                           (2e24, 20.0, 10.0), (2e27, 3000.0, None)])
def test_dkappa_dne_matches_fd(orc, ne_val, te, cl):
    import torch

    from sloth_amd.adjoint import unpack
    from sloth_amd.core import engine

    lwl = 1064e-9
    omega = 2 * np.pi * C / lwl
    shape = (2, 2, 2)
    c = np.float32(np.linspace(0, 1e-3, 2))
    ne = np.full(shape, ne_val, np.float32)
    g = engine.PackedGrid(ne, c, c, c, omega, inv_brems=True, Te=te, Z=1.0, coulomb_log=cl)
    adj = torch.zeros(g.nbytes // 4, dtype=torch.float32, device="cuda")
    adj.view(-1, 4)[:, 3] = 1.0
    cfg = type("cfg", (), {"Te": te, "Z": 1.0, "coulomb_log": cl})
    got = unpack(adj, g, torch.from_numpy(ne).cuda(), cfg).double().cpu().numpy()[0, 0, 0]
    x = float(np.float32(ne_val))
    d = 1e-6 * x
    fd = (orc.kappa(np.array([x + d]), te, 1.0, omega, coulomb_log=cl)[0]
          - orc.kappa(np.array([x - d]), te, 1.0, omega, coulomb_log=cl)[0]) / (2 * d)
    assert got == pytest.approx(fd, rel=1e-5)


def _crossed_voxels(s0, coords, n_pick, seed):
    """n_pick distinct voxels drawn uniformly over the (ray, plane) pairs of the beam: ray
    r's column at its start, offset transversally by 0 or 1 cell (the cell's corners) or,
    on one axis only, by -1 or 2 (the gradient stencil of a corner reaches that far), at
    a random plane along the probing axis (z).  The rays are nearly straight, so such a
    voxel's density moves the loss; the draw does not look at the adjoint's output, so a
    dropped contribution shows up as a mismatch."""
    rng = np.random.default_rng(seed)
    lo = [float(c[0]) for c in coords]
    d = [(float(c[-1]) - float(c[0])) / (len(c) - 1) for c in coords]
    n = [len(c) for c in coords]
    picked = set()
    while len(picked) < n_pick:
        r = rng.integers(s0.shape[1])
        di, dj = (int(v) for v in rng.integers(-1, 3, size=2))
        if di not in (0, 1) and dj not in (0, 1):
            continue
        i = int((float(s0[0, r]) - lo[0]) / d[0]) + di
        j = int((float(s0[1, r]) - lo[1]) / d[1]) + dj
        k = int(rng.integers(1, n[2] - 1))
        if 0 <= i < n[0] and 0 <= j < n[1]:
            picked.add((i, j, k))
    return sorted(picked)


def _oracle_loss(orc, ne64, c64, s64, nsub, h32, omega, tp, te_, cl, pos_scale):
    grads = orc.gradients(ne64, *c64, omega)
    kap = orc.kappa(ne64, 20.0, 1.0, omega, coulomb_log=cl)
    saves, _ = orc.trace(grads, kap, c64, s64, nsub, h32)
    fin = saves[-1]
    lp = ((fin[0:2] - tp) ** 2).sum(0) / pos_scale ** 2
    le = (np.exp(-fin[6]) - te_) ** 2
    return float(np.mean(lp + le))


@pytest.mark.parametrize("cl", [10.0, None])
def test_adjoint_matches_oracle_central_fd(orc, cl):
    import torch

    from sloth_amd.adjoint import detector_loss, trace_states
    from sloth_amd.core import engine

    ne, coords, s0, _ = _setup(cl=cl)
    target_ne, _, _, _ = _setup(seed=77)
    lwl = 532e-9
    omega = 2 * np.pi * C / lwl
    c64 = [c.astype(np.float64) for c in coords]
    s64 = s0.astype(np.float64)
    pos_scale = 1e-4
    # targets: the oracle's trace through a second seeded grid
    g = engine.PackedGrid(ne, *coords, omega, inv_brems=True, Te=20.0, Z=1.0, coulomb_log=cl)
    T = np.sqrt(8.0) * 4e-3 / C
    nsub, h = engine.schedule(np.array([0.0, T]), engine.nominal_step(g, 0.5))
    h32 = h.astype(np.float32).astype(np.float64)
    tg = orc.gradients(target_ne.astype(np.float64), *c64, omega)
    tk = orc.kappa(target_ne.astype(np.float64), 20.0, 1.0, omega, coulomb_log=cl)
    tsv, _ = orc.trace(tg, tk, c64, s64, nsub, h32)
    tp, te_ = tsv[-1][0:2], np.exp(-tsv[-1][6])

    ne_t = torch.from_numpy(ne).cuda().requires_grad_(True)
    state, tau, _ = trace_states(ne_t, s0, *coords, lwl=lwl, inv_brems=True, Te=20.0, Z=1.0,
                                 coulomb_log=cl, save_times=np.array([0.0, T]))
    loss = detector_loss(state, tau, torch.from_numpy(tp).cuda(), torch.from_numpy(te_).cuda(),
                         pos_scale=pos_scale)
    loss.backward()
    grad = ne_t.grad.double().cpu().numpy()
    L0 = _oracle_loss(orc, ne.astype(np.float64), c64, s64, nsub, h32, omega, tp, te_, cl,
                      pos_scale)
    assert float(loss.detach()) == pytest.approx(L0, rel=1e-4)

    pick = _crossed_voxels(s0, coords, 64, seed=5)
    fd, ad = [], []
    for idx in pick:
        x = float(ne[idx])
        d = 1e-4 * x
        up = ne.astype(np.float64)
        up[idx] = x + d
        lo = ne.astype(np.float64)
        lo[idx] = x - d
        fd.append((_oracle_loss(orc, up, c64, s64, nsub, h32, omega, tp, te_, cl, pos_scale)
                   - _oracle_loss(orc, lo, c64, s64, nsub, h32, omega, tp, te_, cl, pos_scale))
                  / (2 * d))
        ad.append(grad[idx])
    fd, ad = np.array(fd), np.array(ad)
    rel = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    print("adjoint vs FD: normwise", rel, "worst", np.max(np.abs(ad - fd)) / np.abs(fd).max(),
          "of the largest")
    # the sample is informative: most drawn voxels move the loss
    assert np.mean(np.abs(fd) > 1e-3 * np.abs(fd).max()) >= 0.5
    assert rel <= 2e-3
    assert np.all(np.abs(ad - fd) <= 2e-2 * np.abs(fd) + 2e-3 * np.abs(fd).max())


def test_adjoint_independent_of_checkpoint_interval():
    """K = 1, 16, one interval for the whole schedule, and K = 0 (the forward records
    the trajectory and the backward replays nothing) give the same gradient."""
    import torch

    from sloth_amd.adjoint import trace_states

    ne, coords, s0, cl = _setup(rays=2000)
    grads = []
    for K in (1, 16, 10_000, 0):
        ne_t = torch.from_numpy(ne).cuda().requires_grad_(True)
        state, tau, _ = trace_states(ne_t, s0, *coords, inv_brems=True, Te=20.0, Z=1.0,
                                     coulomb_log=cl, ckpt_every=K)
        (state[0].double().sum() * 1e3 + tau.sum()).backward()
        grads.append(ne_t.grad.double().cpu().numpy())
    for g in grads[1:]:
        assert np.linalg.norm(g - grads[0]) <= 1e-5 * np.linalg.norm(grads[0])


def test_forward_without_backward_keeps_no_tape():
    """With grad mode off, trace_states(ckpt_every=0) records no trajectory tape (its peak
    memory stays below the tape's size) and returns what the recording forward returns
    (up to the fp64 rounding of straight moves split at the one checkpoint's end)."""
    import torch

    from sloth_amd import _native as N
    from sloth_amd.adjoint import trace_states

    ne, coords, s0, cl = _setup(rays=20000)
    kw = dict(inv_brems=True, Te=20.0, Z=1.0, coulomb_log=cl, ckpt_every=0)
    ne_t = torch.from_numpy(ne).cuda()
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    with torch.no_grad():
        st0, ta0, n0 = trace_states(ne_t, s0, *coords, **kw)
    torch.cuda.synchronize()
    peak = torch.cuda.max_memory_allocated() - base
    x = ne_t.clone().requires_grad_(True)
    st1, ta1, n1 = trace_states(x, s0, *coords, **kw)
    assert st1.grad_fn is not None
    sched = int(n1.item()) // s0.shape[1] + 1   # at least the in-box steps per ray
    tape = int(N.lib().slt_trajectory_tape_bytes(s0.shape[1], sched))
    print(f"no-grad forward peak {peak} B, tape of one ray-step per in-box step {tape} B")
    assert peak < tape
    assert int(n0.item()) == int(n1.item())
    assert torch.allclose(st0, st1.detach(), rtol=1e-6, atol=0)
    assert torch.allclose(ta0, ta1.detach(), rtol=1e-12, atol=0)


def _line_update(orc, ne64, c64, grads, kap, idx, omega, cl):
    """Recompute, in place, the gradient entries and kappa that depend on ne64[idx]."""
    coeff = orc.gradient_coeff(omega)
    for a in range(3):
        sl = list(idx)
        sl[a] = slice(None)
        line = ne64[tuple(sl)] * coeff
        grads[a][tuple(sl)] = np.gradient(line, c64[a], edge_order=1)
    kap[idx] = orc.kappa(np.array([ne64[idx]]), 20.0, 1.0, omega, coulomb_log=cl)[0]


def test_adjoint_full_size_grid_subset_fd(orc):
    """Config-A grid (256^3 Kolmogorov, 10 mm, 532 nm, absorption): dL/dn_e of the loss of a
    random 4000-ray subset of the 2e6-ray beam against oracle central differences."""
    import torch

    from sloth_amd.adjoint import detector_loss, trace_states
    from sloth_amd.core import engine
    from sloth_amd.core.beam import Beam
    from sloth_amd.fields import kolmogorov_field

    L, n, lwl, cl = 10e-3, 256, 532e-9, 10.0
    ne, coords = kolmogorov_field(n, L, seed=11, n0=2e24, frac=0.3, l_min=0.2e-3)
    tne, _ = kolmogorov_field(n, L, seed=12, n0=2e24, frac=0.3, l_min=0.2e-3)
    s_all = Beam(2_000_000, 4e-3, 0.0, L / 2, beam_type="square", seeded=True).s0
    s0 = s_all[:, np.random.default_rng(3).choice(s_all.shape[1], 4000, replace=False)]
    omega = 2 * np.pi * C / lwl
    c64 = [c.astype(np.float64) for c in coords]
    s64 = s0.astype(np.float64)
    g = engine.PackedGrid(ne, *coords, omega, inv_brems=True, Te=20.0, Z=1.0, coulomb_log=cl)
    T = np.sqrt(8.0) * L / C
    nsub, h = engine.schedule(np.array([0.0, T]), engine.nominal_step(g, 0.5))
    h32 = h.astype(np.float32).astype(np.float64)
    tg = orc.gradients(tne.astype(np.float64), *c64, omega)
    tk = orc.kappa(tne.astype(np.float64), 20.0, 1.0, omega, coulomb_log=cl)
    tsv, _ = orc.trace(tg, tk, c64, s64, nsub, h32)
    tp, te_ = tsv[-1][0:2], np.exp(-tsv[-1][6])
    del tg, tk
    pos_scale = 1e-4

    ne_t = torch.from_numpy(ne).cuda().requires_grad_(True)
    state, tau, _ = trace_states(ne_t, s0, *coords, lwl=lwl, inv_brems=True, Te=20.0, Z=1.0,
                                 coulomb_log=cl, save_times=np.array([0.0, T]))
    detector_loss(state, tau, torch.from_numpy(tp).cuda(), torch.from_numpy(te_).cuda(),
                  pos_scale=pos_scale).backward()
    grad = ne_t.grad.double().cpu().numpy()

    ne64 = ne.astype(np.float64)
    grads = list(orc.gradients(ne64, *c64, omega))
    kap = orc.kappa(ne64, 20.0, 1.0, omega, coulomb_log=cl)

    def loss():
        saves, _ = orc.trace(grads, kap, c64, s64, nsub, h32)
        fin = saves[-1]
        return float(np.mean(((fin[0:2] - tp) ** 2).sum(0) / pos_scale ** 2
                             + (np.exp(-fin[6]) - te_) ** 2))

    pick = _crossed_voxels(s0, coords, 64, seed=9)
    fd, ad = [], []
    for idx in pick:
        x = ne64[idx]
        d = 1e-4 * x
        vals = []
        for v in (x + d, x - d):
            ne64[idx] = v
            _line_update(orc, ne64, c64, grads, kap, idx, omega, cl)
            vals.append(loss())
        ne64[idx] = x
        _line_update(orc, ne64, c64, grads, kap, idx, omega, cl)
        fd.append((vals[0] - vals[1]) / (2 * d))
        ad.append(grad[idx])
    fd, ad = np.array(fd), np.array(ad)
    rel = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    print("full-size adjoint vs FD: normwise", rel, "worst",
          np.max(np.abs(ad - fd)) / np.abs(fd).max(), "of the largest")
    assert np.mean(np.abs(fd) > 1e-3 * np.abs(fd).max()) >= 0.5
    assert rel <= 2e-3
    assert np.all(np.abs(ad - fd) <= 2e-2 * np.abs(fd) + 2e-3 * np.abs(fd).max())


def test_adjoint_unsorted_rays_many_cells_per_flush():
    """Rays in their seeded random order (sort=False): the 64 lanes of a wavefront are
    scattered over ~290 beam columns of the 24^3 grid, so a collective flush (all lanes
    cross a cell face along the probing axis together) holds ~57 distinct cells, more
    than the 16 LDS slots, and the lanes beyond them scatter their own moments.  The
    gradient must equal the sorted run's (same contributions in another summation order)
    and the oracle's central differences on crossed voxels."""
    import torch

    from sloth_amd.adjoint import trace_states

    ne, coords, s0, cl = _setup(rays=3000)
    grads = []
    for sort in (True, False):
        ne_t = torch.from_numpy(ne).cuda().requires_grad_(True)
        state, tau, _ = trace_states(ne_t, s0, *coords, inv_brems=True, Te=20.0, Z=1.0,
                                     coulomb_log=cl, sort=sort)
        (state[0].double().sum() * 1e3 + state[1].double().sum() * 1e3 + tau.sum()).backward()
        grads.append(ne_t.grad.double().cpu().numpy())
    rel = np.linalg.norm(grads[1] - grads[0]) / np.linalg.norm(grads[0])
    print("unsorted vs sorted adjoint: normwise", rel)
    assert rel <= 1e-5
    assert np.array_equal(grads[0] != 0, grads[1] != 0)


def _smooth_direction(ne, coords, seed):
    """ne times a seeded smooth modulation m (|m| <= 1, three random low-k plane waves):
    ne (1 +- eps m) stays non-negative where the field is clipped to zero."""
    rng = np.random.default_rng(seed)
    x, y, z = (c.astype(np.float64) for c in coords)
    L = float(x[-1] - x[0])
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    m = np.zeros(ne.shape)
    for _ in range(3):
        k = rng.integers(1, 5, size=3) * (2 * np.pi / L)
        m += rng.uniform(0.5, 1.0) * np.sin(k[0] * X + k[1] * Y + k[2] * Z
                                            + rng.uniform(0, 2 * np.pi))
    m /= np.abs(m).max()
    return ne.astype(np.float64) * m


def test_adjoint_full_config_directional_derivative():
    """Config A in full: 256^3 Kolmogorov grid (10 mm), 2e6 rays at 532 nm, absorption,
    checkpoints every 16 steps; loss = MSE of detector positions and energies against the
    trace through a second seeded grid.  <dL/dn_e, delta> must equal the central
    difference (L(ne + eps delta) - L(ne - eps delta)) / (2 eps) of the GPU forward for
    three seeded smooth directions.  Positions are close to linear in ne at these
    deflections and the loss is quadratic in them, so the central difference is
    nearly exact even at eps = 1 %; every ray and every wavefront flush contributes."""
    import torch

    from sloth_amd.adjoint import detector_loss, trace_states
    from sloth_amd.core.beam import Beam
    from sloth_amd.fields import kolmogorov_field

    L, n, lwl, cl = 10e-3, 256, 532e-9, 10.0
    ne, coords = kolmogorov_field(n, L, seed=11, n0=2e24, frac=0.3, l_min=0.2e-3)
    tne, _ = kolmogorov_field(n, L, seed=12, n0=2e24, frac=0.3, l_min=0.2e-3)
    s0 = Beam(2_000_000, 4e-3, 0.0, L / 2, beam_type="square", seeded=True).s0
    T = np.sqrt(8.0) * L / C
    kw = dict(lwl=lwl, inv_brems=True, Te=20.0, Z=1.0, coulomb_log=cl,
              save_times=np.array([0.0, T]), ckpt_every=16)
    pos_scale = 1e-4
    with torch.no_grad():
        ts, ttau, _ = trace_states(torch.from_numpy(tne).cuda(), s0, *coords, **kw)
        tp, te_ = ts[0:2].double(), torch.exp(-ttau)
    del tne

    def loss_of(ne_d):
        with torch.no_grad():
            st, ta, _ = trace_states(ne_d, s0, *coords, **kw)
            return float(detector_loss(st, ta, tp, te_, pos_scale=pos_scale))

    ne_t = torch.from_numpy(ne).cuda().requires_grad_(True)
    state, tau, total = trace_states(ne_t, s0, *coords, **kw)
    assert int(total.item()) > 1.0e9                       # 2e6 rays x ~510 in-box steps
    detector_loss(state, tau, tp, te_, pos_scale=pos_scale).backward()
    grad = ne_t.grad.double()
    del state, tau
    eps = 1e-2
    for seed in (1, 2, 3):
        delta = _smooth_direction(ne, coords, seed)
        d_t = torch.from_numpy(delta).cuda()
        ad = float((grad * d_t).sum())
        base = torch.from_numpy(ne.astype(np.float64)).cuda()
        lp = loss_of((base + eps * d_t).float())
        lm = loss_of((base - eps * d_t).float())
        fd = (lp - lm) / (2 * eps)
        rel = abs(ad - fd) / abs(fd)
        print(f"config A direction {seed}: <grad, delta> {ad:.6e}, central difference "
              f"{fd:.6e}, relative {rel:.2e}", flush=True)
        assert rel <= 2e-3


def test_adjoint_full_config_trajectory_tape_equals_checkpoints():
    """Config A (256^3 Kolmogorov, 2e6 rays, absorption) with ckpt_every = 0: the
    forward records the whole trajectory (1443 schedule steps, a 92 GB tape) and the
    backward walks it with no replay.  Final states and tau equal the checkpointed
    forward's up to fp64 rounding (a straight move outside the grid is split at
    checkpoint ends in the checkpointed run and made in one piece here), and the
    gradient equals the K = 16 one up to that and the order of the float atomics."""
    import torch

    from sloth_amd.adjoint import detector_loss, trace_states
    from sloth_amd.core.beam import Beam
    from sloth_amd.fields import kolmogorov_field

    L, n, lwl, cl = 10e-3, 256, 532e-9, 10.0
    ne, coords = kolmogorov_field(n, L, seed=11, n0=2e24, frac=0.3, l_min=0.2e-3)
    tne, _ = kolmogorov_field(n, L, seed=12, n0=2e24, frac=0.3, l_min=0.2e-3)
    s0 = Beam(2_000_000, 4e-3, 0.0, L / 2, beam_type="square", seeded=True).s0
    kw = dict(lwl=lwl, inv_brems=True, Te=20.0, Z=1.0, coulomb_log=cl)
    with torch.no_grad():
        ts, ttau, _ = trace_states(torch.from_numpy(tne).cuda(), s0, *coords, **kw)
        tp, te_ = ts[0:2].double(), torch.exp(-ttau)
    del tne, ts, ttau
    out = {}
    for K in (16, 0):
        ne_t = torch.from_numpy(ne).cuda().requires_grad_(True)
        state, tau, total = trace_states(ne_t, s0, *coords, ckpt_every=K, **kw)
        detector_loss(state, tau, tp, te_, pos_scale=1e-4).backward()
        out[K] = (state.detach().cpu(), tau.detach().cpu(), int(total.item()),
                  ne_t.grad.double().cpu())
        del state, tau, ne_t
        torch.cuda.empty_cache()
    ds = float(((out[0][0] - out[16][0]).abs().amax(1) / out[16][0].abs().amax(1)).max())
    dt = float((out[0][1] - out[16][1]).abs().max() / out[16][1].abs().max())
    print(f"config A, trajectory tape vs checkpoints: state {ds:.2e}, tau {dt:.2e} of the "
          f"largest; in-box steps {out[0][2]} vs {out[16][2]}", flush=True)
    assert ds <= 1e-6 and dt <= 1e-9
    assert abs(out[0][2] - out[16][2]) <= 1e-6 * out[16][2] and out[16][2] > 1.0e9
    g0, g16 = out[0][3], out[16][3]
    rel = float((g0 - g16).norm() / g16.norm())
    print(f"config A, trajectory tape vs checkpoints: gradient normwise {rel:.2e}", flush=True)
    assert rel <= 1e-5


def test_adjoint_nonuniform_grid_matches_oracle_central_fd(orc):
    """Non-uniform axes (every interior grid point moved by up to 0.3 of a cell, the ends
    fixed): the cells are found by the knot search, the Jacobian carries 1 / cell width,
    and dL/dn_e matches central differences of the FP64 oracle's loss (whose interpolation
    searches the axes) on 64 crossed voxels, with checkpoints every 16 steps and with the
    trajectory recorded by the forward."""
    import torch

    from sloth_amd.adjoint import detector_loss, trace_states
    from sloth_amd.core import engine

    ne, coords, s0, cl = _setup()
    target_ne, _, _, _ = _setup(seed=77)
    rng = np.random.default_rng(8)
    nu = []
    for c in coords:
        c = c.astype(np.float64)
        d = (c[-1] - c[0]) / (len(c) - 1)
        c[1:-1] += 0.3 * d * rng.uniform(-1, 1, len(c) - 2)
        nu.append(c.astype(np.float32))
    assert all(np.all(np.diff(c) > 0) for c in nu)
    lwl = 532e-9
    omega = 2 * np.pi * C / lwl
    c64 = [c.astype(np.float64) for c in nu]
    s64 = s0.astype(np.float64)
    pos_scale = 1e-4
    g = engine.PackedGrid(ne, *nu, omega, inv_brems=True, Te=20.0, Z=1.0, coulomb_log=cl)
    assert not g.desc.uniform
    T = np.sqrt(8.0) * 4e-3 / C
    nsub, h = engine.schedule(np.array([0.0, T]), engine.nominal_step(g, 0.5))
    h32 = h.astype(np.float32).astype(np.float64)
    tg = orc.gradients(target_ne.astype(np.float64), *c64, omega)
    tk = orc.kappa(target_ne.astype(np.float64), 20.0, 1.0, omega, coulomb_log=cl)
    tsv, _ = orc.trace(tg, tk, c64, s64, nsub, h32)
    tp, te_ = tsv[-1][0:2], np.exp(-tsv[-1][6])
    grads = {}
    for K in (16, 0):
        ne_t = torch.from_numpy(ne).cuda().requires_grad_(True)
        state, tau, _ = trace_states(ne_t, s0, *nu, lwl=lwl, inv_brems=True, Te=20.0, Z=1.0,
                                     coulomb_log=cl, save_times=np.array([0.0, T]),
                                     ckpt_every=K)
        loss = detector_loss(state, tau, torch.from_numpy(tp).cuda(),
                             torch.from_numpy(te_).cuda(), pos_scale=pos_scale)
        loss.backward()
        grads[K] = ne_t.grad.double().cpu().numpy()
    L0 = _oracle_loss(orc, ne.astype(np.float64), c64, s64, nsub, h32, omega, tp, te_, cl,
                      pos_scale)
    assert float(loss.detach()) == pytest.approx(L0, rel=1e-4)
    rel_k = np.linalg.norm(grads[0] - grads[16]) / np.linalg.norm(grads[16])
    assert rel_k <= 1e-5
    pick = _crossed_voxels(s0, nu, 64, seed=5)
    fd, ad = [], []
    for idx in pick:
        x = float(ne[idx])
        d = 1e-4 * x
        up = ne.astype(np.float64)
        up[idx] = x + d
        lo = ne.astype(np.float64)
        lo[idx] = x - d
        fd.append((_oracle_loss(orc, up, c64, s64, nsub, h32, omega, tp, te_, cl, pos_scale)
                   - _oracle_loss(orc, lo, c64, s64, nsub, h32, omega, tp, te_, cl, pos_scale))
                  / (2 * d))
        ad.append(grads[16][idx])
    fd, ad = np.array(fd), np.array(ad)
    rel = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    print("non-uniform adjoint vs FD: normwise", rel, "worst",
          np.max(np.abs(ad - fd)) / np.abs(fd).max(), "of the largest; K = 0 vs 16", rel_k)
    assert np.mean(np.abs(fd) > 1e-3 * np.abs(fd).max()) >= 0.5
    assert rel <= 2e-3
    assert np.all(np.abs(ad - fd) <= 2e-2 * np.abs(fd) + 2e-3 * np.abs(fd).max())

"""Adjoint (k4) checks: transposed stencil identity, d kappa / d n_e, and dL/dn_e against
central finite differences of the FP64 oracle's loss on 64 random voxels.

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

    rng = np.random.default_rng(5)
    cand = np.argwhere(np.abs(grad) > 0.05 * np.abs(grad).max())
    pick = cand[rng.choice(len(cand), 64, replace=False)]
    fd, ad = [], []
    for idx in map(tuple, pick):
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
    print("adjoint vs FD: normwise", rel, "worst", np.max(np.abs(ad - fd) / np.abs(fd)))
    assert rel <= 2e-3
    assert np.all(np.abs(ad - fd) <= 2e-2 * np.abs(fd) + 2e-3 * np.abs(fd).max())


def test_adjoint_independent_of_checkpoint_interval():
    import torch

    from sloth_amd.adjoint import trace_states

    ne, coords, s0, cl = _setup(rays=2000)
    grads = []
    for K in (1, 16, 10_000):
        ne_t = torch.from_numpy(ne).cuda().requires_grad_(True)
        state, tau, _ = trace_states(ne_t, s0, *coords, inv_brems=True, Te=20.0, Z=1.0,
                                     coulomb_log=cl, ckpt_every=K)
        (state[0].double().sum() * 1e3 + tau.sum()).backward()
        grads.append(ne_t.grad.double().cpu().numpy())
    for g in grads[1:]:
        assert np.linalg.norm(g - grads[0]) <= 1e-5 * np.linalg.norm(grads[0])


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

    rng = np.random.default_rng(9)
    cand = np.argwhere(np.abs(grad) > 0.05 * np.abs(grad).max())
    pick = cand[rng.choice(len(cand), 64, replace=False)]
    fd, ad = [], []
    for idx in map(tuple, pick):
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
    print("full-size adjoint vs FD: normwise", rel, "worst", np.max(np.abs(ad - fd) / np.abs(fd)))
    assert rel <= 2e-3

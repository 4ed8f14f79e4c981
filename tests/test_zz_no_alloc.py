"""slt_adjoint_noalloc (the reverse pass with its tape in the library's own device
array) against slt_adjoint (tape passed in by the caller), through the C ABI.

Two calls in a row on one GPU, the second with fewer rays and another
checkpoint interval, so the second reuses the array the first left behind.
The voxel adjoints are sums of float atomics whose order varies from run to
run, so the two entry points agree to fp32 summation-order rounding, not bit
for bit: normwise 1e-5, elementwise 1e-4 of the largest entry.
"""
import ctypes

import numpy as np
import pytest

from _parity import C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


@pytest.fixture(scope="module")
def setup():
    from sloth_amd.core import engine
    from sloth_amd.core.beam import Beam
    from sloth_amd.fields import kolmogorov_field

    L = 4e-3
    ne, coords = kolmogorov_field(40, L, seed=21, n0=5e24, frac=0.3, l_min=0.4e-3)
    omega = 2 * np.pi * C / 532e-9
    grid = engine.PackedGrid(ne, *coords, omega, inv_brems=True, Te=20.0, Z=1.0,
                             coulomb_log=10.0)
    s_all = Beam(30000, 1.6e-3, 0.0, L / 2, beam_type="square", seeded=True).s0
    T = np.sqrt(8.0) * L / C
    return grid, s_all, T


def _run(setup, n, K, seed):
    import torch

    from sloth_amd import _native as N
    from sloth_amd.core import engine

    grid, s_all, T = setup
    L = N.lib()
    dev = grid.device
    rays = engine.as_rays(s_all[:, :n], dev)
    order = engine.ray_order(grid, rays)
    nsub, h = engine.schedule(np.array([0.0, T]), engine.nominal_step(grid, 0.5))
    sched = int(nsub.astype(np.int64).sum())
    nsub_d = torch.from_numpy(nsub).to(dev)
    h_d = torch.from_numpy(h.astype(np.float32)).to(dev)
    ck = torch.empty(int(L.slt_checkpoint_bytes(n, sched, K)), dtype=torch.uint8, device=dev)
    state = torch.empty((6, n), dtype=torch.float32, device=dev)
    tau = torch.zeros(n, dtype=torch.float64, device=dev)

    a = N.AdjointArgs()
    a.n_rays = n
    a.s0 = rays.data_ptr()
    a.order = order.data_ptr()
    a.n_seg = nsub.size
    a.ckpt_every = K
    a.seg_nsub = nsub_d.data_ptr()
    a.seg_h = h_d.data_ptr()
    a.sched_steps = sched
    a.checkpoints = ck.data_ptr()
    a.final_state = state.data_ptr()
    a.final_tau = tau.data_ptr()
    stream = N.stream_handle()
    N.check(L.slt_forward_ckpt(grid.voxels.data_ptr(), ctypes.byref(grid.desc), ctypes.byref(a),
                               stream))

    g = torch.Generator(device="cpu").manual_seed(seed)
    seed_state = torch.randn((6, n), generator=g, dtype=torch.float32).to(dev)
    seed_tau = torch.randn(n, generator=g, dtype=torch.float64).to(dev)
    a.seed_state = seed_state.data_ptr()
    a.seed_tau = seed_tau.data_ptr()

    adj_ref = torch.zeros(grid.nbytes // 4, dtype=torch.float32, device=dev)
    tape = torch.empty(int(L.slt_tape_bytes(n, K)), dtype=torch.uint8, device=dev)
    a.grid_adj = adj_ref.data_ptr()
    a.tape = tape.data_ptr()
    N.check(L.slt_adjoint(grid.voxels.data_ptr(), ctypes.byref(grid.desc), ctypes.byref(a),
                          stream))

    adj_na = torch.zeros(grid.nbytes // 4, dtype=torch.float32, device=dev)
    a.grid_adj = adj_na.data_ptr()
    a.tape = None
    N.check(L.slt_adjoint_noalloc(grid.voxels.data_ptr(), ctypes.byref(grid.desc),
                                  ctypes.byref(a), stream))
    torch.cuda.synchronize()
    return adj_ref.double().cpu().numpy(), adj_na.double().cpu().numpy()


def _compare(ref, got):
    scale = np.abs(ref).max()
    assert scale > 0
    assert np.all(np.isfinite(got))
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    worst = np.abs(got - ref).max() / scale
    print(f"noalloc vs caller tape: normwise {rel:.3g}, worst {worst:.3g} of max", flush=True)
    assert rel <= 1e-5
    assert worst <= 1e-4
    # the same cells receive contributions
    assert np.array_equal(ref != 0, got != 0)


def test_noalloc_matches_adjoint_two_calls(setup):
    print("noalloc call 1: 30000 rays, K=16", flush=True)
    ref1, got1 = _run(setup, 30000, 16, seed=1)
    _compare(ref1, got1)
    print("noalloc call 2: 7000 rays, K=8", flush=True)
    ref2, got2 = _run(setup, 7000, 8, seed=2)
    _compare(ref2, got2)
    # the second call's result is its own, not what the first left in the scratch
    assert np.linalg.norm(ref2 - ref1) > 0.1 * np.linalg.norm(ref1)


def test_noalloc_rejects_beyond_scratch(setup):
    """A tape larger than the scratch is refused before any device work."""
    import torch

    from sloth_amd import _native as N

    grid, s_all, _ = setup
    L = N.lib()
    K = 16
    # one ray more than the scratch holds, by the library's own tape size
    n = N.SLT_ADJ_SCRATCH_BYTES // int(L.slt_tape_bytes(1, K)) + 1
    assert int(L.slt_tape_bytes(n, K)) > N.SLT_ADJ_SCRATCH_BYTES
    dev = grid.device
    a = N.AdjointArgs()
    a.n_rays = n
    a.s0 = grid.voxels.data_ptr()          # never read: refused first
    a.n_seg = 1
    a.ckpt_every = K
    one = torch.ones(1, dtype=torch.int32, device=dev)
    a.seg_nsub = one.data_ptr()
    a.seg_h = grid.voxels.data_ptr()
    a.sched_steps = 1
    a.checkpoints = grid.voxels.data_ptr()
    a.seed_state = grid.voxels.data_ptr()
    a.grid_adj = grid.voxels.data_ptr()
    rc = L.slt_adjoint_noalloc(grid.voxels.data_ptr(), ctypes.byref(grid.desc), ctypes.byref(a),
                               N.stream_handle())
    assert rc == N.SLT_ERR_UNSUPPORTED
    assert "scratch" in N.last_error()
    a.n_rays = 10
    a.tape = grid.voxels.data_ptr()
    rc = L.slt_adjoint_noalloc(grid.voxels.data_ptr(), ctypes.byref(grid.desc), ctypes.byref(a),
                               N.stream_handle())
    assert rc == N.SLT_ERR_INVALID
    assert "tape must be NULL" in N.last_error()

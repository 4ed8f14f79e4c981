"""Differentiable ray trace: d(loss on final ray states)/d(n_e) on the GPU.

``trace_states(ne, s0, x, y, z, ...)`` returns the final ray states (6, N)
float32 and optical depths (N,) float64 like ``engine.trace``; gradients of
any torch loss built from them flow back to ``ne`` through the hand-written
adjoint kernels (``slt_forward_ckpt`` stores the ray state every
``ckpt_every`` steps, ``slt_adjoint`` runs the transposed RK4 steps interval by
interval, ``slt_adjoint_unpack`` applies the transposed gradient stencil and
d kappa / d n_e).  ``ckpt_every=0`` trades memory for time: the forward records
every move of the trajectory (32 bytes per ray and step, held until the
backward) and the backward replays nothing.  A forward that no backward can
follow (grad mode off, or ``ne`` not requiring grad) keeps neither.  The
gradient is that of the discrete fixed-step RK4 map (the one the forward
computes), so it matches central finite differences of the same trace.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as N
from .constants import C_LIGHT, E_CHARGE, EPSILON_0
from .core import engine


@dataclass
class _Cfg:
    coords: tuple
    omega: float
    probing_direction: str
    inv_brems: bool
    Te: object
    Z: object
    coulomb_log: float | None
    save_times: np.ndarray
    step_fraction: float
    ckpt_every: int
    sort: bool
    record: bool = True   # False when no backward can follow (no checkpoints, no tape)


def _args(grid, rays, order, nsub_d, h_d, sched, ck, K):
    a = N.AdjointArgs()
    a.n_rays = rays.shape[1]
    a.s0 = rays.data_ptr()
    a.order = N.ptr(order)
    a.n_seg = nsub_d.numel()
    a.ckpt_every = K
    a.seg_nsub = nsub_d.data_ptr()
    a.seg_h = h_d.data_ptr()
    a.sched_steps = sched
    a.checkpoints = ck.data_ptr()
    return a


class _TraceFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ne, s0, cfg: _Cfg):
        dev = ne.device
        L = N.lib()
        ne_d = ne.detach().to(torch.float32).contiguous()
        grid = engine.PackedGrid(ne_d, *cfg.coords, cfg.omega,
                                 probing_direction=cfg.probing_direction,
                                 inv_brems=cfg.inv_brems, Te=cfg.Te, Z=cfg.Z,
                                 coulomb_log=cfg.coulomb_log, device=dev)
        rays = engine.as_rays(s0, dev)
        n = rays.shape[1]
        nsub, h = engine.schedule(cfg.save_times, engine.nominal_step(grid, cfg.step_fraction))
        sched = int(nsub.astype(np.int64).sum())
        K = int(cfg.ckpt_every)
        if not cfg.record:
            # no backward can follow: one checkpoint (the start) and no tape
            K = max(1, min(sched, 2 ** 31 - 1))
        order = engine.ray_order(grid, rays) if cfg.sort and n > 0 else None
        nsub_d = engine.const_device(nsub, dev)
        h_d = engine.const_device(h.astype(np.float32), dev)
        ck = torch.empty(max(int(L.slt_checkpoint_bytes(n, sched, K)), 1), dtype=torch.uint8,
                         device=dev)
        traj = None
        if K == 0:
            nb = int(L.slt_trajectory_tape_bytes(n, sched))
            if n > 0 and sched > 0 and nb == 0:
                raise ValueError(f"a trajectory tape of {n} rays x {sched} steps is too large")
            traj = torch.empty(max(nb, 1), dtype=torch.uint8, device=dev)
        state = torch.empty((6, n), dtype=torch.float32, device=dev)
        tau = torch.zeros(n, dtype=torch.float64, device=dev)
        total = torch.zeros(1, dtype=torch.int64, device=dev)
        a = _args(grid, rays, order, nsub_d, h_d, sched, ck, K)
        a.final_state = state.data_ptr()
        a.final_tau = tau.data_ptr()
        a.total_steps = total.data_ptr()
        a.tape = N.ptr(traj)
        N.check(L.slt_forward_ckpt(grid.voxels.data_ptr(), ctypes.byref(grid.desc),
                                   ctypes.byref(a), N.stream_handle()))
        ctx.cfg = cfg
        ctx.grid = grid
        ctx.saved = (ne_d, rays, order, nsub_d, h_d, ck, sched, K, total, traj)
        ctx.mark_non_differentiable(total)
        return state, tau, total

    @staticmethod
    def backward(ctx, g_state, g_tau, _g_total):
        cfg = ctx.cfg
        grid = ctx.grid
        ne_d, rays, order, nsub_d, h_d, ck, sched, K, _, traj = ctx.saved
        dev = ne_d.device
        L = N.lib()
        n = rays.shape[1]
        if g_state is None:
            g_state = torch.zeros((6, n), dtype=torch.float32, device=dev)
        seed = g_state.to(torch.float32).contiguous()
        seed_tau = g_tau.to(torch.float64).contiguous() if g_tau is not None else None
        adj = torch.zeros(grid.nbytes // 4, dtype=torch.float32, device=dev)
        tape = traj if K == 0 else torch.empty(max(int(L.slt_tape_bytes(n, K)), 1),
                                               dtype=torch.uint8, device=dev)
        a = _args(grid, rays, order, nsub_d, h_d, sched, ck, K)
        a.seed_state = seed.data_ptr()
        a.seed_tau = N.ptr(seed_tau)
        a.grid_adj = adj.data_ptr()
        a.tape = tape.data_ptr()
        N.check(L.slt_adjoint(grid.voxels.data_ptr(), ctypes.byref(grid.desc), ctypes.byref(a),
                              N.stream_handle()))
        dne = unpack(adj, grid, ne_d, cfg)
        return dne, None, None


def unpack(adj: torch.Tensor, grid: engine.PackedGrid, ne_d: torch.Tensor, cfg) -> torch.Tensor:
    """dL/dn_e (nx, ny, nz) from a packed voxel adjoint."""
    dev = ne_d.device
    cd = [engine.const_device(c, dev) for c in grid.coords]
    te_d = z_d = None
    if grid.inv_brems:
        te_d = engine.device_f32(cfg.Te, dev).reshape(-1)
        z_d = engine.device_f32(cfg.Z, dev).reshape(-1)
    dne = torch.empty(grid.shape, dtype=torch.float32, device=dev)
    cl = float(cfg.coulomb_log) if cfg.coulomb_log else 0.0
    N.check(N.lib().slt_adjoint_unpack(
        adj.data_ptr(), ctypes.byref(grid.desc), ne_d.data_ptr(), cd[0].data_ptr(),
        cd[1].data_ptr(), cd[2].data_ptr(), engine.grad_coeff(grid.omega), N.ptr(te_d),
        te_d.numel() if te_d is not None else 0, N.ptr(z_d), z_d.numel() if z_d is not None else 0,
        grid.omega, E_CHARGE, EPSILON_0, C_LIGHT, cl, dne.data_ptr(), N.stream_handle()))
    return dne


def trace_states(ne: torch.Tensor, s0, x, y, z, *, lwl: float = 532e-9,
                 probing_direction: str = "z", inv_brems: bool = False, Te=None, Z=None,
                 coulomb_log=None, save_times=None, depth=None, step_fraction: float = 0.5,
                 ckpt_every: int = 16, sort: bool = True):
    """Final ray states and optical depths, differentiable with respect to ``ne``.

    The horizon is ``save_times[-1]`` if given, else ``depth / c`` if given, else
    the reference solve horizon sqrt(8) L / c with L the grid extent along the
    probing axis.  Returns (state (6, N) float32, tau (N,) float64, total
    in-box ray-steps (1,) int64).

    ``ckpt_every`` K > 0 stores each ray's state every K steps (60 bytes per ray
    and checkpoint) and replays each interval in the backward.  K = 0 stores the
    whole trajectory instead (32 bytes per ray and schedule step, e.g. 92 GB for
    2e6 rays over 1443 steps, held from the forward to the backward) and the
    backward replays nothing.
    """
    if int(ckpt_every) < 0:
        raise ValueError(f"ckpt_every must be >= 0, got {ckpt_every}")
    N.require_device()
    coords = tuple(engine.host_f32(c).reshape(-1) for c in (x, y, z))
    ax = engine.axis_index(probing_direction)
    if save_times is None:
        if depth is None:
            span = float(coords[ax][-1]) - float(coords[ax][0])
            T = np.sqrt(8.0) * span / C_LIGHT
        else:
            T = float(depth) / C_LIGHT
        save_times = np.array([0.0, T])
    cfg = _Cfg(coords, 2 * np.pi * C_LIGHT / lwl, probing_direction, inv_brems, Te, Z,
               coulomb_log, np.asarray(save_times, dtype=np.float64), step_fraction,
               int(ckpt_every), sort,
               record=torch.is_grad_enabled() and bool(getattr(ne, "requires_grad", False)))
    return _TraceFn.apply(ne, s0, cfg)


def detector_loss(state, tau, target_pos, target_energy, *, probing_direction="z",
                  pos_scale: float = 1e-3, energy_weight: float = 1.0):
    """Mean over rays of ((p - p_t)^2 summed over the two transverse axes) / pos_scale^2
    + energy_weight (exp(-tau) - E_t)^2: the MSE on detector positions and energies."""
    ax = engine.axis_index(probing_direction)
    tr = [a for a in range(3) if a != ax]
    pos = state[tr].to(torch.float64)
    e = torch.exp(-tau)
    lp = ((pos - target_pos) ** 2).sum(0) / pos_scale ** 2
    le = energy_weight * (e - target_energy) ** 2
    return (lp + le).mean()

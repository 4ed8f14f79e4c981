"""Device-resident packed grids and RK4 ray traces (the layer under ``propagator``).

``PackedGrid`` runs k1 (gradient stencil + inverse-bremsstrahlung rate + 16-byte
voxel pack) once per density field and keeps the voxels in HBM; ``trace`` runs
k2 (the RK4 integrator) over a (6, N) ray tensor with a save schedule.  Both go
through the C ABI of ``include/sloth_hip.h``; PyTorch only provides device
memory and the stream.
"""
from __future__ import annotations

import ctypes
from collections import OrderedDict
from dataclasses import dataclass

import numpy as np
import torch

from .. import _native as N
from ..constants import ALPHA_CRIT, C_LIGHT, E_CHARGE, EPSILON_0

AXES = {"x": 0, "y": 1, "z": 2}


def axis_index(probing_direction: str) -> int:
    if probing_direction not in AXES:
        raise AssertionError("Invalid entry for probing_direction! Use x, y or z.")
    return AXES[probing_direction]


def grad_coeff(omega: float) -> float:
    """-0.5 c^2 / (3.14207787e-4 omega^2) in float64 (src/core/propagator.py:88)."""
    return float(np.float64(-0.5) * float(C_LIGHT) ** 2 / (ALPHA_CRIT * float(omega) ** 2))


def critical_density(omega: float) -> float:
    """n_cr = 3.14207787e-4 omega^2 [m^-3] (src/core/propagator.py:73)."""
    return ALPHA_CRIT * float(omega) ** 2


def host_f32(a) -> np.ndarray:
    if isinstance(a, torch.Tensor):
        return a.detach().to("cpu", torch.float32).contiguous().numpy()
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


_STAGED_MAX_BYTES = 1 << 20


def to_device(a, device, stream=None) -> torch.Tensor:
    """Host array -> device tensor, queued on ``stream`` (default: the current one).

    Arrays up to 1 MiB (coordinates, schedules, bin edges) go through pinned
    memory without blocking the host; torch's caching host allocator keeps the
    staging block until the copy has run.  A plain ``.to(device)`` would wait
    for everything queued on the stream before it, which stalls a host that is
    feeding several streams.  Larger arrays take the ordinary (blocking) copy.
    """
    arr = np.ascontiguousarray(a)
    if not arr.flags.writeable:     # torch wraps only writable arrays
        arr = arr.copy()
    t = torch.from_numpy(arr)
    with torch.cuda.stream(stream):
        if t.numel() * t.element_size() <= _STAGED_MAX_BYTES:
            return t.pin_memory().to(device, non_blocking=True)
        return t.to(device)


_CONST: "OrderedDict[tuple, tuple]" = OrderedDict()  # key -> (tensor, upload event)
_CONST_MAX = 512


def const_device(a, device, stream=None) -> torch.Tensor:
    """A read-only device copy of a small host array, shared by every caller that passes
    the same dtype, shape and bytes.

    For coordinates, step schedules, bin edges and scalar Te/Z. Requests of a serving mix
    reuse a handful of these, and each small host-to-device copy is a blit kernel of tens
    of microseconds. Callers must not write to the result. Arrays above 1 MiB are
    copied as ``to_device`` does. The upload's completion is recorded as an event. A
    later hit, on whatever stream, makes that stream wait for the event (a GPU-side
    dependency, no host sync). An entry stays alive until evicted. Eviction first
    waits for the device, so a kernel queued on any stream has finished reading it.
    """
    arr = np.ascontiguousarray(a)
    if arr.nbytes > _STAGED_MAX_BYTES:
        return to_device(arr, device, stream)
    dev = torch.device(device)
    key = (dev.type, dev.index, arr.dtype.str, arr.shape, arr.tobytes())
    use = stream if stream is not None else torch.cuda.current_stream(dev)
    hit = _CONST.get(key)
    if hit is not None:
        _CONST.move_to_end(key)
        t, uploaded = hit
        use.wait_event(uploaded)
        return t
    t = to_device(arr, dev, use)
    uploaded = torch.cuda.Event()
    uploaded.record(use)
    _CONST[key] = (t, uploaded)
    if len(_CONST) > _CONST_MAX:
        torch.cuda.synchronize(dev)
        _CONST.popitem(last=False)
    return t


def _small_f32(a, device, stream=None) -> torch.Tensor:
    """device_f32 for read-only inputs: host values go through const_device."""
    if isinstance(a, torch.Tensor):
        return device_f32(a, device, stream)
    return const_device(host_f32(a), device, stream)


def device_f32(a, device, stream=None) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.detach().to(device=device, dtype=torch.float32).contiguous()
    return to_device(host_f32(a), device, stream)


def _cl(coulomb_log) -> float:
    return float(coulomb_log) if coulomb_log is not None and coulomb_log > 0 else 0.0


class PackedGrid:
    """A density field packed for tracing: 16 B per grid point, resident in HBM.

    Voxel (i, j, k) holds (d/dx, d/dy, d/dz) of ``coeff * ne`` and, when
    ``inv_brems``, the inverse-bremsstrahlung rate kappa [1/s].  The probing
    axis is the fastest-varying axis of the layout.
    """

    def __init__(self, ne, x, y, z, omega, *, probing_direction="z", inv_brems=False,
                 Te=None, Z=None, coulomb_log=None, device=None, stream=None):
        dev = device if device is not None else N.require_device()
        L = N.lib()
        self.omega = float(omega)
        self.axis = axis_index(probing_direction)
        self.probing_direction = probing_direction
        self.inv_brems = bool(inv_brems)
        self.coords = [host_f32(c).reshape(-1) for c in (x, y, z)]
        nx, ny, nz = (c.size for c in self.coords)
        ne_d = device_f32(ne, dev, stream)
        if tuple(ne_d.shape) != (nx, ny, nz):
            raise AssertionError(
                f"ne has shape {tuple(ne_d.shape)} but the coordinates give {(nx, ny, nz)}")
        desc = N.GridDesc()
        xs, ys, zs = self.coords
        N.check(L.slt_grid_desc_init(xs.ctypes.data, ys.ctypes.data, zs.ctypes.data, nx, ny, nz,
                                     self.axis, 1 if self.inv_brems else 0, ctypes.byref(desc)))
        if not desc.uniform:
            # cells between the grid points: their positions in grid units, on the device
            kn = np.empty(nx + ny + nz, dtype=np.float32)
            N.check(L.slt_grid_knots(xs.ctypes.data, ys.ctypes.data, zs.ctypes.data,
                                     ctypes.byref(desc), kn.ctypes.data))
            self.knots = const_device(kn, dev, stream)
            desc.knots = self.knots.data_ptr()
        else:
            self.knots = None
        self.desc = desc
        self.shape = (nx, ny, nz)
        self.nbytes = int(L.slt_grid_bytes(ctypes.byref(desc)))
        self.voxels = torch.empty(self.nbytes // 4, dtype=torch.float32, device=dev)
        cd = [const_device(c, dev, stream) for c in self.coords]
        te_d = z_d = None
        if self.inv_brems:
            if Te is None or Z is None:
                raise AssertionError("Te and Z must both be provided when inv_brems=True.")
            te_d = _small_f32(Te, dev, stream).reshape(-1)
            z_d = _small_f32(Z, dev, stream).reshape(-1)
        s = N.stream_handle(stream)
        N.check(L.slt_grid_pack(
            ne_d.data_ptr(), cd[0].data_ptr(), cd[1].data_ptr(), cd[2].data_ptr(),
            grad_coeff(omega), N.ptr(te_d), te_d.numel() if te_d is not None else 0,
            N.ptr(z_d), z_d.numel() if z_d is not None else 0, self.omega, E_CHARGE, EPSILON_0,
            C_LIGHT, _cl(coulomb_log), ctypes.byref(desc), self.voxels.data_ptr(),
            None, None, None, None, s))
        self.device = dev
        self.spacing = np.array([1.0 / float(desc.inv_d[a]) for a in range(3)])
        # the smallest cell width (the nominal spacing on a uniform grid)
        self._min_cell = float(self.spacing.min()) if desc.uniform else float(
            min(np.diff(c.astype(np.float64)).min() for c in self.coords))

    @property
    def min_spacing(self) -> float:
        return self._min_cell

    def voxel_view(self) -> torch.Tensor:
        """(nx, ny, nz, 4) view in (x, y, z) index order (strided, no copy)."""
        st = [int(self.desc.stride[a]) * 4 for a in range(3)]
        return torch.as_strided(self.voxels, (*self.shape, 4), (*st, 1))


def grid_fields(ne, x, y, z, omega, *, Te=None, Z=None, coulomb_log=None, want_kappa=False,
                device=None):
    """Planar fp32 (dndx, dndy, dndz[, kappa]) grids from k1 (no packing)."""
    dev = device if device is not None else N.require_device()
    L = N.lib()
    coords = [host_f32(c).reshape(-1) for c in (x, y, z)]
    nx, ny, nz = (c.size for c in coords)
    ne_d = device_f32(ne, dev)
    if tuple(ne_d.shape) != (nx, ny, nz):
        raise AssertionError(
            f"ne has shape {tuple(ne_d.shape)} but the coordinates give {(nx, ny, nz)}")
    desc = N.GridDesc()
    N.check(L.slt_grid_desc_init(coords[0].ctypes.data, coords[1].ctypes.data,
                                 coords[2].ctypes.data, nx, ny, nz, 2, 1 if want_kappa else 0,
                                 ctypes.byref(desc)))
    outs = [torch.empty((nx, ny, nz), dtype=torch.float32, device=dev)
            for _ in range(4 if want_kappa else 3)]
    cd = [const_device(c, dev) for c in coords]
    te_d = z_d = None
    if want_kappa:
        te_d = device_f32(Te, dev).reshape(-1)
        z_d = device_f32(Z, dev).reshape(-1)
    N.check(L.slt_grid_pack(
        ne_d.data_ptr(), cd[0].data_ptr(), cd[1].data_ptr(), cd[2].data_ptr(), grad_coeff(omega),
        N.ptr(te_d), te_d.numel() if te_d is not None else 0, N.ptr(z_d),
        z_d.numel() if z_d is not None else 0, float(omega), E_CHARGE, EPSILON_0, C_LIGHT,
        _cl(coulomb_log), ctypes.byref(desc), None, outs[0].data_ptr(), outs[1].data_ptr(),
        outs[2].data_ptr(), outs[3].data_ptr() if want_kappa else None, N.stream_handle()))
    return tuple(outs)


def kappa_field(ne, Te, Z, omega, *, coulomb_log=None, device=None) -> torch.Tensor:
    """Element-wise inverse-bremsstrahlung rate [1/s] (fp64 math, fp32 result)."""
    dev = device if device is not None else N.require_device()
    ne_d = device_f32(ne, dev)
    shape = ne_d.shape
    ne1 = ne_d.reshape(-1)
    te_d = device_f32(Te, dev).reshape(-1)
    z_d = device_f32(Z, dev).reshape(-1)
    if te_d.numel() not in (1, ne1.numel()) or z_d.numel() not in (1, ne1.numel()):
        if ne1.numel() == 1:
            # scalar density against Te/Z grids: broadcast the density
            ne1 = ne1.expand(max(te_d.numel(), z_d.numel())).contiguous()
            shape = (te_d if te_d.numel() > 1 else z_d).shape
        else:
            raise ValueError("Te and Z must be scalars or have the shape of ne")
    out = torch.empty(ne1.numel(), dtype=torch.float32, device=dev)
    N.check(N.lib().slt_kappa(ne1.data_ptr(), ne1.numel(), te_d.data_ptr(), te_d.numel(),
                              z_d.data_ptr(), z_d.numel(), float(omega), E_CHARGE, EPSILON_0,
                              C_LIGHT, _cl(coulomb_log), out.data_ptr(), N.stream_handle()))
    return out.reshape(shape)


def schedule(save_times, h_nominal: float):
    """Split [t_0, t_{S-1}] at the save times into RK4 steps of at most h_nominal.

    Segment s (between save s and s+1) gets nsub[s] = ceil(dt_s / h_nominal)
    equal steps of h[s] = dt_s / nsub[s], so every save time is hit exactly.
    """
    t = np.asarray(save_times, dtype=np.float64).reshape(-1)
    if t.size < 2:
        raise ValueError("need at least two save times")
    dt = np.diff(t)
    if np.any(dt < 0):
        raise ValueError("save times must be non-decreasing")
    nsub = np.ceil(dt / h_nominal - 1e-9).astype(np.int64)
    nsub = np.maximum(nsub, 0)
    h = np.where(nsub > 0, dt / np.maximum(nsub, 1), 0.0)
    return nsub.astype(np.int32), h


def nominal_step(grid: PackedGrid, step_fraction: float) -> float:
    """RK4 time step: step_fraction of the smallest cell edge, at the speed of light."""
    if not step_fraction > 0:
        raise ValueError("step_fraction must be positive")
    return step_fraction * grid.min_spacing / C_LIGHT


@dataclass
class TraceResult:
    states: torch.Tensor           # (S, 6, N) float32: x, y, z, vx, vy, vz at each save
    tau: torch.Tensor | None       # (S, N) float64 optical depth (inv_brems only)
    path: torch.Tensor | None      # (N,) float64 path length [m]
    steps: torch.Tensor            # (N,) int32 in-box RK4 steps per ray
    total_steps: torch.Tensor      # (1,) int64 sum of steps
    order: torch.Tensor | None     # processing order used (Morton sort) or None
    nsub: np.ndarray
    h: np.ndarray
    # True: columns are in processing order (column t is the caller's ray order[t]);
    # False: columns are in the caller's order
    permuted: bool = False


def ray_order(grid: PackedGrid, s0: torch.Tensor, stream=None) -> torch.Tensor:
    """Z-order of the rays' transverse start positions (int32 permutation)."""
    n = s0.shape[1]
    # the keys, the kernel that writes them and the sort that reads them all go
    # on `stream`, so the sort cannot run ahead of the keys
    with torch.cuda.stream(stream):
        keys = torch.empty(n, dtype=torch.int32, device=s0.device)
        N.check(N.lib().slt_morton_keys(s0.data_ptr(), n, ctypes.byref(grid.desc),
                                        keys.data_ptr(), N.stream_handle(stream)))
        return torch.sort(keys, stable=True).indices.to(torch.int32)


def _check_order(order: torch.Tensor, n: int) -> None:
    if order.dtype != torch.int32 or order.shape != (n,) or not order.is_contiguous():
        raise AssertionError("order must be a contiguous int32 permutation of the rays")


def ray_work(n: int, device) -> torch.Tensor:
    """Device workspace for gather_rays / unpermute_rays over n rays (32 B per ray)."""
    return torch.empty(max(int(N.lib().slt_ray_work_bytes(n)), 1), dtype=torch.uint8,
                       device=device)


def gather_rays(s0: torch.Tensor, order: torch.Tensor, stream=None,
                work: torch.Tensor | None = None, out: torch.Tensor | None = None) -> torch.Tensor:
    """(6, N) start states in processing order: column t is the caller's ray order[t]."""
    n = s0.shape[1]
    _check_order(order, n)
    with torch.cuda.stream(stream):
        if work is None:
            work = ray_work(n, s0.device)
        if out is None:
            out = torch.empty_like(s0)
        N.check(N.lib().slt_gather_rays(s0.data_ptr(), n, order.data_ptr(), out.data_ptr(),
                                        work.data_ptr(), N.stream_handle(stream)))
    return out


def unpermute_rays(src: torch.Tensor | None, src64: torch.Tensor | None, order: torch.Tensor,
                   stream=None, work: torch.Tensor | None = None):
    """Results in processing order back to the caller's: returns (dst, dst64) with
    dst[:, order[t]] = src[:, t] for the fp32 rows of `src` (r, N) and
    dst64[order[t]] = src64[t] for an optional fp64 row."""
    n = order.numel()
    _check_order(order, n)
    rows = 0 if src is None else src.shape[0]
    if src is not None:
        if src.dtype != torch.float32 or src.dim() != 2 or src.shape[1] != n:
            raise AssertionError("src must be an (r, N) float32 tensor")
        src = src.contiguous()
    if src64 is not None:
        if src64.dtype != torch.float64 or src64.shape != (n,):
            raise AssertionError("src64 must be an (N,) float64 tensor")
        src64 = src64.contiguous()
    with torch.cuda.stream(stream):
        if work is None:
            work = ray_work(n, order.device)
        dst = torch.empty_like(src) if src is not None else None
        dst64 = torch.empty_like(src64) if src64 is not None else None
        N.check(N.lib().slt_unpermute_rays(N.ptr(src), rows, N.ptr(src64), n, order.data_ptr(),
                                           N.ptr(dst), N.ptr(dst64), work.data_ptr(),
                                           N.stream_handle(stream)))
    return dst, dst64


def as_rays(s0, device) -> torch.Tensor:
    """(6, N) float32 contiguous device tensor from any array-like."""
    t = device_f32(s0, device)
    if t.dim() != 2 or t.shape[0] != 6:
        raise AssertionError("Expected a (6, N) matrix of pre-created rays.")
    return t


def trace(grid: PackedGrid, s0, save_times, *, step_fraction: float = 0.5,
          save_all: bool = True, want_path: bool = True, sort: bool = True,
          order: torch.Tensor | None = None, gather: bool = False,
          stream=None) -> TraceResult:
    """Integrate rays through ``grid`` with fixed-step RK4, saving at ``save_times``.

    ``save_times`` are real times [s] starting at 0 (the initial state).
    ``gather=True`` (with a processing order) gathers the start states into that
    order first and traces them in place, so that every state read and write is
    coalesced; the result's columns then stay in processing order
    (``permuted=True``), for ``jones_device(..., order=result.order)`` to put back.
    """
    rays = as_rays(s0, grid.device)
    n = rays.shape[1]
    nsub, h = schedule(save_times, nominal_step(grid, step_fraction))
    if int(nsub.astype(np.int64).sum()) >= 2 ** 31:
        raise NotImplementedError("more than 2^31 RK4 steps per ray")
    dev = grid.device
    nseg = nsub.size
    S = nseg + 1 if save_all else 1
    states = torch.empty((S, 6, n), dtype=torch.float32, device=dev)
    tau = torch.empty((S, n), dtype=torch.float64, device=dev) if grid.inv_brems else None
    path = torch.empty(n, dtype=torch.float64, device=dev) if want_path else None
    steps = torch.empty(n, dtype=torch.int32, device=dev)
    total = torch.zeros(1, dtype=torch.int64, device=dev)
    if n == 0:
        return TraceResult(states, tau, path, steps, total, None, nsub, h)
    if order is None and sort:
        order = ray_order(grid, rays, stream)
    permuted = gather and order is not None
    if permuted:
        rays = gather_rays(rays, order, stream)
    seg_nsub = const_device(nsub, dev, stream)
    seg_h = const_device(h.astype(np.float32), dev, stream)
    a = N.TraceArgs()
    a.n_rays = n
    a.s0 = rays.data_ptr()
    a.order = None if permuted else N.ptr(order)
    a.n_seg = nseg
    a.save_every = 1 if save_all else 0
    a.seg_nsub = seg_nsub.data_ptr()
    a.seg_h = seg_h.data_ptr()
    a.save_state = states.data_ptr()
    a.save_tau = N.ptr(tau)
    a.path = N.ptr(path)
    a.steps = steps.data_ptr()
    a.total_steps = total.data_ptr()
    N.check(N.lib().slt_trace(grid.voxels.data_ptr(), ctypes.byref(grid.desc), ctypes.byref(a),
                              N.stream_handle(stream)))
    return TraceResult(states, tau, path, steps, total, order, nsub, h, permuted)

"""Continuous-batching ray-trace serving on one GPU.

Requests (``RequestSpec``) arrive over time.  The engine

* admits them in arrival order while a slot is free and their HBM footprint
  fits the budget, letting a request that fits pass an older one that waits
  for memory unless that one has waited longer than an age cap (a request
  larger than the whole budget is rejected at once),
* sets each admitted request up on a side stream (field synthesis, k1 pack,
  beam, Morton order, resumable ray state) while earlier requests keep
  running,
* advances all set-up requests together in *step batches*: one persistent
  ``slt_advance`` launch per batch walks a device-resident job table in which
  every request has its own step range; the launch is captured once in a HIP
  graph and replayed after each table update,
* finalises a request (Jones vector, exp(-tau), detector image) on the
  compute stream as soon as its last batch is enqueued, and reports its
  latency from GPU timestamps.

A request's result does not depend on what it shares batches with: every ray
runs the same per-ray integration as the offline ``slt_trace`` path, only
split at batch boundaries.
"""
from __future__ import annotations

import ctypes
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _native as N
from ..core import engine
from ..processing.diagnostics import histogram2d_device
from ..shared.propagation import jones_device
from .requests import RequestSpec, materialize


# ---------------------------------------------------------------------------
# Admission (host only)
# ---------------------------------------------------------------------------

class Admission:
    """Admission by free slots and HBM footprint, in arrival order, without head-of-line
    blocking: an arrived request that fits is admitted even when an older one is waiting
    for memory, unless the oldest waiting request has waited longer than ``max_wait``
    seconds -- then nothing behind it is admitted until it is, so running requests drain
    and free its memory.  A request larger than the whole budget is rejected at once."""

    def __init__(self, budget_bytes: int, max_slots: int, max_wait: float = 0.05):
        if max_slots < 1:
            raise ValueError("max_slots must be >= 1")
        if max_wait < 0:
            raise ValueError("max_wait must be >= 0")
        self.budget = int(budget_bytes)
        self.max_slots = int(max_slots)
        self.max_wait = float(max_wait)
        self.used = 0
        self.slots = 0

    def admit(self, queue: deque, now: float):
        """Pop the arrived requests that fit.  Returns (admitted, rejected)."""
        admitted, rejected = [], []
        i = 0
        while i < len(queue) and queue[i].arrival <= now:
            req = queue[i]
            fp = req.footprint_bytes()
            if fp > self.budget:
                rejected.append(req)
                del queue[i]
                continue
            if self.slots >= self.max_slots:
                break
            if self.used + fp <= self.budget:
                del queue[i]
                self.used += fp
                self.slots += 1
                admitted.append(req)
                continue
            # req waits for memory; later arrivals may pass it while it is young
            if now - req.arrival > self.max_wait:
                break
            i += 1
        return admitted, rejected

    def release(self, req: RequestSpec):
        self.used -= req.footprint_bytes()
        self.slots -= 1
        assert self.used >= 0 and self.slots >= 0


@dataclass
class Response:
    rid: int
    status: str                     # "ok" | "rejected" | "error"
    message: str = ""
    rf: torch.Tensor | None = None  # (4, N) float32
    amp: torch.Tensor | None = None  # (N,) float64 exp(-tau)
    image: torch.Tensor | None = None  # (bins, bins) counts or exp(-tau)-weighted sums
    path: torch.Tensor | None = None
    state: torch.Tensor | None = None  # (6, N) float32 final x, y, z, vx, vy, vz
    tau: torch.Tensor | None = None    # (N,) float64 final optical depth (absorption only)
    steps: torch.Tensor | None = None  # (N,) int32 in-box RK4 steps per ray
    ray_steps: int = 0
    arrival: float = 0.0
    completed: float = math.nan     # seconds since the run started (GPU clock)
    n_rays: int = 0

    @property
    def latency(self) -> float:
        return self.completed - self.arrival


@dataclass
class _Active:
    req: RequestSpec
    mat: object
    order: torch.Tensor
    state: torch.Tensor
    seg_nsub: torch.Tensor
    seg_h: torch.Tensor
    sched: int
    inbox_estimate: int
    out_state: torch.Tensor
    out_tau: torch.Tensor
    out_path: torch.Tensor
    out_steps: torch.Tensor
    counter: torch.Tensor
    ready: torch.cuda.Event
    cursor: int = 0
    finished: bool = False
    done: torch.cuda.Event | None = None
    result: dict = field(default_factory=dict)


def _f(t):
    return t.data_ptr() if t is not None else None


class ServingEngine:
    """Continuous-batching engine.

    ``blocks``: workgroups of the persistent step-batch kernel; 1024 = one resident
    round on MI355X (256 CUs x 4 SIMDs x 4 wavefronts at its register count).
    ``max_wait``: seconds an arrived request may wait for memory while younger requests
    that fit are admitted past it; after that, admission holds everything behind it.
    ``batch_work``: target ray-steps per batch (the batch length in steps is this
    over the rays resident, at most 4096); smaller batches admit new requests sooner,
    but each batch costs ~0.35 ms of unit dealing and state traffic whatever its
    length.  On the serving mix, 4e9 (at most ~30 ms of work) gave higher throughput
    and lower p50/p99 latency than 5e8-2e9, and longer batches gained no more.
    """
    def __init__(self, *, max_slots: int = 16, hbm_budget: int | None = None,
                 blocks: int = 1024, batch_work: float = 4e9, use_graph: bool = True,
                 step_fraction: float = 0.5, keep_outputs: bool = True, max_wait: float = 0.05,
                 device=None):
        self.dev = device if device is not None else N.require_device()
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info(self.dev)
            hbm_budget = int(0.8 * free)
        self.max_slots = min(int(max_slots), N.SLT_MAX_JOBS)
        self.budget = int(hbm_budget)
        if blocks < 8 or blocks % 8:
            raise ValueError("blocks must be a positive multiple of 8")
        self.blocks = int(blocks)
        self.batch_work = float(batch_work)
        # rays that give every wavefront of the persistent launch two 64-ray units
        self.fill_rays = 2 * 64 * (self.blocks * 4)
        self.step_fraction = step_fraction
        self.keep_outputs = keep_outputs
        self.max_wait = float(max_wait)
        self.main = torch.cuda.Stream(self.dev)
        self.setup_stream = torch.cuda.Stream(self.dev)
        tb = ctypes.sizeof(N.JobTable)
        self.table_dev = torch.zeros(tb, dtype=torch.uint8, device=self.dev)
        self._ring = [torch.zeros(tb, dtype=torch.uint8).pin_memory() for _ in range(4)]
        self._ring_ev = [None] * 4
        self._ring_i = 0
        self.graph = None
        self.use_graph = use_graph
        self.batches = 0
        L = N.lib()
        empty = N.JobTable()
        N.check(L.slt_job_table_prepare(ctypes.byref(empty)))
        with torch.cuda.stream(self.main):
            self._upload(empty)
            N.check(L.slt_advance(self.table_dev.data_ptr(), self.blocks, N.stream_handle()))
        self.main.synchronize()
        if use_graph:
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                N.check(L.slt_advance(self.table_dev.data_ptr(), self.blocks, N.stream_handle()))
            torch.cuda.synchronize(self.dev)

    # -- table upload through a ring of pinned buffers ------------------------
    def _upload(self, table: N.JobTable):
        i = self._ring_i
        self._ring_i = (i + 1) % len(self._ring)
        ev = self._ring_ev[i]
        if ev is not None:
            ev.synchronize()
        buf = self._ring[i]
        ctypes.memmove(buf.data_ptr(), ctypes.addressof(table), ctypes.sizeof(table))
        self.table_dev.copy_(buf, non_blocking=True)
        e = torch.cuda.Event()
        e.record(torch.cuda.current_stream())
        self._ring_ev[i] = e

    # -- request setup on the side stream -------------------------------------
    def _setup(self, req: RequestSpec) -> _Active:
        with torch.cuda.stream(self.setup_stream):
            mat = materialize(req, self.dev)
            n = mat.rays.shape[1]
            nsub, h = engine.schedule(mat.save_times, engine.nominal_step(mat.grid, self.step_fraction))
            order = engine.ray_order(mat.grid, mat.rays)
            # the state rows in Morton order, from rays gathered into it (coalesced init);
            # the step batches then write their outputs in the same order
            rays_sorted = engine.gather_rays(mat.rays, order)
            state = torch.empty(int(N.lib().slt_ray_state_bytes(n)), dtype=torch.uint8,
                                device=self.dev)
            N.check(N.lib().slt_ray_state_init(state.data_ptr(), n, rays_sorted.data_ptr(),
                                               None, ctypes.byref(mat.grid.desc),
                                               N.stream_handle()))
            del rays_sorted
            seg_nsub = engine.const_device(nsub, self.dev)
            seg_h = engine.const_device(h.astype(np.float32), self.dev)
            out_state = torch.empty((6, n), dtype=torch.float32, device=self.dev)
            out_tau = torch.zeros(n, dtype=torch.float64, device=self.dev)
            out_path = torch.empty(n, dtype=torch.float64, device=self.dev)
            out_steps = torch.empty(n, dtype=torch.int32, device=self.dev)
            counter = torch.zeros(1, dtype=torch.int64, device=self.dev)
            ready = torch.cuda.Event()
            ready.record(self.setup_stream)
        for t in (mat.grid.voxels, state, order, seg_nsub, seg_h, out_state, out_tau, out_path,
                  out_steps, counter):
            t.record_stream(self.main)
        sched = int(nsub.astype(np.int64).sum())
        # steps a ray needs to cross the box along the probing axis; after that the
        # remaining steps are straight-line moves and finish in the batch that
        # reaches this estimate
        ax = mat.grid.axis
        span = float(mat.grid.coords[ax][-1]) - float(mat.grid.coords[ax][0])
        inbox = int(math.ceil(span / (299792458.0 * float(h[0])))) + 4 if h.size and h[0] > 0 else sched
        return _Active(req, mat, order, state, seg_nsub, seg_h, sched, min(inbox, sched),
                       out_state, out_tau, out_path, out_steps, counter, ready)

    def _job(self, a: _Active, begin: int, end: int) -> N.Job:
        j = N.Job()
        j.packed = a.mat.grid.voxels.data_ptr()
        j.grid = a.mat.grid.desc
        j.state = a.state.data_ptr()
        j.order = None  # outputs in state-row (Morton) order; finish_outputs puts them back
        j.n_rays = a.mat.rays.shape[1]
        j.seg_nsub = a.seg_nsub.data_ptr()
        j.seg_h = a.seg_h.data_ptr()
        j.n_seg = a.seg_nsub.numel()
        j.sched_steps = a.sched
        j.step_begin = begin
        j.step_end = end
        j.out_state = a.out_state.data_ptr()
        j.out_tau = _f(a.out_tau) if a.mat.grid.inv_brems else None
        j.out_path = a.out_path.data_ptr()
        j.out_steps = a.out_steps.data_ptr()
        j.step_counter = a.counter.data_ptr()
        return j

    def _finalize(self, a: _Active):
        inv = a.mat.grid.inv_brems
        a.result = finish_outputs(a.out_state, a.out_tau if inv else None, a.out_path,
                                  a.out_steps, a.order, a.mat.grid.probing_direction,
                                  a.mat.edges, keep=self.keep_outputs)
        a.done = torch.cuda.Event(enable_timing=True)
        a.done.record(self.main)

    def _launch_batch(self, runnable, inflight):
        """Queue one step batch over the set-up requests on the compute stream."""
        n_rays = sum(a.mat.rays.shape[1] for a in runnable)
        k = int(max(16, min(4096, self.batch_work / max(n_rays, 1))))
        table = N.JobTable()
        table.n_jobs = len(runnable)
        spans = []
        for i, a in enumerate(runnable):
            end = min(a.sched, a.cursor + k)
            if end >= a.inbox_estimate:
                # the rays have crossed the box by the end of this batch: the rest of the
                # schedule is one straight-line move per ray, so finish it here
                end = a.sched
            table.jobs[i] = self._job(a, a.cursor, end)
            spans.append(end)
        N.check(N.lib().slt_job_table_prepare(ctypes.byref(table)))
        with torch.cuda.stream(self.main):
            self._upload(table)
            if self.graph is not None:
                self.graph.replay()
            else:
                N.check(N.lib().slt_advance(self.table_dev.data_ptr(), self.blocks,
                                            N.stream_handle()))
            self.batches += 1
            ev = torch.cuda.Event()
            ev.record(self.main)
            inflight.append(ev)
            for a, end in zip(runnable, spans):
                a.cursor = end
                if end >= a.sched:
                    a.finished = True
                    self._finalize(a)

    def _worth_launching(self, runnable, active, pending, inflight) -> bool:
        """Whether to queue a batch over ``runnable`` now.  With no batch left on the GPU,
        always.  Behind a running batch, only a batch that gives every resident wavefront
        work: a smaller one (requests trickling out of setup one at a time) waits, so the
        requests still in setup can join it instead of each running in a batch that leaves
        most of the GPU idle while it drains."""
        if not any(not e.query() for e in inflight):
            return True
        if not pending and all(a.finished or a.ready.query() for a in active):
            return True
        rays = sum(a.mat.rays.shape[1] for a in runnable)
        return rays >= self.fill_rays

    # -- main loop ------------------------------------------------------------
    def run(self, requests, *, realtime: bool = True, max_lookahead: int = 2):
        """Serve ``requests`` (arrival times in seconds from the start).

        Returns (responses by rid, stats dict).  With ``realtime=False`` every
        request is treated as already arrived.
        """
        queue = deque(sorted(requests, key=lambda r: (r.arrival if realtime else 0.0, r.rid)))
        if not realtime:
            for r in queue:
                r.arrival = 0.0
        adm = Admission(self.budget, self.max_slots, self.max_wait)
        responses: dict[int, Response] = {}
        active: list[_Active] = []
        pending = deque()   # admitted, not yet set up
        inflight = deque()
        torch.cuda.synchronize(self.dev)
        t_start = time.perf_counter()
        start_ev = torch.cuda.Event(enable_timing=True)
        start_ev.record(self.main)
        total_steps = 0
        while queue or active or pending:
            now = time.perf_counter() - t_start
            admitted, rejected = adm.admit(queue, now)
            for r in rejected:
                responses[r.rid] = Response(r.rid, "rejected",
                                            f"footprint {r.footprint_bytes()} B exceeds the HBM "
                                            f"budget {self.budget} B", arrival=r.arrival)
            pending.extend(admitted)
            progressed = False
            runnable = [a for a in active if not a.finished and a.ready.query()]
            runnable = runnable[:N.SLT_MAX_JOBS]
            if runnable and self._worth_launching(runnable, active, pending, inflight):
                self._launch_batch(runnable, inflight)
                progressed = True
            # One setup per pass, after the batch is queued: the GPU integrates
            # while the host builds the next request's launches.
            if pending:
                r = pending.popleft()
                try:
                    active.append(self._setup(r))
                except (ValueError, NotImplementedError, N.SlothError) as e:
                    adm.release(r)
                    responses[r.rid] = Response(r.rid, "error", str(e), arrival=r.arrival)
                progressed = True
            for a in [a for a in active if a.finished and a.done is not None and a.done.query()]:
                steps = int(a.counter.item())
                total_steps += steps
                r = a.req
                responses[r.rid] = Response(
                    r.rid, "ok", rf=a.result["rf"], amp=a.result["amp"],
                    image=a.result["image"], path=a.result["path"], state=a.result["state"],
                    tau=a.result["tau"], steps=a.result["steps"], ray_steps=steps,
                    arrival=r.arrival, completed=start_ev.elapsed_time(a.done) / 1e3,
                    n_rays=a.mat.rays.shape[1])
                active.remove(a)
                adm.release(r)
                progressed = True
            while len(inflight) > max_lookahead:
                inflight.popleft().synchronize()
            if not progressed:
                if active:
                    time.sleep(20e-6)
                elif queue:
                    wait = queue[0].arrival - (time.perf_counter() - t_start)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
        torch.cuda.synchronize(self.dev)
        makespan = max([r.completed for r in responses.values() if r.status == "ok"] or [0.0])
        lat = np.array([r.latency for r in responses.values() if r.status == "ok"])
        stats = {
            "requests": len(responses),
            "ok": int(sum(r.status == "ok" for r in responses.values())),
            "rejected": int(sum(r.status == "rejected" for r in responses.values())),
            "ray_steps": total_steps,
            "makespan_s": makespan,
            "ray_steps_per_s": total_steps / makespan if makespan > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else math.nan,
            "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else math.nan,
            "batches": self.batches,
        }
        return responses, stats


def finish_outputs(state, tau, path, steps, order, probing_direction, edges, *, keep=True):
    """A request's results from its final ray states.

    state (6, N), tau (N,) or None, path (N,), steps (N,) int32 are in processing order:
    column t belongs to the caller's ray order[t] (order None: the caller's order).  The
    Jones vectors and the detector image are made in that order, where a wavefront's rays
    share detector cells; with ``keep`` the state, tau, path and steps are put back in the
    caller's order (two 32-byte records per ray) and rf / amp come from the caller-order
    state, which gives the same values as in processing order.
    """
    rf, amp = jones_device(state, probing_direction, tau)
    counts, wsum = histogram2d_device(rf[0], rf[2], edges, edges, weights=amp)
    out = {"image": wsum if wsum is not None else counts, "rf": None, "amp": None,
           "path": None, "state": None, "tau": None, "steps": None}
    if not keep:
        return out
    if order is not None:
        work = engine.ray_work(order.numel(), state.device)
        state, tau = engine.unpermute_rays(state, tau, order, work=work)
        st, path = engine.unpermute_rays(steps.view(torch.float32).unsqueeze(0), path, order,
                                         work=work)
        steps = st[0].view(torch.int32)
        rf, amp = jones_device(state, probing_direction, tau)
    out.update(rf=rf, amp=amp, path=path, state=state, tau=tau, steps=steps)
    return out


def run_offline(requests, *, step_fraction: float = 0.5, device=None):
    """The same requests one after another through the offline path (slt_trace).

    Returns (results by rid, stats).
    """
    dev = device if device is not None else N.require_device()
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    out = {}
    total = 0
    for req in requests:
        mat = materialize(req, dev)
        res = engine.trace(mat.grid, mat.rays, mat.save_times, step_fraction=step_fraction,
                           save_all=False, want_path=True, gather=True)
        fin = finish_outputs(res.states[0], res.tau[0] if res.tau is not None else None,
                             res.path, res.steps, res.order if res.permuted else None,
                             mat.grid.probing_direction, mat.edges)
        out[req.rid] = {"rf": fin["rf"], "amp": fin["amp"], "image": fin["image"],
                        "path": fin["path"], "steps": res.total_steps, "state": fin["state"],
                        "steps_per_ray": fin["steps"], "nsub": res.nsub, "h": res.h}
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    for v in out.values():
        v["steps"] = int(v["steps"].item())
        total += v["steps"]
    return out, {"ray_steps": total, "wall_s": wall, "ray_steps_per_s": total / wall}

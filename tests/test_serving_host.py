"""Serving host logic without a GPU: admission and the job-table checks of the C ABI."""
import ctypes
from collections import deque

import numpy as np
import pytest

from sloth_amd import _native as N
from sloth_amd.serving.engine import Admission
from sloth_amd.serving.requests import BeamSpec, FieldSpec, RequestSpec


def _req(rid, side=64, rays=10_000, arrival=0.0):
    return RequestSpec(rid, FieldSpec(side=side), BeamSpec(n_rays=rays), arrival=arrival)


def test_admission_fifo_budget_slots_and_rejection():
    a, b, c = _req(0, arrival=0.0), _req(1, arrival=0.5), _req(2, side=128, arrival=0.6)
    budget = a.footprint_bytes() + c.footprint_bytes()
    adm = Admission(budget, max_slots=4)
    q = deque([a, b, c])
    got, rej = adm.admit(q, now=0.1)          # only a has arrived
    assert [r.rid for r in got] == [0] and not rej and [r.rid for r in q] == [1, 2]
    got, rej = adm.admit(q, now=1.0)          # b fits beside a; c must wait for memory
    assert [r.rid for r in got] == [1] and not rej and [r.rid for r in q] == [2]
    adm.release(a)
    got, rej = adm.admit(q, now=1.0)
    assert [r.rid for r in got] == [2] and not rej and not q
    assert adm.used == b.footprint_bytes() + c.footprint_bytes() and adm.slots == 2


def test_admission_rejects_oversized_and_keeps_going():
    big = _req(0, side=384, rays=4_000_000)
    small = _req(1)
    adm = Admission(small.footprint_bytes() * 2, max_slots=1)
    q = deque([big, small])
    got, rej = adm.admit(q, now=0.0)
    assert rej == [big] and got == [small]
    q2 = deque([_req(2)])
    assert adm.admit(q2, 0.0) == ([], [])  # slot limit
    adm.release(small)
    assert adm.admit(q2, 0.0)[0][0].rid == 2


def test_admission_passes_a_blocked_head_until_its_age_cap():
    """A request that fits is admitted past an older one waiting for memory (no head-of-line
    blocking) while that one is younger than max_wait; once it is older, nothing behind it
    is admitted, and it goes first as soon as memory frees."""
    running = _req(0, side=128)
    head = _req(1, side=192, arrival=0.0)           # does not fit beside `running`
    s1, s2 = _req(2, arrival=0.01), _req(3, arrival=0.02)
    budget = head.footprint_bytes() + s1.footprint_bytes() // 2
    adm = Admission(budget, max_slots=8, max_wait=0.05)
    assert adm.admit(deque([running]), 0.0)[0] == [running]
    q = deque([head, s1, s2])
    got, rej = adm.admit(q, now=0.03)               # head waits 30 ms: small ones pass it
    assert [r.rid for r in got] == [2, 3] and not rej and list(q) == [head]
    adm.release(s1)
    adm.release(s2)
    s3 = _req(4, arrival=0.04)
    q.append(s3)
    got, _ = adm.admit(q, now=0.06)                 # head has waited 60 ms: hold s3
    assert got == [] and list(q) == [head, s3]
    adm.release(running)                            # memory frees: head first, then s3 fits
    got, _ = adm.admit(q, now=0.07)
    assert [r.rid for r in got] == [1] and list(q) == [s3]
    assert adm.used == head.footprint_bytes() and adm.slots == 1
    # an oversized request anywhere in the arrived queue is rejected without blocking
    huge = _req(5, side=384, rays=4_000_000, arrival=0.08)
    q = deque([huge, _req(6, arrival=0.09)])
    got, rej = Admission(head.footprint_bytes(), 4).admit(q, now=0.1)
    assert rej == [huge] and [r.rid for r in got] == [6]
    with pytest.raises(ValueError):
        Admission(1, 1, max_wait=-1.0)


def _grid_desc(n=8):
    c = np.float32(np.linspace(0, 1e-3, n))
    d = N.GridDesc()
    assert N.lib().slt_grid_desc_init(c.ctypes.data, c.ctypes.data, c.ctypes.data, n, n, n, 2, 1,
                                      ctypes.byref(d)) == 0
    return d


def _valid_job(n_rays=1000):
    j = N.Job()
    j.packed = 0x1000
    j.grid = _grid_desc()
    j.state = 0x2000
    j.n_rays = n_rays
    j.seg_nsub = 0x3000
    j.seg_h = 0x4000
    j.n_seg = 1
    j.sched_steps = 100
    j.step_begin = 0
    j.step_end = 50
    return j


def test_job_table_prepare_units_and_errors():
    L = N.lib()
    t = N.JobTable()
    t.n_jobs = 3
    for i, n in enumerate((1000, 256, 1)):
        t.jobs[i] = _valid_job(n)
    t.next_unit[3][0] = 7
    t.waves_done[0] = 9
    assert L.slt_job_table_prepare(ctypes.byref(t)) == N.SLT_OK
    # one unit = 64 rays (a wavefront); the work queues start empty
    assert list(t.unit_begin[:4]) == [0, 16, 20, 21] and t.n_units == 21
    assert t.next_unit[3][0] == 0 and t.waves_done[0] == 0
    # the queues sit on their own 128-byte lines ahead of the job list
    assert N.JobTable.next_unit.offset == 0 and N.JobTable.waves_done.offset == 1024
    assert N.JobTable.jobs.offset % 8 == 0
    # a finishing batch without an output buffer is refused
    t.jobs[1].step_end = 100
    assert L.slt_job_table_prepare(ctypes.byref(t)) == N.SLT_ERR_INVALID
    assert "job 1" in N.last_error() and "out_state" in N.last_error()
    t.jobs[1].step_end = 50
    # a step range past the schedule is refused
    t.jobs[2].step_end = 101
    assert L.slt_job_table_prepare(ctypes.byref(t)) == N.SLT_ERR_INVALID
    t.jobs[2].step_end = 10
    # a grid with a plane stride beyond 24 bits is unsupported by the kernel
    t.jobs[0].grid.stride[0] = 1 << 24
    assert L.slt_job_table_prepare(ctypes.byref(t)) == N.SLT_ERR_UNSUPPORTED
    t.n_jobs = N.SLT_MAX_JOBS + 1
    assert L.slt_job_table_prepare(ctypes.byref(t)) == N.SLT_ERR_INVALID
    with pytest.raises(ValueError):
        N.check(N.SLT_ERR_INVALID)


def test_batch_launch_policy():
    """A batch goes out at once when the GPU has none left or nothing is still in setup;
    behind a running batch only when it fills every resident wavefront."""
    from types import SimpleNamespace as NS

    from sloth_amd.serving.engine import ServingEngine

    class Ev:
        def __init__(self, done):
            self.done = done

        def query(self):
            return self.done

    def act(rays, ready=True, finished=False):
        return NS(mat=NS(rays=np.zeros((6, rays), np.float32)), ready=Ev(ready),
                  finished=finished)

    eng = NS(fill_rays=2 * 64 * 1024 * 4)
    go = lambda run, active, pending, inflight: ServingEngine._worth_launching(  # noqa: E731
        eng, run, active, pending, inflight)
    small, big = act(10_000), act(600_000)
    setting_up = act(50_000, ready=False)
    # nothing on the GPU: always
    assert go([small], [small, setting_up], deque(), deque())
    assert go([small], [small], deque([_req(9)]), deque([Ev(True)]))
    # a batch still running, requests still in setup: only a batch that fills the GPU
    running = deque([Ev(True), Ev(False)])
    assert not go([small], [small, setting_up], deque(), running)
    assert not go([small], [small], deque([_req(9)]), running)
    assert go([big], [big, setting_up], deque(), running)
    assert go([small, big], [small, big, setting_up], deque(), running)
    # a batch still running, but every request is set up: no reason to wait
    assert go([small], [small, act(5, finished=True)], deque(), running)

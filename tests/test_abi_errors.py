"""C-ABI argument checks, on the CPU: every entry point of include/sloth_hip.h rejects bad
arguments with its status code and a message before it touches device memory, and returns
SLT_OK without a launch for zero rays.  (Compute paths run in the gpu tests.)

The pointers passed as "device" buffers here are never dereferenced: each call fails its
checks, or has nothing to do, before any use of them.
"""
import ctypes

import numpy as np
import pytest

from sloth_amd import _native as N

FAKE = 0x100000  # a non-null "device" pointer that the checks below never reach


@pytest.fixture(scope="module")
def L():
    return N.lib()


def _err(L):
    msg = L.slt_last_error()
    return msg.decode() if msg else ""


def _desc(L, n=(8, 9, 10), axis=2, nchan=4):
    c = [np.float32(np.linspace(-1e-3, 1e-3, k)) for k in n]
    d = N.GridDesc()
    rc = L.slt_grid_desc_init(c[0].ctypes.data, c[1].ctypes.data, c[2].ctypes.data, *n, axis,
                              1 if nchan == 4 else 0, ctypes.byref(d))
    assert rc == N.SLT_OK, _err(L)
    return d


def _copy(d):
    e = N.GridDesc()
    ctypes.memmove(ctypes.addressof(e), ctypes.addressof(d), ctypes.sizeof(d))
    return e


def _expect(L, rc, code, name):
    assert rc == code, (rc, _err(L))
    if code != N.SLT_OK:
        assert _err(L), "an error status must come with a message"
        if name:
            assert name in _err(L) or "grid" in _err(L), _err(L)
    else:
        assert _err(L) == "", "a successful call clears the last error"


def test_grid_pack_rejections(L):
    g = _desc(L)
    args = dict(ne=FAKE, cx=FAKE, cy=FAKE, cz=FAKE, te=FAKE, n_te=1, zz=FAKE, n_z=1,
                packed=FAKE, pdx=None)

    def call(g=g, **kw):
        a = {**args, **kw}
        return L.slt_grid_pack(a["ne"], a["cx"], a["cy"], a["cz"], -1.0, a["te"], a["n_te"],
                               a["zz"], a["n_z"], 1e15, 1.6e-19, 8.85e-12, 3e8, 0.0,
                               ctypes.byref(g), a["packed"], a["pdx"], None, None, None, None)

    _expect(L, call(ne=None), N.SLT_ERR_INVALID, "slt_grid_pack")
    _expect(L, call(te=FAKE, zz=None), N.SLT_ERR_INVALID, "slt_grid_pack")
    _expect(L, call(n_te=5), N.SLT_ERR_INVALID, "slt_grid_pack")
    _expect(L, call(packed=None), N.SLT_ERR_INVALID, "slt_grid_pack")   # no output at all
    short = _copy(g)
    short.n[0] = 1
    _expect(L, call(g=short), N.SLT_ERR_INVALID, "slt_grid_pack")
    flat = _copy(g)
    flat.hi[1] = flat.lo[1]
    _expect(L, call(g=flat), N.SLT_ERR_INVALID, "slt_grid_pack")


def test_kappa_rejections(L):
    def call(n, n_te=1, ne=FAKE, out=FAKE):
        return L.slt_kappa(ne, n, FAKE, n_te, FAKE, 1, 1e15, 1.6e-19, 8.85e-12, 3e8, 0.0, out,
                           None)

    _expect(L, call(-1), N.SLT_ERR_INVALID, "slt_kappa")
    _expect(L, call(10, n_te=3), N.SLT_ERR_INVALID, "slt_kappa")
    _expect(L, call(10, ne=None), N.SLT_ERR_INVALID, "slt_kappa")
    _expect(L, call(0), N.SLT_OK, None)


def test_trace_rejections(L):
    g = _desc(L)

    def targs(**kw):
        a = N.TraceArgs()
        a.n_rays, a.s0, a.n_seg, a.seg_nsub, a.seg_h, a.save_state = 10, FAKE, 1, FAKE, FAKE, FAKE
        for k, v in kw.items():
            setattr(a, k, v)
        return a

    _expect(L, L.slt_trace(None, ctypes.byref(g), ctypes.byref(targs()), None),
            N.SLT_ERR_INVALID, None)
    _expect(L, L.slt_trace(FAKE, ctypes.byref(g), None, None), N.SLT_ERR_INVALID, "slt_trace")
    for bad in (dict(n_rays=-1), dict(s0=None), dict(n_seg=0), dict(save_state=None),
                dict(seg_h=None)):
        _expect(L, L.slt_trace(FAKE, ctypes.byref(g), ctypes.byref(targs(**bad)), None),
                N.SLT_ERR_INVALID, "slt_trace")
    big = _copy(g)
    big.n[0], big.n[1], big.n[2] = 2048, 1024, 1024      # 2^31 voxels
    _expect(L, L.slt_trace(FAKE, ctypes.byref(big), ctypes.byref(targs()), None),
            N.SLT_ERR_UNSUPPORTED, None)
    wide = _copy(g)
    wide.stride[0] = 1 << 24
    _expect(L, L.slt_trace(FAKE, ctypes.byref(wide), ctypes.byref(targs()), None),
            N.SLT_ERR_UNSUPPORTED, None)
    _expect(L, L.slt_trace(FAKE, ctypes.byref(g), ctypes.byref(targs(n_rays=0)), None),
            N.SLT_OK, None)


def test_jones_histogram_morton_rejections(L):
    _expect(L, L.slt_jones(FAKE, 10, -1, 2, FAKE, None, None, None), N.SLT_ERR_INVALID,
            "slt_jones")
    _expect(L, L.slt_jones(FAKE, 10, 10, 3, FAKE, None, None, None), N.SLT_ERR_INVALID,
            "slt_jones")
    _expect(L, L.slt_jones(None, 10, 10, 2, FAKE, None, None, None), N.SLT_ERR_INVALID,
            "slt_jones")
    _expect(L, L.slt_jones(FAKE, 5, 10, 2, FAKE, None, None, None), N.SLT_ERR_INVALID,
            "slt_jones")   # row stride below n
    _expect(L, L.slt_jones(None, 0, 0, 2, None, None, None, None), N.SLT_OK, None)

    def hist(n=10, nxb=4, xe=FAKE, counts=FAKE, wsum=None, weights=None, xs=1):
        return L.slt_histogram2d(FAKE, xs, FAKE, 1, 0, n, None, None, xe, nxb, FAKE, 4, weights,
                                 counts, wsum, None)

    _expect(L, hist(nxb=0), N.SLT_ERR_INVALID, "slt_histogram2d")
    _expect(L, hist(xe=None), N.SLT_ERR_INVALID, "slt_histogram2d")
    _expect(L, hist(counts=None), N.SLT_ERR_INVALID, "slt_histogram2d")
    _expect(L, hist(wsum=FAKE), N.SLT_ERR_INVALID, "slt_histogram2d")   # sums need weights
    _expect(L, hist(xs=-1), N.SLT_ERR_INVALID, "slt_histogram2d")
    _expect(L, hist(n=-1), N.SLT_ERR_INVALID, "slt_histogram2d")
    _expect(L, hist(n=0), N.SLT_OK, None)

    g = _desc(L)
    _expect(L, L.slt_morton_keys(FAKE, 10, None, FAKE, None), N.SLT_ERR_INVALID,
            "slt_morton_keys")
    _expect(L, L.slt_morton_keys(None, 10, ctypes.byref(g), FAKE, None), N.SLT_ERR_INVALID,
            "slt_morton_keys")
    bad = _copy(g)
    bad.probing_axis = 3
    _expect(L, L.slt_morton_keys(FAKE, 10, ctypes.byref(bad), FAKE, None), N.SLT_ERR_INVALID,
            "slt_morton_keys")
    _expect(L, L.slt_morton_keys(None, 0, ctypes.byref(g), None, None), N.SLT_OK, None)


def test_ray_permutation_rejections(L):
    other, work = FAKE + (1 << 20), FAKE + (1 << 40)
    assert L.slt_ray_work_bytes(10) == 320 and L.slt_ray_work_bytes(0) == 0
    assert L.slt_ray_work_bytes(-3) == 0
    _expect(L, L.slt_gather_rays(FAKE, -1, FAKE, other, work, None), N.SLT_ERR_INVALID,
            "slt_gather_rays")
    for s0, order, out, w in ((None, FAKE, other, work), (FAKE, None, other, work),
                              (FAKE, other, None, work), (FAKE, other, FAKE, None)):
        _expect(L, L.slt_gather_rays(s0, 10, order, out, w, None), N.SLT_ERR_INVALID,
                "slt_gather_rays")
    # the records would overwrite rays before they are read
    _expect(L, L.slt_gather_rays(FAKE, 10, other, FAKE + 1000, FAKE + 100, None),
            N.SLT_ERR_INVALID, "overlaps")
    _expect(L, L.slt_gather_rays(None, 0, None, None, None, None), N.SLT_OK, None)

    def unperm(src=FAKE, rows=4, src64=None, n=10, order=other, dst=FAKE + 4096, dst64=None,
               w=work):
        return L.slt_unpermute_rays(src, rows, src64, n, order, dst, dst64, w, None)

    for bad in (dict(rows=9), dict(rows=7, src64=FAKE + 8192, dst64=FAKE + 9000),
                dict(rows=0), dict(rows=-1), dict(src64=FAKE + 8192),        # no dst64
                dict(dst64=FAKE + 8192), dict(order=None), dict(w=None), dict(src=None),
                dict(dst=None), dict(n=-1)):
        _expect(L, unperm(**bad), N.SLT_ERR_INVALID, "slt_unpermute_rays")
    _expect(L, unperm(w=FAKE + 8), N.SLT_ERR_INVALID, "overlaps")
    _expect(L, unperm(n=0), N.SLT_OK, None)


def test_rtm_and_beam_rejections(L):
    ops = (N.Optic * (N.SLT_MAX_OPTICS + 1))()
    _expect(L, L.slt_rtm(FAKE, 0, 10, 10, 1.0, ops, N.SLT_MAX_OPTICS + 1, FAKE, 10, None),
            N.SLT_ERR_INVALID, "slt_rtm")
    _expect(L, L.slt_rtm(FAKE, 0, 10, 10, 1.0, None, 1, FAKE, 10, None), N.SLT_ERR_INVALID,
            "slt_rtm")
    ops[0].kind = 99
    _expect(L, L.slt_rtm(FAKE, 0, 10, 10, 1.0, ops, 1, FAKE, 10, None), N.SLT_ERR_INVALID,
            "slt_rtm")
    ops[0].kind = N.SLT_OPTIC_TRAVEL
    _expect(L, L.slt_rtm(None, 0, 10, 10, 1.0, ops, 1, FAKE, 10, None), N.SLT_ERR_INVALID,
            "slt_rtm")
    _expect(L, L.slt_rtm(FAKE, 0, 10, 10, 1.0, ops, 1, FAKE, 5, None), N.SLT_ERR_INVALID,
            "slt_rtm")
    _expect(L, L.slt_rtm(None, 0, 0, 0, 1.0, ops, 1, None, 0, None), N.SLT_OK, None)

    def beam(n=10, kind=N.SLT_BEAM_SQUARE, axis=2, s0=FAKE):
        return L.slt_beam(n, kind, axis, 1e-3, 1e-3, 0.0, 5e-3, 42, s0, None)

    _expect(L, beam(kind=7), N.SLT_ERR_INVALID, "slt_beam")
    _expect(L, beam(axis=-1), N.SLT_ERR_INVALID, "slt_beam")
    _expect(L, beam(s0=None), N.SLT_ERR_INVALID, "slt_beam")
    _expect(L, beam(n=-1), N.SLT_ERR_INVALID, "slt_beam")
    _expect(L, beam(n=0, s0=None), N.SLT_OK, None)


def test_serving_state_and_advance_rejections(L):
    g = _desc(L)
    _expect(L, L.slt_ray_state_init(FAKE, 10, FAKE, None, None, None), N.SLT_ERR_INVALID,
            "slt_ray_state_init")
    _expect(L, L.slt_ray_state_init(None, 10, FAKE, None, ctypes.byref(g), None),
            N.SLT_ERR_INVALID, "slt_ray_state_init")
    short = _copy(g)
    short.n[2] = 1
    _expect(L, L.slt_ray_state_init(FAKE, 10, FAKE, None, ctypes.byref(short), None),
            N.SLT_ERR_INVALID, "slt_ray_state_init")
    _expect(L, L.slt_ray_state_init(None, 0, None, None, ctypes.byref(g), None), N.SLT_OK, None)
    _expect(L, L.slt_advance(None, 1024, None), N.SLT_ERR_INVALID, "slt_advance")
    for blocks in (0, 4, 12, -8):
        _expect(L, L.slt_advance(FAKE, blocks, None), N.SLT_ERR_INVALID, "slt_advance")


def test_adjoint_rejections(L):
    g = _desc(L)

    def aargs(**kw):
        a = N.AdjointArgs()
        a.n_rays, a.s0, a.n_seg, a.ckpt_every = 10, FAKE, 1, 16
        a.seg_nsub, a.seg_h, a.sched_steps, a.checkpoints = FAKE, FAKE, 100, FAKE
        a.final_state, a.seed_state, a.grid_adj, a.tape = FAKE, FAKE, FAKE, FAKE
        for k, v in kw.items():
            setattr(a, k, v)
        return a

    for fn, name in ((L.slt_forward_ckpt, "slt_forward_ckpt"), (L.slt_adjoint, "slt_adjoint")):
        for bad in (dict(ckpt_every=-1), dict(n_seg=0), dict(checkpoints=None), dict(s0=None),
                    dict(n_rays=-1), dict(seg_h=None), dict(sched_steps=-1)):
            _expect(L, fn(FAKE, ctypes.byref(g), ctypes.byref(aargs(**bad)), None),
                    N.SLT_ERR_INVALID, name)
        _expect(L, fn(None, ctypes.byref(g), ctypes.byref(aargs()), None), N.SLT_ERR_INVALID,
                None)
        _expect(L, fn(FAKE, ctypes.byref(g), ctypes.byref(aargs(n_rays=0)), None), N.SLT_OK,
                None)
    _expect(L, L.slt_forward_ckpt(FAKE, ctypes.byref(g),
                                  ctypes.byref(aargs(final_state=None)), None),
            N.SLT_ERR_INVALID, "slt_forward_ckpt")
    for missing in ("seed_state", "grid_adj", "tape"):
        _expect(L, L.slt_adjoint(FAKE, ctypes.byref(g), ctypes.byref(aargs(**{missing: None})),
                                 None), N.SLT_ERR_INVALID, "slt_adjoint")
    # ckpt_every = 0: the forward records the trajectory, so it needs the tape too, and
    # a schedule longer than a tape's record index is refused
    _expect(L, L.slt_forward_ckpt(FAKE, ctypes.byref(g),
                                  ctypes.byref(aargs(ckpt_every=0, tape=None)), None),
            N.SLT_ERR_INVALID, "a->tape")
    _expect(L, L.slt_adjoint(FAKE, ctypes.byref(g),
                             ctypes.byref(aargs(ckpt_every=0, tape=None)), None),
            N.SLT_ERR_INVALID, "tape are required")
    for fn in (L.slt_forward_ckpt, L.slt_adjoint):
        _expect(L, fn(FAKE, ctypes.byref(g), ctypes.byref(aargs(ckpt_every=0,
                                                                sched_steps=2 ** 31)), None),
                N.SLT_ERR_UNSUPPORTED, "moves per ray")

    # the allocation-free reverse pass: its tape is the library's, so a caller's
    # tape and a tape beyond the scratch are refused before any device work
    noalloc = L.slt_adjoint_noalloc
    _expect(L, noalloc(FAKE, ctypes.byref(g), ctypes.byref(aargs()), None), N.SLT_ERR_INVALID,
            "tape must be NULL")
    for missing in ("seed_state", "grid_adj"):
        _expect(L, noalloc(FAKE, ctypes.byref(g),
                           ctypes.byref(aargs(tape=None, **{missing: None})), None),
                N.SLT_ERR_INVALID, "slt_adjoint_noalloc")
    for K in (1, 4, 16, 1000):
        n = N.SLT_ADJ_SCRATCH_BYTES // int(L.slt_tape_bytes(1, K)) + 1
        assert int(L.slt_tape_bytes(n, K)) > N.SLT_ADJ_SCRATCH_BYTES
        _expect(L, noalloc(FAKE, ctypes.byref(g),
                           ctypes.byref(aargs(tape=None, n_rays=n, ckpt_every=K)), None),
                N.SLT_ERR_UNSUPPORTED, "scratch")
    for n, K in ((2 ** 37, 2 ** 30),):
        _expect(L, noalloc(FAKE, ctypes.byref(g),
                           ctypes.byref(aargs(tape=None, n_rays=n, ckpt_every=K)), None),
                N.SLT_ERR_UNSUPPORTED, "scratch")
    _expect(L, noalloc(FAKE, ctypes.byref(g), ctypes.byref(aargs(tape=None, n_rays=0)), None),
            N.SLT_OK, None)
    _expect(L, noalloc(FAKE, ctypes.byref(g), ctypes.byref(aargs(tape=None, ckpt_every=0)),
                       None), N.SLT_ERR_UNSUPPORTED, "trajectory tape")

    def unpack(adj=FAKE, te=None, zz=None, n_te=0, g=g):
        return L.slt_adjoint_unpack(adj, ctypes.byref(g), FAKE, FAKE, FAKE, FAKE, -1.0, te, n_te,
                                    zz, 1, 1e15, 1.6e-19, 8.85e-12, 3e8, 0.0, FAKE, None)

    _expect(L, unpack(adj=None), N.SLT_ERR_INVALID, "slt_adjoint_unpack")
    _expect(L, unpack(te=FAKE, n_te=1), N.SLT_ERR_INVALID, "slt_adjoint_unpack")   # Te, no Z
    _expect(L, unpack(te=FAKE, zz=FAKE, n_te=7), N.SLT_ERR_INVALID, "slt_adjoint_unpack")
    short = _copy(g)
    short.n[1] = 1
    _expect(L, unpack(g=short), N.SLT_ERR_INVALID, "slt_adjoint_unpack")


def test_nonuniform_grids_only_where_supported(L):
    """A non-uniform grid needs its knots for slt_trace and the adjoint entry points;
    the serving step batches, built on index arithmetic, refuse it before any device
    work."""
    g = _desc(L)
    nu = _copy(g)
    nu.uniform, nu.knots = 0, None
    a = N.TraceArgs()
    a.n_rays, a.s0, a.n_seg, a.seg_nsub, a.seg_h, a.save_state = 0, FAKE, 1, FAKE, FAKE, FAKE
    _expect(L, L.slt_trace(FAKE, ctypes.byref(nu), ctypes.byref(a), None), N.SLT_ERR_INVALID,
            "knots")
    nu.knots = FAKE
    _expect(L, L.slt_trace(FAKE, ctypes.byref(nu), ctypes.byref(a), None), N.SLT_OK, None)
    aa = N.AdjointArgs()
    aa.n_rays, aa.s0, aa.n_seg, aa.ckpt_every = 10, FAKE, 1, 16
    aa.seg_nsub, aa.seg_h, aa.sched_steps, aa.checkpoints = FAKE, FAKE, 100, FAKE
    aa.final_state, aa.seed_state, aa.grid_adj, aa.tape = FAKE, FAKE, FAKE, FAKE
    nu.knots = None
    for fn in (L.slt_forward_ckpt, L.slt_adjoint):
        _expect(L, fn(FAKE, ctypes.byref(nu), ctypes.byref(aa), None), N.SLT_ERR_INVALID,
                "knots")
    aa.tape = None
    _expect(L, L.slt_adjoint_noalloc(FAKE, ctypes.byref(nu), ctypes.byref(aa), None),
            N.SLT_ERR_INVALID, "knots")
    # with its knots the grid is accepted (no rays: nothing is launched)
    nu.knots = FAKE
    aa.n_rays = 0
    _expect(L, L.slt_adjoint_noalloc(FAKE, ctypes.byref(nu), ctypes.byref(aa), None),
            N.SLT_OK, None)
    aa.tape = FAKE
    for fn in (L.slt_forward_ckpt, L.slt_adjoint):
        _expect(L, fn(FAKE, ctypes.byref(nu), ctypes.byref(aa), None), N.SLT_OK, None)
    t = N.JobTable()
    t.n_jobs = 1
    j = t.jobs[0]
    j.packed, j.grid, j.state, j.n_rays = FAKE, nu, FAKE, 10
    j.seg_nsub, j.seg_h, j.n_seg, j.sched_steps, j.step_begin, j.step_end = FAKE, FAKE, 1, 10, 0, 10
    j.out_state, j.out_path, j.out_steps = FAKE, FAKE, FAKE
    _expect(L, L.slt_job_table_prepare(ctypes.byref(t)), N.SLT_ERR_UNSUPPORTED,
            "uniformly spaced")


def test_sizes_of_invalid_requests_are_zero(L):
    assert L.slt_ray_state_bytes(-1) == 0
    assert L.slt_checkpoint_bytes(-1, 10, 16) == 0
    assert L.slt_checkpoint_bytes(10, 10, -1) == 0
    assert L.slt_tape_bytes(10, 0) == 0
    # ckpt_every = 0: a record count per ray, and a tape of one record per ray and step
    assert L.slt_checkpoint_bytes(10, 10, 0) == 4 * 10
    assert L.slt_trajectory_tape_bytes(10, 7) == 32 * 10 * 7
    assert L.slt_trajectory_tape_bytes(2_000_000, 1443) == 92_352_000_000
    assert L.slt_trajectory_tape_bytes(-1, 7) == 0
    assert L.slt_trajectory_tape_bytes(10, -1) == 0
    assert L.slt_trajectory_tape_bytes(10, 2 ** 31) == 0
    assert L.slt_trajectory_tape_bytes(2 ** 40, 2 ** 30) == 0  # beyond int64
    assert L.slt_ray_state_bytes(10) > 0
    # one 32-byte record per ray and recorded step; at most a fixed number of steps are recorded
    assert L.slt_tape_bytes(10, 1) == 32 * 10
    assert L.slt_tape_bytes(10, 16) % (32 * 10) == 0 and L.slt_tape_bytes(10, 16) <= 16 * 32 * 10


def test_valu_calibrate_rejections(L):
    for kind, blocks, iters, out in ((0, 0, 10, FAKE), (1, 8, 0, FAKE), (2, 8, 10, None),
                                     (0, -1, 10, FAKE), (6, 8, 10, FAKE), (-1, 8, 10, FAKE)):
        _expect(L, L.slt_valu_calibrate(kind, blocks, iters, out, None), N.SLT_ERR_INVALID,
                "slt_valu_calibrate")

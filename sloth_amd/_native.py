"""ctypes binding of libsloth_hip.so (the C ABI declared in include/sloth_hip.h).

The product path has no CPU fallback: if the library is missing, or no HIP
device is visible when a kernel is called, these helpers raise.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libsloth_hip.so")

SLT_OK = 0
SLT_ABI_VERSION = 3
SLT_ERR_INVALID = -1
SLT_ERR_HIP = -2
SLT_ERR_UNSUPPORTED = -3

# Every symbol include/sloth_hip.h declares (tests check the library exports them).
EXPORTED = (
    "slt_version", "slt_last_error", "slt_grid_desc_init", "slt_grid_knots", "slt_grid_bytes",
    "slt_grid_pack",
    "slt_kappa", "slt_trace", "slt_jones", "slt_histogram2d", "slt_morton_keys",
    "slt_ray_state_bytes", "slt_ray_state_init", "slt_job_table_prepare", "slt_advance",
    "slt_checkpoint_bytes", "slt_tape_bytes", "slt_forward_ckpt", "slt_adjoint",
    "slt_adjoint_unpack", "slt_rtm", "slt_beam", "slt_adjoint_noalloc", "slt_valu_calibrate",
    "slt_ray_work_bytes", "slt_gather_rays", "slt_unpermute_rays", "slt_trajectory_tape_bytes",
)

SLT_MAX_JOBS = 64
SLT_ADJ_SCRATCH_BYTES = 128 << 20
SLT_CALIB_CHAINS = 16
SLT_CALIB_FORMS = {"fma_sgpr": 0, "fma_vgpr": 1, "fmac": 2, "add_f32": 3, "mul_f32": 4,
                   "fma_f64": 5}
SLT_WORK_QUEUES = 8
SLT_MAX_OPTICS = 32
SLT_OPTIC_TRAVEL = 0
SLT_OPTIC_LENS = 1
SLT_OPTIC_CIRCULAR_APERTURE = 2
SLT_OPTIC_CIRCULAR_STOP = 3
SLT_OPTIC_RECT_APERTURE = 4
SLT_OPTIC_KNIFE_EDGE = 5
SLT_BEAM_CIRCULAR, SLT_BEAM_SQUARE, SLT_BEAM_RECTANGULAR, SLT_BEAM_LINEAR = 0, 1, 2, 3


class GridDesc(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32 * 3),
        ("nchan", ctypes.c_int32),
        ("stride", ctypes.c_int64 * 3),
        ("lo", ctypes.c_float * 3),
        ("hi", ctypes.c_float * 3),
        ("inv_d", ctypes.c_float * 3),
        ("probing_axis", ctypes.c_int32),
        ("uniform", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("knots", ctypes.c_void_p),
    ]


class TraceArgs(ctypes.Structure):
    _fields_ = [
        ("n_rays", ctypes.c_int64),
        ("s0", ctypes.c_void_p),
        ("order", ctypes.c_void_p),
        ("n_seg", ctypes.c_int32),
        ("save_every", ctypes.c_int32),
        ("seg_nsub", ctypes.c_void_p),
        ("seg_h", ctypes.c_void_p),
        ("save_state", ctypes.c_void_p),
        ("save_tau", ctypes.c_void_p),
        ("path", ctypes.c_void_p),
        ("steps", ctypes.c_void_p),
        ("total_steps", ctypes.c_void_p),
    ]


class AdjointArgs(ctypes.Structure):
    _fields_ = [
        ("n_rays", ctypes.c_int64),
        ("s0", ctypes.c_void_p),
        ("order", ctypes.c_void_p),
        ("n_seg", ctypes.c_int32),
        ("ckpt_every", ctypes.c_int32),
        ("seg_nsub", ctypes.c_void_p),
        ("seg_h", ctypes.c_void_p),
        ("sched_steps", ctypes.c_int64),
        ("checkpoints", ctypes.c_void_p),
        ("final_state", ctypes.c_void_p),
        ("final_tau", ctypes.c_void_p),
        ("seed_state", ctypes.c_void_p),
        ("seed_tau", ctypes.c_void_p),
        ("grid_adj", ctypes.c_void_p),
        ("tape", ctypes.c_void_p),
        ("total_steps", ctypes.c_void_p),
    ]


class Job(ctypes.Structure):
    _fields_ = [
        ("packed", ctypes.c_void_p),
        ("grid", GridDesc),
        ("state", ctypes.c_void_p),
        ("order", ctypes.c_void_p),
        ("n_rays", ctypes.c_int64),
        ("seg_nsub", ctypes.c_void_p),
        ("seg_h", ctypes.c_void_p),
        ("n_seg", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("sched_steps", ctypes.c_int64),
        ("step_begin", ctypes.c_int64),
        ("step_end", ctypes.c_int64),
        ("out_state", ctypes.c_void_p),
        ("out_tau", ctypes.c_void_p),
        ("out_path", ctypes.c_void_p),
        ("out_steps", ctypes.c_void_p),
        ("step_counter", ctypes.c_void_p),
    ]


class Optic(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("axis", ctypes.c_int32),
                ("a", ctypes.c_double), ("b", ctypes.c_double)]


class JobTable(ctypes.Structure):
    _fields_ = [
        ("next_unit", (ctypes.c_uint32 * 32) * SLT_WORK_QUEUES),
        ("waves_done", ctypes.c_uint32 * 32),
        ("n_jobs", ctypes.c_int32),
        ("n_units", ctypes.c_int32),
        ("unit_begin", ctypes.c_int32 * (SLT_MAX_JOBS + 1)),
        ("reserved", ctypes.c_int32),
        ("jobs", Job * SLT_MAX_JOBS),
    ]


_lib = None


class SlothError(RuntimeError):
    """A HIP runtime failure reported by libsloth_hip."""


def lib():
    """Load libsloth_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -m sloth_amd.build, or __graft_entry__.build())")
    # torch first: the library then binds to the HIP runtime torch has loaded,
    # so both share one runtime (and one device context) in the process.
    import torch  # noqa: F401

    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    L.slt_version.restype = ctypes.c_int
    L.slt_last_error.restype = ctypes.c_char_p
    if L.slt_version() != SLT_ABI_VERSION:
        raise ImportError(f"{LIB_PATH} has ABI version {L.slt_version()}, these bindings "
                          f"expect {SLT_ABI_VERSION}: rebuild the HIP library")
    L.slt_grid_desc_init.argtypes = [vp, vp, vp, i32, i32, i32, i32, i32, ctypes.POINTER(GridDesc)]
    L.slt_grid_knots.argtypes = [vp, vp, vp, ctypes.POINTER(GridDesc), vp]
    L.slt_grid_bytes.argtypes = [ctypes.POINTER(GridDesc)]
    L.slt_grid_bytes.restype = i64
    L.slt_grid_pack.argtypes = [vp, vp, vp, vp, dbl, vp, i64, vp, i64, dbl, dbl, dbl, dbl, dbl,
                                ctypes.POINTER(GridDesc), vp, vp, vp, vp, vp, vp]
    L.slt_kappa.argtypes = [vp, i64, vp, i64, vp, i64, dbl, dbl, dbl, dbl, dbl, vp, vp]
    L.slt_trace.argtypes = [vp, ctypes.POINTER(GridDesc), ctypes.POINTER(TraceArgs), vp]
    L.slt_jones.argtypes = [vp, i64, i64, i32, vp, vp, vp, vp]
    L.slt_ray_work_bytes.argtypes = [i64]
    L.slt_ray_work_bytes.restype = i64
    L.slt_gather_rays.argtypes = [vp, i64, vp, vp, vp, vp]
    L.slt_unpermute_rays.argtypes = [vp, i32, vp, i64, vp, vp, vp, vp, vp]
    L.slt_histogram2d.argtypes = [vp, i64, vp, i64, i32, i64, vp, vp, vp, i32, vp, i32, vp, vp, vp,
                                  vp]
    L.slt_morton_keys.argtypes = [vp, i64, ctypes.POINTER(GridDesc), vp, vp]
    L.slt_rtm.argtypes = [vp, i32, i64, i64, dbl, ctypes.POINTER(Optic), i32, vp, i64, vp]
    L.slt_beam.argtypes = [i64, i32, i32, dbl, dbl, dbl, dbl, ctypes.c_uint64, vp, vp]
    L.slt_ray_state_bytes.argtypes = [i64]
    L.slt_ray_state_bytes.restype = i64
    L.slt_ray_state_init.argtypes = [vp, i64, vp, vp, ctypes.POINTER(GridDesc), vp]
    L.slt_job_table_prepare.argtypes = [ctypes.POINTER(JobTable)]
    L.slt_advance.argtypes = [vp, i32, vp]
    L.slt_checkpoint_bytes.argtypes = [i64, i64, i32]
    L.slt_checkpoint_bytes.restype = i64
    L.slt_tape_bytes.argtypes = [i64, i32]
    L.slt_tape_bytes.restype = i64
    L.slt_trajectory_tape_bytes.argtypes = [i64, i64]
    L.slt_trajectory_tape_bytes.restype = i64
    L.slt_forward_ckpt.argtypes = [vp, ctypes.POINTER(GridDesc), ctypes.POINTER(AdjointArgs), vp]
    L.slt_adjoint.argtypes = [vp, ctypes.POINTER(GridDesc), ctypes.POINTER(AdjointArgs), vp]
    L.slt_adjoint_noalloc.argtypes = [vp, ctypes.POINTER(GridDesc), ctypes.POINTER(AdjointArgs),
                                      vp]
    L.slt_valu_calibrate.argtypes = [i32, i32, i32, vp, vp]
    L.slt_adjoint_unpack.argtypes = [vp, ctypes.POINTER(GridDesc), vp, vp, vp, vp, dbl, vp, i64,
                                     vp, i64, dbl, dbl, dbl, dbl, dbl, vp, vp]
    for name in EXPORTED:
        getattr(L, name).restype = getattr(L, name).restype or ctypes.c_int
    _lib = L
    return L


def last_error() -> str:
    msg = lib().slt_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> None:
    """Raise the Python exception matching a libsloth status code."""
    if rc == SLT_OK:
        return
    msg = last_error() or what
    if rc == SLT_ERR_INVALID:
        raise ValueError(msg)
    if rc == SLT_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise SlothError(msg)


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_handle(stream=None) -> int | None:
    """hipStream_t of a torch stream (current stream by default)."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream or None


def require_device():
    """The HIP path is the only path: fail loudly without a visible GPU."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("sloth_amd needs a HIP device (MI355X); none is visible")
    lib()
    return torch.device("cuda", torch.cuda.current_device())

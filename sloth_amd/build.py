"""Build libsloth_hip.so (gfx950) in-tree with hipcc.

Usage: ``python -m sloth_amd.build`` (also called by ``__graft_entry__.build()``).
The library lands in ``sloth_amd/_lib/libsloth_hip.so``; it is a build product
(git-ignored) that the Python package loads with ctypes.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libsloth_hip.so")
SOURCES = ["abi.hip", "grid_pack.hip", "trace.hip", "detector.hip", "optics.hip", "beam.hip",
           "adjoint.hip", "calib.hip"]
ARCH = os.environ.get("SLOTH_OFFLOAD_ARCH", "gfx950")

# Packed-FP32 VALU forms are disabled for all device code (some of their forms
# have returned wrong results on this hardware).  FMAs are formed only inside
# one source expression (-ffp-contract=on), so the offline, serving and adjoint
# kernels, which inline the same integration code into different control flow,
# round identically and give bit-equal results.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-fast-math", "-ffp-contract=on", "-Wall", "-Wno-unused-function",
    "-munsafe-fp-atomics",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

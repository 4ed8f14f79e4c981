"""sloth_amd -- MI355X-native (gfx950 HIP) geometric-optics ray tracer.

Drop-in for the hot path of MAGPIE-ICL/SLOTH: ``core.propagator.solve`` /
``trace_and_save_depths``, ``precompute_gradients``, ``kappa_inv_brems``, the
``ScalarDomain`` attribute protocol, ``Beam`` and ``Diagnostic.histogram``.
Kernels live in ``csrc/`` and are reached through the C ABI of
``include/sloth_hip.h`` (ctypes), with PyTorch-ROCm tensors as device buffers.
"""
__version__ = "0.1.0"

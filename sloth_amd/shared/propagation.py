"""Ray state -> geometric Jones vector, on the GPU.

Mirrors ``shared.propagation.ray_to_Jonesvector`` (src/shared/propagation.py:4-90).
"""
from __future__ import annotations

import torch

from .. import _native as N
from ..core.engine import axis_index, device_f32, unpermute_rays


def jones_device(state: torch.Tensor, probing_direction: str = "z", tau: torch.Tensor | None = None,
                 order: torch.Tensor | None = None):
    """(6, N) (or (>=6, N)) float32 device state -> ((4, N) float32, amp (N,) float64 | None).

    With ``order`` the state (and tau) are in processing order, column t being the
    caller's ray order[t], and rf / amp come out in the caller's order.
    """
    if state.dim() != 2 or state.shape[0] < 6:
        raise AssertionError("expected a (6, N) ray state")
    st = state if state.is_contiguous() else state.contiguous()
    n = st.shape[1]
    rf = torch.empty((4, n), dtype=torch.float32, device=st.device)
    amp = None
    if tau is not None:
        tau = tau.to(torch.float64).contiguous()
        amp = torch.empty(n, dtype=torch.float64, device=st.device)
    N.check(N.lib().slt_jones(st.data_ptr(), n, n, axis_index(probing_direction), rf.data_ptr(),
                              N.ptr(tau), N.ptr(amp), N.stream_handle()))
    if order is not None:
        rf, amp = unpermute_rays(rf, amp, order)
    return rf, amp


def ray_to_Jonesvector(rays, *, ne_extent=None, probing_direction="z", keep_current_plane=False):
    """Reference-compatible wrapper: returns ((4, N) array, None).

    ``keep_current_plane=False`` back-projects each ray to the plane
    ``ne_extent`` of the probing axis before taking its transverse position
    (src/shared/propagation.py:36-78); the angles do not depend on it.
    """
    dev = N.require_device()
    if ne_extent is None and not keep_current_plane:
        keep_current_plane = True
    st = device_f32(rays, dev)
    if st.shape[0] > 6:
        st = st[:6].contiguous()
    if not keep_current_plane:
        ax = axis_index(probing_direction)
        t_bp = (st[ax] - float(ne_extent)) / st[3 + ax]
        st = st.clone()
        for a in range(3):
            if a != ax:
                st[a] = st[a] - st[3 + a] * t_bp
    rf, _ = jones_device(st, probing_direction)
    return rf.cpu().numpy(), None

"""Drop-in ``core.propagator`` on MI355X.

Same names, arguments, return shapes and exception types as the reference
(src/core/propagator.py), with the ray integration done by the HIP kernels
of ``libsloth_hip.so``:

* ``precompute_gradients``  (:39-93)   -> k1 planar outputs
* ``kappa_inv_brems``       (:116-164) -> element-wise kernel
* ``solve``                 (:364-929) -> k1 pack + k2 RK4 trace + Jones
* ``trace_and_save_depths`` (:932-1234)-> same, with depth save points

Integrator: the reference uses diffrax Tsit5 with a PID step controller
(rtol 1e-3, atol 1e-5).  Here every ray takes fixed RK4 steps of
``step_fraction`` (default 0.5) of the smallest cell edge at speed c, landing
exactly on every save time.  ``rtol``/``atol``/``jitted``/``parallelise`` are
accepted for call compatibility and have no effect.  Results agree with the
reference within its own rtol (the reference's known-answer tests pass
unchanged) and with the FP64 oracle to 1e-5.
"""
from __future__ import annotations

import pickle
import time
import warnings

import numpy as np
import torch

from .. import _native as N
from ..constants import C_LIGHT
from . import engine
from .beam import Beam
from ..shared.propagation import jones_device


def omega_pe(ne):
    """Electron plasma frequency [rad/s] for ne in cm^-3 (NRL p.28), propagator.py:31-33."""
    return 5.64e4 * np.sqrt(ne)


def n_refrac(ne, omega):
    """Plasma refractive index for ne in m^-3, propagator.py:36-37."""
    return np.sqrt(1.0 - (omega_pe(ne * 1e-6) / omega) ** 2)


def _warn_overcritical(ne, omega, stacklevel=3):
    """One UserWarning when max(ne) >= n_cr (src/core/propagator.py:72-84)."""
    ncr = engine.critical_density(omega)
    if isinstance(ne, torch.Tensor):
        ne_max = float(ne.max())
    else:
        ne_max = float(np.max(np.asarray(ne)))
    if ne_max >= ncr:
        warnings.warn(
            f"precompute_gradients: {ne_max:.3g} m⁻³ cells at or above critical "
            f"density ncr = {ncr:.3g} m⁻³ (ne/ncr = {ne_max / ncr:.2g}).  "
            "This causes large gradient magnitudes that prevent ODE convergence.  "
            "Check that ne is in m⁻³ (not cm⁻³ — convert with ne_cc * 1e6) and "
            "that your domain does not include over-critical cells.",
            stacklevel=stacklevel,
        )


def precompute_gradients(ne, x, y, z, omega):
    """(dndx, dndy, dndz) of ``-0.5 c^2 ne / (3.14207787e-4 omega^2)``.

    numpy.gradient semantics with coordinate arrays (2nd-order interior,
    1st-order edges); fp64 arithmetic on the GPU, fp32 results as device
    tensors.  src/core/propagator.py:39-93.
    """
    dev = N.require_device()
    ne_d = engine.device_f32(ne, dev)
    _warn_overcritical(ne_d, omega)
    return engine.grid_fields(ne_d, x, y, z, omega, device=dev)


def kappa_inv_brems(ne, Te, Z, omega, *, coulomb_log=None):
    """Inverse-bremsstrahlung amplitude absorption rate [1/s] (NRL p.58).

    src/core/propagator.py:116-164.  ``coulomb_log`` (extension) replaces the
    reference's clamped ``max(2, ln(v_the / (omega_max b_min)))`` when given.
    Returns a device tensor with the shape of ``ne`` (0-d for scalars).
    """
    return engine.kappa_field(ne, Te, Z, omega, coulomb_log=coulomb_log)


def _domain_grid(domain, omega, inv_brems):
    ne = domain.ne
    _warn_overcritical(ne, omega, stacklevel=4)
    return engine.PackedGrid(
        ne, domain.x, domain.y, domain.z, omega,
        probing_direction=domain.probing_direction, inv_brems=inv_brems,
        Te=getattr(domain, "Te", None), Z=getattr(domain, "Z", None),
        coulomb_log=getattr(domain, "coulomb_log", None))


def _beam_rays(beam, domain):
    if isinstance(beam, Beam):
        raise AssertionError(
            "\nThis function does not take in the direct output of the Beam object, pass either "
            "Beam.s0 rays, or the parameters passed to be Beam here as a tuple if batching rays.")
    if isinstance(beam, tuple):
        np_total = getattr(domain, "Np_total", None)
        assert np_total is not None, (
            "\nScalarDomain.Np_total must be set when passing beam as a tuple. "
            "Pass Np=<number_of_rays> to ScalarDomain(...).")
        return Beam(np_total, beam_size=beam[0], divergence=beam[1], ne_extent=beam[2],
                    probing_direction=beam[3], beam_type=beam[4], seeded=beam[5]).s0
    shape = getattr(beam, "shape", None)
    assert shape is not None and len(shape) == 2, "\nExpected a matrix of pre-created rays."
    return beam


def solve(beam, ScalarDomain, probing_depth, *, parallelise=True, jitted=True,
          save_points_per_region=2, memory_debug=False, lwl=1064e-9, keep_domain=False,
          return_raw_results=False, verbose=True, step_fraction=0.5):
    """Trace rays for ``sqrt(8) * probing_depth / c`` and return their Jones vector.

    Returns ``(rf, amp, duration)`` like src/core/propagator.py:364-929:
    ``rf`` (4, N) float32 (or (S, 4, N) when ``save_points_per_region > 2``),
    ``amp = exp(-tau)`` (N,) when ``ScalarDomain.inv_brems`` else None, and the
    wall time of the trace in seconds (grid packing excluded, as the reference
    excludes ``precompute_gradients``).  ``return_raw_results=True`` returns the
    engine's ``TraceResult`` in place of ``rf``.
    """
    omega = 2 * np.pi * (C_LIGHT / lwl)
    inv_brems = bool(getattr(ScalarDomain, "inv_brems", False))
    if verbose:
        print("\nNumber of domain batches:", getattr(ScalarDomain, "region_count", 1))
        print("Number of ray batches:", getattr(ScalarDomain, "ray_batch_count", 1))
    s0 = _beam_rays(beam, ScalarDomain)
    dev = N.require_device()
    grid = _domain_grid(ScalarDomain, omega, inv_brems)
    rays = engine.as_rays(s0, dev)
    n_save = max(int(save_points_per_region), 2)
    t_end = np.sqrt(8.0) * float(probing_depth) / C_LIGHT
    save_times = np.linspace(0.0, t_end, n_save)
    torch.cuda.synchronize()
    start = time.time()
    # raw results keep the caller's ray order; otherwise the rays are traced in
    # processing order and the Jones epilogue puts them back
    res = engine.trace(grid, rays, save_times, step_fraction=step_fraction,
                       save_all=n_save > 2, want_path=False, gather=not return_raw_results)
    torch.cuda.synchronize()
    duration = np.float64(time.time() - start)
    if verbose:
        print("\nCompleted ray trace in", np.round(duration, 3), "seconds.")
    if return_raw_results:
        return res, None, duration
    pd = ScalarDomain.probing_direction
    back = res.order if res.permuted else None
    if n_save == 2:
        rf, amp = jones_device(res.states[-1], pd, res.tau[-1] if inv_brems else None,
                               order=back)
        rf = rf.cpu().numpy()
        if inv_brems:
            return rf, amp.cpu().numpy(), duration
        return rf, None, duration
    rfs = [jones_device(res.states[j], pd, order=back)[0] for j in range(res.states.shape[0])]
    return torch.stack(rfs).cpu().numpy(), None, duration


def trace_and_save_depths(beam, ScalarDomain, step, depth_max, output_path, *,
                          lwl=1064e-9, jones_components=None, jitted=True,
                          rtol=1e-3, atol=1e-5, verbose=True, step_fraction=0.5):
    """Trace and save the Jones vector every ``step`` metres of depth up to ``depth_max``.

    Returns the dict of src/core/propagator.py:932-1234: ``depth_saves``,
    ``jvec`` (list of (n_comp, N) arrays), ``jones_components`` and, with
    inverse bremsstrahlung, ``amplitude`` (list of (N,) arrays of exp(-tau)).
    Save k is the ray state at real time ``depth_saves[k] / c``.
    """
    omega = 2 * np.pi * (C_LIGHT / lwl)
    if isinstance(beam, tuple):
        s0 = _beam_rays(beam, ScalarDomain)
    else:
        s0 = beam
    if step <= 0 or depth_max <= 0:
        raise ValueError("step and depth_max must be positive.")
    if jones_components is None or (isinstance(jones_components, str) and jones_components == "all"):
        comp = [0, 1, 2, 3]
    elif isinstance(jones_components, str) and jones_components == "position":
        comp = [0, 2]
    elif isinstance(jones_components, str) and jones_components == "angle":
        comp = [1, 3]
    else:
        comp = list(jones_components)
        invalid = [i for i in comp if i not in (0, 1, 2, 3)]
        if invalid:
            raise ValueError(f"jones_components indices {invalid} are out of range. "
                             f"Valid indices are 0, 1, 2, 3.")
    pd = ScalarDomain.probing_direction
    dir_idx = ["x", "y", "z"].index(pd)
    trace_depth = float(np.asarray(ScalarDomain.lengths)[dir_idx])
    if trace_depth < depth_max:
        raise ValueError(
            f"Domain length in probing direction '{pd}' ({trace_depth:.6g} m) "
            f"is smaller than the requested depth_max ({depth_max:.6g} m). "
            f"Increase the domain size to at least depth_max.")
    n_saves = int(np.round(depth_max / step)) + 1
    depth_saves = np.linspace(0.0, depth_max, n_saves)
    inv_brems = bool(getattr(ScalarDomain, "inv_brems", False))
    if verbose:
        print(f"\ntrace_and_save_depths: saving at {n_saves} depth(s) from 0 to "
              f"{depth_max * 1e3:.4g} mm (step = {step * 1e6:.4g} µm, probing direction = "
              f"'{pd}', jones_components = {comp}, inv_brems = {inv_brems}).")
    dev = N.require_device()
    grid = _domain_grid(ScalarDomain, omega, inv_brems)
    rays = engine.as_rays(s0, dev)
    n = rays.shape[1]
    if n_saves >= 2:
        res = engine.trace(grid, rays, depth_saves / C_LIGHT, step_fraction=step_fraction,
                           save_all=True, want_path=False, gather=True)
        states, taus = res.states, res.tau
        back = res.order if res.permuted else None
    else:
        states = rays.unsqueeze(0)
        taus = torch.zeros((1, n), dtype=torch.float64, device=dev) if inv_brems else None
        back = None
    jvec_list, amp_list = [], ([] if inv_brems else None)
    for j in range(n_saves):
        rf, amp = jones_device(states[j], pd, taus[j] if inv_brems else None, order=back)
        jvec_list.append(rf.cpu().numpy()[comp, :])
        if inv_brems:
            amp_list.append(amp.cpu().numpy())
    result = {"depth_saves": depth_saves, "jvec": jvec_list, "jones_components": comp}
    if inv_brems:
        result["amplitude"] = amp_list
    if output_path is not None:
        with open(output_path, "wb") as fh:
            pickle.dump(result, fh)
        if verbose:
            print(f"trace_and_save_depths: results saved to '{output_path}'.")
    return result

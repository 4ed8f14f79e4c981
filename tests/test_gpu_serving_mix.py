"""The serving loop at the serving mix's own sizes, tied to the FP64 oracle.

Eight requests of the benchmark's seeded serving mix (grid sides 64-384, rays
1.7e4-3.8e6, 532 and 1064 nm, sine fields without and Kolmogorov fields with
absorption), including a 384^3 Kolmogorov grid and requests of 2.8e6-3.8e6
rays, run through ``ServingEngine`` with its default launch shape and batch
length.  Each served request must equal the offline path (``slt_trace``) and
the same request served alone (1e-12 relative), and a seeded 2000-ray subset
of each must match the FP64 oracle traced through the same field with the same
rays (the parity bar of 1e-5, the same exception rule as every oracle test).
One request is what one reference ``solve`` call computes
(src/core/propagator.py:364-929): these rays through this field to the solve
horizon.
"""
import numpy as np
import pytest

from _parity import C, assert_parity

pytestmark = pytest.mark.gpu

# rids of bench.serving_mix(48): 29 kolmogorov 384^3 8.8e5 rays; 37 kolmogorov 256^3
# 3.8e6 rays; 23 sine 384^3 3.4e6 rays (1064 nm); 2 sine 128^3 2.8e6 rays (1064 nm);
# 7 kolmogorov 64^3 3.6e5 rays; 19 kolmogorov 128^3 1.7e4 rays; 4 kolmogorov 192^3
# 7.1e4 rays (1064 nm); 33 kolmogorov 384^3 5.2e4 rays (1064 nm)
MIX_RIDS = (29, 37, 23, 2, 7, 19, 4, 33)


@pytest.fixture(scope="module", autouse=True)
def _dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _requests():
    import bench

    mix = {r.rid: r for r in bench.serving_mix(48)}
    return [mix[rid] for rid in MIX_RIDS]


def _close(a, b, rtol=1e-12):
    a = a.detach().cpu().numpy().astype(np.float64)
    b = b.detach().cpu().numpy().astype(np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    err = np.max(np.abs(a - b)) / scale
    assert err <= rtol, err


def test_serving_mix_sizes_offline_alone_and_oracle(orc):
    import torch

    from sloth_amd.serving import ServingEngine, run_offline
    from sloth_amd.serving.requests import _coords, beam_on_device, field_on_device

    reqs = _requests()
    sides = {r.field.side for r in reqs}
    assert 384 in sides and max(r.beam.n_rays for r in reqs) >= 2_000_000
    assert any(r.field.kind == "kolmogorov" and r.field.side == 384 for r in reqs)
    print("serving", len(reqs), "requests,", sum(r.beam.n_rays for r in reqs), "rays", flush=True)
    resp, stats = ServingEngine().run(_requests(), realtime=False)
    assert stats["ok"] == len(reqs), stats
    print("served:", stats["batches"], "batches", flush=True)
    off, off_stats = run_offline(_requests())
    assert stats["ray_steps"] == off_stats["ray_steps"]
    dev = torch.device("cuda", 0)
    for r in reqs:
        got, ref = resp[r.rid], off[r.rid]
        assert got.status == "ok" and got.ray_steps == ref["steps"]
        _close(got.state, ref["state"])
        _close(got.rf, ref["rf"])
        _close(got.path, ref["path"])
        if r.inv_brems:
            _close(got.amp, ref["amp"])
        assert (got.image.double() - ref["image"].double()).abs().sum().item() <= 2.0
        assert int(got.steps.long().sum().item()) == got.ray_steps
        # the same request served alone
        alone, _ = ServingEngine(max_slots=1).run([_requests()[MIX_RIDS.index(r.rid)]],
                                                  realtime=False)
        _close(alone[r.rid].state, got.state)
        _close(alone[r.rid].path, got.path)

        # a seeded 2000-ray subset against the FP64 oracle on the same field and rays
        ne = field_on_device(r.field, dev).cpu().numpy()
        s0 = beam_on_device(r.beam, r.field.length, dev).cpu().numpy()
        coords = _coords(r.field)
        idx = np.sort(np.random.default_rng(r.rid).choice(s0.shape[1], 2000, replace=False))
        nsub, h = ref["nsub"], ref["h"]      # the schedule both paths integrated
        omega = 2 * np.pi * C / r.lwl
        c64 = [c.astype(np.float64) for c in coords]
        ne64 = ne.astype(np.float64)
        del ne
        grads = orc.gradients(ne64, *c64, omega)
        kap = (orc.kappa(ne64, r.Te, r.Z, omega, coulomb_log=r.coulomb_log)
               if r.inv_brems else None)
        del ne64
        saves, steps, gap = orc.trace(grads, kap, c64, s0[:, idx].astype(np.float64), nsub,
                                      h.astype(np.float32).astype(np.float64), face_gap=True)
        del grads, kap
        gpu = {"states": got.state.cpu().numpy().astype(np.float64)[None][:, :, idx],
               "tau": got.tau.cpu().numpy()[None][:, idx] if r.inv_brems else None,
               "path": got.path.cpu().numpy()[idx],
               "steps": got.steps.cpu().numpy().astype(np.int64)[idx]}
        oref = {"saves": saves[-1:], "steps": steps, "face_gap": gap}
        print(f"request {r.rid} ({r.field.kind} {r.field.side}^3, {r.beam.n_rays} rays, "
              f"{r.lwl * 1e9:.0f} nm): oracle parity",
              assert_parity(gpu, oref, coords), flush=True)

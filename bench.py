"""Throughput benchmark: ray-steps/s on the 512^3 turbulent box with absorption.

Workload (one GPU): 512^3 fp32 grid over a 20 mm cube, n_e = max(0, 2e24 (1 + 0.3 g))
with g a unit-variance Kolmogorov (k^-11/3) Gaussian random field between
k = 2 pi / 20 mm and 2 pi / 0.2 mm; Te = 20 eV, Z = 1, Coulomb log 10;
1e7 rays at 532 nm from a seeded collimated square beam (half-width 8 mm)
entering at z = -10 mm; fixed-step RK4 at half a cell; integration horizon
sqrt(8) * 20 mm / c (the reference ``solve`` horizon).

One step = one pass of the traced path over the batch: Morton keys + sort,
RK4 trace (FP64 tau and path), Jones epilogue with exp(-tau), and a 256 x 256
detector histogram.  The packed grid and the rays are resident in HBM before
timing starts.  value = in-box ray-steps per second (a ray-step is one RK4
step with a stage point inside the grid box).

Prints one JSON line (rank 0).  ``roofline`` is the trace kernel's binding
resource, VALU issue: its VALU lane-instructions per second (instructions per
ray-step from the committed rocprofv3 PMC summary, times this run's ray-steps
over the kernel's average duration from HIP events on its stream) against the
rate of a calibration kernel of known VALU count (independent fp32 FMA
chains, no memory traffic) timed in the same run.  ``hbm_roofline`` is the
measured HBM traffic of one launch (FETCH_SIZE + WRITE_SIZE, PMC summary)
over the same duration against 8 TB/s, beside the compulsory bytes (packed
grid once, per-ray I/O once).  ``cpu_baseline`` is the FP64 C oracle built
with -O3 -march=native on this host (OpenMP, all cores) on a bounded
subsample of the same rays; ``cpu_baseline_numpy`` is the oracle's NumPy
restatement (RK4 vectorised over rays) in a process pool over the same cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

C = 299792458.0
HBM_PEAK_GBS = 8000.0
# algorithmic gather bytes of a ray-step: 4 RHS x 8 corners x 4 fp32 channels; served
# from per-lane registers and L2/MALL, not HBM (reported, not a roof)
BYTES_PER_STEP = 4 * 8 * 4 * 4
# compulsory HBM bytes of one trace launch: the packed grid (16 B per voxel) once, and
# per ray its start state (24 B, already gathered into processing order) read, final
# state (24 B), tau (8 B), path (8 B) and step count (4 B) written
VOXEL_BYTES = 16
RAY_IO_BYTES = 24 + 24 + 8 + 8 + 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


PMC_SUMMARY = os.path.join("profiles", "trace_kernel_turbulent512_pmc.md")


def profile_value(key):
    """A figure of the committed rocprofv3 --pmc summary of the trace kernel (or None)."""
    p = os.path.join(ROOT, PMC_SUMMARY)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith(key + ":"):
            try:
                return float(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def traffic_from_profiles():
    """HBM bytes per trace launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE)."""
    return profile_value("traffic_bytes_per_launch")


def valu_calibration(device, n_cu, iters=200_000, reps=3):
    """VALU-issue rates measured on this GPU: slt_valu_calibrate for every instruction
    form (SLT_CALIB_CHAINS independent chains per lane, eight wavefronts per SIMD), each
    timed with HIP events, best of `reps` launches.  Returns ({form: lane-instructions per
    second}, {form: ms})."""
    import torch

    from sloth_amd import _native as N

    L = N.lib()
    blocks = 8 * n_cu                       # 256-lane workgroups: 8 wavefronts per SIMD
    out = torch.empty(blocks * 256, dtype=torch.float32, device=device)
    stream = torch.cuda.current_stream(device)
    lane_instr = blocks * 4 * N.SLT_CALIB_CHAINS * iters * 64.0
    rates, times = {}, {}
    for name, kind in N.SLT_CALIB_FORMS.items():
        N.check(L.slt_valu_calibrate(kind, blocks, 1000, out.data_ptr(), N.stream_handle()))
        best = None
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            N.check(L.slt_valu_calibrate(kind, blocks, iters, out.data_ptr(),
                                         N.stream_handle()))
            b.record(stream)
            b.synchronize()
            ms = a.elapsed_time(b)
            best = ms if best is None else min(best, ms)
        rates[name] = lane_instr / (best * 1e-3)
        times[name] = best
    return rates, times


def valu_roofline(steps_per_s_kernel, rates, times=None):
    """VALU issue of the trace kernel, its binding resource, in TOP/s of VALU
    lane-instructions (a wave64 instruction = 64): the PMC summary's lane-instructions per
    ray-step (SQ_INSTS_VALU x 64 / ray-steps) times the run's ray-steps per second of
    kernel time, against the fastest instruction form of the calibration kernels measured
    in the same run (``rates``: lane-instructions per second per form)."""
    lane_instr = profile_value("valu_lane_instr_per_ray_step")
    if not lane_instr or not rates:
        return None
    form = max(rates, key=rates.get)
    achieved = steps_per_s_kernel * lane_instr / 1e12
    peak = rates[form] / 1e12
    return {"bound": "valu_issue", "achieved": round(achieved, 3), "peak": round(peak, 3),
            "unit": "TOP/s", "frac": round(achieved / peak, 4),
            "op": "VALU lane-instruction (one wave64 VALU instruction = 64)",
            "lane_instr_per_ray_step": lane_instr, "instr_source": PMC_SUMMARY,
            "peak_source": f"slt_valu_calibrate, fastest form ({form}), 16 independent chains "
                           "per lane, 8 wavefronts per SIMD, timed in this run",
            "calibration_tops": {k: round(v / 1e12, 3) for k, v in rates.items()},
            "calibration_ms": {k: round(v, 3) for k, v in times.items()} if times else None}


def hbm_roofline(kernel_s, n_voxels, n_rays):
    """Measured HBM traffic of one trace launch (PMC summary) over the kernel's duration,
    against the 8 TB/s peak, with the compulsory bytes beside it."""
    traffic = traffic_from_profiles()
    compulsory = VOXEL_BYTES * n_voxels + RAY_IO_BYTES * n_rays
    if not traffic or not kernel_s:
        return None
    achieved = traffic / kernel_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "compulsory_bytes": compulsory,
            "traffic_over_compulsory": round(traffic / compulsory, 3)}


def cpu_baseline(ne, coords, s0, nsub, h, omega, target_s):
    from oracle import oracle as O

    # the timed CPU restatement: the oracle's C source compiled for this host with
    # -O3 -march=native (the checker build keeps its portable, contraction-free flags)
    flags = O.load_native()
    cores = len(os.sched_getaffinity(0))
    c64 = [c.astype(np.float64) for c in coords]
    ne64 = ne.astype(np.float64)
    grads = O.gradients(ne64, *c64, omega)
    kap = O.kappa(ne64, 20.0, 1.0, omega, coulomb_log=10.0)
    h64 = np.asarray(h, dtype=np.float32).astype(np.float64)
    rng = np.random.default_rng(1)
    # calibrate on 2000 rays, then size the sample for ~target_s seconds
    pick = rng.choice(s0.shape[1], 2000, replace=False)
    t0 = time.perf_counter()
    _, st = O.trace(grads, kap, c64, s0[:, pick].astype(np.float64), nsub, h64, nthreads=cores,
                    native=True)
    dt = time.perf_counter() - t0
    rate = st.sum() / dt
    n = int(min(s0.shape[1], max(2000, target_s * rate / max(st.mean(), 1))))
    pick = rng.choice(s0.shape[1], n, replace=False)
    t0 = time.perf_counter()
    _, st = O.trace(grads, kap, c64, s0[:, pick].astype(np.float64), nsub, h64, nthreads=cores,
                    native=True)
    dt = time.perf_counter() - t0
    del grads, kap
    return {"value": float(st.sum() / dt), "unit": "ray-steps/s", "cores": cores, "kind": "port",
            "sample": f"{n} of the workload's rays (seeded random subset), same grid and "
                      f"schedule, FP64 C oracle (gcc {flags}, OpenMP on {cores} threads), "
                      f"{dt:.1f} s"}


def cpu_baseline_numpy(ne, coords, s0, nsub, h, omega, target_s):
    """The oracle's NumPy restatement (RK4 vectorised over rays, as a vmapped solver on a
    CPU) on the host's cores.  It runs in a child interpreter (oracle/np_baseline.py)
    whose process pool is forked from a process that never opened the GPU."""
    import subprocess
    import tempfile

    cores = len(os.sched_getaffinity(0))
    rng = np.random.default_rng(3)
    pick = rng.choice(s0.shape[1], min(s0.shape[1], 1024 * cores), replace=False)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "np_baseline.npz")
        np.savez(path, ne=np.asarray(ne, dtype=np.float32), x=coords[0], y=coords[1],
                 z=coords[2], s0=s0[:, pick], nsub=np.asarray(nsub),
                 h=np.asarray(h, dtype=np.float32), omega=np.float64(omega))
        r = subprocess.run([sys.executable, "-m", "oracle.np_baseline", path, str(cores),
                            str(target_s)], cwd=ROOT, capture_output=True, text=True,
                           timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"NumPy CPU baseline failed ({r.returncode}): {r.stderr[-2000:]}")
    res = json.loads(r.stdout.strip().splitlines()[-1])
    return {"value": float(res["value"]), "unit": "ray-steps/s", "cores": int(res["workers"]),
            "kind": "port",
            "sample": f"{res['rays']} of the workload's rays (seeded random subset), "
                      f"{res['width']} per worker process, every ray through the whole "
                      f"schedule, same grid; FP64 NumPy RK4 vectorised over rays, "
                      f"{res['seconds']:.1f} s"}


def serving_mix(n_requests, seed=12345):
    """The serving mix: grid side from {64,128,192,256,384}, rays log-uniform in [1e4, 4e6],
    532 or 1064 nm, sine field (no absorption) or Kolmogorov field (absorption)."""
    from sloth_amd.serving import BeamSpec, FieldSpec, RequestSpec

    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n_requests):
        side = int(rng.choice([64, 128, 192, 256, 384]))
        rays = int(np.exp(rng.uniform(np.log(1e4), np.log(4e6))))
        lwl = float(rng.choice([532e-9, 1064e-9]))
        if rng.random() < 0.5:
            f = FieldSpec(kind="sine", side=side, length=10e-3, seed=int(rng.integers(1 << 30)),
                          n0=1e24, frac=0.1, wavelength=2e-3)
            inv = False
        else:
            f = FieldSpec(kind="kolmogorov", side=side, length=10e-3,
                          seed=int(rng.integers(1 << 30)), n0=2e24, frac=0.3, l_min=0.2e-3)
            inv = True
        reqs.append(RequestSpec(rid=i, field=f, beam=BeamSpec(n_rays=rays, half_width=4e-3,
                                                               seed=int(rng.integers(1 << 30))),
                                lwl=lwl, inv_brems=inv, Te=20.0, Z=1.0, coulomb_log=10.0,
                                bins=256))
    return reqs


def serving_bench(n_requests):
    from sloth_amd.serving import ServingEngine, run_offline

    from sloth_amd.serving import BeamSpec, FieldSpec, RequestSpec

    reqs = serving_mix(n_requests)
    # rocFFT builds its kernels at run time per transform size: warm every grid side
    # and field kind once so neither leg below pays for compilation
    warm = [RequestSpec(rid=i, field=FieldSpec(kind=k, side=s, length=10e-3),
                        beam=BeamSpec(n_rays=1000), inv_brems=(k == "kolmogorov"))
            for i, (k, s) in enumerate((k, s) for k in ("sine", "kolmogorov")
                                       for s in (64, 128, 192, 256, 384))]
    run_offline(warm)
    ServingEngine(max_slots=len(warm)).run(warm, realtime=False)
    # best of two for both legs (the first pass also grows the allocator's pools)
    off = max((run_offline(serving_mix(n_requests))[1] for _ in range(2)),
              key=lambda d: d["ray_steps_per_s"])
    log(f"[serving] offline batch: {off['ray_steps_per_s']:.3e} ray-steps/s over "
        f"{off['wall_s']:.2f} s")
    sat = max((ServingEngine().run(serving_mix(n_requests), realtime=False)[1]
               for _ in range(2)), key=lambda d: d["ray_steps_per_s"])
    log(f"[serving] saturated: {sat['ray_steps_per_s']:.3e} ray-steps/s, {sat['batches']} batches")
    capacity = n_requests / off["wall_s"]
    rate = 0.7 * capacity
    # latency percentiles over a longer stream of the same mix (own seed): p99 of 4 x 48
    # requests rests on ~2 samples above it rather than on the single worst of 48
    n_poisson = 4 * n_requests
    mix = serving_mix(n_poisson, seed=54321)
    rng = np.random.default_rng(777)
    t = 0.0
    for r in mix:
        t += rng.exponential(1.0 / rate)
        r.arrival = t
    _, poi = ServingEngine().run(mix, realtime=True)
    log(f"[serving] poisson {rate:.1f} req/s, {n_poisson} requests: "
        f"p50 {poi['p50_latency_s'] * 1e3:.1f} ms, "
        f"p99 {poi['p99_latency_s'] * 1e3:.1f} ms")
    return {
        "requests": n_requests,
        "mix": "side {64,128,192,256,384}, rays log-uniform [1e4,4e6], 532/1064 nm, "
               "sine (no absorption) or Kolmogorov (absorption, CL 10), 10 mm box, 256^2 detector",
        "offline_ray_steps_per_s": off["ray_steps_per_s"],
        "saturated_ray_steps_per_s": sat["ray_steps_per_s"],
        "saturated_over_offline": sat["ray_steps_per_s"] / off["ray_steps_per_s"],
        "poisson_rate_hz": rate,
        "poisson_load": 0.7,
        "poisson_requests": n_poisson,
        "p50_latency_s": poi["p50_latency_s"],
        "p99_latency_s": poi["p99_latency_s"],
        "poisson_ray_steps_per_s": poi["ray_steps_per_s"],
        "total_ray_steps": off["ray_steps"],
    }


def adjoint_bench(n_rays=2_000_000, n=256):
    """Config A: 256^3 Kolmogorov (10 mm), 2e6 rays at 532 nm with absorption; loss = MSE of
    detector positions and energies against a trace through a second seeded grid.  Times
    the forward and the backward with checkpoints every 16 steps, then with the whole
    trajectory recorded by the forward (ckpt_every = 0, no replay in the backward)."""
    import torch

    from sloth_amd.adjoint import detector_loss, trace_states
    from sloth_amd.fields import kolmogorov_field
    from sloth_amd.core.beam import Beam

    dev = torch.device("cuda", 0)
    L = 10e-3
    ne, coords = kolmogorov_field(n, L, n0=2e24, frac=0.3, l_min=0.2e-3, seed=21, device=dev)
    tne, _ = kolmogorov_field(n, L, n0=2e24, frac=0.3, l_min=0.2e-3, seed=22, device=dev)
    s0 = torch.from_numpy(Beam(n_rays, 4e-3, 0.0, L / 2, beam_type="square",
                               seeded=True).s0).to(dev)
    kw = dict(lwl=532e-9, inv_brems=True, Te=20.0, Z=1.0, coulomb_log=10.0)
    with torch.no_grad():
        ts, tt, _ = trace_states(torch.from_numpy(tne).to(dev), s0, *coords, **kw)
    tp, te = ts[0:2].double(), torch.exp(-tt)
    del ts, tt
    ne_t = torch.from_numpy(ne).to(dev)
    res = {"workload": "kolmogorov_256_2e6_rays_absorption"}
    for K in (16, 0):
        times = []
        for rep in range(3):
            x = ne_t.clone().requires_grad_(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st, tau, total = trace_states(x, s0, *coords, ckpt_every=K, **kw)
            loss = detector_loss(st, tau, tp, te, pos_scale=1e-4)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            loss.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            times.append((t1 - t0, t2 - t1))
            del st, tau, loss
        fwd, bwd = min(t[0] for t in times), min(t[1] for t in times)
        steps = int(total.item())
        log(f"[adjoint] ckpt_every {K}: forward {fwd * 1e3:.1f} ms, backward {bwd * 1e3:.1f} ms, "
            f"{steps} ray-steps")
        leg = {"forward_ms": fwd * 1e3, "backward_ms": bwd * 1e3,
               "backward_ray_steps_per_s": steps / bwd, "ckpt_every": K,
               "grad_norm": float(x.grad.double().norm())}
        if K:
            res.update(ray_steps=steps, forward_ckpt_ms=leg["forward_ms"],
                       backward_ms=leg["backward_ms"],
                       backward_ray_steps_per_s=leg["backward_ray_steps_per_s"], ckpt_every=K,
                       grad_norm=leg["grad_norm"])
        else:
            res["trajectory_tape"] = leg
        del x, total
        torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--rays", type=int, default=10_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--shard-parts", type=int, default=8,
                    help="also time one GPU's 1/N share of the rays (0 or 1: skip)")
    ap.add_argument("--serving-requests", type=int, default=48,
                    help="requests in the serving mix (0 skips the serving leg)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB first (PMC calibration)")
    ap.add_argument("--valu-calibration-only", action="store_true",
                    help="run only the VALU calibration kernel (PMC calibration) and exit")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint leg (profiling) and print its figures")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the traced path runs on a single MI355X "
              "(runs on several GPUs are not launched)", file=sys.stderr)
        sys.exit(2)

    import torch

    from sloth_amd.core import engine
    from sloth_amd.core.beam import Beam
    from sloth_amd.fields import kolmogorov_field
    from sloth_amd.processing.diagnostics import histogram2d_device
    from sloth_amd.shared.propagation import jones_device

    dev = torch.device("cuda", 0)
    if args.copy_calibration_mib > 0:
        # a plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE counters
        # scale the trace kernel's counters to bytes
        src = torch.ones(args.copy_calibration_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.adjoint_only:
        log(json.dumps(adjoint_bench()))
        return
    if args.valu_calibration_only:
        rates, ms = valu_calibration(dev,
                                     torch.cuda.get_device_properties(0).multi_processor_count)
        for k in rates:
            log(f"[bench] VALU calibration {k}: {rates[k] / 1e12:.2f} T lane-instr/s, "
                f"{ms[k]:.3f} ms")
        return
    L = 20e-3
    n = args.grid
    lwl = 532e-9
    omega = 2 * np.pi * C / lwl
    log(f"[bench] field {n}^3 ...")
    ne, coords = kolmogorov_field(n, L, n0=2e24, frac=0.3, l_max=L, l_min=0.2e-3, seed=0,
                                  device=dev)
    log("[bench] beam ...")
    s0_host = Beam(args.rays, 8e-3, 0.0, L / 2, beam_type="square", seeded=True).s0
    grid = engine.PackedGrid(ne, *coords, omega, inv_brems=True, Te=20.0, Z=1.0,
                             coulomb_log=10.0)
    s0 = torch.from_numpy(s0_host).to(dev)
    T = np.sqrt(8.0) * L / C
    save_times = np.array([0.0, T])
    edges = np.linspace(-L / 2, L / 2, 257)
    stream = torch.cuda.current_stream()

    work = engine.ray_work(args.rays, dev)

    def step(events=None, s0=s0, work=work):
        # Morton order, start states gathered into it, the trace over the gathered rays
        # (coalesced state reads and writes), the Jones epilogue and the detector
        # histogram in that order, then rf and amp back in the caller's ray order
        order = engine.ray_order(grid, s0)
        s0_sorted = engine.gather_rays(s0, order, work=work)
        if events is not None:
            events[0].record(stream)
        res = engine.trace(grid, s0_sorted, save_times, save_all=False, want_path=True,
                           sort=False)
        if events is not None:
            events[1].record(stream)
        rf_sorted, amp_sorted = jones_device(res.states[0], "z", res.tau[0])
        counts, wsum = histogram2d_device(rf_sorted[0], rf_sorted[2], edges, edges,
                                          weights=amp_sorted)
        rf, amp = engine.unpermute_rays(rf_sorted, amp_sorted, order, work=work)
        return res, counts, wsum, rf, amp

    log("[bench] warmup ...")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    log(f"[bench] timing {args.steps} steps ...")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    results = []
    for k in range(args.steps):
        results.append(step(evs[k]))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    res, counts = results[-1][0], results[-1][1]
    per_pass = [int(r[0].total_steps.item()) for r in results]
    total_steps = per_pass[0]
    assert all(p == total_steps for p in per_pass), per_pass
    del results
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    ms_per_step = wall * 1e3 / args.steps
    value = total_steps / (ms_per_step * 1e-3)
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    rates, calib_ms = valu_calibration(dev, n_cu)
    roofline = valu_roofline(total_steps / (kern_ms * 1e-3), rates, calib_ms) or {}
    roofline.update({"traffic": traffic_from_profiles(),
                     "kernel": "slt::trace_kernel<true,true>", "kernel_ms": round(kern_ms, 3)})
    hbm = hbm_roofline(kern_ms * 1e-3, n ** 3, args.rays)
    gathers = {"bytes_per_ray_step": BYTES_PER_STEP,
               "per_launch": total_steps * BYTES_PER_STEP,
               "model": "4 RHS x 8 corners x 4 fp32 channels, served from per-lane "
                        "registers and L2/MALL (not HBM traffic)"}
    log(f"[bench] {ms_per_step:.2f} ms/step, trace kernel {kern_ms:.2f} ms, "
        f"{value:.3e} ray-steps/s, detector sum {int(counts.sum())}")
    del res, counts
    shard = None
    if args.shard_parts > 1:
        # one GPU's share of a run over shard_parts GPUs: a contiguous range of the
        # Morton-ordered rays (the middle one), through the same per-GPU step on the same
        # grid; the all-reduce of the detector image that would follow is not part of it
        parts = args.shard_parts
        order = engine.ray_order(grid, s0)
        lo = (parts // 2) * args.rays // parts
        hi = (parts // 2 + 1) * args.rays // parts
        s0_share = s0[:, order[lo:hi].long()].contiguous()
        work_share = engine.ray_work(hi - lo, dev)
        step(s0=s0_share, work=work_share)
        torch.cuda.synchronize()
        sh_evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                  for _ in range(args.steps)]
        t0 = time.perf_counter()
        outs = [step(ev, s0=s0_share, work=work_share) for ev in sh_evs]
        torch.cuda.synchronize()
        sh_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        sh_kern = float(np.mean([a.elapsed_time(b) for a, b in sh_evs]))
        sh_steps = int(outs[-1][0].total_steps.item())
        del outs
        shard = {"parts": parts, "rays": hi - lo, "ray_steps": sh_steps, "ms_per_step": sh_ms,
                 "kernel_ms": sh_kern,
                 "ray_steps_per_s": sh_steps / (sh_ms * 1e-3),
                 "over_full_batch": sh_steps / (sh_ms * 1e-3) / value,
                 "note": "one GPU's share (contiguous Morton range) of the same workload on "
                         "the same grid, full per-GPU step, no collective"}
        log(f"[bench] 1/{parts} share: {hi - lo} rays, {sh_ms:.2f} ms/step (trace kernel "
            f"{sh_kern:.2f} ms), "
            f"{shard['ray_steps_per_s']:.3e} ray-steps/s ({shard['over_full_batch']:.3f} of "
            f"the full batch)")
        del s0_share, work_share, order
    cpu = cpu_np = None
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        del s0
        torch.cuda.empty_cache()
        nsub, h = engine.schedule(save_times, engine.nominal_step(grid, 0.5))
        cpu = cpu_baseline(ne, coords, s0_host, nsub, h, omega, args.cpu_seconds)
        log(f"[bench] cpu {cpu['value']:.3e} ray-steps/s on {cpu['cores']} cores")
        cpu_np = cpu_baseline_numpy(ne, coords, s0_host, nsub, h, omega, args.cpu_seconds)
        log(f"[bench] cpu (NumPy) {cpu_np['value']:.3e} ray-steps/s on {cpu_np['cores']} "
            f"processes")
    serving = None
    if args.serving_requests > 0:
        log("[bench] serving mix ...")
        torch.cuda.empty_cache()
        serving = serving_bench(args.serving_requests)
    adjoint = None
    if not args.no_adjoint:
        log("[bench] adjoint ...")
        torch.cuda.empty_cache()
        adjoint = adjoint_bench()
    out = {
        "metric": "ray-steps/s (512^3 grid, 1e7 rays)",
        "value": value,
        "unit": "ray-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"turbulent_box_{n}_absorption", "grid": [n, n, n],
                   "box_mm": 20, "rays": args.rays, "lambda_nm": 532, "beam": "square 8 mm half-width",
                   "Te_eV": 20, "Z": 1, "coulomb_log": 10, "integrator": "rk4, 0.5 cell",
                   "horizon": "sqrt(8) L / c", "ray_steps_per_pass": total_steps,
                   "accumulators": "fp64 tau and path"},
        "rhs_evals_per_s": 4.0 * value,   # RK4: 4 RHS evaluations per ray-step
        "roofline": roofline,
        "hbm_roofline": hbm,
        "algorithmic_gathers": gathers,
        "shard": shard,
        "cpu_baseline": cpu,
        "cpu_baseline_numpy": cpu_np,
        "serving": serving,
        "adjoint": adjoint,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

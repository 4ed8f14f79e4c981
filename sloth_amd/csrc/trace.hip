// k2: fixed-step RK4 ray integrator, one ray per lane.
//
// RHS (src/core/propagator.py:168-243):  dr/dt = v,  dv/dt = g(r),  dtau/dt = kappa(r)
// where g and kappa are trilinear interpolants of the packed voxel grid that
// are 0 outside the closed grid box (src/core/interpolator.py:80-104).
//
// Design (gfx950):
//  * Rays arrive sorted along a Z-order curve of their transverse start
//    position, so the 64 lanes of a wavefront sit in a few neighbouring grid
//    columns and step through the same cells at the same time.
//  * Each lane keeps the cell it last touched in VGPRs as the 8 coefficients
//    of its trilinear polynomial (per channel: f = k0 + kx tx + ky ty + kz tz
//    + kxy tx ty + kxz tx tz + kyz ty tz + kxyz tx ty tz), built from the 8
//    corner voxels (8 x 16 B) when a stage point enters a new cell.  A ray
//    crosses a cell every ~2 steps along the probing axis (the fastest-varying
//    axis of the packed layout), so the gathers per RK4 step are ~1/8 of the
//    4 x 8 corner loads the algorithm names, and an evaluation is 7 FMAs per
//    channel.
//  * The state is in grid units (position (x - lo) / d, velocity v / d, with
//    the packed gradient pre-scaled by 1 / d), so every stage point is its own
//    grid coordinate: the cell test is a subtraction, with no per-stage scaling.
//    Position = fp64 base + fp32 offset: the offset takes each step's
//    increment and is folded into the base every 8 steps, so rounding stays at
//    the offset's (small) magnitude.  Velocity = fixed start velocity + fp32
//    change.  tau and path length accumulate in fp64.  All RHS arithmetic is
//    fp32.
//  * Steps whose straight segment cannot reach the box are exact linear moves
//    (the RHS is identically 0 there); a ray outside the box skips all of them
//    at once.  Only steps with a stage point inside the box are counted.
//  * Blocks are remapped so each XCD integrates one contiguous range of the
//    sorted rays (its own L2 holds that range's grid columns).
#include <math.h>
#include <stdio.h>
#include <stddef.h>
#include <string.h>

#include "integrate.h"

namespace slt {

struct TraceK {
    DevGrid G;
    int64_t n;
    const float* __restrict__ s0;
    const int32_t* __restrict__ order;
    int nseg;
    int save_every;
    const int32_t* __restrict__ seg_nsub;
    const float* __restrict__ seg_h;
    float* __restrict__ save_state;
    double* __restrict__ save_tau;
    double* __restrict__ path;
    uint32_t* __restrict__ steps;
    unsigned long long* __restrict__ total;
    int nblocks;
};

// Five wavefronts per SIMD (96 VGPRs): the few values that do not fit spill
// only outside the step loop, and the extra wavefront hides more of the
// gather latency than the spills cost (50.5 vs 52.7 ms for 1e10 ray-steps).
template <bool KAPPA, bool PATH, bool UNI>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
trace_kernel(TraceK K) {
    const int lb = xcd_block(blockIdx.x, gridDim.x, K.nblocks);
    if (lb < 0) return;
    const int64_t t = (int64_t)lb * blockDim.x + threadIdx.x;
    uint32_t nsteps = 0;
    if (t < K.n) {
        const int64_t n = K.n;
        const int64_t ray = K.order ? (int64_t)K.order[t] : t;
        RayState s;
        state_from_rays(K.G, K.s0, n, ray, s);
        SaveSink sink{K.save_state, K.save_tau, n, ray, K.save_every, K.nseg, K.s0 + 3 * n, n,
                      ray};
        sink.put(0, s, K.G);
        int64_t total = 0;
        for (int q = 0; q < K.nseg; ++q) total += K.seg_nsub[q];
        integrate<KAPPA, PATH, UNI>(K.G, s, K.seg_nsub, K.seg_h, K.nseg, 0, total, sink);
        if (PATH && K.path) K.path[ray] = s.path;
        nsteps = inbox_steps(s);
        if (K.steps) K.steps[ray] = nsteps;
    }
    if (K.total) add_steps(K.total, nsteps);
}

// ---------------------------------------------------------------------------
// Resumable ray state (serving): structure-of-arrays in one buffer.
//   rb[3] f64 | tau f64 | path f64 | off[3] f32 | v0[3] f32 | dv[3] f32 | nout u32 | fold u32
// rb, off and dv are in the grid units of the job's grid; v0 is the start
// velocity in m/s (converted on load, and the base of the saved velocity).
// ---------------------------------------------------------------------------
struct StateView {
    SLT_GLOBAL double* rb[3];
    SLT_GLOBAL double* tau;
    SLT_GLOBAL double* path;
    SLT_GLOBAL float* off[3];
    SLT_GLOBAL float* v0[3];
    SLT_GLOBAL float* dv[3];
    SLT_GLOBAL uint32_t* nout;
    SLT_GLOBAL uint32_t* fold;
};

__host__ __device__ inline int64_t state_bytes(int64_t n) { return n * (5 * 8 + 9 * 4 + 2 * 4); }

__device__ __forceinline__ StateView state_view(void* base, int64_t n) {
    StateView v;
    SLT_GLOBAL double* d = to_global(static_cast<double*>(base));
    for (int c = 0; c < 3; ++c) v.rb[c] = d + c * n;
    v.tau = d + 3 * n;
    v.path = d + 4 * n;
    SLT_GLOBAL float* f = reinterpret_cast<SLT_GLOBAL float*>(d + 5 * n);
    for (int c = 0; c < 3; ++c) {
        v.off[c] = f + c * n;
        v.v0[c] = f + (3 + c) * n;
        v.dv[c] = f + (6 + c) * n;
    }
    SLT_GLOBAL uint32_t* u = reinterpret_cast<SLT_GLOBAL uint32_t*>(f + 9 * n);
    v.nout = u;
    v.fold = u + n;
    return v;
}

__device__ __forceinline__ void state_load(const StateView& v, int64_t i, const DevGrid& G,
                                           RayState& s) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        s.rb[c] = v.rb[c][i];
        s.off[c] = v.off[c][i];
        s.v0[c] = velocity_units(G, c, v.v0[c][i]);
        s.dv[c] = v.dv[c][i];
    }
    s.tau = v.tau[i];
    s.path = v.path[i];
    s.nout = v.nout[i];
    s.fold = v.fold[i];
}

__device__ __forceinline__ void state_store(const StateView& v, int64_t i, const RayState& s) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        v.rb[c][i] = s.rb[c];
        v.off[c][i] = s.off[c];
        v.dv[c][i] = s.dv[c];
    }
    v.tau[i] = s.tau;
    v.path[i] = s.path;
    v.nout[i] = s.nout;
    v.fold[i] = s.fold;
}

__global__ void __launch_bounds__(256) state_init_kernel(DevGrid G, void* base, int64_t n,
                                                         const float* __restrict__ s0,
                                                         const int32_t* __restrict__ order) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t ray = order ? (int64_t)order[t] : t;
    RayState s;
    state_from_rays(G, s0, n, ray, s);
    const StateView v = state_view(base, n);
    state_store(v, t, s);
#pragma unroll
    for (int c = 0; c < 3; ++c) v.v0[c][t] = s0[(3 + c) * n + ray];
}

// Work unit = 64 consecutive (sorted) rays of one job: one wavefront.  The
// units are dealt into SLT_WORK_QUEUES contiguous queues; a wavefront takes
// units one at a time from its XCD's queue (workgroup index mod 8, the order
// in which workgroups are dispatched to XCDs) and, once that queue is empty,
// from the others.  A batch thus ends when its work does, however unevenly the
// jobs' per-step costs (absorption or not, grid size) fall across the queues.
template <bool KAPPA>
__device__ void advance_unit(const slt_job& J, int64_t t) {
    uint32_t nsteps = 0;
    if (t < J.n_rays) {
        DevGrid G;
        make_dev_grid_device(J.packed, &J.grid, &G);
        RayState s;
        state_load(state_view(J.state, J.n_rays), t, G, s);
        const uint32_t before = inbox_steps(s);
        const bool finishing = J.step_end >= J.sched_steps;
        const int64_t ray = J.order ? (int64_t)J.order[t] : t;
        // the state's v0 rows (m/s) follow rb[3], tau, path (f64) and off[3] (f32)
        const float* v0m = reinterpret_cast<const float*>(static_cast<const char*>(J.state) +
                                                          J.n_rays * (5 * 8 + 3 * 4));
        SaveSink sink{finishing ? J.out_state : nullptr, J.out_tau, J.n_rays, ray, 0, J.n_seg,
                      v0m, J.n_rays, t};
        integrate<KAPPA, true>(G, s, J.seg_nsub, J.seg_h, J.n_seg, J.step_begin, J.step_end,
                               sink);
        // Recompute the row's 14 addresses for the store instead of keeping the
        // load's live through the integration (the empty asm hides that they match).
        int64_t row = t;
        asm volatile("" : "+v"(row));
        state_store(state_view(J.state, J.n_rays), row, s);
        if (finishing) {
            if (J.out_path) J.out_path[ray] = s.path;
            if (J.out_steps) J.out_steps[ray] = inbox_steps(s);
        }
        nsteps = inbox_steps(s) - before;
    }
    if (J.step_counter) add_steps(J.step_counter, nsteps);
}

// Index of the job that owns unit u: the last j with unit_begin[j] <= u.
__device__ __forceinline__ int job_of_unit(const slt_job_table* __restrict__ T, int u) {
    int lo = 0, hi = T->n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (T->unit_begin[mid] <= u)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// The table is read through T and its work queues are updated through
// `queues` (the same allocation, disjoint bytes): with the two kept apart the
// compiler knows the job entries do not change during the launch and reads
// them with scalar loads into SGPRs.  Four wavefronts per SIMD: the few values
// that do not fit in 128 VGPRs spill only around a unit, not inside the step
// loop.  (At five, the offline kernel's occupancy, it spills inside the step
// loop and the serving mix ran 3-4 % slower.)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
advance_kernel(const slt_job_table* __restrict__ T, uint32_t* __restrict__ queues) {
    const int n_units = T->n_units;
    const int per_queue = (n_units + SLT_WORK_QUEUES - 1) / SLT_WORK_QUEUES;
    const int home = (int)(blockIdx.x % SLT_WORK_QUEUES);
    const int lane = (int)(threadIdx.x & 63);
    for (int k = 0; k < SLT_WORK_QUEUES; ++k) {
        const int q = (home + k) % SLT_WORK_QUEUES;
        const int lo = q * per_queue;
        const int len = min(n_units - lo, per_queue);
        if (len <= 0) continue;
        for (;;) {
            // every lane of the wavefront is active here
            uint32_t i = 0;
            if (lane == 0) i = atomicAdd(&queues[q * 32], 1u);
            i = (uint32_t)__builtin_amdgcn_readfirstlane((int)i);
            if (i >= (uint32_t)len) break;
            const int u = lo + (int)i;
            const int j = job_of_unit(T, u);
            const slt_job& J = T->jobs[j];
            if (J.step_end <= J.step_begin) continue;
            const int64_t t = (int64_t)(u - T->unit_begin[j]) * 64 + lane;
            if (J.grid.nchan >= 4)
                advance_unit<true>(J, t);
            else
                advance_unit<false>(J, t);
        }
    }
    // The last wavefront to drain the queues resets them, so the table can be
    // launched again without a new upload.
    if (lane == 0) {
        __threadfence();
        const uint32_t waves = gridDim.x * (blockDim.x / 64);
        if (atomicAdd(&queues[SLT_WORK_QUEUES * 32], 1u) == waves - 1) {
            for (int q = 0; q < SLT_WORK_QUEUES; ++q) atomicExch(&queues[q * 32], 0u);
            atomicExch(&queues[SLT_WORK_QUEUES * 32], 0u);
        }
    }
}

}  // namespace slt

using namespace slt;

extern "C" int slt_trace(const void* packed, const slt_grid_desc* g, const slt_trace_args* a,
                         slt_stream_t stream) {
    clear_error();
    TraceK K;
    int rc = make_dev_grid(packed, g, &K.G, /*allow_nonuniform=*/true);
    if (rc) return rc;
    if (!a || a->n_rays < 0) {
        set_error("slt_trace: invalid arguments");
        return SLT_ERR_INVALID;
    }
    if (a->n_rays == 0) return SLT_OK;
    if (!a->s0 || !a->seg_nsub || !a->seg_h || !a->save_state || a->n_seg < 1) {
        set_error("slt_trace: s0, seg_nsub, seg_h, save_state and n_seg >= 1 are required");
        return SLT_ERR_INVALID;
    }
    if (a->n_rays > (int64_t)INT32_MAX * 64) {
        set_error("slt_trace: too many rays");
        return SLT_ERR_UNSUPPORTED;
    }
    K.n = a->n_rays;
    K.s0 = a->s0;
    K.order = a->order;
    K.nseg = a->n_seg;
    K.save_every = a->save_every ? 1 : 0;
    K.seg_nsub = a->seg_nsub;
    K.seg_h = a->seg_h;
    K.save_state = a->save_state;
    K.save_tau = a->save_tau;
    K.path = a->path;
    K.steps = a->steps;
    K.total = a->total_steps;
    const int64_t nb64 = (a->n_rays + 255) / 256;
    if (nb64 > INT32_MAX / 2) {
        set_error("slt_trace: too many rays");
        return SLT_ERR_UNSUPPORTED;
    }
    K.nblocks = (int)nb64;
    const int grid = xcd_pad(K.nblocks);
    const bool kap = g->nchan >= 4;
    const bool pth = a->path != nullptr;
    hipStream_t s = (hipStream_t)stream;
    const dim3 gr(grid), bl(256);
    if (g->uniform) {
        if (kap && pth)
            hipLaunchKernelGGL((trace_kernel<true, true, true>), gr, bl, 0, s, K);
        else if (kap)
            hipLaunchKernelGGL((trace_kernel<true, false, true>), gr, bl, 0, s, K);
        else if (pth)
            hipLaunchKernelGGL((trace_kernel<false, true, true>), gr, bl, 0, s, K);
        else
            hipLaunchKernelGGL((trace_kernel<false, false, true>), gr, bl, 0, s, K);
    } else {
        if (kap && pth)
            hipLaunchKernelGGL((trace_kernel<true, true, false>), gr, bl, 0, s, K);
        else if (kap)
            hipLaunchKernelGGL((trace_kernel<true, false, false>), gr, bl, 0, s, K);
        else if (pth)
            hipLaunchKernelGGL((trace_kernel<false, true, false>), gr, bl, 0, s, K);
        else
            hipLaunchKernelGGL((trace_kernel<false, false, false>), gr, bl, 0, s, K);
    }
    return check_launch("slt_trace");
}

extern "C" int64_t slt_ray_state_bytes(int64_t n_rays) {
    return n_rays < 0 ? 0 : state_bytes(n_rays);
}

extern "C" int slt_ray_state_init(void* state, int64_t n_rays, const float* s0,
                                  const int32_t* order, const slt_grid_desc* g,
                                  slt_stream_t stream) {
    clear_error();
    if (n_rays < 0 || !g || (n_rays > 0 && (!state || !s0))) {
        set_error("slt_ray_state_init: invalid arguments");
        return SLT_ERR_INVALID;
    }
    for (int a = 0; a < 3; ++a) {
        if (g->n[a] < 2 || !(g->hi[a] > g->lo[a])) {
            set_error("slt_ray_state_init: grid axis %d needs >= 2 points and hi > lo", a);
            return SLT_ERR_INVALID;
        }
    }
    if (n_rays == 0) return SLT_OK;
    DevGrid G = {};
    fill_grid_units(g, &G);
    const int64_t blocks = (n_rays + 255) / 256;
    hipLaunchKernelGGL(state_init_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, G, state, n_rays, s0, order);
    return check_launch("slt_ray_state_init");
}

extern "C" int slt_job_table_prepare(slt_job_table* t) {
    clear_error();
    if (!t || t->n_jobs < 0 || t->n_jobs > SLT_MAX_JOBS) {
        set_error("slt_job_table_prepare: n_jobs must be in [0, %d]", SLT_MAX_JOBS);
        return SLT_ERR_INVALID;
    }
    int64_t units = 0;
    for (int j = 0; j < t->n_jobs; ++j) {
        slt_job& J = t->jobs[j];
        DevGrid G;
        int rc = make_dev_grid(J.packed, &J.grid, &G);
        if (rc) {
            char buf[400];
            snprintf(buf, sizeof(buf), "job %d: %s", j, slt_last_error());
            set_error("%s", buf);
            return rc;
        }
        if (J.n_rays < 0 || (J.n_rays > 0 && !J.state) || !J.seg_nsub || !J.seg_h ||
            J.n_seg < 1 || J.step_begin < 0 || J.step_end < J.step_begin ||
            J.step_end > J.sched_steps) {
            set_error("job %d: invalid rays, state, schedule or step range [%lld, %lld] of %lld",
                      j, (long long)J.step_begin, (long long)J.step_end,
                      (long long)J.sched_steps);
            return SLT_ERR_INVALID;
        }
        if (J.step_end == J.sched_steps && J.n_rays > 0 && !J.out_state) {
            set_error("job %d: finishing step batch needs out_state", j);
            return SLT_ERR_INVALID;
        }
        t->unit_begin[j] = (int32_t)units;
        units += (J.n_rays + 63) / 64;
        if (units > INT32_MAX / 2) {
            set_error("slt_job_table_prepare: too many rays in one table");
            return SLT_ERR_UNSUPPORTED;
        }
    }
    t->unit_begin[t->n_jobs] = (int32_t)units;
    t->n_units = (int32_t)units;
    memset(t->next_unit, 0, sizeof(t->next_unit));
    memset(t->waves_done, 0, sizeof(t->waves_done));
    return SLT_OK;
}

extern "C" int slt_advance(slt_job_table* device_table, int32_t blocks,
                           slt_stream_t stream) {
    clear_error();
    if (!device_table || blocks < kXcds || blocks % kXcds) {
        set_error("slt_advance: need a device table and a positive multiple of %d blocks", kXcds);
        return SLT_ERR_INVALID;
    }
    static_assert(offsetof(slt_job_table, next_unit) == 0 &&
                      offsetof(slt_job_table, waves_done) == SLT_WORK_QUEUES * 32 * 4,
                  "advance_kernel addresses the work queues as uint32 lines from the table base");
    hipLaunchKernelGGL(advance_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       device_table, &device_table->next_unit[0][0]);
    return check_launch("slt_advance");
}

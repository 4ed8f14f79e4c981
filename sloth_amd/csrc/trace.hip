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
//  * Position = fp64 base + fp32 offset: the offset takes each step's
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

#include "sloth_internal.h"

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

// Trilinear polynomial of the cached cell, one float4 (gx, gy, gz, kappa) per
// coefficient.
struct Cell {
    int id;  // voxel index of corner (i, j, k); -1 = none cached
    float4 k0, kx, ky, kz, kxy, kxz, kyz, kxyz;
};

__device__ __forceinline__ float4 f4sub(float4 a, float4 b) {
    return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
}

__device__ __forceinline__ void load_cell(const DevGrid& G, int id, Cell& c) {
    const float4* __restrict__ p = G.vox + id;
    const float4 c000 = p[0];
    const float4 c001 = p[G.s2];
    const float4 c010 = p[G.s1];
    const float4 c011 = p[G.s1 + G.s2];
    const float4 c100 = p[G.s0];
    const float4 c101 = p[G.s0 + G.s2];
    const float4 c110 = p[G.s0 + G.s1];
    const float4 c111 = p[G.s0 + G.s1 + G.s2];
    const float4 dz00 = f4sub(c001, c000), dz01 = f4sub(c011, c010);
    const float4 dz10 = f4sub(c101, c100), dz11 = f4sub(c111, c110);
    c.k0 = c000;
    c.kz = dz00;
    c.kx = f4sub(c100, c000);
    c.ky = f4sub(c010, c000);
    c.kxy = f4sub(f4sub(c110, c010), c.kx);
    c.kyz = f4sub(dz01, dz00);
    c.kxz = f4sub(dz10, dz00);
    c.kxyz = f4sub(f4sub(dz11, dz10), c.kyz);
    c.id = id;
}

__device__ __forceinline__ float tri(float tx, float ty, float tz, float k0, float kx, float ky,
                                     float kz, float kxy, float kxz, float kyz, float kxyz) {
    const float a = fmaf(tx, kxyz, kyz);
    const float b = fmaf(ty, a, fmaf(tx, kxz, kz));
    const float c = fmaf(tx, kxy, ky);
    return fmaf(tz, b, fmaf(ty, c, fmaf(tx, kx, k0)));
}

__device__ __forceinline__ bool in_box(const DevGrid& G, float x, float y, float z) {
    return (x >= G.lo0) & (x <= G.hi0) & (y >= G.lo1) & (y <= G.hi1) & (z >= G.lo2) &
           (z <= G.hi2);
}

// Interpolated (gx, gy, gz, kappa) at (x, y, z); zero and false outside the
// closed box.  CHECK = false when the caller already knows the point is inside.
template <bool KAPPA, bool CHECK>
__device__ __forceinline__ bool eval_rhs(const DevGrid& G, float x, float y, float z, Cell& cc,
                                         float4& out) {
    if (CHECK && !in_box(G, x, y, z)) {
        out = make_float4(0.f, 0.f, 0.f, 0.f);
        return false;
    }
    const float u0 = (x - G.lo0) * G.id0;
    const float u1 = (y - G.lo1) * G.id1;
    const float u2 = (z - G.lo2) * G.id2;
    const int i0 = min((int)u0, G.n0 - 2);  // u >= 0 inside the box: truncation == floor
    const int i1 = min((int)u1, G.n1 - 2);
    const int i2 = min((int)u2, G.n2 - 2);
    const float tx = u0 - (float)i0;
    const float ty = u1 - (float)i1;
    const float tz = u2 - (float)i2;
    const int id = (int)(__umul24(i0, G.s0) + __umul24(i1, G.s1) + __umul24(i2, G.s2));
    if (id != cc.id) load_cell(G, id, cc);
    out.x = tri(tx, ty, tz, cc.k0.x, cc.kx.x, cc.ky.x, cc.kz.x, cc.kxy.x, cc.kxz.x, cc.kyz.x,
                cc.kxyz.x);
    out.y = tri(tx, ty, tz, cc.k0.y, cc.kx.y, cc.ky.y, cc.kz.y, cc.kxy.y, cc.kxz.y, cc.kyz.y,
                cc.kxyz.y);
    out.z = tri(tx, ty, tz, cc.k0.z, cc.kx.z, cc.ky.z, cc.kz.z, cc.kxy.z, cc.kxz.z, cc.kyz.z,
                cc.kxyz.z);
    out.w = KAPPA ? tri(tx, ty, tz, cc.k0.w, cc.kx.w, cc.ky.w, cc.kz.w, cc.kxy.w, cc.kxz.w,
                        cc.kyz.w, cc.kxyz.w)
                  : 0.f;
    return true;
}

// Number of whole steps of length h (at most kmax) that a ray outside the box
// can take as straight moves without any stage point touching the box.
// Conservative: under-estimating only means an exact RK4 step with zero RHS.
__device__ __forceinline__ int ballistic_steps(const DevGrid& G, const float r[3],
                                               const float v[3], float h, int kmax) {
    const float lo[3] = {G.lo0, G.lo1, G.lo2};
    const float hi[3] = {G.hi0, G.hi1, G.hi2};
    float tmin = 0.f, tmax = INFINITY;
    bool never = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (v[a] == 0.f) {
            never |= (r[a] < lo[a]) | (r[a] > hi[a]);
        } else {
            const float inv = 1.f / v[a];
            const float ta = (lo[a] - r[a]) * inv, tb = (hi[a] - r[a]) * inv;
            tmin = fmaxf(tmin, fminf(ta, tb));
            tmax = fminf(tmax, fmaxf(ta, tb));
        }
    }
    if (never || tmax < tmin || tmax < 0.f) return kmax;
    const float m = tmin / h;
    const float safe = m - 1e-4f * m - 2.f;  // margin for fp32 rounding of tmin and m
    if (!(safe >= 1.f)) return 0;
    return safe >= (float)kmax ? kmax : (int)safe;
}

// Per-ray integration state.  In the offline kernel it lives in VGPRs for the
// whole trajectory; the serving kernel loads and stores it around each step batch.
struct RayState {
    double rb[3];     // position base [m]
    float off[3];     // position offset since the last fold into rb
    float v0[3];      // start velocity [m/s]
    float dv[3];      // velocity change since the start
    double tau;       // optical depth
    double path;      // path length [m]
    uint32_t steps;   // in-box RK4 steps
    uint32_t fold;    // RK4 steps since the start (fold phase of off into rb)
};

__device__ __forceinline__ void state_from_rays(const float* __restrict__ s0, int64_t n,
                                                int64_t ray, RayState& s) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        s.rb[c] = (double)s0[c * n + ray];
        s.off[c] = 0.f;
        s.v0[c] = s0[(3 + c) * n + ray];
        s.dv[c] = 0.f;
    }
    s.tau = 0.0;
    s.path = 0.0;
    s.steps = 0;
    s.fold = 0;
}

// Writes the ray state at save points: save_state is (S, 6, n), save_tau (S, n).
struct SaveSink {
    float* __restrict__ state;
    double* __restrict__ tau;
    int64_t n;
    int64_t ray;
    int every;  // 1: slot = save index; 0: only the last save, in slot 0
    int last;   // index of the last save point (= n_seg)

    __device__ __forceinline__ void put(int idx, const RayState& s) const {
        if (!state || (!every && idx != last)) return;
        const int64_t slot = every ? idx : 0;
        float* base = state + slot * 6 * n;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            base[c * n + ray] = (float)(s.rb[c] + (double)s.off[c]);
            base[(3 + c) * n + ray] = s.v0[c] + s.dv[c];
        }
        if (tau) tau[slot * n + ray] = s.tau;
    }
};

// Advances one ray from global step g_begin to g_end of the save schedule
// (segment s has seg_nsub[s] steps of seg_h[s] seconds and ends at save s+1).
// Saves of segments that end within [g_begin, g_end] go to the sink.
template <bool KAPPA, bool PATH>
__device__ __forceinline__ void integrate(const DevGrid& G, RayState& s,
                                          const int32_t* __restrict__ seg_nsub,
                                          const float* __restrict__ seg_h, int nseg,
                                          int64_t g_begin, int64_t g_end, const SaveSink& sink) {
    Cell cc;
    cc.id = -1;
    float rbf[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) rbf[c] = (float)s.rb[c];
    int64_t seg_lo = 0;
    for (int seg = 0; seg < nseg && seg_lo <= g_end; ++seg) {
        const int nsub = seg_nsub[seg];
        const int64_t seg_hi = seg_lo + nsub;
        if (seg_hi >= g_begin) {
            const float h = seg_h[seg];
            const float h6 = h * (1.0f / 6.0f);
            const float hh = 0.5f * h;
            const float hh6 = h * h6;
            int k = (int)(g_begin > seg_lo ? g_begin - seg_lo : 0);
            const int kend = (int)((g_end < seg_hi ? g_end : seg_hi) - seg_lo);
            while (k < kend) {
                float rh[3], vh[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    rh[c] = rbf[c] + s.off[c];
                    vh[c] = s.v0[c] + s.dv[c];
                }
                const bool in1 = in_box(G, rh[0], rh[1], rh[2]);
                if (!in1) {
                    const int m = ballistic_steps(G, rh, vh, h, kend - k);
                    if (m > 0) {
                        const double T = (double)m * (double)h;
                        double sp2 = 0.0;
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const double vc = (double)s.v0[c] + (double)s.dv[c];
                            s.rb[c] += T * vc;
                            rbf[c] = (float)s.rb[c];
                            sp2 += vc * vc;
                        }
                        if (PATH) s.path += T * sqrt(sp2);
                        k += m;
                        continue;
                    }
                }
                // One RK4 step.
                float4 A1 = make_float4(0.f, 0.f, 0.f, 0.f), A2, A3, A4;
                bool in = in1;
                if (in1) eval_rhs<KAPPA, false>(G, rh[0], rh[1], rh[2], cc, A1);
                const float v2x = fmaf(hh, A1.x, vh[0]), v2y = fmaf(hh, A1.y, vh[1]),
                            v2z = fmaf(hh, A1.z, vh[2]);
                in |= eval_rhs<KAPPA, true>(G, fmaf(hh, vh[0], rh[0]), fmaf(hh, vh[1], rh[1]),
                                            fmaf(hh, vh[2], rh[2]), cc, A2);
                const float v3x = fmaf(hh, A2.x, vh[0]), v3y = fmaf(hh, A2.y, vh[1]),
                            v3z = fmaf(hh, A2.z, vh[2]);
                in |= eval_rhs<KAPPA, true>(G, fmaf(hh, v2x, rh[0]), fmaf(hh, v2y, rh[1]),
                                            fmaf(hh, v2z, rh[2]), cc, A3);
                const float v4x = fmaf(h, A3.x, vh[0]), v4y = fmaf(h, A3.y, vh[1]),
                            v4z = fmaf(h, A3.z, vh[2]);
                in |= eval_rhs<KAPPA, true>(G, fmaf(h, v3x, rh[0]), fmaf(h, v3y, rh[1]),
                                            fmaf(h, v3z, rh[2]), cc, A4);
                if (PATH) {
                    const float s1 = __builtin_amdgcn_sqrtf(
                        fmaf(vh[0], vh[0], fmaf(vh[1], vh[1], vh[2] * vh[2])));
                    const float s2 = __builtin_amdgcn_sqrtf(fmaf(v2x, v2x, fmaf(v2y, v2y, v2z * v2z)));
                    const float s3 = __builtin_amdgcn_sqrtf(fmaf(v3x, v3x, fmaf(v3y, v3y, v3z * v3z)));
                    const float s4 = __builtin_amdgcn_sqrtf(fmaf(v4x, v4x, fmaf(v4y, v4y, v4z * v4z)));
                    s.path += (double)(h6 * (s1 + 2.f * (s2 + s3) + s4));
                }
                // r += h v + h^2/6 (A1 + A2 + A3)   (= h/6 (v1 + 2 v2 + 2 v3 + v4))
                // v += h/6 (A1 + 2 A2 + 2 A3 + A4)
                s.off[0] += fmaf(h, vh[0], hh6 * (A1.x + A2.x + A3.x));
                s.off[1] += fmaf(h, vh[1], hh6 * (A1.y + A2.y + A3.y));
                s.off[2] += fmaf(h, vh[2], hh6 * (A1.z + A2.z + A3.z));
                s.dv[0] += h6 * (A1.x + 2.f * (A2.x + A3.x) + A4.x);
                s.dv[1] += h6 * (A1.y + 2.f * (A2.y + A3.y) + A4.y);
                s.dv[2] += h6 * (A1.z + 2.f * (A2.z + A3.z) + A4.z);
                if (KAPPA) s.tau += (double)(h6 * (A1.w + 2.f * (A2.w + A3.w) + A4.w));
                s.steps += in ? 1u : 0u;
                ++k;
                if ((++s.fold & 7u) == 0) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        s.rb[c] += (double)s.off[c];
                        s.off[c] = 0.f;
                        rbf[c] = (float)s.rb[c];
                    }
                }
            }
            if (seg_hi <= g_end) sink.put(seg + 1, s);
        }
        seg_lo = seg_hi;
    }
}

__device__ __forceinline__ void add_steps(unsigned long long* total, uint32_t nsteps) {
    unsigned long long v = nsteps;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(total, v);
}

template <bool KAPPA, bool PATH>
__global__ void __launch_bounds__(256) trace_kernel(TraceK K) {
    const int lb = xcd_block(blockIdx.x, gridDim.x, K.nblocks);
    if (lb < 0) return;
    const int64_t t = (int64_t)lb * blockDim.x + threadIdx.x;
    uint32_t nsteps = 0;
    if (t < K.n) {
        const int64_t n = K.n;
        const int64_t ray = K.order ? (int64_t)K.order[t] : t;
        RayState s;
        state_from_rays(K.s0, n, ray, s);
        SaveSink sink{K.save_state, K.save_tau, n, ray, K.save_every, K.nseg};
        sink.put(0, s);
        int64_t total = 0;
        for (int q = 0; q < K.nseg; ++q) total += K.seg_nsub[q];
        integrate<KAPPA, PATH>(K.G, s, K.seg_nsub, K.seg_h, K.nseg, 0, total, sink);
        if (PATH && K.path) K.path[ray] = s.path;
        if (K.steps) K.steps[ray] = s.steps;
        nsteps = s.steps;
    }
    if (K.total) add_steps(K.total, nsteps);
}

// ---------------------------------------------------------------------------
// Resumable ray state (serving): structure-of-arrays in one buffer.
//   rb[3] f64 | tau f64 | path f64 | off[3] f32 | v0[3] f32 | dv[3] f32 | steps u32 | fold u32
// ---------------------------------------------------------------------------
struct StateView {
    double* rb[3];
    double* tau;
    double* path;
    float* off[3];
    float* v0[3];
    float* dv[3];
    uint32_t* steps;
    uint32_t* fold;
};

__host__ __device__ inline int64_t state_bytes(int64_t n) { return n * (5 * 8 + 9 * 4 + 2 * 4); }

__device__ __forceinline__ StateView state_view(void* base, int64_t n) {
    StateView v;
    double* d = static_cast<double*>(base);
    for (int c = 0; c < 3; ++c) v.rb[c] = d + c * n;
    v.tau = d + 3 * n;
    v.path = d + 4 * n;
    float* f = reinterpret_cast<float*>(d + 5 * n);
    for (int c = 0; c < 3; ++c) {
        v.off[c] = f + c * n;
        v.v0[c] = f + (3 + c) * n;
        v.dv[c] = f + (6 + c) * n;
    }
    uint32_t* u = reinterpret_cast<uint32_t*>(f + 9 * n);
    v.steps = u;
    v.fold = u + n;
    return v;
}

__device__ __forceinline__ void state_load(const StateView& v, int64_t i, RayState& s) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        s.rb[c] = v.rb[c][i];
        s.off[c] = v.off[c][i];
        s.v0[c] = v.v0[c][i];
        s.dv[c] = v.dv[c][i];
    }
    s.tau = v.tau[i];
    s.path = v.path[i];
    s.steps = v.steps[i];
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
    v.steps[i] = s.steps;
    v.fold[i] = s.fold;
}

__global__ void __launch_bounds__(256) state_init_kernel(void* base, int64_t n,
                                                         const float* __restrict__ s0,
                                                         const int32_t* __restrict__ order) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t ray = order ? (int64_t)order[t] : t;
    RayState s;
    state_from_rays(s0, n, ray, s);
    const StateView v = state_view(base, n);
    state_store(v, t, s);
#pragma unroll
    for (int c = 0; c < 3; ++c) v.v0[c][t] = s.v0[c];
}

// Work unit = 256 consecutive (sorted) rays of one job.  Units are split into
// kXcds contiguous chunks; block b (on XCD b % kXcds) walks chunk b % kXcds.
template <bool KAPPA>
__device__ void advance_unit(const slt_job& J, int64_t first) {
    const int64_t t = first + threadIdx.x;
    uint32_t nsteps = 0;
    if (t < J.n_rays) {
        DevGrid G;
        make_dev_grid_device(J.packed, &J.grid, &G);
        const StateView v = state_view(J.state, J.n_rays);
        RayState s;
        state_load(v, t, s);
        const uint32_t before = s.steps;
        const bool finishing = J.step_end >= J.sched_steps;
        const int64_t ray = J.order ? (int64_t)J.order[t] : t;
        SaveSink sink{finishing ? J.out_state : nullptr, J.out_tau, J.n_rays, ray, 0, J.n_seg};
        integrate<KAPPA, true>(G, s, J.seg_nsub, J.seg_h, J.n_seg, J.step_begin, J.step_end,
                               sink);
        state_store(v, t, s);
        if (finishing) {
            if (J.out_path) J.out_path[ray] = s.path;
            if (J.out_steps) J.out_steps[ray] = s.steps;
        }
        nsteps = s.steps - before;
    }
    if (J.step_counter) add_steps(J.step_counter, nsteps);
}

__global__ void __launch_bounds__(256) advance_kernel(const slt_job_table* __restrict__ T) {
    const int n_units = T->n_units;
    const int n_jobs = T->n_jobs;
    const int chunk = kXcds > 1 ? (int)(blockIdx.x % kXcds) : 0;
    const int per_chunk = (n_units + kXcds - 1) / kXcds;
    const int u_lo = chunk * per_chunk;
    const int u_hi = min(n_units, u_lo + per_chunk);
    const int stride = (int)(gridDim.x / kXcds);
    for (int u = u_lo + (int)(blockIdx.x / kXcds); u < u_hi; u += stride) {
        int j = 0;
        while (j + 1 < n_jobs && T->unit_begin[j + 1] <= u) ++j;
        const slt_job& J = T->jobs[j];
        if (J.step_end <= J.step_begin) continue;
        const int64_t first = (int64_t)(u - T->unit_begin[j]) * 256;
        if (J.grid.nchan >= 4)
            advance_unit<true>(J, first);
        else
            advance_unit<false>(J, first);
    }
}

}  // namespace slt

using namespace slt;

extern "C" int slt_trace(const void* packed, const slt_grid_desc* g, const slt_trace_args* a,
                         slt_stream_t stream) {
    clear_error();
    TraceK K;
    int rc = make_dev_grid(packed, g, &K.G);
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
    if (kap && pth)
        hipLaunchKernelGGL((trace_kernel<true, true>), dim3(grid), dim3(256), 0, s, K);
    else if (kap)
        hipLaunchKernelGGL((trace_kernel<true, false>), dim3(grid), dim3(256), 0, s, K);
    else if (pth)
        hipLaunchKernelGGL((trace_kernel<false, true>), dim3(grid), dim3(256), 0, s, K);
    else
        hipLaunchKernelGGL((trace_kernel<false, false>), dim3(grid), dim3(256), 0, s, K);
    return check_launch("slt_trace");
}

extern "C" int64_t slt_ray_state_bytes(int64_t n_rays) {
    return n_rays < 0 ? 0 : state_bytes(n_rays);
}

extern "C" int slt_ray_state_init(void* state, int64_t n_rays, const float* s0,
                                  const int32_t* order, slt_stream_t stream) {
    clear_error();
    if (n_rays < 0 || (n_rays > 0 && (!state || !s0))) {
        set_error("slt_ray_state_init: invalid arguments");
        return SLT_ERR_INVALID;
    }
    if (n_rays == 0) return SLT_OK;
    const int64_t blocks = (n_rays + 255) / 256;
    hipLaunchKernelGGL(state_init_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, state, n_rays, s0, order);
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
        units += (J.n_rays + 255) / 256;
        if (units > INT32_MAX / 2) {
            set_error("slt_job_table_prepare: too many rays in one table");
            return SLT_ERR_UNSUPPORTED;
        }
    }
    t->unit_begin[t->n_jobs] = (int32_t)units;
    t->n_units = (int32_t)units;
    return SLT_OK;
}

extern "C" int slt_advance(const slt_job_table* device_table, int32_t blocks,
                           slt_stream_t stream) {
    clear_error();
    if (!device_table || blocks < kXcds || blocks % kXcds) {
        set_error("slt_advance: need a device table and a positive multiple of %d blocks", kXcds);
        return SLT_ERR_INVALID;
    }
    hipLaunchKernelGGL(advance_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       device_table);
    return check_launch("slt_advance");
}

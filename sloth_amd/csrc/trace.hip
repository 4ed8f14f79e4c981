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

template <bool KAPPA, bool PATH>
__global__ void __launch_bounds__(256) trace_kernel(TraceK K) {
    const int lb = xcd_block(blockIdx.x, gridDim.x, K.nblocks);
    if (lb < 0) return;
    const int64_t t = (int64_t)lb * blockDim.x + threadIdx.x;
    const bool valid = t < K.n;
    const DevGrid& G = K.G;
    uint32_t nsteps = 0;
    if (valid) {
        const int64_t n = K.n;
        const int64_t ray = K.order ? (int64_t)K.order[t] : t;
        double rb[3];           // position base [m]
        float rbf[3];           // rb rounded to fp32
        float off[3] = {0.f, 0.f, 0.f};  // position offset since the last fold
        float v0[3];            // start velocity [m/s]
        float dv[3] = {0.f, 0.f, 0.f};   // velocity change since the start
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            rbf[c] = K.s0[c * n + ray];
            rb[c] = (double)rbf[c];
            v0[c] = K.s0[(3 + c) * n + ray];
        }
        double tau = 0.0, path = 0.0;
        Cell cc;
        cc.id = -1;
        uint32_t since_fold = 0;
        if (K.save_every) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                K.save_state[c * n + ray] = rbf[c];
                K.save_state[(3 + c) * n + ray] = v0[c];
            }
            if (K.save_tau) K.save_tau[ray] = 0.0;
        }
        for (int seg = 0; seg < K.nseg; ++seg) {
            const float h = K.seg_h[seg];
            const int nsub = K.seg_nsub[seg];
            const float h6 = h * (1.0f / 6.0f);
            const float hh = 0.5f * h;
            const float hh6 = h * h6;
            int k = 0;
            while (k < nsub) {
                float rh[3], vh[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    rh[c] = rbf[c] + off[c];
                    vh[c] = v0[c] + dv[c];
                }
                const bool in1 = in_box(G, rh[0], rh[1], rh[2]);
                if (!in1) {
                    const int m = ballistic_steps(G, rh, vh, h, nsub - k);
                    if (m > 0) {
                        const double T = (double)m * (double)h;
                        double sp2 = 0.0;
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const double vc = (double)v0[c] + (double)dv[c];
                            rb[c] += T * vc;
                            rbf[c] = (float)rb[c];
                            sp2 += vc * vc;
                        }
                        if (PATH) path += T * sqrt(sp2);
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
                    path += (double)(h6 * (s1 + 2.f * (s2 + s3) + s4));
                }
                // r += h v + h^2/6 (A1 + A2 + A3)   (= h/6 (v1 + 2 v2 + 2 v3 + v4))
                // v += h/6 (A1 + 2 A2 + 2 A3 + A4)
                off[0] += fmaf(h, vh[0], hh6 * (A1.x + A2.x + A3.x));
                off[1] += fmaf(h, vh[1], hh6 * (A1.y + A2.y + A3.y));
                off[2] += fmaf(h, vh[2], hh6 * (A1.z + A2.z + A3.z));
                dv[0] += h6 * (A1.x + 2.f * (A2.x + A3.x) + A4.x);
                dv[1] += h6 * (A1.y + 2.f * (A2.y + A3.y) + A4.y);
                dv[2] += h6 * (A1.z + 2.f * (A2.z + A3.z) + A4.z);
                if (KAPPA) tau += (double)(h6 * (A1.w + 2.f * (A2.w + A3.w) + A4.w));
                nsteps += in ? 1u : 0u;
                ++k;
                if ((++since_fold & 7u) == 0) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        rb[c] += (double)off[c];
                        off[c] = 0.f;
                        rbf[c] = (float)rb[c];
                    }
                }
            }
            if (K.save_every || seg == K.nseg - 1) {
                const int64_t slot = K.save_every ? (int64_t)(seg + 1) : 0;
                float* base = K.save_state + slot * 6 * n;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    base[c * n + ray] = (float)(rb[c] + (double)off[c]);
                    base[(3 + c) * n + ray] = v0[c] + dv[c];
                }
                if (K.save_tau) K.save_tau[slot * n + ray] = tau;
            }
        }
        if (PATH && K.path) K.path[ray] = path;
        if (K.steps) K.steps[ray] = nsteps;
    }
    if (K.total) {
        unsigned long long v = nsteps;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(K.total, v);
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

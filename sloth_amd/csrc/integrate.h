// Per-ray RK4 integration core shared by the offline trace, the serving step
// batches and the adjoint (checkpointed forward and tape replay).
#pragma once

#include <math.h>

#include "sloth_internal.h"

namespace slt {

// Trilinear polynomial of a cell, one float4 (gx, gy, gz, kappa) per
// coefficient.
struct CellPoly {
    float4 k0, kx, ky, kz, kxy, kxz, kyz, kxyz;
};

// Cell cached by the forward integration: polynomial + the cell's lower
// corner in grid units (i, j, k as floats; NaN = none cached) + whether the
// cell touches a face of the grid.  A stage point whose local coordinates
// u - corner all lie in [0, 1) is in the cached cell, so the common case needs
// no voxel index; if the cell is interior the point is also a whole cell away
// from every face and needs no box test.  Points in a face cell still take the
// box test in physical coordinates, so a point an ulp outside a face is outside
// whatever u rounds to.
struct FwdCell : CellPoly {
    float c0, c1, c2;
    bool face;
    float w0, w1, w2;  // non-uniform grids: 1 / (cell width in grid units) per axis
};

// Cell of u along an axis with grid points kn[0..n) in grid units (ascending,
// kn[0] = 0, kn[n-1] = n-1): the largest i in [0, n-2] with kn[i] <= u.  In grid
// units floor(u) is the cell of the linspace, so the search starts there and walks
// up to kKnotWalk cells each way (enough for grids whose points stray a few cells
// from the linspace), then falls back to a binary search of what is left.
constexpr int kKnotWalk = 4;
__device__ __forceinline__ int knot_cell(const float* __restrict__ kn, int n, float u) {
    int i = min(max((int)u, 0), n - 2);
    int lo = 0, hi = n - 2;  // the answer stays in [lo, hi]
#pragma unroll 1
    for (int k = 0; k < kKnotWalk; ++k) {
        if (kn[i] > u) {
            hi = i - 1;
            i = max(i - 1, lo);
        } else if (i < n - 2 && kn[i + 1] <= u) {
            lo = i + 1;
            i = lo;
        } else {
            return i;
        }
    }
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (kn[mid] <= u)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ float4 f4sub(float4 a, float4 b) {
    return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
}

typedef float slt_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 voxel(const SLT_GLOBAL slt_f4* __restrict__ p, int i) {
    const slt_f4 v = p[i];
    return make_float4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ float4 voxel_at(const SLT_GLOBAL char* p) {
    const slt_f4 v = *reinterpret_cast<const SLT_GLOBAL slt_f4*>(p);
    return make_float4(v.x, v.y, v.z, v.w);
}

// The cell polynomial of corner voxel id (its 8 voxels).  A packed grid below
// 4 GiB is addressed as one uniform base per corner plus the lane's 32-bit byte
// offset, which each load takes as its VGPR offset: no 64-bit address
// arithmetic per lane.  Larger grids take 64-bit addresses.
__device__ __forceinline__ void load_cell(const DevGrid& G, int id, CellPoly& c) {
    float4 c000, c001, c010, c011, c100, c101, c110, c111;
    if (G.off32) {
        const uint32_t off = (uint32_t)id << 4;
        const SLT_GLOBAL char* b = reinterpret_cast<const SLT_GLOBAL char*>(to_global(G.vox));
        c000 = voxel_at(b + off);
        c001 = voxel_at(b + G.b2 + off);
        c010 = voxel_at(b + G.b1 + off);
        c011 = voxel_at(b + (G.b1 + G.b2) + off);
        c100 = voxel_at(b + G.b0 + off);
        c101 = voxel_at(b + (G.b0 + G.b2) + off);
        c110 = voxel_at(b + (G.b0 + G.b1) + off);
        c111 = voxel_at(b + (G.b0 + G.b1 + G.b2) + off);
    } else {
        const SLT_GLOBAL slt_f4* __restrict__ p =
            to_global(reinterpret_cast<const slt_f4*>(G.vox)) + id;
        c000 = voxel(p, 0);
        c001 = voxel(p, G.s2);
        c010 = voxel(p, G.s1);
        c011 = voxel(p, G.s1 + G.s2);
        c100 = voxel(p, G.s0);
        c101 = voxel(p, G.s0 + G.s2);
        c110 = voxel(p, G.s0 + G.s1);
        c111 = voxel(p, G.s0 + G.s1 + G.s2);
    }
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
}

__device__ __forceinline__ float tri(float tx, float ty, float tz, float k0, float kx, float ky,
                                     float kz, float kxy, float kxz, float kyz, float kxyz) {
    const float a = fmaf(tx, kxyz, kyz);
    const float b = fmaf(ty, a, fmaf(tx, kxz, kz));
    const float c = fmaf(tx, kxy, ky);
    return fmaf(tz, b, fmaf(ty, c, fmaf(tx, kx, k0)));
}

// The closed box in grid units: 0 <= u <= n - 1 on every axis.
__device__ __forceinline__ bool in_box(const DevGrid& G, float u0, float u1, float u2) {
    return (u0 >= 0.f) & (u0 <= G.ub0) & (u1 >= 0.f) & (u1 <= G.ub1) & (u2 >= 0.f) &
           (u2 <= G.ub2);
}

// (gx, gy, gz, kappa) of the cell polynomial at local coordinates t.
template <bool KAPPA>
__device__ __forceinline__ void poly_eval(const CellPoly& cc, float tx, float ty, float tz,
                                          float4& out) {
    out.x = tri(tx, ty, tz, cc.k0.x, cc.kx.x, cc.ky.x, cc.kz.x, cc.kxy.x, cc.kxz.x, cc.kyz.x,
                cc.kxyz.x);
    out.y = tri(tx, ty, tz, cc.k0.y, cc.kx.y, cc.ky.y, cc.kz.y, cc.kxy.y, cc.kxz.y, cc.kyz.y,
                cc.kxyz.y);
    out.z = tri(tx, ty, tz, cc.k0.z, cc.kx.z, cc.ky.z, cc.kz.z, cc.kxy.z, cc.kxz.z, cc.kyz.z,
                cc.kxyz.z);
    out.w = KAPPA ? tri(tx, ty, tz, cc.k0.w, cc.kx.w, cc.ky.w, cc.kz.w, cc.kxy.w, cc.kxz.w,
                        cc.kyz.w, cc.kxyz.w)
                  : 0.f;
}

// t in [0, 1) (false for -0.0, negatives and NaN): one unsigned compare of the
// bits.  For t = u - corner this holds exactly when floor(u) == corner (the
// subtraction is exact there), so the fast path picks the same cell as the
// index computation and the result does not depend on which cell was cached
// (a point on the upper face of the top cell takes the index path).
__device__ __forceinline__ unsigned unit_interval(float t) {
    return __float_as_uint(t) < 0x3f800000u ? 1u : 0u;
}

// Interpolated (gx, gy, gz, kappa) at the grid-unit point (u0, u1, u2); zero and
// false outside the closed box.  CHECK = false when the caller already knows the
// point is inside.  UNI: cells are unit intervals of u (uniform grid); else they
// lie between the axis' knots and t is the fraction of the cell's width.
template <bool KAPPA, bool CHECK, bool UNI = true>
__device__ __forceinline__ bool eval_rhs(const DevGrid& G, float u0, float u1, float u2,
                                         FwdCell& cc, float4& out) {
    float tx = UNI ? u0 - cc.c0 : (u0 - cc.c0) * cc.w0;
    float ty = UNI ? u1 - cc.c1 : (u1 - cc.c1) * cc.w1;
    float tz = UNI ? u2 - cc.c2 : (u2 - cc.c2) * cc.w2;
    bool inside = true;
    if (!(unit_interval(tx) & unit_interval(ty) & unit_interval(tz))) {
        if (CHECK && !in_box(G, u0, u1, u2)) {
            inside = false;
        } else {
            int i0, i1, i2;
            if (UNI) {
                // inside the box u >= 0, so truncation is floor
                i0 = min((int)u0, G.n0 - 2);
                i1 = min((int)u1, G.n1 - 2);
                i2 = min((int)u2, G.n2 - 2);
                cc.c0 = (float)i0;
                cc.c1 = (float)i1;
                cc.c2 = (float)i2;
                tx = u0 - cc.c0;
                ty = u1 - cc.c1;
                tz = u2 - cc.c2;
            } else {
                i0 = knot_cell(G.kn0, G.n0, u0);
                i1 = knot_cell(G.kn1, G.n1, u1);
                i2 = knot_cell(G.kn2, G.n2, u2);
                cc.c0 = G.kn0[i0];
                cc.c1 = G.kn1[i1];
                cc.c2 = G.kn2[i2];
                cc.w0 = 1.0f / (G.kn0[i0 + 1] - cc.c0);
                cc.w1 = 1.0f / (G.kn1[i1 + 1] - cc.c1);
                cc.w2 = 1.0f / (G.kn2[i2 + 1] - cc.c2);
                tx = (u0 - cc.c0) * cc.w0;
                ty = (u1 - cc.c1) * cc.w1;
                tz = (u2 - cc.c2) * cc.w2;
            }
            cc.face = (min(min(i0, i1), i2) == 0) | (i0 == G.n0 - 2) | (i1 == G.n1 - 2) |
                      (i2 == G.n2 - 2);
            load_cell(G, (int)(__umul24(i0, G.s0) + __umul24(i1, G.s1) + __umul24(i2, G.s2)),
                      cc);
        }
    } else if (CHECK && cc.face && !in_box(G, u0, u1, u2)) {
        inside = false;
    }
    // evaluated for every point; a point outside the box then gets zero in a branch
    // that is skipped unless a lane needs it
    poly_eval<KAPPA>(cc, tx, ty, tz, out);
    if (CHECK && __builtin_expect(!inside, 0)) out = make_float4(0.f, 0.f, 0.f, 0.f);
    return inside;
}

// Number of whole steps of length h (at most kmax) that a ray outside the box
// can take as straight moves without any stage point touching the box.
// Conservative: under-estimating only means an exact RK4 step with zero RHS.
__device__ __forceinline__ int ballistic_steps(const DevGrid& G, const float r[3],
                                               const float v[3], float h, int kmax) {
    const float lo[3] = {0.f, 0.f, 0.f};
    const float hi[3] = {G.ub0, G.ub1, G.ub2};
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
    double rb[3];     // position base [grid units]
    float off[3];     // position offset since the last fold into rb
    float v0[3];      // start velocity [grid units/s]
    float dv[3];      // velocity change since the start
    double tau;       // optical depth
    double path;      // path length [m]
    uint32_t nout;    // RK4 steps with no stage point in the box
    uint32_t fold;    // RK4 steps since the start (also the fold phase of off into rb)
};

// In-box RK4 steps of a ray so far.
__device__ __forceinline__ uint32_t inbox_steps(const RayState& s) { return s.fold - s.nout; }

// Start velocity component c in grid units/s (fp32), as the state holds it.
__device__ __forceinline__ float velocity_units(const DevGrid& G, int c, float v) {
    return (float)((double)v * G.isp[c]);
}

__device__ __forceinline__ void state_from_rays(const DevGrid& G, const float* __restrict__ s0,
                                                int64_t n, int64_t ray, RayState& s) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        s.rb[c] = ((double)s0[c * n + ray] - G.org[c]) * G.isp[c];
        s.off[c] = 0.f;
        s.v0[c] = velocity_units(G, c, s0[(3 + c) * n + ray]);
        s.dv[c] = 0.f;
    }
    s.tau = 0.0;
    s.path = 0.0;
    s.nout = 0;
    s.fold = 0;
}

// Writes the ray state at save points, in metres and m/s: save_state is
// (S, 6, n), save_tau (S, n).  The velocity written is the caller's start
// velocity (v0m[c * v0_ld + v0_i], m/s) plus the change, so a ray that meets no
// gradient keeps its input velocity bit for bit.
struct SaveSink {
    float* __restrict__ state;
    double* __restrict__ tau;
    int64_t n;
    int64_t ray;
    int every;  // 1: slot = save index; 0: only the last save, in slot 0
    int last;   // index of the last save point (= n_seg)
    const float* __restrict__ v0m;
    int64_t v0_ld;
    int64_t v0_i;

    __device__ __forceinline__ void put(int idx, const RayState& s, const DevGrid& G) const {
        if (!state || (!every && idx != last)) return;
        const int64_t slot = every ? idx : 0;
        float* base = state + slot * 6 * n;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            base[c * n + ray] = (float)(G.org[c] + (s.rb[c] + (double)s.off[c]) * G.dsp[c]);
            base[(3 + c) * n + ray] =
                (float)((double)v0m[c * v0_ld + v0_i] + (double)s.dv[c] * G.dsp[c]);
        }
        if (tau) tau[slot * n + ray] = s.tau;
    }
};

// Records nothing (the forward passes).  The adjoint replays intervals with a
// tape that records every move: rk(position, velocity, h) before an RK4 step
// and straight(T) for a straight-line move of duration T.
struct NoTape {
    __device__ __forceinline__ void rk(const float*, const float*, float) {}
    __device__ __forceinline__ void straight(float) {}
};

// Advances one ray from global step g_begin to g_end of the save schedule
// (segment s has seg_nsub[s] steps of seg_h[s] seconds and ends at save s+1).
// Saves of segments that end within [g_begin, g_end] go to the sink.
template <bool KAPPA, bool PATH, class Tape, bool UNI = true>
__device__ __forceinline__ void integrate(const DevGrid& G, RayState& s,
                                          const int32_t* __restrict__ seg_nsub,
                                          const float* __restrict__ seg_h, int nseg,
                                          int64_t g_begin, int64_t g_end, const SaveSink& sink,
                                          Tape& tape) {
    FwdCell cc;
    cc.c0 = cc.c1 = cc.c2 = __builtin_nanf("");
    cc.face = true;
    cc.w0 = cc.w1 = cc.w2 = 1.0f;
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
                // The first stage doubles as the box test of the step's start point:
                // zero RHS and false outside the box.
                float4 A1, A2, A3, A4;
                const bool in1 = eval_rhs<KAPPA, true, UNI>(G, rh[0], rh[1], rh[2], cc, A1);
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
                            const double vm = vc * G.dsp[c];  // m/s
                            sp2 += vm * vm;
                        }
                        if (PATH) s.path += T * sqrt(sp2);
                        tape.straight((float)T);
                        k += m;
                        continue;
                    }
                }
                // One RK4 step.
                tape.rk(rh, vh, h);
                bool in = in1;
                const float v2x = fmaf(hh, A1.x, vh[0]), v2y = fmaf(hh, A1.y, vh[1]),
                            v2z = fmaf(hh, A1.z, vh[2]);
                in |= eval_rhs<KAPPA, true, UNI>(G, fmaf(hh, vh[0], rh[0]), fmaf(hh, vh[1], rh[1]),
                                            fmaf(hh, vh[2], rh[2]), cc, A2);
                const float v3x = fmaf(hh, A2.x, vh[0]), v3y = fmaf(hh, A2.y, vh[1]),
                            v3z = fmaf(hh, A2.z, vh[2]);
                in |= eval_rhs<KAPPA, true, UNI>(G, fmaf(hh, v2x, rh[0]), fmaf(hh, v2y, rh[1]),
                                            fmaf(hh, v2z, rh[2]), cc, A3);
                in |= eval_rhs<KAPPA, true, UNI>(G, fmaf(h, v3x, rh[0]), fmaf(h, v3y, rh[1]),
                                            fmaf(h, v3z, rh[2]), cc, A4);
                // r += h v + h^2/6 (A1 + A2 + A3)   (= h/6 (v1 + 2 v2 + 2 v3 + v4))
                // v += h/6 (A1 + 2 A2 + 2 A3 + A4)
                // A2 + A3 enters both updates
                const float sx = A2.x + A3.x, sy = A2.y + A3.y, sz = A2.z + A3.z;
                const float dx = fmaf(h, vh[0], hh6 * (A1.x + sx));
                const float dy = fmaf(h, vh[1], hh6 * (A1.y + sy));
                const float dz = fmaf(h, vh[2], hh6 * (A1.z + sz));
                s.off[0] += dx;
                s.off[1] += dy;
                s.off[2] += dz;
                // path length: the polyline through the step end points (in metres)
                if (PATH) {
                    const float ex = dx * G.sp0, ey = dy * G.sp1, ez = dz * G.sp2;
                    s.path += (double)__builtin_amdgcn_sqrtf(fmaf(ex, ex, fmaf(ey, ey, ez * ez)));
                }
                s.dv[0] += h6 * (A1.x + 2.f * sx + A4.x);
                s.dv[1] += h6 * (A1.y + 2.f * sy + A4.y);
                s.dv[2] += h6 * (A1.z + 2.f * sz + A4.z);
                if (KAPPA) s.tau += (double)(h6 * (A1.w + 2.f * (A2.w + A3.w) + A4.w));
                // every RK4 step advances fold; the rare step with no stage point in the
                // box is counted on its own (in-box steps = fold - nout)
                if (__builtin_expect(!in, 0)) ++s.nout;
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
            if (seg_hi <= g_end) sink.put(seg + 1, s, G);
        }
        seg_lo = seg_hi;
    }
}

template <bool KAPPA, bool PATH, bool UNI = true>
__device__ __forceinline__ void integrate(const DevGrid& G, RayState& s,
                                          const int32_t* __restrict__ seg_nsub,
                                          const float* __restrict__ seg_h, int nseg,
                                          int64_t g_begin, int64_t g_end, const SaveSink& sink) {
    NoTape none;
    integrate<KAPPA, PATH, NoTape, UNI>(G, s, seg_nsub, seg_h, nseg, g_begin, g_end, sink,
                                        none);
}

__device__ __forceinline__ void add_steps(unsigned long long* total, uint32_t nsteps) {
    unsigned long long v = nsteps;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(total, v);
}


}  // namespace slt

// Internal declarations shared by the HIP translation units of libsloth_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sloth_hip.h"

namespace slt {

// Thread-local error message (see slt_last_error).
void set_error(const char* fmt, ...);
void clear_error();

// Returns SLT_OK after a successful launch, else records the HIP error.
int check_launch(const char* what);

// Device view of a packed grid.  Indices fit in 32 bits and every stride and
// extent in 24 (so voxel offsets use 24-bit multiplies): the host rejects
// larger grids.
//
// The integrators work in grid units: a position is u = (x - lo) / d per axis
// (d = (hi - lo) / (n - 1), the cell size), so node i sits at u = i and the
// box is [0, n - 1]; velocities are v / d and the packed gradient channels
// hold the gradient times 1 / d (slt_grid_pack).  A stage point is then its own
// grid coordinate and needs no scaling before the cell lookup.  Conversions to
// and from metres happen once per ray (start state, save points) in fp64.
struct DevGrid {
    const float4* __restrict__ vox;
    int n0, n1, n2;
    int s0, s1, s2;
    int64_t b0, b1, b2;   // strides in bytes
    int off32;            // packed grid below 4 GiB: byte offsets fit in 32 bits
    float ub0, ub1, ub2;  // upper box face in grid units: n - 1 (the lower one is 0)
    float sp0, sp1, sp2;  // cell size [m] (path length in fp32)
    double org[3];        // grid origin lo [m]
    double dsp[3];        // cell size d [m]
    double isp[3];        // 1 / d [1/m]
    int nchan;
    // non-uniform grid: each axis' grid points in grid units (device), else null
    const float* __restrict__ kn0;
    const float* __restrict__ kn1;
    const float* __restrict__ kn2;
};

// Unit conversion fields of a DevGrid from a descriptor (host or device).
__host__ __device__ inline void fill_grid_units(const slt_grid_desc* g, DevGrid* d) {
    const int n[3] = {g->n[0], g->n[1], g->n[2]};
    for (int a = 0; a < 3; ++a) {
        const double lo = (double)g->lo[a], hi = (double)g->hi[a];
        const double m = (double)(n[a] - 1);
        d->org[a] = lo;
        d->dsp[a] = (hi - lo) / m;
        d->isp[a] = m / (hi - lo);
    }
    d->ub0 = (float)(n[0] - 1); d->ub1 = (float)(n[1] - 1); d->ub2 = (float)(n[2] - 1);
    d->sp0 = (float)d->dsp[0]; d->sp1 = (float)d->dsp[1]; d->sp2 = (float)d->dsp[2];
}

// Byte strides and the 32-bit addressing flag of a DevGrid.
__host__ __device__ inline void fill_grid_strides(const slt_grid_desc* g, DevGrid* d) {
    d->b0 = g->stride[0] * 16; d->b1 = g->stride[1] * 16; d->b2 = g->stride[2] * 16;
    const int64_t total = (int64_t)g->n[0] * g->n[1] * g->n[2];
    d->off32 = total * 16 <= (int64_t)0xFFFFFFFF ? 1 : 0;
}

// 1 / cell size of axis a, as fill_grid_units computes it (the pack's scale).
__host__ __device__ inline double grid_inv_spacing(const slt_grid_desc* g, int a) {
    return (double)(g->n[a] - 1) / ((double)g->hi[a] - (double)g->lo[a]);
}

// Kernels built with index arithmetic take uniform grids only; slt_trace also
// takes a non-uniform one (allow_nonuniform) whose knots are set.
inline int make_dev_grid(const void* packed, const slt_grid_desc* g, DevGrid* d,
                         bool allow_nonuniform = false) {
    if (!packed || !g) {
        set_error("null grid pointer");
        return SLT_ERR_INVALID;
    }
    if (!g->uniform && !allow_nonuniform) {
        set_error("this entry point takes uniformly spaced grids only (linspace axes)");
        return SLT_ERR_UNSUPPORTED;
    }
    if (!g->uniform && !g->knots) {
        set_error("non-uniform grid without knots (set slt_grid_desc.knots from slt_grid_knots)");
        return SLT_ERR_INVALID;
    }
    const int64_t total = (int64_t)g->n[0] * g->n[1] * g->n[2];
    if (total >= (int64_t(1) << 31)) {
        set_error("grid of %lld voxels exceeds the 2^31 limit", (long long)total);
        return SLT_ERR_UNSUPPORTED;
    }
    for (int a = 0; a < 3; ++a) {
        if (g->stride[a] >= (int64_t(1) << 24) || g->n[a] >= (1 << 24)) {
            set_error("grid axis %d: stride %lld / extent %d must stay below 2^24", a,
                      (long long)g->stride[a], g->n[a]);
            return SLT_ERR_UNSUPPORTED;
        }
        if (g->n[a] < 2 || !(g->hi[a] > g->lo[a])) {
            set_error("grid axis %d: need >= 2 points and hi > lo", a);
            return SLT_ERR_INVALID;
        }
    }
    d->vox = static_cast<const float4*>(packed);
    d->n0 = g->n[0]; d->n1 = g->n[1]; d->n2 = g->n[2];
    d->s0 = (int)g->stride[0]; d->s1 = (int)g->stride[1]; d->s2 = (int)g->stride[2];
    fill_grid_strides(g, d);
    fill_grid_units(g, d);
    d->nchan = g->nchan;
    d->kn0 = g->uniform ? nullptr : g->knots;
    d->kn1 = g->uniform ? nullptr : g->knots + g->n[0];
    d->kn2 = g->uniform ? nullptr : g->knots + g->n[0] + g->n[1];
    return SLT_OK;
}

// Device-side fill of a DevGrid from a descriptor the host already checked
// (slt_job_table_prepare).
__host__ __device__ inline void make_dev_grid_device(const void* packed, const slt_grid_desc* g,
                                                     DevGrid* d) {
    d->vox = static_cast<const float4*>(packed);
    d->n0 = g->n[0]; d->n1 = g->n[1]; d->n2 = g->n[2];
    d->s0 = (int)g->stride[0]; d->s1 = (int)g->stride[1]; d->s2 = (int)g->stride[2];
    fill_grid_strides(g, d);
    fill_grid_units(g, d);
    d->nchan = g->nchan;
    d->kn0 = d->kn1 = d->kn2 = nullptr;  // the host check admits uniform grids only
}

// Device memory seen as the global address space.  A pointer that a kernel
// reads out of memory (a job table) is generic to the compiler, which then
// emits flat loads; through these the accesses become global loads, as for
// pointers passed as kernel arguments.
#define SLT_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const SLT_GLOBAL T* to_global(const T* p) {
    return (const SLT_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ SLT_GLOBAL T* to_global(T* p) {
    return (SLT_GLOBAL T*)p;
}

// Number of XCDs on MI355X; workgroups are dispatched to XCDs round-robin by
// workgroup id, so block b runs on XCD (b % kXcds).
constexpr int kXcds = 8;

// Maps the hardware block id to a logical block id so that each XCD works on
// one contiguous range of logical blocks (rays are spatially sorted, so a
// contiguous range shares grid cache lines in that XCD's L2).  nb_padded must
// be a multiple of kXcds; returns -1 for padding blocks.
__device__ inline int xcd_block(int bid, int nb_padded, int nb) {
    const int per = nb_padded / kXcds;
    const int logical = (bid % kXcds) * per + bid / kXcds;
    return logical < nb ? logical : -1;
}

inline int xcd_pad(int nb) { return (nb + kXcds - 1) / kXcds * kXcds; }

}  // namespace slt

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
struct DevGrid {
    const float4* __restrict__ vox;
    int n0, n1, n2;
    int s0, s1, s2;
    float lo0, lo1, lo2;
    float hi0, hi1, hi2;
    float id0, id1, id2;
    float ul0, ul1, ul2;  // -lo * inv_d: the grid coordinate is u = fma(x, inv_d, ul)
    int nchan;
};

inline int make_dev_grid(const void* packed, const slt_grid_desc* g, DevGrid* d) {
    if (!packed || !g) {
        set_error("null grid pointer");
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
    }
    d->vox = static_cast<const float4*>(packed);
    d->n0 = g->n[0]; d->n1 = g->n[1]; d->n2 = g->n[2];
    d->s0 = (int)g->stride[0]; d->s1 = (int)g->stride[1]; d->s2 = (int)g->stride[2];
    d->lo0 = g->lo[0]; d->lo1 = g->lo[1]; d->lo2 = g->lo[2];
    d->hi0 = g->hi[0]; d->hi1 = g->hi[1]; d->hi2 = g->hi[2];
    d->id0 = g->inv_d[0]; d->id1 = g->inv_d[1]; d->id2 = g->inv_d[2];
    d->ul0 = -(d->lo0 * d->id0); d->ul1 = -(d->lo1 * d->id1); d->ul2 = -(d->lo2 * d->id2);
    d->nchan = g->nchan;
    return SLT_OK;
}

// Device-side fill of a DevGrid from a descriptor the host already checked
// (slt_job_table_prepare).
__host__ __device__ inline void make_dev_grid_device(const void* packed, const slt_grid_desc* g,
                                                     DevGrid* d) {
    d->vox = static_cast<const float4*>(packed);
    d->n0 = g->n[0]; d->n1 = g->n[1]; d->n2 = g->n[2];
    d->s0 = (int)g->stride[0]; d->s1 = (int)g->stride[1]; d->s2 = (int)g->stride[2];
    d->lo0 = g->lo[0]; d->lo1 = g->lo[1]; d->lo2 = g->lo[2];
    d->hi0 = g->hi[0]; d->hi1 = g->hi[1]; d->hi2 = g->hi[2];
    d->id0 = g->inv_d[0]; d->id1 = g->inv_d[1]; d->id2 = g->inv_d[2];
    d->ul0 = -(d->lo0 * d->id0); d->ul1 = -(d->lo1 * d->id1); d->ul2 = -(d->lo2 * d->id2);
    d->nchan = g->nchan;
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

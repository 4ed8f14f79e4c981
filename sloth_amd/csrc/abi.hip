// Error handling, host-side grid descriptors and ray ordering for libsloth_hip.so.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "sloth_internal.h"

namespace slt {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = 0; }

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return SLT_ERR_HIP;
    }
    return SLT_OK;
}

// Spreads the low 15 bits of v to the even bit positions of a 30-bit word.
__device__ inline uint32_t spread15(uint32_t v) {
    v &= 0x7FFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__device__ inline uint32_t quantise15(float p, float lo, float hi) {
    float u = (p - lo) / (hi - lo);
    u = fminf(fmaxf(u, 0.0f), 1.0f);  // NaN -> 0
    return (uint32_t)(u * 32767.0f);
}

__global__ void __launch_bounds__(256) morton_kernel(const float* __restrict__ s0, int64_t n,
                                                     int ax_a, int ax_b, float lo_a, float hi_a,
                                                     float lo_b, float hi_b,
                                                     int32_t* __restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t qa = quantise15(s0[ax_a * n + i], lo_a, hi_a);
    const uint32_t qb = quantise15(s0[ax_b * n + i], lo_b, hi_b);
    keys[i] = (int32_t)(spread15(qa) | (spread15(qb) << 1));
}

// Per-ray 32-byte records for permuting rays.  A scattered 4-byte access moves
// a whole 128-byte line, so permuting six SoA rows element by element moves six
// lines per ray; a record moves one.  Each permutation is therefore a coalesced
// transpose between the SoA rows and records in one order plus one record
// access per ray at its position in the other order.
struct __attribute__((aligned(32))) RayRec {
    float v[8];
};

// rec[i] = the six start-state rows of ray i (caller order, coalesced).
__global__ void __launch_bounds__(256) rays_to_rec_kernel(const float* __restrict__ s0, int64_t n,
                                                          RayRec* __restrict__ rec) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    RayRec r;
#pragma unroll
    for (int c = 0; c < 6; ++c) r.v[c] = s0[c * n + i];
    r.v[6] = 0.f;
    r.v[7] = 0.f;
    rec[i] = r;
}

// Row t of the output takes the caller's ray order[t]; an entry outside [0, n)
// gives a NaN row.
__global__ void __launch_bounds__(256) gather_rec_kernel(const RayRec* __restrict__ rec,
                                                         int64_t n,
                                                         const int32_t* __restrict__ order,
                                                         float* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t i = order[t];
    RayRec r;
    if (i >= 0 && i < n) {
        r = rec[i];
    } else {
#pragma unroll
        for (int c = 0; c < 6; ++c) r.v[c] = __builtin_nanf("");
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) out[c * n + t] = r.v[c];
}

// Column t of `rows` fp32 rows (and of the fp64 row src64, in the two slots after
// them) goes to record order[t]; entries outside [0, n) are skipped.
__global__ void __launch_bounds__(256) scatter_rec_kernel(const float* __restrict__ src, int rows,
                                                          const double* __restrict__ src64,
                                                          int64_t n,
                                                          const int32_t* __restrict__ order,
                                                          RayRec* __restrict__ rec) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t o = order[t];
    if (o < 0 || o >= n) return;
    RayRec r;
#pragma unroll
    for (int c = 0; c < 8; ++c) r.v[c] = c < rows ? src[c * n + t] : 0.f;
    if (src64) {
        const float2 b = __builtin_bit_cast(float2, src64[t]);
#pragma unroll
        for (int c = 0; c < 7; ++c)
            if (c == rows) {
                r.v[c] = b.x;
                r.v[c + 1] = b.y;
            }
    }
    rec[o] = r;
}

// The records back to rows (caller order, coalesced).
__global__ void __launch_bounds__(256) rec_to_rows_kernel(const RayRec* __restrict__ rec,
                                                          int rows, int64_t n,
                                                          float* __restrict__ dst,
                                                          double* __restrict__ dst64) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const RayRec r = rec[i];
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (c < rows) dst[c * n + i] = r.v[c];
    if (dst64) {
        float2 b = make_float2(0.f, 0.f);
#pragma unroll
        for (int c = 0; c < 7; ++c)
            if (c == rows) b = make_float2(r.v[c], r.v[c + 1]);
        dst64[i] = __builtin_bit_cast(double, b);
    }
}

}  // namespace slt

using namespace slt;

extern "C" {

int slt_version(void) { return SLT_ABI_VERSION; }

const char* slt_last_error(void) { return g_err; }

int slt_grid_desc_init(const float* x, const float* y, const float* z, int32_t nx, int32_t ny,
                       int32_t nz, int32_t probing_axis, int32_t with_kappa,
                       slt_grid_desc* out) {
    clear_error();
    if (!x || !y || !z || !out) {
        set_error("slt_grid_desc_init: null pointer");
        return SLT_ERR_INVALID;
    }
    if (probing_axis < 0 || probing_axis > 2) {
        set_error("slt_grid_desc_init: probing_axis must be 0, 1 or 2 (got %d)", probing_axis);
        return SLT_ERR_INVALID;
    }
    const float* c[3] = {x, y, z};
    const int32_t n[3] = {nx, ny, nz};
    int uniform = 1;
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 2) {
            set_error("slt_grid_desc_init: axis %d has %d points (need >= 2)", a, n[a]);
            return SLT_ERR_INVALID;
        }
        const double lo = c[a][0], hi = c[a][n[a] - 1];
        if (!(hi > lo) || !isfinite(lo) || !isfinite(hi)) {
            set_error("slt_grid_desc_init: axis %d coordinates must be finite and ascending", a);
            return SLT_ERR_INVALID;
        }
        for (int i = 1; i < n[a]; ++i) {
            if (!((double)c[a][i] > (double)c[a][i - 1])) {
                set_error("slt_grid_desc_init: axis %d must be strictly ascending "
                          "(point %d is not above point %d)", a, i, i - 1);
                return SLT_ERR_INVALID;
            }
        }
        // uniform: every point within max(1e-4 cell, 4 ulp) of the linspace
        const double d = (hi - lo) / (n[a] - 1);
        const double amax = fabs(lo) > fabs(hi) ? fabs(lo) : fabs(hi);
        const double ulp = amax * 1.1920928955078125e-07;
        const double tol = fmax(1e-4 * d, 4.0 * ulp);
        for (int i = 0; i < n[a] && uniform; ++i)
            if (!(fabs((double)c[a][i] - (lo + i * d)) <= tol)) uniform = 0;
        out->n[a] = n[a];
        out->lo[a] = c[a][0];
        out->hi[a] = c[a][n[a] - 1];
        out->inv_d[a] = (float)((n[a] - 1) / ((double)out->hi[a] - (double)out->lo[a]));
    }
    // Probing axis fastest, the other two in x, y, z order (slowest first).
    int others[2], k = 0;
    for (int a = 0; a < 3; ++a)
        if (a != probing_axis) others[k++] = a;
    out->stride[probing_axis] = 1;
    out->stride[others[1]] = n[probing_axis];
    out->stride[others[0]] = (int64_t)n[probing_axis] * n[others[1]];
    out->nchan = with_kappa ? 4 : 3;
    out->probing_axis = probing_axis;
    out->uniform = uniform;
    out->reserved = 0;
    out->knots = nullptr;
    return SLT_OK;
}

int slt_grid_knots(const float* x, const float* y, const float* z, const slt_grid_desc* g,
                   float* knots) {
    clear_error();
    if (!x || !y || !z || !g || !knots) {
        set_error("slt_grid_knots: null pointer");
        return SLT_ERR_INVALID;
    }
    const float* c[3] = {x, y, z};
    int64_t off = 0;
    for (int a = 0; a < 3; ++a) {
        if (g->n[a] < 2 || !(g->hi[a] > g->lo[a])) {
            set_error("slt_grid_knots: axis %d needs >= 2 points and hi > lo", a);
            return SLT_ERR_INVALID;
        }
        const double lo = g->lo[a], isp = grid_inv_spacing(g, a);
        for (int i = 0; i < g->n[a]; ++i) knots[off + i] = (float)(((double)c[a][i] - lo) * isp);
        off += g->n[a];
    }
    return SLT_OK;
}

int64_t slt_grid_bytes(const slt_grid_desc* g) {
    if (!g) return 0;
    return (int64_t)g->n[0] * g->n[1] * g->n[2] * 16;
}

int slt_morton_keys(const float* s0, int64_t n, const slt_grid_desc* g, int32_t* keys,
                    slt_stream_t stream) {
    clear_error();
    if (n < 0 || !g || g->probing_axis < 0 || g->probing_axis > 2 ||
        (n > 0 && (!s0 || !keys))) {
        set_error("slt_morton_keys: invalid arguments");
        return SLT_ERR_INVALID;
    }
    if (n == 0) return SLT_OK;
    int ax[2], k = 0;
    for (int a = 0; a < 3; ++a)
        if (a != g->probing_axis) ax[k++] = a;
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(morton_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, s0, n, ax[0], ax[1], g->lo[ax[0]], g->hi[ax[0]],
                       g->lo[ax[1]], g->hi[ax[1]], keys);
    return check_launch("slt_morton_keys");
}

int64_t slt_ray_work_bytes(int64_t n) { return n > 0 ? n * (int64_t)sizeof(RayRec) : 0; }

// [a, a + na) and [b, b + nb) share a byte
static bool overlaps(const void* a, int64_t na, const void* b, int64_t nb) {
    const char* pa = static_cast<const char*>(a);
    const char* pb = static_cast<const char*>(b);
    return pa < pb + nb && pb < pa + na;
}

int slt_gather_rays(const float* s0, int64_t n, const int32_t* order, float* s0_sorted,
                    void* work, slt_stream_t stream) {
    clear_error();
    if (n < 0 || (n > 0 && (!s0 || !order || !s0_sorted || !work))) {
        set_error("slt_gather_rays: invalid arguments");
        return SLT_ERR_INVALID;
    }
    if (n == 0) return SLT_OK;
    const int64_t wb = slt_ray_work_bytes(n);
    if (overlaps(work, wb, s0, 24 * n) || overlaps(work, wb, s0_sorted, 24 * n)) {
        set_error("slt_gather_rays: work overlaps the rays");
        return SLT_ERR_INVALID;
    }
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    hipStream_t st = (hipStream_t)stream;
    RayRec* rec = static_cast<RayRec*>(work);
    hipLaunchKernelGGL(rays_to_rec_kernel, grid, block, 0, st, s0, n, rec);
    hipLaunchKernelGGL(gather_rec_kernel, grid, block, 0, st, rec, n, order, s0_sorted);
    return check_launch("slt_gather_rays");
}

int slt_unpermute_rays(const float* src, int32_t rows, const double* src64, int64_t n,
                       const int32_t* order, float* dst, double* dst64, void* work,
                       slt_stream_t stream) {
    clear_error();
    if (n < 0 || rows < 0 || rows + (src64 ? 2 : 0) > 8 || (rows == 0 && !src64) ||
        (!src64) != (!dst64) ||
        (n > 0 && ((rows > 0 && (!src || !dst)) || !order || !work))) {
        set_error("slt_unpermute_rays: invalid arguments");
        return SLT_ERR_INVALID;
    }
    if (n == 0) return SLT_OK;
    const int64_t wb = slt_ray_work_bytes(n);
    if ((rows > 0 && (overlaps(work, wb, src, 4 * rows * n) ||
                      overlaps(work, wb, dst, 4 * rows * n))) ||
        (src64 && (overlaps(work, wb, src64, 8 * n) || overlaps(work, wb, dst64, 8 * n)))) {
        set_error("slt_unpermute_rays: work overlaps the data");
        return SLT_ERR_INVALID;
    }
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    hipStream_t st = (hipStream_t)stream;
    RayRec* rec = static_cast<RayRec*>(work);
    hipLaunchKernelGGL(scatter_rec_kernel, grid, block, 0, st, src, (int)rows, src64, n, order,
                       rec);
    hipLaunchKernelGGL(rec_to_rows_kernel, grid, block, 0, st, rec, (int)rows, n, dst, dst64);
    return check_launch("slt_unpermute_rays");
}

}  // extern "C"

// Jones-vector epilogue and k3, the detector histogram.
//
// The histogram reproduces numpy.histogram2d's binning exactly: each value is
// promoted to fp64, located with numpy's searchsorted(edges, x, side='right')
// semantics against the caller's fp64 edge array (an arithmetic guess
// corrected by at most a few comparisons), and a value equal to the last edge
// is moved into the last bin (numpy/lib/_histograms_impl.py histogramdd).
// Counts are uint32 atomics (order-independent, so bit-exact); weighted sums
// are fp64 atomics (equal to numpy's sequential sums up to summation order,
// which also depends on the visiting order and on how a wavefront combines
// its lanes).
#include <math.h>

#include "sloth_internal.h"

namespace slt {

__global__ void __launch_bounds__(256) jones_kernel(const float* __restrict__ st, int64_t ld,
                                                    int64_t n, int axis, float* __restrict__ rf,
                                                    const double* __restrict__ tau,
                                                    double* __restrict__ amp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = st[i], y = st[ld + i], z = st[2 * ld + i];
    const float vx = st[3 * ld + i], vy = st[4 * ld + i], vz = st[5 * ld + i];
    float p1, a1, p2, a2;
    if (axis == 0) {
        p1 = y; a1 = atanf(vy / vx); p2 = z; a2 = atanf(vz / vx);
    } else if (axis == 1) {
        p1 = x; a1 = atanf(vx / vy); p2 = z; a2 = atanf(vz / vy);
    } else {
        p1 = x; a1 = atanf(vx / vz); p2 = y; a2 = atanf(vy / vz);
    }
    rf[i] = p1;
    rf[n + i] = a1;
    rf[2 * n + i] = p2;
    rf[3 * n + i] = a2;
    if (amp && tau) amp[i] = exp(-tau[i]);
}

// numpy bin index in [0, nb+1]: 0 = below, nb+1 = above (dropped), else bin+1.
__device__ __forceinline__ int np_bin(double v, const double* __restrict__ e, int nb) {
    if (!(v >= e[0])) return 0;  // below range, or NaN
    if (v >= e[nb]) return v == e[nb] ? nb : nb + 1;
    // guess, then correct so that e[b] <= v < e[b+1]
    int b = (int)((v - e[0]) / (e[nb] - e[0]) * (double)nb);
    b = b < 0 ? 0 : (b > nb - 1 ? nb - 1 : b);
    while (b > 0 && v < e[b]) --b;
    while (b < nb - 1 && v >= e[b + 1]) ++b;
    return b + 1;
}

// Sum of v over the 64 lanes of the wavefront (every lane gets the total).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Cells per wavefront step whose entries are combined into one atomic each.
constexpr int kHistCombine = 4;

// One histogram entry per visited ray (ray order[i] at step i when ORDERED, else
// ray i).  Each wavefront combines the entries it makes for its first
// kHistCombine distinct cells into a single atomic per cell (count = popcount of
// the lanes, weight = their wave sum): with rays in Morton order a wavefront's 64
// rays fall in one or two cells, so the device-scope atomics on the same address
// drop ~30-fold.  The lanes of any further cells (rays in no useful order
// rarely share a cell) add their own entries.
template <typename T, bool ORDERED>
__global__ void __launch_bounds__(256) hist_kernel(const T* __restrict__ x, int64_t xs,
                                                   const T* __restrict__ y, int64_t ys,
                                                   int64_t n, const int32_t* __restrict__ order,
                                                   const uint8_t* __restrict__ mask,
                                                   const double* __restrict__ xe, int nxb,
                                                   const double* __restrict__ ye, int nyb,
                                                   const double* __restrict__ w,
                                                   uint32_t* __restrict__ counts,
                                                   double* __restrict__ wsum) {
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // wave-uniform trip count: every lane of a wavefront runs the same iterations
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < n;
         base += stride) {
        const int64_t i = base + lane;
        int64_t j = i;
        bool ok = i < n;
        if (ORDERED && ok) {
            j = order[i];
            ok = j >= 0 && j < n;  // a non-permutation misbins but never reads out of bounds
        }
        if (ok && mask) ok = mask[j] != 0;
        int64_t cell = -1;
        if (ok) {
            const int bx = np_bin((double)x[j * xs], xe, nxb);
            const int by = np_bin((double)y[j * ys], ye, nyb);
            if (bx >= 1 && bx <= nxb && by >= 1 && by <= nyb)
                cell = (int64_t)(by - 1) * nxb + (bx - 1);  // H.T layout
        }
        const double wj = (wsum && cell >= 0) ? w[j] : 0.0;
        const int c32 = (int)cell;  // cells < nxb * nyb < 2^31
        uint64_t pending = __ballot(cell >= 0);
        for (int k = 0; pending && k < kHistCombine; ++k) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const int c = __builtin_amdgcn_readlane(c32, leader);
            const uint64_t same = __ballot(c32 == c);
            if (counts && lane == leader) atomicAdd(counts + c, (uint32_t)__popcll(same));
            if (wsum) {
                const double s = wave_sum(c32 == c ? wj : 0.0);
                if (lane == leader) atomicAdd(wsum + c, s);
            }
            pending &= ~same;
        }
        if ((pending >> lane) & 1ull) {
            if (counts) atomicAdd(counts + c32, 1u);
            if (wsum) atomicAdd(wsum + c32, wj);
        }
    }
}

}  // namespace slt

using namespace slt;

extern "C" int slt_jones(const float* state, int64_t ld, int64_t n, int32_t probing_axis,
                         float* rf, const double* tau, double* amp, slt_stream_t stream) {
    clear_error();
    if (n < 0 || probing_axis < 0 || probing_axis > 2 || (n > 0 && (!state || !rf)) || ld < n) {
        set_error("slt_jones: invalid arguments");
        return SLT_ERR_INVALID;
    }
    if (n == 0) return SLT_OK;
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(jones_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       state, ld, n, probing_axis, rf, tau, amp);
    return check_launch("slt_jones");
}

extern "C" int slt_histogram2d(const void* x, int64_t x_stride, const void* y,
                               int64_t y_stride, int32_t is_f64, int64_t n, const int32_t* order,
                               const uint8_t* mask, const double* xedges, int32_t nxb,
                               const double* yedges, int32_t nyb, const double* weights,
                               uint32_t* counts, double* wsum, slt_stream_t stream) {
    clear_error();
    if (n < 0 || nxb < 1 || nyb < 1 || (int64_t)nxb * nyb > INT32_MAX || !xedges || !yedges ||
        (!counts && !wsum) || (n > 0 && (!x || !y || (wsum && !weights))) || x_stride < 0 ||
        y_stride < 0) {
        set_error("slt_histogram2d: invalid arguments");
        return SLT_ERR_INVALID;
    }
    if (n == 0) return SLT_OK;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    const dim3 grid((unsigned)blocks), block(256);
    hipStream_t st = (hipStream_t)stream;
    if (is_f64) {
        const double* xd = static_cast<const double*>(x);
        const double* yd = static_cast<const double*>(y);
        if (order)
            hipLaunchKernelGGL((hist_kernel<double, true>), grid, block, 0, st, xd, x_stride, yd,
                               y_stride, n, order, mask, xedges, nxb, yedges, nyb, weights,
                               counts, wsum);
        else
            hipLaunchKernelGGL((hist_kernel<double, false>), grid, block, 0, st, xd, x_stride, yd,
                               y_stride, n, order, mask, xedges, nxb, yedges, nyb, weights,
                               counts, wsum);
    } else {
        const float* xf = static_cast<const float*>(x);
        const float* yf = static_cast<const float*>(y);
        if (order)
            hipLaunchKernelGGL((hist_kernel<float, true>), grid, block, 0, st, xf, x_stride, yf,
                               y_stride, n, order, mask, xedges, nxb, yedges, nyb, weights,
                               counts, wsum);
        else
            hipLaunchKernelGGL((hist_kernel<float, false>), grid, block, 0, st, xf, x_stride, yf,
                               y_stride, n, order, mask, xedges, nxb, yedges, nyb, weights,
                               counts, wsum);
    }
    return check_launch("slt_histogram2d");
}

// k5: ray-transfer-matrix optics between the plasma and the detector.
//
// A diagnostic (shadowgraphy, schlieren, refractometry) carries every ray
// (x, theta, y, phi) through a chain of thin lenses, free travel and
// apertures before the detector (src/processing/diagnostics.py:28-128 for the
// elements, :284-406 for the chains).  The reference applies each element to
// the whole (4, N) array with a 4x4 matrix product or a boolean mask; here one
// lane walks the whole chain for its ray in fp64 registers, so the chain costs
// one read and one write of the rays whatever its length.
//
// Semantics kept from the reference:
//  * positions are scaled from m to mm first (diagnostics.py:16-20) in the
//    input's precision (a float32 input is scaled in float32, as numpy does);
//  * a linear element is the product of its 4x4 matrix with the ray, each
//    output component accumulated over the four inputs in order with fused
//    multiply-adds, so a rejected ray (all NaN) stays NaN through it, as in
//    the matrix product;
//  * a rejected ray becomes NaN in all four components (r[:, filt] = None);
//  * comparisons use x*x and y*y rounded separately, as numpy's x**2 does.
#include <math.h>

#include "sloth_internal.h"

namespace slt {

struct OpticsK {
    const void* in;
    int in_f64;
    int64_t in_ld;
    int64_t n;
    double pos_scale;
    double* out;
    int64_t out_ld;
    int n_ops;
    // Per element: kind, axis and the parameters prepared on the host
    // (the matrix entries for linear elements, squared radii for masks).
    int kind[SLT_MAX_OPTICS];
    int axis[SLT_MAX_OPTICS];
    double p0[SLT_MAX_OPTICS];
    double p1[SLT_MAX_OPTICS];
};

// 4x4 product of a block-diagonal [[a, b], [c, d]] (x block) and
// [[e, f], [g, h]] (y block) matrix with the ray, accumulated as a dense
// 4x4 product (the zero entries included, so NaN and inf propagate as in the
// reference's matmul).
__device__ __forceinline__ void matvec(double r[4], double a, double b, double c, double d,
                                       double e, double f, double g, double h) {
    const double m[4][4] = {{a, b, 0.0, 0.0}, {c, d, 0.0, 0.0}, {0.0, 0.0, e, f},
                            {0.0, 0.0, g, h}};
    double o[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double acc = m[i][0] * r[0];
        acc = fma(m[i][1], r[1], acc);
        acc = fma(m[i][2], r[2], acc);
        acc = fma(m[i][3], r[3], acc);
        o[i] = acc;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = o[i];
}

__global__ void __launch_bounds__(256) optics_kernel(OpticsK K) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= K.n) return;
    double r[4];
    if (K.in_f64) {
        const double* p = static_cast<const double*>(K.in);
#pragma unroll
        for (int c = 0; c < 4; ++c) r[c] = p[c * K.in_ld + t];
        r[0] = r[0] * K.pos_scale;
        r[2] = r[2] * K.pos_scale;
    } else {
        const float* p = static_cast<const float*>(K.in);
        const float s = (float)K.pos_scale;
        const float x = p[t] * s;
        const float y = p[2 * K.in_ld + t] * s;
        r[0] = (double)x;
        r[1] = (double)p[K.in_ld + t];
        r[2] = (double)y;
        r[3] = (double)p[3 * K.in_ld + t];
    }
    for (int k = 0; k < K.n_ops; ++k) {
        const double a = K.p0[k], b = K.p1[k];
        bool reject = false;
        switch (K.kind[k]) {
            case SLT_OPTIC_TRAVEL:  // [[1, d], [0, 1]] on both axes
                matvec(r, 1.0, a, 0.0, 1.0, 1.0, a, 0.0, 1.0);
                break;
            case SLT_OPTIC_LENS:  // [[1, 0], [-1/f, 1]]; a = -1/f1, b = -1/f2
                matvec(r, 1.0, 0.0, a, 1.0, 1.0, 0.0, b, 1.0);
                break;
            case SLT_OPTIC_CIRCULAR_APERTURE: {  // a = R^2
                const double x2 = r[0] * r[0];
                const double y2 = r[2] * r[2];
                reject = x2 + y2 > a;
                break;
            }
            case SLT_OPTIC_CIRCULAR_STOP: {  // a = R^2
                const double x2 = r[0] * r[0];
                const double y2 = r[2] * r[2];
                reject = x2 + y2 < a;
                break;
            }
            case SLT_OPTIC_RECT_APERTURE: {  // a = Lx^2, b = Ly^2
                const double x2 = r[0] * r[0];
                const double y2 = r[2] * r[2];
                reject = (x2 > a) && (y2 > b);
                break;
            }
            case SLT_OPTIC_KNIFE_EDGE: {  // a = offset, b = direction sign
                const double v = K.axis[k] == 0 ? r[0] : r[2];
                reject = b > 0.0 ? v > a : v < a;
                break;
            }
            default:
                break;
        }
        if (reject) r[0] = r[1] = r[2] = r[3] = __builtin_nan("");
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) K.out[c * K.out_ld + t] = r[c];
}

}  // namespace slt

using namespace slt;

extern "C" int slt_rtm(const void* rays, int32_t rays_f64, int64_t ld, int64_t n,
                       double pos_scale, const slt_optic* ops, int32_t n_ops, double* out,
                       int64_t out_ld, slt_stream_t stream) {
    clear_error();
    if (n < 0 || n_ops < 0 || n_ops > SLT_MAX_OPTICS || (n_ops > 0 && !ops)) {
        set_error("slt_rtm: need n >= 0 and 0 <= n_ops <= %d", SLT_MAX_OPTICS);
        return SLT_ERR_INVALID;
    }
    OpticsK K;
    K.in = rays;
    K.in_f64 = rays_f64 ? 1 : 0;
    K.in_ld = ld;
    K.n = n;
    K.pos_scale = pos_scale;
    K.out = out;
    K.out_ld = out_ld;
    K.n_ops = n_ops;
    for (int k = 0; k < n_ops; ++k) {
        const slt_optic& o = ops[k];
        K.kind[k] = o.kind;
        K.axis[k] = 0;
        switch (o.kind) {
            case SLT_OPTIC_TRAVEL:
                K.p0[k] = o.a;
                K.p1[k] = 0.0;
                break;
            case SLT_OPTIC_LENS:
                // the matrix entry -1 / f as the reference builds it
                K.p0[k] = -1.0 / o.a;
                K.p1[k] = -1.0 / o.b;
                break;
            case SLT_OPTIC_CIRCULAR_APERTURE:
            case SLT_OPTIC_CIRCULAR_STOP:
                K.p0[k] = o.a * o.a;
                K.p1[k] = 0.0;
                break;
            case SLT_OPTIC_RECT_APERTURE:
                K.p0[k] = o.a * o.a;
                K.p1[k] = o.b * o.b;
                break;
            case SLT_OPTIC_KNIFE_EDGE:
                if ((o.axis != 0 && o.axis != 2) || !(o.b > 0.0 || o.b < 0.0)) {
                    set_error("slt_rtm: knife edge %d needs axis 0 (x) or 2 (y) and a "
                              "direction < 0 or > 0", k);
                    return SLT_ERR_INVALID;
                }
                K.axis[k] = o.axis;
                K.p0[k] = o.a;
                K.p1[k] = o.b;
                break;
            default:
                set_error("slt_rtm: element %d has unknown kind %d", k, o.kind);
                return SLT_ERR_INVALID;
        }
    }
    if (n == 0) return SLT_OK;
    if (!rays || !out || ld < n || out_ld < n) {
        set_error("slt_rtm: rays and out must be 4 x n with row strides >= n");
        return SLT_ERR_INVALID;
    }
    const int64_t blocks = (n + 255) / 256;
    if (blocks > INT32_MAX) {
        set_error("slt_rtm: too many rays");
        return SLT_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(optics_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       K);
    return check_launch("slt_rtm");
}

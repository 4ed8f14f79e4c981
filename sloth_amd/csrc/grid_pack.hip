// k1: refraction-gradient stencil, inverse-bremsstrahlung rate and 16-byte
// voxel packing.  fp64 arithmetic, one rounding to fp32 at the store.  Reads 7
// neighbouring fp32 densities per point (6 of them L1/L2 hits) and writes one
// 16-byte voxel: 20 B of HBM traffic per point.
//
// pack_march_kernel (axis 0 of at most kMarchTableMax points): one thread per
// (j, k) column point marches through a run of planes along axis 0, keeping
// the axis-0 neighbours in registers, so a point reads one new density from
// HBM; the numpy.gradient stencil coefficients are built once (axis 0 in LDS,
// axes 1 and 2 per thread) instead of per point, and with scalar Te and Z the
// density-independent factors of kappa are computed once per thread.  Every
// value is the same fp64 expression as in the per-point pack_kernel below
// (used for longer axes 0), so both kernels write identical bits.

#include <math.h>

#include "sloth_internal.h"

namespace slt {

struct PackArgs {
    const float* __restrict__ ne;
    const float* __restrict__ c0;
    const float* __restrict__ c1;
    const float* __restrict__ c2;
    double coeff;
    const float* __restrict__ te;
    int64_t n_te;
    const float* __restrict__ zz;
    int64_t n_z;
    double omega, e_charge, epsilon_0, c_light, coulomb_log;
    int n0, n1, n2;
    int64_t s0, s1, s2;  // packed strides
    double isp[3];       // per-axis scale of the packed gradient channels (1 / cell size)
    float4* __restrict__ packed;
    float* __restrict__ pdx;
    float* __restrict__ pdy;
    float* __restrict__ pdz;
    float* __restrict__ pk;
};

// numpy/jnp.gradient along one axis with a coordinate array (edge_order=1):
// src/core/propagator.py:90-92.
__device__ inline double grad_axis(const float* __restrict__ ne, const float* __restrict__ c,
                                   int64_t lin, int64_t stride, int i, int n, double coeff) {
    const double f0 = (double)ne[lin] * coeff;
    if (i == 0) {
        const double f1 = (double)ne[lin + stride] * coeff;
        return (f1 - f0) / ((double)c[1] - (double)c[0]);
    }
    if (i == n - 1) {
        const double fm = (double)ne[lin - stride] * coeff;
        return (f0 - fm) / ((double)c[n - 1] - (double)c[n - 2]);
    }
    const double fm = (double)ne[lin - stride] * coeff;
    const double fp = (double)ne[lin + stride] * coeff;
    const double dx1 = (double)c[i] - (double)c[i - 1];
    const double dx2 = (double)c[i + 1] - (double)c[i];
    const double a = -dx2 / (dx1 * (dx1 + dx2));
    const double b = (dx2 - dx1) / (dx1 * dx2);
    const double cc = dx1 / (dx2 * (dx1 + dx2));
    return a * fm + b * f0 + cc * fp;
}

// NRL inverse-bremsstrahlung amplitude absorption rate [1/s]:
// src/core/propagator.py:147-164.
__device__ inline double kappa_ib(double ne, double te, double z, const PackArgs& A) {
    const double pi = 3.14159265358979311600;
    const double ne_cc = ne * 1e-6;
    const double v_the = 4.19e5 * sqrt(te);
    const double o_pe = 5.64e4 * sqrt(ne_cc);
    const double o_max = o_pe > A.omega ? o_pe : A.omega;
    const double b_min = z * A.e_charge / (4.0 * pi * A.epsilon_0 * te);
    double cl = log(v_the / (o_max * b_min));
    if (!(cl >= 2.0)) cl = 2.0;
    if (A.coulomb_log > 0.0) cl = A.coulomb_log;
    const double r = ne_cc / A.omega;
    return 3.1e-5 * z * A.c_light * (r * r) * cl * pow(te, -1.5);
}

// Stencil coefficients of grad_axis for point i of an axis of n points:
// interior points (a, b, c) of (f[i-1], f[i], f[i+1]); the two end points keep
// their one-sided denominators in b (a = c = 0).
__device__ inline void axis_coefs(const float* __restrict__ c, int n, int i, double& a,
                                  double& b, double& cc) {
    a = 0.0;
    cc = 0.0;
    if (i == 0) {
        b = (double)c[1] - (double)c[0];
    } else if (i == n - 1) {
        b = (double)c[n - 1] - (double)c[n - 2];
    } else {
        const double dx1 = (double)c[i] - (double)c[i - 1];
        const double dx2 = (double)c[i + 1] - (double)c[i];
        a = -dx2 / (dx1 * (dx1 + dx2));
        b = (dx2 - dx1) / (dx1 * dx2);
        cc = dx1 / (dx2 * (dx1 + dx2));
    }
}

// The density-independent factors of kappa_ib for scalar Te and Z, in the same
// operations: the rate is ((k1 (r r)) cl) te^-1.5 with cl = cl_under when the
// plasma frequency does not exceed omega (or an override is set).
struct KappaConst {
    double v_the, b_min, k1, p, cl_under;
};

__device__ inline KappaConst kappa_const(double te, double z, const PackArgs& A) {
    const double pi = 3.14159265358979311600;
    KappaConst K;
    K.v_the = 4.19e5 * sqrt(te);
    K.b_min = z * A.e_charge / (4.0 * pi * A.epsilon_0 * te);
    K.k1 = 3.1e-5 * z * A.c_light;
    K.p = pow(te, -1.5);
    double cl = log(K.v_the / (A.omega * K.b_min));
    if (!(cl >= 2.0)) cl = 2.0;
    if (A.coulomb_log > 0.0) cl = A.coulomb_log;
    K.cl_under = cl;
    return K;
}

__device__ inline double kappa_scalar(double ne, const KappaConst& K, const PackArgs& A) {
    const double ne_cc = ne * 1e-6;
    double cl = K.cl_under;
    if (!(A.coulomb_log > 0.0)) {
        const double o_pe = 5.64e4 * sqrt(ne_cc);
        if (o_pe > A.omega) {
            cl = log(K.v_the / (o_pe * K.b_min));
            if (!(cl >= 2.0)) cl = 2.0;
        }
    }
    const double r = ne_cc / A.omega;
    return K.k1 * (r * r) * cl * K.p;
}

// grad_axis on the three densities around a point (vm, v0, vp at i - 1, i, i + 1;
// the one-sided ends ignore the missing side), with axis_coefs' coefficients.
__device__ inline double grad_pts(float vm, float v0, float vp, int i, int n, double coeff,
                                  double a, double b, double cc) {
    const double f0 = (double)v0 * coeff;
    if (i == 0) {
        const double f1 = (double)vp * coeff;
        return (f1 - f0) / b;
    }
    if (i == n - 1) {
        const double fm = (double)vm * coeff;
        return (f0 - fm) / b;
    }
    const double fm = (double)vm * coeff;
    const double fp = (double)vp * coeff;
    return a * fm + b * f0 + cc * fp;
}

constexpr int kMarchPlanes = 32;   // planes of axis 0 one thread marches through
constexpr int kMarchTableMax = 2048;  // n0 for the LDS table of axis-0 coefficients (48 KiB)

// One thread per (j, k) column point and run of kMarchPlanes planes along axis 0
// (the slowest axis of ne): the axis-0 neighbours stay in a rolling window of
// registers, so each point reads one new density from HBM, and the next plane's
// five loads are issued before the current plane is computed.  The axis-1/2
// stencil coefficients are fixed per thread; the axis-0 ones come from LDS.
template <int KMODE>  // 0: no kappa, 1: scalar Te and Z, 2: Te/Z grids
__global__ void __launch_bounds__(256) pack_march_kernel(PackArgs A) {
    extern __shared__ double xt[];  // a | b | c of axis 0, n0 each
    const int n0 = A.n0, n1 = A.n1, n2 = A.n2;
    for (int e = threadIdx.x; e < n0; e += blockDim.x)
        axis_coefs(A.c0, n0, e, xt[e], xt[n0 + e], xt[2 * n0 + e]);
    __syncthreads();
    const uint32_t cols = (uint32_t)n1 * (uint32_t)n2;
    const uint32_t nrun = (uint32_t)((n0 + kMarchPlanes - 1) / kMarchPlanes);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cols * nrun) return;
    const uint32_t run = t / cols;
    const uint32_t col = t - run * cols;
    const int j = (int)(col / (uint32_t)n2);
    const int k = (int)(col - (uint32_t)j * (uint32_t)n2);
    const int i0 = (int)run * kMarchPlanes;
    const int i1 = min(n0, i0 + kMarchPlanes);
    double ya, yb, yc, za, zb, zc;
    axis_coefs(A.c1, n1, j, ya, yb, yc);
    axis_coefs(A.c2, n2, k, za, zb, zc);
    KappaConst K = {};
    if (KMODE == 1) K = kappa_const((double)A.te[0], (double)A.zz[0], A);
    const int64_t sx = (int64_t)cols;
    // in-plane neighbour offsets, clamped at the faces (the one-sided formulas
    // there do not use the missing side)
    const int64_t oym = j > 0 ? -(int64_t)n2 : 0, oyp = j < n1 - 1 ? (int64_t)n2 : 0;
    const int64_t ozm = k > 0 ? -1 : 0, ozp = k < n2 - 1 ? 1 : 0;
    const float* __restrict__ p = A.ne + (int64_t)i0 * sx + col;  // plane i0
    float xm = i0 > 0 ? p[-sx] : 0.f;
    float x0 = p[0];
    float xp = i0 + 1 < n0 ? p[sx] : 0.f;
    float ym = p[oym], yp = p[oyp], zm = p[ozm], zp = p[ozp];
    const int64_t obase = (int64_t)j * A.s1 + (int64_t)k * A.s2;
    for (int i = i0; i < i1; ++i) {
        // the next plane's loads first
        float nxp = 0.f, nym = 0.f, nyp = 0.f, nzm = 0.f, nzp = 0.f;
        const float* __restrict__ q = p + sx;
        if (i + 1 < i1) {
            nym = q[oym];
            nyp = q[oyp];
            nzm = q[ozm];
            nzp = q[ozp];
            if (i + 2 < n0) nxp = q[sx];
        }
        const double gx = grad_pts(xm, x0, xp, i, n0, A.coeff, xt[i], xt[n0 + i], xt[2 * n0 + i]);
        const double gy = grad_pts(ym, x0, yp, j, n1, A.coeff, ya, yb, yc);
        const double gz = grad_pts(zm, x0, zp, k, n2, A.coeff, za, zb, zc);
        const int64_t lin = (int64_t)i * sx + col;
        double kap = 0.0;
        if (KMODE == 1) {
            kap = kappa_scalar((double)x0, K, A);
        } else if (KMODE == 2) {
            const double te = (double)A.te[A.n_te == 1 ? 0 : lin];
            const double z = (double)A.zz[A.n_z == 1 ? 0 : lin];
            kap = kappa_ib((double)x0, te, z, A);
        }
        if (A.packed)
            A.packed[(int64_t)i * A.s0 + obase] =
                make_float4((float)(gx * A.isp[0]), (float)(gy * A.isp[1]),
                            (float)(gz * A.isp[2]), (float)kap);
        if (A.pdx) A.pdx[lin] = (float)gx;
        if (A.pdy) A.pdy[lin] = (float)gy;
        if (A.pdz) A.pdz[lin] = (float)gz;
        if (A.pk) A.pk[lin] = (float)kap;
        xm = x0;
        x0 = xp;
        xp = nxp;
        ym = nym;
        yp = nyp;
        zm = nzm;
        zp = nzp;
        p = q;
    }
}

__global__ void __launch_bounds__(256) pack_kernel(PackArgs A) {
    const int64_t total = (int64_t)A.n0 * A.n1 * A.n2;
    const int64_t sy = A.n2, sx = (int64_t)A.n1 * A.n2;
    for (int64_t lin = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; lin < total;
         lin += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(lin / sx);
        const int64_t rem = lin - (int64_t)i * sx;
        const int j = (int)(rem / sy);
        const int k = (int)(rem - (int64_t)j * sy);
        const double gx = grad_axis(A.ne, A.c0, lin, sx, i, A.n0, A.coeff);
        const double gy = grad_axis(A.ne, A.c1, lin, sy, j, A.n1, A.coeff);
        const double gz = grad_axis(A.ne, A.c2, lin, 1, k, A.n2, A.coeff);
        double kap = 0.0;
        if (A.te) {
            const double te = (double)A.te[A.n_te == 1 ? 0 : lin];
            const double z = (double)A.zz[A.n_z == 1 ? 0 : lin];
            kap = kappa_ib((double)A.ne[lin], te, z, A);
        }
        // packed: the gradient in grid units (per cell, as the integrators step);
        // planar: the physical gradient
        if (A.packed)
            A.packed[i * A.s0 + j * A.s1 + k * A.s2] =
                make_float4((float)(gx * A.isp[0]), (float)(gy * A.isp[1]),
                            (float)(gz * A.isp[2]), (float)kap);
        if (A.pdx) A.pdx[lin] = (float)gx;
        if (A.pdy) A.pdy[lin] = (float)gy;
        if (A.pdz) A.pdz[lin] = (float)gz;
        if (A.pk) A.pk[lin] = (float)kap;
    }
}

__global__ void __launch_bounds__(256) kappa_kernel(PackArgs A, int64_t n, float* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double te = (double)A.te[A.n_te == 1 ? 0 : i];
        const double z = (double)A.zz[A.n_z == 1 ? 0 : i];
        out[i] = (float)kappa_ib((double)A.ne[i], te, z, A);
    }
}

}  // namespace slt

using namespace slt;

extern "C" int slt_kappa(const float* ne, int64_t n, const float* te, int64_t n_te,
                         const float* zz, int64_t n_z, double omega, double e_charge,
                         double epsilon_0, double speed_of_light, double coulomb_log,
                         float* out, slt_stream_t stream) {
    clear_error();
    if (n < 0 || (n > 0 && (!ne || !te || !zz || !out)) || (n_te != 1 && n_te != n) ||
        (n_z != 1 && n_z != n)) {
        set_error("slt_kappa: invalid arguments (Te/Z must have 1 or n elements)");
        return SLT_ERR_INVALID;
    }
    if (n == 0) return SLT_OK;
    PackArgs A = {};
    A.ne = ne;
    A.te = te; A.n_te = n_te;
    A.zz = zz; A.n_z = n_z;
    A.omega = omega; A.e_charge = e_charge; A.epsilon_0 = epsilon_0; A.c_light = speed_of_light;
    A.coulomb_log = coulomb_log;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(kappa_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, A,
                       n, out);
    return check_launch("slt_kappa");
}

extern "C" int slt_grid_pack(const float* ne, const float* cx, const float* cy, const float* cz,
                             double grad_coeff, const float* te, int64_t n_te, const float* zz,
                             int64_t n_z, double omega, double e_charge, double epsilon_0,
                             double speed_of_light, double coulomb_log, const slt_grid_desc* g,
                             void* packed, float* planar_dx, float* planar_dy, float* planar_dz,
                             float* planar_kappa, slt_stream_t stream) {
    clear_error();
    if (!ne || !cx || !cy || !cz || !g) {
        set_error("slt_grid_pack: null input pointer");
        return SLT_ERR_INVALID;
    }
    const int64_t total = (int64_t)g->n[0] * g->n[1] * g->n[2];
    if (g->n[0] < 2 || g->n[1] < 2 || g->n[2] < 2) {
        set_error("slt_grid_pack: every axis needs >= 2 points");
        return SLT_ERR_INVALID;
    }
    const bool want_kappa = te != nullptr || zz != nullptr;
    if (want_kappa) {
        if (!te || !zz) {
            set_error("slt_grid_pack: Te and Z must both be given for kappa");
            return SLT_ERR_INVALID;
        }
        if ((n_te != 1 && n_te != total) || (n_z != 1 && n_z != total)) {
            set_error("slt_grid_pack: Te/Z must have 1 or nx*ny*nz elements (got %lld, %lld)",
                      (long long)n_te, (long long)n_z);
            return SLT_ERR_INVALID;
        }
    }
    if (!packed && !planar_dx && !planar_dy && !planar_dz && !planar_kappa) {
        set_error("slt_grid_pack: no output requested");
        return SLT_ERR_INVALID;
    }
    PackArgs A;
    A.ne = ne; A.c0 = cx; A.c1 = cy; A.c2 = cz;
    A.coeff = grad_coeff;
    A.te = want_kappa ? te : nullptr; A.n_te = n_te;
    A.zz = want_kappa ? zz : nullptr; A.n_z = n_z;
    A.omega = omega; A.e_charge = e_charge; A.epsilon_0 = epsilon_0; A.c_light = speed_of_light;
    A.coulomb_log = coulomb_log;
    A.n0 = g->n[0]; A.n1 = g->n[1]; A.n2 = g->n[2];
    A.s0 = g->stride[0]; A.s1 = g->stride[1]; A.s2 = g->stride[2];
    for (int a = 0; a < 3; ++a) {
        if (!(g->hi[a] > g->lo[a])) {
            set_error("slt_grid_pack: grid axis %d needs hi > lo", a);
            return SLT_ERR_INVALID;
        }
        A.isp[a] = grid_inv_spacing(g, a);
    }
    A.packed = static_cast<float4*>(packed);
    A.pdx = planar_dx; A.pdy = planar_dy; A.pdz = planar_dz; A.pk = planar_kappa;
    if (g->n[0] <= kMarchTableMax) {
        const int64_t threads = (int64_t)g->n[1] * g->n[2] *
                                ((g->n[0] + kMarchPlanes - 1) / kMarchPlanes);
        const dim3 grid((unsigned)((threads + 255) / 256));
        const size_t lds = (size_t)g->n[0] * 3 * sizeof(double);
        if (!A.te)
            hipLaunchKernelGGL(pack_march_kernel<0>, grid, dim3(256), lds, (hipStream_t)stream, A);
        else if (A.n_te == 1 && A.n_z == 1)
            hipLaunchKernelGGL(pack_march_kernel<1>, grid, dim3(256), lds, (hipStream_t)stream, A);
        else
            hipLaunchKernelGGL(pack_march_kernel<2>, grid, dim3(256), lds, (hipStream_t)stream, A);
    } else {
        int64_t blocks = (total + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(pack_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                           A);
    }
    return check_launch("slt_grid_pack");
}

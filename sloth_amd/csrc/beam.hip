// k6: probe-beam synthesis on the device with a counter-based generator.
//
// The reference draws its rays on the host from NumPy's RandomState
// (src/core/beam.py:57-346); that stream is sequential, so the drop-in Beam
// keeps it on the host.  This kernel is the opt-in alternative for callers
// that want the beam made in HBM (the serving loop): ray i takes its random
// numbers from Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC 2011) with counter (i, word, 0, 0) and the
// 64-bit seed as key, so every ray is independent of the launch shape and of
// the number of rays.  The distributions are the reference's: circular beams
// take t = 2 pi U, a radius with density 2r (numpy's power(2): sqrt(U)),
// phi = pi U and chi = divergence * N(0, 1) (Box-Muller); square and
// rectangular beams take two U(-1, 1) positions; linear beams one, in the x-z
// plane.  Positions and velocities follow the reference's layout per probing
// axis.
#include <math.h>

#include "sloth_internal.h"

namespace slt {

struct Philox {
    uint32_t c[4];
};

// Philox4x32 with 10 rounds: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
// increments 0x9E3779B9 / 0xBB67AE85.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
        const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// A 32-bit draw as a uniform double in (0, 1): (top 24 bits + 1/2) / 2^24.
__device__ __forceinline__ double unit(uint32_t x) {
    return ((double)(x >> 8) + 0.5) * (1.0 / 16777216.0);
}

struct BeamK {
    int64_t n;
    int type;   // 0 circular, 1 square, 2 rectangular, 3 linear
    int axis;   // probing axis
    double a, b, div, z0;
    uint32_t k0, k1;
    float* s0;
};

__global__ void __launch_bounds__(256) beam_kernel(BeamK K) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K.n) return;
    uint32_t w0[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0u, 0u};
    uint32_t w1[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), 1u, 0u};
    philox4x32_10(w0, K.k0, K.k1);
    philox4x32_10(w1, K.k0, K.k1);
    const double PI = 3.14159265358979311600;
    // chi = div * N(0,1) by Box-Muller from the second counter's first two words
    const double g = sqrt(-2.0 * log(unit(w1[0]))) * cos(2.0 * PI * unit(w1[1]));
    const double chi = K.div * g;
    double p, q, phi;
    if (K.type == 0) {
        const double t = 2.0 * PI * unit(w0[0]);
        const double r = K.a * sqrt(unit(w0[1]));
        phi = PI * unit(w0[2]);
        p = r * cos(t);
        q = r * sin(t);
    } else if (K.type == 1 || K.type == 2) {
        const double t = 2.0 * unit(w0[0]) - 1.0;
        const double u = 2.0 * unit(w0[1]) - 1.0;
        phi = PI * unit(w0[2]);
        p = K.a * u;
        q = K.b * t;
    } else {
        p = K.a * (2.0 * unit(w0[0]) - 1.0);
        q = 0.0;
        phi = 0.0;
    }
    const double c = 299792458.0;
    const double vl = c * cos(chi), vt1 = c * sin(chi) * cos(phi), vt2 = c * sin(chi) * sin(phi);
    double x, y, z, vx, vy, vz;
    if (K.type == 3) {  // always along z in the x-z plane
        x = p; y = 0.0; z = K.z0; vx = c * sin(chi); vy = 0.0; vz = c * cos(chi);
    } else if (K.axis == 0) {
        x = K.z0; y = p; z = q; vx = vl; vy = vt1; vz = vt2;
    } else if (K.axis == 2) {
        x = p; y = q; z = K.z0; vx = vt1; vy = vt2; vz = vl;
    } else {
        x = p; y = K.z0; z = q; vx = vt1; vy = vl; vz = vt2;
    }
    float* s = K.s0;
    s[i] = (float)x;
    s[K.n + i] = (float)y;
    s[2 * K.n + i] = (float)z;
    s[3 * K.n + i] = (float)vx;
    s[4 * K.n + i] = (float)vy;
    s[5 * K.n + i] = (float)vz;
}

}  // namespace slt

using namespace slt;

extern "C" int slt_beam(int64_t n, int32_t beam_type, int32_t probing_axis, double size_a,
                        double size_b, double divergence, double ne_extent, uint64_t seed,
                        float* s0, slt_stream_t stream) {
    clear_error();
    if (n < 0 || beam_type < SLT_BEAM_CIRCULAR || beam_type > SLT_BEAM_LINEAR ||
        probing_axis < 0 || probing_axis > 2 || (n > 0 && !s0)) {
        set_error("slt_beam: invalid arguments");
        return SLT_ERR_INVALID;
    }
    if (n == 0) return SLT_OK;
    BeamK K;
    K.n = n;
    K.type = beam_type;
    K.axis = probing_axis;
    K.a = size_a;
    K.b = beam_type == SLT_BEAM_SQUARE ? size_a : size_b;
    K.div = divergence;
    K.z0 = -ne_extent;
    K.k0 = (uint32_t)seed;
    K.k1 = (uint32_t)(seed >> 32);
    K.s0 = s0;
    const int64_t blocks = (n + 255) / 256;
    if (blocks > INT32_MAX) {
        set_error("slt_beam: too many rays");
        return SLT_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(beam_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, K);
    return check_launch("slt_beam");
}

// VALU issue calibration: kernels whose wave64 VALU instruction count is known
// exactly and which touch memory only to store one value per lane.
//
// Each lane runs SLT_CALIB_CHAINS independent chains of one instruction form for
// `iters` iterations, so every wavefront issues SLT_CALIB_CHAINS * iters
// instructions of that form (plus a fixed few dozen VALU instructions around
// the loop); with eight wavefronts per SIMD and sixteen independent chains per
// lane, issue -- not dependency latency -- limits it.  The forms differ in
// encoding and operand source, which is what sets their issue rate; the
// fastest form's rate (instructions per second over the launch, from HIP
// events) is the VALU-issue peak the trace kernel's rate is compared against,
// measured on the same GPU in the same run.
#include "sloth_internal.h"

namespace slt {

// One step of a chain, per instruction form (the encodings are those hipcc
// -O3 emits for gfx950 here):
//   FMA_SGPR  a = a * b + c, b and c uniform     v_fma_f32 (VOP3, SGPR operand)
//   FMA_VGPR  a = a * b + c, b and c per lane    v_fma_f32 (VOP3, VGPR operands)
//   FMAC      a = b * c + a, b and c per lane    v_fmac_f32 (VOP2)
//   ADD       a = a + b, b per lane              v_add_f32 (VOP2)
//   MUL       a = a * b, b per lane              v_mul_f32 (VOP2)
//   FMA64     d = d * e + f, e and f per lane    v_fma_f64
template <int KIND>
__device__ __forceinline__ void chain_step(float& a, double& d, float b, float c, double e,
                                           double f) {
    if (KIND == SLT_CALIB_FMA_SGPR || KIND == SLT_CALIB_FMA_VGPR) a = fmaf(a, b, c);
    if (KIND == SLT_CALIB_FMAC) a = fmaf(b, c, a);
    if (KIND == SLT_CALIB_ADD) a = a + b;
    if (KIND == SLT_CALIB_MUL) a = a * b;
    if (KIND == SLT_CALIB_FMA64) d = fma(d, e, f);
}

template <int KIND>
__global__ void __launch_bounds__(256) valu_calib_kernel(float bs, float cs, int iters,
                                                         float* __restrict__ out) {
    const float lane = (float)threadIdx.x;
    // per-lane operands (VGPRs) for every form but FMA_SGPR
    const float b = KIND == SLT_CALIB_FMA_SGPR ? bs : bs + lane * 1e-9f;
    const float c = KIND == SLT_CALIB_FMA_SGPR ? cs : cs + lane * 1e-9f;
    const double e = 0.999 + (double)lane * 1e-12, f = 1e-3 + (double)lane * 1e-12;
    float a[SLT_CALIB_CHAINS];
    double da[SLT_CALIB_CHAINS];
#pragma unroll
    for (int k = 0; k < SLT_CALIB_CHAINS; ++k) {
        a[k] = (float)(threadIdx.x + k) * 1e-3f;
        da[k] = (double)(threadIdx.x + k) * 1e-3;
    }
    // eight iterations per trip through the loop: 128 instructions of the form per
    // loop branch, so the scalar loop control takes almost no issue cycles
#pragma unroll 8
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < SLT_CALIB_CHAINS; ++k) chain_step<KIND>(a[k], da[k], b, c, e, f);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < SLT_CALIB_CHAINS; ++k) {
        if (KIND == SLT_CALIB_FMA64) s += (float)da[k];
        else s += a[k];
    }
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace slt

using namespace slt;

extern "C" int slt_valu_calibrate(int32_t kind, int32_t blocks, int32_t iters, float* out,
                                  slt_stream_t stream) {
    clear_error();
    if (blocks < 1 || iters < 1 || !out) {
        set_error("slt_valu_calibrate: need blocks >= 1, iters >= 1 and an output buffer");
        return SLT_ERR_INVALID;
    }
    const hipStream_t s = (hipStream_t)stream;
    const dim3 g(blocks), b(256);
    switch (kind) {
        case SLT_CALIB_FMA_SGPR:
            hipLaunchKernelGGL(valu_calib_kernel<SLT_CALIB_FMA_SGPR>, g, b, 0, s, 0.999f, 1e-3f,
                               iters, out);
            break;
        case SLT_CALIB_FMA_VGPR:
            hipLaunchKernelGGL(valu_calib_kernel<SLT_CALIB_FMA_VGPR>, g, b, 0, s, 0.999f, 1e-3f,
                               iters, out);
            break;
        case SLT_CALIB_FMAC:
            hipLaunchKernelGGL(valu_calib_kernel<SLT_CALIB_FMAC>, g, b, 0, s, 1e-4f, 1e-3f, iters,
                               out);
            break;
        case SLT_CALIB_ADD:
            hipLaunchKernelGGL(valu_calib_kernel<SLT_CALIB_ADD>, g, b, 0, s, 1e-6f, 0.f, iters,
                               out);
            break;
        case SLT_CALIB_MUL:
            hipLaunchKernelGGL(valu_calib_kernel<SLT_CALIB_MUL>, g, b, 0, s, 0.9999999f, 0.f,
                               iters, out);
            break;
        case SLT_CALIB_FMA64:
            hipLaunchKernelGGL(valu_calib_kernel<SLT_CALIB_FMA64>, g, b, 0, s, 0.f, 0.f, iters,
                               out);
            break;
        default:
            set_error("slt_valu_calibrate: unknown instruction form %d", kind);
            return SLT_ERR_INVALID;
    }
    return check_launch("slt_valu_calibrate");
}

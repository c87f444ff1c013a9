// Calibration kernels of known VALU content for the blend's utilisation
// figures (libubs_probe.so, measurement only; not part of the C ABI of the
// rasteriser).  Each thread runs `iters` rounds of 8 independent dependency
// chains of one instruction kind, so the SIMDs issue that kind back to back:
//   kind 0: v_fma_f32 (full-rate VALU);
//   kind 1: v_exp_f32 (the transcendental unit, as the blend's exp2/log2).
// The instruction counts per dispatch come from the same counters as the
// blend's (SQ_INSTS_VALU, SQ_INSTS_VALU_TRANS_F32); their rate over the
// dispatch's duration is the peak issue rate the blend is measured against.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

template <int KIND>
__global__ __launch_bounds__(256) void valu_probe_kernel(int iters, float seed, float *__restrict__ out) {
    float a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = seed + (float)(threadIdx.x * 8 + k) * 1e-7f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if constexpr (KIND == 0)
                    a[k] = __builtin_fmaf(a[k], 0.999999f, 1e-7f);
                else
                    a[k] = __builtin_amdgcn_exp2f(a[k] * -0.5f);
            }
        }
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += a[k];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" {

// Launches `blocks` workgroups of 256 threads, each thread `iters` x 128
// instructions of the kind; out: blocks * 256 floats.  Returns hipError_t.
__attribute__((visibility("default"))) int ubs_probe_valu(int kind, int blocks, int iters, float *out, void *stream) {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (kind == 0)
        valu_probe_kernel<0><<<blocks, 256, 0, st>>>(iters, 0.25f, out);
    else
        valu_probe_kernel<1><<<blocks, 256, 0, st>>>(iters, 0.25f, out);
    return (int)hipGetLastError();
}

}  // extern "C"

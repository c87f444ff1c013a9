// Per-primitive activation math of a D-dim Beta scene (BetaModel's getters,
// scene/beta_model.py:36-52, :103-121, and the view query :675-690), shared
// by the activation kernels and the fused Beta projection.
#pragma once
#pragma clang fp contract(off)

#include "ubs_common.h"

namespace ubs {

__device__ __forceinline__ float softplus_f(float x) { return x > 20.f ? x : log1pf(expf(x)); }
// d softplus / dx = sigmoid(x), 1 above the threshold (torch's softplus_backward)
__device__ __forceinline__ float softplus_grad(float x) {
    if (x > 20.f) return 1.f;
    const float z = expf(x);
    return z / (z + 1.f);
}
__device__ __forceinline__ float sigmoid_f(float x) { return 1.f / (1.f + expf(-x)); }

// R(q / |q|), row-major; the product formula of quat_to_rotmat.
__device__ __forceinline__ void quat_rot(const float (&q)[4], float (&R)[9]) {
    const float inv = 1.f / fmaxf(sqrtf((q[0] * q[0] + q[1] * q[1]) + (q[2] * q[2] + q[3] * q[3])), 1e-12f);
    const float w = q[0] * inv, x = q[1] * inv, y = q[2] * inv, z = q[3] * inv;
    R[0] = 1.f - 2.f * (y * y + z * z); R[1] = 2.f * (x * y - w * z);       R[2] = 2.f * (x * z + w * y);
    R[3] = 2.f * (x * y + w * z);       R[4] = 1.f - 2.f * (x * x + z * z); R[5] = 2.f * (y * z - w * x);
    R[6] = 2.f * (x * z - w * y);       R[7] = 2.f * (y * z + w * x);       R[8] = 1.f - 2.f * (x * x + y * y);
}

// VJP of quat_rot: the gradient g[9] of R (row-major) -> the gradient of the
// raw quaternion q[4], through the normalisation.
__device__ __forceinline__ void quat_rot_vjp(const float (&q)[4], const float (&g)[9], float (&vq)[4]) {
    const float len = fmaxf(sqrtf((q[0] * q[0] + q[1] * q[1]) + (q[2] * q[2] + q[3] * q[3])), 1e-12f);
    const float inv = 1.f / len;
    const float w = q[0] * inv, x = q[1] * inv, y = q[2] * inv, z = q[3] * inv;
    // dL/d(w,x,y,z) of the unit quaternion
    const float gw = 2.f * (x * (g[7] - g[5]) + y * (g[2] - g[6]) + z * (g[3] - g[1]));
    const float gx = 2.f * (-2.f * x * (g[4] + g[8]) + y * (g[1] + g[3]) + z * (g[2] + g[6]) + w * (g[7] - g[5]));
    const float gy = 2.f * (x * (g[1] + g[3]) - 2.f * y * (g[0] + g[8]) + z * (g[5] + g[7]) + w * (g[2] - g[6]));
    const float gz = 2.f * (x * (g[2] + g[6]) + y * (g[5] + g[7]) - 2.f * z * (g[0] + g[4]) + w * (g[3] - g[1]));
    // through the normalisation: (g - u (u . g)) / |q|
    const float dot = (w * gw + x * gx) + (y * gy + z * gz);
    vq[0] = (gw - w * dot) * inv;
    vq[1] = (gx - x * dot) * inv;
    vq[2] = (gy - y * dot) * inv;
    vq[3] = (gz - z * dot) * inv;
}

// Unit query direction (xyz - cam) / |xyz - cam| (scene/beta_model.py:675-690).
// Camera centre -(R^T t) of a row-major world->camera [4,4] matrix: one FMA
// chain per coordinate, the same wherever the centre is taken (in the fused
// projection or by ubs_camera_centers), so every path queries from the same
// bits.
__host__ __device__ inline void camera_centre(const float *vm, float (&cc)[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
        cc[k] = -__builtin_fmaf(vm[8 + k], vm[11], __builtin_fmaf(vm[4 + k], vm[7], vm[k] * vm[3]));
}

__device__ __forceinline__ void view_dir(float p0, float p1, float p2, const float *__restrict__ cam,
                                         float (&q)[3]) {
    const float d0 = p0 - cam[0], d1 = p1 - cam[1], d2 = p2 - cam[2];
    const float nrm = sqrtf((d0 * d0 + d1 * d1) + d2 * d2);
    q[0] = d0 / nrm;
    q[1] = d1 / nrm;
    q[2] = d2 / nrm;
}

}  // namespace ubs

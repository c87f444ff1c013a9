// Per-(camera, primitive) EWA projection math, shared by the projection
// kernels and the fused Beta projection: world->camera, near/far cull,
// Sigma_c = R Sigma R^T, perspective (or orthographic) Jacobian with the
// +-0.3 tan(fov) clamp, Sigma_2D = J Sigma_c J^T + eps2d I, conic = inverse,
// 1-sigma radius = ceil(sqrt(lambda_max)), off-image cull, and the VJP chain.
// Every product follows the row-dot-column order with ascending k
// ((a0 b0 + a1 b1) + a2 b2) and contraction is off, so radii, means2d,
// depths and conics are bit-identical to the float oracle.  The functions
// take plain pointers so that they serve both rows of global tensors and
// register arrays of the caller.
#pragma once
#pragma clang fp contract(off)

#include "ubs_common.h"

namespace ubs {

// Below this fraction of c00 c11 the fp32 determinant c00 c11 - c01^2 of the
// blurred 2-D covariance has lost more than 15 bits to cancellation (its
// relative error bound is 2^-23 c00 c11 / det); there the projection takes the
// determinant, the culling test, the inverse and the compensation in FP64.
// Elongated primitives reach it (det / (c00 c11) ~ 1.2 / lambda_max for a
// needle at 45 degrees), and so do conditioned covariances whose mirrored
// upper triangle is indefinite (the conditioning is not symmetric for unequal
// Beta factors), where det approaches 0 from above.
constexpr float kDetFp64Ratio = 0.00390625f;  // 2^-8

struct Geom {
    float R[3][3], t[3], pc[3], Sw[3][3], Sc[3][3];
    float fx, fy, cx, cy, W, H;
    float lxp, lxn, lyp, lyn, rz, rz2, tx, ty;
    float J[2][3], cov2d[2][2];
};

// reference: utils.cuh:13-36 (quaternion -> rotation; rsqrt replaced by the
// correctly rounded 1/sqrt so the result is reproducible)
__device__ __forceinline__ void quat_to_R(const float *q, float (&R)[3][3]) {
    float w = q[0], x = q[1], y = q[2], z = q[3];
    const float inv = 1.f / sqrtf(((x * x + y * y) + z * z) + w * w);
    x *= inv; y *= inv; z *= inv; w *= inv;
    const float x2 = x * x, y2 = y * y, z2 = z * z, xy = x * y, xz = x * z, yz = y * z;
    const float wx = w * x, wy = w * y, wz = w * z;
    R[0][0] = 1.f - 2.f * (y2 + z2); R[1][0] = 2.f * (xy + wz); R[2][0] = 2.f * (xz - wy);
    R[0][1] = 2.f * (xy - wz); R[1][1] = 1.f - 2.f * (x2 + z2); R[2][1] = 2.f * (yz + wx);
    R[0][2] = 2.f * (xz + wy); R[1][2] = 2.f * (yz - wx); R[2][2] = 1.f - 2.f * (x2 + y2);
}

// covars layout: UBS_COVAR_TRIU6 -> [N,6] (c00 c01 c02 c11 c12 c22);
// UBS_COVAR_FULL33 -> [N,3,3] row-major of which only the upper triangle is
// read (the gather rendering.py:54-56 does before projecting).
__device__ __forceinline__ void load_cov(const float *__restrict__ covars, const float *__restrict__ quats,
                                         const float *__restrict__ scales, uint32_t g, float (&S)[3][3],
                                         int cov33) {
    if (covars) {
        float c[6];
        if (cov33) {
            const float *m = covars + (size_t)g * 9;
            c[0] = m[0]; c[1] = m[1]; c[2] = m[2]; c[3] = m[4]; c[4] = m[5]; c[5] = m[8];
        } else {
            const float *m = covars + (size_t)g * 6;
#pragma unroll
            for (int k = 0; k < 6; ++k) c[k] = m[k];
        }
        S[0][0] = c[0]; S[0][1] = c[1]; S[0][2] = c[2];
        S[1][0] = c[1]; S[1][1] = c[3]; S[1][2] = c[4];
        S[2][0] = c[2]; S[2][1] = c[4]; S[2][2] = c[5];
    } else {
        float R[3][3], M[3][3];
        quat_to_R(quats + (size_t)g * 4, R);
        const float *s = scales + (size_t)g * 3;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) M[r][c] = R[r][c] * s[c];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) S[r][c] = (M[r][0] * M[c][0] + M[r][1] * M[c][1]) + M[r][2] * M[c][2];
    }
}

// reference: fully_fused_projection_fwd.cu:56-134, utils.cuh:181-211,252-292,374-413
__device__ __forceinline__ void geom(const float *__restrict__ means, const float *__restrict__ covars,
                                     const float *__restrict__ quats, const float *__restrict__ scales,
                                     const float *__restrict__ vm, const float *__restrict__ K, uint32_t g,
                                     int32_t W, int32_t H, bool ortho, int cov33, Geom &G) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) G.R[r][c] = vm[r * 4 + c];
        G.t[r] = vm[r * 4 + 3];
    }
    const float p0 = means[(size_t)g * 3], p1 = means[(size_t)g * 3 + 1], p2 = means[(size_t)g * 3 + 2];
#pragma unroll
    for (int r = 0; r < 3; ++r) G.pc[r] = ((G.R[r][0] * p0 + G.R[r][1] * p1) + G.R[r][2] * p2) + G.t[r];
    load_cov(covars, quats, scales, g, G.Sw, cov33);
    float M[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) M[r][c] = (G.R[r][0] * G.Sw[0][c] + G.R[r][1] * G.Sw[1][c]) + G.R[r][2] * G.Sw[2][c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) G.Sc[r][c] = (M[r][0] * G.R[c][0] + M[r][1] * G.R[c][1]) + M[r][2] * G.R[c][2];
    G.fx = K[0]; G.fy = K[4]; G.cx = K[2]; G.cy = K[5];
    G.W = (float)W; G.H = (float)H;
    const float x = G.pc[0], y = G.pc[1], z = G.pc[2];
    if (ortho) {
        G.J[0][0] = G.fx; G.J[0][1] = 0.f; G.J[0][2] = 0.f;
        G.J[1][0] = 0.f; G.J[1][1] = G.fy; G.J[1][2] = 0.f;
    } else {
        const float tfx = (0.5f * G.W) / G.fx, tfy = (0.5f * G.H) / G.fy;
        G.lxp = (G.W - G.cx) / G.fx + 0.3f * tfx;
        G.lxn = G.cx / G.fx + 0.3f * tfx;
        G.lyp = (G.H - G.cy) / G.fy + 0.3f * tfy;
        G.lyn = G.cy / G.fy + 0.3f * tfy;
        G.rz = 1.f / z;
        G.rz2 = G.rz * G.rz;
        G.tx = z * fminf(G.lxp, fmaxf(-G.lxn, x * G.rz));
        G.ty = z * fminf(G.lyp, fmaxf(-G.lyn, y * G.rz));
        G.J[0][0] = G.fx * G.rz; G.J[0][1] = 0.f; G.J[0][2] = -G.fx * G.tx * G.rz2;
        G.J[1][0] = 0.f; G.J[1][1] = G.fy * G.rz; G.J[1][2] = -G.fy * G.ty * G.rz2;
    }
    float A[2][3];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) A[r][c] = (G.J[r][0] * G.Sc[0][c] + G.J[r][1] * G.Sc[1][c]) + G.J[r][2] * G.Sc[2][c];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) G.cov2d[r][c] = (A[r][0] * G.J[c][0] + A[r][1] * G.J[c][1]) + A[r][2] * G.J[c][2];
}

// Outputs of one (camera, primitive) projection.
struct ProjOut {
    int32_t radius;
    float m2[2], depth, conic[3], comp;
};

// Projects primitive g for camera (vm, K); false when it is culled (near/far,
// degenerate 2-D covariance, radius_clip, off-image), which the caller
// records as radius 0.
// reference: fully_fused_projection_fwd.cu:20-177
__device__ __forceinline__ bool proj_fwd_body(const float *__restrict__ means, const float *__restrict__ covars,
                                              int cov33, const float *__restrict__ quats,
                                              const float *__restrict__ scales, const float *__restrict__ vm,
                                              const float *__restrict__ K, uint32_t g, int32_t W, int32_t H,
                                              float eps2d, float near_plane, float far_plane, float radius_clip,
                                              bool ortho, ProjOut &out) {
    const float *p = means + (size_t)g * 3;
    const float z = ((vm[8] * p[0] + vm[9] * p[1]) + vm[10] * p[2]) + vm[11];
    if (z < near_plane || z > far_plane) return false;
    Geom G;
    geom(means, covars, quats, scales, vm, K, g, W, H, ortho, cov33, G);
    float m2x, m2y;
    if (ortho) {
        m2x = G.fx * G.pc[0] + G.cx;
        m2y = G.fy * G.pc[1] + G.cy;
    } else {
        m2x = G.fx * G.pc[0] * G.rz + G.cx;
        m2y = G.fy * G.pc[1] * G.rz + G.cy;
    }
    float c00 = G.cov2d[0][0], c01 = G.cov2d[0][1], c10 = G.cov2d[1][0], c11 = G.cov2d[1][1];
    const float o00 = c00, o11 = c11;
    const float det_orig = c00 * c11 - c01 * c10;
    c00 += eps2d;
    c11 += eps2d;
    float det = c00 * c11 - c01 * c10;
    float cmp, i00, i01, i11;
    if (det < kDetFp64Ratio * (c00 * c11)) {
        // near-singular (or indefinite) 2-D covariance: the culling test, the
        // inverse and the compensation from an FP64 determinant of the same
        // fp32 entries (the products are exact in FP64, one rounding)
        const double d64 = (double)c00 * (double)c11 - (double)c01 * (double)c10;
        if (!(d64 > 0.0)) return false;
        const double o64 = (double)o00 * (double)o11 - (double)c01 * (double)c10;
        cmp = sqrtf(fmaxf(0.f, (float)(o64 / d64)));
        i00 = (float)((double)c11 / d64);
        i01 = (float)(-(double)c01 / d64);
        i11 = (float)((double)c00 / d64);
        det = (float)d64;
    } else {
        cmp = sqrtf(fmaxf(0.f, det_orig / det));
        if (det <= 0.f) return false;
        const float invd = 1.f / det;
        i00 = c11 * invd;
        i01 = -c01 * invd;
        i11 = c00 * invd;
    }
    const float b = 0.5f * (c00 + c11);
    const float v1 = b + sqrtf(fmaxf(0.01f, b * b - det));
    const float radius = ceilf(sqrtf(v1));
    if (radius <= radius_clip) return false;
    if (m2x + radius <= 0.f || m2x - radius >= (float)W || m2y + radius <= 0.f || m2y - radius >= (float)H)
        return false;
    out.radius = f2i_sat(radius);
    out.m2[0] = m2x;
    out.m2[1] = m2y;
    out.depth = G.pc[2];
    out.conic[0] = i00;
    out.conic[1] = i01;
    out.conic[2] = i11;
    out.comp = cmp;
    return true;
}

// VJP of proj_fwd_body for a visible (camera, primitive): from the saved
// conic cn[3] and the gradients of the 2-D mean vm2[2], depth, conic vc[3]
// (and compensation, when v_comp is not null) to the gradients of the 3-D
// mean, of the symmetric 3x3 covariance vS (before the triangle fold) and of
// the camera's rotation vR and translation vt.
// reference: fully_fused_projection_bwd.cu:50-239 with utils.cuh:38-62,
// 98-136, 214-249, 294-372, 385-435, 451-502.
__device__ __forceinline__ void proj_bwd_body(const float *__restrict__ means, const float *__restrict__ covars,
                                              int cov33, const float *__restrict__ quats,
                                              const float *__restrict__ scales, const float *__restrict__ vm,
                                              const float *__restrict__ K, uint32_t g, int32_t W, int32_t H,
                                              float eps2d, bool ortho, const float *__restrict__ cn,
                                              const float *__restrict__ comp, const float *__restrict__ vm2,
                                              float v_depth, const float *__restrict__ vc,
                                              const float *__restrict__ v_comp, float (&vmean)[3],
                                              float (&vS)[3][3], float (&vR)[3][3], float (&vt)[3]) {
    const float P[2][2] = {{cn[0], cn[1]}, {cn[1], cn[2]}};
    const float vP[2][2] = {{vc[0], vc[1] * 0.5f}, {vc[1] * 0.5f, vc[2]}};
    float vcov2[2][2];
    {
        float tmp[2][2];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int s = 0; s < 2; ++s) tmp[r][s] = P[r][0] * vP[0][s] + P[r][1] * vP[1][s];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int s = 0; s < 2; ++s) vcov2[r][s] = -(tmp[r][0] * P[0][s] + tmp[r][1] * P[1][s]);
    }
    if (v_comp) {
        const float cmpv = *comp, vcp = *v_comp;
        const float detc = P[0][0] * P[1][1] - P[0][1] * P[1][0];
        const float vsq = vcp * 0.5f / (cmpv + 1e-6f);
        const float om = 1.f - cmpv * cmpv;
        vcov2[0][0] += vsq * (om * P[0][0] - eps2d * detc);
        vcov2[0][1] += vsq * (om * P[0][1]);
        vcov2[1][0] += vsq * (om * P[1][0]);
        vcov2[1][1] += vsq * (om * P[1][1] - eps2d * detc);
    }
    Geom G;
    geom(means, covars, quats, scales, vm, K, g, W, H, ortho, cov33, G);
    float vSc[3][3], vpc[3] = {0.f, 0.f, 0.f};
    {
        float tmp[3][2];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int s = 0; s < 2; ++s) tmp[r][s] = G.J[0][r] * vcov2[0][s] + G.J[1][r] * vcov2[1][s];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int s = 0; s < 3; ++s) vSc[r][s] = tmp[r][0] * G.J[0][s] + tmp[r][1] * G.J[1][s];
    }
    const float x = G.pc[0], y = G.pc[1];
    if (ortho) {
        vpc[0] += G.fx * vm2[0];
        vpc[1] += G.fy * vm2[1];
    } else {
        const float rz = G.rz, rz2 = G.rz2, rz3 = rz2 * rz;
        vpc[0] += G.fx * rz * vm2[0];
        vpc[1] += G.fy * rz * vm2[1];
        vpc[2] += -(G.fx * x * vm2[0] + G.fy * y * vm2[1]) * rz2;
        float vJ[2][3];
        {
            float JS[2][3], JSt[2][3];
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    JS[r][s] = G.J[r][0] * G.Sc[0][s] + G.J[r][1] * G.Sc[1][s] + G.J[r][2] * G.Sc[2][s];
                    JSt[r][s] = G.J[r][0] * G.Sc[s][0] + G.J[r][1] * G.Sc[s][1] + G.J[r][2] * G.Sc[s][2];
                }
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int s = 0; s < 3; ++s)
                    vJ[r][s] = (vcov2[r][0] * JSt[0][s] + vcov2[r][1] * JSt[1][s]) +
                               (vcov2[0][r] * JS[0][s] + vcov2[1][r] * JS[1][s]);
        }
        if (x * rz <= G.lxp && x * rz >= -G.lxn) vpc[0] += -G.fx * rz2 * vJ[0][2];
        else vpc[2] += -G.fx * rz3 * vJ[0][2] * G.tx;
        if (y * rz <= G.lyp && y * rz >= -G.lyn) vpc[1] += -G.fy * rz2 * vJ[1][2];
        else vpc[2] += -G.fy * rz3 * vJ[1][2] * G.ty;
        vpc[2] += -G.fx * rz2 * vJ[0][0] - G.fy * rz2 * vJ[1][1] + 2.f * G.fx * G.tx * rz3 * vJ[0][2] +
                  2.f * G.fy * G.ty * rz3 * vJ[1][2];
    }
    vpc[2] += v_depth;
    const float p[3] = {means[(size_t)g * 3], means[(size_t)g * 3 + 1], means[(size_t)g * 3 + 2]};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int s = 0; s < 3; ++s) vR[r][s] = vpc[r] * p[s];
        vt[r] = vpc[r];
        vmean[r] = G.R[0][r] * vpc[0] + G.R[1][r] * vpc[1] + G.R[2][r] * vpc[2];
    }
    {
        float RS[3][3], RSt[3][3], tmp[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                RS[r][s] = G.R[r][0] * G.Sw[0][s] + G.R[r][1] * G.Sw[1][s] + G.R[r][2] * G.Sw[2][s];
                RSt[r][s] = G.R[r][0] * G.Sw[s][0] + G.R[r][1] * G.Sw[s][1] + G.R[r][2] * G.Sw[s][2];
            }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int s = 0; s < 3; ++s)
                vR[r][s] += (vSc[r][0] * RSt[0][s] + vSc[r][1] * RSt[1][s] + vSc[r][2] * RSt[2][s]) +
                            (vSc[0][r] * RS[0][s] + vSc[1][r] * RS[1][s] + vSc[2][r] * RS[2][s]);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int s = 0; s < 3; ++s)
                tmp[r][s] = G.R[0][r] * vSc[0][s] + G.R[1][r] * vSc[1][s] + G.R[2][r] * vSc[2][s];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int s = 0; s < 3; ++s)
                vS[r][s] = tmp[r][0] * G.R[0][s] + tmp[r][1] * G.R[1][s] + tmp[r][2] * G.R[2][s];
    }
}

}  // namespace ubs

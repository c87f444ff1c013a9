// EWA projection of (conditioned) 3-D Beta primitives on gfx950.
//
// Forward: one thread per (camera, primitive); world->camera, near/far cull,
// Sigma_c = R Sigma R^T, perspective (or orthographic) Jacobian with the
// +-0.3 tan(fov) clamp, Sigma_2D = J Sigma_c J^T + eps2d I, conic = inverse,
// 1-sigma radius = ceil(sqrt(lambda_max)), off-image cull.
// Every product follows the row-dot-column order with ascending k
// ((a0 b0 + a1 b1) + a2 b2) and contraction is off, so radii, means2d,
// depths and conics are bit-identical to the float oracle.
//
// Backward: the chain of VJPs of fully_fused_projection_bwd.cu:50-239;
// per-primitive gradients are stored directly for one camera and
// accumulated with hardware float atomics for several.
#pragma clang fp contract(off)

#include "ubs_projection_math.h"

namespace ubs {

// reference: fully_fused_projection_fwd.cu:20-177
__global__ __launch_bounds__(kThreads) void proj_fwd_kernel(uint32_t C, uint32_t N, const float *__restrict__ means,
                                const float *__restrict__ covars, int cov33, const float *__restrict__ quats,
                                const float *__restrict__ scales, const float *__restrict__ viewmats,
                                const float *__restrict__ Ks, int32_t W, int32_t H, float eps2d,
                                float near_plane, float far_plane, float radius_clip, int ortho,
                                int32_t *__restrict__ radii, float *__restrict__ means2d,
                                float *__restrict__ depths, float *__restrict__ conics,
                                float *__restrict__ comp) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (uint64_t)C * N) return;
    const uint32_t c = (uint32_t)(idx / N), g = (uint32_t)(idx % N);
    const float *vm = viewmats + (size_t)c * 16, *K = Ks + (size_t)c * 9;
    ProjOut o;
    if (!proj_fwd_body(means, covars, cov33, quats, scales, vm, K, g, W, H, eps2d, near_plane, far_plane,
                       radius_clip, ortho != 0, o)) {
        radii[idx] = 0;
        return;
    }
    radii[idx] = o.radius;
    means2d[idx * 2] = o.m2[0];
    means2d[idx * 2 + 1] = o.m2[1];
    depths[idx] = o.depth;
    conics[idx * 3] = o.conic[0];
    conics[idx * 3 + 1] = o.conic[1];
    conics[idx * 3 + 2] = o.conic[2];
    if (comp) comp[idx] = o.comp;
}

// reference: fully_fused_projection_bwd.cu:50-239 with utils.cuh:38-62,
// 98-136, 214-249, 294-372, 385-435, 451-502.
__global__ __launch_bounds__(kThreads) void proj_bwd_kernel(uint32_t C, uint32_t N, const float *__restrict__ means,
                                const float *__restrict__ covars, int cov33, const float *__restrict__ quats,
                                const float *__restrict__ scales, const float *__restrict__ viewmats,
                                const float *__restrict__ Ks, int32_t W, int32_t H, float eps2d, int ortho,
                                const int32_t *__restrict__ radii, const float *__restrict__ conics,
                                const float *__restrict__ comp, const float *__restrict__ v_means2d,
                                const float *__restrict__ v_depths, const float *__restrict__ v_conics,
                                const float *__restrict__ v_comp, float *__restrict__ v_means,
                                float *__restrict__ v_covars, float *__restrict__ v_quats,
                                float *__restrict__ v_scales, float *__restrict__ v_viewmats) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in_range = idx < (uint64_t)C * N;
    const bool direct = (C == 1);  // one camera: plain stores, no atomics
    const uint32_t c = in_range ? (uint32_t)(idx / N) : 0u;
    const uint32_t g = in_range ? (uint32_t)(idx % N) : 0u;
    const bool active = in_range && radii[idx] > 0;

    float vmean[3] = {0.f, 0.f, 0.f};
    float vS[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    float vR[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    float vt[3] = {0.f, 0.f, 0.f};
    if (active) {
        const float *vm = viewmats + (size_t)c * 16, *K = Ks + (size_t)c * 9;
        proj_bwd_body(means, covars, cov33, quats, scales, vm, K, g, W, H, eps2d, ortho != 0, conics + idx * 3,
                      v_comp ? comp + idx : nullptr, v_means2d + idx * 2, v_depths[idx], v_conics + idx * 3,
                      v_comp ? v_comp + idx : nullptr, vmean, vS, vR, vt);
    }
    if (in_range && (active || direct)) {
        // per-primitive outputs
        float *om = v_means + (size_t)g * 3;
        if (direct) {
            om[0] = vmean[0]; om[1] = vmean[1]; om[2] = vmean[2];
        } else {
            atomic_add_f32(om, vmean[0]); atomic_add_f32(om + 1, vmean[1]); atomic_add_f32(om + 2, vmean[2]);
        }
        if (covars) {
            const float o6[6] = {vS[0][0], vS[0][1] + vS[1][0], vS[0][2] + vS[2][0], vS[1][1],
                                 vS[1][2] + vS[2][1], vS[2][2]};
            if (cov33) {
                // gradient of the upper-triangle gather: upper entries get o6, lower stay 0
                constexpr int kSlot[6] = {0, 1, 2, 4, 5, 8};
                float *oc = v_covars + (size_t)g * 9;
                if (direct) {
                    oc[3] = 0.f;
                    oc[6] = 0.f;
                    oc[7] = 0.f;
                }
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    if (direct) oc[kSlot[k]] = o6[k];
                    else atomic_add_f32(oc + kSlot[k], o6[k]);
                }
            } else {
                float *oc = v_covars + (size_t)g * 6;
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    if (direct) oc[k] = o6[k];
                    else atomic_add_f32(oc + k, o6[k]);
                }
            }
        } else {
            float vq[4] = {0.f, 0.f, 0.f, 0.f}, vs3[3] = {0.f, 0.f, 0.f};
            if (active) {
                const float *q = quats + (size_t)g * 4, *s = scales + (size_t)g * 3;
                float R[3][3], M[3][3], vM[3][3], vRq[3][3];
                quat_to_R(q, R);
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) M[r][cc] = R[r][cc] * s[cc];
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc)
                        vM[r][cc] = (vS[r][0] + vS[0][r]) * M[0][cc] + (vS[r][1] + vS[1][r]) * M[1][cc] +
                                    (vS[r][2] + vS[2][r]) * M[2][cc];
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) vRq[r][cc] = vM[r][cc] * s[cc];
                float w = q[0], xq = q[1], yq = q[2], zq = q[3];
                const float inv = 1.f / sqrtf(((xq * xq + yq * yq) + zq * zq) + w * w);
                xq *= inv; yq *= inv; zq *= inv; w *= inv;
#define V(col, row) vRq[row][col]
                const float vqn[4] = {
                    2.f * (xq * (V(1, 2) - V(2, 1)) + yq * (V(2, 0) - V(0, 2)) + zq * (V(0, 1) - V(1, 0))),
                    2.f * (-2.f * xq * (V(1, 1) + V(2, 2)) + yq * (V(0, 1) + V(1, 0)) + zq * (V(0, 2) + V(2, 0)) +
                           w * (V(1, 2) - V(2, 1))),
                    2.f * (xq * (V(0, 1) + V(1, 0)) - 2.f * yq * (V(0, 0) + V(2, 2)) + zq * (V(1, 2) + V(2, 1)) +
                           w * (V(2, 0) - V(0, 2))),
                    2.f * (xq * (V(0, 2) + V(2, 0)) + yq * (V(1, 2) + V(2, 1)) - 2.f * zq * (V(0, 0) + V(1, 1)) +
                           w * (V(0, 1) - V(1, 0)))};
#undef V
                const float qn[4] = {w, xq, yq, zq};
                const float dot = vqn[0] * qn[0] + vqn[1] * qn[1] + vqn[2] * qn[2] + vqn[3] * qn[3];
#pragma unroll
                for (int k = 0; k < 4; ++k) vq[k] = (vqn[k] - dot * qn[k]) * inv;
#pragma unroll
                for (int j = 0; j < 3; ++j) vs3[j] = R[0][j] * vM[0][j] + R[1][j] * vM[1][j] + R[2][j] * vM[2][j];
            }
            float *oq = v_quats + (size_t)g * 4, *os = v_scales + (size_t)g * 3;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (direct) oq[k] = vq[k];
                else atomic_add_f32(oq + k, vq[k]);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (direct) os[k] = vs3[k];
                else atomic_add_f32(os + k, vs3[k]);
            }
        }
    }
    if (v_viewmats) {
        // per-camera sums; a wave may straddle two cameras, so reduce per
        // camera id present in the wave (at most a few distinct ids)
        // (the loop is wave-uniform: every lane runs every iteration, so the
        // cross-lane sums always see a full wave)
        bool pending = in_range;
        while (true) {
            const uint64_t left = __ballot(pending);
            if (left == 0) break;
            const int leader = __builtin_ctzll(left);
            const uint32_t leader_cam = (uint32_t)__builtin_amdgcn_readlane((int)c, leader);
            const bool mine = pending && (c == leader_cam);
            float vals[12];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
#pragma unroll
                for (int s = 0; s < 3; ++s) vals[r * 4 + s] = mine ? vR[r][s] : 0.f;
                vals[r * 4 + 3] = mine ? vt[r] : 0.f;
            }
            float tot[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) tot[k] = wave_sum(vals[k]);
            if (lane_id() == 0) {
#pragma unroll
                for (int k = 0; k < 12; ++k) atomic_add_f32(v_viewmats + (size_t)leader_cam * 16 + k, tot[k]);
            }
            if (mine) pending = false;
        }
    }
}

}  // namespace ubs

using namespace ubs;

extern "C" {

int ubs_fully_fused_projection_fwd(uint32_t C, uint32_t N, const float *means, const float *covars,
                                   int covar_layout, const float *quats, const float *scales, const float *viewmats,
                                   const float *Ks, int32_t W, int32_t H, float eps2d, float near_plane,
                                   float far_plane, float radius_clip, int ortho, int32_t *radii,
                                   float *means2d, float *depths, float *conics, float *comp, void *stream) {
    const uint64_t total = (uint64_t)C * N;
    if (total == 0) return UBS_OK;
    UBS_REQUIRE(total < (1ull << 31), "C*N must be < 2^31");
    UBS_REQUIRE(means && viewmats && Ks && radii && means2d && depths && conics, "null pointer");
    UBS_REQUIRE(covars || (quats && scales), "covars or (quats, scales) required");
    UBS_REQUIRE(covar_layout == UBS_COVAR_TRIU6 || covar_layout == UBS_COVAR_FULL33, "bad covar_layout %d",
                covar_layout);
    proj_fwd_kernel<<<grid_for((int64_t)total), kThreads, 0, as_stream(stream)>>>(
        C, N, means, covars, covar_layout == UBS_COVAR_FULL33, covars ? nullptr : quats, covars ? nullptr : scales, viewmats, Ks, W, H, eps2d,
        near_plane, far_plane, radius_clip, ortho, radii, means2d, depths, conics, comp);
    UBS_LAUNCH_CHECK("proj_fwd_kernel");
    return UBS_OK;
}

int ubs_fully_fused_projection_bwd(uint32_t C, uint32_t N, const float *means, const float *covars,
                                   int covar_layout, const float *quats, const float *scales, const float *viewmats,
                                   const float *Ks, int32_t W, int32_t H, float eps2d, int ortho,
                                   const int32_t *radii, const float *conics, const float *comp,
                                   const float *v_means2d, const float *v_depths, const float *v_conics,
                                   const float *v_comp, float *v_means, float *v_covars, float *v_quats,
                                   float *v_scales, float *v_viewmats, void *stream) {
    const uint64_t total = (uint64_t)C * N;
    hipStream_t st = as_stream(stream);
    UBS_REQUIRE(total < (1ull << 31), "C*N must be < 2^31");
    UBS_REQUIRE(covars || (quats && scales), "covars or (quats, scales) required");
    UBS_REQUIRE(!v_comp || comp, "v_compensations needs compensations");
    if (v_viewmats) {
        int e = zero_async(v_viewmats, (size_t)C * 16 * sizeof(float), st);
        if (e) return e;
    }
    if (N == 0) return UBS_OK;
    UBS_REQUIRE(v_means && (covars ? v_covars != nullptr : (v_quats && v_scales)), "null gradient output");
    if (C != 1) {
        int e = zero_async(v_means, (size_t)N * 3 * sizeof(float), st);
        if (!e && covars)
            e = zero_async(v_covars, (size_t)N * (covar_layout == UBS_COVAR_FULL33 ? 9 : 6) * sizeof(float), st);
        if (!e && !covars) e = zero_async(v_quats, (size_t)N * 4 * sizeof(float), st);
        if (!e && !covars) e = zero_async(v_scales, (size_t)N * 3 * sizeof(float), st);
        if (e) return e;
    }
    if (total == 0) return UBS_OK;
    UBS_REQUIRE(means && viewmats && Ks && radii && conics && v_means2d && v_depths && v_conics, "null pointer");
    UBS_REQUIRE(covar_layout == UBS_COVAR_TRIU6 || covar_layout == UBS_COVAR_FULL33, "bad covar_layout %d",
                covar_layout);
    proj_bwd_kernel<<<grid_for((int64_t)total), kThreads, 0, st>>>(
        C, N, means, covars, covar_layout == UBS_COVAR_FULL33, covars ? nullptr : quats, covars ? nullptr : scales, viewmats, Ks, W, H, eps2d,
        ortho, radii, conics, comp, v_means2d, v_depths, v_conics, v_comp, v_means, covars ? v_covars : nullptr,
        covars ? nullptr : v_quats, covars ? nullptr : v_scales, v_viewmats);
    UBS_LAUNCH_CHECK("proj_bwd_kernel");
    return UBS_OK;
}

}  // extern "C"

// Parameter activations, view/time query and quaternion rotation of a D-dim
// Beta scene in one pass per primitive, and the matching backward.
//
// What BetaModel does with separate torch elementwise ops before it calls the
// rasteriser (scene/beta_model.py:36-52 activations, :103-121 getters,
// :675-690 query):
//   mean    = cat(xyz, mean_rest)                 [N,D]
//   scale   = softplus(scale_raw)                 [N,D]   (torch: x if x > 20)
//   opacity = sigmoid(opacity_raw)                [N,1]
//   beta    = 4 exp(beta_raw)                     [N,D-2]
//   query   = (xyz - cam) / |xyz - cam| (, t)      [N,D-3] (D = 6 or 7)
//   rot     = R(q / |q|)  from (w,x,y,z) quats    [N,3,3] (optional)
// The query carries no gradient (the conditioning backward drops it,
// _wrapper.py:611), so the backward is elementwise per leaf.  One thread per
// primitive; every input row is read once and every output row written once.
#pragma clang fp contract(off)

#include "ubs_activate_math.h"

namespace ubs {

template <int D>
__global__ __launch_bounds__(kThreads) void activate_fwd_kernel(
    int64_t N, const float *__restrict__ xyz, const float *__restrict__ mean_rest, const float *__restrict__ scale_raw,
    const float *__restrict__ opacity_raw, const float *__restrict__ beta_raw, const float *__restrict__ quats,
    const float *__restrict__ cam, float timestamp, float *__restrict__ mean, float *__restrict__ scale,
    float *__restrict__ opacity, float *__restrict__ beta, float *__restrict__ query, float *__restrict__ rot) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const float p0 = xyz[n * 3], p1 = xyz[n * 3 + 1], p2 = xyz[n * 3 + 2];
    float *m = mean + n * D;
    m[0] = p0; m[1] = p1; m[2] = p2;
#pragma unroll
    for (int k = 3; k < D; ++k) m[k] = mean_rest[n * (D - 3) + (k - 3)];
#pragma unroll
    for (int k = 0; k < D; ++k) scale[n * D + k] = softplus_f(scale_raw[n * D + k]);
    opacity[n] = sigmoid_f(opacity_raw[n]);
#pragma unroll
    for (int k = 0; k < D - 2; ++k) beta[n * (D - 2) + k] = 4.f * expf(beta_raw[n * (D - 2) + k]);
    if (query) {
        float qd[3];
        view_dir(p0, p1, p2, cam, qd);
        float *q = query + n * (D - 3);
        q[0] = qd[0]; q[1] = qd[1]; q[2] = qd[2];
        if constexpr (D == 7) q[3] = timestamp;
    }
    if (rot) {
        float qv[4], R[9];
#pragma unroll
        for (int k = 0; k < 4; ++k) qv[k] = quats[n * 4 + k];
        quat_rot(qv, R);
#pragma unroll
        for (int k = 0; k < 9; ++k) rot[n * 9 + k] = R[k];
    }
}

template <int D>
__global__ __launch_bounds__(kThreads) void activate_bwd_kernel(
    int64_t N, const float *__restrict__ scale_raw, const float *__restrict__ opacity_raw,
    const float *__restrict__ beta_raw, const float *__restrict__ quats, const float *__restrict__ v_mean,
    const float *__restrict__ v_scale, const float *__restrict__ v_opacity, const float *__restrict__ v_beta,
    const float *__restrict__ v_rot, float *__restrict__ v_xyz, float *__restrict__ v_mean_rest,
    float *__restrict__ v_scale_raw, float *__restrict__ v_opacity_raw, float *__restrict__ v_beta_raw,
    float *__restrict__ v_quats) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) v_xyz[n * 3 + k] = v_mean ? v_mean[n * D + k] : 0.f;
#pragma unroll
    for (int k = 3; k < D; ++k) v_mean_rest[n * (D - 3) + (k - 3)] = v_mean ? v_mean[n * D + k] : 0.f;
#pragma unroll
    for (int k = 0; k < D; ++k)
        v_scale_raw[n * D + k] = v_scale ? v_scale[n * D + k] * softplus_grad(scale_raw[n * D + k]) : 0.f;
    {
        const float o = sigmoid_f(opacity_raw[n]);
        v_opacity_raw[n] = v_opacity ? v_opacity[n] * (o * (1.f - o)) : 0.f;
    }
#pragma unroll
    for (int k = 0; k < D - 2; ++k) {
        const int64_t i = n * (D - 2) + k;
        v_beta_raw[i] = v_beta ? v_beta[i] * (4.f * expf(beta_raw[i])) : 0.f;
    }
    if (v_quats) {
        float q[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = quats[n * 4 + k];
        float g[9], vq[4];
#pragma unroll
        for (int k = 0; k < 9; ++k) g[k] = v_rot ? v_rot[n * 9 + k] : 0.f;
        quat_rot_vjp(q, g, vq);
#pragma unroll
        for (int k = 0; k < 4; ++k) v_quats[n * 4 + k] = vq[k];
    }
}

template <int D>
static int launch_activate_fwd(int64_t N, const float *xyz, const float *mean_rest, const float *scale_raw,
                               const float *opacity_raw, const float *beta_raw, const float *quats, const float *cam,
                               float t, float *mean, float *scale, float *opacity, float *beta, float *query,
                               float *rot, hipStream_t st) {
    activate_fwd_kernel<D><<<grid_for(N), kThreads, 0, st>>>(N, xyz, mean_rest, scale_raw, opacity_raw, beta_raw,
                                                             quats, cam, t, mean, scale, opacity, beta, query, rot);
    UBS_LAUNCH_CHECK("activate_fwd_kernel");
    return UBS_OK;
}

template <int D>
static int launch_activate_bwd(int64_t N, const float *scale_raw, const float *opacity_raw, const float *beta_raw,
                               const float *quats, const float *v_mean, const float *v_scale, const float *v_opacity,
                               const float *v_beta, const float *v_rot, float *v_xyz, float *v_mean_rest,
                               float *v_scale_raw, float *v_opacity_raw, float *v_beta_raw, float *v_quats,
                               hipStream_t st) {
    activate_bwd_kernel<D><<<grid_for(N), kThreads, 0, st>>>(N, scale_raw, opacity_raw, beta_raw, quats, v_mean,
                                                             v_scale, v_opacity, v_beta, v_rot, v_xyz, v_mean_rest,
                                                             v_scale_raw, v_opacity_raw, v_beta_raw, v_quats);
    UBS_LAUNCH_CHECK("activate_bwd_kernel");
    return UBS_OK;
}

}  // namespace ubs

using namespace ubs;

extern "C" {

int ubs_beta_activate_fwd(int64_t N, int64_t D, const float *xyz, const float *mean_rest, const float *scale_raw,
                          const float *opacity_raw, const float *beta_raw, const float *quats, const float *cam_pos,
                          float timestamp, float *mean, float *scale, float *opacity, float *beta, float *query,
                          float *rot, void *stream) {
    UBS_REQUIRE(N >= 0 && N < (1ll << 31), "N out of range");
    UBS_REQUIRE(D >= 4 && D <= 11, "D must be in [4, 11] (got %lld)", (long long)D);
    UBS_REQUIRE(!query || D == 6 || D == 7, "the view/time query is defined for D = 6 or 7 (got %lld)",
                (long long)D);
    UBS_REQUIRE(!rot || quats, "rot needs quats");
    if (N == 0) return UBS_OK;
    UBS_REQUIRE(xyz && mean_rest && scale_raw && opacity_raw && beta_raw && mean && scale && opacity && beta &&
                    (!query || cam_pos),
                "null pointer");
    hipStream_t st = as_stream(stream);
    switch (D) {
#define UBS_ACT_CASE(K)                                                                                        \
    case K:                                                                                                    \
        return launch_activate_fwd<K>(N, xyz, mean_rest, scale_raw, opacity_raw, beta_raw, quats, cam_pos,    \
                                      timestamp, mean, scale, opacity, beta, query, rot, st);
        UBS_ACT_CASE(4) UBS_ACT_CASE(5) UBS_ACT_CASE(6) UBS_ACT_CASE(7) UBS_ACT_CASE(8) UBS_ACT_CASE(9)
        UBS_ACT_CASE(10) UBS_ACT_CASE(11)
#undef UBS_ACT_CASE
    }
    return UBS_ERR_INVALID;
}

int ubs_beta_activate_bwd(int64_t N, int64_t D, const float *scale_raw, const float *opacity_raw,
                          const float *beta_raw, const float *quats, const float *v_mean, const float *v_scale,
                          const float *v_opacity, const float *v_beta, const float *v_rot, float *v_xyz,
                          float *v_mean_rest, float *v_scale_raw, float *v_opacity_raw, float *v_beta_raw,
                          float *v_quats, void *stream) {
    UBS_REQUIRE(N >= 0 && N < (1ll << 31), "N out of range");
    UBS_REQUIRE(D >= 4 && D <= 11, "D must be in [4, 11] (got %lld)", (long long)D);
    UBS_REQUIRE(!v_quats || quats, "v_quats needs quats");
    if (N == 0) return UBS_OK;
    UBS_REQUIRE(scale_raw && opacity_raw && beta_raw && v_xyz && v_mean_rest && v_scale_raw && v_opacity_raw &&
                    v_beta_raw,
                "null pointer");
    hipStream_t st = as_stream(stream);
    switch (D) {
#define UBS_ACT_CASE(K)                                                                                        \
    case K:                                                                                                    \
        return launch_activate_bwd<K>(N, scale_raw, opacity_raw, beta_raw, quats, v_mean, v_scale, v_opacity,  \
                                      v_beta, v_rot, v_xyz, v_mean_rest, v_scale_raw, v_opacity_raw,           \
                                      v_beta_raw, v_quats, st);
        UBS_ACT_CASE(4) UBS_ACT_CASE(5) UBS_ACT_CASE(6) UBS_ACT_CASE(7) UBS_ACT_CASE(8) UBS_ACT_CASE(9)
        UBS_ACT_CASE(10) UBS_ACT_CASE(11)
#undef UBS_ACT_CASE
    }
    return UBS_ERR_INVALID;
}

}  // extern "C"

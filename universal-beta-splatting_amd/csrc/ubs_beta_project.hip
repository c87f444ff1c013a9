// Fused Beta projection on gfx950: from the leaf parameters of a D-dim Beta
// scene (D = 6 static, 7 dynamic) straight to the per-view 2-D primitives the
// tile intersection and the blend consume, for C views in one launch.
//
// One thread per primitive.  The per-primitive part -- activations
// (scene/beta_model.py:36-52, :103-121), the rotation (unit quaternion or the
// first-order l_triangle skew of l_triagnle_to_rotmat_fwd.cu:16-35) and
// Sigma = L L^T (rot_scale_l_triangle_to_covar_fwd.cu:28-187) -- runs once,
// in registers; then for every view: the query (beta_model.py:675-690),
// the conditioning on it (cond_mean_convariance_opacity_fwd.cu:106-290) and
// the EWA projection of the conditioned 3-D primitive
// (fully_fused_projection_fwd.cu:20-177, reading the upper triangle of the
// conditioned covariance as rendering.py:54-56 gathers it).
//
// The reference runs these as ~45 torch launches and 4 kernels per view with
// [N,D,D] covariances, [N,3,3] rotations and conditioned [N,3,3] matrices
// round-tripping through HBM; here the leaves (140 B per 6-D primitive) are
// read once for all views and only the 2-D records (40 B per view) written.
//
// The backward recomputes the forward per view in registers and chains the
// VJPs of the same functions (projection -> triangle fold -> conditioning),
// sums the per-view gradients of the mean, covariance, opacity and Beta
// parameters in registers, then applies the covariance, rotation and
// activation VJPs once: each leaf gradient is written exactly once, no
// atomics.  Every stage calls the same device functions as the separate
// kernels (ubs_*_math.h), so with one view the fused results are
// bit-identical to the drop-in chain.
#pragma clang fp contract(off)

#include "ubs_activate_math.h"
#include "ubs_cond_math.h"
#include "ubs_covar_math.h"
#include "ubs_projection_math.h"

namespace ubs {

// Per-primitive state shared by every view.
template <int D>
struct BetaPrim {
    float mean[D], scale[D], beta[D - 2], opac;
    float R[9];
    float L[D][D];
    uint32_t P[D];
    float Sig[D * D];
};

template <int D, bool QUAT>
__device__ __forceinline__ void beta_prim_setup(int64_t n, const float *__restrict__ xyz,
                                                const float *__restrict__ mean_rest,
                                                const float *__restrict__ scale_raw,
                                                const float *__restrict__ ltri,
                                                const float *__restrict__ opacity_raw,
                                                const float *__restrict__ beta_raw,
                                                const float *__restrict__ quats, BetaPrim<D> &bp) {
    constexpr int FULL_M = D * (D - 1) / 2;
#pragma unroll
    for (int k = 0; k < 3; ++k) bp.mean[k] = xyz[n * 3 + k];
#pragma unroll
    for (int k = 3; k < D; ++k) bp.mean[k] = mean_rest[n * (D - 3) + (k - 3)];
#pragma unroll
    for (int k = 0; k < D; ++k) bp.scale[k] = softplus_f(scale_raw[n * D + k]);
    bp.opac = sigmoid_f(opacity_raw[n]);
#pragma unroll
    for (int k = 0; k < D - 2; ++k) bp.beta[k] = 4.f * expf(beta_raw[n * (D - 2) + k]);
    const float *lt = ltri + n * FULL_M;
    if constexpr (QUAT) {
        float q[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = quats[n * 4 + k];
        quat_rot(q, bp.R);
    } else {
        const float a0 = lt[0], a1 = lt[1], a2 = lt[2];
        bp.R[0] = 1.f; bp.R[1] = a0;  bp.R[2] = a1;
        bp.R[3] = -a0; bp.R[4] = 1.f; bp.R[5] = a2;
        bp.R[6] = -a1; bp.R[7] = -a2; bp.R[8] = 1.f;
    }
    float xr[FULL_M - 3];
#pragma unroll
    for (int k = 0; k < FULL_M - 3; ++k) xr[k] = lt[3 + k];
    load_L<D>(bp.R, bp.scale, xr, nullptr, FULL_M - 3, true, bp.L, bp.P);
    covar_fwd_body<D>(bp.L, bp.P, false, bp.Sig);
}

// Query of view c: unit direction from the camera centre (cam_pos, or taken
// from the view matrix when cam_pos is null) (+ the view's time for 7-D).
template <int D>
__device__ __forceinline__ void beta_query(const BetaPrim<D> &bp, const float *__restrict__ cam_pos,
                                           const float *__restrict__ timestamps,
                                           const float *__restrict__ viewmats, uint32_t c, float (&q)[D - 3]) {
    float qd[3], cc[3];
    if (!cam_pos) camera_centre(viewmats + (size_t)c * 16, cc);
    view_dir(bp.mean[0], bp.mean[1], bp.mean[2], cam_pos ? cam_pos + (size_t)c * 3 : cc, qd);
    q[0] = qd[0];
    q[1] = qd[1];
    q[2] = qd[2];
    if constexpr (D == 7) q[3] = timestamps ? timestamps[c] : 0.f;
}

// View c of a primitive: its query, the conditioning on it and the EWA
// projection of the conditioned 3-D primitive, written at row entry idx.
template <int D>
__device__ __forceinline__ void beta_view_fwd(const BetaPrim<D> &bp, uint32_t c, int64_t idx,
                                              const float *__restrict__ cam_pos,
                                              const float *__restrict__ timestamps,
                                              const float *__restrict__ viewmats, const float *__restrict__ Ks,
                                              int32_t W, int32_t H, const int32_t *__restrict__ image_wh,
                                              float eps2d, float near_plane, float far_plane, float radius_clip,
                                              int ortho, int32_t *__restrict__ radii, float *__restrict__ means2d,
                                              float *__restrict__ depths, float *__restrict__ conics,
                                              float *__restrict__ opacities, float *__restrict__ betas,
                                              const float *__restrict__ rgb_n = nullptr,
                                              float4 *__restrict__ rec = nullptr) {
    constexpr int CQ = D - 3;
    float q[CQ];
    beta_query<D>(bp, cam_pos, timestamps, viewmats, c, q);
    CondRegs<CQ> st;
    cond_setup<CQ>(bp.mean, bp.Sig, bp.beta + 1, q, st);
    float m3[3], V9[9], oc;
    cond_fwd_body<CQ>(st, bp.mean, bp.opac, m3, V9, &oc);
    opacities[idx] = oc;
    betas[idx] = bp.beta[0];
    const int32_t Wc = image_wh ? image_wh[2 * c] : W, Hc = image_wh ? image_wh[2 * c + 1] : H;
    ProjOut o;
    if (!proj_fwd_body(m3, V9, 1, nullptr, nullptr, viewmats + (size_t)c * 16, Ks + (size_t)c * 9, 0u, Wc, Hc,
                       eps2d, near_plane, far_plane, radius_clip, ortho != 0, o)) {
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
    if (rec) {  // the blend's 64-byte record of the entry (what its packing writes)
        rec[4 * idx] = make_float4(o.m2[0], o.m2[1], o.conic[0], o.conic[1]);
        rec[4 * idx + 1] = make_float4(o.conic[2], oc, bp.beta[0], 0.f);
        rec[4 * idx + 2] = make_float4(rgb_n[0], rgb_n[1], rgb_n[2], 0.f);
    }
}

template <int D, bool QUAT>
__global__ __launch_bounds__(kThreads) void beta_project_fwd_kernel(
    int64_t N, uint32_t C, const float *__restrict__ xyz, const float *__restrict__ mean_rest,
    const float *__restrict__ scale_raw, const float *__restrict__ ltri, const float *__restrict__ opacity_raw,
    const float *__restrict__ beta_raw, const float *__restrict__ quats, const float *__restrict__ cam_pos,
    const float *__restrict__ timestamps, const float *__restrict__ viewmats, const float *__restrict__ Ks,
    int32_t W, int32_t H, const int32_t *__restrict__ image_wh, float eps2d, float near_plane, float far_plane,
    float radius_clip, int ortho, int32_t *__restrict__ radii, float *__restrict__ means2d,
    float *__restrict__ depths, float *__restrict__ conics, float *__restrict__ opacities,
    float *__restrict__ betas, int64_t row_stride, const float *__restrict__ rgb = nullptr,
    float4 *__restrict__ rec = nullptr) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    BetaPrim<D> bp;
    beta_prim_setup<D, QUAT>(n, xyz, mean_rest, scale_raw, ltri, opacity_raw, beta_raw, quats, bp);
    const float *rgb_n = rec ? rgb + 3 * n : nullptr;
    for (uint32_t c = 0; c < C; ++c)
        beta_view_fwd<D>(bp, c, (int64_t)c * row_stride + n, cam_pos, timestamps, viewmats, Ks, W, H, image_wh,
                         eps2d, near_plane, far_plane, radius_clip, ortho, radii, means2d, depths, conics,
                         opacities, betas, rgb_n, rec);
}

// Views of several resident scenes in one batch: view c projects the scene
// table[c] names into row c (entries c * row_stride + n).  The launch for D
// takes the views of D-dim scenes and the empty slots (N = 0) when D = 6;
// entries past a view's N get radius 0 (culled).  A primitive's
// view-independent state is kept while consecutive views share the scene.
// The colours of each view's scene are copied into its row as well, so the
// blend reads one [C, row_stride, 3] array.
template <int D>
__global__ __launch_bounds__(kThreads) void beta_project_fwd_table_kernel(
    uint32_t C, int64_t row_stride, const ubs_view_scene *__restrict__ table, const float *__restrict__ cam_pos,
    const float *__restrict__ timestamps, const float *__restrict__ viewmats, const float *__restrict__ Ks,
    const int32_t *__restrict__ image_wh, float eps2d, float near_plane, float far_plane, float radius_clip,
    int32_t *__restrict__ radii, float *__restrict__ means2d, float *__restrict__ depths,
    float *__restrict__ conics, float *__restrict__ opacities, float *__restrict__ betas,
    float *__restrict__ colors) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= row_stride) return;
    BetaPrim<D> bp;
    const float *held = nullptr;  // the scene whose state bp holds
    for (uint32_t c = 0; c < C; ++c) {
        const ubs_view_scene vs = table[c];
        if ((vs.N > 0 ? vs.D : 6) != D) continue;
        const int64_t idx = (int64_t)c * row_stride + n;
        if (n >= vs.N) {
            radii[idx] = 0;
            continue;
        }
        if (vs.xyz != held) {
            if (vs.quats)
                beta_prim_setup<D, true>(n, vs.xyz, vs.mean_rest, vs.scale_raw, vs.l_triangle, vs.opacity_raw,
                                         vs.beta_raw, vs.quats, bp);
            else
                beta_prim_setup<D, false>(n, vs.xyz, vs.mean_rest, vs.scale_raw, vs.l_triangle, vs.opacity_raw,
                                          vs.beta_raw, nullptr, bp);
            held = vs.xyz;
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) colors[idx * 3 + k] = vs.rgb[n * 3 + k];
        beta_view_fwd<D>(bp, c, idx, cam_pos, timestamps, viewmats, Ks, 0, 0, image_wh, eps2d, near_plane,
                         far_plane, radius_clip, 0, radii, means2d, depths, conics, opacities, betas);
    }
}

// Minimum waves per SIMD of the 7-D backward.  Unbounded it once took 317
// registers (one wave per SIMD), and held to two waves it spilled 55 values
// to scratch yet ran faster (0.262 vs 0.291 ms at 1M primitives,
// 1352x1014); with the covariance sums in LDS and the leaves read again
// after the view loop it fits two waves in 227 registers, no scratch
// (0.227 ms).  The 6-D kernel fits two waves by itself (225 registers).
#ifndef UBS_PROJ_BWD_WAVES_7
#define UBS_PROJ_BWD_WAVES_7 2
#endif
#ifndef UBS_PROJ_BWD_WAVES_6
#define UBS_PROJ_BWD_WAVES_6 1
#endif

template <int D, bool QUAT>
__global__ __launch_bounds__(kThreads, D == 7 ? UBS_PROJ_BWD_WAVES_7 : UBS_PROJ_BWD_WAVES_6) void beta_project_bwd_kernel(
    int64_t N, uint32_t C, const float *__restrict__ xyz, const float *__restrict__ mean_rest,
    const float *__restrict__ scale_raw, const float *__restrict__ ltri, const float *__restrict__ opacity_raw,
    const float *__restrict__ beta_raw, const float *__restrict__ quats, const float *__restrict__ cam_pos,
    const float *__restrict__ timestamps, const float *__restrict__ viewmats, const float *__restrict__ Ks,
    int32_t W, int32_t H, const int32_t *__restrict__ image_wh, float eps2d, int ortho, int exact,
    const int32_t *__restrict__ radii, const float *__restrict__ conics, const float *__restrict__ v_means2d,
    const float *__restrict__ v_depths, const float *__restrict__ v_conics, const float *__restrict__ v_opacities,
    const float *__restrict__ v_betas, float *__restrict__ v_xyz, float *__restrict__ v_mean_rest,
    float *__restrict__ v_scale_raw, float *__restrict__ v_ltri, float *__restrict__ v_opacity_raw,
    float *__restrict__ v_beta_raw, float *__restrict__ v_quats) {
    constexpr int CQ = D - 3;
    constexpr int FULL_M = D * (D - 1) / 2;
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    BetaPrim<D> bp;
    beta_prim_setup<D, QUAT>(n, xyz, mean_rest, scale_raw, ltri, opacity_raw, beta_raw, quats, bp);
    // gradients of the activated mean, covariance, opacity and Beta
    // parameters, summed over the views
    // The covariance VJP reads its gradient G only as G[a][b] + G[b][a]
    // (Sigma = L L^T), so the views' gradients are summed as that symmetric
    // part, upper triangle only: gU[a][b] = sum of G[a][b] + G[b][a] for
    // a < b and of G[a][a] on the diagonal (28 sums instead of 49 for 7-D).
    constexpr int NU = D * (D + 1) / 2;
    // 7-D: the NU sums live in LDS, one column per thread (stride kThreads:
    // the 64 lanes of a wave hit 64 consecutive words), not in registers.
    // The view loop is the kernel's register peak, and 28 fewer live values
    // there leave it 227 registers at two waves per SIMD, no scratch (0.227
    // against 0.229 ms with them in registers at 256, at 1M primitives,
    // 1352x1014).  6-D keeps them in registers (same time for one view, 1%
    // faster for eight).
    constexpr bool kLdsSums = D == 7;
    __shared__ float s_gU[kLdsSums ? NU : 1][kLdsSums ? kThreads : 1];
    float gUr[kLdsSums ? 1 : NU];
    auto gU = [&](int u) -> float & {
        if constexpr (kLdsSums)
            return s_gU[u][threadIdx.x];
        else
            return gUr[u];
    };
    float gmean[D], gop = 0.f, gbeta[D - 2];
#pragma unroll
    for (int k = 0; k < D; ++k) gmean[k] = 0.f;
#pragma unroll
    for (int k = 0; k < NU; ++k) gU(k) = 0.f;
#pragma unroll
    for (int k = 0; k < D - 2; ++k) gbeta[k] = 0.f;
    for (uint32_t c = 0; c < C; ++c) {
        const int64_t idx = (int64_t)c * N + n;
        // a culled view has no 2-D record, so every gradient it could pass is 0
        if (radii[idx] <= 0) continue;
        float q[CQ];
        beta_query<D>(bp, cam_pos, timestamps, viewmats, c, q);
        CondRegs<CQ> st;
        cond_setup<CQ>(bp.mean, bp.Sig, bp.beta + 1, q, st);
        float m3[3], V9[9], oc;
        cond_fwd_body<CQ>(st, bp.mean, bp.opac, m3, V9, &oc);
        const int32_t Wc = image_wh ? image_wh[2 * c] : W, Hc = image_wh ? image_wh[2 * c + 1] : H;
        float vmean[3], vS[3][3], vR[3][3], vt[3];
        proj_bwd_body(m3, V9, 1, nullptr, nullptr, viewmats + (size_t)c * 16, Ks + (size_t)c * 9, 0u, Wc, Hc, eps2d,
                      ortho != 0, conics + idx * 3, nullptr, v_means2d + idx * 2, v_depths ? v_depths[idx] : 0.f,
                      v_conics + idx * 3,
                      nullptr, vmean, vS, vR, vt);
        // the projection reads the upper triangle of the conditioned
        // covariance, so its gradient is the folded upper triangle
        const float Gv[9] = {vS[0][0], vS[0][1] + vS[1][0], vS[0][2] + vS[2][0],
                             0.f,      vS[1][1],            vS[1][2] + vS[2][1],
                             0.f,      0.f,                 vS[2][2]};
        float gm[D], gV[D * D], go, gb[CQ];
        cond_bwd_body<CQ>(st, bp.opac, vmean, Gv, v_opacities[idx], exact, gm, gV, &go, gb);
#pragma unroll
        for (int k = 0; k < D; ++k) gmean[k] += gm[k];
        {
            int u = 0;
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = a; b < D; ++b, ++u)
                    gU(u) += a == b ? gV[a * D + a] : gV[a * D + b] + gV[b * D + a];
        }
        gop += go;
        gbeta[0] += v_betas[idx];
#pragma unroll
        for (int k = 0; k < CQ; ++k) gbeta[1 + k] += gb[k];
    }
    // the mean, opacity and Beta gradients are final after the view loop:
    // written before the covariance VJP, whose peak then holds 13 values fewer
#pragma unroll
    for (int k = 0; k < 3; ++k) v_xyz[n * 3 + k] = gmean[k];
#pragma unroll
    for (int k = 3; k < D; ++k) v_mean_rest[n * (D - 3) + (k - 3)] = gmean[k];
    {
        const float o = sigmoid_f(opacity_raw[n]);
        v_opacity_raw[n] = gop * (o * (1.f - o));
    }
#pragma unroll
    for (int k = 0; k < D - 2; ++k) {
        const int64_t i = n * (D - 2) + k;
        v_beta_raw[i] = gbeta[k] * (4.f * expf(beta_raw[i]));
    }
    // covariance, rotation and activation VJPs, once per primitive.  The
    // rotation, scales and L factor are set up again here rather than kept
    // through the view loop (which reads only the mean, Sigma, opacity and
    // Beta parameters): the same values, ~70 fewer registers live in the loop
    // (the leaves are read again through an index the compiler cannot
    // match with the first reads, so no loaded value is kept live through
    // the loop in their place)
    BetaPrim<D> bq;
    int64_t n2 = n;
    if constexpr (D == 7) {
        asm volatile("" : "+v"(n2));
        beta_prim_setup<D, QUAT>(n2, xyz, mean_rest, scale_raw, ltri, opacity_raw, beta_raw, quats, bq);
    } else {
        bq = bp;  // 6-D fits its registers without the set-up again
    }
    // a G whose G[a][b] + G[b][a] is gU exactly: the triangle above the
    // diagonal, zeros below, the diagonal as summed (x + 0 and d + d are exact)
    float gSig[D * D];
    {
        int u = 0;
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = a; b < D; ++b, ++u) {
                gSig[a * D + b] = gU(u);
                if (b != a) gSig[b * D + a] = 0.f;
            }
    }
    float dR[9], dS[D];
    covar_bwd_body<D>(bq.R, bq.scale, bq.L, bq.P, false, gSig, true, nullptr, FULL_M - 3, dR, dS,
                      v_ltri + n * FULL_M);
    if constexpr (QUAT) {
        float qv[4], vq[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) qv[k] = quats[n2 * 4 + k];
        quat_rot_vjp(qv, dR, vq);
#pragma unroll
        for (int k = 0; k < 4; ++k) v_quats[n * 4 + k] = vq[k];
    } else {
        // l_triagnle_to_rotmat_bwd.cu:16-25
        v_ltri[n * FULL_M + 0] = dR[1] - dR[3];
        v_ltri[n * FULL_M + 1] = dR[2] - dR[6];
        v_ltri[n * FULL_M + 2] = dR[5] - dR[7];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) v_scale_raw[n * D + k] = dS[k] * softplus_grad(scale_raw[n2 * D + k]);
}

}  // namespace ubs

namespace ubs {

// Camera centres [C,3] of the view matrices, as the fused projection takes
// them when it is given none.
__global__ void camera_centers_kernel(uint32_t C, const float *__restrict__ viewmats, float *__restrict__ out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    float cc[3];
    camera_centre(viewmats + (size_t)c * 16, cc);
    out[3 * c] = cc[0];
    out[3 * c + 1] = cc[1];
    out[3 * c + 2] = cc[2];
}

}  // namespace ubs

using namespace ubs;

extern "C" {

static int beta_project_fwd_impl(int64_t N, int64_t D, uint32_t C, const float *xyz, const float *mean_rest,
                                 const float *scale_raw, const float *l_triangle, const float *opacity_raw,
                                 const float *beta_raw, const float *quats, const float *cam_pos,
                                 const float *timestamps, const float *viewmats, const float *Ks, int32_t width,
                                 int32_t height, const int32_t *image_wh, float eps2d, float near_plane,
                                 float far_plane, float radius_clip, int ortho, int32_t *radii, float *means2d,
                                 float *depths, float *conics, float *opacities, float *betas, int64_t row_stride,
                                 void *stream, const float *rgb = nullptr, void *records = nullptr);

int ubs_beta_project_fwd(int64_t N, int64_t D, uint32_t C, const float *xyz, const float *mean_rest,
                         const float *scale_raw, const float *l_triangle, const float *opacity_raw,
                         const float *beta_raw, const float *quats, const float *cam_pos, const float *timestamps,
                         const float *viewmats, const float *Ks, int32_t width, int32_t height,
                         const int32_t *image_wh, float eps2d, float near_plane, float far_plane, float radius_clip,
                         int ortho, int32_t *radii, float *means2d, float *depths, float *conics, float *opacities,
                         float *betas, void *stream) {
    return beta_project_fwd_impl(N, D, C, xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats,
                                 cam_pos, timestamps, viewmats, Ks, width, height, image_wh, eps2d, near_plane,
                                 far_plane, radius_clip, ortho, radii, means2d, depths, conics, opacities, betas, N,
                                 stream);
}

int ubs_beta_project_fwd_rec(int64_t N, int64_t D, uint32_t C, const float *xyz, const float *mean_rest,
                             const float *scale_raw, const float *l_triangle, const float *opacity_raw,
                             const float *beta_raw, const float *quats, const float *cam_pos, const float *timestamps,
                             const float *viewmats, const float *Ks, int32_t width, int32_t height,
                             const int32_t *image_wh, float eps2d, float near_plane, float far_plane,
                             float radius_clip, int ortho, int32_t *radii, float *means2d, float *depths,
                             float *conics, float *opacities, float *betas, const float *rgb, void *records,
                             void *stream) {
    UBS_REQUIRE(records && rgb, "null pointer");
    return beta_project_fwd_impl(N, D, C, xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats,
                                 cam_pos, timestamps, viewmats, Ks, width, height, image_wh, eps2d, near_plane,
                                 far_plane, radius_clip, ortho, radii, means2d, depths, conics, opacities, betas, N,
                                 stream, rgb, records);
}

int ubs_beta_project_fwd_rows(int64_t N, int64_t D, uint32_t C, const float *xyz, const float *mean_rest,
                              const float *scale_raw, const float *l_triangle, const float *opacity_raw,
                              const float *beta_raw, const float *quats, const float *cam_pos,
                              const float *timestamps, const float *viewmats, const float *Ks, int32_t width,
                              int32_t height, const int32_t *image_wh, float eps2d, float near_plane,
                              float far_plane, float radius_clip, int ortho, int32_t *radii, float *means2d,
                              float *depths, float *conics, float *opacities, float *betas, int64_t row_stride,
                              void *stream) {
    UBS_REQUIRE(row_stride >= N, "row_stride %lld < N %lld", (long long)row_stride, (long long)N);
    return beta_project_fwd_impl(N, D, C, xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats,
                                 cam_pos, timestamps, viewmats, Ks, width, height, image_wh, eps2d, near_plane,
                                 far_plane, radius_clip, ortho, radii, means2d, depths, conics, opacities, betas,
                                 row_stride, stream);
}

static int beta_project_fwd_impl(int64_t N, int64_t D, uint32_t C, const float *xyz, const float *mean_rest,
                                 const float *scale_raw, const float *l_triangle, const float *opacity_raw,
                                 const float *beta_raw, const float *quats, const float *cam_pos,
                                 const float *timestamps, const float *viewmats, const float *Ks, int32_t width,
                                 int32_t height, const int32_t *image_wh, float eps2d, float near_plane,
                                 float far_plane, float radius_clip, int ortho, int32_t *radii, float *means2d,
                                 float *depths, float *conics, float *opacities, float *betas, int64_t row_stride,
                                 void *stream, const float *rgb, void *records) {
    UBS_REQUIRE(D == 6 || D == 7, "the fused Beta projection is defined for D = 6 or 7 (got %lld)", (long long)D);
    UBS_REQUIRE(!records || (rgb && row_stride == N), "records need rgb and row_stride == N");
    UBS_REQUIRE(N >= 0 && (uint64_t)row_stride * C < (1ull << 31), "C*N must be < 2^31");
    if (N == 0 || C == 0) return UBS_OK;
    UBS_REQUIRE(xyz && mean_rest && scale_raw && l_triangle && opacity_raw && beta_raw && viewmats && Ks && radii &&
                    means2d && depths && conics && opacities && betas,
                "null pointer");
    hipStream_t st = as_stream(stream);
    const unsigned grid = grid_for(N);
#define UBS_BP_FWD(DD, QQ)                                                                                       \
    beta_project_fwd_kernel<DD, QQ><<<grid, kThreads, 0, st>>>(                                                  \
        N, C, xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats, cam_pos, timestamps, viewmats, \
        Ks, width, height, image_wh, eps2d, near_plane, far_plane, radius_clip, ortho, radii, means2d, depths,    \
        conics, opacities, betas, row_stride, rgb, static_cast<float4 *>(records))
    if (D == 6) {
        if (quats) UBS_BP_FWD(6, true);
        else UBS_BP_FWD(6, false);
    } else {
        if (quats) UBS_BP_FWD(7, true);
        else UBS_BP_FWD(7, false);
    }
#undef UBS_BP_FWD
    UBS_LAUNCH_CHECK("beta_project_fwd_kernel");
    return UBS_OK;
}

int ubs_camera_centers(uint32_t C, const float *viewmats, float *centres, void *stream) {
    if (C == 0) return UBS_OK;
    UBS_REQUIRE(viewmats && centres, "null pointer");
    camera_centers_kernel<<<(C + 63) / 64, 64, 0, as_stream(stream)>>>(C, viewmats, centres);
    UBS_LAUNCH_CHECK("camera_centers_kernel");
    return UBS_OK;
}

int ubs_beta_project_fwd_table(uint32_t C, int64_t row_stride, const ubs_view_scene *table, const float *cam_pos,
                               const float *timestamps, const float *viewmats, const float *Ks,
                               const int32_t *image_wh, float eps2d, float near_plane, float far_plane,
                               float radius_clip, int32_t *radii, float *means2d, float *depths, float *conics,
                               float *opacities, float *betas, float *colors, void *stream) {
    UBS_REQUIRE(row_stride >= 0 && (uint64_t)row_stride * C < (1ull << 31), "C*row_stride must be < 2^31");
    if (row_stride == 0 || C == 0) return UBS_OK;
    UBS_REQUIRE(table && viewmats && Ks && image_wh && radii && means2d && depths && conics && opacities && betas &&
                    colors,
                "null pointer");
    hipStream_t st = as_stream(stream);
    const unsigned grid = grid_for(row_stride);
    beta_project_fwd_table_kernel<6><<<grid, kThreads, 0, st>>>(
        C, row_stride, table, cam_pos, timestamps, viewmats, Ks, image_wh, eps2d, near_plane, far_plane, radius_clip,
        radii, means2d, depths, conics, opacities, betas, colors);
    UBS_LAUNCH_CHECK("beta_project_fwd_table_kernel<6>");
    beta_project_fwd_table_kernel<7><<<grid, kThreads, 0, st>>>(
        C, row_stride, table, cam_pos, timestamps, viewmats, Ks, image_wh, eps2d, near_plane, far_plane, radius_clip,
        radii, means2d, depths, conics, opacities, betas, colors);
    UBS_LAUNCH_CHECK("beta_project_fwd_table_kernel<7>");
    return UBS_OK;
}

int ubs_beta_project_bwd(int64_t N, int64_t D, uint32_t C, const float *xyz, const float *mean_rest,
                         const float *scale_raw, const float *l_triangle, const float *opacity_raw,
                         const float *beta_raw, const float *quats, const float *cam_pos, const float *timestamps,
                         const float *viewmats, const float *Ks, int32_t width, int32_t height,
                         const int32_t *image_wh, float eps2d, int ortho, const int32_t *radii, const float *conics,
                         const float *v_means2d, const float *v_depths, const float *v_conics,
                         const float *v_opacities, const float *v_betas, float *v_xyz, float *v_mean_rest,
                         float *v_scale_raw, float *v_l_triangle, float *v_opacity_raw, float *v_beta_raw,
                         float *v_quats, void *stream) {
    UBS_REQUIRE(D == 6 || D == 7, "the fused Beta projection is defined for D = 6 or 7 (got %lld)", (long long)D);
    UBS_REQUIRE(N >= 0 && (uint64_t)N * C < (1ull << 31), "C*N must be < 2^31");
    UBS_REQUIRE(!quats || v_quats, "quats given without v_quats");
    if (N == 0) return UBS_OK;
    UBS_REQUIRE(xyz && mean_rest && scale_raw && l_triangle && opacity_raw && beta_raw && v_xyz && v_mean_rest &&
                    v_scale_raw && v_l_triangle && v_opacity_raw && v_beta_raw,
                "null pointer");
    UBS_REQUIRE(C == 0 || (viewmats && Ks && radii && conics && v_means2d && v_conics && v_opacities && v_betas),
                "null pointer");
    hipStream_t st = as_stream(stream);
    const unsigned grid = grid_for(N);
    const int exact = cond_backward_exact();
#define UBS_BP_BWD(DD, QQ)                                                                                       \
    beta_project_bwd_kernel<DD, QQ><<<grid, kThreads, 0, st>>>(                                                  \
        N, C, xyz, mean_rest, scale_raw, l_triangle, opacity_raw, beta_raw, quats, cam_pos, timestamps, viewmats, \
        Ks, width, height, image_wh, eps2d, ortho, exact, radii, conics, v_means2d, v_depths, v_conics,           \
        v_opacities, v_betas, v_xyz, v_mean_rest, v_scale_raw, v_l_triangle, v_opacity_raw, v_beta_raw, v_quats)
    if (D == 6) {
        if (quats) UBS_BP_BWD(6, true);
        else UBS_BP_BWD(6, false);
    } else {
        if (quats) UBS_BP_BWD(7, true);
        else UBS_BP_BWD(7, false);
    }
#undef UBS_BP_BWD
    UBS_LAUNCH_CHECK("beta_project_bwd_kernel");
    return UBS_OK;
}

}  // extern "C"

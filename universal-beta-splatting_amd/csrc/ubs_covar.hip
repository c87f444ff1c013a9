// Covariance construction of Beta primitives on gfx950:
//   l_triangle_to_rotmat  (R = I + A, first-order skew rotation)
//   rot_scale_l_triangle_to_covar  (Sigma = L L^T)
// One thread per primitive; D is a template parameter so L stays in VGPRs.
// Summation order is fixed (documented beside each sum) and FP contraction is
// off, so the forward is bit-identical to the float oracle.
#pragma clang fp contract(off)

#include "ubs_covar_math.h"

namespace ubs {

// reference: l_triagnle_to_rotmat_fwd.cu:16-35
__global__ void rotmat_fwd_kernel(int64_t N, const float *__restrict__ l, float *__restrict__ R) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float a0 = l[i * 3 + 0], a1 = l[i * 3 + 1], a2 = l[i * 3 + 2];
    float *r = R + i * 9;
    r[0] = 1.f; r[1] = a0;  r[2] = a1;
    r[3] = -a0; r[4] = 1.f; r[5] = a2;
    r[6] = -a1; r[7] = -a2; r[8] = 1.f;
}

// reference: l_triagnle_to_rotmat_bwd.cu:16-25
__global__ void rotmat_bwd_kernel(int64_t N, const float *__restrict__ g, float *__restrict__ o) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float *G = g + i * 9;
    o[i * 3 + 0] = G[1] - G[3];
    o[i * 3 + 1] = G[2] - G[6];
    o[i * 3 + 2] = G[5] - G[7];
}

template <int D>
__global__ __launch_bounds__(kThreads) void covar_fwd_kernel(int64_t N, const float *__restrict__ rot, const float *__restrict__ scale,
                                 const float *__restrict__ ltri, const int32_t *__restrict__ rest_i,
                                 const int32_t *__restrict__ rest_j, int M, int spatial,
                                 float *__restrict__ out) {
    __shared__ RestLds rl;
    if (threadIdx.x < kMaxRest) {
        const bool ok = (int)threadIdx.x < M;
        rl.r[threadIdx.x] = ok ? (int8_t)rest_i[threadIdx.x] : (int8_t)-1;
        rl.c[threadIdx.x] = ok ? (int8_t)rest_j[threadIdx.x] : (int8_t)-1;
    }
    __syncthreads();
    const bool canon = rest_is_canonical<D>(rl, M);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    constexpr int FULL_M = D * (D - 1) / 2;
    float L[D][D];
    uint32_t P[D];
    load_L<D>(rot + i * 9, scale + i * D, ltri + i * FULL_M + 3, &rl, M, canon, L, P);
    const bool sp = spatial || D == 3;
    covar_fwd_body<D>(L, P, sp, out + i * (sp ? 9 : D * D));
}

template <int D>
__global__ __launch_bounds__(kThreads) void covar_bwd_kernel(int64_t N, const float *__restrict__ rot, const float *__restrict__ scale,
                                 const float *__restrict__ ltri, const int32_t *__restrict__ rest_i,
                                 const int32_t *__restrict__ rest_j, int M, int spatial,
                                 const float *__restrict__ vcov, float *__restrict__ vrot,
                                 float *__restrict__ vscale, float *__restrict__ vltri) {
    __shared__ RestLds rl;
    if (threadIdx.x < kMaxRest) {
        const bool ok = (int)threadIdx.x < M;
        rl.r[threadIdx.x] = ok ? (int8_t)rest_i[threadIdx.x] : (int8_t)-1;
        rl.c[threadIdx.x] = ok ? (int8_t)rest_j[threadIdx.x] : (int8_t)-1;
    }
    __syncthreads();
    const bool canon = rest_is_canonical<D>(rl, M);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    constexpr int FULL_M = D * (D - 1) / 2;
    const bool sp = spatial || D == 3;
    const int dim = sp ? 3 : D;
    const float *R = rot + i * 9;
    const float *s = scale + i * D;
    float L[D][D];
    uint32_t P[D];
    load_L<D>(R, s, ltri + i * FULL_M + 3, &rl, M, canon, L, P);
    covar_bwd_body<D>(R, s, L, P, sp, vcov + i * dim * dim, canon, &rl, M, vrot + i * 9, vscale + i * D,
                      vltri ? vltri + i * FULL_M : nullptr);
}

template <int D>
static void launch_covar_fwd(int64_t N, const float *rot, const float *scale, const float *ltri,
                             const int32_t *ri, const int32_t *rj, int M, int spatial, float *out,
                             hipStream_t st) {
    covar_fwd_kernel<D><<<grid_for(N), kThreads, 0, st>>>(N, rot, scale, ltri, ri, rj, M, spatial, out);
}

template <int D>
static void launch_covar_bwd(int64_t N, const float *rot, const float *scale, const float *ltri,
                             const int32_t *ri, const int32_t *rj, int M, int spatial, const float *g,
                             float *vr, float *vs, float *vl, hipStream_t st) {
    covar_bwd_kernel<D><<<grid_for(N), kThreads, 0, st>>>(N, rot, scale, ltri, ri, rj, M, spatial, g, vr, vs, vl);
}

}  // namespace ubs

using namespace ubs;

#define UBS_DISPATCH_D(D, FN, ...)                                             \
    switch (D) {                                                               \
    case 3: FN<3>(__VA_ARGS__); break;                                         \
    case 4: FN<4>(__VA_ARGS__); break;                                         \
    case 5: FN<5>(__VA_ARGS__); break;                                         \
    case 6: FN<6>(__VA_ARGS__); break;                                         \
    case 7: FN<7>(__VA_ARGS__); break;                                         \
    case 8: FN<8>(__VA_ARGS__); break;                                         \
    case 9: FN<9>(__VA_ARGS__); break;                                         \
    case 10: FN<10>(__VA_ARGS__); break;                                       \
    case 11: FN<11>(__VA_ARGS__); break;                                       \
    default: break;                                                            \
    }

extern "C" {

int ubs_l_triangle_to_rotmat_fwd(int64_t N, const float *l, float *R, void *stream) {
    UBS_REQUIRE(N >= 0, "N must be >= 0");
    if (N == 0) return UBS_OK;
    UBS_REQUIRE(l && R, "null pointer");
    rotmat_fwd_kernel<<<grid_for(N), kThreads, 0, as_stream(stream)>>>(N, l, R);
    UBS_LAUNCH_CHECK("rotmat_fwd_kernel");
    return UBS_OK;
}

int ubs_l_triangle_to_rotmat_bwd(int64_t N, const float *g, float *o, void *stream) {
    UBS_REQUIRE(N >= 0, "N must be >= 0");
    if (N == 0) return UBS_OK;
    UBS_REQUIRE(g && o, "null pointer");
    rotmat_bwd_kernel<<<grid_for(N), kThreads, 0, as_stream(stream)>>>(N, g, o);
    UBS_LAUNCH_CHECK("rotmat_bwd_kernel");
    return UBS_OK;
}

int ubs_rot_scale_l_triangle_to_covar_fwd(int64_t N, int64_t D, const float *rot, const float *scale,
                                          const float *ltri, const int32_t *ri, const int32_t *rj,
                                          int64_t M, int spatial, float *covar, void *stream) {
    UBS_REQUIRE(D >= 3 && D <= 11, "D must be in [3, 11], got %lld", (long long)D);
    UBS_REQUIRE(M >= 0 && M <= D * (D - 1) / 2 - 3 + (D == 3 ? 3 : 0) && M <= kMaxRest,
                "M_rest=%lld does not fit D=%lld", (long long)M, (long long)D);
    UBS_REQUIRE(N >= 0, "N must be >= 0");
    if (N == 0) return UBS_OK;
    UBS_REQUIRE(rot && scale && covar && (ltri || D == 3) && (M == 0 || (ri && rj)), "null pointer");
    UBS_DISPATCH_D(D, launch_covar_fwd, N, rot, scale, ltri, ri, rj, (int)M, spatial, covar, as_stream(stream));
    UBS_LAUNCH_CHECK("covar_fwd_kernel");
    return UBS_OK;
}

int ubs_rot_scale_l_triangle_to_covar_bwd(int64_t N, int64_t D, const float *rot, const float *scale,
                                          const float *ltri, const int32_t *ri, const int32_t *rj,
                                          int64_t M, int spatial, const float *g, float *vr,
                                          float *vs, float *vl, void *stream) {
    UBS_REQUIRE(D >= 3 && D <= 11, "D must be in [3, 11], got %lld", (long long)D);
    UBS_REQUIRE(M >= 0 && M <= kMaxRest && (D == 3 || M <= D * (D - 1) / 2 - 3),
                "M_rest=%lld does not fit D=%lld", (long long)M, (long long)D);
    UBS_REQUIRE(N >= 0, "N must be >= 0");
    if (N == 0) return UBS_OK;
    UBS_REQUIRE(rot && scale && g && vr && vs && (vl || D == 3) && (ltri || D == 3) && (M == 0 || (ri && rj)),
                "null pointer");
    UBS_DISPATCH_D(D, launch_covar_bwd, N, rot, scale, ltri, ri, rj, (int)M, spatial, g, vr, vs, vl,
                   as_stream(stream));
    UBS_LAUNCH_CHECK("covar_bwd_kernel");
    return UBS_OK;
}

}  // extern "C"

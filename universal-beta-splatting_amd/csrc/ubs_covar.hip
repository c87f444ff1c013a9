// Covariance construction of Beta primitives on gfx950:
//   l_triangle_to_rotmat  (R = I + A, first-order skew rotation)
//   rot_scale_l_triangle_to_covar  (Sigma = L L^T)
// One thread per primitive; D is a template parameter so L stays in VGPRs.
// Summation order is fixed (documented beside each sum) and FP contraction is
// off, so the forward is bit-identical to the float oracle.
#pragma clang fp contract(off)

#include "ubs_common.h"

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

constexpr int kMaxRest = 64;

// Rest pattern staged in LDS: row of each entry, column, validity.
struct RestLds {
    int8_t r[kMaxRest];
    int8_t c[kMaxRest];
};

// True (for the whole block) when the rest pattern is the canonical one of
// scene/beta_model.py:69-73: rows 3..D-1 list columns 0..r-1 in order.
template <int D>
__device__ __forceinline__ bool rest_is_canonical(const RestLds &rl, int M) {
    bool ok = (M == D * (D - 1) / 2 - 3);
    if (ok && (int)threadIdx.x < M) {
        int m = 0, r_want = 3, c_want = 0;
        // position -> (row, col) of the canonical enumeration
        for (int r = 3; r < D; ++r)
            for (int c = 0; c < r; ++c, ++m)
                if (m == (int)threadIdx.x) {
                    r_want = r;
                    c_want = c;
                }
        ok = rl.r[threadIdx.x] == r_want && rl.c[threadIdx.x] == c_want;
    }
    return __syncthreads_and(ok) != 0;
}

template <int D>
__device__ __forceinline__ void load_L(const float *__restrict__ R, const float *__restrict__ s,
                                       const float *__restrict__ x_rest, const RestLds &rl, int M, bool canon,
                                       float (&L)[D][D], uint32_t (&present)[D]) {
#pragma unroll
    for (int r = 0; r < D; ++r) {
        present[r] = 0u;
#pragma unroll
        for (int c = 0; c < D; ++c) L[r][c] = 0.f;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            L[r][c] = R[r * 3 + c] * s[c];
            present[r] |= 1u << c;
        }
#pragma unroll
    for (int r = 3; r < D; ++r) {
        L[r][r] = s[r];
        present[r] |= 1u << r;
    }
    if (canon) {
        // rows 3..D-1 hold all their strictly-lower entries, in order
        int m = 0;
#pragma unroll
        for (int r = 3; r < D; ++r)
#pragma unroll
            for (int c = 0; c < r; ++c, ++m) {
                L[r][c] = x_rest[m];
                present[r] |= 1u << c;
            }
        return;
    }
    for (int m = 0; m < M; ++m) {
        const int r = rl.r[m], c = rl.c[m];
        if (r < 3 || r >= D || c >= r) continue;
        const float v = x_rest[m];
        // scatter into the register tile with a compile-time index sweep
#pragma unroll
        for (int rr = 3; rr < D; ++rr)
#pragma unroll
            for (int cc = 0; cc < rr; ++cc)
                if (rr == r && cc == c) {
                    L[rr][cc] = v;
                    present[rr] |= 1u << cc;
                }
    }
}

// reference: rot_scale_l_triangle_to_covar_fwd.cu:28-187 (same summation
// order: 3x3 block ((a+b)+c); cross block from 0 over row-c entries with
// k<3; diagonal s_r^2 then the row entries; off-diagonal from 0 over k<c
// then + s_c x_rc).
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
    load_L<D>(rot + i * 9, scale + i * D, ltri + i * FULL_M + 3, rl, M, canon, L, P);
    if (spatial || D == 3) {
        float *o = out + i * 9;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = r; c < 3; ++c) {
                const float v = (L[r][0] * L[c][0] + L[r][1] * L[c][1]) + L[r][2] * L[c][2];
                o[r * 3 + c] = v;
                o[c * 3 + r] = v;
            }
        return;
    }
    float *o = out + i * D * D;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = r; c < 3; ++c) {
            const float v = (L[r][0] * L[c][0] + L[r][1] * L[c][1]) + L[r][2] * L[c][2];
            o[r * D + c] = v;
            o[c * D + r] = v;
        }
#pragma unroll
    for (int c = 3; c < D; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (P[c] & (1u << k)) acc += L[r][k] * L[c][k];
            o[r * D + c] = acc;
            o[c * D + r] = acc;
        }
#pragma unroll
    for (int r = 3; r < D; ++r) {
        float acc = L[r][r] * L[r][r];
#pragma unroll
        for (int k = 0; k < r; ++k)
            if (P[r] & (1u << k)) acc += L[r][k] * L[r][k];
        o[r * D + r] = acc;
    }
#pragma unroll
    for (int r = 4; r < D; ++r)
#pragma unroll
        for (int c = 3; c < r; ++c) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < c; ++k)
                if (P[c] & (1u << k)) acc += ((P[r] & (1u << k)) ? L[r][k] : 0.f) * L[c][k];
            acc += L[c][c] * ((P[r] & (1u << c)) ? L[r][c] : 0.f);
            o[r * D + c] = acc;
            o[c * D + r] = acc;
        }
}

// reference: rot_scale_l_triangle_to_covar_bwd.cu:48-234 -- dL = (G + G^T) L
// restricted to the parameterised entries of L.
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
    load_L<D>(R, s, ltri + i * FULL_M + 3, rl, M, canon, L, P);
    const float *G = vcov + i * dim * dim;
    auto S = [&](int a, int b) { return G[a * dim + b] + G[b * dim + a]; };
    float *dR = vrot + i * 9, *dS = vscale + i * D, *dL = vltri + i * FULL_M;
    float H[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) H[r][c] = (S(r, 0) * L[0][c] + S(r, 1) * L[1][c]) + S(r, 2) * L[2][c];
    if (!sp) {
#pragma unroll
        for (int c = 3; c < D; ++c)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (!(P[c] & (1u << k))) continue;
#pragma unroll
                for (int r = 0; r < 3; ++r) H[r][k] += S(r, c) * L[c][k];
            }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int r = 0; r < 3; ++r) dR[r * 3 + c] = H[r][c] * s[c];
        dS[c] = (H[0][c] * R[0 * 3 + c] + H[1][c] * R[1 * 3 + c]) + H[2][c] * R[2 * 3 + c];
    }
    if (sp) {
#pragma unroll
        for (int k = 3; k < D; ++k) dS[k] = 0.f;
        for (int k = 0; k < FULL_M; ++k) dL[k] = 0.f;
        return;
    }
    dL[0] = 0.f;
    dL[1] = 0.f;
    dL[2] = 0.f;
    if (canon) {
        // same sums as the generic loop below, with compile-time indices
        int m = 0;
#pragma unroll
        for (int r = 3; r < D; ++r)
#pragma unroll
            for (int c = 0; c < r; ++c, ++m) {
                float g = 0.f;
                if (c < 3) {
                    g += S(r, 0) * L[0][c];
                    g += S(r, 1) * L[1][c];
                    g += S(r, 2) * L[2][c];
#pragma unroll
                    for (int j = 3; j < D; ++j)
                        if (j > c) g += S(r, j) * L[j][c];
                } else {
                    g += S(r, c) * L[c][c];
#pragma unroll
                    for (int j = c + 1; j < D; ++j) g += S(r, j) * L[j][c];
                }
                dL[3 + m] = g;
            }
    }
    for (int m = 0; m < (canon ? 0 : M); ++m) {
        const int r = rl.r[m], c = rl.c[m];
        float g = 0.f;
        if (r >= 3 && r < D && c < r) {
#pragma unroll
            for (int rr = 3; rr < D; ++rr) {
                if (rr != r) continue;
                if (c < 3) {
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) {
                        if (cc != c) continue;
                        g += S(rr, 0) * L[0][cc];
                        g += S(rr, 1) * L[1][cc];
                        g += S(rr, 2) * L[2][cc];
#pragma unroll
                        for (int j = 3; j < D; ++j)
                            if (P[j] & (1u << cc)) g += S(rr, j) * L[j][cc];
                    }
                } else {
#pragma unroll
                    for (int cc = 3; cc < D; ++cc) {
                        if (cc != c) continue;
                        g += S(rr, cc) * L[cc][cc];
#pragma unroll
                        for (int j = cc + 1; j < D; ++j)
                            if (P[j] & (1u << cc)) g += S(rr, j) * L[j][cc];
                    }
                }
            }
        }
        dL[3 + m] = g;
    }
    for (int m = 3 + M; m < FULL_M; ++m) dL[m] = 0.f;
#pragma unroll
    for (int r = 3; r < D; ++r) {
        float ds = S(r, r) * L[r][r];
#pragma unroll
        for (int j = r + 1; j < D; ++j)
            if (P[j] & (1u << r)) ds += S(r, j) * L[j][r];
        dS[r] = ds;
    }
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

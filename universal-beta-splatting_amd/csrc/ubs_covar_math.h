// Per-primitive covariance math of a D-dim Beta primitive, shared by the
// covariance kernels and the fused Beta projection:
//   L  = [R diag(s_0..2) | 0; rest entries | s_3..s_{D-1} on the diagonal]
//   Sigma = L L^T, and its VJP dL = (G + G^T) L restricted to the
//   parameterised entries of L.
// The functions take plain pointers so that they serve both a row of a global
// tensor and a register array of the caller (all indices are compile-time
// after inlining, so local arrays stay in VGPRs).  Summation orders are fixed
// and FP contraction is off (bit-identical to the float oracle).
#pragma once
#pragma clang fp contract(off)

#include "ubs_common.h"

namespace ubs {

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

// L from the rotation R (row-major 3x3), the activated scales s[D] and the
// rest entries x_rest (l_triangle without its first three entries).
// present[r] has bit c set for every parameterised entry L[r][c].
template <int D>
__device__ __forceinline__ void load_L(const float *__restrict__ R, const float *__restrict__ s,
                                       const float *__restrict__ x_rest, const RestLds *rl, int M, bool canon,
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
        const int r = rl->r[m], c = rl->c[m];
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

// Sigma = L L^T into o: the 3x3 spatial block (row-major, 9 floats) when sp,
// else the full D x D matrix.
// reference: rot_scale_l_triangle_to_covar_fwd.cu:28-187 (same summation
// order: 3x3 block ((a+b)+c); cross block from 0 over row-c entries with
// k<3; diagonal s_r^2 then the row entries; off-diagonal from 0 over k<c
// then + s_c x_rc).
template <int D>
__device__ __forceinline__ void covar_fwd_body(const float (&L)[D][D], const uint32_t (&P)[D], bool sp,
                                               float *__restrict__ o) {
    if (sp || D == 3) {
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

// VJP of covar_fwd_body: G is the gradient of Sigma (3x3 when sp, else
// D x D, row-major); writes dR[9], dS[D] and dL[D(D-1)/2] (whose first three
// entries, the rotation parameters, get 0; dL may be null when sp).
// reference: rot_scale_l_triangle_to_covar_bwd.cu:48-234
template <int D>
__device__ __forceinline__ void covar_bwd_body(const float *__restrict__ R, const float *__restrict__ s,
                                               const float (&L)[D][D], const uint32_t (&P)[D], bool sp,
                                               const float *__restrict__ G, bool canon, const RestLds *rl, int M,
                                               float *__restrict__ dR, float *__restrict__ dS,
                                               float *__restrict__ dL) {
    constexpr int FULL_M = D * (D - 1) / 2;
    sp = sp || D == 3;
    const int dim = sp ? 3 : D;
    auto S = [&](int a, int b) { return G[a * dim + b] + G[b * dim + a]; };
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
        if (dL)
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
        const int r = rl->r[m], c = rl->c[m];
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
    if (!canon)
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

}  // namespace ubs

// Per-primitive conditioning math of a D-dim Beta primitive on its C = D-3
// query dimensions, shared by the conditioning kernels and the fused Beta
// projection.  Every small matrix (v12, v22^-1, Cholesky factor, ...) lives
// in VGPRs with compile-time indices, including the pivoting Gauss-Jordan
// inverse used for C != 3.  The functions take plain pointers so that they
// serve both rows of global tensors and register arrays of the caller.
// m_cond and V_cond are pure +,-,*,/ in the oracle's operation order with
// contraction off (bit-identical); the opacity factor uses tanh/pow.
#pragma once
#pragma clang fp contract(off)

#include "ubs_common.h"

namespace ubs {

// Process-wide choice of the opacity-path Cholesky backward (set by
// ubs_set_cond_backward_exact): 0 = the reference's solve, 1 = exact.
int cond_backward_exact();

template <int C> struct CondRegs {
    float x[C], beta[C], badj[C];
    float v11[9], v12[3][C], v21[C][3], v22[C][C], i22[C][C];
    float r_[3][C], rb[3][C];
    float Lc[C][C], y[C];
};

// reference: cond_mean_convariance_opacity_fwd.cu:75-104 (cofactor inverse)
__device__ __forceinline__ void inv3(const float (&a)[3][3], float (&o)[3][3]) {
    const float c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    const float c01 = -(a[1][0] * a[2][2] - a[1][2] * a[2][0]);
    const float c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    const float c10 = -(a[0][1] * a[2][2] - a[0][2] * a[2][1]);
    const float c11 = a[0][0] * a[2][2] - a[0][2] * a[2][0];
    const float c12 = -(a[0][0] * a[2][1] - a[0][1] * a[2][0]);
    const float c20 = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    const float c21 = -(a[0][0] * a[1][2] - a[0][2] * a[1][0]);
    const float c22 = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    float det = (a[0][0] * c00 + a[0][1] * c01) + a[0][2] * c02;
    if (det == 0.f) det = 1e-20f;
    const float id = 1.f / det;
    o[0][0] = c00 * id; o[0][1] = c10 * id; o[0][2] = c20 * id;
    o[1][0] = c01 * id; o[1][1] = c11 * id; o[1][2] = c21 * id;
    o[2][0] = c02 * id; o[2][1] = c12 * id; o[2][2] = c22 * id;
}

// reference: cond_mean_convariance_opacity_fwd.cu:26-73 (Gauss-Jordan with
// partial pivoting, zero pivot -> 1e-20).  Row swaps are predicated moves so
// the augmented matrix never leaves registers.
template <int C>
__device__ __forceinline__ void inv_gj(const float (&A)[C][C], float (&inv)[C][C]) {
    float aug[C][2 * C];
#pragma unroll
    for (int r = 0; r < C; ++r)
#pragma unroll
        for (int c = 0; c < 2 * C; ++c) aug[r][c] = (c < C) ? A[r][c] : (c - C == r ? 1.f : 0.f);
#pragma unroll
    for (int col = 0; col < C; ++col) {
        int piv = col;
        float maxabs = fabsf(aug[col][col]);
#pragma unroll
        for (int r = col + 1; r < C; ++r) {
            const float v = fabsf(aug[r][col]);
            if (v > maxabs) {
                maxabs = v;
                piv = r;
            }
        }
#pragma unroll
        for (int r = col + 1; r < C; ++r) {
            if (piv == r) {
#pragma unroll
                for (int c = 0; c < 2 * C; ++c) {
                    const float t = aug[col][c];
                    aug[col][c] = aug[r][c];
                    aug[r][c] = t;
                }
            }
        }
        float diag = aug[col][col];
        if (diag == 0.f) diag = 1e-20f;
        const float invd = 1.f / diag;
#pragma unroll
        for (int c = 0; c < 2 * C; ++c) aug[col][c] *= invd;
#pragma unroll
        for (int r = 0; r < C; ++r) {
            if (r == col) continue;
            const float f = aug[r][col];
            if (f != 0.f) {
#pragma unroll
                for (int c = 0; c < 2 * C; ++c) aug[r][c] -= f * aug[col][c];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < C; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) inv[r][c] = aug[r][C + c];
}

template <int C>
__device__ __forceinline__ void cond_setup(const float *__restrict__ m, const float *__restrict__ V,
                                           const float *__restrict__ b, const float *__restrict__ q,
                                           CondRegs<C> &st) {
    constexpr int D = C + 3;
#pragma unroll
    for (int j = 0; j < C; ++j) {
        st.x[j] = q[j] - m[3 + j];
        st.beta[j] = b[j];
        const float t = b[j] * 0.25f;
        st.badj[j] = (t < 1.f) ? t : 1.f;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) st.v11[r * 3 + c] = V[r * D + c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) st.v12[r][c] = V[r * D + 3 + c];
#pragma unroll
    for (int r = 0; r < C; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) st.v21[r][c] = V[(3 + r) * D + c];
#pragma unroll
    for (int r = 0; r < C; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) st.v22[r][c] = V[(3 + r) * D + 3 + c];
    if constexpr (C == 3) {
        inv3(st.v22, st.i22);
    } else {
        inv_gj<C>(st.v22, st.i22);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < C; ++k) acc += st.v12[r][k] * st.i22[k][c];
            st.r_[r][c] = acc;
            st.rb[r][c] = acc * st.badj[c];
        }
    // Cholesky of v22 (guard: non-positive pivot -> 1e-20, zero divisor -> 1e-20)
#pragma unroll
    for (int r = 0; r < C; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) st.Lc[r][c] = 0.f;
#pragma unroll
    for (int i = 0; i < C; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            float sum = st.v22[i][j];
#pragma unroll
            for (int k = 0; k < j; ++k) sum -= st.Lc[i][k] * st.Lc[j][k];
            if (i == j) {
                if (sum <= 0.f) sum = 1e-20f;
                st.Lc[i][j] = sqrtf(sum);
            } else {
                float den = st.Lc[j][j];
                if (den == 0.f) den = 1e-20f;
                st.Lc[i][j] = sum / den;
            }
        }
#pragma unroll
    for (int i = 0; i < C; ++i) {
        float sum = st.x[i];
#pragma unroll
        for (int k = 0; k < i; ++k) sum -= st.Lc[i][k] * st.y[k];
        float den = st.Lc[i][i];
        if (den == 0.f) den = 1e-20f;
        st.y[i] = sum / den;
    }
}

constexpr float kOneMinusEps = 1.0f - 1.1920928955078125e-07f;  // 1 - FLT_EPSILON

// m_cond [3], V_cond [3x3] and the conditioned opacity from the set-up st.
// reference: cond_mean_convariance_opacity_fwd.cu:106-290
template <int C>
__device__ __forceinline__ void cond_fwd_body(const CondRegs<C> &st, const float *__restrict__ m, float opac,
                                              float *__restrict__ om, float *__restrict__ oV,
                                              float *__restrict__ oo) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        float acc = 0.f;
#pragma unroll
        for (int c = 0; c < C; ++c) acc += st.rb[r][c] * st.x[c];
        om[r] = m[r] + acc;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < C; ++k) acc += st.rb[r][k] * st.v21[k][s];
            oV[r * 3 + s] = st.v11[r * 3 + s] - acc;
        }
    float oc = 1.f;
#pragma unroll
    for (int k = 0; k < C; ++k) {
        float d = tanhf(st.y[k] * st.y[k]);
        if (d < 0.f) d = 0.f;
        if (d > kOneMinusEps) d = kOneMinusEps;
        oc *= powf(1.f - d, st.beta[k]);
    }
    *oo = opac * oc;
}


// Gradients of (m_cond, V_cond, conditioned opacity) -> (means [D], covars
// [D x D], opacity, betas [C]) from the set-up st.  exact selects the
// opacity-path Cholesky backward (0: the reference's solve).
// reference: cond_mean_convariance_opacity_bwd.cu:101-563 (mean/covariance
// path through v22^-1, opacity path through the Cholesky factor with the
// standard Cholesky backward, symmetrised).
template <int C>
__device__ __forceinline__ void cond_bwd_body(const CondRegs<C> &st, float opac, const float *__restrict__ vm,
                                              const float *__restrict__ vV, float gO, int exact,
                                              float *__restrict__ gm, float *__restrict__ gV,
                                              float *__restrict__ go, float *__restrict__ gb) {
    constexpr int D = C + 3;
    float gVl[D][D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) gVl[r][c] = 0.f;
    float gml[D], gbl[C];
    const float gM[3] = {vm[0], vm[1], vm[2]};
    float Gv[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int s = 0; s < 3; ++s) Gv[r][s] = vV[r * 3 + s];
#pragma unroll
    for (int r = 0; r < 3; ++r) gml[r] = gM[r];
    float Gr[3][C], gx[C];
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) Gr[r][c] = gM[r] * st.x[c];
#pragma unroll
    for (int c = 0; c < C; ++c) gx[c] = (st.rb[0][c] * gM[0] + st.rb[1][c] * gM[1]) + st.rb[2][c] * gM[2];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int s = 0; s < 3; ++s) gVl[r][s] += Gv[r][s];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const float a0 = st.v21[c][0], a1 = st.v21[c][1], a2 = st.v21[c][2];
#pragma unroll
        for (int r = 0; r < 3; ++r) Gr[r][c] += -((Gv[r][0] * a0 + Gv[r][1] * a1) + Gv[r][2] * a2);
    }
#pragma unroll
    for (int rr = 0; rr < C; ++rr) {
        const float b0 = st.rb[0][rr], b1 = st.rb[1][rr], b2 = st.rb[2][rr];
#pragma unroll
        for (int s = 0; s < 3; ++s) gVl[3 + rr][s] += -((b0 * Gv[0][s] + b1 * Gv[1][s]) + b2 * Gv[2][s]);
    }
    float dba[C], G_r[3][C];
#pragma unroll
    for (int c = 0; c < C; ++c)
        dba[c] = (Gr[0][c] * st.r_[0][c] + Gr[1][c] * st.r_[1][c]) + Gr[2][c] * st.r_[2][c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) G_r[r][c] = Gr[r][c] * st.badj[c];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < C; ++k) acc += G_r[r][k] * st.i22[c][k];
            gVl[r][3 + c] += acc;
        }
    float Gi22[C][C];
#pragma unroll
    for (int r = 0; r < C; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < 3; ++k) acc += st.v12[k][r] * G_r[k][c];
            Gi22[r][c] = acc;
        }
    // ---- opacity path
    float dv[C], base[C], oc = 1.f;
#pragma unroll
    for (int k = 0; k < C; ++k) {
        float d = tanhf(st.y[k] * st.y[k]);
        if (d < 0.f) d = 0.f;
        if (d > kOneMinusEps) d = kOneMinusEps;
        dv[k] = d;
        base[k] = 1.f - d;
        oc *= powf(base[k], st.beta[k]);
    }
    *go = gO * oc;
    const float goc = gO * opac;
    float g_y[C], a[C];
#pragma unroll
    for (int k = 0; k < C; ++k) {
        float g_d = 0.f;
        gbl[k] = 0.f;
        if (base[k] > 0.f) {
            gbl[k] += goc * oc * logf(base[k]);
            if (dv[k] > 0.f && dv[k] < kOneMinusEps) g_d += goc * (-oc * st.beta[k] / base[k]);
        }
        g_y[k] = g_d * (2.f * st.y[k] * (1.f - dv[k] * dv[k]));
    }
#pragma unroll
    for (int k = C - 1; k >= 0; --k) {
        float sum = g_y[k];
#pragma unroll
        for (int j = k + 1; j < C; ++j) sum -= st.Lc[j][k] * a[j];
        float den = st.Lc[k][k];
        if (den == 0.f) den = 1e-20f;
        a[k] = sum / den;
    }
#pragma unroll
    for (int k = 0; k < C; ++k) gx[k] += a[k];
    // Cholesky backward: S = Phi(L^T gL), G = L^-T S L^-1, gA = (G + G^T)/2
    float S[C][C], U[C][C], Gm[C][C];
#pragma unroll
    for (int r = 0; r < C; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < C; ++k) {
                const float gl = (c <= k) ? -(a[k] * st.y[c]) : 0.f;
                acc += st.Lc[k][r] * gl;
            }
            S[r][c] = acc;
        }
#pragma unroll
    for (int r = 0; r < C; ++r) {
#pragma unroll
        for (int c = r + 1; c < C; ++c) S[r][c] = 0.f;
        S[r][r] *= 0.5f;
    }
#pragma unroll
    for (int col = 0; col < C; ++col)
#pragma unroll
        for (int k = C - 1; k >= 0; --k) {
            float sum = S[k][col];
#pragma unroll
            for (int j = k + 1; j < C; ++j) sum -= st.Lc[j][k] * U[j][col];
            float den = st.Lc[k][k];
            if (den == 0.f) den = 1e-20f;
            U[k][col] = sum / den;
        }
    if (exact) {
        // G = U L^-1 (back substitution with L^T, row by row)
#pragma unroll
        for (int row = 0; row < C; ++row)
#pragma unroll
            for (int k = C - 1; k >= 0; --k) {
                float sum = U[row][k];
#pragma unroll
                for (int j = k + 1; j < C; ++j) sum -= st.Lc[j][k] * Gm[row][j];
                float den = st.Lc[k][k];
                if (den == 0.f) den = 1e-20f;
                Gm[row][k] = sum / den;
            }
    } else {
        // the reference's solve (cond_mean_convariance_opacity_bwd.cu:508-518),
        // a forward substitution with L that yields L^-1 U^T
#pragma unroll
        for (int col = 0; col < C; ++col)
#pragma unroll
            for (int k = 0; k < C; ++k) {
                float sum = U[col][k];
#pragma unroll
                for (int j = 0; j < k; ++j) sum -= st.Lc[k][j] * Gm[j][col];
                float den = st.Lc[k][k];
                if (den == 0.f) den = 1e-20f;
                Gm[k][col] = sum / den;
            }
    }
#pragma unroll
    for (int r = 0; r < C; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) gVl[3 + r][3 + c] += 0.5f * (Gm[r][c] + Gm[c][r]);
#pragma unroll
    for (int k = 0; k < C; ++k) gml[3 + k] = -gx[k];
#pragma unroll
    for (int k = 0; k < C; ++k) gbl[k] += dba[k] * ((st.beta[k] < 4.f) ? 0.25f : 0.f);
    float tmp[C][C];
#pragma unroll
    for (int r = 0; r < C; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < C; ++k) acc += Gi22[r][k] * st.i22[c][k];
            tmp[r][c] = acc;
        }
#pragma unroll
    for (int r = 0; r < C; ++r)
#pragma unroll
        for (int c = 0; c < C; ++c) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < C; ++k) acc += st.i22[k][r] * tmp[k][c];
            gVl[3 + r][3 + c] += -acc;
        }
#pragma unroll
    for (int k = 0; k < D; ++k) gm[k] = gml[k];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) gV[r * D + c] = gVl[r][c];
#pragma unroll
    for (int k = 0; k < C; ++k) gb[k] = gbl[k];
}


}  // namespace ubs

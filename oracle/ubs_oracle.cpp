// TEST INFRASTRUCTURE ONLY -- the oracle.
//
// CPU restatement of the Beta-splat tile rasteriser that universal-beta-splatting
// reaches through its gsplat fork.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load this library, and only as the checker; the
// product path (libubs_hip.so) never links or calls it.
//
// Every function is instantiated for float (bit-exact twin of the HIP kernels
// for the integer/structural stages) and double (the FP64 oracle).  The
// float instantiation follows one fixed operation order which the HIP kernels
// also follow; the build uses -ffp-contract=off so no multiply-add is fused.
// Citations are to files of the reference (relative to its root).
//
// Parity status per function (pinned = checked by a test that does not reuse
// this file's formula; see tests/test_oracle_*.py):
//   ora_rotmat_*           pinned (closed form R = I + A, FD)
//   ora_covar_*            pinned (dense numpy L@L.T, torch autograd FP64)
//   ora_cond_*             pinned (Schur complement / cholesky via numpy,
//                          torch autograd FP64, central finite differences)
//   ora_proj_*             pinned (on-axis closed form, culling cases,
//                          torch autograd FP64 of a dense restatement)
//   ora_isect_* / offsets  pinned (python brute force, numpy stable argsort,
//                          searchsorted)
//   ora_raster_*           pinned (1- and 2-primitive closed forms, T cutoff,
//                          central finite differences in FP64)
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <typename T> struct M3 { T m[3][3]; };  // m[row][col]
template <typename T> struct M2 { T m[2][2]; };

// --------------------------------------------------------------------------
// l_triangle_to_rotmat: R = I + A, A skew from (a01, a02, a12).
// reference: submodules/gsplat/cuda/csrc/l_triagnle_to_rotmat_fwd.cu:16-35,
//            l_triagnle_to_rotmat_bwd.cu:16-25
// --------------------------------------------------------------------------
template <typename T>
void rotmat_fwd(int64_t N, const T *l, T *R) {
    for (int64_t i = 0; i < N; ++i) {
        const T a0 = l[i * 3 + 0], a1 = l[i * 3 + 1], a2 = l[i * 3 + 2];
        T *r = R + i * 9;
        r[0] = T(1); r[1] = a0;   r[2] = a1;
        r[3] = -a0;  r[4] = T(1); r[5] = a2;
        r[6] = -a1;  r[7] = -a2;  r[8] = T(1);
    }
}

template <typename T>
void rotmat_bwd(int64_t N, const T *vR, T *vl) {
    for (int64_t i = 0; i < N; ++i) {
        const T *g = vR + i * 9;
        vl[i * 3 + 0] = g[1] - g[3];
        vl[i * 3 + 1] = g[2] - g[6];
        vl[i * 3 + 2] = g[5] - g[7];
    }
}

// --------------------------------------------------------------------------
// rot_scale_l_triangle_to_covar: Sigma = L L^T.
//   L[0:3,0:3] = R diag(s0..s2); L[i,i] = s_i (i>=3);
//   L[rest_i[m], rest_j[m]] = l_tri[3+m].
// reference: rot_scale_l_triangle_to_covar_fwd.cu:28-187 (summation order:
// 3x3 block ((a+b)+c); cross block ascending k from 0; diagonal starts at
// s_r^2 then adds the row's entries; off-diagonal ascending k then + s_c x_rc)
// The dense L form below reproduces that order term by term for rest
// patterns grouped by row with ascending columns (torch.tril_indices order,
// scene/beta_model.py:69-73).
// --------------------------------------------------------------------------
constexpr int kMaxD = 11;

template <typename T>
void build_L(int64_t D, const T *R, const T *s, const T *ltri, const int32_t *ri,
             const int32_t *rj, int64_t M, T L[kMaxD][kMaxD], bool present[kMaxD][kMaxD]) {
    for (int r = 0; r < kMaxD; ++r)
        for (int c = 0; c < kMaxD; ++c) { L[r][c] = T(0); present[r][c] = false; }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) { L[r][c] = R[r * 3 + c] * s[c]; present[r][c] = true; }
    for (int r = 3; r < D; ++r) { L[r][r] = s[r]; present[r][r] = true; }
    for (int64_t m = 0; m < M; ++m) {
        const int r = ri[m], c = rj[m];
        if (r >= D || c >= r || r < 3) continue;
        L[r][c] = ltri[3 + m];
        present[r][c] = true;
    }
}

template <typename T>
void covar_fwd(int64_t N, int64_t D, const T *rot, const T *scale, const T *ltri,
               const int32_t *ri, const int32_t *rj, int64_t M, bool spatial, T *out) {
    const int64_t full_M = D * (D - 1) / 2;
    for (int64_t i = 0; i < N; ++i) {
        const T *R = rot + i * 9, *s = scale + i * D, *lt = ltri + i * full_M;
        T L[kMaxD][kMaxD];
        bool P[kMaxD][kMaxD];
        build_L<T>(D, R, s, lt, ri, rj, M, L, P);
        if (spatial || D == 3) {
            T *o = out + i * 9;
            for (int r = 0; r < 3; ++r)
                for (int c = r; c < 3; ++c) {
                    const T v = (L[r][0] * L[c][0] + L[r][1] * L[c][1]) + L[r][2] * L[c][2];
                    o[r * 3 + c] = v;
                    o[c * 3 + r] = v;
                }
            continue;
        }
        T *o = out + i * D * D;
        for (int r = 0; r < 3; ++r)
            for (int c = r; c < 3; ++c) {
                const T v = (L[r][0] * L[c][0] + L[r][1] * L[c][1]) + L[r][2] * L[c][2];
                o[r * D + c] = v;
                o[c * D + r] = v;
            }
        for (int c = 3; c < D; ++c)
            for (int r = 0; r < 3; ++r) {
                T acc = T(0);
                for (int k = 0; k < 3; ++k)
                    if (P[c][k]) acc += L[r][k] * L[c][k];
                o[r * D + c] = acc;
                o[c * D + r] = acc;
            }
        for (int r = 3; r < D; ++r) {
            T acc = s[r] * s[r];
            for (int k = 0; k < r; ++k)
                if (P[r][k]) acc += L[r][k] * L[r][k];
            o[r * D + r] = acc;
        }
        for (int r = 4; r < D; ++r)
            for (int c = 3; c < r; ++c) {
                T acc = T(0);
                for (int k = 0; k < c; ++k)
                    if (P[c][k]) acc += (P[r][k] ? L[r][k] : T(0)) * L[c][k];
                acc += s[c] * (P[r][c] ? L[r][c] : T(0));
                o[r * D + c] = acc;
                o[c * D + r] = acc;
            }
    }
}

// Backward: dL = (G + G^T) L restricted to the parameterised entries.
// reference: rot_scale_l_triangle_to_covar_bwd.cu:48-234
template <typename T>
void covar_bwd(int64_t N, int64_t D, const T *rot, const T *scale, const T *ltri,
               const int32_t *ri, const int32_t *rj, int64_t M, bool spatial, const T *vcov,
               T *vrot, T *vscale, T *vltri) {
    const int64_t full_M = D * (D - 1) / 2;
    const int dim = (spatial || D == 3) ? 3 : (int)D;
    for (int64_t i = 0; i < N; ++i) {
        const T *R = rot + i * 9, *s = scale + i * D, *lt = ltri + i * full_M;
        const T *G = vcov + i * dim * dim;
        T L[kMaxD][kMaxD];
        bool P[kMaxD][kMaxD];
        build_L<T>(D, R, s, lt, ri, rj, M, L, P);
        auto S = [&](int a, int b) { return G[a * dim + b] + G[b * dim + a]; };
        T *dR = vrot + i * 9, *dS = vscale + i * D, *dL = vltri + i * full_M;
        for (int k = 0; k < 9; ++k) dR[k] = T(0);
        for (int k = 0; k < D; ++k) dS[k] = T(0);
        for (int k = 0; k < full_M; ++k) dL[k] = T(0);
        // H = S_EE Lxyz (+ S_EF L_FE)
        T H[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                H[r][c] = (S(r, 0) * L[0][c] + S(r, 1) * L[1][c]) + S(r, 2) * L[2][c];
        if (dim == D) {
            for (int c = 3; c < D; ++c)
                for (int k = 0; k < 3; ++k) {
                    if (!P[c][k]) continue;
                    for (int r = 0; r < 3; ++r) H[r][k] += S(r, c) * L[c][k];
                }
        }
        for (int c = 0; c < 3; ++c) {
            for (int r = 0; r < 3; ++r) dR[r * 3 + c] = H[r][c] * s[c];
            dS[c] = (H[0][c] * R[0 * 3 + c] + H[1][c] * R[1 * 3 + c]) + H[2][c] * R[2 * 3 + c];
        }
        if (dim != D) continue;
        for (int64_t m = 0; m < M; ++m) {
            const int r = ri[m], c = rj[m];
            if (r >= D || c >= r || r < 3) continue;
            T g = T(0);
            if (c < 3) {
                g += S(r, 0) * L[0][c];
                g += S(r, 1) * L[1][c];
                g += S(r, 2) * L[2][c];
                for (int j = 3; j < D; ++j)
                    if (P[j][c]) g += S(r, j) * L[j][c];
            } else {
                g += S(r, c) * s[c];
                for (int j = c + 1; j < D; ++j)
                    if (P[j][c]) g += S(r, j) * L[j][c];
            }
            dL[3 + m] = g;
        }
        for (int r = 3; r < D; ++r) {
            T ds = S(r, r) * s[r];
            for (int j = r + 1; j < D; ++j)
                if (P[j][r]) ds += S(r, j) * L[j][r];
            dS[r] = ds;
        }
    }
}

// --------------------------------------------------------------------------
// cond_mean_convariance_opacity.
// reference: cond_mean_convariance_opacity_fwd.cu:26-290 (3x3 cofactor inverse
// for C==3, Gauss-Jordan with partial pivoting otherwise, Cholesky guard
// s<=0 -> 1e-20), _bwd.cu:101-563
// --------------------------------------------------------------------------
constexpr int kMaxC = 8;

template <typename T> T eps_of() { return std::numeric_limits<T>::epsilon(); }

template <typename T>
void inv3(const T *a, T *o) {
    const T c00 = a[4] * a[8] - a[5] * a[7];
    const T c01 = -(a[3] * a[8] - a[5] * a[6]);
    const T c02 = a[3] * a[7] - a[4] * a[6];
    const T c10 = -(a[1] * a[8] - a[2] * a[7]);
    const T c11 = a[0] * a[8] - a[2] * a[6];
    const T c12 = -(a[0] * a[7] - a[1] * a[6]);
    const T c20 = a[1] * a[5] - a[2] * a[4];
    const T c21 = -(a[0] * a[5] - a[2] * a[3]);
    const T c22 = a[0] * a[4] - a[1] * a[3];
    T det = (a[0] * c00 + a[1] * c01) + a[2] * c02;
    if (det == T(0)) det = T(1e-20f);
    const T id = T(1) / det;
    o[0] = c00 * id; o[1] = c10 * id; o[2] = c20 * id;
    o[3] = c01 * id; o[4] = c11 * id; o[5] = c21 * id;
    o[6] = c02 * id; o[7] = c12 * id; o[8] = c22 * id;
}

template <typename T>
void inv_gj(const T *A, T *inv, int n) {
    T aug[kMaxC][2 * kMaxC];
    for (int r = 0; r < kMaxC; ++r)
        for (int c = 0; c < 2 * kMaxC; ++c) aug[r][c] = T(0);
    for (int r = 0; r < n; ++r) {
        for (int c = 0; c < n; ++c) aug[r][c] = A[r * n + c];
        aug[r][n + r] = T(1);
    }
    for (int col = 0; col < n; ++col) {
        int piv = col;
        T maxabs = std::fabs(aug[piv][col]);
        for (int r = col + 1; r < n; ++r) {
            const T v = std::fabs(aug[r][col]);
            if (v > maxabs) { maxabs = v; piv = r; }
        }
        if (piv != col)
            for (int c = 0; c < 2 * n; ++c) std::swap(aug[col][c], aug[piv][c]);
        T diag = aug[col][col];
        if (diag == T(0)) diag = T(1e-20f);
        const T invd = T(1) / diag;
        for (int c = 0; c < 2 * n; ++c) aug[col][c] *= invd;
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const T f = aug[r][col];
            if (f != T(0))
                for (int c = 0; c < 2 * n; ++c) aug[r][c] -= f * aug[col][c];
        }
    }
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) inv[r * n + c] = aug[r][n + c];
}

template <typename T> struct CondState {
    int C;
    T x[kMaxC], beta[kMaxC], badj[kMaxC];
    T v11[9], v12[3 * kMaxC], v21[kMaxC * 3], v22[kMaxC * kMaxC], i22[kMaxC * kMaxC];
    T r_[3 * kMaxC], rb[3 * kMaxC];
    T Lc[kMaxC][kMaxC], y[kMaxC];
};

template <typename T>
void cond_setup(int64_t D, const T *m, const T *V, const T *b, const T *q, CondState<T> &st) {
    const int C = int(D - 3);
    st.C = C;
    for (int j = 0; j < C; ++j) {
        st.x[j] = q[j] - m[3 + j];
        st.beta[j] = b[j];
        const T t = b[j] * T(0.25f);
        st.badj[j] = (t < T(1)) ? t : T(1);
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) st.v11[r * 3 + c] = V[r * D + c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < C; ++c) st.v12[r * C + c] = V[r * D + 3 + c];
    for (int r = 0; r < C; ++r)
        for (int c = 0; c < 3; ++c) st.v21[r * 3 + c] = V[(3 + r) * D + c];
    for (int r = 0; r < C; ++r)
        for (int c = 0; c < C; ++c) st.v22[r * C + c] = V[(3 + r) * D + 3 + c];
    if (C == 3) inv3<T>(st.v22, st.i22);
    else inv_gj<T>(st.v22, st.i22, C);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < C; ++c) {
            T acc = T(0);
            for (int k = 0; k < C; ++k) acc += st.v12[r * C + k] * st.i22[k * C + c];
            st.r_[r * C + c] = acc;
            st.rb[r * C + c] = acc * st.badj[c];
        }
    for (int r = 0; r < kMaxC; ++r)
        for (int c = 0; c < kMaxC; ++c) st.Lc[r][c] = T(0);
    for (int i = 0; i < C; ++i)
        for (int j = 0; j <= i; ++j) {
            T sum = st.v22[i * C + j];
            for (int k = 0; k < j; ++k) sum -= st.Lc[i][k] * st.Lc[j][k];
            if (i == j) {
                if (sum <= T(0)) sum = T(1e-20f);
                st.Lc[i][j] = std::sqrt(sum);
            } else {
                T den = st.Lc[j][j];
                if (den == T(0)) den = T(1e-20f);
                st.Lc[i][j] = sum / den;
            }
        }
    for (int i = 0; i < C; ++i) {
        T sum = st.x[i];
        for (int k = 0; k < i; ++k) sum -= st.Lc[i][k] * st.y[k];
        T den = st.Lc[i][i];
        if (den == T(0)) den = T(1e-20f);
        st.y[i] = sum / den;
    }
}

template <typename T>
void cond_fwd(int64_t N, int64_t D, const T *means, const T *covars, const T *opac,
              const T *betas, const T *query, T *om, T *oV, T *oo) {
    const int C = int(D - 3);
    for (int64_t i = 0; i < N; ++i) {
        CondState<T> st;
        cond_setup<T>(D, means + i * D, covars + i * D * D, betas + i * C, query + i * C, st);
        const T *m = means + i * D;
        for (int r = 0; r < 3; ++r) {
            T acc = T(0);
            for (int c = 0; c < C; ++c) acc += st.rb[r * C + c] * st.x[c];
            om[i * 3 + r] = m[r] + acc;
        }
        for (int r = 0; r < 3; ++r)
            for (int s = 0; s < 3; ++s) {
                T acc = T(0);
                for (int k = 0; k < C; ++k) acc += st.rb[r * C + k] * st.v21[k * 3 + s];
                oV[i * 9 + r * 3 + s] = st.v11[r * 3 + s] - acc;
            }
        T oc = T(1);
        const T upper = T(1) - eps_of<T>();
        for (int k = 0; k < C; ++k) {
            T d = std::tanh(st.y[k] * st.y[k]);
            if (d < T(0)) d = T(0);
            if (d > upper) d = upper;
            oc *= std::pow(T(1) - d, st.beta[k]);
        }
        oo[i] = opac[i] * oc;
    }
}

template <typename T>
void cond_bwd(int64_t N, int64_t D, const T *means, const T *covars, const T *opac,
              const T *betas, const T *query, const T *vm, const T *vV, const T *vo, T *gm,
              T *gV, T *go, T *gb, bool exact) {
    const int C = int(D - 3);
    for (int64_t i = 0; i < N; ++i) {
        CondState<T> st;
        cond_setup<T>(D, means + i * D, covars + i * D * D, betas + i * C, query + i * C, st);
        T *gmi = gm + i * D, *gVi = gV + i * D * D, *gbi = gb + i * C;
        for (int k = 0; k < D; ++k) gmi[k] = T(0);
        for (int k = 0; k < D * D; ++k) gVi[k] = T(0);
        for (int k = 0; k < C; ++k) gbi[k] = T(0);
        const T *gM = vm + i * 3, *Gv = vV + i * 9;
        const T gO = vo[i];
        for (int r = 0; r < 3; ++r) gmi[r] += gM[r];
        T Gr[3 * kMaxC], gx[kMaxC];
        for (int c = 0; c < C; ++c)
            for (int r = 0; r < 3; ++r) Gr[r * C + c] = gM[r] * st.x[c];
        for (int c = 0; c < C; ++c)
            gx[c] = (st.rb[0 * C + c] * gM[0] + st.rb[1 * C + c] * gM[1]) + st.rb[2 * C + c] * gM[2];
        for (int r = 0; r < 3; ++r)
            for (int s = 0; s < 3; ++s) gVi[r * D + s] += Gv[r * 3 + s];
        for (int c = 0; c < C; ++c) {
            const T a0 = st.v21[c * 3 + 0], a1 = st.v21[c * 3 + 1], a2 = st.v21[c * 3 + 2];
            for (int r = 0; r < 3; ++r)
                Gr[r * C + c] += -((Gv[r * 3 + 0] * a0 + Gv[r * 3 + 1] * a1) + Gv[r * 3 + 2] * a2);
        }
        for (int rr = 0; rr < C; ++rr) {
            const T b0 = st.rb[0 * C + rr], b1 = st.rb[1 * C + rr], b2 = st.rb[2 * C + rr];
            for (int s = 0; s < 3; ++s)
                gVi[(3 + rr) * D + s] += -((b0 * Gv[0 * 3 + s] + b1 * Gv[1 * 3 + s]) + b2 * Gv[2 * 3 + s]);
        }
        T dba[kMaxC], G_r[3 * kMaxC];
        for (int c = 0; c < C; ++c)
            dba[c] = (Gr[0 * C + c] * st.r_[0 * C + c] + Gr[1 * C + c] * st.r_[1 * C + c]) +
                     Gr[2 * C + c] * st.r_[2 * C + c];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < C; ++c) G_r[r * C + c] = Gr[r * C + c] * st.badj[c];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < C; ++c) {
                T acc = T(0);
                for (int k = 0; k < C; ++k) acc += G_r[r * C + k] * st.i22[c * C + k];
                gVi[r * D + 3 + c] += acc;
            }
        T Gi22[kMaxC * kMaxC];
        for (int r = 0; r < C; ++r)
            for (int c = 0; c < C; ++c) {
                T acc = T(0);
                for (int k = 0; k < 3; ++k) acc += st.v12[k * C + r] * G_r[k * C + c];
                Gi22[r * C + c] = acc;
            }
        // opacity path
        T dv[kMaxC], base[kMaxC], oc = T(1);
        const T upper = T(1) - eps_of<T>();
        for (int k = 0; k < C; ++k) {
            T d = std::tanh(st.y[k] * st.y[k]);
            if (d < T(0)) d = T(0);
            if (d > upper) d = upper;
            dv[k] = d;
            base[k] = T(1) - d;
            oc *= std::pow(base[k], st.beta[k]);
        }
        go[i] = gO * oc;
        const T goc = gO * opac[i];
        T g_d[kMaxC], g_y[kMaxC], a[kMaxC];
        for (int k = 0; k < C; ++k) {
            g_d[k] = T(0);
            if (base[k] > T(0)) {
                gbi[k] += goc * oc * std::log(base[k]);
                if (dv[k] > T(0) && dv[k] < upper) g_d[k] += goc * (-oc * st.beta[k] / base[k]);
            }
        }
        for (int k = 0; k < C; ++k) g_y[k] = g_d[k] * (T(2) * st.y[k] * (T(1) - dv[k] * dv[k]));
        for (int k = C - 1; k >= 0; --k) {
            T sum = g_y[k];
            for (int j = k + 1; j < C; ++j) sum -= st.Lc[j][k] * a[j];
            T den = st.Lc[k][k];
            if (den == T(0)) den = T(1e-20f);
            a[k] = sum / den;
        }
        for (int k = 0; k < C; ++k) gx[k] += a[k];
        T gL[kMaxC][kMaxC], S[kMaxC][kMaxC], U[kMaxC][kMaxC], Gm[kMaxC][kMaxC];
        for (int r = 0; r < kMaxC; ++r)
            for (int c = 0; c < kMaxC; ++c) { gL[r][c] = S[r][c] = U[r][c] = Gm[r][c] = T(0); }
        for (int r = 0; r < C; ++r)
            for (int c = 0; c <= r; ++c) gL[r][c] = -(a[r] * st.y[c]);
        for (int r = 0; r < C; ++r)
            for (int c = 0; c < C; ++c) {
                T acc = T(0);
                for (int k = 0; k < C; ++k) acc += st.Lc[k][r] * gL[k][c];
                S[r][c] = acc;
            }
        for (int r = 0; r < C; ++r) {
            for (int c = r + 1; c < C; ++c) S[r][c] = T(0);
            S[r][r] *= T(0.5f);
        }
        for (int col = 0; col < C; ++col)
            for (int k = C - 1; k >= 0; --k) {
                T sum = S[k][col];
                for (int j = k + 1; j < C; ++j) sum -= st.Lc[j][k] * U[j][col];
                T den = st.Lc[k][k];
                if (den == T(0)) den = T(1e-20f);
                U[k][col] = sum / den;
            }
        if (exact) {
            // G = U L^-1: each row w of G solves w L = u, i.e. L^T w^T = u^T
            // (back substitution with the upper-triangular L^T)
            for (int row = 0; row < C; ++row)
                for (int k = C - 1; k >= 0; --k) {
                    T sum = U[row][k];
                    for (int j = k + 1; j < C; ++j) sum -= st.Lc[j][k] * Gm[row][j];
                    T den = st.Lc[k][k];
                    if (den == T(0)) den = T(1e-20f);
                    Gm[row][k] = sum / den;
                }
        } else {
            // the reference's solve (cond_mean_convariance_opacity_bwd.cu:508-518):
            // a forward substitution with L, which yields L^-1 U^T instead of
            // U L^-1
            for (int col = 0; col < C; ++col)
                for (int k = 0; k < C; ++k) {
                    T sum = U[col][k];
                    for (int j = 0; j < k; ++j) sum -= st.Lc[k][j] * Gm[j][col];
                    T den = st.Lc[k][k];
                    if (den == T(0)) den = T(1e-20f);
                    Gm[k][col] = sum / den;
                }
        }
        for (int r = 0; r < C; ++r)
            for (int c = 0; c < C; ++c)
                gVi[(3 + r) * D + 3 + c] += T(0.5f) * (Gm[r][c] + Gm[c][r]);
        for (int k = 0; k < C; ++k) gmi[3 + k] += -gx[k];
        for (int k = 0; k < C; ++k) gbi[k] += dba[k] * ((st.beta[k] < T(4)) ? T(0.25f) : T(0));
        T tmp[kMaxC * kMaxC];
        for (int r = 0; r < C; ++r)
            for (int c = 0; c < C; ++c) {
                T acc = T(0);
                for (int k = 0; k < C; ++k) acc += Gi22[r * C + k] * st.i22[c * C + k];
                tmp[r * C + c] = acc;
            }
        for (int r = 0; r < C; ++r)
            for (int c = 0; c < C; ++c) {
                T acc = T(0);
                for (int k = 0; k < C; ++k) acc += st.i22[k * C + r] * tmp[k * C + c];
                gVi[(3 + r) * D + 3 + c] += -acc;
            }
    }
}

// --------------------------------------------------------------------------
// fully_fused_projection (EWA).  reference: fully_fused_projection_fwd.cu:20-177,
// utils.cuh:13-96 (quaternion->covariance), :181-211 (ortho), :252-292 (persp),
// :374-466 (world->cam, 2x2 inverse, blur).  Products follow GLM's row-dot-
// column order with ascending k: ((a0 b0 + a1 b1) + a2 b2).
// --------------------------------------------------------------------------
template <typename T>
void quat_to_R(const T *q, T R[3][3]) {
    T w = q[0], x = q[1], y = q[2], z = q[3];
    const T inv = T(1) / std::sqrt(((x * x + y * y) + z * z) + w * w);
    x *= inv; y *= inv; z *= inv; w *= inv;
    const T x2 = x * x, y2 = y * y, z2 = z * z, xy = x * y, xz = x * z, yz = y * z;
    const T wx = w * x, wy = w * y, wz = w * z;
    R[0][0] = T(1) - T(2) * (y2 + z2); R[1][0] = T(2) * (xy + wz); R[2][0] = T(2) * (xz - wy);
    R[0][1] = T(2) * (xy - wz); R[1][1] = T(1) - T(2) * (x2 + z2); R[2][1] = T(2) * (yz + wx);
    R[0][2] = T(2) * (xz + wy); R[1][2] = T(2) * (yz - wx); R[2][2] = T(1) - T(2) * (x2 + y2);
}

template <typename T>
void quat_scale_covar(const T *q, const T *s, T S[3][3]) {
    T R[3][3], Mm[3][3];
    quat_to_R<T>(q, R);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Mm[r][c] = R[r][c] * s[c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            S[r][c] = (Mm[r][0] * Mm[c][0] + Mm[r][1] * Mm[c][1]) + Mm[r][2] * Mm[c][2];
}

template <typename T> struct ProjGeom {
    T R[3][3], t[3], pc[3], Sw[3][3], Sc[3][3];
    T fx, fy, cx, cy, W, H;
    T lxp, lxn, lyp, lyn, rz, rz2, tx, ty;
    T J[2][3], cov2d[2][2];
};

template <typename T>
void load_cov(const T *covars, const T *quats, const T *scales, int64_t g, T S[3][3]) {
    if (covars) {
        const T *c = covars + g * 6;
        S[0][0] = c[0]; S[0][1] = c[1]; S[0][2] = c[2];
        S[1][0] = c[1]; S[1][1] = c[3]; S[1][2] = c[4];
        S[2][0] = c[2]; S[2][1] = c[4]; S[2][2] = c[5];
    } else {
        quat_scale_covar<T>(quats + g * 4, scales + g * 3, S);
    }
}

template <typename T>
void proj_geom(const T *means, const T *covars, const T *quats, const T *scales, const T *vm,
               const T *K, int64_t g, int32_t W, int32_t H, bool ortho, ProjGeom<T> &G) {
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) G.R[r][c] = vm[r * 4 + c];
        G.t[r] = vm[r * 4 + 3];
    }
    const T *p = means + g * 3;
    for (int r = 0; r < 3; ++r) G.pc[r] = ((G.R[r][0] * p[0] + G.R[r][1] * p[1]) + G.R[r][2] * p[2]) + G.t[r];
    load_cov<T>(covars, quats, scales, g, G.Sw);
    T Mm[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            Mm[r][c] = (G.R[r][0] * G.Sw[0][c] + G.R[r][1] * G.Sw[1][c]) + G.R[r][2] * G.Sw[2][c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            G.Sc[r][c] = (Mm[r][0] * G.R[c][0] + Mm[r][1] * G.R[c][1]) + Mm[r][2] * G.R[c][2];
    G.fx = K[0]; G.fy = K[4]; G.cx = K[2]; G.cy = K[5];
    G.W = T(W); G.H = T(H);
    const T x = G.pc[0], y = G.pc[1], z = G.pc[2];
    if (ortho) {
        G.J[0][0] = G.fx; G.J[0][1] = T(0); G.J[0][2] = T(0);
        G.J[1][0] = T(0); G.J[1][1] = G.fy; G.J[1][2] = T(0);
    } else {
        const T tfx = (T(0.5f) * G.W) / G.fx, tfy = (T(0.5f) * G.H) / G.fy;
        G.lxp = (G.W - G.cx) / G.fx + T(0.3f) * tfx;
        G.lxn = G.cx / G.fx + T(0.3f) * tfx;
        G.lyp = (G.H - G.cy) / G.fy + T(0.3f) * tfy;
        G.lyn = G.cy / G.fy + T(0.3f) * tfy;
        G.rz = T(1) / z;
        G.rz2 = G.rz * G.rz;
        G.tx = z * std::min(G.lxp, std::max(-G.lxn, x * G.rz));
        G.ty = z * std::min(G.lyp, std::max(-G.lyn, y * G.rz));
        G.J[0][0] = G.fx * G.rz; G.J[0][1] = T(0); G.J[0][2] = -G.fx * G.tx * G.rz2;
        G.J[1][0] = T(0); G.J[1][1] = G.fy * G.rz; G.J[1][2] = -G.fy * G.ty * G.rz2;
    }
    T A[2][3];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c)
            A[r][c] = (G.J[r][0] * G.Sc[0][c] + G.J[r][1] * G.Sc[1][c]) + G.J[r][2] * G.Sc[2][c];
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c)
            G.cov2d[r][c] = (A[r][0] * G.J[c][0] + A[r][1] * G.J[c][1]) + A[r][2] * G.J[c][2];
}

// float->int32 as the GPU converts it: truncation, saturating, NaN -> 0.
template <typename T> int32_t f2i_sat(T v) {
    if (!(v == v)) return 0;
    if (v >= T(2147483647.0)) return 2147483647;
    if (v <= T(-2147483648.0)) return (-2147483647 - 1);
    return (int32_t)v;
}

template <typename T>
void proj_fwd(int64_t C, int64_t N, const T *means, const T *covars, const T *quats,
              const T *scales, const T *viewmats, const T *Ks, int32_t W, int32_t H, T eps2d,
              T near_plane, T far_plane, T radius_clip, int ortho, int32_t *radii, T *means2d,
              T *depths, T *conics, T *comp) {
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < C * N; ++idx) {
        const int64_t c = idx / N, g = idx % N;
        const T *vm = viewmats + c * 16, *K = Ks + c * 9;
        // early z-cull needs only the camera-space mean
        T pc2 = T(0);
        {
            const T *p = means + g * 3;
            pc2 = ((vm[8] * p[0] + vm[9] * p[1]) + vm[10] * p[2]) + vm[11];
        }
        if (pc2 < near_plane || pc2 > far_plane) { radii[idx] = 0; continue; }
        ProjGeom<T> G;
        proj_geom<T>(means, covars, quats, scales, vm, K, g, W, H, ortho != 0, G);
        T m2x, m2y;
        if (ortho) { m2x = G.fx * G.pc[0] + G.cx; m2y = G.fy * G.pc[1] + G.cy; }
        else { m2x = G.fx * G.pc[0] * G.rz + G.cx; m2y = G.fy * G.pc[1] * G.rz + G.cy; }
        T c00 = G.cov2d[0][0], c01 = G.cov2d[0][1], c10 = G.cov2d[1][0], c11 = G.cov2d[1][1];
        const T o00 = c00, o11 = c11;
        const T det_orig = c00 * c11 - c01 * c10;
        c00 += eps2d;
        c11 += eps2d;
        T det = c00 * c11 - c01 * c10;
        T cmp, i00, i01, i11;
        // The build's FP64 branch for a near-singular blurred 2-D covariance
        // (det < 2^-8 c00 c11; the reference inverts in fp32 throughout,
        // fully_fused_projection_fwd.cu:137-150, utils.cuh:438-448): the
        // determinant of the same fp32 entries in FP64, then cull, inverse and
        // compensation from it, each rounded once to T.
        if (det < T(0.00390625f) * (c00 * c11)) {
            const double d64 = (double)c00 * (double)c11 - (double)c01 * (double)c10;
            if (!(d64 > 0.0)) { radii[idx] = 0; continue; }
            const double o64 = (double)o00 * (double)o11 - (double)c01 * (double)c10;
            cmp = std::sqrt(std::max(T(0), T(o64 / d64)));
            i00 = T((double)c11 / d64);
            i01 = T(-(double)c01 / d64);
            i11 = T((double)c00 / d64);
            det = T(d64);
        } else {
            cmp = std::sqrt(std::max(T(0), det_orig / det));
            if (det <= T(0)) { radii[idx] = 0; continue; }
            const T invd = T(1) / det;
            i00 = c11 * invd;
            i01 = -c01 * invd;
            i11 = c00 * invd;
        }
        const T b = T(0.5f) * (c00 + c11);
        const T v1 = b + std::sqrt(std::max(T(0.01f), b * b - det));
        const T radius = std::ceil(std::sqrt(v1));
        if (radius <= radius_clip) { radii[idx] = 0; continue; }
        if (m2x + radius <= T(0) || m2x - radius >= T(W) || m2y + radius <= T(0) ||
            m2y - radius >= T(H)) {
            radii[idx] = 0;
            continue;
        }
        radii[idx] = f2i_sat(radius);
        means2d[idx * 2] = m2x;
        means2d[idx * 2 + 1] = m2y;
        depths[idx] = G.pc[2];
        conics[idx * 3] = i00;
        conics[idx * 3 + 1] = i01;
        conics[idx * 3 + 2] = i11;
        if (comp) comp[idx] = cmp;
    }
}

// Backward.  reference: fully_fused_projection_bwd.cu:50-239, utils.cuh:38-62
// (quat vjp), :98-136 (quat-scale vjp), :214-249 (ortho vjp), :294-372 (persp
// vjp), :385-435 (world->cam vjp), :451-502 (inverse / blur vjp).
template <typename T>
void proj_bwd(int64_t C, int64_t N, const T *means, const T *covars, const T *quats,
              const T *scales, const T *viewmats, const T *Ks, int32_t W, int32_t H, T eps2d,
              int ortho, const int32_t *radii, const T *conics, const T *comp, const T *v_means2d,
              const T *v_depths, const T *v_conics, const T *v_comp, T *v_means, T *v_covars,
              T *v_quats, T *v_scales, T *v_viewmats) {
    for (int64_t idx = 0; idx < C * N; ++idx) {
        if (radii[idx] <= 0) continue;
        const int64_t c = idx / N, g = idx % N;
        const T *vm = viewmats + c * 16, *K = Ks + c * 9;
        const T *cn = conics + idx * 3, *vc = v_conics + idx * 3;
        const T P[2][2] = {{cn[0], cn[1]}, {cn[1], cn[2]}};
        const T vP[2][2] = {{vc[0], vc[1] * T(0.5f)}, {vc[1] * T(0.5f), vc[2]}};
        T vcov2[2][2];
        {
            T tmp[2][2];
            for (int r = 0; r < 2; ++r)
                for (int s = 0; s < 2; ++s) tmp[r][s] = P[r][0] * vP[0][s] + P[r][1] * vP[1][s];
            for (int r = 0; r < 2; ++r)
                for (int s = 0; s < 2; ++s) vcov2[r][s] = -(tmp[r][0] * P[0][s] + tmp[r][1] * P[1][s]);
        }
        if (v_comp) {
            const T cmpv = comp[idx], vcp = v_comp[idx];
            const T detc = P[0][0] * P[1][1] - P[0][1] * P[1][0];
            const T vsq = vcp * T(0.5) / (cmpv + T(1e-6));
            const T om = T(1) - cmpv * cmpv;
            vcov2[0][0] += vsq * (om * P[0][0] - eps2d * detc);
            vcov2[0][1] += vsq * (om * P[0][1]);
            vcov2[1][0] += vsq * (om * P[1][0]);
            vcov2[1][1] += vsq * (om * P[1][1] - eps2d * detc);
        }
        ProjGeom<T> G;
        proj_geom<T>(means, covars, quats, scales, vm, K, g, W, H, ortho != 0, G);
        T vSc[3][3] = {{0}}, vpc[3] = {0, 0, 0};
        const T *vm2 = v_means2d + idx * 2;
        // v_Sc += J^T vcov2 J
        {
            T tmp[3][2];
            for (int r = 0; r < 3; ++r)
                for (int s = 0; s < 2; ++s) tmp[r][s] = G.J[0][r] * vcov2[0][s] + G.J[1][r] * vcov2[1][s];
            for (int r = 0; r < 3; ++r)
                for (int s = 0; s < 3; ++s) vSc[r][s] += tmp[r][0] * G.J[0][s] + tmp[r][1] * G.J[1][s];
        }
        const T x = G.pc[0], y = G.pc[1];
        if (ortho) {
            vpc[0] += G.fx * vm2[0];
            vpc[1] += G.fy * vm2[1];
        } else {
            const T rz = G.rz, rz2 = G.rz2, rz3 = rz2 * rz;
            vpc[0] += G.fx * rz * vm2[0];
            vpc[1] += G.fy * rz * vm2[1];
            vpc[2] += -(G.fx * x * vm2[0] + G.fy * y * vm2[1]) * rz2;
            // v_J = vcov2 J Sc^T + vcov2^T J Sc
            T vJ[2][3];
            {
                T JS[2][3], JSt[2][3];
                for (int r = 0; r < 2; ++r)
                    for (int s = 0; s < 3; ++s) {
                        JS[r][s] = G.J[r][0] * G.Sc[0][s] + G.J[r][1] * G.Sc[1][s] + G.J[r][2] * G.Sc[2][s];
                        JSt[r][s] = G.J[r][0] * G.Sc[s][0] + G.J[r][1] * G.Sc[s][1] + G.J[r][2] * G.Sc[s][2];
                    }
                for (int r = 0; r < 2; ++r)
                    for (int s = 0; s < 3; ++s)
                        vJ[r][s] = (vcov2[r][0] * JSt[0][s] + vcov2[r][1] * JSt[1][s]) +
                                   (vcov2[0][r] * JS[0][s] + vcov2[1][r] * JS[1][s]);
            }
            if (x * rz <= G.lxp && x * rz >= -G.lxn) vpc[0] += -G.fx * rz2 * vJ[0][2];
            else vpc[2] += -G.fx * rz3 * vJ[0][2] * G.tx;
            if (y * rz <= G.lyp && y * rz >= -G.lyn) vpc[1] += -G.fy * rz2 * vJ[1][2];
            else vpc[2] += -G.fy * rz3 * vJ[1][2] * G.ty;
            vpc[2] += -G.fx * rz2 * vJ[0][0] - G.fy * rz2 * vJ[1][1] +
                      T(2) * G.fx * G.tx * rz3 * vJ[0][2] + T(2) * G.fy * G.ty * rz3 * vJ[1][2];
        }
        vpc[2] += v_depths[idx];
        const T *p = means + g * 3;
        T vR[3][3], vt[3], vmean[3], vS[3][3];
        for (int r = 0; r < 3; ++r) {
            for (int s = 0; s < 3; ++s) vR[r][s] = vpc[r] * p[s];
            vt[r] = vpc[r];
            vmean[r] = G.R[0][r] * vpc[0] + G.R[1][r] * vpc[1] + G.R[2][r] * vpc[2];
        }
        // v_R += vSc R S^T + vSc^T R S ; v_S += R^T vSc R
        {
            T RS[3][3], RSt[3][3], tmp[3][3];
            for (int r = 0; r < 3; ++r)
                for (int s = 0; s < 3; ++s) {
                    RS[r][s] = G.R[r][0] * G.Sw[0][s] + G.R[r][1] * G.Sw[1][s] + G.R[r][2] * G.Sw[2][s];
                    RSt[r][s] = G.R[r][0] * G.Sw[s][0] + G.R[r][1] * G.Sw[s][1] + G.R[r][2] * G.Sw[s][2];
                }
            for (int r = 0; r < 3; ++r)
                for (int s = 0; s < 3; ++s)
                    vR[r][s] += (vSc[r][0] * RSt[0][s] + vSc[r][1] * RSt[1][s] + vSc[r][2] * RSt[2][s]) +
                                (vSc[0][r] * RS[0][s] + vSc[1][r] * RS[1][s] + vSc[2][r] * RS[2][s]);
            for (int r = 0; r < 3; ++r)
                for (int s = 0; s < 3; ++s)
                    tmp[r][s] = G.R[0][r] * vSc[0][s] + G.R[1][r] * vSc[1][s] + G.R[2][r] * vSc[2][s];
            for (int r = 0; r < 3; ++r)
                for (int s = 0; s < 3; ++s)
                    vS[r][s] = tmp[r][0] * G.R[0][s] + tmp[r][1] * G.R[1][s] + tmp[r][2] * G.R[2][s];
        }
        for (int r = 0; r < 3; ++r) v_means[g * 3 + r] += vmean[r];
        if (covars) {
            T *o = v_covars + g * 6;
            o[0] += vS[0][0];
            o[1] += vS[0][1] + vS[1][0];
            o[2] += vS[0][2] + vS[2][0];
            o[3] += vS[1][1];
            o[4] += vS[1][2] + vS[2][1];
            o[5] += vS[2][2];
        } else {
            const T *q = quats + g * 4, *s = scales + g * 3;
            T R[3][3], Mm[3][3], vM[3][3], vRq[3][3];
            quat_to_R<T>(q, R);
            for (int r = 0; r < 3; ++r)
                for (int cc = 0; cc < 3; ++cc) Mm[r][cc] = R[r][cc] * s[cc];
            for (int r = 0; r < 3; ++r)
                for (int cc = 0; cc < 3; ++cc)
                    vM[r][cc] = (vS[r][0] + vS[0][r]) * Mm[0][cc] + (vS[r][1] + vS[1][r]) * Mm[1][cc] +
                                (vS[r][2] + vS[2][r]) * Mm[2][cc];
            for (int r = 0; r < 3; ++r)
                for (int cc = 0; cc < 3; ++cc) vRq[r][cc] = vM[r][cc] * s[cc];
            T w = q[0], xq = q[1], yq = q[2], zq = q[3];
            const T inv = T(1) / std::sqrt(((xq * xq + yq * yq) + zq * zq) + w * w);
            xq *= inv; yq *= inv; zq *= inv; w *= inv;
            // GLM v_R[c][r] == vRq[r][c]
            auto V = [&](int col, int row) { return vRq[row][col]; };
            T vqn[4] = {
                T(2) * (xq * (V(1, 2) - V(2, 1)) + yq * (V(2, 0) - V(0, 2)) + zq * (V(0, 1) - V(1, 0))),
                T(2) * (T(-2) * xq * (V(1, 1) + V(2, 2)) + yq * (V(0, 1) + V(1, 0)) +
                        zq * (V(0, 2) + V(2, 0)) + w * (V(1, 2) - V(2, 1))),
                T(2) * (xq * (V(0, 1) + V(1, 0)) - T(2) * yq * (V(0, 0) + V(2, 2)) +
                        zq * (V(1, 2) + V(2, 1)) + w * (V(2, 0) - V(0, 2))),
                T(2) * (xq * (V(0, 2) + V(2, 0)) + yq * (V(1, 2) + V(2, 1)) -
                        T(2) * zq * (V(0, 0) + V(1, 1)) + w * (V(0, 1) - V(1, 0)))};
            const T qn[4] = {w, xq, yq, zq};
            const T dot = vqn[0] * qn[0] + vqn[1] * qn[1] + vqn[2] * qn[2] + vqn[3] * qn[3];
            for (int k = 0; k < 4; ++k) v_quats[g * 4 + k] += (vqn[k] - dot * qn[k]) * inv;
            // v_scale[j] = sum_i R[i][j] * vM[i][j]  (GLM R[0][0]*v_M[0][0] + R[0][1]*v_M[0][1] ...
            // indexes column j of both matrices)
            for (int j = 0; j < 3; ++j)
                v_scales[g * 3 + j] += R[0][j] * vM[0][j] + R[1][j] * vM[1][j] + R[2][j] * vM[2][j];
        }
        if (v_viewmats) {
            T *o = v_viewmats + c * 16;
            for (int r = 0; r < 3; ++r) {
                for (int s = 0; s < 3; ++s) o[r * 4 + s] += vR[r][s];
                o[r * 4 + 3] += vt[r];
            }
        }
    }
}

// --------------------------------------------------------------------------
// Tile intersection, sort and ranges.  reference: isect_tiles.cu:16-97 (count
// and emit), :137-138 (bit widths), :224-278 (stable radix sort on bits
// [0, 32+tile_bits+cam_bits)), :287-333 (offsets).
// --------------------------------------------------------------------------
inline uint32_t sat_u32(float v) {
    // CUDA's float->uint32 conversion saturates (negative and NaN -> 0).
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}

inline uint32_t bit_width(uint32_t v) {
    uint32_t b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
}

void tile_box(float mx, float my, int32_t rad, uint32_t ts, uint32_t tw, uint32_t th,
              uint32_t &x0, uint32_t &y0, uint32_t &x1, uint32_t &y1) {
    const float tr = float(rad) / float(ts);
    const float tx = mx / float(ts), ty = my / float(ts);
    x0 = std::min(sat_u32(std::floor(tx - tr)), tw);
    y0 = std::min(sat_u32(std::floor(ty - tr)), th);
    x1 = std::min(sat_u32(std::ceil(tx + tr)), tw);
    y1 = std::min(sat_u32(std::ceil(ty + tr)), th);
}

// --------------------------------------------------------------------------
// Forward blend.  reference: rasterize_to_pixels_fwd.cu:39-190.
//   sigma = (a dx^2 + c dy^2) + 2 b dx dy; skip sigma<0 or sigma>=1;
//   alpha = min(0.999, o (1-sigma)^beta); stop (exclusive) when T(1-alpha)<=1e-4.
// margin (optional, FP64 use): per pixel, the smallest |next_T - 1e-4| / 1e-4
// seen at any evaluated primitive and the smallest |sigma| of a primitive
// that was accepted or rejected for sigma<0 -- pixels with tiny margins are
// where an FP32 and an FP64 evaluation can take different branches.
// --------------------------------------------------------------------------
template <typename T>
void raster_fwd(int64_t C, int64_t N, int64_t n_isects, int64_t D, const T *means2d,
                const T *conics, const T *colors, const T *opac, const T *betas, const T *bg,
                const uint8_t *masks, int32_t W, int32_t H, int32_t ts, int32_t tw, int32_t th,
                const int32_t *offsets, const int32_t *flat, T *out_c, T *out_a, int32_t *last,
                double *margin_T, double *margin_s) {
    const T thr = std::is_same<T, float>::value ? T(1e-4f) : T(1e-4);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t blk = 0; blk < C * th * tw; ++blk) {
        const int64_t c = blk / (th * tw);
        const int32_t tile = int32_t(blk % (th * tw));
        const int32_t ty = tile / tw, tx = tile % tw;
        const int32_t start = offsets[c * th * tw + tile];
        const int32_t end = (c == C - 1 && tile == tw * th - 1) ? (int32_t)n_isects
                                                                 : offsets[c * th * tw + tile + 1];
        for (int32_t py_ = 0; py_ < ts; ++py_)
            for (int32_t px_ = 0; px_ < ts; ++px_) {
                const int32_t i = ty * ts + py_, j = tx * ts + px_;
                if (i >= H || j >= W) continue;
                const int64_t pix = c * H * W + (int64_t)i * W + j;
                if (masks && !masks[c * th * tw + tile]) {
                    for (int k = 0; k < D; ++k) out_c[pix * D + k] = bg ? bg[c * D + k] : T(0);
                    out_a[pix] = T(0);
                    last[pix] = 0;
                    continue;
                }
                const T px = T(j) + T(0.5f), py = T(i) + T(0.5f);
                T Tr = T(1);
                int32_t cur = 0;
                double mT = 1e300, mS = 1e300;
                std::vector<T> acc(D, T(0));
                for (int32_t k = start; k < end; ++k) {
                    const int32_t g = flat[k];
                    const T *cn = conics + (int64_t)g * 3;
                    const T dx = means2d[(int64_t)g * 2] - px, dy = means2d[(int64_t)g * 2 + 1] - py;
                    const T sig = (cn[0] * dx * dx + cn[2] * dy * dy) + T(2) * cn[1] * dx * dy;
                    if (margin_s && sig < T(1)) mS = std::min(mS, std::fabs((double)sig));
                    if (sig < T(0) || sig >= T(1)) continue;
                    const T alpha = std::min(T(0.999f), opac[g] * std::pow(T(1) - sig, betas[g]));
                    const T nT = Tr * (T(1) - alpha);
                    if (margin_T) mT = std::min(mT, std::fabs((double)nT - 1e-4) / 1e-4);
                    if (nT <= thr) break;
                    const T vis = alpha * Tr;
                    for (int kk = 0; kk < D; ++kk) acc[kk] += colors[(int64_t)g * D + kk] * vis;
                    cur = k;
                    Tr = nT;
                }
                out_a[pix] = T(1) - Tr;
                for (int kk = 0; kk < D; ++kk) out_c[pix * D + kk] = bg ? acc[kk] + Tr * bg[c * D + kk] : acc[kk];
                last[pix] = cur;
                if (margin_T) margin_T[pix] = mT;
                if (margin_s) margin_s[pix] = mS;
            }
    }
}

// Backward blend.  reference: rasterize_to_pixels_bwd.cu:50-275.  Transmittance
// is recovered from T_final = 1 - render_alpha by T /= (1 - alpha) front-to-back
// in reverse; gradient accumulation is in double, per thread, and reduced in a
// fixed order (deterministic).
template <typename T>
void raster_bwd(int64_t C, int64_t N, int64_t n_isects, int64_t D, const T *means2d,
                const T *conics, const T *colors, const T *opac, const T *betas, const T *bg,
                const uint8_t *masks, int32_t W, int32_t H, int32_t ts, int32_t tw, int32_t th,
                const int32_t *offsets, const int32_t *flat, const T *render_a,
                const int32_t *last, const T *v_c, const T *v_a, const T *T_final_opt,
                T *v_means2d, T *v_conics, T *v_colors, T *v_opac, T *v_betas) {
    const int64_t CN = C * N;
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    const int64_t per = 2 + 3 + D + 2;
    std::vector<std::vector<double>> part(nthr);
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        std::vector<double> &acc = part[tid];
        acc.assign(CN * per, 0.0);
#pragma omp for schedule(static)
        for (int64_t blk = 0; blk < C * th * tw; ++blk) {
            const int64_t c = blk / (th * tw);
            const int32_t tile = int32_t(blk % (th * tw));
            if (masks && !masks[c * th * tw + tile]) continue;
            const int32_t ty = tile / tw, tx = tile % tw;
            const int32_t start = offsets[c * th * tw + tile];
            const int32_t end = (c == C - 1 && tile == tw * th - 1) ? (int32_t)n_isects
                                                                     : offsets[c * th * tw + tile + 1];
            std::vector<T> buf(D), vrc(D);
            for (int32_t py_ = 0; py_ < ts; ++py_)
                for (int32_t px_ = 0; px_ < ts; ++px_) {
                    const int32_t i = ty * ts + py_, j = tx * ts + px_;
                    if (i >= H || j >= W) continue;
                    const int64_t pix = c * H * W + (int64_t)i * W + j;
                    const T px = T(j) + T(0.5f), py = T(i) + T(0.5f);
                    const T Tf = T_final_opt ? T_final_opt[pix] : T(1) - render_a[pix];
                    T Tr = Tf;
                    const int32_t binf = last[pix];
                    for (int kk = 0; kk < D; ++kk) { buf[kk] = T(0); vrc[kk] = v_c[pix * D + kk]; }
                    const T vra = v_a[pix];
                    for (int32_t k = end - 1; k >= start; --k) {
                        if (k > binf) continue;
                        const int32_t g = flat[k];
                        const T *cn = conics + (int64_t)g * 3;
                        const T o = opac[g], be = betas[g];
                        const T dx = means2d[(int64_t)g * 2] - px, dy = means2d[(int64_t)g * 2 + 1] - py;
                        const T sig = (cn[0] * dx * dx + cn[2] * dy * dy) + T(2) * cn[1] * dx * dy;
                        if (sig < T(0) || sig >= T(1)) continue;
                        const T vis = std::pow(T(1) - sig, be);
                        const T alpha = std::min(T(0.999f), o * vis);
                        const T ra = T(1) / (T(1) - alpha);
                        Tr *= ra;
                        const T fac = alpha * Tr;
                        double *a = acc.data() + (int64_t)g * per;
                        T valpha = T(0);
                        for (int kk = 0; kk < D; ++kk) {
                            a[5 + kk] += double(fac * vrc[kk]);
                            valpha += (colors[(int64_t)g * D + kk] * Tr - buf[kk] * ra) * vrc[kk];
                        }
                        valpha += Tf * ra * vra;
                        if (bg) {
                            T s = T(0);
                            for (int kk = 0; kk < D; ++kk) s += bg[c * D + kk] * vrc[kk];
                            valpha += -Tf * ra * s;
                        }
                        if (o * vis <= T(0.999f)) {
                            const T vsig = -valpha * o * be * std::pow(T(1) - sig, be - T(1));
                            a[2] += double(dx * dx * vsig);
                            a[3] += double(T(2) * dx * dy * vsig);
                            a[4] += double(dy * dy * vsig);
                            a[0] += double(T(2) * vsig * (cn[0] * dx + cn[1] * dy));
                            a[1] += double(T(2) * vsig * (cn[1] * dx + cn[2] * dy));
                            a[5 + D] += double(vis * valpha);
                            a[6 + D] += double(valpha * o * std::pow(T(1) - sig, be) * std::log(T(1) - sig));
                        }
                        for (int kk = 0; kk < D; ++kk) buf[kk] += colors[(int64_t)g * D + kk] * fac;
                    }
                }
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < CN; ++g) {
        std::vector<double> s(per);
        const int64_t pp = per;
        for (int64_t q = 0; q < pp; ++q) s[q] = 0.0;
        for (int t = 0; t < nthr; ++t)
            for (int64_t q = 0; q < pp; ++q) s[q] += part[t][g * pp + q];
        v_means2d[g * 2] = T(s[0]);
        v_means2d[g * 2 + 1] = T(s[1]);
        v_conics[g * 3] = T(s[2]);
        v_conics[g * 3 + 1] = T(s[3]);
        v_conics[g * 3 + 2] = T(s[4]);
        for (int kk = 0; kk < D; ++kk) v_colors[g * D + kk] = T(s[5 + kk]);
        v_opac[g] = T(s[5 + D]);
        v_betas[g] = T(s[6 + D]);
    }
}

}  // namespace

extern "C" {

void ora_rotmat_fwd_f32(int64_t N, const float *l, float *R) { rotmat_fwd<float>(N, l, R); }
void ora_rotmat_fwd_f64(int64_t N, const double *l, double *R) { rotmat_fwd<double>(N, l, R); }
void ora_rotmat_bwd_f32(int64_t N, const float *v, float *o) { rotmat_bwd<float>(N, v, o); }
void ora_rotmat_bwd_f64(int64_t N, const double *v, double *o) { rotmat_bwd<double>(N, v, o); }

void ora_covar_fwd_f32(int64_t N, int64_t D, const float *rot, const float *scale, const float *lt,
                       const int32_t *ri, const int32_t *rj, int64_t M, int spatial, float *out) {
    covar_fwd<float>(N, D, rot, scale, lt, ri, rj, M, spatial != 0, out);
}
void ora_covar_fwd_f64(int64_t N, int64_t D, const double *rot, const double *scale, const double *lt,
                       const int32_t *ri, const int32_t *rj, int64_t M, int spatial, double *out) {
    covar_fwd<double>(N, D, rot, scale, lt, ri, rj, M, spatial != 0, out);
}
void ora_covar_bwd_f32(int64_t N, int64_t D, const float *rot, const float *scale, const float *lt,
                       const int32_t *ri, const int32_t *rj, int64_t M, int spatial, const float *g,
                       float *vr, float *vs, float *vl) {
    covar_bwd<float>(N, D, rot, scale, lt, ri, rj, M, spatial != 0, g, vr, vs, vl);
}
void ora_covar_bwd_f64(int64_t N, int64_t D, const double *rot, const double *scale, const double *lt,
                       const int32_t *ri, const int32_t *rj, int64_t M, int spatial, const double *g,
                       double *vr, double *vs, double *vl) {
    covar_bwd<double>(N, D, rot, scale, lt, ri, rj, M, spatial != 0, g, vr, vs, vl);
}

void ora_cond_fwd_f32(int64_t N, int64_t D, const float *m, const float *V, const float *o,
                      const float *b, const float *q, float *om, float *oV, float *oo) {
    cond_fwd<float>(N, D, m, V, o, b, q, om, oV, oo);
}
void ora_cond_fwd_f64(int64_t N, int64_t D, const double *m, const double *V, const double *o,
                      const double *b, const double *q, double *om, double *oV, double *oo) {
    cond_fwd<double>(N, D, m, V, o, b, q, om, oV, oo);
}
// exact = 0: the reference's Cholesky-backward solve; 1: the exact gradient.
void ora_cond_bwd_f32(int64_t N, int64_t D, const float *m, const float *V, const float *o,
                      const float *b, const float *q, const float *vm, const float *vV,
                      const float *vo, float *gm, float *gV, float *go, float *gb, int exact) {
    cond_bwd<float>(N, D, m, V, o, b, q, vm, vV, vo, gm, gV, go, gb, exact != 0);
}
void ora_cond_bwd_f64(int64_t N, int64_t D, const double *m, const double *V, const double *o,
                      const double *b, const double *q, const double *vm, const double *vV,
                      const double *vo, double *gm, double *gV, double *go, double *gb, int exact) {
    cond_bwd<double>(N, D, m, V, o, b, q, vm, vV, vo, gm, gV, go, gb, exact != 0);
}

void ora_proj_fwd_f32(int64_t C, int64_t N, const float *means, const float *covars,
                      const float *quats, const float *scales, const float *viewmats,
                      const float *Ks, int32_t W, int32_t H, float eps2d, float near_plane,
                      float far_plane, float radius_clip, int ortho, int32_t *radii,
                      float *means2d, float *depths, float *conics, float *comp) {
    proj_fwd<float>(C, N, means, covars, quats, scales, viewmats, Ks, W, H, eps2d, near_plane,
                    far_plane, radius_clip, ortho, radii, means2d, depths, conics, comp);
}
void ora_proj_fwd_f64(int64_t C, int64_t N, const double *means, const double *covars,
                      const double *quats, const double *scales, const double *viewmats,
                      const double *Ks, int32_t W, int32_t H, double eps2d, double near_plane,
                      double far_plane, double radius_clip, int ortho, int32_t *radii,
                      double *means2d, double *depths, double *conics, double *comp) {
    proj_fwd<double>(C, N, means, covars, quats, scales, viewmats, Ks, W, H, eps2d, near_plane,
                     far_plane, radius_clip, ortho, radii, means2d, depths, conics, comp);
}
void ora_proj_bwd_f32(int64_t C, int64_t N, const float *means, const float *covars,
                      const float *quats, const float *scales, const float *viewmats,
                      const float *Ks, int32_t W, int32_t H, float eps2d, int ortho,
                      const int32_t *radii, const float *conics, const float *comp,
                      const float *v_means2d, const float *v_depths, const float *v_conics,
                      const float *v_comp, float *v_means, float *v_covars, float *v_quats,
                      float *v_scales, float *v_viewmats) {
    proj_bwd<float>(C, N, means, covars, quats, scales, viewmats, Ks, W, H, eps2d, ortho, radii,
                    conics, comp, v_means2d, v_depths, v_conics, v_comp, v_means, v_covars,
                    v_quats, v_scales, v_viewmats);
}
void ora_proj_bwd_f64(int64_t C, int64_t N, const double *means, const double *covars,
                      const double *quats, const double *scales, const double *viewmats,
                      const double *Ks, int32_t W, int32_t H, double eps2d, int ortho,
                      const int32_t *radii, const double *conics, const double *comp,
                      const double *v_means2d, const double *v_depths, const double *v_conics,
                      const double *v_comp, double *v_means, double *v_covars, double *v_quats,
                      double *v_scales, double *v_viewmats) {
    proj_bwd<double>(C, N, means, covars, quats, scales, viewmats, Ks, W, H, eps2d, ortho, radii,
                     conics, comp, v_means2d, v_depths, v_conics, v_comp, v_means, v_covars,
                     v_quats, v_scales, v_viewmats);
}

// Pass 1: tiles per primitive and the int64 inclusive cumulative sum (what
// torch::cumsum returns in isect_tiles.cu:180).  Returns n_isects.
int64_t ora_isect_count(int64_t C, int64_t N, const float *means2d, const int32_t *radii,
                        uint32_t ts, uint32_t tw, uint32_t th, int32_t *tiles_per_gauss,
                        int64_t *cum) {
    int64_t acc = 0;
    for (int64_t idx = 0; idx < C * N; ++idx) {
        int32_t n = 0;
        if (radii[idx] > 0) {
            uint32_t x0, y0, x1, y1;
            tile_box(means2d[idx * 2], means2d[idx * 2 + 1], radii[idx], ts, tw, th, x0, y0, x1, y1);
            n = int32_t((y1 - y0) * (x1 - x0));
        }
        tiles_per_gauss[idx] = n;
        acc += n;
        cum[idx] = acc;
    }
    return acc;
}

// Pass 2 (+ optional stable sort): keys cam|tile|depth-bits and flatten ids.
void ora_isect_emit(int64_t C, int64_t N, const float *means2d, const int32_t *radii,
                    const float *depths, const int64_t *cum, uint32_t ts, uint32_t tw,
                    uint32_t th, int sort, int64_t *isect_ids, int32_t *flatten_ids) {
    const uint32_t tile_bits = bit_width(tw * th);
    const uint32_t cam_bits = bit_width((uint32_t)C);
    for (int64_t idx = 0; idx < C * N; ++idx) {
        if (radii[idx] <= 0) continue;
        uint32_t x0, y0, x1, y1;
        tile_box(means2d[idx * 2], means2d[idx * 2 + 1], radii[idx], ts, tw, th, x0, y0, x1, y1);
        const int64_t cid = idx / N;
        int32_t dbits;
        std::memcpy(&dbits, depths + idx, 4);
        const int64_t denc = (int64_t)dbits;
        int64_t cur = idx == 0 ? 0 : cum[idx - 1];
        for (uint32_t i = y0; i < y1; ++i)
            for (uint32_t j = x0; j < x1; ++j) {
                const int64_t tile = (int64_t)i * tw + j;
                isect_ids[cur] = (cid << (32 + tile_bits)) | (tile << 32) | denc;
                flatten_ids[cur] = (int32_t)idx;
                ++cur;
            }
    }
    const int64_t n = (C * N > 0) ? cum[C * N - 1] : 0;
    if (!sort || n == 0) return;
    const uint32_t end_bit = 32 + tile_bits + cam_bits;
    const uint64_t mask = end_bit >= 64 ? ~0ull : ((1ull << end_bit) - 1ull);
    std::vector<int64_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) {
        return ((uint64_t)isect_ids[a] & mask) < ((uint64_t)isect_ids[b] & mask);
    });
    std::vector<int64_t> k2(n);
    std::vector<int32_t> v2(n);
    for (int64_t i = 0; i < n; ++i) { k2[i] = isect_ids[perm[i]]; v2[i] = flatten_ids[perm[i]]; }
    std::memcpy(isect_ids, k2.data(), n * 8);
    std::memcpy(flatten_ids, v2.data(), n * 4);
}

// Stable sort of (key, value) pairs on bits [begin_bit, end_bit).
void ora_sort_pairs(int64_t n, int64_t *keys, int32_t *vals, int begin_bit, int end_bit) {
    const uint64_t hi = end_bit >= 64 ? ~0ull : ((1ull << end_bit) - 1ull);
    const uint64_t lo = begin_bit <= 0 ? 0ull : ((1ull << begin_bit) - 1ull);
    const uint64_t mask = hi & ~lo;
    std::vector<int64_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int64_t a, int64_t b) {
        return ((uint64_t)keys[a] & mask) < ((uint64_t)keys[b] & mask);
    });
    std::vector<int64_t> k2(n);
    std::vector<int32_t> v2(n);
    for (int64_t i = 0; i < n; ++i) { k2[i] = keys[perm[i]]; v2[i] = vals[perm[i]]; }
    std::memcpy(keys, k2.data(), n * 8);
    std::memcpy(vals, v2.data(), n * 4);
}

// offsets[c*T + t] = first sorted index whose (cam, tile) >= (c, t).
void ora_offsets(int64_t n, const int64_t *isect_ids, int64_t C, uint32_t tw, uint32_t th,
                 int32_t *offsets) {
    const int64_t T = (int64_t)tw * th;
    const uint32_t tile_bits = bit_width(tw * th);
    const int64_t tot = C * T;
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t hi = (int64_t)((uint64_t)isect_ids[i] >> 32);
        const int64_t id = (hi >> tile_bits) * T + (hi & ((1ll << tile_bits) - 1));
        while (k <= id && k < tot) offsets[k++] = (int32_t)i;
    }
    while (k < tot) offsets[k++] = (int32_t)n;
}

void ora_raster_fwd_f32(int64_t C, int64_t N, int64_t n_isects, int64_t D, const float *means2d,
                        const float *conics, const float *colors, const float *opac,
                        const float *betas, const float *bg, const uint8_t *masks, int32_t W,
                        int32_t H, int32_t ts, int32_t tw, int32_t th, const int32_t *offsets,
                        const int32_t *flat, float *out_c, float *out_a, int32_t *last,
                        double *margin_T, double *margin_s) {
    raster_fwd<float>(C, N, n_isects, D, means2d, conics, colors, opac, betas, bg, masks, W, H, ts,
                      tw, th, offsets, flat, out_c, out_a, last, margin_T, margin_s);
}
void ora_raster_fwd_f64(int64_t C, int64_t N, int64_t n_isects, int64_t D, const double *means2d,
                        const double *conics, const double *colors, const double *opac,
                        const double *betas, const double *bg, const uint8_t *masks, int32_t W,
                        int32_t H, int32_t ts, int32_t tw, int32_t th, const int32_t *offsets,
                        const int32_t *flat, double *out_c, double *out_a, int32_t *last,
                        double *margin_T, double *margin_s) {
    raster_fwd<double>(C, N, n_isects, D, means2d, conics, colors, opac, betas, bg, masks, W, H,
                       ts, tw, th, offsets, flat, out_c, out_a, last, margin_T, margin_s);
}
void ora_raster_bwd_f32(int64_t C, int64_t N, int64_t n_isects, int64_t D, const float *means2d,
                        const float *conics, const float *colors, const float *opac,
                        const float *betas, const float *bg, const uint8_t *masks, int32_t W,
                        int32_t H, int32_t ts, int32_t tw, int32_t th, const int32_t *offsets,
                        const int32_t *flat, const float *render_a, const int32_t *last,
                        const float *v_c, const float *v_a, const float *T_final, float *v_means2d,
                        float *v_conics, float *v_colors, float *v_opac, float *v_betas) {
    raster_bwd<float>(C, N, n_isects, D, means2d, conics, colors, opac, betas, bg, masks, W, H, ts,
                      tw, th, offsets, flat, render_a, last, v_c, v_a, T_final, v_means2d,
                      v_conics, v_colors, v_opac, v_betas);
}
void ora_raster_bwd_f64(int64_t C, int64_t N, int64_t n_isects, int64_t D, const double *means2d,
                        const double *conics, const double *colors, const double *opac,
                        const double *betas, const double *bg, const uint8_t *masks, int32_t W,
                        int32_t H, int32_t ts, int32_t tw, int32_t th, const int32_t *offsets,
                        const int32_t *flat, const double *render_a, const int32_t *last,
                        const double *v_c, const double *v_a, const double *T_final,
                        double *v_means2d, double *v_conics, double *v_colors, double *v_opac,
                        double *v_betas) {
    raster_bwd<double>(C, N, n_isects, D, means2d, conics, colors, opac, betas, bg, masks, W, H,
                       ts, tw, th, offsets, flat, render_a, last, v_c, v_a, T_final, v_means2d,
                       v_conics, v_colors, v_opac, v_betas);
}

int ora_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"

// Conditioning of a D-dim Beta primitive on its C = D-3 query dimensions
// (view direction for 6-D, view direction + time for 7-D) on gfx950.
//
// One thread per primitive, C a template parameter: every small matrix
// (v12, v22^-1, Cholesky factor, ...) lives in VGPRs with compile-time
// indices, including the pivoting Gauss-Jordan inverse used for C != 3.
// m_cond and V_cond are pure +,-,*,/ and follow the oracle's operation order
// with contraction off (bit-identical); the opacity factor uses tanh/pow and
// is compared with a tolerance.
#pragma clang fp contract(off)

#include "ubs_cond_math.h"

namespace ubs {

// reference: cond_mean_convariance_opacity_fwd.cu:106-290
template <int C>
__global__ __launch_bounds__(kThreads) void cond_fwd_kernel(int64_t N, const float *__restrict__ means, const float *__restrict__ covars,
                                const float *__restrict__ opac, const float *__restrict__ betas,
                                const float *__restrict__ query, float *__restrict__ om,
                                float *__restrict__ oV, float *__restrict__ oo) {
    constexpr int D = C + 3;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    CondRegs<C> st;
    const float *m = means + i * D;
    cond_setup<C>(m, covars + i * D * D, betas + i * C, query + i * C, st);
    cond_fwd_body<C>(st, m, opac[i], om + i * 3, oV + i * 9, oo + i);
}

// reference: cond_mean_convariance_opacity_bwd.cu:101-563
template <int C>
__global__ __launch_bounds__(kThreads) void cond_bwd_kernel(int64_t N, const float *__restrict__ means, const float *__restrict__ covars,
                                const float *__restrict__ opac, const float *__restrict__ betas,
                                const float *__restrict__ query, const float *__restrict__ vm,
                                const float *__restrict__ vV, const float *__restrict__ vo,
                                float *__restrict__ gm, float *__restrict__ gV, float *__restrict__ go,
                                float *__restrict__ gb, int exact) {
    constexpr int D = C + 3;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    CondRegs<C> st;
    cond_setup<C>(means + i * D, covars + i * D * D, betas + i * C, query + i * C, st);
    cond_bwd_body<C>(st, opac[i], vm + i * 3, vV + i * 9, vo[i], exact, gm + i * D, gV + i * D * D, go + i,
                     gb + i * C);
}

template <int C>
static void launch_cond_fwd(int64_t N, const float *m, const float *V, const float *o, const float *b,
                            const float *q, float *om, float *oV, float *oo, hipStream_t st) {
    cond_fwd_kernel<C><<<grid_for(N), kThreads, 0, st>>>(N, m, V, o, b, q, om, oV, oo);
}

template <int C>
static void launch_cond_bwd(int64_t N, const float *m, const float *V, const float *o, const float *b,
                            const float *q, const float *vm, const float *vV, const float *vo,
                            float *gm, float *gV, float *go, float *gb, int exact, hipStream_t st) {
    cond_bwd_kernel<C><<<grid_for(N), kThreads, 0, st>>>(N, m, V, o, b, q, vm, vV, vo, gm, gV, go, gb, exact);
}

// 0: the reference's gradient (default, drop-in parity); 1: the exact one.
static int g_exact_cholesky_backward = 0;

int cond_backward_exact() { return g_exact_cholesky_backward; }

}  // namespace ubs

using namespace ubs;

#define UBS_DISPATCH_C(Cv, FN, ...)                                            \
    switch (Cv) {                                                              \
    case 1: FN<1>(__VA_ARGS__); break;                                         \
    case 2: FN<2>(__VA_ARGS__); break;                                         \
    case 3: FN<3>(__VA_ARGS__); break;                                         \
    case 4: FN<4>(__VA_ARGS__); break;                                         \
    case 5: FN<5>(__VA_ARGS__); break;                                         \
    case 6: FN<6>(__VA_ARGS__); break;                                         \
    case 7: FN<7>(__VA_ARGS__); break;                                         \
    case 8: FN<8>(__VA_ARGS__); break;                                         \
    default: break;                                                            \
    }

extern "C" {

int ubs_cond_mean_convariance_opacity_fwd(int64_t N, int64_t D, const float *m, const float *V,
                                          const float *o, const float *b, const float *q, float *om,
                                          float *oV, float *oo, void *stream) {
    UBS_REQUIRE(D >= 4 && D <= 11, "D must be in [4, 11], got %lld", (long long)D);
    UBS_REQUIRE(N >= 0, "N must be >= 0");
    if (N == 0) return UBS_OK;
    UBS_REQUIRE(m && V && o && b && q && om && oV && oo, "null pointer");
    UBS_DISPATCH_C((int)(D - 3), launch_cond_fwd, N, m, V, o, b, q, om, oV, oo, as_stream(stream));
    UBS_LAUNCH_CHECK("cond_fwd_kernel");
    return UBS_OK;
}

int ubs_cond_mean_convariance_opacity_bwd(int64_t N, int64_t D, const float *m, const float *V,
                                          const float *o, const float *b, const float *q,
                                          const float *vm, const float *vV, const float *vo,
                                          float *gm, float *gV, float *go, float *gb, void *stream) {
    UBS_REQUIRE(D >= 4 && D <= 11, "D must be in [4, 11], got %lld", (long long)D);
    UBS_REQUIRE(N >= 0, "N must be >= 0");
    if (N == 0) return UBS_OK;
    UBS_REQUIRE(m && V && o && b && q && vm && vV && vo && gm && gV && go && gb, "null pointer");
    UBS_DISPATCH_C((int)(D - 3), launch_cond_bwd, N, m, V, o, b, q, vm, vV, vo, gm, gV, go, gb,
                   g_exact_cholesky_backward, as_stream(stream));
    UBS_LAUNCH_CHECK("cond_bwd_kernel");
    return UBS_OK;
}

// Chooses the opacity-path Cholesky backward of the conditioning op:
// 0 = the reference's formula (it solves with L where L^T is needed, so its
// gradient w.r.t. the conditioning covariance block is not the true one),
// 1 = the exact gradient.  Process-wide; not part of the reference interface.
int ubs_set_cond_backward_exact(int enable) {
    g_exact_cholesky_backward = enable ? 1 : 0;
    return UBS_OK;
}

}  // extern "C"

// Optimiser and MCMC step after the backward on gfx950: Adam over the Beta
// scene's leaf tensors in one launch, the covariance-shaped position noise,
// and the relocation / addition of primitives with their opacity split.
//
// Adam is torch.optim.Adam as BetaModel.training_setup configures it
// (scene/beta_model.py:239-276: one group per leaf with its own learning
// rate, betas (0.9, 0.999), eps 1e-15, no weight decay), in the operation
// order of torch's multi-tensor Adam:
//   m = m + (g - m) (1 - beta1)          (lerp, weight < 0.5 form)
//   v = v beta2;  v = v + (1 - beta2) (g g)
//   p = p + (-lr / bc1) (m / (sqrt(v) / sqrt(bc2) + eps))
// with bc1 = 1 - beta1^t, bc2 = 1 - beta2^t computed in double on the host
// per tensor (each tensor keeps its own step count, as torch's state does).
// The leaves are HBM-bound elementwise work: the launch covers all tensors
// as one flat range of 2048-element chunks (a block finds its tensor in the
// chunk prefix table), float4 loads where a tensor is 16-byte aligned; each
// element reads p, g, m, v once and writes p, m, v once (28 B).
//
// The noise is train.py:159-168: xyz += Sigma_xyz (z (1 - o)^100 scale),
// scale = noise_lr * xyz_lr, Sigma_xyz = L L^T with L = R diag(s0, s1, s2)
// the spatial block of the covariance (get_xyz_covariance,
// beta_model.py:143-152 and rot_scale_l_triangle_to_covar_fwd.cu:29-51),
// R = I + skew(l_triangle[:3]) (l_triagnle_to_rotmat_fwd.cu:16-35).  The
// normal draws z come in as a tensor, so a step is reproducible.
//
// Relocation follows relocate_gs / add_new_gs / _update_params
// (beta_model.py:553-658) given the sampled rows: every destination row
// takes its source row's leaves, the opacity split
// o' = clamp(1 - (1 - sigmoid(o_raw))^(1 / (ratio + 1)), 0.005, 1 - 2^-23)
// stored as logit(o') = log(o' / (1 - o')); then every source row takes the
// split opacity too and its Adam moments are zeroed
// (replace_tensors_to_optimizer(inds=...), :506-537).  Destination rows are
// never source rows (dead rows are not sampled; added rows are new), so the
// first launch reads sources that no thread writes; the second launch
// writes the sources (duplicates write equal values).
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>

#include "ubs_activate_math.h"

namespace ubs {

constexpr int kMaxAdamTensors = 8;
constexpr int kAdamChunk = 2048;  // elements per block: 256 threads x 2 float4

struct AdamArgs {
    float *p[kMaxAdamTensors];
    const float *g[kMaxAdamTensors];
    float *m[kMaxAdamTensors];
    float *v[kMaxAdamTensors];
    int64_t n[kMaxAdamTensors];
    int64_t blk_end[kMaxAdamTensors];  // exclusive prefix end of each tensor's chunks
    float neg_step[kMaxAdamTensors];   // -lr / bc1
    float bc2_sqrt[kMaxAdamTensors];
    int vec[kMaxAdamTensors];          // all four pointers 16-byte aligned
    int n_tensors;
    float eps, beta1c, beta2, beta2c;  // beta1c = 1 - beta1, beta2c = 1 - beta2
};

struct AdamScalars {
    float neg_step, bc2_sqrt, eps, beta1c, beta2, beta2c;
};

__device__ __forceinline__ void adam_elem(const AdamScalars &s, float &p, float g, float &m, float &v) {
    m = m + (g - m) * s.beta1c;
    v = v * s.beta2;
    v = v + s.beta2c * (g * g);
    const float denom = sqrtf(v) / s.bc2_sqrt + s.eps;
    p = p + s.neg_step * (m / denom);
}

__global__ __launch_bounds__(256) void adam_kernel(AdamArgs a) {
    const int64_t b = blockIdx.x;
    int t = 0;
    while (t + 1 < a.n_tensors && b >= a.blk_end[t]) ++t;
    const int64_t first_blk = t == 0 ? 0 : a.blk_end[t - 1];
    const int64_t base = (b - first_blk) * kAdamChunk;
    const int64_t n = a.n[t];
    float *__restrict__ P = a.p[t];
    const float *__restrict__ G = a.g[t];
    float *__restrict__ M = a.m[t];
    float *__restrict__ V = a.v[t];
    const AdamScalars s{a.neg_step[t], a.bc2_sqrt[t], a.eps, a.beta1c, a.beta2, a.beta2c};
#pragma unroll
    for (int k = 0; k < kAdamChunk / 1024; ++k) {
        const int64_t i = base + ((int64_t)k * 256 + threadIdx.x) * 4;
        if (i >= n) break;
        if (a.vec[t] && i + 4 <= n) {
            float4 p = *reinterpret_cast<const float4 *>(P + i);
            const float4 g = *reinterpret_cast<const float4 *>(G + i);
            float4 m = *reinterpret_cast<const float4 *>(M + i);
            float4 v = *reinterpret_cast<const float4 *>(V + i);
            adam_elem(s, p.x, g.x, m.x, v.x);
            adam_elem(s, p.y, g.y, m.y, v.y);
            adam_elem(s, p.z, g.z, m.z, v.z);
            adam_elem(s, p.w, g.w, m.w, v.w);
            *reinterpret_cast<float4 *>(P + i) = p;
            *reinterpret_cast<float4 *>(M + i) = m;
            *reinterpret_cast<float4 *>(V + i) = v;
        } else {
            const int64_t e = std::min<int64_t>(i + 4, n);
            for (int64_t j = i; j < e; ++j) {
                float p = P[j], m = M[j], v = V[j];
                adam_elem(s, p, G[j], m, v);
                P[j] = p;
                M[j] = m;
                V[j] = v;
            }
        }
    }
}

// xyz[n] += Sigma_xyz[n] (z[n] (1 - sigmoid(o_raw[n]))^100 scale)
__global__ __launch_bounds__(256) void xyz_noise_kernel(int64_t N, int64_t D, float *__restrict__ xyz,
                                                        const float *__restrict__ l_triangle,
                                                        const float *__restrict__ scale_raw,
                                                        const float *__restrict__ opacity_raw,
                                                        const float *__restrict__ z, float scale) {
    const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    const int64_t M = D * (D - 1) / 2;
    const float a0 = l_triangle[n * M], a1 = l_triangle[n * M + 1], a2 = l_triangle[n * M + 2];
    const float R[9] = {1.f, a0, a1, -a0, 1.f, a2, -a1, -a2, 1.f};
    float s[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] = softplus_f(scale_raw[n * D + k]);
    float L[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) L[r * 3 + c] = R[r * 3 + c] * s[c];
    float S[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            S[r * 3 + c] = (L[r * 3] * L[c * 3] + L[r * 3 + 1] * L[c * 3 + 1]) + L[r * 3 + 2] * L[c * 3 + 2];
    const float o = sigmoid_f(opacity_raw[n]);
    const float f = powf(1.f - o, 100.f);
    float w[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) w[k] = ((z[n * 3 + k] * f) * scale);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const float d = (S[r * 3] * w[0] + S[r * 3 + 1] * w[1]) + S[r * 3 + 2] * w[2];
        xyz[n * 3 + r] = xyz[n * 3 + r] + d;
    }
}

constexpr int kMaxLeaves = 8;

struct RelocArgs {
    float *leaf[kMaxLeaves];
    float *exp_avg[kMaxLeaves];
    float *exp_avg_sq[kMaxLeaves];
    int width[kMaxLeaves];
    int n_leaves, opacity_leaf;
    float min_opacity, max_opacity;
};

__device__ __forceinline__ float split_opacity_raw(float raw, int64_t ratio, float lo, float hi) {
    const float o = sigmoid_f(raw);
    const float e = 1.f / (float)(ratio + 1);
    float on = 1.f - powf(1.f - o, e);
    on = fminf(fmaxf(on, lo), hi);
    return logf(on / (1.f - on));
}

// Launch 1: one thread per (pair, leaf column) over all leaves' columns.
__global__ __launch_bounds__(256) void relocate_copy_kernel(RelocArgs a, int64_t K, int total_width,
                                                            const int64_t *__restrict__ src,
                                                            const int64_t *__restrict__ dst,
                                                            const int64_t *__restrict__ ratio) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= K * total_width) return;
    const int64_t k = i / total_width;
    int col = (int)(i - k * total_width);
    int l = 0;
    while (col >= a.width[l]) {
        col -= a.width[l];
        ++l;
    }
    const int64_t s = src[k], d = dst[k];
    const int w = a.width[l];
    const float x = a.leaf[l][s * w + col];
    a.leaf[l][d * w + col] = l == a.opacity_leaf ? split_opacity_raw(x, ratio[k], a.min_opacity, a.max_opacity) : x;
}

// Launch 2: opacity of each source row from its destination, source moments zeroed.
__global__ __launch_bounds__(256) void relocate_sources_kernel(RelocArgs a, int64_t K, int total_width,
                                                               const int64_t *__restrict__ src,
                                                               const int64_t *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= K * total_width) return;
    const int64_t k = i / total_width;
    int col = (int)(i - k * total_width);
    int l = 0;
    while (col >= a.width[l]) {
        col -= a.width[l];
        ++l;
    }
    const int64_t s = src[k];
    const int w = a.width[l];
    if (l == a.opacity_leaf) a.leaf[l][s * w + col] = a.leaf[l][dst[k] * w + col];
    if (a.exp_avg[l]) a.exp_avg[l][s * w + col] = 0.f;
    if (a.exp_avg_sq[l]) a.exp_avg_sq[l][s * w + col] = 0.f;
}

}  // namespace ubs

using namespace ubs;

extern "C" {

int ubs_adam_step(int n_tensors, float *const *params, const float *const *grads, float *const *exp_avgs,
                  float *const *exp_avg_sqs, const int64_t *sizes, const double *lrs, const int64_t *steps,
                  double beta1, double beta2, double eps, void *stream) {
    UBS_REQUIRE(n_tensors >= 0 && n_tensors <= kMaxAdamTensors, "n_tensors must be in [0, %d]", kMaxAdamTensors);
    UBS_REQUIRE(beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0, "betas must be in [0, 1)");
    UBS_REQUIRE(eps >= 0.0, "eps must be >= 0");
    if (n_tensors == 0) return UBS_OK;
    UBS_REQUIRE(params && grads && exp_avgs && exp_avg_sqs && sizes && lrs && steps, "null pointer");
    AdamArgs a{};
    int64_t blocks = 0;
    int used = 0;
    for (int t = 0; t < n_tensors; ++t) {
        UBS_REQUIRE(sizes[t] >= 0, "tensor %d: negative size", t);
        UBS_REQUIRE(steps[t] >= 1, "tensor %d: step must be >= 1 (the count after this update)", t);
        if (sizes[t] == 0) continue;
        UBS_REQUIRE(params[t] && grads[t] && exp_avgs[t] && exp_avg_sqs[t], "tensor %d: null pointer", t);
        const double bc1 = 1.0 - std::pow(beta1, (double)steps[t]);
        const double bc2 = 1.0 - std::pow(beta2, (double)steps[t]);
        a.p[used] = params[t];
        a.g[used] = grads[t];
        a.m[used] = exp_avgs[t];
        a.v[used] = exp_avg_sqs[t];
        a.n[used] = sizes[t];
        a.neg_step[used] = (float)(-(lrs[t] / bc1));
        a.bc2_sqrt[used] = (float)std::sqrt(bc2);
        a.vec[used] = ((reinterpret_cast<uintptr_t>(params[t]) | reinterpret_cast<uintptr_t>(grads[t]) |
                        reinterpret_cast<uintptr_t>(exp_avgs[t]) | reinterpret_cast<uintptr_t>(exp_avg_sqs[t])) &
                       15) == 0;
        blocks += (sizes[t] + kAdamChunk - 1) / kAdamChunk;
        a.blk_end[used] = blocks;
        ++used;
    }
    if (used == 0) return UBS_OK;
    UBS_REQUIRE(blocks <= 0x7fffffff, "too many elements for one launch");
    a.n_tensors = used;
    a.eps = (float)eps;
    a.beta1c = (float)(1.0 - beta1);
    a.beta2 = (float)beta2;
    a.beta2c = (float)(1.0 - beta2);
    adam_kernel<<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(a);
    UBS_LAUNCH_CHECK("adam_kernel");
    return UBS_OK;
}

int ubs_xyz_noise(int64_t N, int64_t D, float *xyz, const float *l_triangle, const float *scale_raw,
                  const float *opacity_raw, const float *normal, float scale, void *stream) {
    UBS_REQUIRE(N >= 0, "N must be >= 0");
    UBS_REQUIRE(D >= 3 && D <= 11, "D must be in [3, 11]");
    if (N == 0) return UBS_OK;
    UBS_REQUIRE(xyz && l_triangle && scale_raw && opacity_raw && normal, "null pointer");
    xyz_noise_kernel<<<grid_for(N), 256, 0, as_stream(stream)>>>(N, D, xyz, l_triangle, scale_raw, opacity_raw,
                                                                 normal, scale);
    UBS_LAUNCH_CHECK("xyz_noise_kernel");
    return UBS_OK;
}

int ubs_mcmc_relocate(int64_t K, int n_leaves, float *const *leaves, const int *widths, int opacity_leaf,
                      float *const *exp_avgs, float *const *exp_avg_sqs, const int64_t *src, const int64_t *dst,
                      const int64_t *ratio, float min_opacity, void *stream) {
    UBS_REQUIRE(K >= 0, "K must be >= 0");
    UBS_REQUIRE(n_leaves >= 1 && n_leaves <= kMaxLeaves, "n_leaves must be in [1, %d]", kMaxLeaves);
    UBS_REQUIRE(opacity_leaf >= 0 && opacity_leaf < n_leaves, "opacity_leaf out of range");
    UBS_REQUIRE(leaves && widths, "null pointer");
    if (K == 0) return UBS_OK;
    UBS_REQUIRE(src && dst && ratio, "null index pointer");
    RelocArgs a{};
    int total = 0;
    for (int l = 0; l < n_leaves; ++l) {
        UBS_REQUIRE(widths[l] >= 0, "leaf %d: negative width", l);
        UBS_REQUIRE(widths[l] == 0 || leaves[l], "leaf %d: null pointer", l);
        a.leaf[l] = leaves[l];
        a.exp_avg[l] = exp_avgs ? exp_avgs[l] : nullptr;
        a.exp_avg_sq[l] = exp_avg_sqs ? exp_avg_sqs[l] : nullptr;
        a.width[l] = widths[l];
        total += widths[l];
    }
    UBS_REQUIRE(widths[opacity_leaf] == 1, "the opacity leaf must have one column");
    a.n_leaves = n_leaves;
    a.opacity_leaf = opacity_leaf;
    a.min_opacity = min_opacity;
    a.max_opacity = 1.f - 1.1920928955078125e-07f;  // 1 - float32 machine epsilon
    const int64_t work = K * total;
    UBS_REQUIRE(work / 256 < 0x7fffffff, "too many rows for one launch");
    relocate_copy_kernel<<<grid_for(work), 256, 0, as_stream(stream)>>>(a, K, total, src, dst, ratio);
    UBS_LAUNCH_CHECK("relocate_copy_kernel");
    relocate_sources_kernel<<<grid_for(work), 256, 0, as_stream(stream)>>>(a, K, total, src, dst);
    UBS_LAUNCH_CHECK("relocate_sources_kernel");
    return UBS_OK;
}

}  // extern "C"

// Image loss of the training step on gfx950: (1 - lambda) L1 + lambda (1 - SSIM)
// between rendered images and targets, forward and backward.
//
// What train.py:117-121 evaluates per view with torch's l1_loss and the
// external fused_ssim, whose arithmetic is the SSIM of utils/loss_utils.py:
// 45-80: per channel, an 11x11 Gaussian window (sigma 1.5) correlated with
// zero padding, mu1 = G*x, mu2 = G*y, s11 = G*x^2 - mu1^2, s22 = G*y^2 - mu2^2,
// s12 = G*xy - mu1 mu2, S = (2 mu1 mu2 + C1)(2 s12 + C2) /
// ((mu1^2 + mu2^2 + C1)(s11 + s22 + C2)), C1 = 0.01^2, C2 = 0.03^2, and the
// loss averages S over every pixel and channel.
//
// Images are [B, H, W, C] float32 (the rasteriser's layout; no permute).
// One workgroup per 16x16 pixel tile and (image, channel): the 26x26 input
// halo is staged in LDS and the window is applied separably (rows, then
// columns).  The forward writes, per element, dL/dmu1, dL/ds11 and dL/ds12
// (the derivatives of the averaged SSIM term w.r.t. the three filtered maps
// the pixel's S depends on) into the workspace, and one partial sum of S and
// of |x - y| per workgroup; a one-block kernel adds those up in a fixed
// order.  The backward correlates the three derivative maps with the same
// window (the adjoint of a symmetric correlation) and adds the L1 term:
//   dL/dx = G*dmu1 + 2 x G*ds11 + y G*ds12 + (1 - lambda) sign(x - y) / n.
// With lambda = 0 both directions are one grid-stride float4 pass.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>

#include "ubs_common.h"

namespace ubs {

constexpr int kLossTile = 16;
constexpr int kWin = 11;
constexpr int kHalo = kWin / 2;                 // 5
constexpr int kSpan = kLossTile + 2 * kHalo;    // 26
constexpr float kC1 = 0.01f * 0.01f;
constexpr float kC2 = 0.03f * 0.03f;

struct LossLayout {
    size_t partial_off, dmaps_off, total;
    int64_t blocks;
};

static LossLayout loss_layout(int64_t B, int64_t H, int64_t W, int64_t C) {
    LossLayout L{};
    const int64_t tx = (W + kLossTile - 1) / kLossTile, ty = (H + kLossTile - 1) / kLossTile;
    L.blocks = std::max<int64_t>(tx * ty * B * C, 1);
    L.partial_off = 0;
    size_t off = (size_t)std::max<int64_t>(L.blocks, 1024) * 2 * sizeof(float);  // also the L1 pass's 1024 blocks
    off = (off + 255) & ~(size_t)255;
    L.dmaps_off = off;
    off += (size_t)3 * B * H * W * C * sizeof(float);
    L.total = (off + 255) & ~(size_t)255;
    return L;
}

// The normalised 11-tap Gaussian of loss_utils.py:26-33 (float weights of
// exp(-(k-5)^2 / 4.5), divided by their float sum).
struct GaussWin {
    float g[kWin];
};

static GaussWin gauss_window() {
    GaussWin w;
    float s = 0.f;
    for (int k = 0; k < kWin; ++k) {
        const double d = (double)(k - kHalo);
        w.g[k] = (float)std::exp(-(d * d) / (2.0 * 1.5 * 1.5));
    }
    for (int k = 0; k < kWin; ++k) s += w.g[k];
    for (int k = 0; k < kWin; ++k) w.g[k] /= s;
    return w;
}

__device__ __forceinline__ int64_t elem(int64_t b, int64_t i, int64_t j, int64_t c, int64_t H, int64_t W,
                                        int64_t C) {
    return ((b * H + i) * W + j) * C + c;
}

// Block sums of two floats (256 threads) -> thread 0.
__device__ __forceinline__ void block_sum2(float &a, float &b, float *lds) {
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[2 * w] = a;
        lds[2 * w + 1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = (lds[0] + lds[2]) + (lds[4] + lds[6]);
        b = (lds[1] + lds[3]) + (lds[5] + lds[7]);
    }
}

template <bool SSIM>
__global__ __launch_bounds__(256) void loss_fwd_kernel(int64_t B, int64_t H, int64_t W, int64_t C,
                                                       const float *__restrict__ x, const float *__restrict__ y,
                                                       GaussWin win, float lambda, float inv_n,
                                                       float *__restrict__ partial, float *__restrict__ dmaps) {
    __shared__ float sx[kSpan][kSpan + 1], sy[kSpan][kSpan + 1];
    __shared__ float h[5][kSpan][kLossTile + 1];
    __shared__ float red[8];
    const int64_t tiles_x = (W + kLossTile - 1) / kLossTile;
    const int64_t bc = blockIdx.y;  // image * C + channel
    const int64_t b = bc / C, c = bc % C;
    const int64_t i0 = (int64_t)(blockIdx.x / tiles_x) * kLossTile, j0 = (int64_t)(blockIdx.x % tiles_x) * kLossTile;
    const int tid = threadIdx.x;
    const int ti = tid / kLossTile, tj = tid % kLossTile;
    const int64_t i = i0 + ti, j = j0 + tj;
    const bool inside = i < H && j < W;
    float s_sum = 0.f, l1_sum = 0.f;
    if constexpr (SSIM) {
        for (int k = tid; k < kSpan * kSpan; k += 256) {
            const int r = k / kSpan, q = k % kSpan;
            const int64_t ii = i0 - kHalo + r, jj = j0 - kHalo + q;
            const bool ok = ii >= 0 && ii < H && jj >= 0 && jj < W;  // zero padding
            sx[r][q] = ok ? x[elem(b, ii, jj, c, H, W, C)] : 0.f;
            sy[r][q] = ok ? y[elem(b, ii, jj, c, H, W, C)] : 0.f;
        }
        __syncthreads();
        // rows: the five products filtered along j for all 26 halo rows
        for (int k = tid; k < kSpan * kLossTile; k += 256) {
            const int r = k / kLossTile, q = k % kLossTile;
            float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, a4 = 0.f;
#pragma unroll
            for (int t = 0; t < kWin; ++t) {
                const float g = win.g[t], u = sx[r][q + t], v = sy[r][q + t];
                a0 += g * u;
                a1 += g * v;
                a2 += g * (u * u);
                a3 += g * (v * v);
                a4 += g * (u * v);
            }
            h[0][r][q] = a0;
            h[1][r][q] = a1;
            h[2][r][q] = a2;
            h[3][r][q] = a3;
            h[4][r][q] = a4;
        }
        __syncthreads();
        float m1 = 0.f, m2 = 0.f, e11 = 0.f, e22 = 0.f, e12 = 0.f;
#pragma unroll
        for (int t = 0; t < kWin; ++t) {
            const float g = win.g[t];
            m1 += g * h[0][ti + t][tj];
            m2 += g * h[1][ti + t][tj];
            e11 += g * h[2][ti + t][tj];
            e22 += g * h[3][ti + t][tj];
            e12 += g * h[4][ti + t][tj];
        }
        if (inside) {
            const float mu11 = m1 * m1, mu22 = m2 * m2, mu12 = m1 * m2;
            const float s11 = e11 - mu11, s22 = e22 - mu22, s12 = e12 - mu12;
            const float A1 = 2.f * mu12 + kC1, A2 = 2.f * s12 + kC2;
            const float B1 = mu11 + mu22 + kC1, B2 = s11 + s22 + kC2;
            const float D = B1 * B2;
            const float S = (A1 * A2) / D;
            s_sum = S;
            // derivatives of S w.r.t. mu1, E[x^2] and E[xy] (loss_utils.py:56-80)
            const float dmu1 = (2.f * m2 * (A2 - A1) - 2.f * m1 * S * (B2 - B1)) / D;
            const float ds11 = -S / B2;
            const float ds12 = 2.f * A1 / D;
            const float k = -lambda * inv_n;  // d loss / d S at every element
            const int64_t e = elem(b, i, j, c, H, W, C), n_el = B * H * W * C;
            dmaps[e] = k * dmu1;
            dmaps[n_el + e] = k * ds11;
            dmaps[2 * n_el + e] = k * ds12;
        }
    }
    if (inside) l1_sum = fabsf(x[elem(b, i, j, c, H, W, C)] - y[elem(b, i, j, c, H, W, C)]);
    block_sum2(s_sum, l1_sum, red);
    if (tid == 0) {
        const int64_t blk = bc * gridDim.x + blockIdx.x;
        partial[2 * blk] = s_sum;
        partial[2 * blk + 1] = l1_sum;
    }
}

// L1 only: a grid-stride pass over float4s (plus a scalar tail), one partial
// sum of |x - y| per workgroup.
constexpr int kL1Blocks = 1024;

__global__ __launch_bounds__(256) void l1_fwd_kernel(int64_t n, const float *__restrict__ x,
                                                     const float *__restrict__ y, float *__restrict__ partial) {
    __shared__ float red[8];
    const int64_t n4 = n / 4;
    const float4 *x4 = reinterpret_cast<const float4 *>(x), *y4 = reinterpret_cast<const float4 *>(y);
    float acc = 0.f, dummy = 0.f;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n4; k += (int64_t)gridDim.x * 256) {
        const float4 a = x4[k], b = y4[k];
        acc += (fabsf(a.x - b.x) + fabsf(a.y - b.y)) + (fabsf(a.z - b.z) + fabsf(a.w - b.w));
    }
    if (blockIdx.x == 0 && threadIdx.x < n - 4 * n4) acc += fabsf(x[4 * n4 + threadIdx.x] - y[4 * n4 + threadIdx.x]);
    block_sum2(dummy, acc, red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = 0.f;
        partial[2 * blockIdx.x + 1] = acc;
    }
}

__global__ __launch_bounds__(256) void l1_bwd_kernel(int64_t n, const float *__restrict__ x,
                                                     const float *__restrict__ y, float inv_n,
                                                     const float *__restrict__ v_loss, float *__restrict__ v_x) {
    const float s = v_loss[0] * inv_n;
    auto sg = [s](float d) { return d > 0.f ? s : (d < 0.f ? -s : 0.f); };
    const int64_t n4 = n / 4;
    const float4 *x4 = reinterpret_cast<const float4 *>(x), *y4 = reinterpret_cast<const float4 *>(y);
    float4 *v4 = reinterpret_cast<float4 *>(v_x);
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n4; k += (int64_t)gridDim.x * 256) {
        const float4 a = x4[k], b = y4[k];
        v4[k] = make_float4(sg(a.x - b.x), sg(a.y - b.y), sg(a.z - b.z), sg(a.w - b.w));
    }
    if (blockIdx.x == 0 && threadIdx.x < n - 4 * n4) {
        const int64_t e = 4 * n4 + threadIdx.x;
        v_x[e] = sg(x[e] - y[e]);
    }
}

// loss = (1 - lambda) sum|x - y| / n + lambda (1 - sum S / n), the block
// partials added in a fixed order (double accumulators).
__global__ void loss_finish_kernel(int64_t blocks, const float *__restrict__ partial, float lambda, double n,
                                   int ssim, float *__restrict__ loss) {
    __shared__ double s[2][256];
    double a = 0.0, l = 0.0;
    for (int64_t k = threadIdx.x; k < blocks; k += 256) {
        a += (double)partial[2 * k];
        l += (double)partial[2 * k + 1];
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s[0][threadIdx.x] += s[0][threadIdx.x + o];
            s[1][threadIdx.x] += s[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double l1 = s[1][0] / n;
        const double ssim_term = ssim ? (1.0 - s[0][0] / n) : 0.0;
        loss[0] = (float)((1.0 - (double)lambda) * l1 + (double)lambda * ssim_term);
    }
}

template <bool SSIM>
__global__ __launch_bounds__(256) void loss_bwd_kernel(int64_t B, int64_t H, int64_t W, int64_t C,
                                                       const float *__restrict__ x, const float *__restrict__ y,
                                                       GaussWin win, float lambda, float inv_n,
                                                       const float *__restrict__ v_loss,
                                                       const float *__restrict__ dmaps, float *__restrict__ v_x) {
    __shared__ float sd[3][kSpan][kSpan + 1];
    __shared__ float h[3][kSpan][kLossTile + 1];
    const int64_t tiles_x = (W + kLossTile - 1) / kLossTile;
    const int64_t bc = blockIdx.y;
    const int64_t b = bc / C, c = bc % C;
    const int64_t i0 = (int64_t)(blockIdx.x / tiles_x) * kLossTile, j0 = (int64_t)(blockIdx.x % tiles_x) * kLossTile;
    const int tid = threadIdx.x;
    const int ti = tid / kLossTile, tj = tid % kLossTile;
    const int64_t i = i0 + ti, j = j0 + tj;
    const bool inside = i < H && j < W;
    const int64_t n_el = B * H * W * C;
    float gm = 0.f, g11 = 0.f, g12 = 0.f;
    if constexpr (SSIM) {
        for (int k = tid; k < kSpan * kSpan; k += 256) {
            const int r = k / kSpan, q = k % kSpan;
            const int64_t ii = i0 - kHalo + r, jj = j0 - kHalo + q;
            const bool ok = ii >= 0 && ii < H && jj >= 0 && jj < W;  // no S outside the image
            const int64_t e = ok ? elem(b, ii, jj, c, H, W, C) : 0;
            sd[0][r][q] = ok ? dmaps[e] : 0.f;
            sd[1][r][q] = ok ? dmaps[n_el + e] : 0.f;
            sd[2][r][q] = ok ? dmaps[2 * n_el + e] : 0.f;
        }
        __syncthreads();
        for (int k = tid; k < kSpan * kLossTile; k += 256) {
            const int r = k / kLossTile, q = k % kLossTile;
            float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
            for (int t = 0; t < kWin; ++t) {
                const float g = win.g[t];
                a0 += g * sd[0][r][q + t];
                a1 += g * sd[1][r][q + t];
                a2 += g * sd[2][r][q + t];
            }
            h[0][r][q] = a0;
            h[1][r][q] = a1;
            h[2][r][q] = a2;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kWin; ++t) {
            const float g = win.g[t];
            gm += g * h[0][ti + t][tj];
            g11 += g * h[1][ti + t][tj];
            g12 += g * h[2][ti + t][tj];
        }
    }
    if (!inside) return;
    const int64_t e = elem(b, i, j, c, H, W, C);
    const float xv = x[e], yv = y[e], dxy = xv - yv;
    const float sgn = dxy > 0.f ? 1.f : (dxy < 0.f ? -1.f : 0.f);
    float g = (1.f - lambda) * inv_n * sgn;
    if constexpr (SSIM) g += gm + 2.f * xv * g11 + yv * g12;
    v_x[e] = v_loss[0] * g;
}

}  // namespace ubs

using namespace ubs;

extern "C" {

size_t ubs_image_loss_workspace_bytes(int64_t B, int64_t H, int64_t W, int64_t C) {
    return loss_layout(B, H, W, C).total;
}

int ubs_image_loss_fwd(int64_t B, int64_t H, int64_t W, int64_t C, const float *img, const float *gt,
                       float lambda_dssim, float *loss, void *workspace, size_t workspace_bytes, void *stream) {
    UBS_REQUIRE(B >= 0 && H >= 0 && W >= 0 && C >= 0, "negative size");
    UBS_REQUIRE(lambda_dssim >= 0.f && lambda_dssim <= 1.f, "lambda_dssim must be in [0, 1]");
    UBS_REQUIRE(loss, "null loss output");
    hipStream_t st = as_stream(stream);
    const int64_t n = B * H * W * C;
    UBS_REQUIRE(n > 0, "empty image");
    UBS_REQUIRE(img && gt, "null pointer");
    const LossLayout L = loss_layout(B, H, W, C);
    if (!workspace || workspace_bytes < L.total) {
        set_error("loss workspace too small (%zu < %zu)", workspace_bytes, L.total);
        return UBS_ERR_WORKSPACE;
    }
    const int64_t tiles = ((W + kLossTile - 1) / kLossTile) * ((H + kLossTile - 1) / kLossTile);
    UBS_REQUIRE(tiles < (1ll << 31) && B * C < 65536, "image batch too large");
    char *ws = static_cast<char *>(workspace);
    float *partial = reinterpret_cast<float *>(ws + L.partial_off);
    float *dmaps = reinterpret_cast<float *>(ws + L.dmaps_off);
    const bool ssim = lambda_dssim > 0.f;
    const bool aligned = (reinterpret_cast<uintptr_t>(img) % 16 == 0) && (reinterpret_cast<uintptr_t>(gt) % 16 == 0);
    if (!ssim && aligned) {
        const unsigned nb = (unsigned)std::min<int64_t>(kL1Blocks, std::max<int64_t>(1, (n / 4 + 255) / 256));
        l1_fwd_kernel<<<nb, 256, 0, st>>>(n, img, gt, partial);
        UBS_LAUNCH_CHECK("l1_fwd_kernel");
        loss_finish_kernel<<<1, 256, 0, st>>>(nb, partial, lambda_dssim, (double)n, 0, loss);
        UBS_LAUNCH_CHECK("loss_finish_kernel");
        return UBS_OK;
    }
    const dim3 grid((unsigned)tiles, (unsigned)(B * C));
    const GaussWin win = gauss_window();
    if (ssim)
        loss_fwd_kernel<true><<<grid, 256, 0, st>>>(B, H, W, C, img, gt, win, lambda_dssim, 1.f / (float)n, partial,
                                                    dmaps);
    else
        loss_fwd_kernel<false><<<grid, 256, 0, st>>>(B, H, W, C, img, gt, win, lambda_dssim, 1.f / (float)n,
                                                     partial, dmaps);
    UBS_LAUNCH_CHECK("loss_fwd_kernel");
    loss_finish_kernel<<<1, 256, 0, st>>>(L.blocks, partial, lambda_dssim, (double)n, ssim ? 1 : 0, loss);
    UBS_LAUNCH_CHECK("loss_finish_kernel");
    return UBS_OK;
}

int ubs_image_loss_bwd(int64_t B, int64_t H, int64_t W, int64_t C, const float *img, const float *gt,
                       float lambda_dssim, const float *v_loss, const void *workspace, size_t workspace_bytes,
                       float *v_img, void *stream) {
    UBS_REQUIRE(B >= 0 && H >= 0 && W >= 0 && C >= 0, "negative size");
    UBS_REQUIRE(lambda_dssim >= 0.f && lambda_dssim <= 1.f, "lambda_dssim must be in [0, 1]");
    hipStream_t st = as_stream(stream);
    const int64_t n = B * H * W * C;
    UBS_REQUIRE(n > 0, "empty image");
    UBS_REQUIRE(img && gt && v_loss && v_img, "null pointer");
    const LossLayout L = loss_layout(B, H, W, C);
    if (!workspace || workspace_bytes < L.total) {
        set_error("loss workspace too small (%zu < %zu)", workspace_bytes, L.total);
        return UBS_ERR_WORKSPACE;
    }
    const int64_t tiles = ((W + kLossTile - 1) / kLossTile) * ((H + kLossTile - 1) / kLossTile);
    UBS_REQUIRE(tiles < (1ll << 31) && B * C < 65536, "image batch too large");
    const float *dmaps = reinterpret_cast<const float *>(static_cast<const char *>(workspace) + L.dmaps_off);
    const bool aligned = (reinterpret_cast<uintptr_t>(img) % 16 == 0) && (reinterpret_cast<uintptr_t>(gt) % 16 == 0) &&
                         (reinterpret_cast<uintptr_t>(v_img) % 16 == 0);
    if (lambda_dssim == 0.f && aligned) {
        const unsigned nb = (unsigned)std::min<int64_t>(kL1Blocks, std::max<int64_t>(1, (n / 4 + 255) / 256));
        l1_bwd_kernel<<<nb, 256, 0, st>>>(n, img, gt, 1.f / (float)n, v_loss, v_img);
        UBS_LAUNCH_CHECK("l1_bwd_kernel");
        return UBS_OK;
    }
    const dim3 grid((unsigned)tiles, (unsigned)(B * C));
    const GaussWin win = gauss_window();
    if (lambda_dssim > 0.f)
        loss_bwd_kernel<true><<<grid, 256, 0, st>>>(B, H, W, C, img, gt, win, lambda_dssim, 1.f / (float)n, v_loss,
                                                    dmaps, v_img);
    else
        loss_bwd_kernel<false><<<grid, 256, 0, st>>>(B, H, W, C, img, gt, win, lambda_dssim, 1.f / (float)n, v_loss,
                                                     dmaps, v_img);
    UBS_LAUNCH_CHECK("loss_bwd_kernel");
    return UBS_OK;
}

}  // extern "C"

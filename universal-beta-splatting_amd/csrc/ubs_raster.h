// Shared pieces of the forward and backward tile blend kernels.
#pragma once

#include "ubs_common.h"

namespace ubs {

// Beta kernel weight (1 - sigma)^beta with the hardware log2/exp2
// (v_log_f32 / v_exp_f32).  The forward and the backward call this same
// function, so the backward recovers transmittance from bit-identical alphas.
__device__ __forceinline__ float beta_pow(float one_minus_sigma, float beta) {
    return __builtin_amdgcn_exp2f(beta * __builtin_amdgcn_logf(one_minus_sigma));
}

// sigma = (a dx^2 + c dy^2) + 2b dx dy, the reference's grouping
// (rasterize_to_pixels_fwd.cu:146-148); b2 = 2 b is exact.
__device__ __forceinline__ float beta_sigma(float a, float b2, float c, float dx, float dy) {
    return (a * dx * dx + c * dy * dy) + b2 * dx * dy;
}

// Workgroup -> (camera, tile).  Workgroups are dispatched round-robin over
// the 8 XCDs (workgroup i runs on XCD i % 8), each with its own L2.  Tiles are
// grouped into blocks of kXcdBlockX x kXcdBlockY, and the blocks are dealt
// round-robin to the XCDs: every XCD gets an even share of an image whose
// work is concentrated in its centre (contiguous bands per XCD left the
// middle XCDs with 1.5x the average at 1080p), while the tiles of one block,
// which share most of their primitives, stay in one L2 (4x2, 2x4 and 4x4
// blocks measured the same as 2x2 at 1080p / 1M).  With an order (below)
// each XCD takes its own tiles heaviest first.
constexpr int kNumXcd = 8;
#ifndef UBS_XCD_BLOCK_X
#define UBS_XCD_BLOCK_X 2
#endif
#ifndef UBS_XCD_BLOCK_Y
#define UBS_XCD_BLOCK_Y 2
#endif
constexpr uint32_t kXcdBlockX = UBS_XCD_BLOCK_X, kXcdBlockY = UBS_XCD_BLOCK_Y;
constexpr uint32_t kXcdBlockTiles = kXcdBlockX * kXcdBlockY;

__host__ __device__ inline uint32_t xcd_blocks(uint32_t C, uint32_t tw, uint32_t th, uint32_t &nbx) {
    nbx = (tw + kXcdBlockX - 1) / kXcdBlockX;
    const uint32_t nby = (th + kXcdBlockY - 1) / kXcdBlockY;
    return C * nbx * nby;
}

// Number of workgroups to launch (some exit at once: block padding).
__host__ __device__ inline uint32_t xcd_grid(uint32_t C, uint32_t tw, uint32_t th) {
    uint32_t nbx;
    const uint32_t blocks = xcd_blocks(C, tw, th, nbx);
    return (blocks + kNumXcd - 1) / kNumXcd * kNumXcd * kXcdBlockTiles;
}

// Tile slots per XCD: the most tiles any XCD is dealt.
__host__ __device__ inline uint32_t xcd_slots(uint32_t C, uint32_t tw, uint32_t th) {
    uint32_t nbx;
    return (xcd_blocks(C, tw, th, nbx) + kNumXcd - 1) / kNumXcd * kXcdBlockTiles;
}

// XCD a tile is dealt to: that of its block.
__host__ __device__ inline uint32_t xcd_of_tile(uint32_t t, uint32_t C, uint32_t tw, uint32_t th) {
    uint32_t nbx;
    const uint32_t blocks = xcd_blocks(C, tw, th, nbx);
    const uint32_t per_cam = blocks / C, cam = t / (tw * th), rem = t % (tw * th);
    const uint32_t k = cam * per_cam + (rem / tw) / kXcdBlockY * nbx + (rem % tw) / kXcdBlockX;
    return k % kNumXcd;
}

// logical = cam * th * tw + ty * tw + tx of workgroup bid, or false.
// order (may be null): per XCD, its tiles in the order to run them (the j-th
// workgroup on XCD x runs tile order[x * xcd_slots + j]; entries past an
// XCD's tiles are ~0u); raster_block_order_kernel writes them heaviest
// first.  Null runs each XCD's blocks in raster order.
__host__ __device__ inline bool xcd_tile(uint32_t bid, uint32_t C, uint32_t tw, uint32_t th, uint32_t &logical,
                                         const uint32_t *order = nullptr) {
    if (order) {
        const uint32_t slots = xcd_slots(C, tw, th), j = bid / kNumXcd;
        if (j >= slots) return false;
        const uint32_t t = order[(size_t)(bid % kNumXcd) * slots + j];
        if (t >= C * tw * th) return false;  // a padding entry
        logical = t;
        return true;
    }
    uint32_t nbx;
    const uint32_t blocks = xcd_blocks(C, tw, th, nbx);
    const uint32_t j = bid / kNumXcd;
    const uint32_t k = bid % kNumXcd + kNumXcd * (j / kXcdBlockTiles);  // block
    const uint32_t r = j % kXcdBlockTiles;                              // tile inside it
    if (k >= blocks) return false;
    const uint32_t per_cam = blocks / C;
    const uint32_t cam = k / per_cam, kb = k % per_cam;
    const uint32_t tx = (kb % nbx) * kXcdBlockX + r % kXcdBlockX;
    const uint32_t ty = (kb / nbx) * kXcdBlockY + r / kXcdBlockX;
    if (tx >= tw || ty >= th) return false;
    logical = (cam * th + ty) * tw + tx;
    return true;
}

// Heaviest tiles first, per XCD.  A tile's blend time grows with its list,
// and the lists of a centre-heavy image differ by ~3.5x between the mean and
// the longest; in raster order the heaviest tiles can start last and run
// alone at the end of the launch.  One workgroup buckets every tile by its
// work (list entries + kTileCost) into 256 buckets and writes, for each XCD,
// the tiles its blocks hold, heaviest bucket first (longest-processing-time-
// first list scheduling, to the bucket width, with each XCD keeping its own
// tiles and so its L2 contents); the order inside a bucket is whatever the
// atomics give (the blend's results do not depend on the order its tiles
// run in).  n_dev as in the blend.
constexpr uint32_t kOrderBuckets = 256;
constexpr uint32_t kOrderThreads = 1024;
#ifndef UBS_ORDER_TILE_COST
#define UBS_ORDER_TILE_COST 64
#endif
constexpr uint32_t kTileCost = UBS_ORDER_TILE_COST;

__host__ __device__ inline size_t raster_order_bytes(uint32_t C, uint32_t tw, uint32_t th) {
    return ((size_t)kNumXcd * xcd_slots(C, tw, th) * sizeof(uint32_t) + 255) & ~(size_t)255;
}

static __global__ __launch_bounds__(kOrderThreads) void raster_block_order_kernel(
    uint32_t C, uint32_t tw, uint32_t th, const int32_t *__restrict__ tile_offsets, int64_t n_isects,
    const int64_t *__restrict__ n_dev, uint32_t *__restrict__ order) {
    // the first kHeld x kOrderThreads tiles stay in registers between the
    // passes (their offsets loaded together); tiles past them are read again
    constexpr uint32_t kHeld = 8;
    constexpr uint32_t kPerXcd = kOrderThreads / kNumXcd;     // 128 threads scan one XCD's buckets
    constexpr uint32_t kPerThread = kOrderBuckets / kPerXcd;  // 2 buckets each
    __shared__ uint32_t s_count[kNumXcd][kOrderBuckets];
    __shared__ uint32_t s_total[kNumXcd];
    __shared__ uint32_t s_wave_sum[kOrderThreads / 64];
    __shared__ uint32_t s_max;
    if (n_dev) n_isects = min(max(*n_dev, (int64_t)0), n_isects);
    const uint32_t tiles = C * tw * th, slots = xcd_slots(C, tw, th);
    auto work_of = [&](uint32_t t) -> uint32_t {
        const int64_t a = tile_offsets[t], b = t + 1 < tiles ? (int64_t)tile_offsets[t + 1] : n_isects;
        return kTileCost + (uint32_t)min(max(b - a, (int64_t)0), (int64_t)(1u << 24));
    };
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    for (uint32_t q = tid; q < kNumXcd * kOrderBuckets; q += kOrderThreads) (&s_count[0][0])[q] = 0;
    if (tid == 0) s_max = 1;
    uint32_t held[kHeld];
#pragma unroll
    for (uint32_t u = 0; u < kHeld; ++u) {
        const uint32_t t = tid + u * kOrderThreads;
        held[u] = t < tiles ? work_of(t) : 0u;
    }
    uint32_t mx = 0;
#pragma unroll
    for (uint32_t u = 0; u < kHeld; ++u) mx = max(mx, held[u]);
    for (uint32_t t = tid + kHeld * kOrderThreads; t < tiles; t += kOrderThreads) mx = max(mx, work_of(t));
    __syncthreads();  // (s_count and s_max initialised)
    atomicMax(&s_max, mx);
    __syncthreads();
    const uint64_t wmax = s_max;
    auto bucket = [&](uint32_t w) -> uint32_t {  // kOrderBuckets - 1 = heaviest
        return (uint32_t)min((uint64_t)w * kOrderBuckets / (wmax + 1), (uint64_t)kOrderBuckets - 1);
    };
    uint32_t key[kHeld];  // xcd << 8 | bucket of the held tiles
#pragma unroll
    for (uint32_t u = 0; u < kHeld; ++u) {
        const uint32_t t = tid + u * kOrderThreads;
        key[u] = t < tiles ? (xcd_of_tile(t, C, tw, th) << 8) | bucket(held[u]) : 0u;
        if (t < tiles) atomicAdd(&s_count[key[u] >> 8][key[u] & 255u], 1u);
    }
    for (uint32_t t = tid + kHeld * kOrderThreads; t < tiles; t += kOrderThreads)
        atomicAdd(&s_count[xcd_of_tile(t, C, tw, th)][bucket(work_of(t))], 1u);
    __syncthreads();
    // first slot of each (XCD, bucket), heaviest bucket at slot 0: the 128
    // threads of XCD x take its buckets two at a time from the heaviest down
    const uint32_t x = tid / kPerXcd, i = tid % kPerXcd;
    uint32_t c[kPerThread], own = 0;
#pragma unroll
    for (uint32_t u = 0; u < kPerThread; ++u) {
        c[u] = s_count[x][kOrderBuckets - 1 - (kPerThread * i + u)];
        own += c[u];
    }
    uint32_t incl = own;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(incl, o, 64);
        if ((int)lane >= o) incl += v;
    }
    if (lane == 63) s_wave_sum[tid >> 6] = incl;
    __syncthreads();
    uint32_t at = incl - own;
    for (uint32_t w = x * (kPerXcd / 64); w < (tid >> 6); ++w) at += s_wave_sum[w];
#pragma unroll
    for (uint32_t u = 0; u < kPerThread; ++u) {
        s_count[x][kOrderBuckets - 1 - (kPerThread * i + u)] = at;
        at += c[u];
    }
    if (i == kPerXcd - 1) s_total[x] = at;  // the XCD's tile count
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < kHeld; ++u) {
        const uint32_t t = tid + u * kOrderThreads;
        if (t < tiles) {
            const uint32_t xt = key[u] >> 8;
            order[(size_t)xt * slots + atomicAdd(&s_count[xt][key[u] & 255u], 1u)] = t;
        }
    }
    for (uint32_t t = tid + kHeld * kOrderThreads; t < tiles; t += kOrderThreads) {
        const uint32_t xt = xcd_of_tile(t, C, tw, th);
        order[(size_t)xt * slots + atomicAdd(&s_count[xt][bucket(work_of(t))], 1u)] = t;
    }
    for (uint32_t q = tid; q < kNumXcd * slots; q += kOrderThreads)  // slots past each XCD's tiles
        if (q % slots >= s_total[q / slots]) order[q] = ~0u;
}

// Half extents of the sigma < 1 ellipse of conic [[a, b], [b, c]]:
// |dx| <= sqrt(c / det), |dy| <= sqrt(a / det), padded by 0.1% + 1e-3 px so
// that every pixel outside the box has fp32 sigma >= 1 (culling never drops
// a pair the per-pixel test would keep).
__device__ __forceinline__ void support_box(float a, float b, float c, float &hx, float &hy) {
    const float det = a * c - b * b;
    if (det > 0.f) {
        hx = sqrtf(c / det) * 1.001f + 1e-3f;
        hy = sqrtf(a / det) * 1.001f + 1e-3f;
    } else {
        hx = __builtin_huge_valf();
        hy = __builtin_huge_valf();
    }
}

// Bit (sy * SB + sx) set when the padded support box of a primitive centred
// at (mx, my) reaches a pixel centre of 8x8 sub-block (sx, sy) of the tile
// whose first pixel is (x0, y0).
template <int SB>
__device__ __forceinline__ uint32_t subblock_mask(float mx, float my, float hx, float hy, float x0, float y0) {
    uint32_t mx_bits = 0, my_bits = 0;
#pragma unroll
    for (int s = 0; s < SB; ++s) {
        const float lo_x = x0 + (float)(s * 8) + 0.5f, hi_x = lo_x + 7.f;
        const float lo_y = y0 + (float)(s * 8) + 0.5f, hi_y = lo_y + 7.f;
        if (mx + hx >= lo_x && mx - hx <= hi_x) mx_bits |= 1u << s;
        if (my + hy >= lo_y && my - hy <= hi_y) my_bits |= 1u << s;
    }
    uint32_t m = 0;
#pragma unroll
    for (int sy = 0; sy < SB; ++sy)
#pragma unroll
        for (int sx = 0; sx < SB; ++sx)
            if (((my_bits >> sy) & 1u) && ((mx_bits >> sx) & 1u)) m |= 1u << (sy * SB + sx);
    return m;
}

// Whether the sigma < 1 ellipse of conic [[a, b], [b, c]] centred at (mx, my)
// reaches the rectangle [lo_x, hi_x] x [lo_y, hi_y] of pixel centres.  The
// quadratic form is convex, so when the centre lies outside the rectangle
// its minimum over the rectangle lies on one of the four edges, where it is
// a 1-D quadratic minimised in closed form and clamped to the edge.  The
// threshold carries a margin of 1e-3 plus 1e-5 of the largest absolute term
// sum over the rectangle, far above the rounding of the per-pixel fp32
// sigma, so a pixel whose sigma the blend would accept is never culled.
__host__ __device__ inline bool ellipse_meets_rect(float mx, float my, float a, float b, float c, float lo_x,
                                                   float hi_x, float lo_y, float hi_y) {
    const float dx0 = mx - hi_x, dx1 = mx - lo_x;  // offsets mean - pixel over the rectangle
    const float dy0 = my - hi_y, dy1 = my - lo_y;
    if (dx0 <= 0.f && dx1 >= 0.f && dy0 <= 0.f && dy1 >= 0.f) return true;
    if (!(a > 0.f && c > 0.f)) return true;
    const float ia = 1.f / a, ic = 1.f / c;
    float best = __builtin_huge_valf();
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const float dx = e ? dx1 : dx0;
        const float dy = fminf(fmaxf(-b * dx * ic, dy0), dy1);
        best = fminf(best, (a * dx * dx + c * dy * dy) + 2.f * b * dx * dy);
    }
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const float dy = e ? dy1 : dy0;
        const float dx = fminf(fmaxf(-b * dy * ia, dx0), dx1);
        best = fminf(best, (a * dx * dx + c * dy * dy) + 2.f * b * dx * dy);
    }
    const float mxx = fmaxf(dx0 * dx0, dx1 * dx1), myy = fmaxf(dy0 * dy0, dy1 * dy1);
    const float sabs = (a * mxx + c * myy) + 2.f * fabsf(b) * sqrtf(mxx * myy);
    return best < 1.001f + 1e-5f * sabs;
}

// subblock_mask refined by the exact ellipse test for every sub-block the
// support box reaches.
template <int SB>
__device__ __forceinline__ uint32_t subblock_mask_exact(float mx, float my, float a, float b, float c, float hx,
                                                        float hy, float x0, float y0) {
    uint32_t m = subblock_mask<SB>(mx, my, hx, hy, x0, y0);
    if (!(hx < __builtin_huge_valf())) return m;
    uint32_t out = m;
#pragma unroll
    for (int sy = 0; sy < SB; ++sy)
#pragma unroll
        for (int sx = 0; sx < SB; ++sx) {
            const uint32_t bit = 1u << (sy * SB + sx);
            if (!(m & bit)) continue;
            const float lo_x = x0 + (float)(sx * 8) + 0.5f, lo_y = y0 + (float)(sy * 8) + 0.5f;
            if (!ellipse_meets_rect(mx, my, a, b, c, lo_x, lo_x + 7.f, lo_y, lo_y + 7.f)) out &= ~bit;
        }
    return out;
}

// Packed per-(camera, primitive) records of the blend: r0 = (x, y, a, b),
// r1 = (c, opacity, beta, 0) for the forward inputs, and the same layout for
// the backward's gradients (v_x, v_y, v_a, v_b | v_c, v_opacity, v_beta, 0).
// One record per intersection gather instead of five arrays.  With at most
// 4 colour channels blended in one pass the record is "wide": 64 bytes, r2 =
// the colours (zero-padded to 4) and r3 unused, so one aligned 64-byte read
// (and, for the gradients, one 64-byte line of atomics) serves a primitive
// where the narrow 32-byte record and a separate 12-byte colour gather touch
// two lines.
constexpr size_t kRecordFloats = 8;
constexpr size_t kWideRecordFloats = 16;

inline size_t raster_workspace_bytes(uint32_t C, uint32_t N, int backward) {
    const size_t one = ((size_t)C * N * kWideRecordFloats * sizeof(float) + 255) & ~(size_t)255;
    return backward ? 2 * one : one;
}

static __global__ void raster_pack_kernel(uint64_t CN, const float *__restrict__ means2d,
                                          const float *__restrict__ conics, const float *__restrict__ opacities,
                                          const float *__restrict__ betas, float4 *__restrict__ rec,
                                          const float *__restrict__ colors = nullptr, uint32_t channels = 0) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= CN) return;
    const float2 xy = reinterpret_cast<const float2 *>(means2d)[i];
    const float4 r0 = make_float4(xy.x, xy.y, conics[3 * i], conics[3 * i + 1]);
    const float4 r1 = make_float4(conics[3 * i + 2], opacities[i], betas[i], 0.f);
    if (colors) {  // wide
        const float *cp = colors + i * channels;
        rec[4 * i] = r0;
        rec[4 * i + 1] = r1;
        rec[4 * i + 2] = make_float4(cp[0], channels > 1 ? cp[1] : 0.f, channels > 2 ? cp[2] : 0.f,
                                     channels > 3 ? cp[3] : 0.f);
    } else {
        rec[2 * i] = r0;
        rec[2 * i + 1] = r1;
    }
}

static __global__ void raster_unpack_grads_kernel(uint64_t CN, const float4 *__restrict__ vrec,
                                                  float *__restrict__ v_means2d, float *__restrict__ v_conics,
                                                  float *__restrict__ v_opacities, float *__restrict__ v_betas,
                                                  float *__restrict__ v_colors = nullptr, uint32_t channels = 0) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= CN) return;
    const uint64_t stride = v_colors ? 4 : 2;  // float4s per record: wide or narrow
    const float4 r0 = vrec[stride * i], r1 = vrec[stride * i + 1];
    reinterpret_cast<float2 *>(v_means2d)[i] = make_float2(r0.x, r0.y);
    v_conics[3 * i] = r0.z;
    v_conics[3 * i + 1] = r0.w;
    v_conics[3 * i + 2] = r1.x;
    v_opacities[i] = r1.y;
    v_betas[i] = r1.z;
    if (v_colors) {
        const float4 r2 = vrec[stride * i + 2];
        const float c4[4] = {r2.x, r2.y, r2.z, r2.w};
        for (uint32_t k = 0; k < channels; ++k) v_colors[i * channels + k] = c4[k];
    }
}

// Transmittance checkpoints (optional): the forward stores every pixel's T
// after each batch of BS = TS^2 entries of its tile's list, the backward
// starts each batch from that value instead of carrying the division
// recovery T / (1 - alpha) across batches.  Batch k of the tile with global
// index L (camera * th * tw + tile) whose list starts at isect index s has
// slot floor(s / BS) + L + k, BS floats wide (one per pixel of the tile).
// Slots are unique: the last slot of a tile [s, e) is at most
// floor(e / BS) + L, one below the next tile's first, floor(e / BS) + L + 1,
// since floor(s / BS) + ceil((e - s) / BS) <= floor(e / BS) + 1.  So
// floor(n_isects / BS) + C th tw slots cover every tile.
__host__ __device__ inline int64_t ckpt_slot(int32_t range_start, uint32_t tile_global, int32_t k, int BS) {
    return (int64_t)(range_start / BS) + (int64_t)tile_global + k;
}

inline size_t ckpt_floats(uint32_t C, int64_t n_isects, uint32_t tw, uint32_t th, uint32_t tile_size) {
    const size_t BS = (size_t)tile_size * tile_size;
    return ((size_t)n_isects / BS + (size_t)C * tw * th + 1) * BS;
}

// Channel chunks the kernels are instantiated for; wider colour vectors are
// blended chunk by chunk (the transmittance does not depend on colour).  The
// widest chunk is bounded by LDS: 32 channels for 8x8 and 16x16 tiles, 4 for
// 32x32 tiles.
constexpr int kChunkSizes[] = {32, 16, 8, 4, 3, 2, 1};

inline int pick_chunk(uint32_t remaining, uint32_t tile_size) {
    const uint32_t cap = tile_size == 32 ? 4u : 32u;
    for (int k : kChunkSizes)
        if ((uint32_t)k <= remaining && (uint32_t)k <= cap) return k;
    return 1;
}

// Whether the blend of `channels` channels uses wide records (one pass over
// all channels, at most 4).
inline bool raster_wide_records(uint32_t channels, uint32_t tile_size) {
    return channels <= 4 && (int)channels == pick_chunk(channels, tile_size);
}

}  // namespace ubs

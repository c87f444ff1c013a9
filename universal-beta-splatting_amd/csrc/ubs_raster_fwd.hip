// Forward tile blend of Beta primitives on gfx950.
//
// One workgroup per (camera, tile), TS x TS threads.  Each wave64 owns an 8x8
// pixel sub-block of the tile (lane l -> pixel (l & 7, l >> 3) of it), which
// is the unit of culling: a batch of TS*TS primitives is staged in LDS, every
// primitive gets a bit mask of the sub-blocks its sigma < 1 ellipse reaches
// (bounding-box test refined by an exact ellipse/rectangle test), and each
// wave compacts the batch to the primitives that touch its sub-block (ballot
// + popcount) before the per-pixel front-to-back loop.
// Pixels stop at the first primitive that would take transmittance to
// <= 1e-4 (exclusive), as rasterize_to_pixels_fwd.cu:150-172 does; a wave
// leaves the batch as soon as all of its pixels are done and the workgroup
// stops loading batches once every pixel is done.
#pragma clang fp contract(off)

#include "ubs_raster.h"

// Read two compacted-list entries and both their records before blending
// either (1), or one entry at a time (0).
#ifndef UBS_FWD_PAIRED_READS
#define UBS_FWD_PAIRED_READS 1
#endif

namespace ubs {

// STATS (measurement builds only; the library's kernels have it off) counts
// into stats[0..4]: (wave, primitive) steps of the blend loop, lane
// evaluations of a pixel still blending, of those inside the primitive's
// support (0 <= sigma < 1), of those that contribute, and staged entries.
template <int TS, int CDIM, bool STATS = false>
__global__ __launch_bounds__(TS *TS) void raster_fwd_kernel(
    uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels, uint32_t coff,
    const float *__restrict__ means2d, const float *__restrict__ conics, const float *__restrict__ colors,
    const float *__restrict__ opacities, const float *__restrict__ betas, const float *__restrict__ backgrounds,
    const uint8_t *__restrict__ masks, uint32_t W, uint32_t H, uint32_t tw, uint32_t th,
    const int32_t *__restrict__ tile_offsets, const int32_t *__restrict__ flatten_ids,
    float *__restrict__ render_colors, float *__restrict__ render_alphas, int32_t *__restrict__ last_ids,
    float *__restrict__ T_final, const float4 *__restrict__ rec, int wide_rec, float *__restrict__ T_ckpt,
    unsigned long long *__restrict__ stats = nullptr, const int64_t *__restrict__ n_dev = nullptr,
    const uint32_t *__restrict__ order = nullptr, unsigned long long *__restrict__ wg_times = nullptr) {
    // n_dev: the intersection count in device memory (at most n_isects, the
    // capacity of the lists), for callers that size the lists on the device
    if (n_dev) n_isects = min(max(*n_dev, (int64_t)0), n_isects);
    constexpr int BS = TS * TS;        // threads = batch size
    constexpr int SB = TS / 8;         // sub-blocks per tile side
    constexpr int NW = BS / 64;        // waves
    __shared__ float4 s_geo[BS];       // x, y, a, 2b
    __shared__ float4 s_geo2[BS];      // c, opacity, beta, (unused)
    // up to 4 channels: one float4 per primitive, addressed with the same
    // byte offset as s_geo / s_geo2 (no per-primitive address arithmetic)
    constexpr bool kColVec = CDIM <= 4;
    __shared__ float4 s_colv[kColVec ? BS : 1];
    __shared__ float s_col[kColVec ? 1 : CDIM][kColVec ? 1 : BS];
    __shared__ uint32_t s_mask[BS];
    // byte offsets (16 x entry) of the wave's primitives; 32-bit so that a
    // pair of them is one 64-bit read with nothing to unpack
    __shared__ __align__(16) uint32_t s_list[NW][BS];

    uint32_t logical;
    if (!xcd_tile(blockIdx.x, C, tw, th, logical, order)) return;
    // (measurement builds: the workgroup's start on the constant-rate clock)
    unsigned long long t_start = 0;
    if constexpr (STATS) t_start = __builtin_amdgcn_s_memrealtime();
    const uint32_t cam = logical / (th * tw);
    const uint32_t tile = logical % (th * tw);
    const uint32_t ty = tile / tw, tx = tile % tw;

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int sx = wave % SB, sy = wave / SB;
    const uint32_t i = ty * TS + sy * 8 + (lane >> 3);
    const uint32_t j = tx * TS + sx * 8 + (lane & 7);
    const bool inside = i < H && j < W;
    const int64_t pix = (int64_t)cam * H * W + (int64_t)i * W + j;
    const bool write_meta = (coff == 0);

    if (masks && !masks[(size_t)cam * th * tw + tile]) {
        if (inside) {
#pragma unroll
            for (int k = 0; k < CDIM; ++k)
                render_colors[pix * channels + coff + k] = backgrounds ? backgrounds[cam * channels + coff + k] : 0.f;
            if (write_meta) {
                render_alphas[pix] = 0.f;
                last_ids[pix] = 0;
                if (T_final) T_final[pix] = 1.f;
            }
        }
        return;
    }

    const int64_t tbase = (int64_t)cam * th * tw;
    // a valid sort gives ordered offsets within [0, n_isects]; the clamps keep
    // the lists of a sort that gave up on a look-back wait inside the buffers
    const int32_t off_hi = (cam == C - 1 && tile == th * tw - 1) ? (int32_t)n_isects : tile_offsets[tbase + tile + 1];
    const int32_t range_start = min(max(tile_offsets[tbase + tile], 0), (int32_t)n_isects);
    const int32_t range_end = min(max(off_hi, range_start), (int32_t)n_isects);
    const float px = (float)j + 0.5f, py = (float)i + 0.5f;
    const float x0 = (float)(tx * TS), y0 = (float)(ty * TS);

    float T = 1.f;
    int32_t cur = 0;
    bool done = !inside;
    float acc[CDIM];
#pragma unroll
    for (int k = 0; k < CDIM; ++k) acc[k] = 0.f;

    uint32_t st_steps = 0, st_live = 0, st_sup = 0, st_take = 0, st_staged = 0;
    const int32_t nb = range_end > range_start ? (range_end - range_start + BS - 1) / BS : 0;
    for (int32_t b = 0; b < nb; ++b) {
        if (__syncthreads_count(done) >= BS) break;
        const int32_t batch_start = range_start + b * BS;
        const int32_t cnt = min(BS, range_end - batch_start);
        if constexpr (STATS) st_staged += (uint32_t)(tid < cnt);
        // ---- stage one primitive per thread
        if (tid < cnt) {
            // an id outside [0, C N) (only from a sort that gave up) stages
            // primitive 0 with an empty sub-block mask: never blended
            const int32_t g_raw = flatten_ids[batch_start + tid];
            const bool g_ok = (uint64_t)(uint32_t)g_raw < (uint64_t)C * N;
            const int32_t g = g_ok ? g_raw : 0;
            float2 xy;
            float ca, cb, cc, op, be;
            float4 cw = make_float4(0.f, 0.f, 0.f, 0.f);
            if (rec && wide_rec) {  // one 64-byte record per (camera, primitive), colours included
                const float4 r0 = rec[4 * (int64_t)g], r1 = rec[4 * (int64_t)g + 1];
                cw = rec[4 * (int64_t)g + 2];
                xy = make_float2(r0.x, r0.y);
                ca = r0.z; cb = r0.w; cc = r1.x; op = r1.y; be = r1.z;
            } else if (rec) {  // one 32-byte record per (camera, primitive)
                const float4 r0 = rec[2 * (int64_t)g], r1 = rec[2 * (int64_t)g + 1];
                xy = make_float2(r0.x, r0.y);
                ca = r0.z; cb = r0.w; cc = r1.x; op = r1.y; be = r1.z;
            } else {
                xy = reinterpret_cast<const float2 *>(means2d)[g];
                ca = conics[(int64_t)g * 3]; cb = conics[(int64_t)g * 3 + 1]; cc = conics[(int64_t)g * 3 + 2];
                op = opacities[g]; be = betas[g];
            }
            s_geo[tid] = make_float4(xy.x, xy.y, ca, 2.f * cb);
            s_geo2[tid] = make_float4(cc, op, be, 0.f);
            const float *cp = colors + (int64_t)g * channels + coff;
            if constexpr (kColVec) {
                s_colv[tid] = (rec && wide_rec) ? cw
                                                : make_float4(cp[0], CDIM > 1 ? cp[1] : 0.f, CDIM > 2 ? cp[2] : 0.f,
                                                              CDIM > 3 ? cp[3] : 0.f);
            } else {
#pragma unroll
                for (int k = 0; k < CDIM; ++k) s_col[k][tid] = cp[k];
            }
            float hx, hy;
            support_box(ca, cb, cc, hx, hy);
            s_mask[tid] = g_ok ? subblock_mask_exact<SB>(xy.x, xy.y, ca, cb, cc, hx, hy, x0, y0) : 0u;
        }
        __syncthreads();
        // ---- per-wave compaction of the batch to this sub-block's primitives
        // (a list in LDS measured faster here than walking ballot bits, whose
        // scalar bookkeeping competes with this short loop body)
        int n_mine = 0;
        for (int c0 = 0; c0 < cnt; c0 += 64) {
            const int e = c0 + lane;
            const bool hit = e < cnt && ((s_mask[e] >> wave) & 1u);
            const uint64_t bal = __ballot(hit);
            if (hit) {
                const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
                s_list[wave][n_mine + __popcll(bal & lt)] = (uint32_t)(e * 16);
            }
            n_mine += __popcll(bal);
        }
        // (s_list[wave] is written and read by this wave only: LDS is in order)
        // ---- front-to-back blend over the compacted list, branch-free per
        // pixel: lanes that skip the primitive blend alpha = 0.  The loop body
        // has no branch at all (the all-done exit is checked once per 8
        // primitives): skipping the body when no lane hits the primitive cost
        // more in scalar mask bookkeeping than it saved (1.39 vs 1.18 ms at
        // 1080p / 1M primitives)
        int cur_off = -1;  // byte offset of this batch's last contributor
        for (int k0 = 0; k0 < n_mine; k0 += 8) {
            if (__ballot(!done) == 0ull) break;  // whole wave done (checked every 8)
            const int kend = min(k0 + 8, n_mine);
            // one primitive of the compacted list: the pixel's front-to-back step
            auto blend_one = [&](const int off, const float4 g1, const float4 g2, const float4 cv) {
                const float dx = g1.x - px, dy = g1.y - py;
                const float sigma = __builtin_fmaf(g1.w * dx, dy, __builtin_fmaf(g1.z * dx, dx, (g2.x * dy) * dy));
                const bool ok = !done && sigma >= 0.f && sigma < 1.f;
                // outside [0, 1) the power is NaN or > 1 and the select drops it
                const float a0 = fminf(0.999f, g2.y * beta_pow(1.f - sigma, g2.z));
                const float alpha = ok ? a0 : 0.f;
                const float nT = T * (1.f - alpha);
                const bool stop = ok && nT <= 1e-4f;  // this primitive would end the pixel: excluded
                const bool take = ok && !stop;
                if constexpr (STATS) {
                    st_steps += 1;
                    st_live += (uint32_t)!done;
                    st_sup += (uint32_t)ok;
                    st_take += (uint32_t)take;
                }
                done = done || stop;
                const float vis = take ? alpha * T : 0.f;
                if constexpr (kColVec) {
                    const float c4[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
                    for (int q = 0; q < CDIM; ++q) acc[q] = __builtin_fmaf(c4[q], vis, acc[q]);
                } else {
#pragma unroll
                    for (int q = 0; q < CDIM; ++q) acc[q] = __builtin_fmaf(s_col[q][off >> 4], vis, acc[q]);
                }
                cur_off = take ? off : cur_off;
                T = take ? nT : T;
            };
            auto rec_at = [&](const int off, float4 &g1, float4 &g2, float4 &cv) {
                g1 = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(s_geo) + off);
                g2 = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(s_geo2) + off);
                cv = kColVec ? *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(s_colv) + off)
                             : make_float4(0.f, 0.f, 0.f, 0.f);
            };
#if UBS_FWD_PAIRED_READS
            // two list entries per LDS read and both records read before either
            // is blended: half the list reads, one LDS round trip per two steps
            int k = k0;
            for (; k + 1 < kend; k += 2) {
                const uint2 both = *reinterpret_cast<const uint2 *>(&s_list[wave][k]);  // k is even
                const int off0 = (int)both.x, off1 = (int)both.y;
                float4 a1, a2, ac, b1, b2, bc;
                rec_at(off0, a1, a2, ac);
                rec_at(off1, b1, b2, bc);
                blend_one(off0, a1, a2, ac);
                blend_one(off1, b1, b2, bc);
            }
            if (k < kend) {
                const int off = s_list[wave][k];
                float4 g1, g2, cv;
                rec_at(off, g1, g2, cv);
                blend_one(off, g1, g2, cv);
            }
#else
            for (int k = k0; k < kend; ++k) {
                const int off = s_list[wave][k];
                float4 g1, g2, cv;
                rec_at(off, g1, g2, cv);
                blend_one(off, g1, g2, cv);
            }
#endif
        }
        if (cur_off >= 0) cur = batch_start + (cur_off >> 4);
        // T after this batch, for a backward that restarts each batch from it
        if (T_ckpt && write_meta) T_ckpt[ckpt_slot(range_start, logical, b, BS) * BS + tid] = T;
    }

    if constexpr (STATS) {
        const uint32_t v[5] = {st_steps, st_live, st_sup, st_take, st_staged};
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            // steps are wave-uniform: one lane's count; the others sum over the wave
            const unsigned long long tot = q == 0 ? (unsigned long long)v[0] : wave_sum_u64(v[q]);
            if (stats && lane == 0 && tot) atomicAdd(stats + q, tot);
        }
        if (wg_times) {  // start, end, tile and the hardware id of the workgroup
            __syncthreads();
            if (tid == 0) {
                unsigned long long *o = wg_times + 4ull * blockIdx.x;
                o[0] = t_start;
                o[1] = __builtin_amdgcn_s_memrealtime();
                o[2] = logical;
                o[3] = (unsigned long long)__builtin_amdgcn_s_getreg((31 << 11) | 4);  // HW_ID
            }
        }
    }
    if (inside) {
#pragma unroll
        for (int k = 0; k < CDIM; ++k)
            render_colors[pix * channels + coff + k] =
                backgrounds ? __builtin_fmaf(T, backgrounds[cam * channels + coff + k], acc[k]) : acc[k];
        if (write_meta) {
            render_alphas[pix] = 1.f - T;
            last_ids[pix] = cur;
            if (T_final) T_final[pix] = T;
        }
    }
}

template <int TS, int CDIM>
static int launch_fwd(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels, uint32_t coff,
                      const float *means2d, const float *conics, const float *colors, const float *opac,
                      const float *betas, const float *bg, const uint8_t *masks, uint32_t W, uint32_t H,
                      uint32_t tw, uint32_t th, const int32_t *offs, const int32_t *flat, float *rc, float *ra,
                      int32_t *last, float *Tf, const float4 *rec, float *ck, const int64_t *n_dev,
                      const uint32_t *order, hipStream_t st) {
    raster_fwd_kernel<TS, CDIM><<<xcd_grid(C, tw, th), TS * TS, 0, st>>>(
        C, N, n_isects, channels, coff, means2d, conics, colors, opac, betas, bg, masks, W, H, tw, th, offs, flat,
        rc, ra, last, Tf, rec, rec && raster_wide_records(channels, TS) ? 1 : 0, ck, nullptr, n_dev, order);
    UBS_LAUNCH_CHECK("raster_fwd_kernel");
    return UBS_OK;
}

template <int TS>
static int launch_fwd_chunk(int cd, uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels, uint32_t coff,
                            const float *means2d, const float *conics, const float *colors, const float *opac,
                            const float *betas, const float *bg, const uint8_t *masks, uint32_t W, uint32_t H,
                            uint32_t tw, uint32_t th, const int32_t *offs, const int32_t *flat, float *rc,
                            float *ra, int32_t *last, float *Tf, const float4 *rec, float *ck, const int64_t *n_dev,
                            const uint32_t *order, hipStream_t st) {
#define UBS_FWD_CASE(K)                                                                                     \
    case K:                                                                                                 \
        return launch_fwd<TS, K>(C, N, n_isects, channels, coff, means2d, conics, colors, opac, betas, bg,  \
                                 masks, W, H, tw, th, offs, flat, rc, ra, last, Tf, rec, ck, n_dev, order, st);
    switch (cd) {
        UBS_FWD_CASE(1)
        UBS_FWD_CASE(2)
        UBS_FWD_CASE(3)
        UBS_FWD_CASE(4)
    case 8:
        if constexpr (TS <= 16)
            return launch_fwd<TS, 8>(C, N, n_isects, channels, coff, means2d, conics, colors, opac, betas, bg,
                                     masks, W, H, tw, th, offs, flat, rc, ra, last, Tf, rec, ck, n_dev, order, st);
        break;
    case 16:
        if constexpr (TS <= 16)
            return launch_fwd<TS, 16>(C, N, n_isects, channels, coff, means2d, conics, colors, opac, betas, bg,
                                      masks, W, H, tw, th, offs, flat, rc, ra, last, Tf, rec, ck, n_dev, order, st);
        break;
    case 32:
        if constexpr (TS <= 16)
            return launch_fwd<TS, 32>(C, N, n_isects, channels, coff, means2d, conics, colors, opac, betas, bg,
                                      masks, W, H, tw, th, offs, flat, rc, ra, last, Tf, rec, ck, n_dev, order, st);
        break;
    default:
        break;
    }
#undef UBS_FWD_CASE
    set_error("internal: bad channel chunk %d", cd);
    return UBS_ERR_INVALID;
}

}  // namespace ubs

using namespace ubs;

static int raster_fwd_impl(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels, const float *means2d,
                           const float *conics, const float *colors, const float *opacities, const float *betas,
                           const float *backgrounds, const uint8_t *masks, uint32_t W, uint32_t H,
                           uint32_t tile_size, uint32_t tw, uint32_t th, const int32_t *tile_offsets,
                           const int32_t *flatten_ids, float *render_colors, float *render_alphas,
                           int32_t *last_ids, float *T_final, float *T_ckpt, size_t ckpt_n, void *workspace,
                           size_t workspace_bytes, void *stream, const int64_t *n_dev = nullptr,
                           bool prepacked = false) {
    hipStream_t st = as_stream(stream);
    UBS_REQUIRE(tile_size == 8 || tile_size == 16 || tile_size == 32, "tile_size must be 8, 16 or 32 (got %u)",
                tile_size);
    UBS_REQUIRE(channels >= 1, "channels must be >= 1");
    UBS_REQUIRE((uint64_t)tw * tile_size >= W && (uint64_t)th * tile_size >= H, "tile grid does not cover the image");
    UBS_REQUIRE(n_isects >= 0 && n_isects < (1ll << 31), "n_isects out of range");
    if ((uint64_t)C * W * H == 0) return UBS_OK;
    UBS_REQUIRE((uint64_t)C * th * tw < (1ull << 31), "too many tiles");
    UBS_REQUIRE(means2d && conics && colors && opacities && betas && tile_offsets && render_colors && render_alphas &&
                    last_ids && (flatten_ids || n_isects == 0),
                "null pointer");
    if (T_ckpt) {
        const size_t need = ckpt_floats(C, n_isects, tw, th, tile_size);
        if (ckpt_n < need) {
            set_error("transmittance checkpoint buffer of %zu floats, %zu needed", ckpt_n, need);
            return UBS_ERR_WORKSPACE;
        }
    }
    const float4 *rec = nullptr;
    const uint32_t *order = nullptr;
    if (workspace) {
        const size_t need = raster_workspace_bytes(C, N, 0);
        if (workspace_bytes < need) {
            set_error("blend workspace of %zu bytes, %zu needed", workspace_bytes, need);
            return UBS_ERR_WORKSPACE;
        }
        const uint64_t CN = (uint64_t)C * N;
        const bool wide = raster_wide_records(channels, tile_size);
        UBS_REQUIRE(!prepacked || (wide && channels == 3), "prepacked records are the wide 3-channel ones");
        if (CN && n_isects) {
            if (!prepacked) {
                raster_pack_kernel<<<grid_for((int64_t)CN), kThreads, 0, st>>>(
                    CN, means2d, conics, opacities, betas, static_cast<float4 *>(workspace), wide ? colors : nullptr,
                    channels);
                UBS_LAUNCH_CHECK("raster_pack_kernel");
            }
            rec = static_cast<const float4 *>(workspace);
            if (workspace_bytes >= need + raster_order_bytes(C, tw, th)) {  // room for the block order
                uint32_t *o = reinterpret_cast<uint32_t *>(static_cast<char *>(workspace) + need);
                raster_block_order_kernel<<<1, kOrderThreads, 0, st>>>(C, tw, th, tile_offsets, n_isects, n_dev, o);
                UBS_LAUNCH_CHECK("raster_block_order_kernel");
                order = o;
            }
        }
    }
    uint32_t coff = 0;
    while (coff < channels) {
        const int cd = pick_chunk(channels - coff, tile_size);
        int e;
        if (tile_size == 16)
            e = launch_fwd_chunk<16>(cd, C, N, n_isects, channels, coff, means2d, conics, colors, opacities, betas,
                                     backgrounds, masks, W, H, tw, th, tile_offsets, flatten_ids, render_colors,
                                     render_alphas, last_ids, T_final, rec, T_ckpt, n_dev, order, st);
        else if (tile_size == 8)
            e = launch_fwd_chunk<8>(cd, C, N, n_isects, channels, coff, means2d, conics, colors, opacities, betas,
                                    backgrounds, masks, W, H, tw, th, tile_offsets, flatten_ids, render_colors,
                                    render_alphas, last_ids, T_final, rec, T_ckpt, n_dev, order, st);
        else
            e = launch_fwd_chunk<32>(cd, C, N, n_isects, channels, coff, means2d, conics, colors, opacities, betas,
                                     backgrounds, masks, W, H, tw, th, tile_offsets, flatten_ids, render_colors,
                                     render_alphas, last_ids, T_final, rec, T_ckpt, n_dev, order, st);
        if (e) return e;
        coff += (uint32_t)cd;
    }
    return UBS_OK;
}

extern "C" int ubs_rasterize_to_pixels_fwd(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels,
                                           const float *means2d, const float *conics, const float *colors,
                                           const float *opacities, const float *betas, const float *backgrounds,
                                           const uint8_t *masks, uint32_t W, uint32_t H, uint32_t tile_size,
                                           uint32_t tw, uint32_t th, const int32_t *tile_offsets,
                                           const int32_t *flatten_ids, float *render_colors, float *render_alphas,
                                           int32_t *last_ids, float *T_final, void *workspace,
                                           size_t workspace_bytes, void *stream) {
    return raster_fwd_impl(C, N, n_isects, channels, means2d, conics, colors, opacities, betas, backgrounds, masks, W,
                           H, tile_size, tw, th, tile_offsets, flatten_ids, render_colors, render_alphas, last_ids,
                           T_final, nullptr, 0, workspace, workspace_bytes, stream);
}

extern "C" int ubs_rasterize_to_pixels_fwd_prepacked(
    uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels, const float *means2d, const float *conics,
    const float *colors, const float *opacities, const float *betas, const float *backgrounds, const uint8_t *masks,
    uint32_t W, uint32_t H, uint32_t tile_size, uint32_t tw, uint32_t th, const int32_t *tile_offsets,
    const int32_t *flatten_ids, float *render_colors, float *render_alphas, int32_t *last_ids, float *T_final,
    void *workspace, size_t workspace_bytes, void *stream) {
    UBS_REQUIRE(workspace, "prepacked records need the workspace that holds them");
    return raster_fwd_impl(C, N, n_isects, channels, means2d, conics, colors, opacities, betas, backgrounds, masks, W,
                           H, tile_size, tw, th, tile_offsets, flatten_ids, render_colors, render_alphas, last_ids,
                           T_final, nullptr, 0, workspace, workspace_bytes, stream, nullptr, true);
}

extern "C" int ubs_rasterize_to_pixels_fwd_dev(uint32_t C, uint32_t N, int64_t n_cap, const int64_t *n_dev,
                                               uint32_t channels, const float *means2d, const float *conics,
                                               const float *colors, const float *opacities, const float *betas,
                                               const float *backgrounds, const uint8_t *masks, uint32_t W,
                                               uint32_t H, uint32_t tile_size, uint32_t tw, uint32_t th,
                                               const int32_t *tile_offsets, const int32_t *flatten_ids,
                                               float *render_colors, float *render_alphas, int32_t *last_ids,
                                               float *T_final, void *workspace, size_t workspace_bytes,
                                               void *stream) {
    UBS_REQUIRE(n_dev, "n_dev is null");
    return raster_fwd_impl(C, N, n_cap, channels, means2d, conics, colors, opacities, betas, backgrounds, masks, W,
                           H, tile_size, tw, th, tile_offsets, flatten_ids, render_colors, render_alphas, last_ids,
                           T_final, nullptr, 0, workspace, workspace_bytes, stream, n_dev);
}

extern "C" int ubs_rasterize_to_pixels_fwd_ckpt(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels,
                                                const float *means2d, const float *conics, const float *colors,
                                                const float *opacities, const float *betas,
                                                const float *backgrounds, const uint8_t *masks, uint32_t W,
                                                uint32_t H, uint32_t tile_size, uint32_t tw, uint32_t th,
                                                const int32_t *tile_offsets, const int32_t *flatten_ids,
                                                float *render_colors, float *render_alphas, int32_t *last_ids,
                                                float *T_final, float *T_ckpt, size_t ckpt_n, void *workspace,
                                                size_t workspace_bytes, void *stream) {
    UBS_REQUIRE(T_ckpt || n_isects == 0, "null checkpoint buffer");
    return raster_fwd_impl(C, N, n_isects, channels, means2d, conics, colors, opacities, betas, backgrounds, masks, W,
                           H, tile_size, tw, th, tile_offsets, flatten_ids, render_colors, render_alphas, last_ids,
                           T_final, T_ckpt, ckpt_n, workspace, workspace_bytes, stream);
}

extern "C" size_t ubs_transmittance_checkpoint_floats(uint32_t C, int64_t n_isects, uint32_t tile_width,
                                                      uint32_t tile_height, uint32_t tile_size) {
    return ckpt_floats(C, n_isects < 0 ? 0 : n_isects, tile_width, tile_height, tile_size);
}

extern "C" size_t ubs_rasterize_workspace_bytes(uint32_t C, uint32_t N, int backward) {
    return raster_workspace_bytes(C, N, backward);
}

extern "C" size_t ubs_rasterize_workspace_bytes_tiles(uint32_t C, uint32_t N, uint32_t tile_width,
                                                      uint32_t tile_height, int backward) {
    return raster_workspace_bytes(C, N, backward) + raster_order_bytes(C, tile_width, tile_height);
}

#ifdef UBS_PROBE_BUILD
// Measurement build only (libubs_probe.so): the 16x16-tile, 3-channel blend
// forward with the STATS counters, on the packed-record path of the library's
// forward; stats: 5 zeroed uint64 on the device.
extern "C" __attribute__((visibility("default"))) int ubs_probe_raster_fwd_stats(
    uint32_t C, uint32_t N, int64_t n_isects, const float *means2d, const float *conics, const float *colors,
    const float *opacities, const float *betas, uint32_t W, uint32_t H, uint32_t tw, uint32_t th,
    const int32_t *tile_offsets, const int32_t *flatten_ids, float *render_colors, float *render_alphas,
    int32_t *last_ids, float *T_final, void *workspace, unsigned long long *stats, void *stream) {
    hipStream_t st = as_stream(stream);
    const uint64_t CN = (uint64_t)C * N;
    raster_pack_kernel<<<grid_for((int64_t)CN), kThreads, 0, st>>>(CN, means2d, conics, opacities, betas,
                                                                   static_cast<float4 *>(workspace), colors, 3);
    raster_fwd_kernel<16, 3, true><<<xcd_grid(C, tw, th), 256, 0, st>>>(
        C, N, n_isects, 3, 0, means2d, conics, colors, opacities, betas, nullptr, nullptr, W, H, tw, th,
        tile_offsets, flatten_ids, render_colors, render_alphas, last_ids, T_final,
        static_cast<const float4 *>(workspace), 1, nullptr, stats);
    return (int)hipGetLastError();
}

// The same forward dealing its tiles through `order` (null: raster order;
// else an order as the library writes it) and writing per workgroup its
// start and end on the constant-rate clock, its tile and its hardware id
// into wg_times [xcd_grid * 4] (zeroed); the records in `workspace` are
// packed here.  stats may be null (no counting atomics: timing only).
extern "C" __attribute__((visibility("default"))) int ubs_probe_raster_fwd_times(
    uint32_t C, uint32_t N, int64_t n_isects, const float *means2d, const float *conics, const float *colors,
    const float *opacities, const float *betas, uint32_t W, uint32_t H, uint32_t tw, uint32_t th,
    const int32_t *tile_offsets, const int32_t *flatten_ids, float *render_colors, float *render_alphas,
    int32_t *last_ids, float *T_final, void *workspace, const uint32_t *order, unsigned long long *stats,
    unsigned long long *wg_times, void *stream) {
    hipStream_t st = as_stream(stream);
    const uint64_t CN = (uint64_t)C * N;
    raster_pack_kernel<<<grid_for((int64_t)CN), kThreads, 0, st>>>(CN, means2d, conics, opacities, betas,
                                                                   static_cast<float4 *>(workspace), colors, 3);
    raster_fwd_kernel<16, 3, true><<<xcd_grid(C, tw, th), 256, 0, st>>>(
        C, N, n_isects, 3, 0, means2d, conics, colors, opacities, betas, nullptr, nullptr, W, H, tw, th,
        tile_offsets, flatten_ids, render_colors, render_alphas, last_ids, T_final,
        static_cast<const float4 *>(workspace), 1, nullptr, stats, nullptr, order, wg_times);
    return (int)hipGetLastError();
}
#endif

// Backward tile blend of Beta primitives on gfx950.
//
// Same workgroup/tile/sub-block layout as the forward.  Batches are walked
// back to front; each wave compacts a batch to the primitives whose support
// box reaches its 8x8 sub-block and whose isect index is at most the wave's
// largest last contributor, then per primitive:
//   * recovers each pixel's transmittance before the primitive from the
//     exact T after the last contributor (kept by the forward) by dividing
//     by (1 - alpha) -- the forward and backward evaluate alpha with the same
//     function, so the divisors are bit-identical;
//   * forms the gradients of rasterize_to_pixels_bwd.cu:200-247, branch-free
//     (lanes without a contribution compute exact zeros);
//   * sums them over the 64 lanes with a reduce-scatter (permlane swaps
//     across the 16-lane rows, bank-masked DPP adds inside them) that leaves
//     each 4-lane bank with different totals (wave_totals_rows), and one lane
//     per total adds it into the primitive's LDS accumulator (ds_add_f32),
//     so the four waves of a tile share one set of partial sums
//     (finishing with half-row sums and two adding lanes per total measured
//     slower: 3.23 vs 3.09 ms, the same-address LDS adds cost more than the
//     DPP steps they save);
// and after the batch the accumulators go to the per-primitive gradients
// with hardware float atomics (one set per (tile, primitive) instead of one
// per (wave, primitive)), 8 lanes on one primitive's 32-byte packed record so
// that an atomic instruction touches 8 records, not 64 scattered floats.
// The T recovery can run in fp64 (template flag) where stability needs it.
#pragma clang fp contract(off)

#include <type_traits>

#include "ubs_raster.h"

namespace ubs {

#ifndef UBS_BWD_WAVES_PER_EU
#define UBS_BWD_WAVES_PER_EU 1
#endif
// STATS (measurement builds only; the library's kernels have it off) counts
// into stats[0..6]: (wave, primitive) evaluations, lane evaluations at or
// before the pixel's last contributor, of those that contribute, pair steps
// that run the wave reduce-scatter, staged entries, and the reduce-scatter
// steps with at most 4 and at most 16 contributing lanes.
template <int TS, int CDIM, bool F64T, bool STATS = false>
__global__ __launch_bounds__(TS *TS, UBS_BWD_WAVES_PER_EU) void raster_bwd_kernel(
    uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels, uint32_t coff,
    const float *__restrict__ means2d, const float *__restrict__ conics, const float *__restrict__ colors,
    const float *__restrict__ opacities, const float *__restrict__ betas, const float *__restrict__ backgrounds,
    const uint8_t *__restrict__ masks, uint32_t W, uint32_t H, uint32_t tw, uint32_t th,
    const int32_t *__restrict__ tile_offsets, const int32_t *__restrict__ flatten_ids,
    const float *__restrict__ render_alphas, const int32_t *__restrict__ last_ids, const float *__restrict__ T_final,
    const float *__restrict__ v_render_colors, const float *__restrict__ v_render_alphas,
    float *__restrict__ v_means2d, float *__restrict__ v_conics, float *__restrict__ v_colors,
    float *__restrict__ v_opacities, float *__restrict__ v_betas, const float4 *__restrict__ rec, int wide_rec,
    float *__restrict__ vrec, const float *__restrict__ T_ckpt, unsigned long long *__restrict__ stats = nullptr,
    const uint32_t *__restrict__ order = nullptr, unsigned long long *__restrict__ wg_times = nullptr) {
    constexpr int BS = TS * TS;
    constexpr int SB = TS / 8;
    constexpr int NW = BS / 64;
    constexpr int NV = 7 + CDIM;  // xy(2) conic(3) opacity beta colour(CDIM)
    constexpr int NG2 = WaveTotals<2 * NV>::kGroups;
    __shared__ float4 s_geo[BS];   // x, y, a, 2b
    __shared__ float4 s_geo2[BS];  // c, opacity, beta, id (bits)
    // up to 4 channels: one float4 per primitive at the stride of s_geo, so one
    // address serves all three LDS reads of a primitive
    constexpr bool kColVec = CDIM <= 4;
    __shared__ float4 s_colv[kColVec ? BS : 1];
    __shared__ float s_col[kColVec ? 1 : CDIM][kColVec ? 1 : BS];
    __shared__ uint32_t s_mask[BS];
    __shared__ float s_acc[NV][BS + 1];  // +1: the four quarter lanes hit distinct banks
    __shared__ int32_t s_wmax[NW];

    uint32_t logical;
    if (!xcd_tile(blockIdx.x, C, tw, th, logical, order)) return;
    // (measurement builds: the workgroup's start on the constant-rate clock)
    unsigned long long t_start = 0;
    if constexpr (STATS) t_start = __builtin_amdgcn_s_memrealtime();
    const uint32_t cam = logical / (th * tw);
    const uint32_t tile = logical % (th * tw);
    const uint32_t ty = tile / tw, tx = tile % tw;
    if (masks && !masks[(size_t)cam * th * tw + tile]) return;

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int sx = wave % SB, sy = wave / SB;
    const uint32_t i = ty * TS + sy * 8 + (lane >> 3);
    const uint32_t j = tx * TS + sx * 8 + (lane & 7);
    const bool inside = i < H && j < W;
    const int64_t pix = (int64_t)cam * H * W + (int64_t)i * W + j;
    const bool first_chunk = (coff == 0);

    const int64_t tbase = (int64_t)cam * th * tw;
    // a valid sort gives ordered offsets within [0, n_isects]; the clamps keep
    // the lists of a sort that gave up on a look-back wait inside the buffers
    const int32_t off_hi = (cam == C - 1 && tile == th * tw - 1) ? (int32_t)n_isects : tile_offsets[tbase + tile + 1];
    const int32_t range_start = min(max(tile_offsets[tbase + tile], 0), (int32_t)n_isects);
    const int32_t range_end = min(max(off_hi, range_start), (int32_t)n_isects);
    const float px = (float)j + 0.5f, py = (float)i + 0.5f;
    const float x0 = (float)(tx * TS), y0 = (float)(ty * TS);

    // per-pixel state
    const float Tf = inside ? (T_final ? T_final[pix] : 1.f - render_alphas[pix]) : 1.f;
    using TT = typename std::conditional<F64T, double, float>::type;
    TT T = (TT)Tf;
    const int32_t bin_final = inside ? last_ids[pix] : 0;
    float vrc[CDIM];
    float bg_dot = 0.f;
#pragma unroll
    for (int k = 0; k < CDIM; ++k) {
        vrc[k] = inside ? v_render_colors[pix * channels + coff + k] : 0.f;
        if (backgrounds) bg_dot = __builtin_fmaf(backgrounds[cam * channels + coff + k], vrc[k], bg_dot);
    }
    const float vra = (inside && first_chunk && v_render_alphas) ? v_render_alphas[pix] : 0.f;  // null: zeros
    const float vra_bg = vra - bg_dot;  // d/d(T_final) of the pixel's outputs
    const float tf_vra = Tf * vra_bg;
    float bufv = 0.f;                   // sum over later primitives of c . v_render_color * alpha T

    // the largest last contributor over the wave bounds the useful range
    {
        const int32_t wm = wave_max_i32(bin_final);
        if (lane == 0) s_wmax[wave] = wm;
    }
    for (int k = tid; k < NV * (BS + 1); k += BS) (&s_acc[0][0])[k] = 0.f;
    __syncthreads();
    const int32_t wave_bin_max = s_wmax[wave];
    // the value each lane adds: after wave_totals_rows, bank b of row Q holds
    // in register k the total of value wave_total_index(row_group(k, b), Q);
    // lane 16 Q + 4 b + j adds register 4 p + j in pass p, so ceil(K / 4)
    // ds_add_f32 (one for up to 5 colour channels) store the 2 NV totals of a
    // pair of primitives
    constexpr int NREG = RowTotals<NG2>::kRegs;
    constexpr int NPASS = (NREG + 3) / 4;
    const int my_j = lane & 3, my_bank = (lane >> 2) & 3;
    // per pass: the lane's accumulator row (byte offset of s_acc[idx % NV][0],
    // -1 for a lane without a total) and whether its total is the pair's
    // second primitive's (then the entry offset t1 - t0 is added)
    int my_row[NPASS];
    uint32_t my_second[NPASS];
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
        const int k = 4 * p + my_j;
        const int grp = k < NREG ? row_group<NG2>(k, my_bank) : -1;
        const int idx = grp >= 0 ? wave_total_index<2 * NV>(grp, lane >> 4) : -1;
        my_row[p] = idx >= 0 ? (idx % NV) * (BS + 1) * (int)sizeof(float) : -1;
        my_second[p] = idx >= NV ? ~0u : 0u;
        asm volatile("" : "+v"(my_second[p]));  // kept as a mask: one AND selects the offset
    }

    uint32_t st_steps = 0, st_live = 0, st_valid = 0, st_rs = 0, st_staged = 0, st_k4 = 0, st_k16 = 0;
    // batches are the forward's, [range_start + k BS, ...), taken last first
    const int32_t nb = range_end > range_start ? (range_end - range_start + BS - 1) / BS : 0;
    for (int32_t b = 0; b < nb; ++b) {
        const int32_t kb = nb - 1 - b;
        const int32_t batch_lo = range_start + kb * BS;
        const int32_t batch_end = min(range_end, batch_lo + BS) - 1;  // isect index of entry t is batch_end - t
        const int32_t cnt = batch_end + 1 - batch_lo;
        if constexpr (STATS) st_staged += (uint32_t)(tid < cnt);
        // with checkpoints, a pixel that has entries here restarts from the
        // forward's T after this batch (the forward ran this batch for it)
        if (T_ckpt && inside && bin_final >= batch_lo)
            T = (TT)T_ckpt[ckpt_slot(range_start, logical, kb, BS) * BS + tid];
        if (tid < cnt) {
            // an id outside [0, C N) (only from a sort that gave up) stages
            // primitive 0 with an empty sub-block mask: never blended
            const int32_t g_raw = flatten_ids[batch_end - tid];
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
            s_geo[tid] = make_float4(xy.x, xy.y, ca, 2.f * cb);  // 2b, as the forward
            s_geo2[tid] = make_float4(cc, op, be, __int_as_float(g));
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
        // ---- back-to-front over this sub-block's primitives of the batch
        // whose isect index is at most the wave's largest last contributor:
        // one ballot per 64 entries, then its set bits in order (the entry
        // index is wave-uniform, so its LDS reads are broadcasts)
        // entry t is at or before this pixel's last contributor iff t >= t_min
        const int32_t t_min = inside ? batch_end - bin_final : INT32_MAX;
        // Gradient terms of this pixel for batch entry t into v[o, o + NV);
        // advances T and bufv.  Branch-free: lanes without a contribution run
        // the same instructions with alpha = 0, (1 - sigma) = 1 and a zero
        // v_alpha, so they produce exact zeros and leave T and bufv unchanged.
        // Returns whether this pixel contributes.
        // the LDS record of batch entry t: geometry, opacity, beta, colour
        struct Rec {
            float4 g1, g2;
            float col[CDIM];
        };
        auto load_rec = [&](const int t) {
            Rec r;
            r.g1 = s_geo[t];
            r.g2 = s_geo2[t];
            if constexpr (kColVec) {
                const float4 cv = s_colv[t];
                const float c4[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
                for (int q = 0; q < CDIM; ++q) r.col[q] = c4[q];
            } else {
#pragma unroll
                for (int q = 0; q < CDIM; ++q) r.col[q] = s_col[q][t];
            }
            return r;
        };
        auto pixel_grads = [&](const Rec &r, const int t, float (&v)[2 * NV], const int o, const bool on) -> bool {
            const float4 g1 = r.g1;
            const float4 g2 = r.g2;
            const float dx = g1.x - px, dy = g1.y - py;
            const float sigma = __builtin_fmaf(g1.w * dx, dy, __builtin_fmaf(g1.z * dx, dx, (g2.x * dy) * dy));
            const bool valid = on && t >= t_min && sigma >= 0.f && sigma < 1.f;
            const float opac = g2.y, beta = g2.z;
            const float oms = valid ? 1.f - sigma : 1.f;
            const float lg = __builtin_amdgcn_logf(oms);  // log2(1 - sigma)
            // (1-sigma)^(beta-1) by one exp2, (1-sigma)^beta from it by one
            // multiply (one transcendental fewer than two exp2)
            const float pm1 = __builtin_amdgcn_exp2f((beta - 1.f) * lg);
            const float vis = pm1 * oms;
            const float ov = opac * vis;
            const float alpha = valid ? fminf(0.999f, ov) : 0.f;
            float ra;
            if constexpr (F64T) {
                const double rad = 1.0 / (1.0 - (double)alpha);
                T = valid ? T * rad : T;
                ra = (float)rad;
            } else {
                ra = __builtin_amdgcn_rcpf(1.f - alpha);
                T = valid ? T * ra : T;
            }
            const float Tc = (float)T;
            const float fac = alpha * Tc;
            // sum_k (c_k T - buf_k ra) v_k = T (c . v) - ra (buf . v), with
            // bufv = buf . v carried as one running sum
            const float(&col)[CDIM] = r.col;
            float cdot = col[0] * vrc[0];
            v[o + 7] = fac * vrc[0];
#pragma unroll
            for (int q = 1; q < CDIM; ++q) {
                v[o + 7 + q] = fac * vrc[q];
                cdot = __builtin_fmaf(col[q], vrc[q], cdot);
            }
            // T (c . v) - ra (buf . v) + ra T_final (v_alpha_out - bg . v)
            const float v_alpha = __builtin_fmaf(Tc, cdot, ra * (tf_vra - bufv));
            bufv = __builtin_fmaf(cdot, fac, bufv);
            // geometry, opacity and beta get gradient only below the 0.999 clamp
            const float va = (valid && ov <= 0.999f) ? v_alpha : 0.f;
            // v_sigma = -opac beta va pm1; the per-primitive factors (-opac beta,
            // the 2 of the symmetric terms, opac ln 2) are applied once per
            // (tile, primitive) at the flush, not per pixel
            const float sg = va * pm1;
            const float u = dx * sg, w = dy * sg;
            v[o + 2] = u * dx;                                // x -opac beta
            v[o + 3] = u * dy;                                // x -2 opac beta
            v[o + 4] = w * dy;                                // x -opac beta
            // v_xy = (a u + b w, b u + c w) summed over the pixels: the conic is
            // per primitive, so the sums of u and w suffice (combined at the flush)
            v[o + 0] = u;
            v[o + 1] = w;
            v[o + 5] = vis * va;
            v[o + 6] = v[o + 5] * lg;                         // x opac ln 2 (lg = log2(1 - sigma))
            return valid;
        };
        // The LDS adds of a pair's totals are issued after the next pair's
        // record reads: LDS operations complete in order, so an add issued
        // before the reads would make every read wait for it (its latency
        // was the largest single cost of this loop).
        bool pend = false;  // wave-uniform: a pair's totals wait to be added
        int pend_t0 = 0, pend_t1 = 0;
        float pend_val[NPASS];
        auto add_pending = [&]() {
            if (!pend) return;
            // byte offsets: the row's, the pair's first entry's, and for the
            // second primitive's totals the distance to its entry
            const uint32_t t0b = (uint32_t)pend_t0 * (uint32_t)sizeof(float);
            const uint32_t db = (uint32_t)(pend_t1 - pend_t0) * (uint32_t)sizeof(float);
#pragma unroll
            for (int p = 0; p < NPASS; ++p) {
                if (my_row[p] >= 0) {
                    const uint32_t off = (uint32_t)my_row[p] + t0b + (my_second[p] & db);
                    atomicAdd(reinterpret_cast<float *>(reinterpret_cast<char *>(&s_acc[0][0]) + off), pend_val[p]);
                }
            }
            pend = false;
        };
        for (int c0 = 0; c0 < cnt; c0 += 64) {
            const int e = c0 + lane;
            uint64_t todo = __ballot(e < cnt && ((s_mask[e] >> wave) & 1u) && (batch_end - e) <= wave_bin_max);
            // two primitives per step: their gradient chains interleave and one
            // reduce-scatter of 2 NV values serves both
            while (todo != 0ull) {
                const int t0 = c0 + __builtin_ctzll(todo);
                todo &= todo - 1ull;
                const bool two = todo != 0ull;
                const int t1 = two ? c0 + __builtin_ctzll(todo) : t0;
                if (two) todo &= todo - 1ull;
                const Rec r0 = load_rec(t0), r1 = load_rec(t1);
                add_pending();
                float v[2 * NV];
                // a lone last primitive runs as a pair with a switched-off partner
                // (exact zeros), which keeps the loop body free of branches
                const bool c0v = pixel_grads(r0, t0, v, 0, true);
                const bool c1v = pixel_grads(r1, t1, v, NV, two);
                if constexpr (STATS) {
                    const bool any = __ballot(c0v || c1v) != 0ull;
                    st_steps += two ? 2u : 1u;
                    st_live += (uint32_t)(t0 >= t_min) + (uint32_t)(two && t1 >= t_min);
                    st_valid += (uint32_t)c0v + (uint32_t)c1v;
                    st_rs += (uint32_t)any;
                    // lanes with a contribution to either primitive of the step
                    const uint32_t k = (uint32_t)__popcll(__ballot(c0v || c1v));
                    st_k4 += (uint32_t)(any && k <= 4);
                    st_k16 += (uint32_t)(any && k <= 16);
                }
                // (nearly every pair has a contribution in the wave; one without
                // adds exact zeros, cheaper than testing every pair for it)
                // wave totals, scattered over the rows and banks; one ds_add_f32
                // per pass adds them into the primitives' accumulators
                float tot[NREG];
                wave_totals_rows<2 * NV>(v, tot);
#pragma unroll
                for (int p = 0; p < NPASS; ++p) {
                    float mine = tot[4 * p];
#pragma unroll
                    for (int k = 4 * p + 1; k < NREG && k < 4 * p + 4; ++k) mine = my_j == k - 4 * p ? tot[k] : mine;
                    pend_val[p] = mine;
                }
                pend_t0 = t0;
                pend_t1 = t1;
                pend = true;
            }
        }
        add_pending();
        __syncthreads();
        // ---- flush this batch's accumulators to the per-primitive gradients
        if (vrec) {
            // packed records: lane -> (entry, value) with 8 consecutive lanes on
            // one primitive's 32-byte record, so one atomic instruction covers 8
            // records (8 floats each) instead of 64 scattered floats; with wide
            // records the colour totals go to the same 64-byte record
            constexpr int kCv = kColVec ? 4 : CDIM;  // lanes per primitive for the colours
            const int64_t rstride = wide_rec ? (int64_t)kWideRecordFloats : (int64_t)kRecordFloats;
            for (int k = tid; k < cnt * 8; k += BS) {
                const int t = k >> 3, q = k & 7;
                const float4 gp = s_geo2[t];  // c, opacity, beta, id
                const float k1 = -gp.y * gp.z;
                float val;
                if (q < 2) {
                    const float4 gg = s_geo[t];  // x, y, a, 2b
                    const float su = s_acc[0][t], sw = s_acc[1][t];
                    const float bb = 0.5f * gg.w;
                    val = (q == 0 ? __builtin_fmaf(gg.z, su, bb * sw) : __builtin_fmaf(bb, su, gp.x * sw)) *
                          (2.f * k1);
                } else {
                    const float acc = s_acc[q < 7 ? q : 0][t];
                    val = q == 2 || q == 4 ? acc * k1
                          : q == 3        ? acc * (2.f * k1)
                          : q == 6        ? acc * (gp.y * 0.69314718055994531f)
                          : q == 5        ? acc
                                          : 0.f;
                }
                if (val != 0.f) atomic_add_f32(vrec + (int64_t)__float_as_int(gp.w) * rstride + q, val);
            }
            for (int k = tid; k < cnt * kCv; k += BS) {
                const int t = k / kCv, q = k % kCv;
                if (q < CDIM) {
                    const float val = s_acc[7 + q][t];
                    const int64_t g = __float_as_int(s_geo2[t].w);
                    if (val != 0.f)
                        atomic_add_f32(wide_rec ? vrec + g * rstride + 8 + q : v_colors + g * channels + coff + q, val);
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < NV; ++q) s_acc[q][tid] = 0.f;
        } else if (tid < cnt) {
            const int32_t g = __float_as_int(s_geo2[tid].w);
            float a[NV];
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                a[q] = s_acc[q][tid];
                s_acc[q][tid] = 0.f;
            }
            const float4 gg = s_geo[tid];   // x, y, a, 2b
            const float4 gp = s_geo2[tid];  // c, opacity, beta, id
            const float k1 = -gp.y * gp.z, k2 = 2.f * k1;
            const float su = a[0], sw = a[1];
            const float bb = 0.5f * gg.w;
            a[0] = __builtin_fmaf(gg.z, su, bb * sw) * k2;
            a[1] = __builtin_fmaf(bb, su, gp.x * sw) * k2;
            a[2] *= k1;
            a[3] *= k2;
            a[4] *= k1;
            a[6] *= gp.y * 0.69314718055994531f;
            if (a[0] != 0.f) atomic_add_f32(v_means2d + (int64_t)g * 2, a[0]);
            if (a[1] != 0.f) atomic_add_f32(v_means2d + (int64_t)g * 2 + 1, a[1]);
            if (a[2] != 0.f) atomic_add_f32(v_conics + (int64_t)g * 3, a[2]);
            if (a[3] != 0.f) atomic_add_f32(v_conics + (int64_t)g * 3 + 1, a[3]);
            if (a[4] != 0.f) atomic_add_f32(v_conics + (int64_t)g * 3 + 2, a[4]);
            if (a[5] != 0.f) atomic_add_f32(v_opacities + g, a[5]);
            if (a[6] != 0.f) atomic_add_f32(v_betas + g, a[6]);
#pragma unroll
            for (int q = 0; q < CDIM; ++q)
                if (a[7 + q] != 0.f) atomic_add_f32(v_colors + (int64_t)g * channels + coff + q, a[7 + q]);
        }
        __syncthreads();
    }
    if constexpr (STATS) {
        const uint32_t v[7] = {st_steps, st_live, st_valid, st_rs, st_staged, st_k4, st_k16};
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            // steps, reduce-scatters and their lane-count classes are wave-uniform
            const unsigned long long tot =
                (q == 0 || q == 3 || q >= 5) ? (unsigned long long)v[q] : wave_sum_u64(v[q]);
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
}

template <int TS, int CDIM>
static int launch_bwd(bool f64t, uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels, uint32_t coff,
                      const float *means2d, const float *conics, const float *colors, const float *opac,
                      const float *betas, const float *bg, const uint8_t *masks, uint32_t W, uint32_t H,
                      uint32_t tw, uint32_t th, const int32_t *offs, const int32_t *flat, const float *ra,
                      const int32_t *last, const float *Tf, const float *vrc, const float *vra, float *vm2,
                      float *vcon, float *vcol, float *vop, float *vbe, const float4 *rec, float *vrec,
                      const float *ck, const uint32_t *order, hipStream_t st) {
    const int wide = rec && raster_wide_records(channels, TS) ? 1 : 0;
    if (f64t)
        raster_bwd_kernel<TS, CDIM, true><<<xcd_grid(C, tw, th), TS * TS, 0, st>>>(
            C, N, n_isects, channels, coff, means2d, conics, colors, opac, betas, bg, masks, W, H, tw, th, offs, flat,
            ra, last, Tf, vrc, vra, vm2, vcon, vcol, vop, vbe, rec, wide, vrec, ck, nullptr, order);
    else
        raster_bwd_kernel<TS, CDIM, false><<<xcd_grid(C, tw, th), TS * TS, 0, st>>>(
            C, N, n_isects, channels, coff, means2d, conics, colors, opac, betas, bg, masks, W, H, tw, th, offs, flat,
            ra, last, Tf, vrc, vra, vm2, vcon, vcol, vop, vbe, rec, wide, vrec, ck, nullptr, order);
    UBS_LAUNCH_CHECK("raster_bwd_kernel");
    return UBS_OK;
}

template <int TS>
static int launch_bwd_chunk(int cd, bool f64t, uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels,
                            uint32_t coff, const float *means2d, const float *conics, const float *colors,
                            const float *opac, const float *betas, const float *bg, const uint8_t *masks, uint32_t W,
                            uint32_t H, uint32_t tw, uint32_t th, const int32_t *offs, const int32_t *flat,
                            const float *ra, const int32_t *last, const float *Tf, const float *vrc,
                            const float *vra, float *vm2, float *vcon, float *vcol, float *vop, float *vbe,
                            const float4 *rec, float *vrec, const float *ck, const uint32_t *order, hipStream_t st) {
#define UBS_BWD_ARGS                                                                                            \
    f64t, C, N, n_isects, channels, coff, means2d, conics, colors, opac, betas, bg, masks, W, H, tw, th, offs,  \
        flat, ra, last, Tf, vrc, vra, vm2, vcon, vcol, vop, vbe, rec, vrec, ck, order, st
    switch (cd) {
    case 1: return launch_bwd<TS, 1>(UBS_BWD_ARGS);
    case 2: return launch_bwd<TS, 2>(UBS_BWD_ARGS);
    case 3: return launch_bwd<TS, 3>(UBS_BWD_ARGS);
    case 4: return launch_bwd<TS, 4>(UBS_BWD_ARGS);
    case 8:
        if constexpr (TS <= 16) return launch_bwd<TS, 8>(UBS_BWD_ARGS);
        break;
    case 16:
        if constexpr (TS <= 16) return launch_bwd<TS, 16>(UBS_BWD_ARGS);
        break;
    case 32:
        if constexpr (TS <= 16) return launch_bwd<TS, 32>(UBS_BWD_ARGS);
        break;
    default:
        break;
    }
#undef UBS_BWD_ARGS
    set_error("internal: bad channel chunk %d", cd);
    return UBS_ERR_INVALID;
}

// fp64 transmittance recovery switch (process-wide, read at each call).
static int g_f64_transmittance = 0;

}  // namespace ubs

using namespace ubs;

extern "C" {

// Selects fp32 (0, default) or fp64 (1) arithmetic for the backward's
// transmittance recovery.  Not part of the reference interface.
int ubs_set_backward_fp64_transmittance(int enable) {
    g_f64_transmittance = enable ? 1 : 0;
    return UBS_OK;
}

static int raster_bwd_impl(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels, const float *means2d,
                           const float *conics, const float *colors, const float *opacities, const float *betas,
                           const float *backgrounds, const uint8_t *masks, uint32_t W, uint32_t H,
                           uint32_t tile_size, uint32_t tw, uint32_t th, const int32_t *tile_offsets,
                           const int32_t *flatten_ids, const float *render_alphas, const int32_t *last_ids,
                           const float *T_final, const float *v_render_colors, const float *v_render_alphas,
                           float *v_means2d, float *v_conics, float *v_colors, float *v_opacities, float *v_betas,
                           const float *T_ckpt, size_t ckpt_n, void *workspace, size_t workspace_bytes,
                           void *stream, const void *fwd_ws = nullptr, size_t fwd_ws_bytes = 0) {
    hipStream_t st = as_stream(stream);
    UBS_REQUIRE(tile_size == 8 || tile_size == 16 || tile_size == 32, "tile_size must be 8, 16 or 32 (got %u)",
                tile_size);
    UBS_REQUIRE(channels >= 1, "channels must be >= 1");
    UBS_REQUIRE((uint64_t)tw * tile_size >= W && (uint64_t)th * tile_size >= H, "tile grid does not cover the image");
    UBS_REQUIRE(n_isects >= 0 && n_isects < (1ll << 31), "n_isects out of range");
    const uint64_t CN = (uint64_t)C * N;
    UBS_REQUIRE(CN == 0 || (v_means2d && v_conics && v_colors && v_opacities && v_betas), "null gradient output");
    const bool packed = workspace != nullptr && n_isects != 0 && (uint64_t)C * W * H != 0 && CN != 0;
    if (fwd_ws) {  // the forward's records and order; this workspace holds the packed gradients only
        UBS_REQUIRE(workspace, "a forward workspace needs a gradient workspace beside it");
        const size_t need = raster_workspace_bytes(C, N, 0), need_fwd = need + raster_order_bytes(C, tw, th);
        if (workspace_bytes < need || fwd_ws_bytes < need_fwd) {
            set_error("blend workspaces of %zu (gradients) and %zu (forward) bytes, %zu and %zu needed",
                      workspace_bytes, fwd_ws_bytes, need, need_fwd);
            return UBS_ERR_WORKSPACE;
        }
    } else if (workspace) {
        const size_t need = raster_workspace_bytes(C, N, 1);
        if (workspace_bytes < need) {
            set_error("blend workspace of %zu bytes, %zu needed", workspace_bytes, need);
            return UBS_ERR_WORKSPACE;
        }
    }
    const float4 *rec = nullptr;
    float *vrec = nullptr;
    const uint32_t *order = nullptr;
    const bool wide = packed && raster_wide_records(channels, tile_size);
    int e = wide ? UBS_OK : zero_async(v_colors, CN * channels * sizeof(float), st);
    if (packed && fwd_ws) {
        // the forward packed these records and wrote the tile order from the
        // same inputs: only the gradient records are cleared here
        vrec = static_cast<float *>(workspace);
        if (!e) e = zero_async(vrec, CN * (wide ? kWideRecordFloats : kRecordFloats) * sizeof(float), st);
        if (e) return e;
        rec = static_cast<const float4 *>(fwd_ws);
        order = reinterpret_cast<const uint32_t *>(static_cast<const char *>(fwd_ws) + raster_workspace_bytes(C, N, 0));
    } else if (packed) {
        // records in the first half of the workspace, packed gradients in the
        // second; the scalar gradient tensors are written by the unpack below
        float4 *r = static_cast<float4 *>(workspace);
        vrec = reinterpret_cast<float *>(static_cast<char *>(workspace) + raster_workspace_bytes(C, N, 0));
        if (!e) e = zero_async(vrec, CN * (wide ? kWideRecordFloats : kRecordFloats) * sizeof(float), st);
        if (e) return e;
        UBS_REQUIRE(means2d && conics && opacities && betas, "null pointer");
        raster_pack_kernel<<<grid_for((int64_t)CN), kThreads, 0, st>>>(CN, means2d, conics, opacities, betas, r,
                                                                       wide ? colors : nullptr, channels);
        UBS_LAUNCH_CHECK("raster_pack_kernel");
        rec = r;
        const size_t base = raster_workspace_bytes(C, N, 1);
        if (tile_offsets && workspace_bytes >= base + raster_order_bytes(C, tw, th)) {  // room for the block order
            uint32_t *o = reinterpret_cast<uint32_t *>(static_cast<char *>(workspace) + base);
            raster_block_order_kernel<<<1, kOrderThreads, 0, st>>>(C, tw, th, tile_offsets, n_isects, nullptr, o);
            UBS_LAUNCH_CHECK("raster_block_order_kernel");
            order = o;
        }
    } else {
        if (!e) e = zero_async(v_means2d, CN * 2 * sizeof(float), st);
        if (!e) e = zero_async(v_conics, CN * 3 * sizeof(float), st);
        if (!e) e = zero_async(v_opacities, CN * sizeof(float), st);
        if (!e) e = zero_async(v_betas, CN * sizeof(float), st);
        if (e) return e;
    }
    if (n_isects == 0 || (uint64_t)C * W * H == 0) return UBS_OK;
    UBS_REQUIRE((uint64_t)C * th * tw < (1ull << 31), "too many tiles");
    UBS_REQUIRE(means2d && conics && colors && opacities && betas && tile_offsets && flatten_ids && last_ids &&
                    (render_alphas || T_final) && v_render_colors,
                "null pointer");
    if (T_ckpt) {
        const size_t need = ckpt_floats(C, n_isects, tw, th, tile_size);
        if (ckpt_n < need) {
            set_error("transmittance checkpoint buffer of %zu floats, %zu needed", ckpt_n, need);
            return UBS_ERR_WORKSPACE;
        }
    }
    const bool f64t = g_f64_transmittance != 0;
    uint32_t coff = 0;
    while (coff < channels) {
        const int cd = pick_chunk(channels - coff, tile_size);
        if (tile_size == 16)
            e = launch_bwd_chunk<16>(cd, f64t, C, N, n_isects, channels, coff, means2d, conics, colors, opacities,
                                     betas, backgrounds, masks, W, H, tw, th, tile_offsets, flatten_ids,
                                     render_alphas, last_ids, T_final, v_render_colors, v_render_alphas, v_means2d,
                                     v_conics, v_colors, v_opacities, v_betas, rec, vrec, T_ckpt, order, st);
        else if (tile_size == 8)
            e = launch_bwd_chunk<8>(cd, f64t, C, N, n_isects, channels, coff, means2d, conics, colors, opacities,
                                    betas, backgrounds, masks, W, H, tw, th, tile_offsets, flatten_ids,
                                    render_alphas, last_ids, T_final, v_render_colors, v_render_alphas, v_means2d,
                                    v_conics, v_colors, v_opacities, v_betas, rec, vrec, T_ckpt, order, st);
        else
            e = launch_bwd_chunk<32>(cd, f64t, C, N, n_isects, channels, coff, means2d, conics, colors, opacities,
                                     betas, backgrounds, masks, W, H, tw, th, tile_offsets, flatten_ids,
                                     render_alphas, last_ids, T_final, v_render_colors, v_render_alphas, v_means2d,
                                     v_conics, v_colors, v_opacities, v_betas, rec, vrec, T_ckpt, order, st);
        if (e) return e;
        coff += (uint32_t)cd;
    }
    if (packed) {
        raster_unpack_grads_kernel<<<grid_for((int64_t)CN), kThreads, 0, st>>>(
            CN, reinterpret_cast<const float4 *>(vrec), v_means2d, v_conics, v_opacities, v_betas,
            wide ? v_colors : nullptr, channels);
        UBS_LAUNCH_CHECK("raster_unpack_grads_kernel");
    }
    return UBS_OK;
}

int ubs_rasterize_to_pixels_bwd(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels, const float *means2d,
                                const float *conics, const float *colors, const float *opacities,
                                const float *betas, const float *backgrounds, const uint8_t *masks, uint32_t W,
                                uint32_t H, uint32_t tile_size, uint32_t tw, uint32_t th,
                                const int32_t *tile_offsets, const int32_t *flatten_ids, const float *render_alphas,
                                const int32_t *last_ids, const float *T_final, const float *v_render_colors,
                                const float *v_render_alphas, float *v_means2d, float *v_conics, float *v_colors,
                                float *v_opacities, float *v_betas, void *workspace, size_t workspace_bytes,
                                void *stream) {
    return raster_bwd_impl(C, N, n_isects, channels, means2d, conics, colors, opacities, betas, backgrounds, masks, W,
                           H, tile_size, tw, th, tile_offsets, flatten_ids, render_alphas, last_ids, T_final,
                           v_render_colors, v_render_alphas, v_means2d, v_conics, v_colors, v_opacities, v_betas,
                           nullptr, 0, workspace, workspace_bytes, stream);
}

int ubs_rasterize_to_pixels_bwd_reuse(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels,
                                      const float *means2d, const float *conics, const float *colors,
                                      const float *opacities, const float *betas, const float *backgrounds,
                                      const uint8_t *masks, uint32_t W, uint32_t H, uint32_t tile_size, uint32_t tw,
                                      uint32_t th, const int32_t *tile_offsets, const int32_t *flatten_ids,
                                      const float *render_alphas, const int32_t *last_ids, const float *T_final,
                                      const float *T_ckpt, size_t ckpt_n, const float *v_render_colors,
                                      const float *v_render_alphas, float *v_means2d, float *v_conics,
                                      float *v_colors, float *v_opacities, float *v_betas, const void *fwd_workspace,
                                      size_t fwd_workspace_bytes, void *workspace, size_t workspace_bytes,
                                      void *stream) {
    UBS_REQUIRE(fwd_workspace, "null forward workspace");
    return raster_bwd_impl(C, N, n_isects, channels, means2d, conics, colors, opacities, betas, backgrounds, masks, W,
                           H, tile_size, tw, th, tile_offsets, flatten_ids, render_alphas, last_ids, T_final,
                           v_render_colors, v_render_alphas, v_means2d, v_conics, v_colors, v_opacities, v_betas,
                           T_ckpt, ckpt_n, workspace, workspace_bytes, stream, fwd_workspace, fwd_workspace_bytes);
}

int ubs_rasterize_to_pixels_bwd_ckpt(uint32_t C, uint32_t N, int64_t n_isects, uint32_t channels,
                                     const float *means2d, const float *conics, const float *colors,
                                     const float *opacities, const float *betas, const float *backgrounds,
                                     const uint8_t *masks, uint32_t W, uint32_t H, uint32_t tile_size, uint32_t tw,
                                     uint32_t th, const int32_t *tile_offsets, const int32_t *flatten_ids,
                                     const float *render_alphas, const int32_t *last_ids, const float *T_final,
                                     const float *T_ckpt, size_t ckpt_n, const float *v_render_colors,
                                     const float *v_render_alphas, float *v_means2d, float *v_conics,
                                     float *v_colors, float *v_opacities, float *v_betas, void *workspace,
                                     size_t workspace_bytes, void *stream) {
    UBS_REQUIRE(T_ckpt || n_isects == 0, "null checkpoint buffer");
    return raster_bwd_impl(C, N, n_isects, channels, means2d, conics, colors, opacities, betas, backgrounds, masks, W,
                           H, tile_size, tw, th, tile_offsets, flatten_ids, render_alphas, last_ids, T_final,
                           v_render_colors, v_render_alphas, v_means2d, v_conics, v_colors, v_opacities, v_betas,
                           T_ckpt, ckpt_n, workspace, workspace_bytes, stream);
}

#ifdef UBS_PROBE_BUILD
// Measurement build only (libubs_probe.so): the 16x16-tile, 3-channel blend
// backward with the STATS counters on the library's packed path (its
// gradients stay packed in the workspace); stats: 7 zeroed uint64.
__attribute__((visibility("default"))) int ubs_probe_raster_bwd_stats(
    uint32_t C, uint32_t N, int64_t n_isects, const float *means2d, const float *conics, const float *colors,
    const float *opacities, const float *betas, uint32_t W, uint32_t H, uint32_t tw, uint32_t th,
    const int32_t *tile_offsets, const int32_t *flatten_ids, const int32_t *last_ids, const float *T_final,
    const float *v_render_colors, const float *v_render_alphas, float *v_colors, void *workspace,
    unsigned long long *stats, void *stream) {
    hipStream_t st = as_stream(stream);
    const uint64_t CN = (uint64_t)C * N;
    float4 *r = static_cast<float4 *>(workspace);
    float *vrec = reinterpret_cast<float *>(static_cast<char *>(workspace) + raster_workspace_bytes(C, N, 0));
    int e = zero_async(v_colors, CN * 3 * sizeof(float), st);
    if (!e) e = zero_async(vrec, CN * kWideRecordFloats * sizeof(float), st);
    if (e) return e;
    raster_pack_kernel<<<grid_for((int64_t)CN), kThreads, 0, st>>>(CN, means2d, conics, opacities, betas, r,
                                                                   colors, 3);
    raster_bwd_kernel<16, 3, false, true><<<xcd_grid(C, tw, th), 256, 0, st>>>(
        C, N, n_isects, 3, 0, means2d, conics, colors, opacities, betas, nullptr, nullptr, W, H, tw, th,
        tile_offsets, flatten_ids, nullptr, last_ids, T_final, v_render_colors, v_render_alphas, nullptr, nullptr,
        v_colors, nullptr, nullptr, r, 1, vrec, nullptr, stats);
    return (int)hipGetLastError();
}

// The same backward through `order` (null: raster order) writing per
// workgroup its start, end, tile and hardware id into wg_times
// [xcd_grid * 4] (zeroed); stats may be null (timing only).
__attribute__((visibility("default"))) int ubs_probe_raster_bwd_times(
    uint32_t C, uint32_t N, int64_t n_isects, const float *means2d, const float *conics, const float *colors,
    const float *opacities, const float *betas, uint32_t W, uint32_t H, uint32_t tw, uint32_t th,
    const int32_t *tile_offsets, const int32_t *flatten_ids, const int32_t *last_ids, const float *T_final,
    const float *v_render_colors, const float *v_render_alphas, float *v_colors, void *workspace,
    const uint32_t *order, unsigned long long *stats, unsigned long long *wg_times, void *stream) {
    hipStream_t st = as_stream(stream);
    const uint64_t CN = (uint64_t)C * N;
    float4 *r = static_cast<float4 *>(workspace);
    float *vrec = reinterpret_cast<float *>(static_cast<char *>(workspace) + raster_workspace_bytes(C, N, 0));
    int e = zero_async(v_colors, CN * 3 * sizeof(float), st);
    if (!e) e = zero_async(vrec, CN * kWideRecordFloats * sizeof(float), st);
    if (e) return e;
    raster_pack_kernel<<<grid_for((int64_t)CN), kThreads, 0, st>>>(CN, means2d, conics, opacities, betas, r,
                                                                   colors, 3);
    raster_bwd_kernel<16, 3, false, true><<<xcd_grid(C, tw, th), 256, 0, st>>>(
        C, N, n_isects, 3, 0, means2d, conics, colors, opacities, betas, nullptr, nullptr, W, H, tw, th,
        tile_offsets, flatten_ids, nullptr, last_ids, T_final, v_render_colors, v_render_alphas, nullptr, nullptr,
        v_colors, nullptr, nullptr, r, 1, vrec, nullptr, stats, order, wg_times);
    return (int)hipGetLastError();
}
#endif

}  // extern "C"

// Tile intersection on gfx950: per-primitive tile count, int64 inclusive scan
// (reduce-then-scan, three launches, no spin waits), key emission and the
// per-(camera, tile) range table.
//
// The tile box follows isect_tiles.cu:54-66 exactly: tile_radius and the
// centre are divided by tile_size in fp32, floor/ceil, converted with CUDA's
// saturating float->uint32 semantics and clamped to the grid.
#pragma clang fp contract(off)

#include "ubs_common.h"

namespace ubs {

constexpr int kScanItems = 4;                        // items per thread
constexpr int kScanTile = kThreads * kScanItems;     // elements per block

__device__ __forceinline__ void tile_box(float mx, float my, int32_t rad, uint32_t ts, uint32_t tw,
                                         uint32_t th, uint32_t &x0, uint32_t &y0, uint32_t &x1,
                                         uint32_t &y1) {
    const float tr = (float)rad / (float)ts;
    const float tx = mx / (float)ts, ty = my / (float)ts;
    x0 = min(sat_u32(floorf(tx - tr)), tw);
    y0 = min(sat_u32(floorf(ty - tr)), th);
    x1 = min(sat_u32(ceilf(tx + tr)), tw);
    y1 = min(sat_u32(ceilf(ty + tr)), th);
}

__device__ __forceinline__ int32_t tile_count(const float *__restrict__ means2d, const int32_t *__restrict__ radii,
                                              uint64_t idx, uint32_t ts, uint32_t tw, uint32_t th) {
    const int32_t r = radii[idx];
    if (r <= 0) return 0;
    uint32_t x0, y0, x1, y1;
    tile_box(means2d[idx * 2], means2d[idx * 2 + 1], r, ts, tw, th, x0, y0, x1, y1);
    return (int32_t)((y1 - y0) * (x1 - x0));
}

// Block-wide exclusive scan of one int64 per thread (256 threads, 4 waves).
__device__ __forceinline__ int64_t block_exclusive_scan(int64_t v, int64_t *lds_wave, int64_t &total) {
    const int lane = (int)lane_id(), wave = threadIdx.x >> 6;
    int64_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t u = __shfl_up(inc, o, 64);
        if (lane >= o) inc += u;
    }
    if (lane == 63) lds_wave[wave] = inc;
    __syncthreads();
    int64_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
        const int64_t s = lds_wave[w];
        if (w < wave) base += s;
        tot += s;
    }
    total = tot;
    return base + inc - v;
}

// max over the block of one uint32 per thread (256 threads); thread 0 gets it
__device__ __forceinline__ uint32_t block_max_u32(uint32_t v, uint32_t *lds_wave) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    if (lane_id() == 0) lds_wave[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < kThreads / 64; ++w) v = max(v, lds_wave[w]);
    return v;
}

// Pass 1a: tiles_per_gauss and one int64 partial sum per block; with depths,
// also the block's bit range of the visible depths as uint32 patterns
// (block_range[2b] = ~min, [2b+1] = max; ~min so that a max serves both).
__global__ void isect_count_kernel(uint64_t total, const float *__restrict__ means2d,
                                   const int32_t *__restrict__ radii, const float *__restrict__ depths,
                                   uint32_t ts, uint32_t tw, uint32_t th, int32_t *__restrict__ tpg,
                                   int64_t *__restrict__ block_sums, uint32_t *__restrict__ block_range,
                                   int64_t *__restrict__ dkeys = nullptr, int32_t *__restrict__ dvals = nullptr,
                                   float4 *__restrict__ prec = nullptr) {
    __shared__ int64_t lds_wave[kThreads / 64];
    __shared__ uint32_t lds_max[kThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    int64_t s = 0;
    uint32_t lo_n = 0u, hi = 0u;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t idx = base + (uint64_t)k * kThreads + threadIdx.x;  // coalesced
        if (idx < total) {
            const int32_t n = tile_count(means2d, radii, idx, ts, tw, th);
            tpg[idx] = n;
            s += n;
            if (depths && radii[idx] > 0) {
                const uint32_t b = __float_as_uint(depths[idx]);
                lo_n = max(lo_n, ~b);
                hi = max(hi, b);
            }
            if (dkeys) {  // the depth-first emission's sort input, as isect_depth_keys_kernel
                dkeys[idx] = radii[idx] > 0 ? (int64_t)__float_as_uint(depths[idx]) : (int64_t)0xFFFFFFFFu;
                dvals[idx] = (int32_t)idx;
            }
            if (prec)  // what the emission reads of the pair, in one 16-byte record
                prec[idx] = make_float4(means2d[idx * 2], means2d[idx * 2 + 1], __int_as_float(radii[idx]),
                                        depths[idx]);
        }
    }
    int64_t tot;
    block_exclusive_scan(s, lds_wave, tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
    if (depths) {  // block-uniform
        lo_n = block_max_u32(lo_n, lds_max);
        __syncthreads();
        hi = block_max_u32(hi, lds_max);
        if (threadIdx.x == 0) {
            block_range[2 * blockIdx.x] = lo_n;
            block_range[2 * blockIdx.x + 1] = hi;
        }
    }
}

// Pass 1b: int64 inclusive cumulative sum (blocked: each thread owns 4
// consecutive elements).  Each block takes its offset by adding the block
// sums before it itself (at most nb int64 reads, from L2), so no scan launch
// runs between the passes; block 0 also writes the grand total to *n_isects
// (when given), the depth range of all blocks to range[0..1] (when given)
// and *status_in (0 when null) to *status_out (when given).
__global__ void isect_cum_kernel(uint64_t total, const int32_t *__restrict__ tpg,
                                 const int64_t *__restrict__ block_sums, int64_t nb, int64_t *__restrict__ cum,
                                 const int32_t *__restrict__ order, int64_t *__restrict__ n_isects,
                                 const uint32_t *__restrict__ block_range, uint32_t *__restrict__ range,
                                 const int64_t *__restrict__ status_in, int64_t *__restrict__ status_out) {
    __shared__ int64_t lds_wave[kThreads / 64];
    __shared__ int64_t lds_off[kThreads / 64];
    __shared__ uint32_t lds_max[kThreads / 64];
    const bool first = blockIdx.x == 0;
    // the block's offset; block 0 sums all nb block sums: the grand total
    const int64_t upto = first && n_isects ? nb : (int64_t)blockIdx.x;
    int64_t part = 0;
    for (int64_t i = threadIdx.x; i < upto; i += kThreads) part += block_sums[i];
    int64_t red;
    block_exclusive_scan(part, lds_off, red);
    const int64_t offset = first ? 0 : red;
    if (first && n_isects && threadIdx.x == 0) *n_isects = red;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    int32_t v[kScanItems];
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        // (an order entry outside [0, total), only from a depth sort that gave
        // up, counts no tiles)
        const uint64_t o = base + k < total ? (order ? (uint64_t)(uint32_t)order[base + k] : base + k) : total;
        v[k] = o < total ? tpg[o] : 0;
        s += v[k];
    }
    int64_t tot;
    int64_t run = block_exclusive_scan(s, lds_wave, tot) + offset;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        run += v[k];
        if (base + k < total) cum[base + k] = run;
    }
    if (first && range) {  // block-uniform
        uint32_t lo_n = 0u, hi = 0u;
        for (int64_t i = threadIdx.x; i < nb; i += kThreads) {
            lo_n = max(lo_n, block_range[2 * i]);
            hi = max(hi, block_range[2 * i + 1]);
        }
        lo_n = block_max_u32(lo_n, lds_max);
        __syncthreads();
        hi = block_max_u32(hi, lds_max);
        if (threadIdx.x == 0) {
            range[0] = lo_n;
            range[1] = hi;
        }
    }
    if (first && status_out && threadIdx.x == 0) *status_out = status_in ? *status_in : 0;
}

// Depth-first emission (an order of the same entries that a stable sort on
// the camera/tile bits alone turns into the full sort order): the
// (camera, primitive) pairs sorted stably by depth, then emitted in that
// order.  Keys: the visible depths' bit patterns, 0xFFFFFFFF for culled
// pairs (they emit nothing); values: the flat ids.
__global__ void isect_depth_keys_kernel(uint64_t total, const int32_t *__restrict__ radii,
                                        const float *__restrict__ depths, int64_t *__restrict__ keys,
                                        int32_t *__restrict__ vals) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    keys[i] = radii[i] > 0 ? (int64_t)__float_as_uint(depths[i]) : (int64_t)0xFFFFFFFFu;
    vals[i] = (int32_t)i;
}

// Block sums of the tile counts taken in `order` (the scan's pass 1 for the
// depth-first emission).
__global__ void isect_ordered_sums_kernel(uint64_t total, const int32_t *__restrict__ tpg,
                                          const int32_t *__restrict__ order, int64_t *__restrict__ block_sums) {
    __shared__ int64_t lds_wave[kThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    int64_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t idx = base + (uint64_t)k * kThreads + threadIdx.x;
        if (idx < total) {
            const uint64_t o = (uint64_t)(uint32_t)order[idx];
            if (o < total) s += tpg[o];
        }
    }
    int64_t tot;
    block_exclusive_scan(s, lds_wave, tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// Pass 2: keys cam | tile | depth bits.  reference: isect_tiles.cu:76-96
// Primitives are taken in flat-id order, or in `order` (depth-first
// emission, with cum the prefix of the counts in that order).
// The 64 primitives of a wave own one contiguous run of the output (their
// counts' prefix from pass 1); the wave writes that run with consecutive
// lanes on consecutive entries: entry k of the run belongs to the last lane
// whose exclusive count prefix is <= k (a 6-step search over the lanes'
// prefixes by lane shuffles) and is that primitive's tile number
// k - prefix in row-major order of its box.  Same entries, same order as
// one thread per primitive writing its own box, with coalesced stores.
__global__ __launch_bounds__(256) void isect_emit_kernel(uint32_t C, uint32_t N, const float *__restrict__ means2d,
                                                         const int32_t *__restrict__ radii,
                                                         const float *__restrict__ depths,
                                                         const int64_t *__restrict__ cum, uint32_t ts, uint32_t tw,
                                                         uint32_t th, uint32_t tile_bits,
                                                         int64_t *__restrict__ isect_ids,
                                                         int32_t *__restrict__ flatten_ids,
                                                         const int32_t *__restrict__ order,
                                                         int64_t cap = INT64_MAX,
                                                         const float4 *__restrict__ prec = nullptr) {
    const uint64_t CN = (uint64_t)C * N;
    const uint64_t pos = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // place in emission order
    const int lane = (int)(threadIdx.x & 63);
    const uint64_t pos0 = pos - (uint64_t)lane;
    if (pos0 >= CN) return;  // wave-uniform
    const bool valid = pos < CN;
    uint64_t idx = !valid ? 0 : order ? (uint64_t)(uint32_t)order[pos] : pos;  // the flat (camera, primitive) id
    const bool in_range = idx < CN;  // (false only for an order from a depth sort that gave up: no tiles)
    if (!in_range) idx = 0;
    // the pair's centre, radius and depth: one 16-byte record, or three arrays
    const float4 pr = prec && valid && in_range ? prec[idx] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int32_t r = !(valid && in_range) ? 0 : prec ? __float_as_int(pr.z) : radii[idx];
    uint32_t x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    if (r > 0)
        tile_box(prec ? pr.x : means2d[idx * 2], prec ? pr.y : means2d[idx * 2 + 1], r, ts, tw, th, x0, y0, x1, y1);
    const uint32_t bw = x1 - x0;
    const uint32_t cnt = r > 0 ? (y1 - y0) * bw : 0u;
    // exclusive prefix of the counts over the wave
    uint32_t inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(inc, o, 64);
        if (lane >= o) inc += u;
    }
    const uint32_t excl = inc - cnt;
    const uint32_t total = __shfl(inc, 63, 64);
    const int64_t wave_start = pos0 == 0 ? 0 : cum[pos0 - 1];
    const int64_t cid = valid ? (int64_t)(idx / N) : 0;
    const float dep = !valid ? 0.f : prec ? pr.w : depths[idx];
    const int64_t key_hi = (cid << (32 + tile_bits)) | (int64_t)__float_as_int(dep);
    const int32_t fid = (int32_t)idx;
    const uint32_t rounds = (total + 63u) / 64u;
    for (uint32_t it = 0; it < rounds; ++it) {  // every lane runs every round: the shuffles see all lanes
        const uint32_t k = it * 64u + (uint32_t)lane;
        int p = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {
            const uint32_t v = __shfl(excl, p + step, 64);
            p = v <= k ? p + step : p;
        }
        const uint32_t pe = __shfl(excl, p, 64);
        const uint32_t pbw = __shfl(bw, p, 64);
        const uint32_t px0 = __shfl(x0, p, 64);
        const uint32_t py0 = __shfl(y0, p, 64);
        const int32_t pfid = __shfl(fid, p, 64);
        const int64_t pkey = __shfl(key_hi, p, 64);
        if (k < total && wave_start + (int64_t)k < cap) {  // entries past a capacity are not stored
            const uint32_t local = k - pe;
            const uint32_t q = local / pbw;
            const int64_t tile = (int64_t)(py0 + q) * tw + (px0 + (local - q * pbw));
            isect_ids[wave_start + k] = pkey | (tile << 32);
            flatten_ids[wave_start + k] = pfid;
        }
    }
}

// offsets[k] = number of sorted isects whose (cam, tile) index is < k.
// reference: isect_tiles.cu:287-333
// n_dev (may be null): the count in device memory, clamped to n_cap (the
// grid's size); a device count of 0 zero-fills the offsets.
constexpr int kOffItems = 4;  // consecutive ids per thread (two 16-byte loads)

__global__ void isect_offsets_kernel(int64_t n_cap, const int64_t *__restrict__ n_dev,
                                     const int64_t *__restrict__ ids, uint32_t tile_bits, int64_t n_tiles,
                                     int64_t total_tiles, int32_t *__restrict__ offsets) {
    const int64_t n = n_dev ? min(max(*n_dev, (int64_t)0), n_cap) : n_cap;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n == 0) {
        for (int64_t k = t; k < total_tiles; k += (int64_t)gridDim.x * blockDim.x) offsets[k] = 0;
        return;
    }
    const int64_t i0 = t * kOffItems;
    if (i0 >= n) return;
    const uint64_t mask = (1ull << tile_bits) - 1ull;
    auto flat = [&](int64_t id) {  // the (camera, tile) index of an isect id
        const uint64_t hi = (uint64_t)id >> 32;
        return (int64_t)(hi >> tile_bits) * n_tiles + (int64_t)(hi & mask);
    };
    int64_t v[kOffItems];
    if (i0 + kOffItems <= n) {
        const longlong2 a = reinterpret_cast<const longlong2 *>(ids + i0)[0];
        const longlong2 b = reinterpret_cast<const longlong2 *>(ids + i0)[1];
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    } else {
#pragma unroll
        for (int j = 0; j < kOffItems; ++j) v[j] = i0 + j < n ? ids[i0 + j] : 0;
    }
    // entry i starts every (camera, tile) after its predecessor's up to its own
    // (entry 0: from the first); the last entry ends the rest with n
    int64_t prev = i0 > 0 ? flat(ids[i0 - 1]) : -1;
#pragma unroll
    for (int j = 0; j < kOffItems; ++j) {
        const int64_t i = i0 + j;
        if (i >= n) break;
        const int64_t cur = flat(v[j]);
        for (int64_t k = prev + 1; k <= cur && k < total_tiles; ++k) offsets[k] = (int32_t)i;
        if (i == n - 1)
            for (int64_t k = cur + 1; k < total_tiles; ++k) offsets[k] = (int32_t)n;
        prev = cur;
    }
}

}  // namespace ubs

using namespace ubs;

extern "C" {

size_t ubs_isect_count_workspace_bytes(uint32_t C, uint32_t N) {
    const uint64_t total = (uint64_t)C * N;
    const uint64_t nb = (total + kScanTile - 1) / kScanTile;
    // per block: the int64 partial sum, then the (~min, max) depth bits
    return (size_t)std::max<uint64_t>(nb, 1) * (sizeof(int64_t) + 2 * sizeof(uint32_t));
}

uint32_t ubs_isect_prefix_bits(uint32_t C, uint32_t tw, uint32_t th) {
    return bit_width_u32(tw * th) + bit_width_u32(C);
}

int ubs_isect_tiles_count(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii, uint32_t ts,
                          uint32_t tw, uint32_t th, int32_t *tpg, int64_t *cum, int64_t *n_isects_dev,
                          void *workspace, size_t ws_bytes, void *stream) {
    return ubs_isect_tiles_count_range(C, N, means2d, radii, nullptr, ts, tw, th, tpg, cum, n_isects_dev, nullptr,
                                       workspace, ws_bytes, stream);
}

static int count_impl(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii, const float *depths,
                      uint32_t ts, uint32_t tw, uint32_t th, int32_t *tpg, int64_t *cum, int64_t *n_isects_dev,
                      uint32_t *range_dev, void *workspace, size_t ws_bytes, void *stream,
                      const int64_t *status_in, int64_t *status_out, int64_t *dkeys = nullptr,
                      int32_t *dvals = nullptr, float4 *prec = nullptr);

int ubs_isect_tiles_count_range(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii,
                                const float *depths, uint32_t ts, uint32_t tw, uint32_t th, int32_t *tpg,
                                int64_t *cum, int64_t *n_isects_dev, uint32_t *range_dev, void *workspace,
                                size_t ws_bytes, void *stream) {
    return count_impl(C, N, means2d, radii, depths, ts, tw, th, tpg, cum, n_isects_dev, range_dev, workspace,
                      ws_bytes, stream, nullptr, nullptr);
}

int ubs_isect_tiles_count_info(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii,
                               const float *depths, uint32_t ts, uint32_t tw, uint32_t th, int32_t *tpg,
                               int64_t *cum, int64_t *info, const int64_t *status_in, int64_t *depth_keys,
                               int32_t *depth_vals, void *pair_records, void *workspace, size_t ws_bytes,
                               void *stream) {
    UBS_REQUIRE(info && depths, "null pointer");
    UBS_REQUIRE(!depth_keys == !depth_vals, "depth_keys and depth_vals go together");
    return count_impl(C, N, means2d, radii, depths, ts, tw, th, tpg, cum, info,
                      reinterpret_cast<uint32_t *>(info + 1), workspace, ws_bytes, stream, status_in, info + 2,
                      depth_keys, depth_vals, static_cast<float4 *>(pair_records));
}

static int count_impl(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii, const float *depths,
                      uint32_t ts, uint32_t tw, uint32_t th, int32_t *tpg, int64_t *cum, int64_t *n_isects_dev,
                      uint32_t *range_dev, void *workspace, size_t ws_bytes, void *stream,
                      const int64_t *status_in, int64_t *status_out, int64_t *dkeys, int32_t *dvals,
                      float4 *prec) {
    hipStream_t st = as_stream(stream);
    const uint64_t total = (uint64_t)C * N;
    UBS_REQUIRE(n_isects_dev, "n_isects_dev is null");
    UBS_REQUIRE(ts > 0, "tile_size must be > 0");
    UBS_REQUIRE(bit_width_u32(tw * th) + bit_width_u32(C) <= 32, "camera and tile ids need more than 32 bits");
    UBS_REQUIRE(!range_dev || depths, "range_dev needs depths");
    if (total == 0) {  // the totals (0), the range (0, 0) and the status word
        isect_cum_kernel<<<1, kThreads, 0, st>>>(0, nullptr, nullptr, 0, nullptr, nullptr, n_isects_dev, nullptr,
                                                 range_dev, status_in, status_out);
        UBS_LAUNCH_CHECK("isect_cum_kernel");
        return UBS_OK;
    }
    UBS_REQUIRE(total < (1ull << 31), "C*N must be < 2^31");
    UBS_REQUIRE(means2d && radii && tpg && cum, "null pointer");
    const uint64_t nb = (total + kScanTile - 1) / kScanTile;
    if (!workspace || ws_bytes < ubs_isect_count_workspace_bytes(C, N)) {
        set_error("isect workspace too small (%zu < %zu)", ws_bytes, ubs_isect_count_workspace_bytes(C, N));
        return UBS_ERR_WORKSPACE;
    }
    int64_t *bs = reinterpret_cast<int64_t *>(workspace);
    uint32_t *br = reinterpret_cast<uint32_t *>(bs + nb);
    const float *dp = range_dev ? depths : nullptr;
    UBS_REQUIRE(!(dkeys || prec) || depths, "depth keys and pair records need depths");
    isect_count_kernel<<<(unsigned)nb, kThreads, 0, st>>>(total, means2d, radii, dp ? dp : (dkeys ? depths : nullptr),
                                                         ts, tw, th, tpg, bs, br, dkeys, dvals, prec);
    UBS_LAUNCH_CHECK("isect_count_kernel");
    isect_cum_kernel<<<(unsigned)nb, kThreads, 0, st>>>(total, tpg, bs, (int64_t)nb, cum, nullptr, n_isects_dev,
                                                        range_dev ? br : nullptr, range_dev, status_in, status_out);
    UBS_LAUNCH_CHECK("isect_cum_kernel");
    return UBS_OK;
}

int ubs_isect_tiles_emit(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii, const float *depths,
                         const int64_t *cum, uint32_t ts, uint32_t tw, uint32_t th, int64_t *isect_ids,
                         int32_t *flatten_ids, void *stream) {
    const uint64_t total = (uint64_t)C * N;
    if (total == 0) return UBS_OK;
    UBS_REQUIRE(total < (1ull << 31), "C*N must be < 2^31");
    UBS_REQUIRE(means2d && radii && depths && cum && isect_ids && flatten_ids, "null pointer");
    const uint32_t tile_bits = bit_width_u32(tw * th);
    isect_emit_kernel<<<grid_for((int64_t)total), kThreads, 0, as_stream(stream)>>>(
        C, N, means2d, radii, depths, cum, ts, tw, th, tile_bits, isect_ids, flatten_ids, nullptr);
    UBS_LAUNCH_CHECK("isect_emit_kernel");
    return UBS_OK;
}

int ubs_isect_depth_keys(uint32_t C, uint32_t N, const int32_t *radii, const float *depths, int64_t *keys,
                         int32_t *vals, void *stream) {
    const uint64_t total = (uint64_t)C * N;
    if (total == 0) return UBS_OK;
    UBS_REQUIRE(total < (1ull << 31), "C*N must be < 2^31");
    UBS_REQUIRE(radii && depths && keys && vals, "null pointer");
    isect_depth_keys_kernel<<<grid_for((int64_t)total), kThreads, 0, as_stream(stream)>>>(total, radii, depths, keys,
                                                                                         vals);
    UBS_LAUNCH_CHECK("isect_depth_keys_kernel");
    return UBS_OK;
}

size_t ubs_isect_ordered_workspace_bytes(uint32_t C, uint32_t N) {
    const uint64_t total = (uint64_t)C * N;
    const uint64_t nb = (total + kScanTile - 1) / kScanTile;
    return (size_t)(std::max<uint64_t>(nb, 1) + 1) * sizeof(int64_t);
}

int ubs_isect_scan_ordered(uint32_t C, uint32_t N, const int32_t *tpg, const int32_t *order, int64_t *cum,
                           void *workspace, size_t ws_bytes, void *stream) {
    hipStream_t st = as_stream(stream);
    const uint64_t total = (uint64_t)C * N;
    if (total == 0) return UBS_OK;
    UBS_REQUIRE(total < (1ull << 31), "C*N must be < 2^31");
    UBS_REQUIRE(tpg && order && cum, "null pointer");
    if (!workspace || ws_bytes < ubs_isect_ordered_workspace_bytes(C, N)) {
        set_error("isect workspace too small (%zu < %zu)", ws_bytes, ubs_isect_ordered_workspace_bytes(C, N));
        return UBS_ERR_WORKSPACE;
    }
    const uint64_t nb = (total + kScanTile - 1) / kScanTile;
    int64_t *bs = reinterpret_cast<int64_t *>(workspace);
    isect_ordered_sums_kernel<<<(unsigned)nb, kThreads, 0, st>>>(total, tpg, order, bs);
    UBS_LAUNCH_CHECK("isect_ordered_sums_kernel");
    isect_cum_kernel<<<(unsigned)nb, kThreads, 0, st>>>(total, tpg, bs, (int64_t)nb, cum, order, nullptr, nullptr,
                                                        nullptr, nullptr, nullptr);
    UBS_LAUNCH_CHECK("isect_cum_kernel");
    return UBS_OK;
}

static int emit_ordered(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii, const float *depths,
                        const int32_t *order, const int64_t *cum, uint32_t ts, uint32_t tw, uint32_t th,
                        int64_t *isect_ids, int32_t *flatten_ids, int64_t cap, void *stream,
                        const float4 *prec = nullptr);

int ubs_isect_tiles_emit_ordered_rec(uint32_t C, uint32_t N, const void *pair_records, const int32_t *order,
                                     const int64_t *cum, uint32_t ts, uint32_t tw, uint32_t th, int64_t *isect_ids,
                                     int32_t *flatten_ids, int64_t cap, void *stream) {
    UBS_REQUIRE(cap >= 0, "cap must be >= 0");
    UBS_REQUIRE(pair_records || (uint64_t)C * N == 0, "null pointer");
    return emit_ordered(C, N, nullptr, nullptr, nullptr, order, cum, ts, tw, th, isect_ids, flatten_ids, cap, stream,
                        static_cast<const float4 *>(pair_records));
}

int ubs_isect_tiles_emit_ordered(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii,
                                 const float *depths, const int32_t *order, const int64_t *cum, uint32_t ts,
                                 uint32_t tw, uint32_t th, int64_t *isect_ids, int32_t *flatten_ids, void *stream) {
    return emit_ordered(C, N, means2d, radii, depths, order, cum, ts, tw, th, isect_ids, flatten_ids, INT64_MAX,
                        stream);
}

int ubs_isect_tiles_emit_ordered_cap(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii,
                                     const float *depths, const int32_t *order, const int64_t *cum, uint32_t ts,
                                     uint32_t tw, uint32_t th, int64_t *isect_ids, int32_t *flatten_ids,
                                     int64_t cap, void *stream) {
    UBS_REQUIRE(cap >= 0, "cap must be >= 0");
    return emit_ordered(C, N, means2d, radii, depths, order, cum, ts, tw, th, isect_ids, flatten_ids, cap, stream);
}

static int emit_ordered(uint32_t C, uint32_t N, const float *means2d, const int32_t *radii, const float *depths,
                        const int32_t *order, const int64_t *cum, uint32_t ts, uint32_t tw, uint32_t th,
                        int64_t *isect_ids, int32_t *flatten_ids, int64_t cap, void *stream, const float4 *prec) {
    const uint64_t total = (uint64_t)C * N;
    if (total == 0) return UBS_OK;
    UBS_REQUIRE(total < (1ull << 31), "C*N must be < 2^31");
    UBS_REQUIRE((prec || (means2d && radii && depths)) && order && cum && isect_ids && flatten_ids, "null pointer");
    const uint32_t tile_bits = bit_width_u32(tw * th);
    isect_emit_kernel<<<grid_for((int64_t)total), kThreads, 0, as_stream(stream)>>>(
        C, N, means2d, radii, depths, cum, ts, tw, th, tile_bits, isect_ids, flatten_ids, order, cap, prec);
    UBS_LAUNCH_CHECK("isect_emit_kernel");
    return UBS_OK;
}

int ubs_isect_offset_encode(int64_t n, const int64_t *ids, uint32_t C, uint32_t tw, uint32_t th, int32_t *offsets,
                            void *stream) {
    hipStream_t st = as_stream(stream);
    const int64_t n_tiles = (int64_t)tw * th, total_tiles = (int64_t)C * n_tiles;
    if (total_tiles == 0) return UBS_OK;
    UBS_REQUIRE(offsets, "null pointer");
    if (n == 0) return zero_async(offsets, (size_t)total_tiles * sizeof(int32_t), st);
    UBS_REQUIRE(ids, "null pointer");
    UBS_REQUIRE(n < (1ll << 31), "n_isects must be < 2^31");
    isect_offsets_kernel<<<grid_for((n + kOffItems - 1) / kOffItems), kThreads, 0, st>>>(
        n, nullptr, ids, bit_width_u32((uint32_t)n_tiles), n_tiles, total_tiles, offsets);
    UBS_LAUNCH_CHECK("isect_offsets_kernel");
    return UBS_OK;
}

int ubs_isect_offset_encode_dev(int64_t n_cap, const int64_t *n_dev, const int64_t *ids, uint32_t C, uint32_t tw,
                                uint32_t th, int32_t *offsets, void *stream) {
    hipStream_t st = as_stream(stream);
    const int64_t n_tiles = (int64_t)tw * th, total_tiles = (int64_t)C * n_tiles;
    if (total_tiles == 0) return UBS_OK;
    UBS_REQUIRE(offsets && n_dev && (ids || n_cap == 0), "null pointer");
    UBS_REQUIRE(n_cap >= 0 && n_cap < (1ll << 31), "n_cap must be in [0, 2^31)");
    isect_offsets_kernel<<<grid_for(std::max<int64_t>((n_cap + kOffItems - 1) / kOffItems, 1)), kThreads, 0, st>>>(
        n_cap, n_dev, ids, bit_width_u32((uint32_t)n_tiles), n_tiles, total_tiles, offsets);
    UBS_LAUNCH_CHECK("isect_offsets_kernel");
    return UBS_OK;
}

}  // extern "C"

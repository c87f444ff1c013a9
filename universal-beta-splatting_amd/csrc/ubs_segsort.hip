// Depth sort inside each tile's segment of the intersection list, on
// gfx950: the second half of the isect-id sort.
//
// The ids (cam << (32 + tile_bits) | tile << 32 | depth bits) must end up in
// the order of a stable LSD radix sort on bits [0, 32 + tile_bits +
// cam_bits), the contract of cub::DeviceRadixSort::SortPairs as
// isect_tiles.cu:224-278 calls it.  That order is reached in two steps: a
// stable sort on the camera/tile bits only (two onesweep passes over HBM),
// which leaves every (camera, tile) segment contiguous with its entries in
// emission order, then this kernel: one workgroup per segment sorts it
// stably on the 32 depth bits.  Equal depths keep emission order in both
// steps, so the result is bit-identical to the full LSD sort, and the
// depth passes touch each pair once from HBM instead of once per byte.
//
// A segment of up to 8192 pairs lives in LDS (depth bits and value); each
// 8-bit pass loads its pairs into registers, ranks them with wave64 ballots
// (stable: each wave a contiguous run, item order inside a lane's stripe,
// lane order), then scatters them back into the same buffer; the last pass
// writes HBM.  One 64 KB buffer lets two 512-thread workgroups share a
// CU.  Longer segments (one tile crossed by more than 8192 primitives) run the
// same ranking over 8192-pair chunks in HBM, the input and output buffers
// serving as ping-pong storage.  Passes whose depth byte is shared by every
// visible primitive (the caller knows them from the depth range) are left
// out, as in the global sort.
#include "ubs_common.h"

namespace ubs {

// 512 threads x 16 pairs: segments of up to 8192 pairs sort in LDS (at
// 1920x1080 with 1M primitives the longest tile list is ~5.2k and a fifth
// of the tiles hold more than 4096 entries, most of the pairs)
constexpr int kSegThreads = 512;
constexpr int kSegWaves = kSegThreads / 64;
constexpr int kSegItems = 16;
constexpr int kSegCap = kSegThreads * kSegItems;  // pairs per chunk
static_assert(kSegThreads >= 256, "one thread per digit");

// Rank the chunk's digits (key >> shift) & 255: rank[i] = number of
// earlier items of this wave (stripe order) with the same digit;
// cnt[d] (the wave's row of counts) ends as the wave's count of digit d.
// The wave's items are [wave_base, wave_end), item i of lane l at
// wave_base + 64 i + l.
__device__ __forceinline__ void rank_chunk(const uint32_t (&key)[kSegItems], int shift, int wave_end, int wave_base,
                                           int lane, volatile uint32_t *cnt, uint32_t (&rank)[kSegItems]) {
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int i = 0; i < kSegItems; ++i) {
        const int idx = wave_base + i * 64 + lane;
        const bool ok = idx < wave_end;
        uint64_t peers = __ballot(ok);
        if (peers == 0ull) {  // wave-uniform: nothing left in this stripe
            rank[i] = 0;
            continue;
        }
        const uint32_t d = (key[i] >> shift) & 255u;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t r = 0;
        if (ok) {
            const uint32_t before = (uint32_t)__popcll(peers & lt_mask);
            const uint32_t c0 = cnt[d];
            r = c0 + before;
            if (before == 0) cnt[d] = c0 + (uint32_t)__popcll(peers);
        }
        rank[i] = r;
    }
}

// Exclusive scan over the 256 digits of v (thread d holds digit d; threads
// from 256 on pass 0); returns the exclusive prefix.  Contains two barriers.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t *s_wsum, int wave, int lane) {
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(inc, o, 64);
        if (lane >= o) inc += u;
    }
    if (lane == 63) s_wsum[wave] = inc;
    __syncthreads();
    uint32_t wbase = 0;
#pragma unroll
    for (int w = 0; w < kSegWaves; ++w)
        if (w < wave) wbase += s_wsum[w];
    __syncthreads();
    return wbase + inc - v;
}

__global__ __launch_bounds__(kSegThreads) void segment_depth_sort_kernel(
    int64_t n, int64_t *__restrict__ keys_a, int32_t *__restrict__ vals_a, int64_t *__restrict__ keys_b,
    int32_t *__restrict__ vals_b, const int32_t *__restrict__ tile_offsets, uint32_t total_tiles,
    uint32_t tiles_per_cam, uint32_t tile_bits, uint32_t skip_passes) {
    __shared__ uint32_t s_k[kSegCap];  // depth bits (the bits above are the segment's U)
    __shared__ int32_t s_v[kSegCap];
    __shared__ uint32_t s_cnt[kSegWaves][256];
    __shared__ uint32_t s_start[4][256];
    __shared__ uint32_t s_run[256];
    __shared__ uint32_t s_wsum[kSegWaves];

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const bool digit_thread = tid < 256;  // thread d owns digit d in the per-digit steps
    const uint32_t cam = blockIdx.x / tiles_per_cam, tile = blockIdx.x % tiles_per_cam;
    const uint64_t U = ((uint64_t)cam << tile_bits) | tile;  // the segment's bits above the depth
    // the segment of this (camera, tile): [offsets[L], offsets[L + 1]) or to n
    const int64_t s = tile_offsets[blockIdx.x];
    const int64_t e = blockIdx.x + 1 < total_tiles ? (int64_t)tile_offsets[blockIdx.x + 1] : n;
    const int64_t len = e - s;
    if (len <= 0) return;

    int shifts[4], npass = 0;
#pragma unroll
    for (int p = 0; p < 4; ++p)
        if (!((skip_passes >> p) & 1u)) shifts[npass++] = 8 * p;
    // a chunk of m pairs is split into contiguous runs of whole 64-pair
    // stripes, one per wave (contiguous runs keep the ranking stable)
    auto wave_run = [&](int m, int &base, int &end) {
        const int per = ((m + kSegThreads - 1) / kSegThreads) * 64;
        base = min(m, wave * per);
        end = min(m, base + per);
    };
    auto clear_counts = [&]() {
        for (int k = tid; k < kSegWaves * 256; k += kSegThreads) (&s_cnt[0][0])[k] = 0;
    };
    // per digit: exclusive prefix of the waves' counts (in place) and the total
    auto wave_prefix = [&]() -> uint32_t {
        uint32_t tot = 0;
        if (digit_thread) {
#pragma unroll
            for (int w = 0; w < kSegWaves; ++w) {
                const uint32_t c = s_cnt[w][tid];
                s_cnt[w][tid] = tot;
                tot += c;
            }
        }
        return tot;
    };

    if (npass == 0) {
        for (int64_t j = tid; j < len; j += kSegThreads) {
            keys_b[s + j] = keys_a[s + j];
            vals_b[s + j] = vals_a[s + j];
        }
        return;
    }

    uint32_t key[kSegItems], rank[kSegItems];
    int32_t val[kSegItems];
    const uint32_t *klo = reinterpret_cast<const uint32_t *>(keys_a);  // low words: the depths

    if (len <= kSegCap) {
        // ---- the whole segment in LDS; each pass loads its items into
        // registers, ranks them, and scatters them back into the same buffer
        const int m = (int)len;
        int wave_base, wave_end;
        wave_run(m, wave_base, wave_end);
        // the first pass takes its items straight from HBM (all loads in flight at once)
#pragma unroll
        for (int i = 0; i < kSegItems; ++i) {
            const int idx = wave_base + i * 64 + lane;
            const bool ok = idx < wave_end;
            key[i] = ok ? klo[2 * (s + idx)] : 0u;
            val[i] = ok ? vals_a[s + idx] : 0;
        }
        for (int p = 0; p < npass; ++p) {
            const int sh = shifts[p];
            clear_counts();
            __syncthreads();  // (also orders the previous pass's scatter before these loads)
            if (p > 0) {
#pragma unroll
                for (int i = 0; i < kSegItems; ++i) {
                    const int idx = wave_base + i * 64 + lane;
                    const bool ok = idx < wave_end;
                    key[i] = ok ? s_k[idx] : 0u;
                    val[i] = ok ? s_v[idx] : 0;
                }
            }
            rank_chunk(key, sh, wave_end, wave_base, lane, s_cnt[wave], rank);
            __syncthreads();  // every item is in registers: the buffer may be overwritten
            const uint32_t tot = wave_prefix();
            const uint32_t start = block_excl_scan256(tot, s_wsum, wave, lane);
            if (digit_thread) s_start[0][tid] = start;
            __syncthreads();
            const bool last = p == npass - 1;
#pragma unroll
            for (int i = 0; i < kSegItems; ++i) {
                const int idx = wave_base + i * 64 + lane;
                if (idx < wave_end) {
                    const uint32_t d = (key[i] >> sh) & 255u;
                    const uint32_t pos = s_start[0][d] + s_cnt[wave][d] + rank[i];  // < m by construction
                    if (last) {
                        keys_b[s + pos] = (int64_t)((U << 32) | key[i]);
                        vals_b[s + pos] = val[i];
                    } else {
                        s_k[pos] = key[i];
                        s_v[pos] = val[i];
                    }
                }
            }
            __syncthreads();  // the scatter's reads of s_start / s_cnt end before the next pass clears them
        }
        return;
    }

    // ---- a long segment: chunks of kSegCap pairs in HBM, A and B as ping-pong.
    // Digit totals of every pass from one read of the segment.
    for (int k = tid; k < 4 * 256; k += kSegThreads) (&s_start[0][0])[k] = 0;
    __syncthreads();
    for (int64_t c0 = 0; c0 < len; c0 += kSegCap) {
#pragma unroll
        for (int i = 0; i < kSegItems; ++i) {
            const int64_t j = c0 + i * kSegThreads + tid;
            key[i] = j < len ? klo[2 * (s + j)] : 0u;
        }
#pragma unroll
        for (int i = 0; i < kSegItems; ++i) {
            if (c0 + i * kSegThreads + tid < len)
                for (int p = 0; p < npass; ++p) atomicAdd(&s_start[p][(key[i] >> shifts[p]) & 255u], 1u);
        }
    }
    __syncthreads();
    int64_t *kin = keys_a, *kout = keys_b;
    int32_t *vin = vals_a, *vout = vals_b;
    for (int p = 0; p < npass; ++p) {
        const int sh = shifts[p];
        const uint32_t start = block_excl_scan256(digit_thread ? s_start[p][tid] : 0u, s_wsum, wave, lane);
        if (digit_thread) {
            s_start[p][tid] = start;
            s_run[tid] = 0;
        }
        const uint32_t *kin_lo = reinterpret_cast<const uint32_t *>(kin);
        for (int64_t c0 = 0; c0 < len; c0 += kSegCap) {
            const int m = (int)std::min<int64_t>(kSegCap, len - c0);
            int wave_base, wave_end;
            wave_run(m, wave_base, wave_end);
            clear_counts();
            __syncthreads();
#pragma unroll
            for (int i = 0; i < kSegItems; ++i) {
                const int idx = wave_base + i * 64 + lane;
                const bool ok = idx < wave_end;
                key[i] = ok ? kin_lo[2 * (s + c0 + idx)] : 0u;
                val[i] = ok ? vin[s + c0 + idx] : 0;
            }
            rank_chunk(key, sh, wave_end, wave_base, lane, s_cnt[wave], rank);
            __syncthreads();
            const uint32_t tot = wave_prefix();
            __syncthreads();
#pragma unroll
            for (int i = 0; i < kSegItems; ++i) {
                const int idx = wave_base + i * 64 + lane;
                if (idx < wave_end) {
                    const uint32_t d = (key[i] >> sh) & 255u;
                    const uint32_t pos = s_start[p][d] + s_run[d] + s_cnt[wave][d] + rank[i];
                    kout[s + pos] = (int64_t)((U << 32) | key[i]);
                    vout[s + pos] = val[i];
                }
            }
            __syncthreads();
            if (digit_thread) s_run[tid] += tot;
        }
        __syncthreads();
        int64_t *tk = kin;
        kin = kout;
        kout = tk;
        int32_t *tv = vin;
        vin = vout;
        vout = tv;
    }
    if (kin != keys_b) {  // an even number of passes left the result in A
        for (int64_t j = tid; j < len; j += kSegThreads) {
            keys_b[s + j] = keys_a[s + j];
            vals_b[s + j] = vals_a[s + j];
        }
    }
}

}  // namespace ubs

using namespace ubs;

extern "C" int ubs_sort_tile_segments(int64_t n, int64_t *keys_in, int32_t *vals_in, int64_t *keys_out,
                                      int32_t *vals_out, const int32_t *tile_offsets, uint32_t C,
                                      uint32_t tile_width, uint32_t tile_height, uint32_t skip_passes,
                                      void *stream) {
    UBS_REQUIRE(n >= 0 && n < (1ll << 31), "n=%lld out of range", (long long)n);
    UBS_REQUIRE(skip_passes < 16u, "skip_passes covers the 4 depth bytes only");
    const uint64_t tiles = (uint64_t)tile_width * tile_height;
    UBS_REQUIRE((uint64_t)C * tiles < (1ull << 31), "too many tiles");
    if (n == 0 || C == 0 || tiles == 0) return UBS_OK;
    UBS_REQUIRE(keys_in && vals_in && keys_out && vals_out && tile_offsets, "null pointer");
    UBS_REQUIRE(keys_in != keys_out && vals_in != vals_out, "input and output buffers must differ");
    const uint32_t tile_bits = bit_width_u32((uint32_t)tiles);
    segment_depth_sort_kernel<<<(unsigned)(C * tiles), kSegThreads, 0, as_stream(stream)>>>(
        n, keys_in, vals_in, keys_out, vals_out, tile_offsets, (uint32_t)(C * tiles), (uint32_t)tiles, tile_bits,
        skip_passes);
    UBS_LAUNCH_CHECK("segment_depth_sort_kernel");
    return UBS_OK;
}

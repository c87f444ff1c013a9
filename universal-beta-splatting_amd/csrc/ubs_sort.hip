// Onesweep LSD radix sort of (int64 key, int32 value) pairs on gfx950.
//
// Stable ordering on key bits [begin_bit, end_bit), the contract of
// cub::DeviceRadixSort::SortPairs as the reference calls it
// (isect_tiles.cu:224-278).  Design:
//   * one histogram launch counts the 8-bit digits of every pass at once
//     (keys read once), a tiny launch turns them into per-pass exclusive
//     digit offsets;
//   * one launch per pass: a workgroup takes the next tile id from an atomic
//     counter (so every tile it waits on is already running), ranks its 6144
//     keys with wave64 ballots (match-by-bits, stable: lane order inside an
//     item, item order inside a wave, wave order inside the tile), publishes
//     its digit counts, stages the tile in LDS in digit order, resolves its
//     global prefix by decoupled look-back over earlier tiles, and writes
//     out runs of equal digits as contiguous (coalesced) segments.
//   * every look-back wait is bounded: past the spin limit it raises the
//     status word and stops waiting (the result is then invalid and the
//     caller is told through ubs_radix_sort_status).
#include <mutex>

#include "ubs_common.h"

namespace ubs {

constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
// 24 keys per thread: fewer, larger tiles amortise the per-tile digit scan
// and look-back (measured 0.86 ms per 12.3M-pair 5-pass sort vs 0.93 ms at
// 16 and 1.2 ms at 28-32, where the LDS staging leaves one workgroup per
// CU; for the 2-pass camera/tile sort 12 to 24 measure alike).  Build with
// -DUBS_SORT_ITEMS=n to try another size.
#ifndef UBS_SORT_ITEMS
#define UBS_SORT_ITEMS 24
#endif
// A sort of fewer keys than kSmallSortN (the C*N depth keys of a view) takes
// fewer keys per thread: at 24 a 1M-key sort has 163 tiles, under one per
// CU, and each pass waits on one workgroup's serial ranking (12 measured
// best of 4 / 8 / 12 at 1M keys, by a few microseconds per pass).
#ifndef UBS_SORT_ITEMS_SMALL
#define UBS_SORT_ITEMS_SMALL 12
#endif
#ifndef UBS_SORT_SMALL_N
#define UBS_SORT_SMALL_N (1 << 22)
#endif
constexpr int kItems = UBS_SORT_ITEMS;
constexpr int kItemsSmall = UBS_SORT_ITEMS_SMALL;
constexpr int64_t kSmallSortN = UBS_SORT_SMALL_N;

inline int sort_tile_keys(int64_t n) { return kSortThreads * (n < kSmallSortN ? kItemsSmall : kItems); }
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;
constexpr uint32_t kSpinLimit = 1u << 22;
// predecessor slots read per look-back round trip
#ifndef UBS_SORT_LOOKBACK
#define UBS_SORT_LOOKBACK 8
#endif
constexpr int kLookback = UBS_SORT_LOOKBACK;

// Digit of pass p.  Byte-aligned (pass p: bits [8p, 8p + 8) of the range)
// when passes may be skipped (the skip mask names depth bytes); otherwise the
// range is split evenly, 14 bits as 7 + 7 rather than 8 + 6 (0.602 against
// 0.606 ms for the isect stage at 1080p / 1M: the first pass writes runs of
// equal digits twice as long).
#ifndef UBS_SORT_BALANCED
#define UBS_SORT_BALANCED 1
#endif
__host__ __device__ inline void pass_digit(int p, int begin_bit, int end_bit, bool balanced, int &shift, int &bits) {
    const int total = end_bit - begin_bit;
    if (UBS_SORT_BALANCED && balanced) {
        const int passes = (total + kRadixBits - 1) / kRadixBits;
        const int base = total / passes, extra = total % passes;
        shift = begin_bit + p * base + (p < extra ? p : extra);
        bits = base + (p < extra ? 1 : 0);
    } else {
        shift = begin_bit + p * kRadixBits;
        bits = total - p * kRadixBits < kRadixBits ? total - p * kRadixBits : kRadixBits;
    }
}

struct SortLayout {
    size_t status_off, hist_off, counter_off, lookback_off, total;
    int passes;
    int64_t tiles;
};

static SortLayout sort_layout(int64_t n, int begin_bit, int end_bit) {
    SortLayout L{};
    L.passes = end_bit > begin_bit ? (end_bit - begin_bit + kRadixBits - 1) / kRadixBits : 0;
    const int tile_keys = sort_tile_keys(n);
    L.tiles = (n + tile_keys - 1) / tile_keys;
    size_t off = 0;
    L.status_off = off;
    off += 256;
    L.hist_off = off;
    off += (size_t)std::max(L.passes, 1) * kRadix * sizeof(uint32_t);
    L.counter_off = off;
    off += 256;
    L.lookback_off = off;
    off += (size_t)std::max(L.passes, 1) * (size_t)std::max<int64_t>(L.tiles, 1) * kRadix * sizeof(uint32_t);
    L.total = (off + 255) & ~(size_t)255;
    return L;
}

__device__ __forceinline__ uint32_t digit_of(int64_t key, int shift, uint32_t mask) {
    return (uint32_t)(((uint64_t)key >> shift) & mask);
}

// Digit counts of every pass in one read of the keys.  A digit shared by the
// whole wave (the constant exponent byte of the depths, a primitive's tile
// bits) is added once by one lane instead of by 64 serialised LDS atomics.
// n_dev (may be null): the key count in device memory, at most n_cap (the
// sizes the grid and the workspace were laid out for); null: n_cap keys.
__device__ __forceinline__ int64_t sort_count(int64_t n_cap, const int64_t *n_dev) {
    return n_dev ? min(max(*n_dev, (int64_t)0), n_cap) : n_cap;
}

// Each wave reads kHistUnroll x 64 consecutive keys per round (all loads
// issued before any is counted: one 8-byte load per lane at a time left the
// kernel short of loads in flight to cover HBM latency).
#ifndef UBS_SORT_HIST_UNROLL
#define UBS_SORT_HIST_UNROLL 8
#endif
constexpr int kHistUnroll = UBS_SORT_HIST_UNROLL;
// most workgroups of the histogram kernel (each adds its counts to the
// global ones with one atomic per nonzero digit): 12.3M keys of 2 passes took
// 50.2 us with 2048, 41.3 with 1024, 48.0 with 512, 72.4 with 256
#ifndef UBS_SORT_HIST_BLOCKS
#define UBS_SORT_HIST_BLOCKS 1024
#endif
#ifndef UBS_SORT_HIST_KEYS_PER_BLOCK
#define UBS_SORT_HIST_KEYS_PER_BLOCK 4096
#endif

__global__ void sort_hist_kernel(int64_t n_cap, const int64_t *__restrict__ n_dev, const int64_t *__restrict__ keys,
                                 int begin_bit, int end_bit, int passes, int balanced, uint32_t *__restrict__ hist) {
    const int64_t n = sort_count(n_cap, n_dev);
    __shared__ uint32_t h[8][kRadix];
    for (int k = threadIdx.x; k < passes * kRadix; k += blockDim.x) h[k / kRadix][k % kRadix] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    constexpr int64_t kChunk = 64 * kHistUnroll;
    const int64_t waves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;  // wave-uniform
    for (int64_t c0 = wave * kChunk; c0 < n; c0 += waves * kChunk) {
        int64_t key[kHistUnroll];
#pragma unroll
        for (int u = 0; u < kHistUnroll; ++u) {
            const int64_t i = c0 + u * 64 + lane;
            key[u] = i < n ? keys[i] : 0;
        }
#pragma unroll
        for (int u = 0; u < kHistUnroll; ++u) {
            const bool ok = c0 + u * 64 + lane < n;
            const uint64_t active = __ballot(ok);
            if (active == 0ull) break;  // (wave-uniform) past the last key
            for (int p = 0; p < passes; ++p) {
                int shift, bits;
                pass_digit(p, begin_bit, end_bit, balanced != 0, shift, bits);
                const uint32_t d = digit_of(key[u], shift, (1u << bits) - 1u);
                const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);  // lane 0 is active when any lane is
                if (__ballot(ok && d == d0) == active) {
                    if (lane == 0) atomicAdd(&h[p][d0], (uint32_t)__popcll(active));
                } else if (ok) {
                    atomicAdd(&h[p][d], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < passes * kRadix; k += blockDim.x) {
        const uint32_t v = h[k / kRadix][k % kRadix];
        if (v) atomicAdd(&hist[k], v);
    }
}

template <int ITEMS>
__global__ __launch_bounds__(kSortThreads) void sort_pass_kernel(
    int64_t n_cap, const int64_t *__restrict__ n_dev, const int64_t *__restrict__ keys_in, int64_t *__restrict__ keys_out,
    const int32_t *__restrict__ vals_in, int32_t *__restrict__ vals_out, int shift, int bits,
    const uint32_t *__restrict__ digit_count, uint32_t *__restrict__ tile_counter, uint32_t *lookback,
    uint32_t *status, uint32_t *flag) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave_tot[kSortWaves];
    __shared__ uint32_t s_wave_cnt[kSortWaves][kRadix];
    __shared__ uint32_t s_digit_start[kRadix];
    __shared__ uint32_t s_out_base[kRadix];
    constexpr int TILE = kSortThreads * ITEMS;
    __shared__ int64_t s_keys[TILE];
    __shared__ int32_t s_vals[TILE];

    const int tid = threadIdx.x, wave = tid >> 6, lane = (int)lane_id();
    const uint32_t mask = (1u << bits) - 1u;
    const int64_t n = sort_count(n_cap, n_dev);
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (int k = tid; k < kSortWaves * kRadix; k += kSortThreads) (&s_wave_cnt[0][0])[k] = 0;
    // the pass's exclusive digit offsets from its digit counts, scanned by
    // every tile (256 counts; no launch of its own): thread d takes digit d
    // (kSortThreads == kRadix)
    const uint32_t my_cnt = digit_count[tid];
    uint32_t incl = my_cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)incl, o, 64);
        if (lane >= o) incl += u;
    }
    if (lane == 63) s_wave_tot[wave] = incl;
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t base = tile * TILE;
    // with the count in device memory the grid covers n_cap keys: the tiles
    // past the last key leave (tile ids are handed out in order, so no tile
    // that a real one waits on is among them)
    if (base >= n) return;
    uint32_t my_base = incl - my_cnt;
    for (int w = 0; w < wave; ++w) my_base += s_wave_tot[w];

    // a pass whose digit every key shares (a constant exponent byte of the
    // depths) is the identity of a stable sort: the tile is copied through
    // (one digit counts all n keys)
    {
        if (__syncthreads_or(my_cnt == (uint32_t)n)) {
            const int64_t end = min(base + (int64_t)TILE, n);
            for (int64_t j = base + tid; j < end; j += kSortThreads) {
                keys_out[j] = keys_in[j];
                vals_out[j] = vals_in[j];
            }
            return;
        }
    }

    // ---- load (striped inside the wave: item i of lane l is wave_base + i*64 + l)
    int64_t key[ITEMS];
    int32_t val[ITEMS];
    uint32_t dig[ITEMS], rank[ITEMS];
    const int64_t wave_base = base + (int64_t)wave * 64 * ITEMS;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int64_t idx = wave_base + i * 64 + lane;
        const bool ok = idx < n;
        key[i] = ok ? keys_in[idx] : 0;
        val[i] = ok ? vals_in[idx] : 0;
        dig[i] = digit_of(key[i], shift, mask);
    }

    // ---- stable rank inside the wave via ballots on the digit bits
    volatile uint32_t *cnt = s_wave_cnt[wave];
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int64_t idx = wave_base + i * 64 + lane;
        const bool ok = idx < n;
        uint64_t peers = __ballot(ok);
        for (int b = 0; b < bits; ++b) {
            const bool bit = (dig[i] >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t r = 0;
        if (ok) {
            const uint32_t before = (uint32_t)__popcll(peers & lt_mask);
            const uint32_t c0 = cnt[dig[i]];
            r = c0 + before;
            if (before == 0) cnt[dig[i]] = c0 + (uint32_t)__popcll(peers);
        }
        rank[i] = r;
    }
    __syncthreads();

    // ---- tile digit totals, per-wave exclusive offsets; publish aggregate
    uint32_t tot = 0;
    {
        const int d = tid;  // kSortThreads == kRadix
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            const uint32_t c = s_wave_cnt[w][d];
            s_wave_cnt[w][d] = tot;
            tot += c;
        }
        uint32_t *slot = lookback + (size_t)tile * kRadix + d;
        __hip_atomic_store(slot, (tile == 0 ? kFlagPrefix : kFlagAgg) | tot, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        s_digit_start[d] = tot;
    }
    __syncthreads();
    // exclusive scan over digits (256 entries, one per thread)
    {
        const int d = tid;
        uint32_t v = s_digit_start[d];
        uint32_t inc = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t u = __shfl_up(inc, o, 64);
            if (lane >= o) inc += u;
        }
        __syncthreads();
        if (lane == 63) s_out_base[wave] = inc;  // wave totals (scratch)
        __syncthreads();
        uint32_t wbase = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w)
            if (w < wave) wbase += s_out_base[w];
        __syncthreads();
        s_digit_start[d] = wbase + inc - v;
    }
    __syncthreads();

    // ---- stage the tile in LDS in digit order
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int64_t idx = wave_base + i * 64 + lane;
        if (idx < n) {
            const uint32_t pos = s_digit_start[dig[i]] + s_wave_cnt[wave][dig[i]] + rank[i];
            s_keys[pos] = key[i];
            s_vals[pos] = val[i];
        }
    }

    // ---- decoupled look-back for this tile's global prefix of digit d
    {
        const int d = tid;
        uint32_t excl = 0;
        if (tile > 0) {
            // look back kLookback tiles per round trip: their slots are read
            // together, then taken in order until an inclusive prefix (done)
            // or a tile that has not published yet (read again from there)
            int64_t j = tile - 1;
            uint32_t spins = 0;
            while (j >= 0) {
                uint32_t v[kLookback];
#pragma unroll
                for (int u = 0; u < kLookback; ++u)
                    v[u] = j - u >= 0 ? __hip_atomic_load(lookback + (size_t)(j - u) * kRadix + d, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT)
                                      : kFlagPrefix;  // (tile 0 publishes a prefix: never reached)
                bool done = false, wait = false;
                int taken = 0;
#pragma unroll
                for (int u = 0; u < kLookback; ++u) {
                    if (done || wait) continue;
                    const uint32_t flag = v[u] & ~kValueMask;
                    if (flag == 0u) {
                        wait = true;
                        continue;
                    }
                    excl += v[u] & kValueMask;
                    ++taken;
                    if (flag == kFlagPrefix) done = true;
                }
                if (done) break;
                j -= taken;
                if (wait) {
                    if (++spins > kSpinLimit) {
                        __hip_atomic_store(status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (flag) __hip_atomic_store(flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __hip_atomic_store(lookback + (size_t)tile * kRadix + d, kFlagPrefix | ((excl + tot) & kValueMask),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_out_base[d] = my_base + excl - s_digit_start[d];
    }
    __syncthreads();

    // ---- write out: consecutive LDS slots of one digit go to consecutive
    // global slots
    const int64_t count = min((int64_t)TILE, n - base);
    for (int j = tid; j < count; j += kSortThreads) {
        const int64_t k = s_keys[j];
        const uint32_t pos = s_out_base[digit_of(k, shift, mask)] + (uint32_t)j;
        if ((int64_t)pos < n) {  // always true unless a look-back wait gave up
            keys_out[pos] = k;
            vals_out[pos] = s_vals[j];
        }
    }
}

// Scratch of ubs_radix_sort_pairs_noalloc: the status word, digit counts,
// tile counters and look-back slots of one sort, in a file-scope device
// array instead of a caller workspace (24 MiB: 16M keys at all 8 passes,
// more keys when fewer passes run).  Each device gets its own copy when the
// library's code object loads there and keeps it until the process ends.
// Host code must not pass the array's name to a kernel (it is a host-side
// stand-in): it uses the device address that hipGetSymbolAddress returns
// once per device, kept in g_scratch_ptr.
constexpr size_t kNoAllocScratchBytes = size_t(24) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// Device address of the current device's scratch (looked up on its first use).
static int scratch_for_current_device(void **out) {
    int dev = 0;
    UBS_HIP_CHECK(hipGetDevice(&dev));
    UBS_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d the scratch table holds", dev,
                kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        UBS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        UBS_REQUIRE(p != nullptr, "hipGetSymbolAddress returned a null scratch address on device %d", dev);
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return UBS_OK;
}

__global__ void copy_status_kernel(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst) {
    if (threadIdx.x == 0) dst[0] = src[0];
}

}  // namespace ubs

using namespace ubs;

extern "C" {

size_t ubs_radix_sort_workspace_bytes(int64_t n, int begin_bit, int end_bit) {
    return sort_layout(n, begin_bit, end_bit).total;
}

int ubs_radix_sort_pairs(int64_t n, int64_t *keys, int64_t *keys_alt, int32_t *vals, int32_t *vals_alt,
                         int begin_bit, int end_bit, void *workspace, size_t ws_bytes, int *result_in_alt,
                         void *stream) {
    return ubs_radix_sort_pairs_skip(n, keys, keys_alt, vals, vals_alt, begin_bit, end_bit, 0u, workspace,
                                     ws_bytes, result_in_alt, stream);
}

static int sort_impl(int64_t n, const int64_t *n_dev, int64_t *keys, int64_t *keys_alt, int32_t *vals,
                     int32_t *vals_alt, int begin_bit, int end_bit, uint32_t skip_passes, void *workspace,
                     size_t ws_bytes, int *result_in_alt, void *stream, uint32_t *flag = nullptr);

int ubs_radix_sort_pairs_skip(int64_t n, int64_t *keys, int64_t *keys_alt, int32_t *vals, int32_t *vals_alt,
                              int begin_bit, int end_bit, uint32_t skip_passes, void *workspace, size_t ws_bytes,
                              int *result_in_alt, void *stream) {
    return sort_impl(n, nullptr, keys, keys_alt, vals, vals_alt, begin_bit, end_bit, skip_passes, workspace,
                     ws_bytes, result_in_alt, stream);
}

int ubs_radix_sort_pairs_flag(int64_t n, int64_t *keys, int64_t *keys_alt, int32_t *vals, int32_t *vals_alt,
                              int begin_bit, int end_bit, uint32_t skip_passes, void *workspace, size_t ws_bytes,
                              int *result_in_alt, uint32_t *flag_dev, void *stream) {
    UBS_REQUIRE(flag_dev, "flag_dev is null");
    return sort_impl(n, nullptr, keys, keys_alt, vals, vals_alt, begin_bit, end_bit, skip_passes, workspace,
                     ws_bytes, result_in_alt, stream, flag_dev);
}

int ubs_radix_sort_pairs_dev(int64_t n_cap, const int64_t *n_dev, int64_t *keys, int64_t *keys_alt, int32_t *vals,
                             int32_t *vals_alt, int begin_bit, int end_bit, void *workspace, size_t ws_bytes,
                             int *result_in_alt, void *stream) {
    UBS_REQUIRE(n_dev, "n_dev is null");
    return sort_impl(n_cap, n_dev, keys, keys_alt, vals, vals_alt, begin_bit, end_bit, 0u, workspace, ws_bytes,
                     result_in_alt, stream);
}

static int sort_impl(int64_t n, const int64_t *n_dev, int64_t *keys, int64_t *keys_alt, int32_t *vals,
                     int32_t *vals_alt, int begin_bit, int end_bit, uint32_t skip_passes, void *workspace,
                     size_t ws_bytes, int *result_in_alt, void *stream, uint32_t *flag) {
    hipStream_t st = as_stream(stream);
    UBS_REQUIRE(result_in_alt, "result_in_alt is null");
    UBS_REQUIRE(begin_bit >= 0 && end_bit <= 64 && begin_bit <= end_bit, "bad bit range [%d, %d)", begin_bit,
                end_bit);
    UBS_REQUIRE(n >= 0 && n <= (int64_t)kValueMask, "n=%lld out of range (max 2^30-1)", (long long)n);
    const SortLayout L = sort_layout(n, begin_bit, end_bit);
    UBS_REQUIRE(L.passes <= 8, "too many passes");
    *result_in_alt = 0;
    if (!workspace || ws_bytes < L.total) {
        set_error("sort workspace too small (%zu < %zu)", ws_bytes, L.total);
        return UBS_ERR_WORKSPACE;
    }
    char *ws = reinterpret_cast<char *>(workspace);
    uint32_t *status = reinterpret_cast<uint32_t *>(ws + L.status_off);
    int e = zero_async(ws, L.total, st);
    if (e) return e;
    if (n == 0 || L.passes == 0) return UBS_OK;
    UBS_REQUIRE(keys && keys_alt && vals && vals_alt, "null pointer");
    uint32_t *hist = reinterpret_cast<uint32_t *>(ws + L.hist_off);
    uint32_t *counters = reinterpret_cast<uint32_t *>(ws + L.counter_off);
    uint32_t *lookback = reinterpret_cast<uint32_t *>(ws + L.lookback_off);
    const unsigned hist_blocks = (unsigned)std::min<int64_t>(
        (n + UBS_SORT_HIST_KEYS_PER_BLOCK - 1) / UBS_SORT_HIST_KEYS_PER_BLOCK, UBS_SORT_HIST_BLOCKS);
    sort_hist_kernel<<<hist_blocks, 256, 0, st>>>(n, n_dev, keys, begin_bit, end_bit, L.passes,
                                                  skip_passes == 0 ? 1 : 0, hist);
    UBS_LAUNCH_CHECK("sort_hist_kernel");
    int64_t *kb[2] = {keys, keys_alt};
    int32_t *vb[2] = {vals, vals_alt};
    int run = 0;  // passes launched so far: the data sits in buffer run & 1
    for (int p = 0; p < L.passes; ++p) {
        // a pass whose digit every key shares is the identity of a stable
        // sort: leaving it out changes nothing
        if ((skip_passes >> p) & 1u) continue;
        int shift, bits;
        pass_digit(p, begin_bit, end_bit, skip_passes == 0, shift, bits);
        auto *pass = n < kSmallSortN ? sort_pass_kernel<kItemsSmall> : sort_pass_kernel<kItems>;
        pass<<<(unsigned)L.tiles, kSortThreads, 0, st>>>(
            n, n_dev, kb[run & 1], kb[(run + 1) & 1], vb[run & 1], vb[(run + 1) & 1], shift, bits,
            hist + (size_t)p * kRadix, counters + p, lookback + (size_t)p * L.tiles * kRadix, status, flag);
        UBS_LAUNCH_CHECK("sort_pass_kernel");
        ++run;
    }
    *result_in_alt = run & 1;
    return UBS_OK;
}

int ubs_radix_sort_pairs_noalloc(int64_t n, int64_t *keys, int64_t *keys_alt, int32_t *vals, int32_t *vals_alt,
                                 int begin_bit, int end_bit, int *result_in_alt, uint32_t *status_dev,
                                 void *stream) {
    UBS_REQUIRE(result_in_alt, "result_in_alt is null");
    UBS_REQUIRE(begin_bit >= 0 && end_bit <= 64 && begin_bit <= end_bit, "bad bit range [%d, %d)", begin_bit,
                end_bit);
    UBS_REQUIRE(n >= 0 && n <= (int64_t)kValueMask, "n=%lld out of range (max 2^30-1)", (long long)n);
    const SortLayout L = sort_layout(n, begin_bit, end_bit);
    UBS_REQUIRE(L.total <= kNoAllocScratchBytes,
                "n=%lld keys over bits [%d, %d) need %zu bytes of sort scratch; the allocation-free sort holds "
                "%zu (at most 16777216 keys at 8 passes); use ubs_radix_sort_pairs with a workspace",
                (long long)n, begin_bit, end_bit, L.total, kNoAllocScratchBytes);
    void *ws = nullptr;
    int e = scratch_for_current_device(&ws);
    if (e) return e;
    e = ubs_radix_sort_pairs_skip(n, keys, keys_alt, vals, vals_alt, begin_bit, end_bit, 0u, ws,
                                  kNoAllocScratchBytes, result_in_alt, stream);
    if (e || !status_dev) return e;
    copy_status_kernel<<<1, 64, 0, as_stream(stream)>>>(
        reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(ws) + L.status_off), status_dev);
    UBS_LAUNCH_CHECK("copy_status_kernel");
    return UBS_OK;
}

int ubs_radix_sort_status(const void *workspace, int *status_out, void *stream) {
    UBS_REQUIRE(workspace && status_out, "null pointer");
    uint32_t v = 0;
    UBS_HIP_CHECK(hipMemcpyAsync(&v, workspace, sizeof(uint32_t), hipMemcpyDeviceToHost, as_stream(stream)));
    UBS_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
    *status_out = (int)v;
    return UBS_OK;
}

}  // extern "C"

// Shared device helpers, error plumbing and launch conventions of the
// gfx950 Beta-splat rasteriser library (libubs_hip.so).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "ubs_capi.h"

namespace ubs {

// ----------------------------------------------------------------------------
// Errors: every C entry point returns a status and records a message that
// ubs_last_error() hands back (thread-local).
// ----------------------------------------------------------------------------
void set_error(const char *fmt, ...);

#define UBS_REQUIRE(cond, ...)                                                 \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ::ubs::set_error(__VA_ARGS__);                                     \
            return UBS_ERR_INVALID;                                            \
        }                                                                      \
    } while (0)

#define UBS_HIP_CHECK(expr)                                                    \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) {                                                \
            ::ubs::set_error("%s failed: %s", #expr, hipGetErrorString(_e));   \
            return UBS_ERR_HIP;                                                \
        }                                                                      \
    } while (0)

// Check for a launch error right after a <<<>>> launch.
#define UBS_LAUNCH_CHECK(name)                                                 \
    do {                                                                       \
        hipError_t _e = hipGetLastError();                                     \
        if (_e != hipSuccess) {                                                \
            ::ubs::set_error("launch of %s failed: %s", name,                  \
                             hipGetErrorString(_e));                           \
            return UBS_ERR_HIP;                                                \
        }                                                                      \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kThreads = 256;  // block size of the per-primitive kernels

inline unsigned grid_for(int64_t n, int threads = kThreads) {
    return (unsigned)((n + threads - 1) / threads);
}

// Zero a float/int buffer with a kernel (no hipMemsetAsync: memset nodes
// have misbehaved inside captured graphs).
int zero_async(void *ptr, size_t bytes, hipStream_t stream);

// ----------------------------------------------------------------------------
// Device helpers
// ----------------------------------------------------------------------------

// float -> uint32 as CUDA converts (saturating, NaN -> 0); the tile-box math
// of isect_tiles.cu:62-66 relies on negative coordinates clamping to 0.
__device__ __forceinline__ uint32_t sat_u32(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)v;
}

// float -> int32, truncating, saturating, NaN -> 0.
__device__ __forceinline__ int32_t f2i_sat(float v) {
    if (!(v == v)) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)v;
}

inline uint32_t bit_width_u32(uint32_t v) {
    uint32_t b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}

// Hardware float atomic add into global memory (global_atomic_add_f32).
__device__ __forceinline__ void atomic_add_f32(float *p, float v) {
    unsafeAtomicAdd(p, v);
}

// Sum over the 64 lanes of a wave; every lane receives the total.
// Quad butterflies and row rotations use DPP (full-rate VALU modifiers);
// the four 16-lane rows are combined through readlane.
__device__ __forceinline__ float wave_sum(float v) {
    v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true));  // quad_perm [1,0,3,2]
    v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, true));  // quad_perm [2,3,0,1]
    v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x124, 0xF, 0xF, true)); // row_ror:4
    v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x128, 0xF, 0xF, true)); // row_ror:8
    const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0));
    const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16));
    const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
    const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 48));
    return (r0 + r1) + (r2 + r3);
}

// Pairwise reduce-scatter of two per-lane values across the wave halves
// (v_permlane32_swap): lanes 0-31 get a[l] + a[l+32], lanes 32-63 get
// b[l-32] + b[l].
__device__ __forceinline__ float swap32_sum(float a, float b) {
    const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    // copy the two results out of the vector before reinterpreting them:
    // hipcc lowers __builtin_bit_cast of a vector element s[1] as a load of
    // s[0], so the bit cast must never see the subscript.
    const unsigned lo = s[0], hi = s[1];
    return __uint_as_float(lo) + __uint_as_float(hi);
}

// Same inside each 32-lane half (v_permlane16_swap): quarter 0 gets
// a[l] + a[l+16], quarter 1 b[l-16] + b[l], quarters 2/3 likewise.
__device__ __forceinline__ float swap16_sum(float a, float b) {
    const auto s = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    const unsigned lo = s[0], hi = s[1];
    return __uint_as_float(lo) + __uint_as_float(hi);
}

// Wave totals of NV per-lane values (wave_totals_rows below) at ~2.4 VALU
// ops per value instead of ~11 for a butterfly per value: two reduce-scatter
// levels (permlane swaps) leave each 16-lane quarter Q with partial sums of
// value wave_total_index<NV>(r, Q) in group register r (or of nothing when
// that returns -1); values go in groups of four, and a trailing group of one
// or two values costs one group-level instead of a padded four.  Must be
// called with all 64 lanes active.
template <int NV>
struct WaveTotals {
    static constexpr int kFull = NV / 4;
    static constexpr int kTail = NV % 4;
    static constexpr int kGroups = kFull + (kTail ? 1 : 0);
};

template <int NV>
__host__ __device__ constexpr int wave_total_index(int r, int q) {
    constexpr int F = WaveTotals<NV>::kFull, T = WaveTotals<NV>::kTail;
    if (r < F || T == 3) {
        const int idx = 4 * r + (((q & 1) << 1) | (q >> 1));
        return idx < NV ? idx : -1;
    }
    if (T == 1) return q == 0 ? 4 * r : -1;
    return q == 0 ? 4 * r : (q == 2 ? 4 * r + 1 : -1);  // T == 2
}

// ---- row stage by reduce-scatter -------------------------------------------
// After the two permlane levels every 16-lane row (quarter) holds, per
// group register, partial sums of one value that still need a 16-lane sum.
// Instead of four DPP adds per register, the rows are reduce-scattered too:
// an all-lane DPP add of one register, then a DPP add of the other with a
// bank mask (banks = 4-lane groups of a row) that rewrites only its half or
// bank pair while the other lanes keep the first result (the tied old
// operand), so two instructions combine two registers into one at each of
// the 8- and 4-lane levels, and two quad steps finish every bank.  After
// row_totals, bank b of register k holds the total of group
// row_group<G>(k, b) (-1: a copy of another bank's total).
// Each DPP instruction is preceded by s_nop 1: a DPP read of a VGPR needs
// two wait states after the VALU write, which the compiler does not insert
// for inline assembly.
// row_ror:N makes lane l of a row read lane (l - N) mod 16.
#define UBS_ADD_DPP(R, OLD, SRC, OTHER, CTRL, BANKS)                                                          \
    asm("s_nop 1\n\tv_add_f32_dpp %0, %1, %2 " CTRL " row_mask:0xf bank_mask:" BANKS                          \
        : "=v"(R) : "v"(SRC), "v"(OTHER), "0"(OLD))
// lanes 8-15: src[l - 8] + other; lanes 0-7: old
__device__ __forceinline__ float add_dpp_hi8(float old, float src, float other) {
    float r;
    UBS_ADD_DPP(r, old, src, other, "row_ror:8", "0xc");
    return r;
}
// banks 1 and 3 (lanes 4-7, 12-15): src[l - 4] + other; banks 0 and 2: old
__device__ __forceinline__ float add_dpp_odd4(float old, float src, float other) {
    float r;
    UBS_ADD_DPP(r, old, src, other, "row_ror:4", "0xa");
    return r;
}
// every lane: x[l + 8 mod 16] + x[l]
__device__ __forceinline__ float add_dpp_all8(float x) {
    float r;
    UBS_ADD_DPP(r, x, x, x, "row_ror:8", "0xf");
    return r;
}
// every lane: x[l + 4 mod 16] + x[l] (lanes 0-3 and 8-11 stay in their half)
__device__ __forceinline__ float add_dpp_all12(float x) {
    float r;
    UBS_ADD_DPP(r, x, x, x, "row_ror:12", "0xf");
    return r;
}
// every lane: x[7 - l] + x[l] inside each 8-lane half of a row (a quad sum
// after it gives the half's total; a row rotate would mix the halves)
__device__ __forceinline__ float add_dpp_half_mirror(float x) {
    float r;
    UBS_ADD_DPP(r, x, x, x, "row_half_mirror", "0xf");
    return r;
}
// every lane: the sum over its quad
__device__ __forceinline__ float add_dpp_quad(float x) {
    float a, b;
    UBS_ADD_DPP(a, x, x, x, "quad_perm:[1,0,3,2]", "0xf");
    UBS_ADD_DPP(b, a, a, a, "quad_perm:[2,3,0,1]", "0xf");
    return b;
}
#undef UBS_ADD_DPP

template <int G>
struct RowTotals {
    static constexpr int kH = (G + 1) / 2;   // registers after the 8-lane level
    static constexpr int kRegs = (kH + 1) / 2;  // after the 4-lane level
};

// group whose total bank b (lanes 4b..4b+3 of a row) of register k holds
// after row_totals: banks 0/2 come from the even 8-lane register, 1/3 from
// the odd one; banks 0/1 from its lower half, 2/3 from its upper half
template <int G>
__host__ __device__ constexpr int row_group(int k, int b) {
    constexpr int H = RowTotals<G>::kH;
    const int h = (b & 1) ? 2 * k + 1 : 2 * k;
    if (h >= H) return -1;
    const int g = (b & 2) ? 2 * h + 1 : 2 * h;
    return g < G ? g : -1;
}

template <int NV>
__device__ __forceinline__ void wave_totals_rows(const float (&v)[NV],
                                                 float (&out)[RowTotals<WaveTotals<NV>::kGroups>::kRegs]) {
    constexpr int F = WaveTotals<NV>::kFull, T = WaveTotals<NV>::kTail;
    constexpr int G = WaveTotals<NV>::kGroups;
    constexpr int H = RowTotals<G>::kH, K = RowTotals<G>::kRegs;
    float g[G];
#pragma unroll
    for (int r = 0; r < F; ++r)
        g[r] = swap16_sum(swap32_sum(v[4 * r], v[4 * r + 1]), swap32_sum(v[4 * r + 2], v[4 * r + 3]));
    if constexpr (T == 3) {
        g[F] = swap16_sum(swap32_sum(v[4 * F], v[4 * F + 1]), swap32_sum(v[4 * F + 2], 0.f));
    } else if constexpr (T == 2) {
        const float u = swap32_sum(v[4 * F], v[4 * F + 1]);
        g[F] = swap16_sum(u, u);
    } else if constexpr (T == 1) {
        const float u = swap32_sum(v[4 * F], v[4 * F]);
        g[F] = swap16_sum(u, u);
    }
    float h[H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
        if (2 * i + 1 < G) {
            // lanes 0-7 from an all-lane add of the even register, then 8-15
            // overwritten from the odd one (the first result is the tied
            // old operand, so no copy is needed)
            h[i] = add_dpp_hi8(add_dpp_all8(g[2 * i]), g[2 * i + 1], g[2 * i + 1]);
        } else {
            h[i] = add_dpp_all8(g[2 * i]);
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        float x;
        if (2 * k + 1 < H) {
            x = add_dpp_odd4(add_dpp_all12(h[2 * k]), h[2 * k + 1], h[2 * k + 1]);
        } else {
            x = add_dpp_half_mirror(h[2 * k]);
        }
        out[k] = add_dpp_quad(x);
    }
}

__device__ __forceinline__ int wave_max_i32(int v) {
    // butterfly through ds_swizzle-free shuffles (used once per tile)
    for (int o = 32; o > 0; o >>= 1) {
        const int w = __shfl_xor(v, o, 64);
        v = v > w ? v : w;
    }
    return v;
}

// Sum of a per-lane count over the wave (measurement counters; not hot).
__device__ __forceinline__ unsigned long long wave_sum_u64(uint32_t v) {
    unsigned long long s = v;
    for (int o = 32; o > 0; o >>= 1) s += (unsigned long long)__shfl_xor((long long)s, o, 64);
    return s;
}

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

}  // namespace ubs

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

__device__ __forceinline__ int wave_max_i32(int v) {
    // butterfly through ds_swizzle-free shuffles (used once per tile)
    for (int o = 32; o > 0; o >>= 1) {
        const int w = __shfl_xor(v, o, 64);
        v = v > w ? v : w;
    }
    return v;
}

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

}  // namespace ubs

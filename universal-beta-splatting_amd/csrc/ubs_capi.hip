// Process-wide plumbing of libubs_hip.so: error strings, version, zero fill.
#include <cstdarg>

#include "ubs_common.h"

namespace ubs {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

__global__ void zero_u32_kernel(uint32_t *__restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = i; k < n; k += stride) p[k] = 0u;
}

__global__ void zero_u8_kernel(uint8_t *__restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

int zero_async(void *ptr, size_t bytes, hipStream_t stream) {
    if (bytes == 0 || ptr == nullptr) return UBS_OK;
    if ((bytes % 4) == 0 && (reinterpret_cast<uintptr_t>(ptr) % 4) == 0) {
        const size_t n = bytes / 4;
        const unsigned blocks = (unsigned)std::min<size_t>((n + kThreads - 1) / kThreads, 8192);
        zero_u32_kernel<<<blocks, kThreads, 0, stream>>>(reinterpret_cast<uint32_t *>(ptr), n);
    } else {
        zero_u8_kernel<<<grid_for((int64_t)bytes), kThreads, 0, stream>>>(reinterpret_cast<uint8_t *>(ptr), bytes);
    }
    UBS_LAUNCH_CHECK("zero_fill");
    return UBS_OK;
}

}  // namespace ubs

extern "C" {

int ubs_version(void) { return 100; }

const char *ubs_last_error(void) { return ubs::g_last_error.c_str(); }

}  // extern "C"

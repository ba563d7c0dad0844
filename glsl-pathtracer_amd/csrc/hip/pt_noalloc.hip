// Environment-map CDF for callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc): pt_env_cdf_noalloc in include/glslpt.h. Same kernels as
// pt_set_env_map, bit-identical results; their scratch (the chunk sums,
// grids, functions and carried states of the parallel scan, and its one-word
// flag) lives in a fixed device array of this file instead of a buffer the
// renderer allocates.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../../include/glslpt.h"
#include "pt_kernels.h"

// pt_renderer.cpp: the message pt_last_error(NULL) reports.
extern "C" __attribute__((visibility("hidden"))) void pt_set_handleless_error(const char* msg);

namespace {

// Layout: the flag word at 0, the chunk scratch from 256 (8-byte aligned).
constexpr size_t FLAG_BYTES = 256;
constexpr size_t SCRATCH_BYTES = FLAG_BYTES + ((ptk::env_cdf_scratch_size(PT_ENV_NOALLOC_MAX_TEXELS) + 255) / 256) * 256;

// Every GPU the library runs on holds its own copy until the process ends.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
void* g_scratch[MAX_DEVICES] = {};  // device address of `scratch` on each device

// Device address of this device's copy (the host-side name of the array is not
// one). Looked up once per device, with that device current.
void* device_scratch(int dev) {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

}  // namespace

extern "C" int pt_env_cdf_noalloc(const float* rgb, int32_t width, int32_t height, float* cdf, float* total,
                                  void* stream) {
    if (!rgb || !cdf || !total || width <= 0 || height <= 0) {
        pt_set_handleless_error("pt_env_cdf_noalloc: null buffer or non-positive size");
        return PT_ERR_INVALID_ARG;
    }
    const long long n = static_cast<long long>(width) * height;
    if (n > PT_ENV_NOALLOC_MAX_TEXELS) {
        pt_set_handleless_error("pt_env_cdf_noalloc: width*height exceeds PT_ENV_NOALLOC_MAX_TEXELS (33554432 texels)");
        return PT_ERR_OVERFLOW;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        pt_set_handleless_error("pt_env_cdf_noalloc: no current HIP device");
        return PT_ERR_NO_DEVICE;
    }
    unsigned char* base = static_cast<unsigned char*>(device_scratch(dev));
    if (!base) {
        pt_set_handleless_error("pt_env_cdf_noalloc: hipGetSymbolAddress of the scratch array failed");
        return PT_ERR_HIP;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    ptk::launch_env_cdf(s, rgb, static_cast<int>(n), cdf, total, reinterpret_cast<int*>(base), base + FLAG_BYTES);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        pt_set_handleless_error(hipGetErrorString(e));
        return PT_ERR_HIP;
    }
    return PT_OK;
}

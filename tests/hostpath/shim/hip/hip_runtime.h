// Host stand-in for <hip/hip_runtime.h>, for the CPU build of the renderer's
// kernel-independent device functions (pt_device.h, pt_path.h, pt_replay.h,
// pt_layout.h) that the host-path tests run against the oracle. Vector types
// and function qualifiers only: no kernels, streams or device memory.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#define __device__
#define __host__
#define __forceinline__ inline
#define __restrict__
struct alignas(16) float4 { float x, y, z, w; };
struct alignas(16) int4 { int x, y, z, w; };
struct alignas(16) uint4 { unsigned x, y, z, w; };
struct int2 { int x, y; };
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline int4 make_int4(int x, int y, int z, int w) { return int4{x, y, z, w}; }
inline uint4 make_uint4(unsigned x, unsigned y, unsigned z, unsigned w) { return uint4{x, y, z, w}; }
inline int2 make_int2(int x, int y) { return int2{x, y}; }
inline int __float_as_int(float f) { int i; std::memcpy(&i, &f, 4); return i; }
inline float __int_as_float(int i) { float f; std::memcpy(&f, &i, 4); return f; }
inline unsigned __float_as_uint(float f) { unsigned i; std::memcpy(&i, &f, 4); return i; }
inline float __uint_as_float(unsigned i) { float f; std::memcpy(&f, &i, 4); return f; }
// The traversal stack's backing-store read is a relaxed atomic load on the device.
#define __HIP_MEMORY_SCOPE_WAVEFRONT 0
#define __hip_atomic_load(p, order, scope) (*(p))
// Launcher declarations in the same headers name a stream type.
typedef struct ihipStream_t* hipStream_t;

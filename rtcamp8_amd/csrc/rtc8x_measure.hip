// rtc8x_measure.hip -- measurement kernels with no counterpart in the renderer: a plain copy (the
// calibration of the HBM byte counters) and a dependent 8-tap gather over the grid's leaf values (the
// memory-latency roofline of the density lookups).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rtc8x_common.hpp"

namespace rtc8x {

__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__device__ __forceinline__ uint32_t mix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

// Each lane performs `steps` lookups.  A lookup reads the 8 corners of a 2x2x2 stencil at a random
// position inside a random leaf (leaf voxel offset x<<6 | y<<3 | z, the NanoVDB leaf order), exactly
// the 8 taps of a trilinear density lookup, as four 8-byte loads of z-adjacent pairs (the fewest
// memory instructions that fetch a stencil); the next lookup's position depends on the values read,
// so lookups are serially dependent while the taps of one lookup are issued together.
__global__ void gather_probe_kernel(const float* __restrict__ values, uint32_t num_leaves, uint32_t steps,
                                    uint32_t seed, float* __restrict__ out)
{
    extern __shared__ float occupancy_limiter[];
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t h = mix32(tid * 0x9e3779b9u ^ seed);
    float acc = 0.0f;
    for (uint32_t s = 0; s < steps; ++s) {
        const uint32_t leaf = (uint32_t)(((uint64_t)h * num_leaves) >> 32);
        const uint32_t r = mix32(h + 0x632be5abu);
        const uint32_t x = (r & 0xffu) % 7u, y = ((r >> 8) & 0xffu) % 7u, z = ((r >> 16) & 0xffu) % 7u;
        const float* b = values + (size_t)leaf * 512 + (x << 6 | y << 3 | z);
        float2 q[4];
        __builtin_memcpy(&q[0], b, 8);
        __builtin_memcpy(&q[1], b + 8, 8);
        __builtin_memcpy(&q[2], b + 64, 8);
        __builtin_memcpy(&q[3], b + 72, 8);
        const float v = ((q[0].x + q[0].y) + (q[1].x + q[1].y)) + ((q[2].x + q[2].y) + (q[3].x + q[3].y));
        acc += v;
        h = mix32(h ^ __float_as_uint(v) ^ s);
    }
    if (acc == -1.0f) occupancy_limiter[threadIdx.x] = acc;  // keeps the dynamic LDS allocation alive
    out[tid] = acc;
}

}  // namespace rtc8x

using namespace rtc8x;

extern "C" rtc8x_status rtc8x_copy_async(void* dst, const void* src, size_t bytes, void* stream)
{
    if (bytes == 0) return RTC8X_OK;
    RTC8X_REQUIRE(dst && src, "rtc8x_copy_async: NULL pointer");
    RTC8X_REQUIRE(((uintptr_t)dst | (uintptr_t)src | bytes) % 16 == 0, "rtc8x_copy_async: needs 16-byte alignment");
    const size_t n = bytes / 16;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 256 * 32);
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, (const uint4*)src,
                       (uint4*)dst, n);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}

extern "C" rtc8x_status rtc8x_gather_probe(void* stream, const float* leaf_values, uint32_t num_leaves,
                                           uint32_t threads, uint32_t steps, uint32_t waves_per_simd, uint32_t seed,
                                           float* out)
{
    RTC8X_REQUIRE(leaf_values && out, "rtc8x_gather_probe: NULL pointer");
    RTC8X_REQUIRE(num_leaves > 0 && threads > 0 && threads % 256 == 0 && steps > 0,
                  "rtc8x_gather_probe: num_leaves and steps must be > 0, threads a multiple of 256");
    RTC8X_REQUIRE(waves_per_simd >= 3 && waves_per_simd <= 8, "rtc8x_gather_probe: waves_per_simd must be in [3, 8]");
    int dev = 0, lds_cu = 0, lds_block = 0;
    RTC8X_HIP(hipGetDevice(&dev));
    RTC8X_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    RTC8X_HIP(hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    // a 256-lane workgroup puts one wave on each of the CU's 4 SIMDs, so W workgroups per CU give W
    // waves per SIMD; the LDS reservation admits W of them and not W + 1
    size_t lds = (size_t)lds_cu / waves_per_simd;
    lds = lds > 1024 ? lds - 1024 : lds;
    RTC8X_REQUIRE(lds <= (size_t)lds_block && (size_t)lds_cu / lds == waves_per_simd,
                  "rtc8x_gather_probe: %u waves per SIMD cannot be set with %d bytes of LDS per CU", waves_per_simd,
                  lds_cu);
    hipLaunchKernelGGL(gather_probe_kernel, dim3(threads / 256), dim3(256), lds, (hipStream_t)stream, leaf_values,
                       num_leaves, steps, seed, out);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}

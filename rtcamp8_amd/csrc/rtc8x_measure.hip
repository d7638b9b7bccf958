// rtc8x_measure.hip -- measurement kernels with no counterpart in the renderer: a plain copy (the
// calibration of the HBM byte counters) and dependent density lookups at random positions of the grid's
// leaves (the memory-latency roofline of the tracer's lookups).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rtc8x_common.hpp"
#include "rtc8x_device.hpp"

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

// Each lane performs `steps` steps of W density lookups.  A lookup is the tracer's own: the trilinear
// stencil at a random position inside a random leaf, fetched in two rounds (leaf-map entries, then the
// voxels) by trilinear_batch, whose W lookups overlap.  The next step's positions depend on the values
// read, so steps are serially dependent.
template <uint32_t W>
__global__ void gather_probe_kernel(const rtc8x_grid g, uint32_t steps, uint32_t seed, float* __restrict__ out)
{
    extern __shared__ float occupancy_limiter[];
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t h = mix32(tid * 0x9e3779b9u ^ seed);
    float acc = 0.0f;
    for (uint32_t s = 0; s < steps; ++s) {
        dev::V3 idx[W];
#pragma unroll
        for (uint32_t k = 0; k < W; ++k) {
            h = mix32(h + 0x632be5abu * (k + 1));
            const uint32_t leaf = (uint32_t)(((uint64_t)h * g.num_leaves) >> 32);
            const uint32_t r = mix32(h ^ 0x5bd1e995u);
            const int32_t* o = g.leaf_origins + 3 * (size_t)leaf;
            idx[k] = dev::V3{(float)o[0] + (float)(r & 0xffu) * (8.0f / 256.0f),
                        (float)o[1] + (float)((r >> 8) & 0xffu) * (8.0f / 256.0f),
                        (float)o[2] + (float)((r >> 16) & 0xffu) * (8.0f / 256.0f)};
        }
        float v[W];
        dev::trilinear_batch<W>(g, idx, v);
        float sum = 0.0f;
#pragma unroll
        for (uint32_t k = 0; k < W; ++k) sum += v[k];
        acc += sum;
        h = mix32(h ^ __float_as_uint(sum) ^ s);
    }
    if (acc == -1.0f) occupancy_limiter[threadIdx.x] = acc;  // keeps the dynamic LDS allocation alive
    out[tid] = acc;
}

}  // namespace rtc8x

using namespace rtc8x;

extern "C" rtc8x_status rtc8x_copy_async(void* dst, const void* src, size_t bytes, void* stream)
try {
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
RTC8X_CATCH_ALL("rtc8x_copy_async")

extern "C" rtc8x_status rtc8x_gather_probe(void* stream, const rtc8x_grid* grid, uint32_t threads, uint32_t steps,
                                           uint32_t width, uint32_t waves_per_simd, uint32_t seed, float* out)
try {
    RTC8X_REQUIRE(grid && out, "rtc8x_gather_probe: NULL pointer");
    const rtc8x_grid& g = *grid;
    RTC8X_REQUIRE(g.num_leaves > 0 && g.leaf_origins && g.leaf_values && g.tile_values,
                  "rtc8x_gather_probe: the grid needs leaves with their origins on the device");
    RTC8X_REQUIRE(threads > 0 && threads % 256 == 0 && steps > 0,
                  "rtc8x_gather_probe: steps must be > 0, threads a multiple of 256");
    RTC8X_REQUIRE(width == 1 || width == 2 || width == 4, "rtc8x_gather_probe: width must be 1, 2 or 4");
    RTC8X_REQUIRE(waves_per_simd >= 2 && waves_per_simd <= 8, "rtc8x_gather_probe: waves_per_simd must be in [2, 8]");
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
    hipStream_t s = (hipStream_t)stream;
    if (width == 1)
        hipLaunchKernelGGL(gather_probe_kernel<1>, dim3(threads / 256), dim3(256), lds, s, g, steps, seed, out);
    else if (width == 2)
        hipLaunchKernelGGL(gather_probe_kernel<2>, dim3(threads / 256), dim3(256), lds, s, g, steps, seed, out);
    else
        hipLaunchKernelGGL(gather_probe_kernel<4>, dim3(threads / 256), dim3(256), lds, s, g, steps, seed, out);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_gather_probe")

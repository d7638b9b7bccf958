// rtc8x_env.hip -- environment importance sampling set-up: the hierarchical importance map
// (computeFirstMipOfEnvIBLImportanceMap / computeMipOfImportanceMap, renderer/gpu_kernels/
// compute_light_probs.cu:161-211, driven as renderer/renderer_scene.h:990-1069) and the two-level
// alias tables over texels (built in FP64 on the host, Vose's method).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "rtc8x_common.hpp"
#include "rtc8x_device.hpp"

namespace rtc8x {
namespace dev {

// mip 0 = luminance * (pi/2) * sin(pi (y + 1/2) / H) inside the image, 0 in the power-of-two padding.
__global__ void env_mip0_kernel(const float* __restrict__ texels, uint32_t w, uint32_t h, uint32_t dim_x, uint32_t dim_y,
                                float* __restrict__ mip0)
{
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= dim_x || y >= dim_y) return;
    float importance = 0.0f;
    if (x < w && y < h) {
        const float4 t = *reinterpret_cast<const float4*>(texels + ((size_t)y * w + x) * 4);
        float correction = RTC8X_PI_F / 2.0f * rtc8x_sinf(RTC8X_PI_F * ((float)y + 0.5f) / (float)h);
        importance = luminance(RGB{t.x, t.y, t.z}) * correction;
    }
    mip0[(size_t)y * dim_x + x] = importance;
}

// 2x2 box reduction (sum of ul, ur, ll, lr, then * 1/4); the 1x1 top level sets the integral.
__global__ void env_mip_kernel(const float* __restrict__ src, uint32_t sdx, uint32_t sdy, float* __restrict__ dst,
                               uint32_t ddx, uint32_t ddy, int top, float top_scale, float* __restrict__ integral)
{
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
    uint32_t ulx = 2 * x, uly = 2 * y;
    float sum = 0.0f;
    sum += (ulx < sdx && uly < sdy) ? src[(size_t)uly * sdx + ulx] : 0.0f;
    sum += (ulx + 1 < sdx && uly < sdy) ? src[(size_t)uly * sdx + ulx + 1] : 0.0f;
    sum += (ulx < sdx && uly + 1 < sdy) ? src[(size_t)(uly + 1) * sdx + ulx] : 0.0f;
    sum += (ulx + 1 < sdx && uly + 1 < sdy) ? src[(size_t)(uly + 1) * sdx + ulx + 1] : 0.0f;
    if (x < ddx && y < ddy) {
        sum *= 0.25f;
        dst[(size_t)y * ddx + x] = sum;
        if (top) *integral = sum * top_scale;
    }
}

}  // namespace dev
}  // namespace rtc8x

using namespace rtc8x;

namespace {

uint32_t next_pow2(uint32_t x)
{
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

uint32_t pow2_exponent(uint32_t x)  // nextPowOf2Exponent (common/common_shared.h:129-133)
{
    if (x == 0) return 0;
    uint32_t e = 0;
    while ((1u << e) < x) ++e;
    return e;
}

// Vose's alias method in FP64; prob_to_pick_first stored as float.
void build_alias(const double* w, uint32_t n, rtc8x_alias_entry* out)
{
    double total = 0.0;
    for (uint32_t i = 0; i < n; ++i) total += w[i];
    std::vector<double> scaled(n);
    std::vector<uint32_t> small, large;
    small.reserve(n);
    large.reserve(n);
    for (uint32_t i = 0; i < n; ++i) {
        scaled[i] = total > 0.0 ? w[i] * (double)n / total : 1.0;
        (scaled[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        uint32_t s = small.back();
        small.pop_back();
        uint32_t l = large.back();
        out[s].prob_to_pick_first = (float)scaled[s];
        out[s].second_index = l;
        scaled[l] = (scaled[l] + scaled[s]) - 1.0;
        if (scaled[l] < 1.0) {
            large.pop_back();
            small.push_back(l);
        }
    }
    for (uint32_t i : large) {
        out[i].prob_to_pick_first = 1.0f;
        out[i].second_index = i;
    }
    for (uint32_t i : small) {  // numerical leftovers
        out[i].prob_to_pick_first = 1.0f;
        out[i].second_index = i;
    }
}

}  // namespace

extern "C" rtc8x_status rtc8x_env_release(rtc8x_env_light* env)
try {
    RTC8X_REQUIRE(env != nullptr, "rtc8x_env_release: env is NULL");
    const void* ptrs[] = {env->mips, env->alias_rows, env->alias_cols, env->alias_weights};
    for (const void* p : ptrs)
        if (p) RTC8X_HIP(hipFree(const_cast<void*>(p)));
    env->mips = nullptr;
    env->alias_rows = nullptr;
    env->alias_cols = nullptr;
    env->alias_weights = nullptr;
    env->num_mip_levels = 0;
    env->map_integral = 0.0f;
    env->alias_integral = 0.0f;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_env_release")

extern "C" rtc8x_status rtc8x_env_build_sampler(rtc8x_env_light* env, void* stream)
try {
    RTC8X_REQUIRE(env != nullptr, "rtc8x_env_build_sampler: env is NULL");
    if (env->kind != RTC8X_ENV_LATLONG) return RTC8X_OK;
    RTC8X_REQUIRE(env->texels && env->width > 0 && env->height > 0 && env->width <= 16384 && env->height <= 16384,
                  "rtc8x_env_build_sampler: bad texture");
    hipStream_t s = (hipStream_t)stream;
    const uint32_t w = env->width, h = env->height;
    // map dimensions (compute_light_probs.cu:166-167)
    const uint32_t dim_x = next_pow2(w);
    const uint32_t dim_y = std::max(next_pow2(h), dim_x >> 1);
    RTC8X_REQUIRE(dim_x == dim_y || dim_x == 2 * dim_y, "rtc8x_env_build_sampler: map %ux%u must be 1:1 or 2:1", dim_x, dim_y);
    const uint32_t nlev = pow2_exponent(dim_x) + 1;
    size_t total = 0;
    std::vector<size_t> offs(nlev);
    for (uint32_t l = 0; l < nlev; ++l) {
        offs[l] = total;
        total += (size_t)std::max(dim_x >> l, 1u) * std::max(dim_y >> l, 1u);
    }
    float* mips = nullptr;
    float* d_integral = nullptr;
    RTC8X_HIP(hipMalloc(&mips, total * sizeof(float)));
    RTC8X_HIP(hipMalloc(&d_integral, sizeof(float)));
    RTC8X_HIP(hipMemsetAsync(d_integral, 0, sizeof(float), s));
    {
        dim3 blk(32, 8), grd((dim_x + 31) / 32, (dim_y + 7) / 8);
        hipLaunchKernelGGL(dev::env_mip0_kernel, grd, blk, 0, s, env->texels, w, h, dim_x, dim_y, mips);
        RTC8X_HIP(hipGetLastError());
    }
    for (uint32_t l = 1; l < nlev; ++l) {
        uint32_t sdx = std::max(dim_x >> (l - 1), 1u), sdy = std::max(dim_y >> (l - 1), 1u);
        uint32_t ddx = std::max(dim_x >> l, 1u), ddy = std::max(dim_y >> l, 1u);
        int top = (l == nlev - 1) ? 1 : 0;
        float top_scale = (float)(1u << (2 * (nlev - 1)));
        dim3 blk(16, 16), grd((ddx + 15) / 16, (ddy + 15) / 16);
        hipLaunchKernelGGL(dev::env_mip_kernel, grd, blk, 0, s, mips + offs[l - 1], sdx, sdy, mips + offs[l], ddx, ddy,
                           top, top_scale, d_integral);
        RTC8X_HIP(hipGetLastError());
    }
    // alias tables from the mip-0 weights (image region), FP64
    std::vector<float> mip0((size_t)dim_x * dim_y);
    float map_integral = 0.0f;
    RTC8X_HIP(hipMemcpyAsync(mip0.data(), mips, mip0.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    RTC8X_HIP(hipMemcpyAsync(&map_integral, d_integral, sizeof(float), hipMemcpyDeviceToHost, s));
    RTC8X_HIP(hipStreamSynchronize(s));
    RTC8X_HIP(hipFree(d_integral));
    if (nlev == 1) map_integral = mip0[0];
    std::vector<float> weights((size_t)w * h);
    std::vector<double> row_sums(h, 0.0), wd(w);
    std::vector<rtc8x_alias_entry> cols((size_t)w * h), rows(h);
    double integral = 0.0;
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            float v = mip0[(size_t)y * dim_x + x];
            weights[(size_t)y * w + x] = v;
            wd[x] = (double)v;
            row_sums[y] += (double)v;
        }
        build_alias(wd.data(), w, cols.data() + (size_t)y * w);
        integral += row_sums[y];
    }
    build_alias(row_sums.data(), h, rows.data());
    RTC8X_REQUIRE(integral > 0.0, "rtc8x_env_build_sampler: environment is black");
    rtc8x_alias_entry *d_rows, *d_cols;
    float* d_w;
    RTC8X_HIP(hipMalloc(&d_rows, rows.size() * sizeof(rtc8x_alias_entry)));
    RTC8X_HIP(hipMalloc(&d_cols, cols.size() * sizeof(rtc8x_alias_entry)));
    RTC8X_HIP(hipMalloc(&d_w, weights.size() * sizeof(float)));
    RTC8X_HIP(hipMemcpy(d_rows, rows.data(), rows.size() * sizeof(rtc8x_alias_entry), hipMemcpyHostToDevice));
    RTC8X_HIP(hipMemcpy(d_cols, cols.data(), cols.size() * sizeof(rtc8x_alias_entry), hipMemcpyHostToDevice));
    RTC8X_HIP(hipMemcpy(d_w, weights.data(), weights.size() * sizeof(float), hipMemcpyHostToDevice));
    env->map_dim_x = dim_x;
    env->map_dim_y = dim_y;
    env->num_mip_levels = nlev;
    env->mips = mips;
    env->map_integral = map_integral;
    env->alias_integral = (float)integral;
    env->alias_rows = d_rows;
    env->alias_cols = d_cols;
    env->alias_weights = d_w;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_env_build_sampler")

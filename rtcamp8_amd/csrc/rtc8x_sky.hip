// rtc8x_sky.hip -- Hosek-Wilkie 2013 sky bake in FP32 (replaces renderer/gpu_kernels/
// analytic_sky_kernels.cu).  Per texel: 16 spectral bands of the sky model (plus the solar disc with
// limb darkening), binned CIE 1931 colour matching to XYZ, linear sRGB, the reference's saturation
// adjustment, and a sea colour below the horizon.  The band states are inputs (the host computes them
// from the model's datasets, renderer/renderer_scene.cpp:570-596).
//
// Compute-bound and run once per frame (3200 x 1600 texels x 16 bands); one lane per texel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "rtc8x_common.hpp"
#include "rtc8x_device.hpp"

namespace rtc8x {
namespace dev {

// ArHosekSkyModel_GetRadianceInternal (ext/hosek_wilkie_2013/ArHosekSkyModel.c:291-304) in float
// The Mie denominator 1 + g^2 - 2 g cos(gamma) is evaluated as (1 - g)^2 + 4 g sin^2(gamma / 2): the
// same value without the FP32 cancellation the original form has around the sun (gamma -> 0, g -> 1).
__device__ __forceinline__ float hw_radiance_internal(const float* c, float theta, float gamma)
{
    const float sh = sinf(0.5f * gamma);
    const float cg = 1.0f - 2.0f * sh * sh;
    const float exp_m = expf(c[4] * gamma);
    const float ray_m = cg * cg;
    const float one_m_g = 1.0f - c[8];
    const float mie_m = (1.0f + cg * cg) / powf(one_m_g * one_m_g + 4.0f * c[8] * (sh * sh), 1.5f);
    const float zenith = sqrtf(cosf(theta));
    return (1.0f + c[0] * expf(c[1] / (cosf(theta) + 0.01f))) *
           (c[2] + c[3] * exp_m + c[5] * ray_m + c[6] * mie_m + c[7] * zenith);
}

// arhosekskymodel_radiance (analytic_sky_kernels.cu:29-66)
__device__ float hw_radiance(const rtc8x_hosek_band& s, float theta, float gamma, float wavelength)
{
    int low_wl = (int)((wavelength - 320.0f) / 40.0f);
    if (low_wl < 0 || low_wl >= 11) return 0.0f;
    float interp = fmodf((wavelength - 320.0f) / 40.0f, 1.0f);
    float val_low = hw_radiance_internal(s.configs[low_wl], theta, gamma) * s.radiances[low_wl] *
                    s.emission_correction_factor_sky[low_wl];
    if (interp < 1e-6f) return val_low;
    float result = (1.0f - interp) * val_low;
    if (low_wl + 1 < 11)
        result += interp * hw_radiance_internal(s.configs[low_wl + 1], theta, gamma) * s.radiances[low_wl + 1] *
                  s.emission_correction_factor_sky[low_wl + 1];
    return result;
}

// arhosekskymodel_sr_internal (analytic_sky_kernels.cu:68-97): 45 pieces of cubic polynomials in elevation
__device__ float hw_sr_internal(const rtc8x_hosek_sky& sky, const rtc8x_hosek_band& s, int turbidity, int wl,
                                float elevation)
{
    constexpr int pieces = 45, order = 4;
    int pos = (int)(powf(2.0f * elevation / RTC8X_PI_F, 1.0f / 3.0f) * pieces);
    if (pos > 44) pos = 44;
    if (pos < 0) pos = 0;
    const float break_x = powf((float)pos / (float)pieces, 3.0f) * (RTC8X_PI_F * 0.5f);
    int t = turbidity - sky.solar_turbidity_base;
    t = t < 0 ? 0 : (t > 1 ? 1 : t);
    const float* coefs = sky.solar_rows + (size_t)wl * 360 + (size_t)t * 180 + (order * (pos + 1) - 1);
    float res = 0.0f;
    const float x = elevation - break_x;
    float x_exp = 1.0f;
    for (int i = 0; i < order; ++i) {
        res += x_exp * *coefs--;
        x_exp *= x;
    }
    return res * s.emission_correction_factor_sun[wl];
}

// arhosekskymodel_solar_radiance_internal2 (analytic_sky_kernels.cu:99-192)
__device__ float hw_solar_internal2(const rtc8x_hosek_sky& sky, const rtc8x_hosek_band& s, float wavelength,
                                    float elevation, float gamma)
{
    int turb_low = (int)s.turbidity - 1;
    float turb_frac = s.turbidity - (float)(turb_low + 1);
    if (turb_low == 9) {
        turb_low = 8;
        turb_frac = 1.0f;
    }
    int wl_low = (int)((wavelength - 320.0f) / 40.0f);
    float wl_frac = fmodf(wavelength, 40.0f) / 40.0f;
    if (wl_low == 10) {
        wl_low = 9;
        wl_frac = 1.0f;
    }
    float direct = (1.0f - turb_frac) * ((1.0f - wl_frac) * hw_sr_internal(sky, s, turb_low, wl_low, elevation) +
                                         wl_frac * hw_sr_internal(sky, s, turb_low, wl_low + 1, elevation)) +
                   turb_frac * ((1.0f - wl_frac) * hw_sr_internal(sky, s, turb_low + 1, wl_low, elevation) +
                                wl_frac * hw_sr_internal(sky, s, turb_low + 1, wl_low + 1, elevation));
    float ld[6];
    for (int i = 0; i < 6; ++i)
        ld[i] = (1.0f - wl_frac) * sky.limb_darkening[6 * wl_low + i] + wl_frac * sky.limb_darkening[6 * (wl_low + 1) + i];
    const float sol_rad_sin = sinf(s.solar_radius);
    const float ar2 = 1.0f / (sol_rad_sin * sol_rad_sin);
    const float singamma = sinf(gamma);
    float sc2 = 1.0f - ar2 * singamma * singamma;
    if (sc2 < 0.0f) sc2 = 0.0f;
    const float sc = sqrtf(sc2);
    const float darkening = ld[0] + ld[1] * sc + ld[2] * powf(sc, 2.0f) + ld[3] * powf(sc, 3.0f) +
                            ld[4] * powf(sc, 4.0f) + ld[5] * powf(sc, 5.0f);
    return direct * darkening;
}

__device__ float hw_solar_radiance(const rtc8x_hosek_sky& sky, const rtc8x_hosek_band& s, float theta, float gamma,
                                   float wavelength)
{
    return hw_solar_internal2(sky, s, wavelength, (RTC8X_PI_F / 2.0f) - theta, gamma) +
           hw_radiance(s, theta, gamma, wavelength);
}

// RGBtoHSV / HSVtoRGB (common/basic_types.h:3114-3163)
__device__ void rgb_to_hsv(RGB c, float& h, float& s, float& v)
{
    float mn = fminf(fminf(c.r, c.g), c.b);
    float mx = fmaxf(fmaxf(c.r, c.g), c.b);
    v = mx;
    float delta = mx - mn;
    if (delta < 1e-5f || mx == 0.0f) {
        h = 0.0f;
        s = 0.0f;
        return;
    }
    s = delta / mx;
    if (c.r >= mx)
        h = (c.g - c.b) / delta;
    else if (c.g >= mx)
        h = 2.0f + (c.b - c.r) / delta;
    else
        h = 4.0f + (c.r - c.g) / delta;
    h /= 6.0f;
    h = fmodf(h + 1.0f, 1.0f);
}

__device__ RGB hsv_to_rgb(float h, float s, float v)
{
    if (s == 0.0f) return RGB{v, v, v};
    h = h - floorf(h);
    int hi = (int)(h * 6.0f);
    float f = h * 6.0f - (float)hi;
    float m = v * (1.0f - s), n = v * (1.0f - s * f), k = v * (1.0f - s * (1.0f - f));
    switch (hi) {
    case 0: return RGB{v, k, m};
    case 1: return RGB{n, v, m};
    case 2: return RGB{m, v, k};
    case 3: return RGB{m, n, v};
    case 4: return RGB{k, m, v};
    case 5: return RGB{v, m, n};
    default: return RGB{0.0f, 0.0f, 0.0f};
    }
}

// Per-texel body of generateArHosekSkyEnvironmentalTexture (analytic_sky_kernels.cu:227-274).
__device__ RGB sky_texel(const rtc8x_hosek_sky& sky, V3 sun, float phi, float theta)
{
    float m_theta = theta;
    const bool lower = theta >= 0.5f * RTC8X_PI_F;
    if (lower) m_theta = RTC8X_PI_F - theta;
    const float st = sinf(m_theta), ct = cosf(m_theta), sp = sinf(phi), cp = cosf(phi);
    const V3 view{-sp * st, ct, cp * st};  // Vector3D::fromPolarYUp
    // Angle to the sun.  The reference takes acos(dot(view, sun)), which in FP32 loses most digits
    // inside the 0.0044 rad solar disc (dot is within 1e-5 of 1); 2 asin(|view - sun| / 2) is the
    // same angle and stays accurate there.
    const float dx = view.x - sun.x, dy = view.y - sun.y, dz = view.z - sun.z;
    const float half_chord = 0.5f * sqrtf(dx * dx + dy * dy + dz * dz);
    const float gamma = 2.0f * asinf(fminf(half_chord, 1.0f));
    float xyz[3] = {0.0f, 0.0f, 0.0f};
    for (int b = 0; b < RTC8X_SKY_BANDS; ++b) {
        const rtc8x_hosek_band& s = sky.bands[b];
        float cw = fminf(fmaxf(sky.center_wavelengths[b], 320.0f), 720.0f);
        float value = (gamma < s.solar_radius && !lower) ? hw_solar_radiance(sky, s, m_theta, gamma, cw)
                                                          : hw_radiance(s, m_theta, gamma, cw);
        xyz[0] += sky.cmf_x[b] * value;
        xyz[1] += sky.cmf_y[b] * value;
        xyz[2] += sky.cmf_z[b] * value;
    }
    xyz[0] /= sky.integral_cmf;
    xyz[1] /= sky.integral_cmf;
    xyz[2] /= sky.integral_cmf;
    // RGBSpectrum::fromXYZ (common/basic_types.h:3000-3012)
    RGB rgb{3.2404542f * xyz[0] + -1.5371385f * xyz[1] + -0.4985314f * xyz[2],
            -0.9692660f * xyz[0] + 1.8760108f * xyz[1] + 0.0415560f * xyz[2],
            0.0556434f * xyz[0] + -0.2040259f * xyz[1] + 1.0572252f * xyz[2]};
    // saturation adjustment for high altitude (analytic_sky_kernels.cu:257-265)
    float max_v = fmaxf(rgb.r, fmaxf(rgb.g, rgb.b));
    rgb = scale(rgb, 1.0f / max_v);
    float h, s, v;
    rgb_to_hsv(rgb, h, s, v);
    float diff_hue = fmodf(h - 2.0f / 3 + 1.0f, 1.0f);
    float diff_angle = 2.0f * RTC8X_PI_F * diff_hue;
    float t = s * powf(0.5f * (1.0f + cosf(diff_angle)), 1.0f);
    s = (1.0f - t) * s + t * 1.0f;
    rgb = scale(hsv_to_rgb(h, s, 1.0f), max_v);
    if (lower) {
        RGB sea = scale(RGB{0.01f, 0.01f, (1.0f - sun.y) * 0.01f + sun.y * 0.1f}, 0.5f);
        float w = powf(cosf(m_theta), 1.0f / 5.0f);
        rgb = add(scale(rgb, 1.0f - w), scale(sea, w));
    }
    return rgb;
}

__global__ void sky_bake_kernel(const rtc8x_hosek_sky sky, V3 sun, uint32_t w, uint32_t h, float* __restrict__ out)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    const float theta = RTC8X_PI_F * ((float)y + 0.5f) / (float)h;
    const float phi = 2.0f * RTC8X_PI_F * ((float)x + 0.5f) / (float)w;
    RGB c = sky_texel(sky, sun, phi, theta);
    *reinterpret_cast<float4*>(out + ((size_t)y * w + x) * 4) = make_float4(c.r, c.g, c.b, 1.0f);
}

__global__ void sky_eval_kernel(const rtc8x_hosek_sky sky, V3 sun, uint32_t n, const float* __restrict__ phi,
                                const float* __restrict__ theta, float* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    RGB c = sky_texel(sky, sun, phi[i], theta[i]);
    out[3 * i] = c.r;
    out[3 * i + 1] = c.g;
    out[3 * i + 2] = c.b;
}

}  // namespace dev
}  // namespace rtc8x

using namespace rtc8x;

static rtc8x_status check_sky(const rtc8x_hosek_sky* sky, const float* sun)
{
    RTC8X_REQUIRE(sky && sun, "rtc8x_sky: NULL argument");
    RTC8X_REQUIRE(sky->bands && sky->solar_rows && sky->limb_darkening, "rtc8x_sky: model tables missing");
    RTC8X_REQUIRE(sky->integral_cmf > 0.0f, "rtc8x_sky: integral_cmf must be positive");
    float l2 = sun[0] * sun[0] + sun[1] * sun[1] + sun[2] * sun[2];
    RTC8X_REQUIRE(std::isfinite(l2) && std::fabs(l2 - 1.0f) < 1e-3f, "rtc8x_sky: sun direction must be unit length");
    return RTC8X_OK;
}

extern "C" rtc8x_status rtc8x_sky_bake(void* stream, const rtc8x_hosek_sky* sky, const float sun_dir[3], uint32_t width,
                                       uint32_t height, float* out_rgba)
try {
    rtc8x_status st = check_sky(sky, sun_dir);
    if (st != RTC8X_OK) return st;
    RTC8X_REQUIRE(out_rgba && width > 0 && height > 0 && width <= 16384 && height <= 16384, "rtc8x_sky_bake: bad output");
    dim3 blk(16, 16), grd((width + 15) / 16, (height + 15) / 16);
    hipLaunchKernelGGL(dev::sky_bake_kernel, grd, blk, 0, (hipStream_t)stream, *sky,
                       dev::V3{sun_dir[0], sun_dir[1], sun_dir[2]}, width, height, out_rgba);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_sky_bake")

extern "C" rtc8x_status rtc8x_sky_eval(void* stream, const rtc8x_hosek_sky* sky, const float sun_dir[3], uint32_t n,
                                       const float* phi, const float* theta, float* out_rgb)
try {
    rtc8x_status st = check_sky(sky, sun_dir);
    if (st != RTC8X_OK) return st;
    RTC8X_REQUIRE(phi && theta && out_rgb, "rtc8x_sky_eval: NULL buffer");
    if (n == 0) return RTC8X_OK;
    hipLaunchKernelGGL(dev::sky_eval_kernel, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)stream, *sky,
                       dev::V3{sun_dir[0], sun_dir[1], sun_dir[2]}, n, phi, theta, out_rgb);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_sky_eval")

// ---------------------------------------------------------------------------------------------------
// The spectral model's tables and its state for any sun elevation and turbidity (host code).
// ---------------------------------------------------------------------------------------------------

struct rtc8x_hosek_model {
    std::vector<double> configs;    // [11][1080]: [albedo 0,1][turbidity 1..10][6 control points][9]
    std::vector<double> radiances;  // [11][120]:  [albedo][turbidity][6]
    std::vector<double> solar;      // [11][1800]: [turbidity][45 pieces][4]
    std::vector<double> limb;       // [11][6]
    float cmf[4][RTC8X_SKY_BANDS];
    float integral_cmf;
    // device copies for the bake (owned; replaced by each rtc8x_hosek_model_state call)
    rtc8x_hosek_band* d_bands = nullptr;
    float* d_solar = nullptr;
    float* d_limb = nullptr;
};

namespace {

constexpr size_t kHosekConfigs = 11 * 1080, kHosekRadiances = 11 * 120, kHosekSolar = 11 * 1800, kHosekLimb = 11 * 6;
constexpr size_t kHosekBytes = 8 + 8 * (kHosekConfigs + kHosekRadiances + kHosekSolar + kHosekLimb) +
                               4 * (4 * RTC8X_SKY_BANDS + 1);

// The quintic Bernstein blend of six control points at x, term by term as the model evaluates it
// (ArHosekSkyModel.c:171-178).
double bernstein5(const double* a, size_t stride, double x)
{
    return std::pow(1.0 - x, 5.0) * a[0] + 5.0 * std::pow(1.0 - x, 4.0) * x * a[stride] +
           10.0 * std::pow(1.0 - x, 3.0) * std::pow(x, 2.0) * a[2 * stride] +
           10.0 * std::pow(1.0 - x, 2.0) * std::pow(x, 3.0) * a[3 * stride] +
           5.0 * (1.0 - x) * std::pow(x, 4.0) * a[4 * stride] + std::pow(x, 5.0) * a[5 * stride];
}

// ArHosekSkyModel_CookConfiguration / _CookRadianceConfiguration (ArHosekSkyModel.c:147-289): bilinear
// in ground albedo (tables 0 and 1) and turbidity (integer levels T and T + 1), Bernstein in the cube
// root of the normalised solar elevation.  `block` is 9 for the configuration, 1 for the radiance.
void cook(const double* table, size_t block, double turbidity, double albedo, double elevation, double* out)
{
    const int it = (int)turbidity;
    const double rem = turbidity - (double)it;
    const double x = std::pow(elevation / (3.14159265358979323846 / 2.0), 1.0 / 3.0);
    const size_t per_albedo = 6 * block * 10;
    for (size_t i = 0; i < block; ++i) {
        const double* lo0 = table + 6 * block * (size_t)(it - 1) + i;
        const double* lo1 = table + per_albedo + 6 * block * (size_t)(it - 1) + i;
        out[i] = (1.0 - albedo) * (1.0 - rem) * bernstein5(lo0, block, x);
        out[i] += albedo * (1.0 - rem) * bernstein5(lo1, block, x);
    }
    if (it == 10) return;
    for (size_t i = 0; i < block; ++i) {
        const double* hi0 = table + 6 * block * (size_t)it + i;
        const double* hi1 = table + per_albedo + 6 * block * (size_t)it + i;
        out[i] += (1.0 - albedo) * rem * bernstein5(hi0, block, x);
        out[i] += albedo * rem * bernstein5(hi1, block, x);
    }
}

void free_device(rtc8x_hosek_model* m)
{
    if (m->d_bands) (void)hipFree(m->d_bands);
    if (m->d_solar) (void)hipFree(m->d_solar);
    if (m->d_limb) (void)hipFree(m->d_limb);
    m->d_bands = nullptr;
    m->d_solar = nullptr;
    m->d_limb = nullptr;
}

}  // namespace

extern "C" rtc8x_status rtc8x_hosek_model_create(const void* data, size_t size, rtc8x_hosek_model** out)
try {
    RTC8X_REQUIRE(data && out, "rtc8x_hosek_model_create: NULL argument");
    RTC8X_REQUIRE(size == kHosekBytes && std::memcmp(data, "RTC8XHW1", 8) == 0,
                  "rtc8x_hosek_model_create: not a model table image (%zu bytes, expected %zu)", size, kHosekBytes);
    auto m = new rtc8x_hosek_model();
    const uint8_t* p = (const uint8_t*)data + 8;
    auto take = [&](std::vector<double>& v, size_t n) {
        v.resize(n);
        std::memcpy(v.data(), p, 8 * n);
        p += 8 * n;
    };
    take(m->configs, kHosekConfigs);
    take(m->radiances, kHosekRadiances);
    take(m->solar, kHosekSolar);
    take(m->limb, kHosekLimb);
    std::memcpy(m->cmf, p, sizeof(m->cmf));
    p += sizeof(m->cmf);
    std::memcpy(&m->integral_cmf, p, 4);
    *out = m;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_hosek_model_create")

extern "C" rtc8x_status rtc8x_hosek_model_destroy(rtc8x_hosek_model* m)
try {
    if (!m) return RTC8X_OK;
    free_device(m);
    delete m;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_hosek_model_destroy")

extern "C" rtc8x_status rtc8x_hosek_bands(const rtc8x_hosek_model* m, float solar_elevation, float turbidity,
                                          rtc8x_hosek_band* out)
try {
    RTC8X_REQUIRE(m && out, "rtc8x_hosek_bands: NULL argument");
    RTC8X_REQUIRE(turbidity >= 1.0f && turbidity <= 10.0f, "rtc8x_hosek_bands: turbidity %g outside [1, 10]",
                  (double)turbidity);
    RTC8X_REQUIRE(solar_elevation >= 0.0f && solar_elevation <= 1.5707964f,
                  "rtc8x_hosek_bands: solar elevation %g outside [0, pi/2]", (double)solar_elevation);
    for (uint32_t b = 0; b < RTC8X_SKY_BANDS; ++b) {
        // ground albedo per band (renderer_scene.cpp:576), then arhosekskymodelstate_alloc_init
        // (ArHosekSkyModel.c:308-345) with every field cast to float (renderer_scene.cpp:579-590)
        const float albedo = 0.5f * (1.0f - ((float)b + 0.5f) / (float)RTC8X_SKY_BANDS);
        rtc8x_hosek_band& s = out[b];
        for (int wl = 0; wl < 11; ++wl) {
            double cfg[9], rad;
            cook(m->configs.data() + (size_t)wl * 1080, 9, turbidity, albedo, solar_elevation, cfg);
            cook(m->radiances.data() + (size_t)wl * 120, 1, turbidity, albedo, solar_elevation, &rad);
            for (int j = 0; j < 9; ++j) s.configs[wl][j] = (float)cfg[j];
            s.radiances[wl] = (float)rad;
            s.emission_correction_factor_sky[wl] = 1.0f;
            s.emission_correction_factor_sun[wl] = 1.0f;
        }
        s.turbidity = turbidity;
        s.solar_radius = (float)((0.51 * (3.14159265358979323846 / 180.0)) / 2.0);
        s.albedo = albedo;
        s.elevation = solar_elevation;
    }
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_hosek_bands")

extern "C" rtc8x_status rtc8x_hosek_model_state(rtc8x_hosek_model* m, float solar_elevation, float turbidity,
                                                rtc8x_hosek_sky* sky, void* stream)
try {
    RTC8X_REQUIRE(m && sky, "rtc8x_hosek_model_state: NULL argument");
    std::vector<rtc8x_hosek_band> bands(RTC8X_SKY_BANDS);
    rtc8x_status st = rtc8x_hosek_bands(m, solar_elevation, turbidity, bands.data());
    if (st != RTC8X_OK) return st;
    // the solar-disc rows the bake reads: turbidity levels base and base + 1 (ArHosekSkyModel.c:700-707)
    int base = (int)turbidity - 1;
    if (base == 9) base = 8;
    std::vector<float> rows((size_t)11 * 2 * 180), limb(kHosekLimb);
    for (int wl = 0; wl < 11; ++wl)
        for (int t = 0; t < 2; ++t)
            for (int k = 0; k < 180; ++k)
                rows[((size_t)wl * 2 + t) * 180 + k] = (float)m->solar[(size_t)wl * 1800 + (size_t)(base + t) * 180 + k];
    for (size_t i = 0; i < kHosekLimb; ++i) limb[i] = (float)m->limb[i];
    hipStream_t s = (hipStream_t)stream;
    RTC8X_HIP(hipStreamSynchronize(s));  // the previous state may still be read by a bake on this stream
    if (!m->d_bands) {
        RTC8X_HIP(hipMalloc((void**)&m->d_bands, sizeof(rtc8x_hosek_band) * RTC8X_SKY_BANDS));
        RTC8X_HIP(hipMalloc((void**)&m->d_solar, sizeof(float) * rows.size()));
        RTC8X_HIP(hipMalloc((void**)&m->d_limb, sizeof(float) * limb.size()));
    }
    RTC8X_HIP(hipMemcpy(m->d_bands, bands.data(), sizeof(rtc8x_hosek_band) * RTC8X_SKY_BANDS, hipMemcpyHostToDevice));
    RTC8X_HIP(hipMemcpy(m->d_solar, rows.data(), sizeof(float) * rows.size(), hipMemcpyHostToDevice));
    RTC8X_HIP(hipMemcpy(m->d_limb, limb.data(), sizeof(float) * limb.size(), hipMemcpyHostToDevice));
    sky->bands = m->d_bands;
    sky->solar_rows = m->d_solar;
    sky->limb_darkening = m->d_limb;
    sky->solar_turbidity_base = base;
    for (uint32_t b = 0; b < RTC8X_SKY_BANDS; ++b) {
        sky->cmf_x[b] = m->cmf[0][b];
        sky->cmf_y[b] = m->cmf[1][b];
        sky->cmf_z[b] = m->cmf[2][b];
        sky->center_wavelengths[b] = m->cmf[3][b];
    }
    sky->integral_cmf = m->integral_cmf;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_hosek_model_state")

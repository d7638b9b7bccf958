/*
 * rtc8_oracle_post.c -- CPU oracle of the post-process pass and the SDR conversion (TEST
 * INFRASTRUCTURE ONLY; see rtc8_oracle.h).
 *
 * oracle_tone_map restates applyToneMap (renderer/gpu_kernels/post_process_kernels.cu:9-40) over the
 * accumulator's mean, followed by saveImage's SDR conversion (common/common_host.cpp:515-570).  The
 * FP32 variant uses rtc8x_detmath.h (bit-identical to the kernel); the FP64 variant uses libm and pins
 * the formulas to the mathematics.
 */
#include <math.h>
#include <stdint.h>

#include "rtc8_oracle.h"
#include "../include/rtc8x_detmath.h"

/* RGBSpectrum::luminance, common/basic_types.h:2981-2985 */
static float lum_f(const float c[3]) { return 0.2126729f * c[0] + 0.7151522f * c[1] + 0.0721750f * c[2]; }
static double lum_d(const double c[3]) { return 0.2126729 * c[0] + 0.7151522 * c[1] + 0.0721750 * c[2]; }

static uint32_t q8(double s)
{
    /* static_cast<uint32_t>(v * 255) clamped to 255 (common_host.cpp:562-565); NaN and negatives -> 0 */
    if (!(s >= 1.0)) return 0u;
    if (s >= 255.0) return 255u;
    return (uint32_t)s;
}

/* sRGB_gamma_s, common/basic_types.h:3180-3186 */
static float gamma_f(float v) { return v <= 0.0031308f ? 12.92f * v : 1.055f * rtc8x_powf(v, (float)(1.0 / 2.4)) - 0.055f; }
static double gamma_d(double v) { return v <= 0.0031308 ? 12.92 * v : 1.055 * pow(v, 1.0 / 2.4) - 0.055; }

static uint32_t sdr_f(float c[3], float alpha, const rtc8x_sdr_config* cfg)
{
    if (cfg->apply_tone_map) {
        int finite = 1;
        for (int k = 0; k < 3; ++k) finite &= (c[k] - c[k] == 0.0f);
        for (int k = 0; k < 3; ++k) c[k] = finite ? (c[k] < 0.0f ? 0.0f : c[k]) : 0.0f;
        float l = lum_f(c);
        float lt = 1.0f - rtc8x_expf(-(cfg->brightness_scale * l));   /* simpleToneMap_s, basic_types.h:3166-3169 */
        float s = l > 0.0f ? lt / l : 0.0f;
        for (int k = 0; k < 3; ++k) c[k] = c[k] * s;
    }
    if (cfg->apply_srgb_gamma)
        for (int k = 0; k < 3; ++k) c[k] = gamma_f(c[k] < 0.0f ? 0.0f : c[k]);
    if (cfg->alpha_override >= 0.0f) alpha = cfg->alpha_override;
    return q8(c[0] * 255.0f) | (q8(c[1] * 255.0f) << 8) | (q8(c[2] * 255.0f) << 16) | (q8(alpha * 255.0f) << 24);
}

static uint32_t sdr_d(double c[3], double alpha, const rtc8x_sdr_config* cfg)
{
    if (cfg->apply_tone_map) {
        int finite = isfinite(c[0]) && isfinite(c[1]) && isfinite(c[2]);
        for (int k = 0; k < 3; ++k) c[k] = finite ? (c[k] < 0.0 ? 0.0 : c[k]) : 0.0;
        double l = lum_d(c);
        double s = l > 0.0 ? (1.0 - exp(-(cfg->brightness_scale * l))) / l : 0.0;
        for (int k = 0; k < 3; ++k) c[k] *= s;
    }
    if (cfg->apply_srgb_gamma)
        for (int k = 0; k < 3; ++k) c[k] = gamma_d(c[k] < 0.0 ? 0.0 : c[k]);
    if (cfg->alpha_override >= 0.0f) alpha = cfg->alpha_override;
    return q8(c[0] * 255.0) | (q8(c[1] * 255.0) << 8) | (q8(c[2] * 255.0) << 16) | (q8(alpha * 255.0) << 24);
}

void oracle_tone_map(const double* accum, uint32_t w, uint32_t h, uint32_t samples, float exponent, float brightness,
                     const rtc8x_sdr_config* cfg, float* out_rgba32f, uint32_t* out_rgba8, int32_t f64)
{
    const double inv = 1.0 / (double)samples;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            const uint32_t sy = cfg->flip_y ? h - 1 - y : y;
            if (!f64) {
                float c[3];
                for (int k = 0; k < 3; ++k) c[k] = rtc8x_powf((float)(accum[3 * i + k] * inv), exponent);
                float l = lum_f(c);
                if (l > 0.0f) {   /* applySimpleToneMap, post_process_kernels.cu:9-18 */
                    float s = (1.0f - rtc8x_expf(-brightness * l)) / l;
                    for (int k = 0; k < 3; ++k) c[k] = c[k] * s;
                }
                if (out_rgba32f)
                    for (int k = 0; k < 4; ++k) out_rgba32f[4 * i + k] = k < 3 ? c[k] : 1.0f;
                if (out_rgba8) out_rgba8[(size_t)sy * w + x] = sdr_f(c, 1.0f, cfg);
            } else {
                double c[3];
                for (int k = 0; k < 3; ++k) c[k] = pow(accum[3 * i + k] * inv, (double)exponent);
                double l = lum_d(c);
                if (l > 0.0) {
                    double s = (1.0 - exp(-(double)brightness * l)) / l;
                    for (int k = 0; k < 3; ++k) c[k] *= s;
                }
                if (out_rgba32f)
                    for (int k = 0; k < 4; ++k) out_rgba32f[4 * i + k] = (float)(k < 3 ? c[k] : 1.0);
                if (out_rgba8) out_rgba8[(size_t)sy * w + x] = sdr_d(c, 1.0, cfg);
            }
        }
}

void oracle_sdr_convert(const float* data, uint32_t w, uint32_t h, uint32_t nch, const rtc8x_sdr_config* cfg,
                        uint32_t* out, int32_t f64)
{
    for (uint32_t y = 0; y < h; ++y) {
        const uint32_t sy = cfg->flip_y ? h - 1 - y : y;   /* common_host.cpp:533 */
        for (uint32_t x = 0; x < w; ++x) {
            const float* d = data + ((size_t)sy * w + x) * nch;
            const float a = nch == 4 ? d[3] : 1.0f;
            if (!f64) {
                float c[3] = {d[0], nch >= 2 ? d[1] : 0.0f, nch >= 3 ? d[2] : 0.0f};
                out[(size_t)y * w + x] = sdr_f(c, a, cfg);
            } else {
                double c[3] = {d[0], nch >= 2 ? d[1] : 0.0, nch >= 3 ? d[2] : 0.0};
                out[(size_t)y * w + x] = sdr_d(c, a, cfg);
            }
        }
    }
}

/* The deterministic exp / pow the post-process uses, exported for accuracy tests against libm. */
float oracle_expf(float x) { return rtc8x_expf(x); }
float oracle_powf(float x, float y) { return rtc8x_powf(x, y); }

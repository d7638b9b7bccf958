/*
 * rtc8_oracle_tracer.c -- CPU oracle tracer (TEST INFRASTRUCTURE ONLY).
 *
 * Instantiates rtc8_oracle_tracer.inc twice: FP32 with the deterministic math of rtc8x_detmath.h
 * (bit-identical to the gfx950 kernels) and FP64 with libm (the converged-image oracle).
 * Threads: one pthread per requested worker, rows handed out dynamically.
 */
#include "rtc8_oracle.h"
#include "../include/rtc8x_detmath.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---------------- FP32 instance */
#define R float
#define SUF(name) name##_f32
#define LOG(x) rtc8x_logf(x)
#define SQRT(x) sqrtf(x)
#define SINCOS(x, s, c) rtc8x_sincosf((x), (s), (c))
#define ACOS(x) rtc8x_acosf(x)
#define ATAN2(y, x) rtc8x_atan2f((y), (x))
#define FMOD(x, y) rtc8x_fmodf((x), (y))
#define FLOOR(x) floorf(x)
#define FMA(a, b, c) fmaf((a), (b), (c))
#define FMIN(a, b) fminf((a), (b))
#define FMAX(a, b) fmaxf((a), (b))
#include "rtc8_oracle_tracer.inc"
#undef R
#undef SUF
#undef LOG
#undef SQRT
#undef SINCOS
#undef ACOS
#undef ATAN2
#undef FMOD
#undef FLOOR
#undef FMA
#undef FMIN
#undef FMAX
#undef UNI
#undef RTC8X_HARD_MAX_PATH_LENGTH

/* ---------------- FP64 instance */
static void sincos_f64(double x, double* s, double* c) { *s = sin(x); *c = cos(x); }
#define R double
#define SUF(name) name##_f64
#define LOG(x) log(x)
#define SQRT(x) sqrt(x)
#define SINCOS(x, s, c) sincos_f64((x), (s), (c))
#define ACOS(x) acos(x)
#define ATAN2(y, x) atan2((y), (x))
#define FMOD(x, y) fmod((x), (y))
#define FLOOR(x) floor(x)
#define FMA(a, b, c) fma((a), (b), (c))
#define FMIN(a, b) fmin((a), (b))
#define FMAX(a, b) fmax((a), (b))
#include "rtc8_oracle_tracer.inc"

int oracle_trace_f32(const rtc8x_frame_params* p, int32_t entry, uint32_t num_samples,
                     uint32_t y0, uint32_t y1, uint32_t nthreads)
{
    return oracle_trace_impl_f32(p, entry, num_samples, 0, 0xFFFFFFFFu, y0, y1, nthreads);
}

int oracle_trace_f64(const rtc8x_frame_params* p, int32_t entry, uint32_t num_samples,
                     uint32_t y0, uint32_t y1, uint32_t nthreads)
{
    return oracle_trace_impl_f64(p, entry, num_samples, 0, 0xFFFFFFFFu, y0, y1, nthreads);
}

int oracle_trace_window_f64(const rtc8x_frame_params* p, uint32_t num_samples, uint32_t x0, uint32_t x1,
                            uint32_t y0, uint32_t y1, uint32_t nthreads)
{
    return oracle_trace_impl_f64(p, RTC8X_ENTRY_PATH_TRACE, num_samples, x0, x1, y0, y1, nthreads);
}

/* Environment sampling hooks for the tests (FP64 instance). */
void oracle_env_sample_f64(const rtc8x_env_light* e, double u0, double u1, double dir[3], double* pdf, double rgb[3])
{
    vec3_f64 d;
    rgb_f64 c = env_sample_f64(e, u0, u1, &d, pdf);
    dir[0] = d.x;
    dir[1] = d.y;
    dir[2] = d.z;
    rgb[0] = c.r;
    rgb[1] = c.g;
    rgb[2] = c.b;
}

void oracle_env_evaluate_f64(const rtc8x_env_light* e, const double dir[3], double* pdf, double rgb[3])
{
    vec3_f64 d;
    d.x = dir[0];
    d.y = dir[1];
    d.z = dir[2];
    rgb_f64 c = env_evaluate_f64(e, d, pdf);
    rgb[0] = c.r;
    rgb[1] = c.g;
    rgb[2] = c.b;
}

/* One segment's tracking and the phase functions on their own, for the closed-form and brute-force
 * tests of the oracle.  The medium is the frame's (grid, density coefficient, majorant = coefficient x
 * grid maximum, built as the tracer's workers build it); ray i starts at o + 0 d with its own PCG state
 * rng[i], which is advanced.  free flight: out[i] = collision distance t, or -1 for an escape
 * (optix_pathtracing_kernels.cu:98-126); transmittance: out[i] = ratio-tracked T
 * (renderer_kernel_common.h:360-383).  lookups counts the density lookups of all rays. */
#define ORACLE_SEGMENT_HOOK(SFX, RT)                                                                            \
    void oracle_track_batch_##SFX(const rtc8x_frame_params* p, int32_t transmittance, uint32_t n, uint64_t* rng, \
                                  const float* o, const float* d, float tmax, double* out, uint64_t* lookups)  \
    {                                                                                                           \
        medium_##SFX m;                                                                                         \
        medium_init_##SFX(&m, p);                                                                               \
        for (uint32_t i = 0; i < n; ++i) {                                                                      \
            vec3_##SFX oo = v3_##SFX((RT)o[3 * i], (RT)o[3 * i + 1], (RT)o[3 * i + 2]);                        \
            vec3_##SFX dd = v3_##SFX((RT)d[3 * i], (RT)d[3 * i + 1], (RT)d[3 * i + 2]);                        \
            if (transmittance) {                                                                                \
                out[i] = (double)transmittance_##SFX(p, &m, &rng[i], oo, dd, (RT)tmax);                        \
            } else {                                                                                            \
                RT t = 0;                                                                                       \
                out[i] = free_flight_##SFX(p, &m, &rng[i], oo, dd, (RT)tmax, &t) ? (double)t : -1.0;            \
            }                                                                                                   \
        }                                                                                                       \
        if (lookups) *lookups = m.lookups;                                                                      \
    }                                                                                                           \
    /* phase sampling (common_renderer_types.h:2063-2097): u holds (u0, u1) pairs; dir the local           \
     * directions, value the phase value returned with them */                                              \
    void oracle_phase_sample_batch_##SFX(int32_t kind, float k, uint32_t n, const double* u, double* dir,      \
                                         double* value)                                                         \
    {                                                                                                           \
        phase_##SFX ph;                                                                                         \
        ph.kind = kind;                                                                                         \
        ph.k = (RT)k;                                                                                           \
        for (uint32_t i = 0; i < n; ++i) {                                                                      \
            vec3_##SFX v;                                                                                       \
            value[i] = (double)phase_sample_##SFX(&ph, (RT)u[2 * i], (RT)u[2 * i + 1], &v);                    \
            dir[3 * i] = (double)v.x;                                                                           \
            dir[3 * i + 1] = (double)v.y;                                                                       \
            dir[3 * i + 2] = (double)v.z;                                                                       \
        }                                                                                                       \
    }                                                                                                           \
    double oracle_phase_value_##SFX(int32_t kind, float k, double cos_theta)                                   \
    {                                                                                                           \
        phase_##SFX ph;                                                                                         \
        ph.kind = kind;                                                                                         \
        ph.k = (RT)k;                                                                                           \
        return (double)phase_value_##SFX(&ph, (RT)cos_theta);                                                   \
    }

ORACLE_SEGMENT_HOOK(f32, float)
ORACLE_SEGMENT_HOOK(f64, double)

/* The hierarchical importance map of a lat-long texture (compute_light_probs.cu:161-211, driven as
 * renderer_scene.h:990-1069), built here from the texels alone, in FP32 in the kernels' order:
 * the map is dim_x x dim_y = the texture's size rounded up to powers of two (dim_y >= dim_x / 2);
 * mip 0 = luminance x (pi/2) sin(pi (y + 1/2) / H) inside the image and 0 in the padding; every
 * coarser level is the 2x2 box sum of the finer one times 1/4; integral = top level x 4^(levels-1).
 * mips (NULL: only the sizes) receives the levels back to back.  Returns the number of levels, 0 when
 * the map is neither 1:1 nor 2:1. */
uint32_t oracle_env_build_mips(const float* texels, uint32_t w, uint32_t h, float* mips, uint32_t* dim_x_out,
                               uint32_t* dim_y_out, float* integral)
{
    uint32_t dx = 1, dy = 1, nlev = 1;
    while (dx < w) dx <<= 1;
    while (dy < h) dy <<= 1;
    if (dy < dx / 2) dy = dx / 2;
    if (!(dx == dy || dx == 2 * dy)) return 0;
    while ((1u << (nlev - 1)) < dx) ++nlev;
    if (dim_x_out) *dim_x_out = dx;
    if (dim_y_out) *dim_y_out = dy;
    if (!mips) return nlev;
    for (uint32_t y = 0; y < dy; ++y)
        for (uint32_t x = 0; x < dx; ++x) {
            float v = 0.0f;
            if (x < w && y < h) {
                const float* t = texels + ((size_t)y * w + x) * 4;
                rgb_f32 c = rgb_make_f32(t[0], t[1], t[2]);
                float correction = RTC8X_PI_F / 2.0f * rtc8x_sinf(RTC8X_PI_F * ((float)y + 0.5f) / (float)h);
                v = luminance_f32(c) * correction;
            }
            mips[(size_t)y * dx + x] = v;
        }
    float* src = mips;
    uint32_t sdx = dx, sdy = dy;
    for (uint32_t l = 1; l < nlev; ++l) {
        uint32_t ddx = sdx > 1 ? sdx / 2 : 1, ddy = sdy > 1 ? sdy / 2 : 1;
        float* dst = src + (size_t)sdx * sdy;
        for (uint32_t y = 0; y < ddy; ++y)
            for (uint32_t x = 0; x < ddx; ++x) {
                uint32_t ux = 2 * x, uy = 2 * y;
                float sum = 0.0f;
                sum += (ux < sdx && uy < sdy) ? src[(size_t)uy * sdx + ux] : 0.0f;
                sum += (ux + 1 < sdx && uy < sdy) ? src[(size_t)uy * sdx + ux + 1] : 0.0f;
                sum += (ux < sdx && uy + 1 < sdy) ? src[(size_t)(uy + 1) * sdx + ux] : 0.0f;
                sum += (ux + 1 < sdx && uy + 1 < sdy) ? src[(size_t)(uy + 1) * sdx + ux + 1] : 0.0f;
                dst[(size_t)y * ddx + x] = sum * 0.25f;
            }
        src = dst;
        sdx = ddx;
        sdy = ddy;
    }
    if (integral) *integral = nlev == 1 ? src[0] : src[0] * (float)(1u << (2 * (nlev - 1)));
    return nlev;
}

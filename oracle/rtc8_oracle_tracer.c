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
    return oracle_trace_impl_f32(p, entry, num_samples, y0, y1, nthreads);
}

int oracle_trace_f64(const rtc8x_frame_params* p, int32_t entry, uint32_t num_samples,
                     uint32_t y0, uint32_t y1, uint32_t nthreads)
{
    return oracle_trace_impl_f64(p, entry, num_samples, y0, y1, nthreads);
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

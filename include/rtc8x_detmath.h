/*
 * rtc8x_detmath.h -- deterministic single-precision math for the volumetric path.
 *
 * The free-flight loop of the delta tracker decides "collide / pass" by comparing
 * FP32 numbers built from log(), and the phase sampler feeds cos/sin into the next
 * ray.  Vendor math libraries (the host libm and the device math library) are free
 * to differ in the last ulp, which would make the number of path vertices differ
 * between a scalar host run and the GPU run on the same PCG stream.  Every function
 * here is therefore written with IEEE-754 basic operations only (+ - * /, sqrt and
 * floor, all correctly rounded on both sides, no contraction), so that the host and
 * the gfx950 device produce bit-identical results.
 *
 * Methods (textbook, constants derived from pi / ln 2 in this file's comments):
 *   log  : x = 2^e * m, m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1),
 *          atanh series s + s^3/3 + ... + s^11/11 (|s| <= 0.1716, truncation < 2^-28).
 *          e*ln2 is added as e*ln2_hi + e*ln2_lo (Cody & Waite split, ln2_hi has the
 *          low 12 mantissa bits cleared so e*ln2_hi is exact).
 *   sin/cos : k = floor(x*2/pi + 1/2), r = x - k*(pi/2) with a three-part Cody-Waite
 *          split of pi/2 (each part has 12 cleared low bits, exact for |k| < 2^12),
 *          Taylor polynomials on |r| <= pi/4 (sin to r^11, cos to r^12).
 *   atan : |x| > 1 -> pi/2 - atan(1/x); x > 2-sqrt(3) -> pi/6 + atan((sqrt3 x - 1)/(x + sqrt3));
 *          Taylor series to t^15 on |t| <= 2-sqrt(3) (truncation < 2^-27).
 *   atan2, acos are built on atan with the usual quadrant rules.
 *   exp  : k = floor(x/ln2 + 1/2), r = x - k*ln2_hi - k*ln2_lo (Cody & Waite, |r| <= 0.35),
 *          Taylor series of e^r to r^8 (truncation < 2^-29 relative), then scaled by 2^k
 *          built from exponent bits (two factors for subnormal results).
 *   pow  : x^y = exp(y log x) for x > 0 (relative error grows with |y log x|: about
 *          (1 + |y log x|) * 2^-22), with the IEEE special cases for 0, +inf and NaN;
 *          negative bases give NaN.
 *
 * The header is plain C99 (the CPU oracle includes it) and HIP C++ (the kernels
 * include it).  It must be compiled without -ffast-math and with -ffp-contract=off.
 * Accuracy is a few ulp; it is checked against libm in the test-suite.
 */
#ifndef RTC8X_DETMATH_H
#define RTC8X_DETMATH_H

#include <stdint.h>
#include <string.h>
#include <math.h>

#if defined(__HIPCC__)
#  define RTC8X_HD __host__ __device__ static inline
#else
#  define RTC8X_HD static inline
#endif

/* Float constants (hex literals are exact). */
#define RTC8X_PI_F        0x1.921fb6p+1f   /* float(pi) = 3.14159274 */
#define RTC8X_2PI_F       0x1.921fb6p+2f   /* 2 * float(pi) */
#define RTC8X_PIO2_F      0x1.921fb6p+0f
#define RTC8X_2OPI_F      0x1.45f306p-1f   /* float(2/pi) */
#define RTC8X_PIO2_HI_F   0x1.92p+0f       /* pi/2 with the low 12 mantissa bits cleared */
#define RTC8X_PIO2_MID_F  0x1.fb4p-12f     /* next 12 bits of pi/2 */
#define RTC8X_PIO2_LO_F   0x1.4442d2p-24f  /* remainder, float-rounded */
#define RTC8X_LN2_HI_F    0x1.62ep-1f      /* ln 2 with the low 12 mantissa bits cleared */
#define RTC8X_LN2_LO_F    0x1.0bfbe8p-15f  /* ln 2 - ln2_hi, float-rounded */
#define RTC8X_SQRT2_F     0x1.6a09e6p+0f
#define RTC8X_SQRT3_F     0x1.bb67aep+0f
#define RTC8X_2MSQRT3_F   0x1.126146p-2f   /* 2 - sqrt(3) */
#define RTC8X_PIO6_F      0x1.0c1524p-1f

RTC8X_HD uint32_t rtc8x_f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
RTC8X_HD float rtc8x_u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* Natural logarithm for x > 0 (returns -inf for 0, NaN for x < 0 or NaN, +inf for +inf). */
RTC8X_HD float rtc8x_logf(float x)
{
    if (!(x > 0.0f)) {
        if (x == 0.0f) return -rtc8x_u2f(0x7f800000u);
        return rtc8x_u2f(0x7fc00000u);
    }
    if (x == rtc8x_u2f(0x7f800000u)) return x;
    uint32_t bits = rtc8x_f2u(x);
    int32_t e = 0;
    if (bits < 0x00800000u) {             /* subnormal: scale by 2^24 (exact) */
        x = x * 16777216.0f;
        bits = rtc8x_f2u(x);
        e = -24;
    }
    e += (int32_t)(bits >> 23) - 127;
    float m = rtc8x_u2f((bits & 0x007fffffu) | 0x3f800000u);   /* [1, 2) */
    if (m > RTC8X_SQRT2_F) { m = m * 0.5f; e += 1; }            /* [0.7071, 1.4142] */
    float num = m - 1.0f;                                        /* exact (Sterbenz) */
    float den = m + 1.0f;
    float s = num / den;
    float s2 = s * s;
    float p = 0x1.745d18p-4f;                   /* 1/11 */
    p = p * s2 + 0x1.c71c72p-4f;                /* 1/9 */
    p = p * s2 + 0x1.24924ap-3f;                /* 1/7 */
    p = p * s2 + 0x1.99999ap-3f;                /* 1/5 */
    p = p * s2 + 0x1.555556p-2f;                /* 1/3 */
    float lm = s + s * (s2 * p);                /* atanh(s) */
    lm = lm + lm;
    float fe = (float)e;
    return fe * RTC8X_LN2_HI_F + (lm + fe * RTC8X_LN2_LO_F);
}

#define RTC8X_INVLN2_F    0x1.715476p+0f   /* float(1/ln 2) */

/* e^x for all float x (NaN in, NaN out; overflow to +inf above 88.72284, underflow to 0). */
RTC8X_HD float rtc8x_expf(float x)
{
    if (x != x) return x;
    if (x > 88.72284f) return rtc8x_u2f(0x7f800000u);
    if (x < -103.97208f) return 0.0f;
    float kf = floorf(x * RTC8X_INVLN2_F + 0.5f);
    float r = x - kf * RTC8X_LN2_HI_F;      /* exact: |k| <= 150 and ln2_hi has 12 free bits */
    r = r - kf * RTC8X_LN2_LO_F;
    float p = 0x1.a01a02p-16f;                /* 1/8! */
    p = p * r + 0x1.a01a02p-13f;              /* 1/7! */
    p = p * r + 0x1.6c16c2p-10f;              /* 1/6! */
    p = p * r + 0x1.111112p-7f;               /* 1/5! */
    p = p * r + 0x1.555556p-5f;               /* 1/4! */
    p = p * r + 0x1.555556p-3f;               /* 1/3! */
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int32_t k = (int32_t)kf;
    if (k >= -126) {
        if (k > 127) return (p * 2.0f) * rtc8x_u2f((uint32_t)(k - 1 + 127) << 23);
        return p * rtc8x_u2f((uint32_t)(k + 127) << 23);
    }
    /* subnormal result: 2^k = 2^(k+64) * 2^-64, both factors normal */
    return (p * rtc8x_u2f((uint32_t)(k + 64 + 127) << 23)) * 0x1p-64f;
}

/* x^y for x >= 0 (NaN for x < 0 or NaN operands; x^0 = 1). */
RTC8X_HD float rtc8x_powf(float x, float y)
{
    if (y == 0.0f) return 1.0f;
    if (x != x || y != y || x < 0.0f) return rtc8x_u2f(0x7fc00000u);
    if (x == 0.0f) return y > 0.0f ? 0.0f : rtc8x_u2f(0x7f800000u);
    if (x == rtc8x_u2f(0x7f800000u)) return y > 0.0f ? x : 0.0f;
    if (x == 1.0f) return 1.0f;
    return rtc8x_expf(y * rtc8x_logf(x));
}

/* sin and cos of x for |x| < 4096 * pi/2 (the phase sampler uses [0, 2pi)). */
RTC8X_HD void rtc8x_sincosf(float x, float* s_out, float* c_out)
{
    float kf = floorf(x * RTC8X_2OPI_F + 0.5f);
    float r = x - kf * RTC8X_PIO2_HI_F;
    r = r - kf * RTC8X_PIO2_MID_F;
    r = r - kf * RTC8X_PIO2_LO_F;
    int32_t k = (int32_t)kf;
    float r2 = r * r;
    /* sin r = r - r^3/3! + r^5/5! - r^7/7! + r^9/9! - r^11/11! */
    float ps = -0x1.ae6456p-26f;                /* -1/11! */
    ps = ps * r2 + 0x1.71de3ap-19f;             /*  1/9!  */
    ps = ps * r2 - 0x1.a01a02p-13f;             /* -1/7!  */
    ps = ps * r2 + 0x1.111112p-7f;              /*  1/5!  */
    ps = ps * r2 - 0x1.555556p-3f;              /* -1/3!  */
    float sr = r + r * (r2 * ps);
    /* cos r = 1 - r^2/2! + r^4/4! - r^6/6! + r^8/8! - r^10/10! + r^12/12! */
    float pc = 0x1.1eed8ep-29f;                 /*  1/12! */
    pc = pc * r2 - 0x1.27e4fcp-22f;             /* -1/10! */
    pc = pc * r2 + 0x1.a01a02p-16f;             /*  1/8!  */
    pc = pc * r2 - 0x1.6c16c2p-10f;             /* -1/6!  */
    pc = pc * r2 + 0x1.555556p-5f;              /*  1/4!  */
    pc = pc * r2 - 0.5f;
    float cr = 1.0f + r2 * pc;
    switch (k & 3) {
    case 0: *s_out = sr;  *c_out = cr;  break;
    case 1: *s_out = cr;  *c_out = -sr; break;
    case 2: *s_out = -sr; *c_out = -cr; break;
    default: *s_out = -cr; *c_out = sr; break;
    }
}

RTC8X_HD float rtc8x_sinf(float x) { float s, c; rtc8x_sincosf(x, &s, &c); return s; }
RTC8X_HD float rtc8x_cosf(float x) { float s, c; rtc8x_sincosf(x, &s, &c); return c; }

/* atan for x >= 0 (finite or +inf). */
RTC8X_HD float rtc8x_atan_pos(float x)
{
    int inv = 0;
    if (x > 1.0f) { x = 1.0f / x; inv = 1; }
    float base = 0.0f;
    if (x > RTC8X_2MSQRT3_F) {
        x = (RTC8X_SQRT3_F * x - 1.0f) / (x + RTC8X_SQRT3_F);
        base = RTC8X_PIO6_F;
    }
    float t2 = x * x;
    /* atan t = t - t^3/3 + t^5/5 - ... - t^15/15 */
    float p = -0x1.111112p-4f;                  /* -1/15 */
    p = p * t2 + 0x1.3b13b2p-4f;                /*  1/13 */
    p = p * t2 - 0x1.745d18p-4f;                /* -1/11 */
    p = p * t2 + 0x1.c71c72p-4f;                /*  1/9  */
    p = p * t2 - 0x1.24924ap-3f;                /* -1/7  */
    p = p * t2 + 0x1.99999ap-3f;                /*  1/5  */
    p = p * t2 - 0x1.555556p-2f;                /* -1/3  */
    float a = base + (x + x * (t2 * p));
    if (inv) a = RTC8X_PIO2_F - a;
    return a;
}

RTC8X_HD float rtc8x_atan2f(float y, float x)
{
    float ay = y < 0.0f ? -y : y;
    float ax = x < 0.0f ? -x : x;
    float a;
    if (ax == 0.0f && ay == 0.0f) {
        a = 0.0f;
    } else if (ax >= ay) {
        a = rtc8x_atan_pos(ay / ax);
    } else {
        a = RTC8X_PIO2_F - rtc8x_atan_pos(ax / ay);
    }
    if (x < 0.0f) a = RTC8X_PI_F - a;
    return y < 0.0f ? -a : a;
}

/* acos on [-1, 1]; the caller clamps. */
RTC8X_HD float rtc8x_acosf(float y)
{
    float s = sqrtf((1.0f - y) * (1.0f + y));
    return rtc8x_atan2f(s, y);
}

/* fmod for the two call patterns of the reference (positive divisor, |x| < 2^24 * y). */
RTC8X_HD float rtc8x_fmodf(float x, float y)
{
    float q = x / y;
    q = q < 0.0f ? -floorf(-q) : floorf(q);
    float r = x - q * y;
    /* Correct the (rare) case where rounding of x/y pushed q one step too far. */
    if (x >= 0.0f && r < 0.0f) r = r + y;
    if (x < 0.0f && r > 0.0f) r = r - y;
    return r;
}

#endif /* RTC8X_DETMATH_H */

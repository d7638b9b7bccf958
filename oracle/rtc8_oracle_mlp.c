/*
 * rtc8_oracle_mlp.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) of the neural radiance cache network:
 * input encoding, forward pass, relative-L2-luminance loss and its weight gradient in FP64, and the
 * Adam + EMA step in FP32.  The network replaces tiny-cuda-nn behind renderer/network_interface.cu,
 * whose exact numerics are not available (the tiny-cuda-nn submodule is not vendored), so this
 * restates the published methods: FullyFusedMLP without biases, RelativeL2Luminance (Mueller et al.
 * 2021, normaliser held constant), Adam (Kingma & Ba 2015) with L2 regularisation, weight EMA.
 *
 * `round_mode` emulates the library's storage precision (0: none, 1: FP16, 2: BF16) on weights,
 * activations, output gradients and pre-activation gradients, so that the only remaining difference
 * to the GPU is the accumulation order of the FP32 MFMA dot products.
 */
#include "rtc8_oracle.h"
#include "../include/rtc8x_detmath.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define W 64
#define LAYER (W * W)

uint32_t oracle_nrc_num_params(uint32_t L) { return L * LAYER + RTC8X_NRC_OUTPUT_PAD * W; }

/* IEEE binary16, round to nearest even (overflow to inf), back to float */
static float round_fp16(float x)
{
    uint32_t u = rtc8x_f2u(x);
    uint32_t sign = u & 0x80000000u;
    uint32_t ax = u & 0x7fffffffu;
    if (ax >= 0x7f800000u) return x;                         /* inf / nan */
    float a = rtc8x_u2f(ax);
    if (a >= 65520.0f) return rtc8x_u2f(sign | 0x7f800000u); /* rounds to inf */
    if (a < 6.103515625e-05f) {                              /* subnormal half: quantum 2^-24 */
        float q = a * 16777216.0f;                           /* exact scaling */
        float r = nearbyintf(q);                             /* round half to even (default mode) */
        return rtc8x_u2f(sign | rtc8x_f2u(r / 16777216.0f));
    }
    uint32_t lsb = (ax >> 13) & 1u;
    uint32_t rounded = (ax + 0xfffu + lsb) & ~0x1fffu;       /* keep 10 mantissa bits, RNE */
    return rtc8x_u2f(sign | rounded);
}

static float round_bf16(float x)
{
    uint32_t u = rtc8x_f2u(x);
    if ((u & 0x7fffffffu) >= 0x7f800000u) return x;
    uint32_t lsb = (u >> 16) & 1u;
    u = (u + 0x7fffu + lsb) & 0xffff0000u;
    return rtc8x_u2f(u);
}

static double rnd(double x, int mode)
{
    if (mode == 1) return (double)round_fp16((float)x);
    if (mode == 2) return (double)round_bf16((float)x);
    return x;
}

/* The library's 64-feature input encoding (include/rtc8x.h, "Neural radiance cache network"). */
void oracle_nrc_encode(const rtc8x_radiance_query* q, uint32_t n, double* out)
{
    const double pi = 3.14159265358979323846;
    for (uint32_t s = 0; s < n; ++s) {
        double* e = out + (size_t)s * 64;
        for (int d = 0; d < 3; ++d)
            for (int f = 0; f < 12; ++f) {
                double t = (double)q[s].position[d] * ldexp(1.0, f - 1);  /* period 2 / 2^f */
                t = t - floor(t);
                e[12 * d + f] = fabs(2.0 * t - 1.0);
            }
        double a[5] = {q[s].normal_phi / pi, q[s].normal_theta / (2 * pi), q[s].vout_phi / pi, q[s].vout_theta / (2 * pi),
                       q[s].roughness};
        for (int j = 0; j < 5; ++j)
            for (int b = 0; b < 4; ++b) {
                double dd = a[j] - (b + 0.5) * 0.25;
                e[36 + 4 * j + b] = exp(-8.0 * dd * dd);
            }
        for (int c = 0; c < 3; ++c) {
            e[56 + c] = q[s].diffuse_reflectance[c];
            e[59 + c] = q[s].specular_reflectance[c];
        }
        e[62] = 1.0;
        e[63] = 0.0;
    }
}

/* ---- HashGrid position encoding (multiresolution hash encoding, Mueller et al. 2022) with the
 * reference's settings (renderer/network_interface.cu:103-111): 16 levels x 2 features, 2^15 entries,
 * base resolution 16, x2 per level.  Level geometry and indexing as include/rtc8x.h states them. */
#define HL RTC8X_NRC_HASH_LEVELS
#define HT (1u << RTC8X_NRC_HASH_LOG2_TABLE)

typedef struct {
    double scale[HL];
    uint32_t stride[HL], offset[HL];
    uint32_t total;
} hash_geom;

static void hash_geometry(hash_geom* g)
{
    uint32_t off = 0;
    for (uint32_t l = 0; l < HL; ++l) {
        double sc = ldexp(1.0, (int)l) * RTC8X_NRC_HASH_BASE_RESOLUTION - 1.0;
        uint64_t r1 = (uint64_t)ceil(sc) + 2;
        uint64_t dense = r1 * r1 * r1;
        uint32_t entries = dense <= HT ? (uint32_t)((dense + 7) / 8 * 8) : HT;
        g->scale[l] = sc;
        g->stride[l] = dense <= HT ? (uint32_t)r1 : 0u;
        g->offset[l] = off;
        off += entries * RTC8X_NRC_HASH_FEATURES;
    }
    g->total = off;
}

/* the 8 corners of level l around position p in [0,1]^3: table float index and trilinear weight */
static void hash_corners(const hash_geom* g, uint32_t l, const double p[3], uint32_t idx[8], double w[8])
{
    /* the cell position in FP32, as the method computes it (p s + 1/2, two roundings): at the fine
     * levels (s up to 524287) FP32 quantises the fraction, and that is part of the encoding */
    uint32_t gi[3];
    double f[3];
    for (int d = 0; d < 3; ++d) {
        float pos = (float)p[d] * (float)g->scale[l];
        pos = pos + 0.5f;
        float fl = floorf(pos);
        f[d] = (double)(pos - fl);
        gi[d] = (uint32_t)fl;
    }
    for (uint32_t c = 0; c < 8; ++c) {
        uint32_t x = gi[0] + (c & 1u), y = gi[1] + ((c >> 1) & 1u), z = gi[2] + (c >> 2);
        w[c] = ((c & 1u) ? f[0] : 1.0 - f[0]) * ((c & 2u) ? f[1] : 1.0 - f[1]) * ((c & 4u) ? f[2] : 1.0 - f[2]);
        uint32_t st = g->stride[l];
        uint32_t e = st ? x + y * st + z * st * st : ((x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u)) & (HT - 1u);
        idx[c] = g->offset[l] + 2u * e;
    }
}

static void query_p01(const rtc8x_radiance_query* q, double p[3])
{
    for (int d = 0; d < 3; ++d) {
        float v = q->position[d];
        v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
        p[d] = (double)v;
    }
}

/* [grid 32 | OneBlob 20 | identity 6 | 1 | 0 x 5] */
static void encode_hash(const rtc8x_radiance_query* q, const float* tables, const hash_geom* g, double* e)
{
    const double pi = 3.14159265358979323846;
    double p[3];
    query_p01(q, p);
    for (uint32_t l = 0; l < HL; ++l) {
        uint32_t idx[8];
        double w[8];
        hash_corners(g, l, p, idx, w);
        double a0 = 0.0, a1 = 0.0;
        for (int c = 0; c < 8; ++c) {
            a0 += w[c] * (double)tables[idx[c]];
            a1 += w[c] * (double)tables[idx[c] + 1];
        }
        e[2 * l] = a0;
        e[2 * l + 1] = a1;
    }
    double a[5] = {q->normal_phi / pi, q->normal_theta / (2 * pi), q->vout_phi / pi, q->vout_theta / (2 * pi), q->roughness};
    for (int j = 0; j < 5; ++j)
        for (int b = 0; b < 4; ++b) {
            double dd = a[j] - (b + 0.5) * 0.25;
            e[32 + 4 * j + b] = exp(-8.0 * dd * dd);
        }
    for (int c = 0; c < 3; ++c) {
        e[52 + c] = q->diffuse_reflectance[c];
        e[55 + c] = q->specular_reflectance[c];
    }
    e[58] = 1.0;
    for (int j = 59; j < 64; ++j) e[j] = 0.0;
}

uint32_t oracle_nrc_num_params_ex(uint32_t L, int32_t encoding)
{
    hash_geom g;
    hash_geometry(&g);
    return oracle_nrc_num_params(L) + (encoding == RTC8X_NRC_HASH_GRID ? g.total : 0u);
}

void oracle_nrc_encode_ex(const float* weights, uint32_t L, int32_t encoding, const rtc8x_radiance_query* q, uint32_t n,
                          double* out)
{
    if (encoding != RTC8X_NRC_HASH_GRID) {
        oracle_nrc_encode(q, n, out);
        return;
    }
    hash_geom g;
    hash_geometry(&g);
    for (uint32_t s = 0; s < n; ++s) encode_hash(&q[s], weights + oracle_nrc_num_params(L), &g, out + (size_t)s * 64);
}

static double luminance(const double* p) { return 0.2126729 * p[0] + 0.7151522 * p[1] + 0.0721750 * p[2]; }

/* Forward + (optionally) backward for n samples.  Returns the summed per-sample loss
 * sum_c (p_c - t_c)^2 / (lum(p)^2 + 0.01); grad (if not NULL) receives the summed per-sample gradient
 * of that loss with respect to the weights (the library's unscaled backward). */
static double run(const float* weights, uint32_t L, int32_t encoding, const rtc8x_radiance_query* q, const float* target,
                  uint32_t n, double* out3, double* grad, int mode)
{
    uint32_t P = oracle_nrc_num_params(L);  /* MLP part; hash tables follow it in `weights` and `grad` */
    uint32_t P_all = oracle_nrc_num_params_ex(L, encoding);
    hash_geom hg;
    hash_geometry(&hg);
    double* w = (double*)malloc(sizeof(double) * P);
    for (uint32_t i = 0; i < P; ++i) w[i] = rnd((double)weights[i], mode);
    double* h = (double*)malloc(sizeof(double) * 64 * (L + 1));
    double* dz = (double*)malloc(sizeof(double) * 64);
    double* dh = (double*)malloc(sizeof(double) * 64);
    double enc[64];
    double total = 0.0;
    if (grad) memset(grad, 0, sizeof(double) * P_all);
    for (uint32_t s = 0; s < n; ++s) {
        if (encoding == RTC8X_NRC_HASH_GRID)
            encode_hash(&q[s], weights + P, &hg, enc);
        else
            oracle_nrc_encode(&q[s], 1, enc);
        for (int j = 0; j < 64; ++j) h[j] = rnd(enc[j], mode);
        for (uint32_t l = 0; l < L; ++l) {
            const double* wl = w + (size_t)l * LAYER;
            const double* hin = h + 64 * l;
            double* hout = h + 64 * (l + 1);
            for (int o = 0; o < W; ++o) {
                double z = 0.0;
                for (int i = 0; i < W; ++i) z += wl[o * W + i] * hin[i];
                hout[o] = rnd(z > 0.0 ? z : 0.0, mode);
            }
        }
        double p[3];
        const double* wo = w + (size_t)L * LAYER;
        for (int o = 0; o < 3; ++o) {
            double z = 0.0;
            for (int i = 0; i < W; ++i) z += wo[o * W + i] * h[64 * L + i];
            p[o] = z;
        }
        if (out3)
            for (int o = 0; o < 3; ++o) out3[3 * (size_t)s + o] = p[o];
        if (!target) continue;
        double lum = luminance(p), denom = lum * lum + 0.01;
        double g[3];
        for (int o = 0; o < 3; ++o) {
            double d = p[o] - (double)target[3 * (size_t)s + o];
            total += d * d / denom;
            g[o] = rnd(2.0 * d / denom, mode);
        }
        if (!grad) continue;
        /* output layer */
        for (int o = 0; o < 3; ++o)
            for (int i = 0; i < W; ++i) grad[(size_t)L * LAYER + o * W + i] += g[o] * h[64 * L + i];
        for (int i = 0; i < W; ++i) {
            double acc = 0.0;
            for (int o = 0; o < 3; ++o) acc += wo[o * W + i] * g[o];
            dh[i] = acc;
        }
        for (int l = (int)L - 1; l >= 0; --l) {
            const double* hout = h + 64 * (l + 1);
            for (int o = 0; o < W; ++o) dz[o] = rnd(hout[o] > 0.0 ? dh[o] : 0.0, mode);
            const double* hin = h + 64 * l;
            for (int o = 0; o < W; ++o)
                for (int i = 0; i < W; ++i) grad[(size_t)l * LAYER + o * W + i] += dz[o] * hin[i];
            if (l == 0 && encoding == RTC8X_NRC_HASH_GRID) {
                /* gradient of the 32 grid features, scattered with the trilinear weights */
                const double* w0 = w;
                double p[3];
                query_p01(&q[s], p);
                for (uint32_t lv = 0; lv < HL; ++lv) {
                    double de[2];
                    for (int f = 0; f < 2; ++f) {
                        double acc = 0.0;
                        for (int o = 0; o < W; ++o) acc += w0[o * W + 2 * lv + f] * dz[o];
                        de[f] = acc;
                    }
                    uint32_t idx[8];
                    double cw[8];
                    hash_corners(&hg, lv, p, idx, cw);
                    for (int c = 0; c < 8; ++c) {
                        grad[P + idx[c]] += cw[c] * de[0];
                        grad[P + idx[c] + 1] += cw[c] * de[1];
                    }
                }
            }
            if (l == 0) break;
            const double* wl = w + (size_t)l * LAYER;
            for (int i = 0; i < W; ++i) {
                double acc = 0.0;
                for (int o = 0; o < W; ++o) acc += wl[o * W + i] * dz[o];
                dh[i] = acc;
            }
        }
    }
    free(w);
    free(h);
    free(dz);
    free(dh);
    return total;
}

void oracle_nrc_forward(const float* weights, uint32_t L, const rtc8x_radiance_query* q, uint32_t n, double* out3,
                        int round_mode)
{
    run(weights, L, RTC8X_NRC_TRIANGLE_WAVE, q, NULL, n, out3, NULL, round_mode);
}

double oracle_nrc_loss(const float* weights, uint32_t L, const rtc8x_radiance_query* q, const float* target, uint32_t n,
                       int round_mode)
{
    return run(weights, L, RTC8X_NRC_TRIANGLE_WAVE, q, target, n, NULL, NULL, round_mode);
}

double oracle_nrc_gradient(const float* weights, uint32_t L, const rtc8x_radiance_query* q, const float* target,
                           uint32_t n, double* grad, int round_mode)
{
    return run(weights, L, RTC8X_NRC_TRIANGLE_WAVE, q, target, n, NULL, grad, round_mode);
}

void oracle_nrc_forward_ex(const float* weights, uint32_t L, int32_t encoding, const rtc8x_radiance_query* q, uint32_t n,
                           double* out3, int round_mode)
{
    run(weights, L, encoding, q, NULL, n, out3, NULL, round_mode);
}

double oracle_nrc_gradient_ex(const float* weights, uint32_t L, int32_t encoding, const rtc8x_radiance_query* q,
                              const float* target, uint32_t n, double* grad, int round_mode)
{
    return run(weights, L, encoding, q, target, n, NULL, grad, round_mode);
}

/* One Adam + EMA step in FP32, in the library's expression order, from the gradient of the mean loss. */
void oracle_nrc_adam(float* master, float* m, float* v, float* ema, const float* grad_mean, uint32_t P, float lr,
                     float beta1, float beta2, float eps, float l2, float decay, uint64_t step)
{
    float bc1 = (float)(1.0 - pow((double)beta1, (double)step));
    float bc2 = (float)(1.0 - pow((double)beta2, (double)step));
    for (uint32_t i = 0; i < P; ++i) {
        float w0 = master[i];
        float g = grad_mean[i] + l2 * w0;
        float mi = beta1 * m[i] + (1.0f - beta1) * g;
        float vi = beta2 * v[i] + (1.0f - beta2) * g * g;
        m[i] = mi;
        v[i] = vi;
        float w1 = w0 - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
        master[i] = w1;
        ema[i] = decay * ema[i] + (1.0f - decay) * w1;
    }
}

float oracle_round_fp16(float x) { return round_fp16(x); }
float oracle_round_bf16(float x) { return round_bf16(x); }

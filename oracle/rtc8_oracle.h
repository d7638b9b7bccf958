/*
 * rtc8_oracle.h -- CPU oracle for the rtc8x volumetric path (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement of rtcamp8's volume-only path-tracing hot path, used by the test-suite as
 * the parity checker and by bench.py as the timed CPU baseline.  Nothing in the shipped library may
 * call into it.  Every function cites the reference file:line it restates (paths relative to the
 * rtcamp8 source tree).
 *
 * FP32 functions use rtc8x_detmath.h so that they are bit-identical to the gfx950 kernels on the
 * same PCG streams; FP64 functions use libm and serve as the converged-image oracle.
 */
#ifndef RTC8_ORACLE_H
#define RTC8_ORACLE_H

#include "../include/rtc8x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* a1: PCG32 XSH-RR with increment 1 (common/common_renderer_types.h:101-125). */
uint32_t oracle_pcg_next(uint64_t* state);
float oracle_pcg_float(uint64_t* state);
void oracle_pcg_fill(uint64_t state, uint32_t n, uint32_t* out_u32, float* out_f32, uint64_t* out_state);
/* mt19937_64 seeding of the per-pixel states (renderer/renderer_main.cpp:86-93). */
void oracle_mt19937_64(uint64_t seed, uint32_t n, uint64_t* out);
void oracle_seed_rng(uint64_t* states, uint32_t width, uint32_t height, uint64_t seed);
/* a2: LCG (renderer/renderer_shared.h:92-109) and the shuffler seeds (renderer_main.cpp:220-229). */
uint32_t oracle_lcg_next(uint32_t* state);
void oracle_lcg_seeds(uint32_t seed, uint32_t n, uint32_t* out);

/* Grid: fBm builder into the NanoVDB tree layout, node offsets, uncached trilinear sampler. */
int oracle_grid_build_fbm(const rtc8x_fbm_desc* desc, rtc8x_grid* out_host);
void oracle_grid_free(rtc8x_grid* g);
float oracle_fbm_value(const rtc8x_fbm_desc* desc, int32_t i, int32_t j, int32_t k);
float oracle_grid_get_value(const rtc8x_grid* g, int32_t i, int32_t j, int32_t k, uint32_t offsets[3]);
void oracle_grid_probe(const rtc8x_grid* g, uint32_t n, const float* world_xyz,
                       int32_t* out_ijk, uint32_t* out_offsets, float* out_value);
/* cell_* and region_* (malloc'd); region_cells 0: the default region size */
void oracle_compute_cell_majorants(rtc8x_grid* g, uint32_t region_cells);
void oracle_compute_region_majorants(rtc8x_grid* g, uint32_t region_cells);  /* region_* from cell_* */
int oracle_ray_clip(const float eye[3], const float dir[3], float tmax, const float bmin[3],
                    const float bmax[3], float* t0, float* t1);

/* Deterministic math wrappers (for accuracy tests against libm). */
float oracle_logf(float x);
void oracle_sincosf(float x, float* s, float* c);
float oracle_atan2f(float y, float x);
float oracle_acosf(float x);

/* Tracer.  Rows [y0, y1) of the image; nthreads worker threads over rows.
 * FP32: bit-exact restatement (accum gets the FP32 contributions summed in FP64).
 * FP64: same algorithm in double precision (the converged-image oracle). */
int oracle_trace_f32(const rtc8x_frame_params* p, int32_t entry, uint32_t num_samples,
                     uint32_t y0, uint32_t y1, uint32_t nthreads);
int oracle_trace_f64(const rtc8x_frame_params* p, int32_t entry, uint32_t num_samples,
                     uint32_t y0, uint32_t y1, uint32_t nthreads);
/* PATH_TRACE of the pixels [x0, x1) x [y0, y1) only, FP64 (a crop of a large frame). */
int oracle_trace_window_f64(const rtc8x_frame_params* p, uint32_t num_samples, uint32_t x0, uint32_t x1,
                            uint32_t y0, uint32_t y1, uint32_t nthreads);
/* One segment's free-flight (transmittance = 0: out = collision distance or -1) or ratio-tracked
 * transmittance (transmittance = 1), one PCG state per ray; and the phase functions' sampling and value. */
void oracle_track_batch_f32(const rtc8x_frame_params* p, int32_t transmittance, uint32_t n, uint64_t* rng,
                            const float* o, const float* d, float tmax, double* out, uint64_t* lookups);
void oracle_track_batch_f64(const rtc8x_frame_params* p, int32_t transmittance, uint32_t n, uint64_t* rng,
                            const float* o, const float* d, float tmax, double* out, uint64_t* lookups);
void oracle_phase_sample_batch_f32(int32_t kind, float k, uint32_t n, const double* u, double* dir, double* value);
void oracle_phase_sample_batch_f64(int32_t kind, float k, uint32_t n, const double* u, double* dir, double* value);
double oracle_phase_value_f32(int32_t kind, float k, double cos_theta);
double oracle_phase_value_f64(int32_t kind, float k, double cos_theta);
/* The hierarchical importance map built from the texels alone (compute_light_probs.cu:161-211). */
uint32_t oracle_env_build_mips(const float* texels, uint32_t w, uint32_t h, float* mips, uint32_t* dim_x,
                               uint32_t* dim_y, float* integral);
/* EnvironmentalLight::sample / evaluate (common_renderer_types.h:988-1015, 2419-2458) in FP64. */
void oracle_env_sample_f64(const rtc8x_env_light* e, double u0, double u1, double dir[3], double* pdf, double rgb[3]);
void oracle_env_evaluate_f64(const rtc8x_env_light* e, const double dir[3], double* pdf, double rgb[3]);


/* NRC data path (renderer/gpu_kernels/nrc_set_up_kernels.cu) on host copies of params->nrc. */
void oracle_nrc_prepare(const rtc8x_frame_params* p, int32_t is_new_sequence);
void oracle_nrc_propagate(const rtc8x_frame_params* p);
void oracle_nrc_shuffle(const rtc8x_frame_params* p);
void oracle_nrc_accumulate(const rtc8x_frame_params* p);

/* NRC network (FP64 forward / gradient, FP32 Adam). round_mode: 0 none, 1 FP16, 2 BF16.  The loss and
 * gradient functions return sums over the n samples (the mean loss is the sum over 3n); Adam takes the
 * gradient of the mean loss. */
uint32_t oracle_nrc_num_params(uint32_t L);
void oracle_nrc_encode(const rtc8x_radiance_query* q, uint32_t n, double* out);
void oracle_nrc_forward(const float* weights, uint32_t L, const rtc8x_radiance_query* q, uint32_t n, double* out3,
                        int round_mode);
double oracle_nrc_loss(const float* weights, uint32_t L, const rtc8x_radiance_query* q, const float* target, uint32_t n,
                       int round_mode);
double oracle_nrc_gradient(const float* weights, uint32_t L, const rtc8x_radiance_query* q, const float* target,
                           uint32_t n, double* grad, int round_mode);
/* Either encoding (rtc8x_nrc_encoding); with the hash grid the tables follow the MLP weights in
 * `weights` / `grad` (oracle_nrc_num_params_ex floats in all). */
uint32_t oracle_nrc_num_params_ex(uint32_t L, int32_t encoding);
void oracle_nrc_encode_ex(const float* weights, uint32_t L, int32_t encoding, const rtc8x_radiance_query* q, uint32_t n,
                          double* out);
void oracle_nrc_forward_ex(const float* weights, uint32_t L, int32_t encoding, const rtc8x_radiance_query* q, uint32_t n,
                           double* out3, int round_mode);
double oracle_nrc_gradient_ex(const float* weights, uint32_t L, int32_t encoding, const rtc8x_radiance_query* q,
                              const float* target, uint32_t n, double* grad, int round_mode);
void oracle_nrc_adam(float* master, float* m, float* v, float* ema, const float* grad_mean, uint32_t P, float lr,
                     float beta1, float beta2, float eps, float l2, float decay, uint64_t step);
float oracle_round_fp16(float x);
float oracle_round_bf16(float x);

/* Post-process (rtc8_oracle_post.c): applyToneMap over the accumulator's mean plus the SDR conversion,
 * and the SDR conversion alone; f64 != 0 evaluates in double with libm. */
void oracle_tone_map(const double* accum, uint32_t w, uint32_t h, uint32_t samples, float exponent, float brightness,
                     const rtc8x_sdr_config* cfg, float* out_rgba32f, uint32_t* out_rgba8, int32_t f64);
void oracle_sdr_convert(const float* data, uint32_t w, uint32_t h, uint32_t nch, const rtc8x_sdr_config* cfg,
                        uint32_t* out, int32_t f64);
float oracle_expf(float x);
float oracle_powf(float x, float y);

#ifdef __cplusplus
}
#endif
#endif

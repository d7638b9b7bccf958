// rtc8x_wavefront.hip -- the plain path tracer (pathTrace entry, optix_pathtracing_kernels.cu:31-448
// with withNRC = genData = false) as a persistent per-lookup state machine.
//
// A lane owns one pixel at a time (all of its samples, so the pixel's PCG stream is consumed in the
// reference's order) and steps its path through three kinds of work: free-flight tracking towards the
// next scattering vertex, ratio tracking of the NEE shadow ray, and vertex events (camera ray, escape,
// scattering, NEE set-up, phase sampling, path end).  Every iteration of the main loop first runs the
// events each lane needs until it wants a density lookup (or has no work left), then performs exactly
// one lookup for every lane that wants one.  The lookups of a wavefront are therefore issued together,
// whatever mix of tracking and shadow rays, path lengths and pixels its lanes hold, and a lane whose
// path ends takes the next pixel from a work counter instead of idling until the wavefront's longest
// path is done.  Per pixel, the sequence of random numbers and of floating-point operations is the
// reference's, so results are bit-identical to the one-lane-one-path form.
#include <hip/hip_runtime.h>

#include <atomic>

#include "rtc8x_common.hpp"
#include "rtc8x_device.hpp"

namespace rtc8x {
namespace dev {

constexpr uint32_t kCounterSlots = 1024;
__device__ uint32_t g_work_counter[kCounterSlots];

__global__ void reset_counter_kernel(uint32_t slot) { g_work_counter[slot] = 0u; }

// Resumable free-flight / ratio tracking segment.  advance() draws the next tentative collision and
// returns true when a density lookup at `t` is needed (with the majorant `mu` that governs it), false
// when the segment is done.  after_*() consume the lookup.
template <int TRACKING>
struct Seg;

// Global majorant, the reference's loops (optix_pathtracing_kernels.cu:110-125,
// renderer_kernel_common.h:371-382).
template <>
struct Seg<RTC8X_TRACKING_GLOBAL_MAJORANT> {
    V3 o, d;
    float t, t1;
    __device__ __forceinline__ bool init(const rtc8x_grid& g, const Medium& m, V3 ro, V3 rd, float tmax)
    {
        float t0;
        if (!(m.majorant > 0.0f)) return false;
        if (!clip(g, ro, rd, tmax, t0, t1)) return false;
        t = fmaxf(0.0f, t0);
        o = ro;
        d = rd;
        return true;
    }
    __device__ __forceinline__ bool advance(const rtc8x_grid&, const Medium& m, Pcg& rng, float& mu)
    {
        t += -rtc8x_logf(1.0f - rng.uni()) / m.majorant;
        if (t > t1) return false;
        mu = m.majorant;
        return true;
    }
    // returns true when the tentative collision is real
    __device__ __forceinline__ bool after_track(Pcg& rng, float sigma, float mu) { return rng.uni() < sigma / mu; }
    __device__ __forceinline__ void after_shadow(Pcg&, float sigma, float mu, float& vis) { vis *= 1.0f - sigma / mu; }
};

// Local majorants over 8^3 cells, the loops of track_local / ratio_local made resumable.
template <>
struct Seg<RTC8X_TRACKING_LOCAL_MAJORANT> {
    V3 o, d;
    float t, t1, tau, mu_cell, tend;
    int32_t axis, in_cell;
    Dda s;
    __device__ __forceinline__ bool init(const rtc8x_grid& g, const Medium& m, V3 ro, V3 rd, float tmax)
    {
        float t0;
        if (!(m.majorant > 0.0f)) return false;
        if (!clip(g, ro, rd, tmax, t0, t1)) return false;
        t = fmaxf(0.0f, t0);
        o = ro;
        d = rd;
        s.init(g, ro, rd, t);
        in_cell = 0;
        return true;
    }
    // first call after init draws tau (as track_local does right after the DDA set-up)
    __device__ __forceinline__ void first_tau(Pcg& rng) { tau = -rtc8x_logf(1.0f - rng.uni()); }
    __device__ __forceinline__ bool advance(const rtc8x_grid& g, const Medium& m, Pcg&, float& mu)
    {
        for (;;) {
            if (!in_cell) {
                axis = s.axis();
                const float tna = s.tnext(axis);
                tend = tna < t1 ? tna : t1;
                mu_cell = cell_majorant(g, m, s);
                in_cell = 1;
            }
            if (mu_cell > 0.0f) {
                float tc = t + tau / mu_cell;
                if (tc < tend) {
                    t = tc;
                    mu = mu_cell;
                    return true;
                }
                tau = tau - mu_cell * (tend - t);
                if (tau < 0.0f) tau = 0.0f;
            }
            if (!(s.tnext(axis) < t1)) return false;
            t = tend;
            if (!s.step(g, axis)) return false;
            in_cell = 0;
        }
    }
    __device__ __forceinline__ bool after_track(Pcg& rng, float sigma, float mu)
    {
        if (rng.uni() < sigma / mu) return true;
        tau = -rtc8x_logf(1.0f - rng.uni());
        return false;
    }
    __device__ __forceinline__ void after_shadow(Pcg& rng, float sigma, float mu, float& vis)
    {
        vis *= 1.0f - sigma / mu;
        tau = -rtc8x_logf(1.0f - rng.uni());
    }
};

enum : uint32_t { kModeWork = 0, kModeTrack = 1, kModeShadow = 2, kModeIdle = 3 };

struct Lane {
    uint32_t pix, lx, ly, sample;
    Pcg rng;
    V3 org, dir, pos;
    RGB tp, contribution, nee_f;
    float dir_pd, init_importance, vis, nee_k1, nee_k2, nee_k3;
    uint32_t path_length, mode;
    bool nee_sun;
};

template <int TRACKING>
__device__ __forceinline__ void start_track(const rtc8x_frame_params& p, const Medium& med, Lane& L, Seg<TRACKING>& sg,
                                            bool& escaped)
{
    escaped = !sg.init(p.grid, med, L.org, L.dir, 1e10f);
    if (!escaped) {
        if constexpr (TRACKING == RTC8X_TRACKING_LOCAL_MAJORANT) sg.first_tau(L.rng);
        L.mode = kModeTrack;
    }
}

template <int TRACKING>
__global__ __launch_bounds__(256) void wavefront_kernel(const rtc8x_frame_params p, uint32_t num_samples, uint32_t slot)
{
    const uint32_t W = p.image_width, H = p.image_height;
    const uint32_t tiles_x = (W + 7) / 8;
    const uint32_t total = tiles_x * ((H + 7) / 8) * 64;
    const uint32_t lane_id = threadIdx.x & 63;
    const uint32_t max_len =
        p.max_path_length ? (p.max_path_length < 4096u ? p.max_path_length : 4096u) : 4096u;
    const uint32_t rr_after = p.rr_after_path_length;
    const float prob_env = (p.sun.enabled && p.env.kind != RTC8X_ENV_NONE) ? 1.0f - p.sun.select_prob : 1.0f;

    Medium med;
    med.coeff = p.density_coeff;
    med.majorant = p.density_coeff * p.grid.max_value;
    med.lookups = 0;
    uint32_t paths = 0, vertices = 0, scatters = 0;

    Lane L;
    L.mode = kModeWork;
    L.pix = 0;
    Seg<TRACKING> sg;
    float mu = 0.0f;

    for (;;) {
        bool want = false;
        // ---- events until this lane wants a lookup or has no work left
        while (L.mode != kModeIdle && !want) {
            if (L.mode == kModeWork) {
                // a new pixel from the work counter, one atomic per wavefront
                const uint64_t ask = __ballot(1);
                const uint32_t leader = (uint32_t)__ffsll((long long)ask) - 1u;
                uint32_t base = 0;
                if (lane_id == leader) base = atomicAdd(&g_work_counter[slot], (uint32_t)__popcll(ask));
                base = __shfl(base, (int)leader, 64);
                const uint32_t w = base + (uint32_t)__popcll(ask & ((1ull << lane_id) - 1ull));
                if (w >= total) {
                    L.mode = kModeIdle;
                    continue;
                }
                const uint32_t tile = w >> 6, within = w & 63u;
                L.lx = (tile % tiles_x) * 8 + (within & 7u);
                L.ly = (tile / tiles_x) * 8 + (within >> 3);
                if (L.lx >= W || L.ly >= H) continue;
                L.pix = L.ly * W + L.lx;
                L.rng.state = p.rng_states[L.pix];
                L.sample = 0;
                L.mode = kModeTrack + 16;  // start a path below
            }
            if (L.mode == kModeTrack + 16) {
                // camera ray (optix_pathtracing_kernels.cu:43-55)
                float px = ((float)L.lx + L.rng.uni()) / (float)W;
                float py = ((float)L.ly + L.rng.uni()) / (float)H;
                const float vh = p.camera.vh, vw = p.camera.vw;
                const float sensor_area = vw * vh;
                L.org = V3{p.camera.position[0], p.camera.position[1], p.camera.position[2]};
                V3 ldir = normalize(V3{vw * (-0.5f + px), vh * (0.5f - py), -1.0f});
                const float* M = p.camera.orientation;
                V3 dir{dot(V3{M[0], M[1], M[2]}, ldir), dot(V3{M[3], M[4], M[5]}, ldir), dot(V3{M[6], M[7], M[8]}, ldir)};
                L.dir = normalize(dir);
                float az = ldir.z < 0.0f ? -ldir.z : ldir.z;
                L.dir_pd = 1.0f / (az * (az * az) * sensor_area);
                L.contribution = RGB{0.0f, 0.0f, 0.0f};
                float rinv = 1.0f / L.dir_pd;
                L.tp = RGB{rinv, rinv, rinv};
                L.init_importance = luminance(L.tp);
                L.path_length = 1;
                bool escaped;
                start_track<TRACKING>(p, med, L, sg, escaped);
                if (escaped) L.mode = kModeTrack + 32;  // escape event below
            }
            if (L.mode == kModeTrack || L.mode == kModeShadow) {
                if (sg.advance(p.grid, med, L.rng, mu)) {
                    want = true;
                    continue;
                }
                L.mode = L.mode == kModeTrack ? kModeTrack + 32 : kModeShadow + 32;  // segment done
            }
            bool after_nee = false, path_end = false;
            if (L.mode == kModeTrack + 32) {
                // escaped to the environment (optix_pathtracing_kernels.cu:130-174)
                if (p.env.kind != RTC8X_ENV_NONE) {
                    float hyp_pd;
                    RGB radiance = env_evaluate(p.env, L.dir, hyp_pd);
                    float mis = 1.0f;
                    if (L.path_length > 1) {
                        float light_pd = prob_env * hyp_pd;
                        float bsdf_pd = L.dir_pd;
                        mis = bsdf_pd * bsdf_pd / (bsdf_pd * bsdf_pd + light_pd * light_pd);
                    }
                    L.contribution = add(L.contribution, mul(L.tp, scale(radiance, mis)));
                }
                path_end = true;
            } else if (L.mode == kModeShadow + 32) {
                // end of the NEE shadow ray (performNextEventEstimation, renderer_kernel_common.h:297-403)
                RGB nee_c{0.0f, 0.0f, 0.0f};
                if (L.vis > 0.0f)
                    nee_c = L.nee_sun ? scale(L.nee_f, L.vis / L.nee_k1)
                                      : scale(L.nee_f, L.vis * L.nee_k1 * L.nee_k2 / L.nee_k3);
                L.contribution = add(L.contribution, mul(L.tp, nee_c));
                after_nee = true;
            } else if (L.mode == kModeTrack + 48) {
                // scattering vertex at sg.t (optix_pathtracing_kernels.cu:176-268)
                scatters++;
                const float th = sg.t;
                L.pos = V3{L.org.x + th * L.dir.x, L.org.y + th * L.dir.y, L.org.z + th * L.dir.z};
                float continue_prob = 1.0f;
                if (L.path_length > rr_after) {
                    float q = luminance(L.tp) / L.init_importance;
                    continue_prob = fminf(q, 1.0f);
                    if (L.rng.uni() >= continue_prob) path_end = true;
                }
                if (!path_end) {
                    L.tp = scale(L.tp, 1.0f / continue_prob);
                    // NEE set-up: light choice and sample, everything of the estimate except visibility
                    const bool have_env = p.env.kind != RTC8X_ENV_NONE, have_sun = p.sun.enabled != 0;
                    float u_light = L.rng.uni();
                    bool use_sun = false;
                    float prob_cur = 1.0f;
                    if (have_env && have_sun) {
                        if (u_light < p.sun.select_prob) {
                            use_sun = true;
                            prob_cur = p.sun.select_prob;
                        } else {
                            prob_cur = 1.0f - p.sun.select_prob;
                        }
                    } else if (have_sun) {
                        use_sun = true;
                    }
                    float u0 = L.rng.uni();
                    float u1 = L.rng.uni();
                    bool zero = !have_env && !have_sun;
                    V3 lpos{0.0f, 0.0f, 0.0f};
                    RGB emittance{0.0f, 0.0f, 0.0f};
                    float area_pd = 0.0f;
                    if (!zero) {
                        if (use_sun) {
                            lpos = V3{p.sun.direction[0], p.sun.direction[1], p.sun.direction[2]};
                            emittance = RGB{p.sun.irradiance[0], p.sun.irradiance[1], p.sun.irradiance[2]};
                        } else {
                            V3 ldir;
                            float ldir_pd;
                            RGB radiance = env_sample(p.env, u0, u1, ldir, ldir_pd);
                            emittance = scale(radiance, RTC8X_PI_F);
                            lpos = ldir;
                            area_pd = ldir_pd;
                            if (area_pd == 0.0f) zero = true;
                            area_pd *= prob_cur;
                        }
                    }
                    V3 sdir{-lpos.x, -lpos.y, -lpos.z};
                    float lp_cos = 0.0f;
                    if (!zero) {
                        V3 lnormal{-lpos.x, -lpos.y, -lpos.z};
                        lp_cos = dot(sdir, lnormal);
                        if (lp_cos <= 0.0f) zero = true;
                    }
                    if (zero) {
                        L.contribution = add(L.contribution, mul(L.tp, RGB{0.0f, 0.0f, 0.0f}));
                        after_nee = true;
                    } else {
                        const Frame fr = make_frame(V3{-L.dir.x, -L.dir.y, -L.dir.z});
                        V3 vin = to_local(fr, V3{-sdir.x, -sdir.y, -sdir.z});
                        float pval = phase_value(p.phase, p.scattering_forwardness, -vin.z);
                        float alb = p.scattering_albedo;
                        RGB fs{alb * pval, alb * pval, alb * pval};
                        L.nee_sun = use_sun;
                        if (use_sun) {
                            L.nee_f = mul(fs, emittance);
                            L.nee_k1 = prob_cur;
                        } else {
                            float bsdf_pd = pval * lp_cos / 1.0f;
                            if (!isfinite(bsdf_pd)) bsdf_pd = 0.0f;
                            float light_pd = area_pd;
                            L.nee_k1 = light_pd * light_pd / (bsdf_pd * bsdf_pd + light_pd * light_pd);
                            L.nee_f = mul(fs, scale(emittance, 1.0f / RTC8X_PI_F));
                            L.nee_k2 = lp_cos * 1.0f / 1.0f;
                            L.nee_k3 = area_pd;
                        }
                        const float trace_len = 2.0f * p.world_radius;
                        V3 lpoint{L.pos.x + trace_len * -sdir.x, L.pos.y + trace_len * -sdir.y,
                                  L.pos.z + trace_len * -sdir.z};
                        L.vis = 1.0f;
                        if (sg.init(p.grid, med, lpoint, sdir, trace_len)) {
                            if constexpr (TRACKING == RTC8X_TRACKING_LOCAL_MAJORANT) sg.first_tau(L.rng);
                            L.mode = kModeShadow;
                            continue;  // ratio tracking
                        }
                        L.mode = kModeShadow + 32;
                        continue;  // unoccluded: finish the estimate at the top of the loop
                    }
                }
            }
            if (after_nee) {
                if (L.path_length >= max_len) {
                    path_end = true;
                } else {
                    // phase sampling (optix_pathtracing_kernels.cu:341-356)
                    float us0 = L.rng.uni();
                    float us1 = L.rng.uni();
                    V3 dl;
                    float pdf = phase_sample(p.phase, p.scattering_forwardness, us0, us1, dl);
                    if (pdf == 0.0f) {
                        path_end = true;
                    } else {
                        const Frame fr = make_frame(V3{-L.dir.x, -L.dir.y, -L.dir.z});
                        const float alb = p.scattering_albedo;
                        RGB fsv{alb * pdf, alb * pdf, alb * pdf};
                        L.dir = from_local(fr, dl);
                        L.tp = mul(L.tp, scale(fsv, 1.0f / pdf));
                        L.org = L.pos;
                        L.dir_pd = pdf;
                        ++L.path_length;
                        bool escaped;
                        start_track<TRACKING>(p, med, L, sg, escaped);
                        if (escaped) L.mode = kModeTrack + 32;
                        continue;
                    }
                }
            }
            if (path_end) {
                paths++;
                vertices += L.path_length;
                if (all_nonneg_finite(L.contribution)) {
                    double* a = p.accum + 3 * (size_t)L.pix;
                    a[0] += (double)L.contribution.r;
                    a[1] += (double)L.contribution.g;
                    a[2] += (double)L.contribution.b;
                }
                if (p.path_vertices) p.path_vertices[L.pix] += L.path_length;
                if (++L.sample < num_samples) {
                    L.mode = kModeTrack + 16;
                } else {
                    p.rng_states[L.pix] = L.rng.state;
                    L.mode = kModeWork;
                }
            }
        }
        if (!__any(want)) break;
        // ---- one density lookup per lane that wants one
        if (want) {
            const float sigma = sample_density(p.grid, med, sg.o, sg.d, sg.t);
            if (L.mode == kModeTrack) {
                if (sg.after_track(L.rng, sigma, mu)) L.mode = kModeTrack + 48;
            } else {
                sg.after_shadow(L.rng, sigma, mu, L.vis);
            }
        }
    }
    if (p.stats) {
        uint64_t v[4] = {paths, vertices, med.lookups, scatters};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint64_t x = v[i];
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
            v[i] = x;
        }
        if (lane_id == 0) {
            atomicAdd((unsigned long long*)&p.stats[RTC8X_STAT_PATHS], (unsigned long long)v[0]);
            atomicAdd((unsigned long long*)&p.stats[RTC8X_STAT_VERTICES], (unsigned long long)v[1]);
            atomicAdd((unsigned long long*)&p.stats[RTC8X_STAT_LOOKUPS], (unsigned long long)v[2]);
            atomicAdd((unsigned long long*)&p.stats[RTC8X_STAT_SCATTER], (unsigned long long)v[3]);
        }
    }
}

template <int TRACKING>
hipError_t launch_wavefront(const rtc8x_frame_params& p, uint32_t num_samples, hipStream_t stream)
{
    static std::atomic<uint32_t> next_slot{0};
    const uint32_t slot = next_slot.fetch_add(1) % kCounterSlots;
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, wavefront_kernel<TRACKING>, 256, 0);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, stream, slot);
    hipLaunchKernelGGL(wavefront_kernel<TRACKING>, dim3((uint32_t)(cus * per_cu)), dim3(256), 0, stream, p, num_samples,
                       slot);
    return hipGetLastError();
}

template hipError_t launch_wavefront<RTC8X_TRACKING_GLOBAL_MAJORANT>(const rtc8x_frame_params&, uint32_t, hipStream_t);
template hipError_t launch_wavefront<RTC8X_TRACKING_LOCAL_MAJORANT>(const rtc8x_frame_params&, uint32_t, hipStream_t);

}  // namespace dev
}  // namespace rtc8x

// rtc8x_trace.hip -- the path-tracing launch (replaces the three OptiX ray-generation programs,
// renderer/gpu_kernels/optix_pathtracing_kernels.cu:31-448, for volume-only scenes).
//
// One lane traces one pixel's paths.  A 256-lane workgroup covers a 16x16 pixel tile (a 64-lane
// wavefront is a 16x4 strip), and workgroup ids are remapped so that the tiles one XCD receives are
// contiguous in the image (the dispatcher deals workgroups round-robin over the 8 XCDs, each with
// its own L2): neighbouring tiles then share that XCD's L2 for the grid's upper/lower tables and
// the leaves near the camera.
#include <hip/hip_runtime.h>

#include <cstring>

#include "rtc8x_common.hpp"
#include "rtc8x_device.hpp"

namespace rtc8x {
namespace dev {

constexpr uint32_t kHardMaxPathLength = 4096u;
constexpr uint32_t kTileW = 16, kTileH = 16;

struct PathCtx {
    Medium med;
    float prob_env;
    uint32_t scatters;
};

// performNextEventEstimation (renderer_kernel_common.h:297-403), environment at infinity + delta sun.
template <int TRACKING>
__device__ __forceinline__ RGB nee(const rtc8x_frame_params& p, PathCtx& c, Pcg& rng, V3 pos, const Frame& fr)
{
    const RGB zero{0.0f, 0.0f, 0.0f};
    float u_light = rng.uni();
    const bool have_env = p.env.kind != RTC8X_ENV_NONE, have_sun = p.sun.enabled != 0;
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
    float u0 = rng.uni();
    float u1 = rng.uni();
    if (!have_env && !have_sun) return zero;

    V3 lpos;
    RGB emittance;
    float area_pd = 0.0f;
    if (use_sun) {
        lpos = V3{p.sun.direction[0], p.sun.direction[1], p.sun.direction[2]};
        emittance = RGB{p.sun.irradiance[0], p.sun.irradiance[1], p.sun.irradiance[2]};
    } else {
        V3 dir;
        float dir_pd;
        RGB radiance = env_sample(p.env, u0, u1, dir, dir_pd);
        emittance = scale(radiance, RTC8X_PI_F);
        lpos = dir;
        area_pd = dir_pd;
        if (area_pd == 0.0f) return zero;
        area_pd *= prob_cur;
    }
    V3 lnormal{-lpos.x, -lpos.y, -lpos.z};
    V3 sdir{-lpos.x, -lpos.y, -lpos.z};
    float lp_cos = dot(sdir, lnormal);
    if (lp_cos <= 0.0f) return zero;
    float trace_len = 2.0f * p.world_radius;
    V3 lpoint{pos.x + trace_len * -sdir.x, pos.y + trace_len * -sdir.y, pos.z + trace_len * -sdir.z};
    float vis = transmittance<TRACKING>(p.grid, c.med, rng, lpoint, sdir, trace_len);
    if (!(vis > 0.0f)) return zero;
    V3 vin = to_local(fr, V3{-sdir.x, -sdir.y, -sdir.z});
    float pval = phase_value(p.phase, p.scattering_forwardness, -vin.z);
    float alb = p.scattering_albedo;
    RGB fs{alb * pval, alb * pval, alb * pval};
    if (use_sun) return scale(mul(fs, emittance), vis / prob_cur);
    float bsdf_pd = pval * lp_cos / 1.0f;
    if (!isfinite(bsdf_pd)) bsdf_pd = 0.0f;
    float light_pd = area_pd;
    float mis = light_pd * light_pd / (bsdf_pd * bsdf_pd + light_pd * light_pd);
    RGB le = scale(emittance, 1.0f / RTC8X_PI_F);
    float g_term = lp_cos * 1.0f / 1.0f;
    return scale(mul(fs, le), vis * mis * g_term / area_pd);
}

// RadianceQuery of a medium vertex (optix_pathtracing_kernels.cu:14-29)
__device__ __forceinline__ void make_query(const rtc8x_frame_params& p, V3 pos, V3 vout, rtc8x_radiance_query* q)
{
    float pp[3] = {pos.x, pos.y, pos.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float num = pp[a] - p.world_aabb_min[a];
        float den = p.world_aabb_max[a] - p.world_aabb_min[a];
        q->position[a] = den != 0.0f ? num / den : 0.0f;
    }
    const float two_pi = 2.0f * RTC8X_PI_F;
    float y = vout.y < -1.0f ? -1.0f : (vout.y > 1.0f ? 1.0f : vout.y);
    float th = rtc8x_acosf(y);
    float ph = rtc8x_fmodf(rtc8x_atan2f(-vout.x, vout.z) + two_pi, two_pi);
    q->normal_phi = th;
    q->normal_theta = ph;
    q->vout_phi = th;
    q->vout_theta = ph;
    q->roughness = 0.0f;
    q->diffuse_reflectance[0] = p.scattering_albedo;
    q->diffuse_reflectance[1] = p.scattering_albedo;
    q->diffuse_reflectance[2] = p.scattering_albedo;
    q->specular_reflectance[0] = q->specular_reflectance[1] = q->specular_reflectance[2] = 0.0f;
}

// One path of pathTrace_generic<withNRC, genData> (optix_pathtracing_kernels.cu:31-434).
// GEN_DATA paths take training-record indices [train_base, train_base + n) (pixel order); with
// COUNT_ONLY nothing is written and the number of record attempts is returned.
template <int ENTRY, int TRACKING, bool COUNT_ONLY>
__device__ uint32_t trace_path(const rtc8x_frame_params& p, PathCtx& c, uint32_t lx, uint32_t ly, uint32_t launch_w,
                               uint32_t launch_h, Pcg& rng, uint32_t train_base, RGB& contribution_out,
                               uint32_t& path_length_out)
{
    constexpr bool with_nrc = ENTRY != RTC8X_ENTRY_PATH_TRACE;
    constexpr bool gen_data = ENTRY == RTC8X_ENTRY_GENERATE_TRAINING_DATA;
    const rtc8x_nrc_buffers& nb = p.nrc;
    const uint32_t linear = ly * launch_w + lx;

    // camera ray (optix_pathtracing_kernels.cu:43-55)
    float px = ((float)lx + rng.uni()) / (float)launch_w;
    float py = ((float)ly + rng.uni()) / (float)launch_h;
    const float vh = p.camera.vh, vw = p.camera.vw;
    const float sensor_area = vw * vh;
    V3 org{p.camera.position[0], p.camera.position[1], p.camera.position[2]};
    V3 ldir = normalize(V3{vw * (-0.5f + px), vh * (0.5f - py), -1.0f});
    const float* M = p.camera.orientation;
    V3 dir{dot(V3{M[0], M[1], M[2]}, ldir), dot(V3{M[3], M[4], M[5]}, ldir), dot(V3{M[6], M[7], M[8]}, ldir)};
    dir = normalize(dir);
    float az = ldir.z < 0.0f ? -ldir.z : ldir.z;
    float dir_pd = 1.0f / (az * (az * az) * sensor_area);

    bool unbiased = false;
    uint32_t prev_idx = RTC8X_INVALID_VERTEX_INDEX;
    RGB prev_local_tp{0.0f, 0.0f, 0.0f};
    uint32_t records = 0;
    if (gen_data) {
        uint32_t tl = (ly % 4u) * 4u + (lx % 4u);
        unbiased = ((tl + nb.offset_to_select_unbiased_path) % 16u) == 0;
    }

    RGB contribution{0.0f, 0.0f, 0.0f};
    float rinv = 1.0f / dir_pd;
    RGB tp{rinv, rinv, rinv};
    const float init_importance = luminance(tp);
    uint32_t path_length = 0;
    float primary_spread = 0.0f, cur_sqrt_spread = 0.0f;
    uint32_t rr_after = p.rr_after_path_length;
    if (gen_data && rr_after < 2) rr_after = 2;
    uint32_t max_len = (ENTRY == RTC8X_ENTRY_PATH_TRACE && p.max_path_length) ? p.max_path_length : kHardMaxPathLength;
    if (max_len > kHardMaxPathLength) max_len = kHardMaxPathLength;

    for (;;) {
        ++path_length;
        float t_hit = 1e10f;
        bool vol = free_flight<TRACKING>(p.grid, c.med, rng, org, dir, 1e10f, t_hit);
        if (!vol) {
            // escaped to the environment (optix_pathtracing_kernels.cu:130-174)
            if (p.env.kind != RTC8X_ENV_NONE) {
                float hyp_pd;
                RGB radiance = env_evaluate(p.env, dir, hyp_pd);
                float mis = 1.0f;
                if (path_length > 1) {
                    float light_pd = c.prob_env * hyp_pd;
                    float bsdf_pd = dir_pd;
                    mis = bsdf_pd * bsdf_pd / (bsdf_pd * bsdf_pd + light_pd * light_pd);
                }
                RGB implicit_cont = scale(radiance, mis);
                if (!gen_data) contribution = add(contribution, mul(tp, implicit_cont));
                if (gen_data && !COUNT_ONLY && prev_idx != RTC8X_INVALID_VERTEX_INDEX) {
                    float* tgt = nb.train_targets[0] + 3 * (size_t)prev_idx;
                    RGB a = mul(prev_local_tp, implicit_cont);
                    tgt[0] = tgt[0] + a.r;
                    tgt[1] = tgt[1] + a.g;
                    tgt[2] = tgt[2] + a.b;
                }
            }
            if (!COUNT_ONLY) {
                if (gen_data) {
                    rtc8x_training_suffix_terminal ti;
                    ti.prev_vertex_index = prev_idx;
                    ti.has_query = 0;
                    ti.path_length = path_length & 0xFFu;
                    ti.is_unbiased = unbiased ? 1u : 0u;
                    nb.train_suffix_terminals[linear] = ti;
                } else if (with_nrc) {
                    rtc8x_terminal_info ti;
                    ti.throughput[0] = tp.r;
                    ti.throughput[1] = tp.g;
                    ti.throughput[2] = tp.b;
                    ti.has_query = (path_length & 0xFFu) << 8;
                    nb.inference_terminals[linear] = ti;
                }
            }
            break;
        }

        // medium interaction (optix_pathtracing_kernels.cu:176-182)
        c.scatters++;
        V3 pos{org.x + t_hit * dir.x, org.y + t_hit * dir.y, org.z + t_hit * dir.z};
        V3 vout{-dir.x, -dir.y, -dir.z};
        Frame fr = make_frame(vout);

        if (with_nrc) {
            // path spread (optix_pathtracing_kernels.cu:200-210)
            float ddx = pos.x - org.x, ddy = pos.y - org.y, ddz = pos.z - org.z;
            float dist2 = ddx * ddx + ddy * ddy + ddz * ddz;
            if (path_length == 1)
                primary_spread = dist2 / (4.0f * RTC8X_PI_F * 1.0f);
            else
                cur_sqrt_spread += sqrtf(dist2 / (dir_pd * 1.0f));
        }

        // Russian roulette (optix_pathtracing_kernels.cu:254-268)
        float continue_prob = 1.0f;
        bool rr_terminated = false;
        if (path_length > rr_after) {
            float q = luminance(tp) / init_importance;
            continue_prob = fminf(q, 1.0f);
            if (rng.uni() >= continue_prob) rr_terminated = true;
        }
        if (!with_nrc && rr_terminated) break;

        if (with_nrc) {
            bool ends_with_cache = rr_terminated;
            const float pt = nb.path_termination_factor;
            if (gen_data) {
                if (!unbiased) ends_with_cache |= cur_sqrt_spread * cur_sqrt_spread > 2.0f * pt * primary_spread;
            } else {
                ends_with_cache |= cur_sqrt_spread * cur_sqrt_spread > pt * primary_spread;
            }
            if (ends_with_cache) {
                if (!COUNT_ONLY) {
                    make_query(p, pos, vout, &nb.inference_queries[linear]);
                    if (gen_data) {
                        rtc8x_training_suffix_terminal ti;
                        ti.prev_vertex_index = prev_idx;
                        ti.has_query = 1;
                        ti.path_length = path_length & 0xFFu;
                        ti.is_unbiased = unbiased ? 1u : 0u;
                        nb.train_suffix_terminals[linear] = ti;
                    } else {
                        rtc8x_terminal_info ti;
                        ti.throughput[0] = tp.r;
                        ti.throughput[1] = tp.g;
                        ti.throughput[2] = tp.b;
                        ti.has_query = 1u | ((path_length & 0xFFu) << 8);
                        nb.inference_terminals[linear] = ti;
                    }
                }
                break;
            }
        }

        tp = scale(tp, 1.0f / continue_prob);
        RGB nee_c = nee<TRACKING>(p, c, rng, pos, fr);
        if (!gen_data) contribution = add(contribution, mul(tp, nee_c));
        if (path_length >= max_len) break;

        // phase sampling (optix_pathtracing_kernels.cu:341-356)
        float us0 = rng.uni();
        float us1 = rng.uni();
        V3 dl;
        float pdf = phase_sample(p.phase, p.scattering_forwardness, us0, us1, dl);
        if (pdf == 0.0f) break;
        const float alb = p.scattering_albedo;
        RGB fsv{alb * pdf, alb * pdf, alb * pdf};
        dir = from_local(fr, dl);
        RGB local_tp = scale(fsv, 1.0f / pdf);
        tp = mul(tp, local_tp);
        org = pos;
        dir_pd = pdf;

        if (gen_data) {
            // training vertex record (optix_pathtracing_kernels.cu:358-410)
            uint32_t idx = train_base + records;
            records++;
            if (COUNT_ONLY) continue;
            rtc8x_radiance_query q;
            make_query(p, pos, vout, &q);
            if (idx < RTC8X_TRAIN_BUFFER_SIZE) {
                nb.train_queries[0][idx] = q;
                rtc8x_training_vertex_info vi;
                vi.local_throughput[0] = local_tp.r;
                vi.local_throughput[1] = local_tp.g;
                vi.local_throughput[2] = local_tp.b;
                vi.prev_vertex_index = prev_idx;
                vi.path_length = path_length & 0xFFu;
                vi.is_unbiased = unbiased ? 1u : 0u;
                nb.train_vertex_info[idx] = vi;
                float* tgt = nb.train_targets[0] + 3 * (size_t)idx;
                tgt[0] = nee_c.r;
                tgt[1] = nee_c.g;
                tgt[2] = nee_c.b;
                prev_local_tp = local_tp;
                prev_idx = idx;
            } else {
                nb.inference_queries[linear] = q;
                rtc8x_training_suffix_terminal ti;
                ti.prev_vertex_index = prev_idx;
                ti.has_query = 1;
                ti.path_length = path_length & 0xFFu;
                ti.is_unbiased = unbiased ? 1u : 0u;
                nb.train_suffix_terminals[linear] = ti;
                break;
            }
        }
    }
    contribution_out = contribution;
    path_length_out = path_length;
    return records;
}

// Maps a workgroup id to a tile so that each XCD's share of the grid is one contiguous run of tiles.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t bid, uint32_t nblocks)
{
    const uint32_t nx = 8;
    uint32_t q = nblocks / nx, r = nblocks % nx;
    uint32_t x = bid % nx, k = bid / nx;
    return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ void add_stats(uint64_t* stats, uint32_t paths, uint32_t vertices, uint32_t lookups,
                                          uint32_t scatters)
{
    // wavefront reduction, then one atomic per wavefront per counter
    uint64_t v[4] = {paths, vertices, lookups, scatters};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        uint64_t x = v[i];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        v[i] = x;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd((unsigned long long*)&stats[RTC8X_STAT_PATHS], (unsigned long long)v[0]);
        atomicAdd((unsigned long long*)&stats[RTC8X_STAT_VERTICES], (unsigned long long)v[1]);
        atomicAdd((unsigned long long*)&stats[RTC8X_STAT_LOOKUPS], (unsigned long long)v[2]);
        atomicAdd((unsigned long long*)&stats[RTC8X_STAT_SCATTER], (unsigned long long)v[3]);
    }
}

template <int ENTRY, int TRACKING, bool COUNT_ONLY>
__global__ __launch_bounds__(256) void trace_kernel(const rtc8x_frame_params p, uint32_t launch_w, uint32_t launch_h,
                                                    uint32_t num_samples, const uint32_t* __restrict__ record_offsets,
                                                    uint32_t* __restrict__ record_counts)
{
    const uint32_t tiles_x = (launch_w + kTileW - 1) / kTileW;
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    const uint32_t lx = tx * kTileW + (threadIdx.x % kTileW);
    const uint32_t ly = ty * kTileH + (threadIdx.x / kTileW);
    const bool active = lx < launch_w && ly < launch_h;

    PathCtx c;
    c.med.coeff = p.density_coeff;
    c.med.majorant = p.density_coeff * p.grid.max_value;
    c.med.lookups = 0;
    c.scatters = 0;
    c.prob_env = (p.sun.enabled && p.env.kind != RTC8X_ENV_NONE) ? 1.0f - p.sun.select_prob : 1.0f;
    uint32_t paths = 0, vertices = 0;

    if (active) {
        uint64_t* rs = &p.rng_states[(size_t)ly * p.image_width + lx];
        Pcg rng{*rs};
        const uint32_t lin = ly * launch_w + lx;
        const uint32_t base = record_offsets ? record_offsets[lin] : 0u;
        double acc[3] = {0.0, 0.0, 0.0};
        if (ENTRY == RTC8X_ENTRY_PATH_TRACE && !COUNT_ONLY) {
            const double* a = p.accum + 3 * (size_t)lin;
            acc[0] = a[0];
            acc[1] = a[1];
            acc[2] = a[2];
        }
        uint32_t vsum = 0;
        for (uint32_t s = 0; s < num_samples; ++s) {
            RGB contrib;
            uint32_t plen;
            uint32_t rec = trace_path<ENTRY, TRACKING, COUNT_ONLY>(p, c, lx, ly, launch_w, launch_h, rng, base, contrib, plen);
            paths++;
            vertices += plen;
            vsum += plen;
            if (COUNT_ONLY) {
                record_counts[lin] = rec;
                continue;
            }
            if (ENTRY == RTC8X_ENTRY_GENERATE_TRAINING_DATA) continue;
            bool ok = all_nonneg_finite(contrib);
            if (ENTRY == RTC8X_ENTRY_PATH_TRACE) {
                if (ok) {
                    acc[0] += (double)contrib.r;
                    acc[1] += (double)contrib.g;
                    acc[2] += (double)contrib.b;
                }
            } else {
                float* pf = p.nrc.per_frame_contribution + 3 * (size_t)lin;
                pf[0] = ok ? contrib.r : 0.0f;
                pf[1] = ok ? contrib.g : 0.0f;
                pf[2] = ok ? contrib.b : 0.0f;
            }
        }
        if (!COUNT_ONLY) {
            *rs = rng.state;
            if (ENTRY == RTC8X_ENTRY_PATH_TRACE) {
                double* a = p.accum + 3 * (size_t)lin;
                a[0] = acc[0];
                a[1] = acc[1];
                a[2] = acc[2];
            }
            if (p.path_vertices) p.path_vertices[lin] += vsum;
        }
    }
    if (p.stats && !COUNT_ONLY) add_stats(p.stats, paths, vertices, c.med.lookups, c.scatters);
}

// Exclusive prefix sum of the per-path record counts (one workgroup, a few hundred thousand entries).
__global__ __launch_bounds__(1024) void record_prefix_kernel(const uint32_t* __restrict__ counts,
                                                             uint32_t* __restrict__ offsets, uint32_t n,
                                                             uint32_t* __restrict__ total_out)
{
    __shared__ uint64_t part[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t chunk = (n + 1023) / 1024;
    const uint32_t b = tid * chunk, e = min(n, b + chunk);
    uint64_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += counts[i];
    part[tid] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint64_t v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint64_t run = part[tid] - s;
    for (uint32_t i = b; i < e; ++i) {
        offsets[i] = run > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)run;
        run += counts[i];
    }
    if (tid == 1023) *total_out = part[1023] > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)part[1023];
}

template <int ENTRY, int TRACKING, bool COUNT_ONLY>
hipError_t launch(const rtc8x_frame_params& p, uint32_t w, uint32_t h, uint32_t num_samples, const uint32_t* offsets,
                  uint32_t* counts, hipStream_t stream)
{
    const uint32_t tiles = ((w + kTileW - 1) / kTileW) * ((h + kTileH - 1) / kTileH);
    hipLaunchKernelGGL((trace_kernel<ENTRY, TRACKING, COUNT_ONLY>), dim3(tiles), dim3(kTileW * kTileH), 0, stream, p,
                       w, h, num_samples, offsets, counts);
    return hipGetLastError();
}

template <int TRACKING>
hipError_t launch_wavefront(const rtc8x_frame_params& p, uint32_t num_samples, hipStream_t stream);

template <int ENTRY, bool COUNT_ONLY>
hipError_t launch_tracking(const rtc8x_frame_params& p, uint32_t w, uint32_t h, uint32_t ns, const uint32_t* offsets,
                           uint32_t* counts, hipStream_t stream)
{
    if (p.tracking == RTC8X_TRACKING_LOCAL_MAJORANT)
        return launch<ENTRY, RTC8X_TRACKING_LOCAL_MAJORANT, COUNT_ONLY>(p, w, h, ns, offsets, counts, stream);
    return launch<ENTRY, RTC8X_TRACKING_GLOBAL_MAJORANT, COUNT_ONLY>(p, w, h, ns, offsets, counts, stream);
}

}  // namespace dev
}  // namespace rtc8x

using namespace rtc8x;

static bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

extern "C" rtc8x_status rtc8x_trace(void* stream, const rtc8x_frame_params* params, int32_t entry, uint32_t num_samples)
{
    RTC8X_REQUIRE(params != nullptr, "rtc8x_trace: params is NULL");
    const rtc8x_frame_params& p = *params;
    RTC8X_REQUIRE(entry >= RTC8X_ENTRY_PATH_TRACE && entry <= RTC8X_ENTRY_PATH_TRACE_WITH_NRC,
                  "rtc8x_trace: unknown entry %d", entry);
    RTC8X_REQUIRE(p.image_width > 0 && p.image_height > 0 && p.image_width <= 16384 && p.image_height <= 16384,
                  "rtc8x_trace: bad image size %ux%u", p.image_width, p.image_height);
    RTC8X_REQUIRE(p.rng_states != nullptr, "rtc8x_trace: rng_states is NULL");
    RTC8X_REQUIRE(num_samples >= 1, "rtc8x_trace: num_samples must be >= 1");
    RTC8X_REQUIRE(entry == RTC8X_ENTRY_PATH_TRACE || num_samples == 1,
                  "rtc8x_trace: the NRC entry points trace one sample per launch");
    RTC8X_REQUIRE(entry != RTC8X_ENTRY_PATH_TRACE || p.accum != nullptr, "rtc8x_trace: accum is NULL");
    RTC8X_REQUIRE(finite3(p.camera.position) && std::isfinite(p.camera.vh) && std::isfinite(p.camera.vw),
                  "rtc8x_trace: camera is not finite");
    RTC8X_REQUIRE(p.phase == RTC8X_PHASE_SCHLICK || p.phase == RTC8X_PHASE_HENYEY_GREENSTEIN,
                  "rtc8x_trace: unknown phase function %d", p.phase);
    RTC8X_REQUIRE(std::fabs(p.scattering_forwardness) < 1.0f, "rtc8x_trace: |forwardness| must be < 1");
    RTC8X_REQUIRE(p.tracking == RTC8X_TRACKING_GLOBAL_MAJORANT || p.tracking == RTC8X_TRACKING_LOCAL_MAJORANT,
                  "rtc8x_trace: unknown tracking mode %d", p.tracking);
    RTC8X_REQUIRE(std::isfinite(p.world_radius) && p.world_radius >= 0.0f, "rtc8x_trace: bad world radius");
    const rtc8x_grid& g = p.grid;
    const float majorant = p.density_coeff * g.max_value;
    if (majorant > 0.0f) {
        RTC8X_REQUIRE(std::isfinite(majorant), "rtc8x_trace: majorant is not finite");
        RTC8X_REQUIRE(g.num_root_tiles > 0 && g.root_tiles && g.upper_tables && g.tile_values,
                      "rtc8x_trace: grid tables missing");
        RTC8X_REQUIRE(finite3(g.world_bbox_min) && finite3(g.world_bbox_max), "rtc8x_trace: grid bbox not finite");
        // Free-flight loops take about majorant * chord steps; refuse inputs whose loops cannot end in
        // reasonable time (the chord is bounded by the bbox diagonal).
        double dx = g.world_bbox_max[0] - g.world_bbox_min[0], dy = g.world_bbox_max[1] - g.world_bbox_min[1],
               dz = g.world_bbox_max[2] - g.world_bbox_min[2];
        double diag = std::sqrt(dx * dx + dy * dy + dz * dz);
        RTC8X_REQUIRE((double)majorant * diag < 1.0e6, "rtc8x_trace: majorant x bbox diagonal = %g is too large",
                      (double)majorant * diag);
        if (p.tracking == RTC8X_TRACKING_LOCAL_MAJORANT)
            RTC8X_REQUIRE(g.cell_majorants && g.cell_dims[0] > 0 && g.cell_dims[1] > 0 && g.cell_dims[2] > 0,
                          "rtc8x_trace: local-majorant tracking needs cell majorants");
    }
    if (p.env.kind == RTC8X_ENV_LATLONG) {
        RTC8X_REQUIRE(p.env.texels && p.env.width > 0 && p.env.height > 0, "rtc8x_trace: env texture missing");
        if (p.env.sampler == RTC8X_ENV_SAMPLER_ALIAS)
            RTC8X_REQUIRE(p.env.alias_rows && p.env.alias_cols && p.env.alias_weights && p.env.alias_integral > 0.0f,
                          "rtc8x_trace: env alias tables missing");
        else
            RTC8X_REQUIRE(p.env.mips && p.env.num_mip_levels >= 1 && p.env.num_mip_levels <= 31 &&
                              p.env.map_integral > 0.0f,
                          "rtc8x_trace: env importance map missing");
    }
    if (p.sun.enabled) {
        RTC8X_REQUIRE(finite3(p.sun.direction) && finite3(p.sun.irradiance), "rtc8x_trace: sun not finite");
        if (p.env.kind != RTC8X_ENV_NONE)
            RTC8X_REQUIRE(p.sun.select_prob > 0.0f && p.sun.select_prob < 1.0f,
                          "rtc8x_trace: sun.select_prob must be in (0,1) when an environment is present");
    }
    hipStream_t s = (hipStream_t)stream;
    if (entry == RTC8X_ENTRY_PATH_TRACE) {
        if (p.tracking == RTC8X_TRACKING_LOCAL_MAJORANT)
            RTC8X_HIP(dev::launch_wavefront<RTC8X_TRACKING_LOCAL_MAJORANT>(p, num_samples, s));
        else
            RTC8X_HIP(dev::launch_wavefront<RTC8X_TRACKING_GLOBAL_MAJORANT>(p, num_samples, s));
        return RTC8X_OK;
    }
    const rtc8x_nrc_buffers& nb = p.nrc;
    RTC8X_REQUIRE(nb.inference_queries && nb.max_num_training_suffixes >= p.image_width * p.image_height,
                  "rtc8x_trace: NRC buffers missing");
    if (entry == RTC8X_ENTRY_PATH_TRACE_WITH_NRC) {
        RTC8X_REQUIRE(nb.inference_terminals && nb.per_frame_contribution, "rtc8x_trace: NRC buffers missing");
        RTC8X_HIP((dev::launch_tracking<RTC8X_ENTRY_PATH_TRACE_WITH_NRC, false>(p, p.image_width, p.image_height, 1,
                                                                                 nullptr, nullptr, s)));
        return RTC8X_OK;
    }
    // GENERATE_TRAINING_DATA: the launch size lives on the device (prepareNRC writes it); read it.
    RTC8X_REQUIRE(nb.train_image_size && nb.num_training_data && nb.train_queries[0] && nb.train_targets[0] &&
                      nb.train_vertex_info && nb.train_suffix_terminals && nb.scratch,
                  "rtc8x_trace: NRC training buffers missing");
    RTC8X_REQUIRE(nb.nrc_buffer_index < 2, "rtc8x_trace: nrc_buffer_index must be 0 or 1");
    int32_t ts[2];
    RTC8X_HIP(hipMemcpyAsync(ts, nb.train_image_size + 2 * nb.nrc_buffer_index, sizeof(ts), hipMemcpyDeviceToHost, s));
    RTC8X_HIP(hipStreamSynchronize(s));
    RTC8X_REQUIRE(ts[0] > 0 && ts[1] > 0 && (uint32_t)ts[0] <= p.image_width && (uint32_t)ts[1] <= p.image_height,
                  "rtc8x_trace: training image size %dx%d out of range", ts[0], ts[1]);
    const uint32_t tw = (uint32_t)ts[0], th = (uint32_t)ts[1], n = tw * th;
    uint32_t* counts = nb.scratch;
    uint32_t* offsets = nb.scratch + p.image_width * p.image_height;
    RTC8X_HIP((dev::launch_tracking<RTC8X_ENTRY_GENERATE_TRAINING_DATA, true>(p, tw, th, 1, nullptr, counts, s)));
    hipLaunchKernelGGL(dev::record_prefix_kernel, dim3(1), dim3(1024), 0, s, counts, offsets, n,
                       nb.num_training_data + nb.nrc_buffer_index);
    RTC8X_HIP(hipGetLastError());
    RTC8X_HIP((dev::launch_tracking<RTC8X_ENTRY_GENERATE_TRAINING_DATA, false>(p, tw, th, 1, offsets, nullptr, s)));
    return RTC8X_OK;
}

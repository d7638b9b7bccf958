// rtc8x_measure.hip -- measurement kernels with no counterpart in the renderer: a plain copy (the
// calibration of the HBM byte counters), dependent density lookups at random positions of the grid's
// leaves, and the tracer's own stepping pattern with its shading taken out (the memory-latency roofline
// of the tracer's lookups).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rtc8x_common.hpp"
#include "rtc8x_device.hpp"

namespace rtc8x {

__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__device__ __forceinline__ uint32_t mix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

// Each lane performs `steps` steps of W density lookups.  A lookup is the tracer's own: the trilinear
// stencil at a random position inside a random leaf, fetched in two rounds (leaf-map entries, then the
// voxels) by trilinear_batch, whose W lookups overlap.  The next step's positions depend on the values
// read, so steps are serially dependent.
template <uint32_t W>
__global__ void gather_probe_kernel(const rtc8x_grid g, uint32_t steps, uint32_t seed, float* __restrict__ out)
{
    extern __shared__ float occupancy_limiter[];
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t h = mix32(tid * 0x9e3779b9u ^ seed);
    float acc = 0.0f;
    for (uint32_t s = 0; s < steps; ++s) {
        dev::V3 idx[W];
#pragma unroll
        for (uint32_t k = 0; k < W; ++k) {
            h = mix32(h + 0x632be5abu * (k + 1));
            const uint32_t leaf = (uint32_t)(((uint64_t)h * g.num_leaves) >> 32);
            const uint32_t r = mix32(h ^ 0x5bd1e995u);
            const int32_t* o = g.leaf_origins + 3 * (size_t)leaf;
            idx[k] = dev::V3{(float)o[0] + (float)(r & 0xffu) * (8.0f / 256.0f),
                        (float)o[1] + (float)((r >> 8) & 0xffu) * (8.0f / 256.0f),
                        (float)o[2] + (float)((r >> 16) & 0xffu) * (8.0f / 256.0f)};
        }
        float v[W];
        dev::trilinear_batch<W>(g, idx, v);
        float sum = 0.0f;
#pragma unroll
        for (uint32_t k = 0; k < W; ++k) sum += v[k];
        acc += sum;
        h = mix32(h ^ __float_as_uint(sum) ^ s);
    }
    if (acc == -1.0f) occupancy_limiter[threadIdx.x] = acc;  // keeps the dynamic LDS allocation alive
    out[tid] = acc;
}

// The tracer's stepping pattern with its shading taken out, as a persistent launch like the tracer's:
// lanes take pixels in 8x8 tiles from a work counter (one atomic per wavefront, the tiles wrapping around
// the image) and trace every one of a pixel's `spp` paths before the next: a jittered camera ray
// (optix_pathtracing_kernels.cu:43-55), delta tracking at the global majorant's spacing, and at every real
// collision a ratio-tracked shadow segment towards `light` (the length the tracer gives it, 2 x the
// world radius) and a phase-function sample of the next direction, until the path escapes.  The global
// walk's tentative collisions are taken two at a time and their lookups issued together
// (density_batch<2>), as the tracer does; the local walk looks up one per round, its faster form here.
// One copy of the lookup code serves both segment kinds.  Every lane stops after `budget`
// lookups, so the launch has no tail; what it leaves out is the tracer's vertex work: NEE light sampling
// and MIS, the environment, Russian roulette, NRC records and accumulation.
// stats[0] += lookups, stats[1] += paths that ended.
__device__ uint32_t g_probe_counter;
__global__ void reset_probe_kernel() { g_probe_counter = 0u; }

// WAVES: the occupancy the registers are allotted for (3: 168 VGPRs, as the tracer's render entries; 4:
// 128), the launch's own occupancy being set by the LDS reservation.
template <int TRACKING, int WAVES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES)))
void march_probe_kernel(const rtc8x_frame_params p, dev::V3 light, uint32_t spp, uint32_t budget,
                                   unsigned long long* __restrict__ stats)
{
    extern __shared__ float occupancy_limiter[];
    const uint8_t* table = nullptr;
    if constexpr (TRACKING == RTC8X_TRACKING_LOCAL_MAJORANT) {
        // the region table in LDS, as the tracer stages it
        __shared__ __align__(16) uint8_t regions[RTC8X_MAX_REGIONS];
        const uint32_t n16 = ((uint32_t)p.grid.region_dims[0] * (uint32_t)p.grid.region_dims[1] *
                                  (uint32_t)p.grid.region_dims[2] + 15u) / 16u;
        const uint4* src = reinterpret_cast<const uint4*>(p.grid.region_majorants);
        for (uint32_t i = threadIdx.x; i < n16; i += blockDim.x) reinterpret_cast<uint4*>(regions)[i] = src[i];
        __syncthreads();
        table = regions;
    }
    const uint32_t W = p.image_width, H = p.image_height;
    const uint32_t tiles_x = (W + 7) / 8, tiles = tiles_x * ((H + 7) / 8);
    const uint32_t lane_id = threadIdx.x & 63;
    const float maj = p.density_coeff * p.grid.max_value;
    const float coeff = p.density_coeff;
    const float trace_len = 2.0f * p.world_radius;
    dev::Medium med;
    med.coeff = coeff;
    med.majorant = maj;
    med.lookups = 0;
    dev::Seg<RTC8X_TRACKING_LOCAL_MAJORANT> sg;
    dev::Pcg rng;
    rng.state = 0;
    dev::V3 o{0.0f, 0.0f, 0.0f}, d{0.0f, 0.0f, 1.0f};
    float t = 0.0f, t1 = -1.0f, vis = 1.0f;
    uint32_t mode = 3;  // 0: start a path, 1: tracking, 2: shadow, 3: needs a pixel
    uint32_t looked = 0, paths = 0, lx = 0, ly = 0, sample = 0;
    dev::V3 pos{0.0f, 0.0f, 0.0f}, dir{0.0f, 0.0f, 1.0f};
    // a segment from so along sd: the global walk keeps (o, d, t, t1), the local one its region walk
    auto start = [&](dev::V3 so, dev::V3 sd, float tmax) -> bool {
        if constexpr (TRACKING == RTC8X_TRACKING_LOCAL_MAJORANT) {
            if (!sg.init(p.grid, med, so, sd, tmax)) return false;
            sg.first_tau(rng);
            return true;
        } else {
            float t0;
            if (!dev::clip(p.grid, so, sd, tmax, t0, t1)) return false;
            o = so;
            d = sd;
            t = fmaxf(0.0f, t0);
            return true;
        }
    };
    for (uint32_t iter = 0; iter < 8u * budget + 256u && looked < budget; ++iter) {
        if (mode == 3) {
            // the next pixel, one atomic per wavefront
            const uint64_t ask = __ballot(1);
            const uint32_t leader = (uint32_t)__ffsll((long long)ask) - 1u;
            uint32_t base = 0;
            if (lane_id == leader) base = atomicAdd(&g_probe_counter, (uint32_t)__popcll(ask));
            base = __shfl(base, (int)leader, 64);
            const uint32_t w = base + (uint32_t)__popcll(ask & ((1ull << lane_id) - 1ull));
            const uint32_t tile = (w >> 6) % tiles, within = w & 63u;
            lx = (tile % tiles_x) * 8 + (within & 7u);
            ly = (tile / tiles_x) * 8 + (within >> 3);
            if (lx >= W || ly >= H) continue;
            rng.state = p.rng_states[(size_t)ly * W + lx] ^ ((uint64_t)(w / (64u * tiles)) << 48);
            sample = 0;
            mode = 0;
            continue;
        }
        if (mode == 0) {
            const float px = ((float)lx + rng.uni()) / (float)W, py = ((float)ly + rng.uni()) / (float)H;
            const dev::V3 ldir = dev::normalize(dev::V3{p.camera.vw * (-0.5f + px), p.camera.vh * (0.5f - py), -1.0f});
            const float* M = p.camera.orientation;
            dir = dev::normalize(dev::V3{dev::dot(dev::V3{M[0], M[1], M[2]}, ldir), dev::dot(dev::V3{M[3], M[4], M[5]}, ldir),
                                         dev::dot(dev::V3{M[6], M[7], M[8]}, ldir)});
            pos = dev::V3{p.camera.position[0], p.camera.position[1], p.camera.position[2]};
            if (start(pos, dir, 1e10f)) {
                mode = 1;
            } else {  // the camera ray misses the volume: the path ends at once
                ++paths;
                mode = ++sample < spp ? 0u : 3u;
            }
            continue;
        }
        uint32_t event = 0;  // 1: real collision at t, 2: segment end
        if constexpr (TRACKING == RTC8X_TRACKING_LOCAL_MAJORANT) {
            // the tracer's region walk: a bounded number of region visits, then the lookup of a
            // tentative collision if one was reached (one lookup per round: the speculative pair the
            // tracer takes measured slower here, 6.84 against 7.53 G lookups/s at 3 waves per SIMD)
            float mu = 0.0f;
            const int r = sg.advance(p.grid, med, rng, mu, table);
            if (r == dev::kAdvanceCollision) {
                const float sig = dev::sample_density(p.grid, med, sg.o, sg.d, sg.t);
                ++looked;
                if (mode == 1) {
                    if (sg.after_track(rng, sig, mu)) event = 1;
                } else {
                    sg.after_shadow(rng, sig, mu, vis);
                }
            } else if (r == dev::kAdvanceEnd) {
                event = 2;
            }
            o = sg.o;
            d = sg.d;
            t = sg.t;
        } else {
            // two tentative collisions (a collision past the segment's end is looked up at the current
            // position, as the tracer does); a tracking segment stops at the first real one
            float tk[2], tl[2], u[2];
            tk[0] = t - rtc8x_logf(1.0f - rng.uni()) / maj;
            tk[1] = tk[0] - rtc8x_logf(1.0f - rng.uni()) / maj;
            u[0] = rng.uni();
            u[1] = rng.uni();
            tl[0] = tk[0] > t1 ? t : tk[0];
            tl[1] = tk[1] > t1 ? t : tk[1];
            float sig[2];
            dev::density_batch<2>(p.grid, coeff, o, d, tl, sig);
            for (int k = 0; k < 2 && event == 0; ++k) {
                if (tk[k] > t1) {
                    event = 2;
                    break;
                }
                ++looked;
                t = tk[k];
                if (mode == 1) {
                    if (u[k] < sig[k] / maj) event = 1;
                } else {
                    vis *= 1.0f - sig[k] / maj;
                }
            }
        }
        if (event == 1) {  // real collision: the shadow segment, ending at the collision
            pos = dev::V3{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z};
            const dev::V3 lp{pos.x + trace_len * light.x, pos.y + trace_len * light.y, pos.z + trace_len * light.z};
            vis = 1.0f;
            mode = 2;
            if (!start(lp, dev::V3{-light.x, -light.y, -light.z}, trace_len)) event = 2;
        }
        if (event == 2) {
            bool ended = mode == 1;  // escaped
            if (!ended) {  // shadow segment done: phase sample and the next tracking segment
                dev::V3 dl;
                const float us0 = rng.uni(), us1 = rng.uni();
                dev::phase_sample(p.phase, p.scattering_forwardness, us0, us1, dl);
                const dev::Frame fr = dev::make_frame(dev::V3{-dir.x, -dir.y, -dir.z});
                dir = dev::from_local(fr, dl);
                if (start(pos, dir, 1e10f)) {
                    mode = 1;
                } else {
                    ended = true;
                }
            }
            if (ended) {
                ++paths;
                mode = ++sample < spp ? 0u : 3u;
            }
        }
    }
    if (vis == -1.0f) occupancy_limiter[threadIdx.x] = vis;  // keeps the dynamic LDS allocation alive
    unsigned long long a = looked, b = paths;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[0], a);
        atomicAdd(&stats[1], b);
    }
}

// The dynamic LDS per 256-lane workgroup that admits `waves_per_simd` workgroups per CU and not one more
// (a 256-lane workgroup puts one wave on each of the CU's 4 SIMDs).
static rtc8x_status occupancy_lds(uint32_t waves_per_simd, size_t* lds_out, const char* who, size_t static_lds = 0)
{
    int dev = 0, lds_cu = 0, lds_block = 0;
    RTC8X_HIP(hipGetDevice(&dev));
    RTC8X_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    RTC8X_HIP(hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    // the workgroup's whole LDS (the kernel's static arrays and the dynamic part) sets the occupancy
    size_t total = (size_t)lds_cu / waves_per_simd;
    total = total > 1024 ? total - 1024 : total;
    if (!(total > static_lds && total <= (size_t)lds_block && (size_t)lds_cu / total == waves_per_simd))
        return fail(RTC8X_ERR_INVALID_ARGUMENT, "%s: %u waves per SIMD cannot be set with %d bytes of LDS per CU", who,
                    waves_per_simd, lds_cu);
    *lds_out = total - static_lds;
    return RTC8X_OK;
}

}  // namespace rtc8x

using namespace rtc8x;

extern "C" rtc8x_status rtc8x_copy_async(void* dst, const void* src, size_t bytes, void* stream)
try {
    if (bytes == 0) return RTC8X_OK;
    RTC8X_REQUIRE(dst && src, "rtc8x_copy_async: NULL pointer");
    RTC8X_REQUIRE(((uintptr_t)dst | (uintptr_t)src | bytes) % 16 == 0, "rtc8x_copy_async: needs 16-byte alignment");
    const size_t n = bytes / 16;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 256 * 32);
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, (const uint4*)src,
                       (uint4*)dst, n);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_copy_async")

extern "C" rtc8x_status rtc8x_gather_probe(void* stream, const rtc8x_grid* grid, uint32_t threads, uint32_t steps,
                                           uint32_t width, uint32_t waves_per_simd, uint32_t seed, float* out)
try {
    RTC8X_REQUIRE(grid && out, "rtc8x_gather_probe: NULL pointer");
    const rtc8x_grid& g = *grid;
    RTC8X_REQUIRE(g.num_leaves > 0 && g.leaf_origins && g.leaf_values && g.tile_values,
                  "rtc8x_gather_probe: the grid needs leaves with their origins on the device");
    RTC8X_REQUIRE(threads > 0 && threads % 256 == 0 && steps > 0,
                  "rtc8x_gather_probe: steps must be > 0, threads a multiple of 256");
    RTC8X_REQUIRE(width == 1 || width == 2 || width == 4, "rtc8x_gather_probe: width must be 1, 2 or 4");
    RTC8X_REQUIRE(waves_per_simd >= 2 && waves_per_simd <= 8, "rtc8x_gather_probe: waves_per_simd must be in [2, 8]");
    size_t lds = 0;
    {
        const rtc8x_status st = occupancy_lds(waves_per_simd, &lds, "rtc8x_gather_probe");
        if (st != RTC8X_OK) return st;
    }
    hipStream_t s = (hipStream_t)stream;
    if (width == 1)
        hipLaunchKernelGGL(gather_probe_kernel<1>, dim3(threads / 256), dim3(256), lds, s, g, steps, seed, out);
    else if (width == 2)
        hipLaunchKernelGGL(gather_probe_kernel<2>, dim3(threads / 256), dim3(256), lds, s, g, steps, seed, out);
    else
        hipLaunchKernelGGL(gather_probe_kernel<4>, dim3(threads / 256), dim3(256), lds, s, g, steps, seed, out);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_gather_probe")

extern "C" rtc8x_status rtc8x_march_probe(void* stream, const rtc8x_frame_params* params, const float light_dir[3],
                                          uint32_t spp, uint32_t budget, uint32_t waves_per_simd, uint64_t* stats)
try {
    RTC8X_REQUIRE(params && light_dir && stats, "rtc8x_march_probe: NULL pointer");
    const rtc8x_frame_params& p = *params;
    RTC8X_REQUIRE(p.image_width > 0 && p.image_height > 0 && p.rng_states,
                  "rtc8x_march_probe: the frame needs an image and its PCG states");
    RTC8X_REQUIRE(p.grid.leaf_values && p.density_coeff * p.grid.max_value > 0.0f,
                  "rtc8x_march_probe: the grid must be on the device with a positive majorant");
    RTC8X_REQUIRE(spp > 0 && budget > 0 && budget <= (1u << 24), "rtc8x_march_probe: spp must be > 0, budget in [1, 2^24]");
    RTC8X_REQUIRE(waves_per_simd >= 2 && waves_per_simd <= 4, "rtc8x_march_probe: waves_per_simd must be in [2, 4]");
    RTC8X_REQUIRE(p.tracking == RTC8X_TRACKING_GLOBAL_MAJORANT || p.tracking == RTC8X_TRACKING_LOCAL_MAJORANT,
                  "rtc8x_march_probe: unknown tracking mode %d", (int)p.tracking);
    const bool local = p.tracking == RTC8X_TRACKING_LOCAL_MAJORANT;
    if (local) {
        const size_t nr = (size_t)std::max(p.grid.region_dims[0], 0) * std::max(p.grid.region_dims[1], 0) *
                          std::max(p.grid.region_dims[2], 0);
        RTC8X_REQUIRE(p.grid.region_majorants && nr > 0 && nr <= RTC8X_MAX_REGIONS && p.grid.region_cells > 0,
                      "rtc8x_march_probe: local tracking needs the grid's region table");
    }
    size_t lds = 0;
    {
        const rtc8x_status st = occupancy_lds(waves_per_simd, &lds, "rtc8x_march_probe", local ? RTC8X_MAX_REGIONS : 0);
        if (st != RTC8X_OK) return st;
    }
    const float n = std::sqrt(light_dir[0] * light_dir[0] + light_dir[1] * light_dir[1] + light_dir[2] * light_dir[2]);
    RTC8X_REQUIRE(n > 0.0f, "rtc8x_march_probe: light direction is zero");
    const dev::V3 l{light_dir[0] / n, light_dir[1] / n, light_dir[2] / n};
    // persistent: waves_per_simd 256-lane workgroups per CU, all resident at once
    int dev = 0, cus = 0;
    RTC8X_HIP(hipGetDevice(&dev));
    RTC8X_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(reset_probe_kernel, dim3(1), dim3(1), 0, s);
    RTC8X_HIP(hipGetLastError());
    const dim3 grid((uint32_t)cus * waves_per_simd), block(256);
    unsigned long long* st = (unsigned long long*)stats;
    if (local && waves_per_simd <= 3)
        hipLaunchKernelGGL((march_probe_kernel<RTC8X_TRACKING_LOCAL_MAJORANT, 3>), grid, block, lds, s, p, l, spp, budget, st);
    else if (local)
        hipLaunchKernelGGL((march_probe_kernel<RTC8X_TRACKING_LOCAL_MAJORANT, 4>), grid, block, lds, s, p, l, spp, budget, st);
    else if (waves_per_simd <= 3)
        hipLaunchKernelGGL((march_probe_kernel<RTC8X_TRACKING_GLOBAL_MAJORANT, 3>), grid, block, lds, s, p, l, spp, budget, st);
    else
        hipLaunchKernelGGL((march_probe_kernel<RTC8X_TRACKING_GLOBAL_MAJORANT, 4>), grid, block, lds, s, p, l, spp, budget, st);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_march_probe")

// rtc8x_nrc_data.hip -- neural-radiance-cache data path between the tracer and the network
// (renderer/gpu_kernels/nrc_set_up_kernels.cu): per-iteration reset and adaptive training size,
// radiance propagation along training suffixes, the training-record shuffle with target statistics,
// and the final composite of inferred radiance into the FP64 accumulation buffer.
//
// All four are bandwidth-light (a few hundred thousand 16-56 B records); one lane per record.
#include <hip/hip_runtime.h>

#include <cmath>

#include "rtc8x_common.hpp"
#include "rtc8x_device.hpp"

namespace rtc8x {
namespace dev {

constexpr uint32_t kShuffleBlock = 256;
constexpr uint32_t kShuffleBlocks = RTC8X_NUM_TRAINING_DATA_PER_FRAME / kShuffleBlock;  // 256

__device__ __forceinline__ RGB load_rgb(const float* p) { return RGB{p[0], p[1], p[2]}; }
__device__ __forceinline__ void store_rgb(float* p, RGB c)
{
    p[0] = c.r;
    p[1] = c.g;
    p[2] = c.b;
}
__device__ __forceinline__ RGB max0(RGB c) { return RGB{fmaxf(c.r, 0.0f), fmaxf(c.g, 0.0f), fmaxf(c.b, 0.0f)}; }
// safeDivide (common/basic_types.h:3094-3101)
__device__ __forceinline__ RGB safe_divide(RGB a, RGB b)
{
    return RGB{b.r != 0.0f ? a.r / b.r : 0.0f, b.g != 0.0f ? a.g / b.g : 0.0f, b.b != 0.0f ? a.b / b.b : 0.0f};
}
__device__ __forceinline__ RGB reflectance(const rtc8x_radiance_query& q)
{
    return RGB{q.diffuse_reflectance[0] + q.specular_reflectance[0], q.diffuse_reflectance[1] + q.specular_reflectance[1],
               q.diffuse_reflectance[2] + q.specular_reflectance[2]};
}
// RadianceQuery::isValid (renderer/renderer_shared.h:56-64)
__device__ __forceinline__ bool query_valid(const rtc8x_radiance_query& q)
{
    bool ok = isfinite(q.position[0]) && isfinite(q.position[1]) && isfinite(q.position[2]) && isfinite(q.normal_phi) &&
              isfinite(q.normal_theta) && isfinite(q.vout_phi) && isfinite(q.vout_theta) && isfinite(q.roughness);
    for (int c = 0; c < 3; ++c) ok = ok && isfinite(q.diffuse_reflectance[c]) && isfinite(q.specular_reflectance[c]);
    return ok;
}

__global__ void prepare_kernel(const rtc8x_nrc_buffers nb, int32_t img_w, int32_t img_h, int32_t is_new_sequence)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb.max_num_training_suffixes) return;
    const uint32_t buf = nb.nrc_buffer_index, prev = (buf + 1) % 2;
    if (i == 0) {
        int32_t nw, nh;
        if (is_new_sequence) {
            nw = img_w / 4;
            nh = img_h / 4;
        } else {
            const uint32_t prev_num = nb.num_training_data[prev];
            const float r = sqrtf((float)prev_num / (float)RTC8X_NUM_TRAINING_DATA_PER_FRAME);
            const int32_t cw = nb.train_image_size[2 * prev], ch = nb.train_image_size[2 * prev + 1];
            nw = (int32_t)fminf((float)cw / r, (float)img_w);
            nh = (int32_t)fminf((float)ch / r, (float)img_h);
        }
        nw = max(nw, 1);
        nh = max(nh, 1);
        nb.train_image_size[2 * buf] = nw;
        nb.train_image_size[2 * buf + 1] = nh;
        nb.num_training_data[buf] = 0;
        for (int c = 0; c < 3; ++c) {
            nb.target_min_max[6 * buf + c] = __builtin_huge_valf();
            nb.target_min_max[6 * buf + 3 + c] = -__builtin_huge_valf();
            nb.target_avg[3 * buf + c] = 0.0f;
        }
    }
    rtc8x_training_suffix_terminal t;
    t.prev_vertex_index = RTC8X_INVALID_VERTEX_INDEX;
    t.has_query = 0;
    t.path_length = 0;
    t.is_unbiased = 0;
    nb.train_suffix_terminals[i] = t;
}

// Launched over every suffix slot; the training size comes from the device (prepareNRC wrote it).
__global__ void propagate_kernel(const rtc8x_nrc_buffers nb, uint32_t max_n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t* ts = nb.train_image_size + 2 * nb.nrc_buffer_index;
    const uint64_t n = (uint64_t)max(ts[0], 0) * (uint64_t)max(ts[1], 0);
    if (i >= max_n || i >= n) return;
    const rtc8x_training_suffix_terminal ti = nb.train_suffix_terminals[i];
    if (ti.prev_vertex_index == RTC8X_INVALID_VERTEX_INDEX) return;
    RGB contribution{0.0f, 0.0f, 0.0f};
    if (ti.has_query) {
        contribution = max0(load_rgb(nb.inferred_radiance + 3 * (size_t)i));
        if (nb.radiance_scale > 0.0f) contribution = scale(contribution, 1.0f / nb.radiance_scale);
        contribution = mul(contribution, reflectance(nb.inference_queries[i]));
    }
    uint32_t idx = ti.prev_vertex_index;
    while (idx != RTC8X_INVALID_VERTEX_INDEX) {
        const rtc8x_training_vertex_info v = nb.train_vertex_info[idx];
        float* tgt = nb.train_targets[0] + 3 * (size_t)idx;
        RGB indirect = mul(RGB{v.local_throughput[0], v.local_throughput[1], v.local_throughput[2]}, contribution);
        contribution = add(load_rgb(tgt), indirect);
        store_rgb(tgt, safe_divide(contribution, reflectance(nb.train_queries[0][idx])));
        idx = v.prev_vertex_index;
    }
}

// Block partials of the target statistics: min RGB, max RGB, sum of target/65536 RGB (fixed tree order).
__global__ __launch_bounds__(kShuffleBlock) void shuffle_kernel(const rtc8x_nrc_buffers nb, float* __restrict__ partials)
{
    const uint32_t i = blockIdx.x * kShuffleBlock + threadIdx.x;
    const uint32_t buf = nb.nrc_buffer_index;
    const uint32_t num = nb.num_training_data[buf];
    __shared__ float smin[3][kShuffleBlock], smax[3][kShuffleBlock], ssum[3][kShuffleBlock];
    RGB target{0.0f, 0.0f, 0.0f};
    if (num > 0) {
        uint32_t state = nb.data_shufflers[i];
        state = ((state * 1103515245u) + 12345u) % (1u << 31);  // LinearCongruentialGenerator::next
        nb.data_shufflers[i] = state;
        const uint32_t dst = state % RTC8X_NUM_TRAINING_DATA_PER_FRAME;
        const uint32_t src = i % num;
        rtc8x_radiance_query q = nb.train_queries[0][src];
        target = load_rgb(nb.train_targets[0] + 3 * (size_t)src);
        if (!query_valid(q)) {
            q.position[0] = q.position[1] = q.position[2] = 0.0f;
            q.normal_phi = q.normal_theta = q.vout_phi = q.vout_theta = q.roughness = 0.0f;
            for (int c = 0; c < 3; ++c) q.diffuse_reflectance[c] = q.specular_reflectance[c] = 0.0f;
        }
        if (!all_nonneg_finite(target)) target = RGB{0.0f, 0.0f, 0.0f};
        const float t[3] = {target.r, target.g, target.b};
        for (int c = 0; c < 3; ++c) {
            smin[c][threadIdx.x] = t[c];
            smax[c][threadIdx.x] = t[c];
            ssum[c][threadIdx.x] = t[c] * (1.0f / (float)RTC8X_NUM_TRAINING_DATA_PER_FRAME);
        }
        RGB scaled = target;
        if (nb.radiance_scale > 0.0f) scaled = scale(scaled, nb.radiance_scale);
        scaled = RGB{fminf(scaled.r, 1e6f), fminf(scaled.g, 1e6f), fminf(scaled.b, 1e6f)};
        nb.train_queries[1][dst] = q;
        store_rgb(nb.train_targets[1] + 3 * (size_t)dst, scaled);
    } else {
        rtc8x_radiance_query q = {};
        nb.train_queries[1][i] = q;
        store_rgb(nb.train_targets[1] + 3 * (size_t)i, RGB{0.0f, 0.0f, 0.0f});
        for (int c = 0; c < 3; ++c) {
            smin[c][threadIdx.x] = __builtin_huge_valf();
            smax[c][threadIdx.x] = -__builtin_huge_valf();
            ssum[c][threadIdx.x] = 0.0f;
        }
    }
    __syncthreads();
    for (uint32_t s = kShuffleBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int c = 0; c < 3; ++c) {
                smin[c][threadIdx.x] = fminf(smin[c][threadIdx.x], smin[c][threadIdx.x + s]);
                smax[c][threadIdx.x] = fmaxf(smax[c][threadIdx.x], smax[c][threadIdx.x + s]);
                ssum[c][threadIdx.x] = ssum[c][threadIdx.x] + ssum[c][threadIdx.x + s];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float* out = partials + 9 * blockIdx.x;
        for (int c = 0; c < 3; ++c) {
            out[c] = smin[c][0];
            out[3 + c] = smax[c][0];
            out[6 + c] = ssum[c][0];
        }
    }
}

__global__ __launch_bounds__(kShuffleBlocks) void shuffle_stats_kernel(const rtc8x_nrc_buffers nb,
                                                                       const float* __restrict__ partials)
{
    __shared__ float smin[3][kShuffleBlocks], smax[3][kShuffleBlocks], ssum[3][kShuffleBlocks];
    const uint32_t t = threadIdx.x;
    for (int c = 0; c < 3; ++c) {
        smin[c][t] = partials[9 * t + c];
        smax[c][t] = partials[9 * t + 3 + c];
        ssum[c][t] = partials[9 * t + 6 + c];
    }
    __syncthreads();
    for (uint32_t s = kShuffleBlocks / 2; s > 0; s >>= 1) {
        if (t < s) {
            for (int c = 0; c < 3; ++c) {
                smin[c][t] = fminf(smin[c][t], smin[c][t + s]);
                smax[c][t] = fmaxf(smax[c][t], smax[c][t + s]);
                ssum[c][t] = ssum[c][t] + ssum[c][t + s];
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        const uint32_t buf = nb.nrc_buffer_index;
        if (nb.num_training_data[buf] == 0) return;
        for (int c = 0; c < 3; ++c) {
            nb.target_min_max[6 * buf + c] = smin[c][0];
            nb.target_min_max[6 * buf + 3 + c] = smax[c][0];
            nb.target_avg[3 * buf + c] = ssum[c][0];
        }
    }
}

// One thread per pixel adds the launch's samples in order, as that many one-sample launches would.
__global__ void accumulate_kernel(const rtc8x_nrc_buffers nb, uint32_t n, uint32_t samples, double* __restrict__ accum)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t q_stride = ((size_t)n + 127u) & ~(size_t)127u;
    double* a = accum + 3 * (size_t)i;
    double s0 = a[0], s1 = a[1], s2 = a[2];
    for (uint32_t k = 0; k < samples; ++k) {
        const size_t pi = (size_t)k * n + i, qi = (size_t)k * q_stride + i;
        const rtc8x_terminal_info ti = nb.inference_terminals[pi];
        const RGB direct = load_rgb(nb.per_frame_contribution + 3 * pi);
        RGB radiance{0.0f, 0.0f, 0.0f};
        if (ti.has_query & 1u) {
            radiance = max0(load_rgb(nb.inferred_radiance + 3 * qi));
            if (nb.radiance_scale > 0.0f) radiance = scale(radiance, 1.0f / nb.radiance_scale);
            radiance = mul(radiance, reflectance(nb.inference_queries[qi]));
        }
        const RGB indirect = mul(RGB{ti.throughput[0], ti.throughput[1], ti.throughput[2]}, radiance);
        const RGB c = add(direct, indirect);
        s0 += (double)c.r;
        s1 += (double)c.g;
        s2 += (double)c.b;
    }
    a[0] = s0;
    a[1] = s1;
    a[2] = s2;
}

}  // namespace dev
}  // namespace rtc8x

using namespace rtc8x;

static rtc8x_status check_nrc(const rtc8x_frame_params* p, const char* what)
{
    RTC8X_REQUIRE(p != nullptr, "%s: params is NULL", what);
    const rtc8x_nrc_buffers& nb = p->nrc;
    RTC8X_REQUIRE(nb.nrc_buffer_index < 2, "%s: nrc_buffer_index must be 0 or 1", what);
    RTC8X_REQUIRE(p->image_width > 0 && p->image_height > 0, "%s: bad image size", what);
    return RTC8X_OK;
}

extern "C" size_t rtc8x_nrc_scratch_bytes(uint32_t width, uint32_t height)
{
    static_assert(dev::kShuffleBlocks == RTC8X_NUM_TRAINING_DATA_PER_FRAME / 256, "scratch layout");
    return nrc_scratch_layout(width, height).total;
}

extern "C" rtc8x_status rtc8x_nrc_prepare(void* stream, const rtc8x_frame_params* p, int32_t is_new_sequence)
try {
    rtc8x_status st = check_nrc(p, "rtc8x_nrc_prepare");
    if (st != RTC8X_OK) return st;
    const rtc8x_nrc_buffers& nb = p->nrc;
    RTC8X_REQUIRE(nb.num_training_data && nb.train_image_size && nb.target_min_max && nb.target_avg &&
                      nb.train_suffix_terminals && nb.max_num_training_suffixes > 0,
                  "rtc8x_nrc_prepare: buffers missing");
    const uint32_t n = nb.max_num_training_suffixes;
    hipLaunchKernelGGL(dev::prepare_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, nb,
                       (int32_t)p->image_width, (int32_t)p->image_height, is_new_sequence);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_prepare")

extern "C" rtc8x_status rtc8x_nrc_propagate(void* stream, const rtc8x_frame_params* p)
try {
    rtc8x_status st = check_nrc(p, "rtc8x_nrc_propagate");
    if (st != RTC8X_OK) return st;
    const rtc8x_nrc_buffers& nb = p->nrc;
    RTC8X_REQUIRE(nb.train_suffix_terminals && nb.train_vertex_info && nb.train_targets[0] && nb.train_queries[0] &&
                      nb.inference_queries && nb.inferred_radiance && nb.train_image_size,
                  "rtc8x_nrc_propagate: buffers missing");
    RTC8X_REQUIRE(nb.max_num_training_suffixes > 0, "rtc8x_nrc_propagate: no suffix slots");
    const uint32_t n = nb.max_num_training_suffixes;
    hipLaunchKernelGGL(dev::propagate_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, nb, n);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_propagate")

extern "C" rtc8x_status rtc8x_nrc_shuffle(void* stream, const rtc8x_frame_params* p)
try {
    rtc8x_status st = check_nrc(p, "rtc8x_nrc_shuffle");
    if (st != RTC8X_OK) return st;
    const rtc8x_nrc_buffers& nb = p->nrc;
    RTC8X_REQUIRE(nb.num_training_data && nb.data_shufflers && nb.train_queries[0] && nb.train_queries[1] &&
                      nb.train_targets[0] && nb.train_targets[1] && nb.target_min_max && nb.target_avg && nb.scratch,
                  "rtc8x_nrc_shuffle: buffers missing");
    float* partials = reinterpret_cast<float*>(nb.scratch + 2 * (size_t)p->image_width * p->image_height);
    hipLaunchKernelGGL(dev::shuffle_kernel, dim3(dev::kShuffleBlocks), dim3(dev::kShuffleBlock), 0, (hipStream_t)stream,
                       nb, partials);
    RTC8X_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::shuffle_stats_kernel, dim3(1), dim3(dev::kShuffleBlocks), 0, (hipStream_t)stream, nb,
                       partials);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_shuffle")

extern "C" rtc8x_status rtc8x_nrc_accumulate(void* stream, const rtc8x_frame_params* p)
try {
    rtc8x_status st = check_nrc(p, "rtc8x_nrc_accumulate");
    if (st != RTC8X_OK) return st;
    const rtc8x_nrc_buffers& nb = p->nrc;
    RTC8X_REQUIRE(nb.inference_terminals && nb.per_frame_contribution && nb.inferred_radiance && nb.inference_queries &&
                      p->accum,
                  "rtc8x_nrc_accumulate: buffers missing");
    const uint32_t n = p->image_width * p->image_height;
    const uint32_t samples = nb.num_samples > 1 ? nb.num_samples : 1u;
    hipLaunchKernelGGL(dev::accumulate_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, nb, n, samples,
                       p->accum);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_accumulate")

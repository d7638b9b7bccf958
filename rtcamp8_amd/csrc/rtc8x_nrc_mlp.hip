// rtc8x_nrc_mlp.hip -- the neural radiance cache network on gfx950 MFMA (replaces tiny-cuda-nn's
// NetworkWithInputEncoding + FullyFusedMLP + RelativeL2Luminance + EMA(Adam) behind
// renderer/network_interface.cu).
//
// Forward (inference and training): a workgroup of 4 wavefronts x 64 samples (training batches of up
// to 32,768 samples: one wavefront, so that they spread over all CUs).  Each lane encodes one query into 64 features and writes them to its wavefront's LDS tile; every layer is then
// 64x64x64 per wavefront = 4 (samples) x 4 (outputs) x 2 (K) v_mfma_f32_16x16x32 {f16,bf16}
// instructions with FP32 accumulation (FP32 weights: 8 v_mfma_f32_16x16x4_f32 per 16x16x32 step).  A operands (activations) come from LDS, B operands (weights,
// [out][in] rows) straight from global memory, where the ~45 KB of weights stay L1/L2-resident for
// the whole launch.  ReLU and the down-conversion are fused into the write-back to LDS.
//
// Backward: a workgroup owns 64 or 256 samples, as the forward pass.  Per layer, from the top:
// dW = dZ^T H (K = the workgroup's samples, operands feature-major in LDS) and dH = dZ W (K = layer width, transposed weights from
// global), with the ReLU mask applied on the way down.  Each workgroup writes its dW to its own slot;
// the slots are summed in a fixed order, so a training step is deterministic.
//
// Optimiser: FP32 master weights, Adam moments and the weight EMA; one lane per weight.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "rtc8x_common.hpp"
#include "rtc8x_device.hpp"

namespace rtc8x {
namespace nrc {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef __bf16 b4 __attribute__((ext_vector_type(4)));

// four values of the working type (one 8- or 16-byte store)
template <typename T>
struct Vec4;
template <>
struct Vec4<_Float16> {
    using type = h4;
};
template <>
struct Vec4<__bf16> {
    using type = b4;
};
template <>
struct Vec4<float> {
    using type = f4;
};

template <typename T>
struct Vec8;
template <>
struct Vec8<_Float16> {
    using type = h8;
};
template <>
struct Vec8<__bf16> {
    using type = b8;
};
template <>
struct Vec8<float> {
    using type = f8;
};

__device__ __forceinline__ f4 mfma(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ f4 mfma(b8 a, b8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
// A 16x16x32 step in FP32.  In the 16x16x32 layouts lane l holds row (l & 15) and the eight K indices
// 8 (l >> 4) + j; v_mfma_f32_16x16x4_f32 takes one K index per lane, (l >> 4), with the same row and
// output layouts.  Instruction j therefore covers the K indices 8 g + j (g = 0..3), and the eight of
// them together cover all 32.
__device__ __forceinline__ f4 mfma(f8 a, f8 b, f4 c)
{
#pragma unroll
    for (int j = 0; j < 8; ++j) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[j], b[j], c, 0, 0, 0);
    return c;
}

constexpr uint32_t kWidth = 64;
constexpr uint32_t kLayer = kWidth * kWidth;             // 4096 weights per hidden layer
constexpr uint32_t kOutPad = RTC8X_NRC_OUTPUT_PAD;       // output rows padded to one 16-row tile
constexpr uint32_t kOutT = 32;                           // transposed output layer: [64 in][32]
constexpr uint32_t kActStride = 72;                      // LDS row stride of a 64-feature row (halves)
constexpr uint32_t kMaxTrain = 65536;
constexpr uint32_t kMaxLayers = 8;

constexpr uint32_t kHashLevels = RTC8X_NRC_HASH_LEVELS;
constexpr uint32_t kHashT = 1u << RTC8X_NRC_HASH_LOG2_TABLE;
constexpr int kEncTriangle = RTC8X_NRC_TRIANGLE_WAVE;
constexpr int kEncHash = RTC8X_NRC_HASH_GRID;

// Per-level geometry of the hash-grid encoding (the scale, the dense stride r + 1 or 0 when the level
// is hashed, and the level's float offset in the tables).
struct HashGeom {
    float scale[kHashLevels];
    uint32_t stride[kHashLevels];
    uint32_t offset[kHashLevels];
};

// Fills the geometry; returns the number of table floats.
inline uint32_t hash_geometry(HashGeom* g)
{
    uint32_t off = 0;
    for (uint32_t l = 0; l < kHashLevels; ++l) {
        const float sc = std::exp2((float)l) * (float)RTC8X_NRC_HASH_BASE_RESOLUTION - 1.0f;
        const uint64_t r1 = (uint64_t)std::ceil(sc) + 2;  // corner coordinates 0..r, r = ceil(s) + 1
        const uint64_t dense = r1 * r1 * r1;
        const uint32_t entries = dense <= kHashT ? (uint32_t)((dense + 7) / 8 * 8) : kHashT;
        g->scale[l] = sc;
        g->stride[l] = dense <= kHashT ? (uint32_t)r1 : 0u;
        g->offset[l] = off;
        off += entries * RTC8X_NRC_HASH_FEATURES;
    }
    return off;
}

// Corner c (bit 0: x, 1: y, 2: z) of level l for the clamped position p: table float index and weight.
__device__ __forceinline__ void hash_corner(const HashGeom& gm, uint32_t l, const uint32_t gi[3], const float f[3],
                                            uint32_t c, uint32_t& idx, float& w)
{
    const uint32_t x = gi[0] + (c & 1u), y = gi[1] + ((c >> 1) & 1u), z = gi[2] + (c >> 2);
    w = ((c & 1u) ? f[0] : 1.0f - f[0]) * ((c & 2u) ? f[1] : 1.0f - f[1]) * ((c & 4u) ? f[2] : 1.0f - f[2]);
    const uint32_t st = gm.stride[l];
    const uint32_t e = st ? x + y * st + z * st * st : ((x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u)) & (kHashT - 1u);
    idx = gm.offset[l] + 2u * e;
}

__device__ __forceinline__ void hash_cell(const HashGeom& gm, uint32_t l, const float p[3], uint32_t gi[3], float f[3])
{
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float pos = p[d] * gm.scale[l] + 0.5f;
        const float fl = floorf(pos);
        f[d] = pos - fl;
        gi[d] = (uint32_t)fl;
    }
}

__device__ __forceinline__ void query_position01(const rtc8x_radiance_query& q, float p[3])
{
#pragma unroll
    for (int d = 0; d < 3; ++d) p[d] = fminf(fmaxf(q.position[d], 0.0f), 1.0f);
}

__host__ __device__ constexpr uint32_t num_params(uint32_t L) { return L * kLayer + kOutPad * kWidth; }
__host__ __device__ constexpr uint32_t num_params_t(uint32_t L) { return L * kLayer + kWidth * kOutT; }

// Input encoding of one RadianceQuery (renderer/renderer_shared.h:46-55), 64 features.
__device__ __forceinline__ void encode(const rtc8x_radiance_query& q, float* e)
{
    // Triangle waves of period 2 / 2^f: tri(x 2^f) = |2 frac(x 2^(f-1)) - 1|.  The base frequency is
    // monotonic on [0, 1]; with period 1 at every frequency the encoding would not tell x from 1 - x.
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float sc = 0.5f;
#pragma unroll
        for (int f = 0; f < 12; ++f) {
            float t = q.position[d] * sc;
            t = t - floorf(t);
            e[12 * d + f] = fabsf(2.0f * t - 1.0f);
            sc = sc * 2.0f;
        }
    }
    const float inv_pi = 1.0f / RTC8X_PI_F, inv_2pi = 1.0f / (2.0f * RTC8X_PI_F);
    // The *_phi fields hold polar angles and the *_theta fields azimuths (see rtc8x_radiance_query).
    const float a[5] = {q.normal_phi * inv_pi, q.normal_theta * inv_2pi, q.vout_phi * inv_pi, q.vout_theta * inv_2pi,
                        q.roughness};
#pragma unroll
    for (int j = 0; j < 5; ++j) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            float d = a[j] - ((float)b + 0.5f) * 0.25f;
            e[36 + 4 * j + b] = expf(-8.0f * d * d);  // Gaussian one-blob kernel, sigma = 1/4 bins
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        e[56 + c] = q.diffuse_reflectance[c];
        e[59 + c] = q.specular_reflectance[c];
    }
    e[62] = 1.0f;
    e[63] = 0.0f;
}

// The grid part of the hash encoding: level l's two features, trilinear over the cell's 8 corners.
__device__ __forceinline__ float2 hash_level(const float p[3], const float* __restrict__ tables, const HashGeom& gm,
                                             uint32_t l)
{
    uint32_t gi[3];
    float f[3];
    hash_cell(gm, l, p, gi, f);
    float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
    for (uint32_t c = 0; c < 8; ++c) {
        uint32_t idx;
        float w;
        hash_corner(gm, l, gi, f, c, idx, w);
        float2 v;
        __builtin_memcpy(&v, tables + idx, sizeof(v));
        a0 += w * v.x;
        a1 += w * v.y;
    }
    return float2{a0, a1};
}

// Features 32..63 of the hash encoding (OneBlob of the angles and roughness, the reflectances, 1, zeros),
// into t[0..31].
__device__ __forceinline__ void encode_hash_tail(const rtc8x_radiance_query& q, float* t)
{
    const float inv_pi = 1.0f / RTC8X_PI_F, inv_2pi = 1.0f / (2.0f * RTC8X_PI_F);
    const float a[5] = {q.normal_phi * inv_pi, q.normal_theta * inv_2pi, q.vout_phi * inv_pi, q.vout_theta * inv_2pi,
                        q.roughness};
#pragma unroll
    for (int j = 0; j < 5; ++j) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            float d = a[j] - ((float)b + 0.5f) * 0.25f;
            t[4 * j + b] = expf(-8.0f * d * d);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        t[20 + c] = q.diffuse_reflectance[c];
        t[23 + c] = q.specular_reflectance[c];
    }
    t[26] = 1.0f;
#pragma unroll
    for (int j = 27; j < 32; ++j) t[j] = 0.0f;
}

// Input encoding with the hash grid: [grid 32 | OneBlob 20 | identity 6 | 1 | 0 x 5].
__device__ __forceinline__ void encode_hash(const rtc8x_radiance_query& q, const float* __restrict__ tables,
                                            const HashGeom& gm, float* e)
{
    float p[3];
    query_position01(q, p);
    for (uint32_t l = 0; l < kHashLevels; ++l) {
        const float2 a = hash_level(p, tables, gm, l);
        e[2 * l] = a.x;
        e[2 * l + 1] = a.y;
    }
    encode_hash_tail(q, e + 32);
}

// The hash encoding of query q (NULL: zeros) straight into an activation row: the grid features go
// to the row level by level, so no per-lane array is indexed at run time.  With BATCH, all levels'
// corner gathers are issued before the row is written, so that they are in flight together: at full
// occupancy other waves hide the gather latency and the level loop is as fast, while a training batch
// of one wave per CU would wait on each level's gathers in turn.
template <typename T, bool BATCH>
__device__ __forceinline__ void encode_hash_row(const rtc8x_radiance_query* q, const float* __restrict__ tables,
                                                const HashGeom& gm, T* row)
{
    using V = typename Vec8<T>::type;
    float t[32];
    if (q) {
        float p[3];
        query_position01(*q, p);
        if constexpr (BATCH) {
            float2 a[kHashLevels];
#pragma unroll
            for (uint32_t l = 0; l < kHashLevels; ++l) a[l] = hash_level(p, tables, gm, l);
#pragma unroll
            for (uint32_t l = 0; l < kHashLevels; ++l) {
                row[2 * l] = (T)a[l].x;
                row[2 * l + 1] = (T)a[l].y;
            }
        } else {
            for (uint32_t l = 0; l < kHashLevels; ++l) {
                const float2 a = hash_level(p, tables, gm, l);
                row[2 * l] = (T)a.x;
                row[2 * l + 1] = (T)a.y;
            }
        }
        encode_hash_tail(*q, t);
    } else {
#pragma unroll
        for (int j = 0; j < 32; ++j) row[j] = (T)0.0f;
#pragma unroll
        for (int j = 0; j < 32; ++j) t[j] = 0.0f;
    }
#pragma unroll
    for (int j = 0; j < 32; j += 8) {
        V v;
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = (T)t[j + k];
        *reinterpret_cast<V*>(&row[32 + j]) = v;
    }
}

template <int ENC>
__device__ __forceinline__ void encode_any(const rtc8x_radiance_query& q, const float* tables, const HashGeom& gm, float* e)
{
    if constexpr (ENC == kEncHash)
        encode_hash(q, tables, gm, e);
    else
        encode(q, e);
}

// With count_dims (a training size w x h on the device), round_up(w h, 128) queries, at most n.
__device__ __forceinline__ uint32_t device_count(uint32_t n, const int32_t* count_dims)
{
    if (!count_dims) return n;
    const uint64_t m = ((uint64_t)max(count_dims[0], 0) * (uint64_t)max(count_dims[1], 0) + 127u) & ~(uint64_t)127u;
    return m < n ? (uint32_t)m : n;
}

constexpr float kHalfMax = 65504.0f;  // the largest finite FP16 value

// The slots of an inference that hold a query.  With terms, the layout of the NRC render entry's
// buffers (sample k of pixel i: query slot k * q_stride + i, terminal k * npx + i) and the terminals'
// query flags; without, every slot below the count.
struct QueryMask {
    const rtc8x_terminal_info* terms;
    uint32_t npx, q_stride;
};
__device__ __forceinline__ bool slot_has_query(const QueryMask& m, uint32_t s)
{
    if (!m.terms) return true;
    const uint32_t k = s / m.q_stride, i = s - k * m.q_stride;
    return i < m.npx && (m.terms[(size_t)k * m.npx + i].has_query & 1u) != 0u;
}

// Training batches up to this size run the training kernels with one wave (64 samples) per workgroup,
// so that a 16,384-sample step spreads over 256 workgroups instead of 64; larger ones with four.
constexpr uint32_t kSmallBatch = 32768;

// The activations a training forward pass saves for the backward pass: per layer, per 64-sample tile,
// feature-major (64 features x 64 samples), so that the backward pass reads its column tiles with
// 16-byte loads.  Samples past the batch end of the last tile hold zeros.
__host__ __device__ __forceinline__ size_t saved_tile(uint32_t l, uint32_t n, uint32_t tile)
{
    return ((size_t)l * ((n + 63) / 64) + tile) * (64 * 64);
}

// The wave's 64 activation rows (LDS, sample-major) into its saved tile: lane c gathers feature c.
template <typename T>
__device__ __forceinline__ void save_tile(const T* A, uint32_t lane, T* __restrict__ tile)
{
    using V = typename Vec8<T>::type;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        V v;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = A[(8 * k + j) * kActStride + lane];
        *reinterpret_cast<V*>(&tile[lane * 64 + 8 * k]) = v;
    }
}

template <typename T, bool TRAIN, int ENC, int WPB>
// (256, 3): the register budget of three 4-wave workgroups per CU, under which the FP16/BF16
// accumulators stay in VGPRs and the kernel runs four waves per SIMD (0.19 ms per 2^21 queries at
// L = 5, against 0.21 ms at three; FP32 stays at two, bound by its LDS tiles).  One-wave workgroups
// (training batches) take the registers that keep all hash levels' gathers in flight.
__global__ __launch_bounds__(64 * WPB, WPB == 1 ? 1 : 3) void forward_kernel(const T* __restrict__ w, uint32_t L,
                                                      const rtc8x_radiance_query* __restrict__ in, uint32_t n,
                                                      float* __restrict__ out, const float* __restrict__ target,
                                                      T* __restrict__ hbuf, T* __restrict__ dout,
                                                      float* __restrict__ loss_part, const float* __restrict__ tables,
                                                      const HashGeom geom, const int32_t* __restrict__ count_dims,
                                                      const float* __restrict__ loss_scale, const QueryMask mask)
{
    using V = typename Vec8<T>::type;
    n = device_count(n, count_dims);
    if (blockIdx.x * (64u * WPB) >= n) return;
    __shared__ __align__(16) T act[WPB][64 * kActStride];
    __shared__ float pred[WPB][64][3];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t s0 = (blockIdx.x * WPB + wave) * 64;
    const uint32_t s = s0 + lane;
    // inference: slots without a query are neither encoded nor written, and a wavefront without any
    // query ends here (the wavefronts are independent: no workgroup barrier follows)
    const bool live = s < n && (TRAIN || slot_has_query(mask, s));
    if (!TRAIN && __ballot(live) == 0) return;
    T* A = act[wave];
    if constexpr (ENC == kEncHash) {
        encode_hash_row<T, WPB == 1>(live ? &in[s] : nullptr, tables, geom, &A[lane * kActStride]);
        if (TRAIN && s0 < n) save_tile(A, lane, hbuf + saved_tile(0, n, s0 / 64));
    } else {
        float e[64];
        if (live) {
            encode_any<ENC>(in[s], tables, geom, e);
        } else {
#pragma unroll
            for (int j = 0; j < 64; ++j) e[j] = 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 64; j += 8) {
            V v;
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = (T)e[j + k];
            *reinterpret_cast<V*>(&A[lane * kActStride + j]) = v;
        }
        if (TRAIN && s0 < n) save_tile(A, lane, hbuf + saved_tile(0, n, s0 / 64));
    }
    const uint32_t r16 = lane & 15, kq = 8 * (lane >> 4);
    for (uint32_t l = 0; l <= L; ++l) {
        const bool last = l == L;
        const T* W = w + (size_t)l * kLayer;
        const int ntiles = last ? 1 : 4;
        f4 acc[4][4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            V a[4];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
                a[mt] = *reinterpret_cast<const V*>(&A[(mt * 16 + r16) * kActStride + ks * 32 + kq]);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                if (nt < ntiles) {
                    const V b = *reinterpret_cast<const V*>(&W[(nt * 16 + r16) * kWidth + ks * 32 + kq]);
                    // weights as the A operand, activations as B: acc[mt][nt][i] is feature
                    // nt * 16 + 4 (lane >> 4) + i of query mt * 16 + (lane & 15)
#pragma unroll
                    for (int mt = 0; mt < 4; ++mt) acc[mt][nt] = mfma(b, a[mt], acc[mt][nt]);
                }
            }
        }
        if (!last) {
            // ReLU and four consecutive features of one query per store
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    typename Vec4<T>::type v;
#pragma unroll
                    for (int i = 0; i < 4; ++i) v[i] = (T)fmaxf(acc[mt][nt][i], 0.0f);
                    *reinterpret_cast<typename Vec4<T>::type*>(&A[(mt * 16 + r16) * kActStride + nt * 16 + 4 * (lane >> 4)]) = v;
                }
            if (TRAIN && s0 < n) save_tile(A, lane, hbuf + saved_tile(l + 1, n, s0 / 64));
        } else if ((lane >> 4) == 0) {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int i = 0; i < 3; ++i) pred[wave][mt * 16 + r16][i] = acc[mt][0][i];
        }
    }
    if (!TRAIN) {
        if (live) {
            out[3 * (size_t)s] = pred[wave][lane][0];
            out[3 * (size_t)s + 1] = pred[wave][lane][1];
            out[3 * (size_t)s + 2] = pred[wave][lane][2];
        }
        return;
    }
    // relative L2 over luminance, the mean over the batch's 3n outputs (the loss tiny-cuda-nn's
    // RelativeL2Luminance computes); per-sample gradient 2 (p - t) / (lum(p)^2 + 0.01) / (3n), times the
    // loss scale before it is rounded to the working type.  A scaled gradient outside FP16's finite range
    // is made infinite in every precision (as the FP16 rounding makes it), so the step is skipped and the
    // loss scale backs off: all precisions take the same steps.
    float loss = 0.0f;
    const float ls = *loss_scale / (3.0f * (float)n);
    if (s < n) {
        const float p0 = pred[wave][lane][0], p1 = pred[wave][lane][1], p2 = pred[wave][lane][2];
        const float lum = (float)0.2126729 * p0 + (float)0.7151522 * p1 + (float)0.0721750 * p2;
        const float denom = lum * lum + 0.01f;
        const float d0 = p0 - target[3 * (size_t)s], d1 = p1 - target[3 * (size_t)s + 1], d2 = p2 - target[3 * (size_t)s + 2];
        loss = (d0 * d0 + d1 * d1 + d2 * d2) / denom;
        T* g = dout + (size_t)s * kOutPad;
        const float x0 = 2.0f * d0 / denom * ls, x1 = 2.0f * d1 / denom * ls, x2 = 2.0f * d2 / denom * ls;
        const bool out = !(fabsf(x0) <= kHalfMax && fabsf(x1) <= kHalfMax && fabsf(x2) <= kHalfMax);
        g[0] = (T)(out ? INFINITY : x0);
        g[1] = (T)x1;
        g[2] = (T)x2;
#pragma unroll
        for (int c = 3; c < (int)kOutPad; ++c) g[c] = (T)0.0f;
    }
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o, 64);
    if (lane == 0) loss_part[blockIdx.x * WPB + wave] = loss;
}

// ---------------------------------------------------------------- FP8 inference
// E4M3 operands on v_mfma_f32_16x16x32_fp8_fp8 with FP32 accumulation.  Weights carry one power-of-two
// scale per layer (quantize_kernel, from the BF16 EMA weights); activations one power-of-two scale per
// wavefront tile and layer, from the tile's largest magnitude.  Both map the largest value to at most
// kFp8Max, inside the range of the E4M3 formats, and are divided out of the FP32 accumulators.
constexpr float kFp8Max = 240.0f;

__device__ __forceinline__ float pow2_scale(float maxabs)
{
    if (!(maxabs > 0.0f) || !isfinite(maxabs)) return 1.0f;
    return exp2f(floorf(log2f(kFp8Max / maxabs)));
}

// 8 floats -> 8 E4M3 bytes (round to nearest even), element k in byte k
__device__ __forceinline__ long pack_fp8(const float (&x)[8])
{
    int lo = 0, hi = 0;
    lo = __builtin_amdgcn_cvt_pk_fp8_f32(x[0], x[1], lo, false);
    lo = __builtin_amdgcn_cvt_pk_fp8_f32(x[2], x[3], lo, true);
    hi = __builtin_amdgcn_cvt_pk_fp8_f32(x[4], x[5], hi, false);
    hi = __builtin_amdgcn_cvt_pk_fp8_f32(x[6], x[7], hi, true);
    return (long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// One workgroup per layer: the layer's scale and its E4M3 weights ([out][in], like the BF16 copy).
__global__ __launch_bounds__(256) void quantize_kernel(const __bf16* __restrict__ winf, uint32_t L,
                                                       uint8_t* __restrict__ w8, float* __restrict__ inv_scale)
{
    const uint32_t l = blockIdx.x;
    const uint32_t count = l < L ? kLayer : kOutPad * kWidth;
    const __bf16* src = winf + (size_t)l * kLayer;
    __shared__ float red[256];
    float m = 0.0f;
    for (uint32_t i = threadIdx.x; i < count; i += 256) m = fmaxf(m, fabsf((float)src[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    const float sc = pow2_scale(red[0]);
    if (threadIdx.x == 0) inv_scale[l] = 1.0f / sc;
    for (uint32_t i = 8 * threadIdx.x; i < count; i += 8 * 256) {
        float x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = (float)src[i + k] * sc;
        *reinterpret_cast<long*>(w8 + (size_t)l * kLayer + i) = pack_fp8(x);
    }
}

// Inference with E4M3 operands: the BF16 forward's structure (encoding into the LDS tile, then per
// layer 4 x 4 x 2 MFMAs per wavefront), with each layer's activations scaled and quantised on the way
// from LDS into the A operands.
template <int ENC>
// (256, 3) as for forward_kernel: four waves per SIMD instead of two (0.20 ms per 2^21 queries at L = 5,
// against 0.29 ms)
__global__ __launch_bounds__(256, 3) void forward_fp8_kernel(const uint8_t* __restrict__ w8, const float* __restrict__ w_inv,
                                                          uint32_t L, const rtc8x_radiance_query* __restrict__ in,
                                                          uint32_t n, float* __restrict__ out,
                                                          const float* __restrict__ tables, const HashGeom geom,
                                                          const int32_t* __restrict__ count_dims, const QueryMask mask)
{
    using T = __bf16;
    using V = typename Vec8<T>::type;
    n = device_count(n, count_dims);
    if (blockIdx.x * 256u >= n) return;
    __shared__ __align__(16) T act[4][64 * kActStride];
    __shared__ float pred[4][64][3];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t s = (blockIdx.x * 4 + wave) * 64 + lane;
    // as in forward_kernel: slots without a query are skipped (their zero rows also leave the tile's
    // activation scales to the queries)
    const bool live = s < n && slot_has_query(mask, s);
    if (__ballot(live) == 0) return;
    T* A = act[wave];
    if constexpr (ENC == kEncHash) {
        encode_hash_row<T, false>(live ? &in[s] : nullptr, tables, geom, &A[lane * kActStride]);
    } else {
        float e[64];
        if (live) {
            encode(in[s], e);
        } else {
#pragma unroll
            for (int j = 0; j < 64; ++j) e[j] = 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 64; j += 8) {
            V v;
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = (T)e[j + k];
            *reinterpret_cast<V*>(&A[lane * kActStride + j]) = v;
        }
    }
    const uint32_t r16 = lane & 15, kq = 8 * (lane >> 4);
    // The hidden layers' activations stay in E4M3, [query][feature] with a 72-byte row, in the first
    // 4.5 KB of the wave's (then consumed) BF16 input tile; sa is their scale.
    uint8_t* A8 = reinterpret_cast<uint8_t*>(A);
    constexpr uint32_t kRow8 = 72;
    float sa = 1.0f;
    for (uint32_t l = 0; l <= L; ++l) {
        const bool last = l == L;
        const uint8_t* W = w8 + (size_t)l * kLayer;
        const int ntiles = last ? 1 : 4;
        long a[2][4];
        if (l == 0) {
            // the encoded input: the tile's largest magnitude (every lane scans its own row), then E4M3
            float m = 0.0f;
#pragma unroll
            for (int j = 0; j < 64; j += 8) {
                const V v = *reinterpret_cast<const V*>(&A[lane * kActStride + j]);
#pragma unroll
                for (int k = 0; k < 8; ++k) m = fmaxf(m, fabsf((float)v[k]));
            }
            for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
            sa = pow2_scale(m);
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    const V v = *reinterpret_cast<const V*>(&A[(mt * 16 + r16) * kActStride + ks * 32 + kq]);
                    float x[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) x[k] = (float)v[k] * sa;
                    a[ks][mt] = pack_fp8(x);
                }
        } else {
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
#pragma unroll
                for (int mt = 0; mt < 4; ++mt)
                    __builtin_memcpy(&a[ks][mt], &A8[(mt * 16 + r16) * kRow8 + ks * 32 + kq], sizeof(long));
        }
        const float unscale = w_inv[l] / sa;
        // Weights as the A operand, activations as B: the product is the layer's output transposed, so
        // acc[mt][nt][i] is feature nt*16 + 4*(lane>>4) + i of query mt*16 + (lane&15).
        f4 acc[4][4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                if (nt < ntiles) {
                    const long b = *reinterpret_cast<const long*>(&W[(nt * 16 + r16) * kWidth + ks * 32 + kq]);
#pragma unroll
                    for (int mt = 0; mt < 4; ++mt)
                        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(b, a[ks][mt], acc[mt][nt], 0, 0, 0);
                }
            }
        }
        if (!last) {
            // ReLU, the next layer's scale from the wave's largest activation, and four E4M3 features
            // per 32-bit store
            float m = 0.0f;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float y = fmaxf(acc[mt][nt][i] * unscale, 0.0f);
                        acc[mt][nt][i] = y;
                        m = fmaxf(m, y);
                    }
            for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
            sa = pow2_scale(m);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    int v = 0;
                    v = __builtin_amdgcn_cvt_pk_fp8_f32(acc[mt][nt][0] * sa, acc[mt][nt][1] * sa, v, false);
                    v = __builtin_amdgcn_cvt_pk_fp8_f32(acc[mt][nt][2] * sa, acc[mt][nt][3] * sa, v, true);
                    __builtin_memcpy(&A8[(mt * 16 + r16) * kRow8 + nt * 16 + 4 * (lane >> 4)], &v, sizeof(int));
                }
        } else if ((lane >> 4) == 0) {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int i = 0; i < 3; ++i) pred[wave][mt * 16 + r16][i] = acc[mt][0][i] * unscale;
        }
    }
    if (live) {
        out[3 * (size_t)s] = pred[wave][lane][0];
        out[3 * (size_t)s + 1] = pred[wave][lane][1];
        out[3 * (size_t)s + 2] = pred[wave][lane][2];
    }
}

// The hash tables' gradient without global atomics.  A level's entries are split in kScatterParts
// ranges and the samples in up to kScatterSlices slices (of at least kScatterMinSlice samples, so that a
// small batch still spreads over 256 workgroups); workgroup (level, range, slice) walks the slice's samples,
// adds the trilinear contributions that land in its range to a copy of the range in LDS, and writes
// the copy to the slice's partial table.  Every table float of every slice is written, so
// grad_reduce_kernel sums the partials into the gradient without clearing it first.  (Global float
// atomics, one per corner, feature and sample, ran at about 2e10 per second: 16.8 M of them made up
// four fifths of a 65,536-sample step.)
constexpr uint32_t kScatterParts = 4;
constexpr uint32_t kScatterSlices = 4;
constexpr uint32_t kScatterMinSlice = 4096;
constexpr uint32_t kScatterRange = kHashT / kScatterParts;  // entries per range (at most)
// The kernel is bound by its LDS float atomics (a third of its wave cycles wait on LDS instructions):
// 16 wavefronts per workgroup instead of 4, or two ranges of a level per workgroup with eight slices
// (half the atomic instructions per workgroup), both left a 16,384-sample step at 40 us.
constexpr uint32_t kScatterThreads = 256;
__device__ __forceinline__ uint32_t level_entries(const HashGeom& gm, uint32_t l, uint32_t p_grid)
{
    return ((l + 1 < kHashLevels ? gm.offset[l + 1] : p_grid) - gm.offset[l]) / RTC8X_NRC_HASH_FEATURES;
}

__global__ __launch_bounds__(kScatterThreads) void grid_scatter_kernel(const rtc8x_radiance_query* __restrict__ in, uint32_t n,
                                                           const float* __restrict__ dE, const HashGeom geom,
                                                           uint32_t p_grid, uint32_t slice_len, float* __restrict__ part)
{
    __shared__ __align__(16) float acc[2 * kScatterRange];
    const uint32_t lv = blockIdx.x / kScatterParts, r = blockIdx.x % kScatterParts, slice = blockIdx.y;
    const uint32_t entries = level_entries(geom, lv, p_grid);
    // A hashed level's workgroup owns a contiguous quarter of its entries (the hash spreads the samples'
    // corners evenly).  A dense level's entries follow space (x + y r + z r^2), and the cloud's samples
    // fill a few of its z-slabs: there a workgroup owns the entries e with e % kScatterParts == r, so that
    // the four share its atomics evenly (contiguous quarters left one workgroup most of level 0's work:
    // 10 of a 16,384-sample step's 91 us).
    const bool dense = geom.stride[lv] != 0u;
    const uint32_t per = (entries + kScatterParts - 1) / kScatterParts;
    const uint32_t e0 = min(entries, r * per), e1 = min(entries, e0 + per);
    const uint32_t owned = dense ? (entries - r + kScatterParts - 1) / kScatterParts : e1 - e0;
    float2* acc2 = reinterpret_cast<float2*>(acc);
    for (uint32_t i = threadIdx.x; i < owned; i += kScatterThreads) acc2[i] = float2{0.0f, 0.0f};
    __syncthreads();
    const uint32_t s1 = min(n, (slice + 1) * slice_len);
    constexpr uint32_t kBatch = 4;  // samples per lane whose loads are issued together
    for (uint32_t sb = slice * slice_len + threadIdx.x; sb < s1; sb += kBatch * kScatterThreads) {
        float p[kBatch][3];
        float2 g[kBatch];
#pragma unroll
        for (uint32_t k = 0; k < kBatch; ++k) {
            const uint32_t sm = min(sb + k * kScatterThreads, s1 - 1);
            query_position01(in[sm], p[k]);
            __builtin_memcpy(&g[k], dE + (size_t)sm * 32 + 2 * lv, sizeof(float2));
        }
#pragma unroll
        for (uint32_t k = 0; k < kBatch; ++k) {
            if (sb + k * kScatterThreads < s1) {
                uint32_t gi[3];
                float f[3];
                hash_cell(geom, lv, p[k], gi, f);
#pragma unroll
                for (uint32_t c = 0; c < 8; ++c) {
                    uint32_t idx;
                    float w;
                    hash_corner(geom, lv, gi, f, c, idx, w);
                    const uint32_t e = (idx - geom.offset[lv]) / 2u;
                    const bool mine = dense ? e % kScatterParts == r : e >= e0 && e < e1;
                    const uint32_t j = dense ? e / kScatterParts : e - e0;
                    if (mine) {
                        atomicAdd(&acc[2 * j], w * g[k].x);
                        atomicAdd(&acc[2 * j + 1], w * g[k].y);
                    }
                }
            }
        }
    }
    __syncthreads();
    float2* out = reinterpret_cast<float2*>(part + (size_t)slice * p_grid + geom.offset[lv]);
    for (uint32_t i = threadIdx.x; i < owned; i += kScatterThreads) out[dense ? i * kScatterParts + r : e0 + i] = acc2[i];
}

template <typename T, int ENC, int WPB>
__global__ __launch_bounds__(64 * WPB) void backward_kernel(const T* __restrict__ wt, uint32_t L, uint32_t n,
                                                            const T* __restrict__ hbuf, const T* __restrict__ dout,
                                                            float* __restrict__ gpart, float* __restrict__ dEout)
{
    using V = typename Vec8<T>::type;
    constexpr uint32_t SB = 64 * WPB;      // samples per workgroup
    constexpr uint32_t CS = SB + 8;        // LDS row stride of a feature-major tile (elements)
    // FP32 with four waves keeps dZ only feature-major, so that the three tiles fit the LDS
    constexpr bool kSampleMajor = sizeof(T) == 2 || WPB == 1;
    __shared__ __align__(16) T dzs[kSampleMajor ? SB * kActStride : 8];   // dZ, sample-major [sample][out]
    __shared__ __align__(16) T dzf[64 * CS];                              // dZ, feature-major [out][sample]
    __shared__ __align__(16) T hf[64 * CS];     // layer input H, feature-major [in][sample]
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint32_t r16 = lane & 15, kq = 8 * (lane >> 4);
    const uint32_t b0 = blockIdx.x * SB;
    const uint32_t tiles = (n + 63) / 64;
    const uint32_t P = num_params(L);
    float* gblk = gpart + (size_t)blockIdx.x * P;
    {
        // the output layer's dZ: thread tid holds sample b0 + tid; columns kOutPad..31 are the zero K
        // padding of the output layer's dH product (columns 32..63 and rows 32..63 are written before
        // they are read)
        T g[kOutPad];
#pragma unroll
        for (uint32_t c = 0; c < kOutPad; ++c) g[c] = (T)0.0f;
        if (b0 + tid < n) __builtin_memcpy(g, dout + (size_t)(b0 + tid) * kOutPad, sizeof(g));
        if constexpr (kSampleMajor) {
#pragma unroll
            for (uint32_t c = 0; c < 32; ++c) dzs[tid * kActStride + c] = c < kOutPad ? g[c] : (T)0.0f;
        }
#pragma unroll
        for (uint32_t c = 0; c < 32; ++c) dzf[c * CS + tid] = c < kOutPad ? g[c] : (T)0.0f;
    }
    __syncthreads();
    for (int l = (int)L; l >= 0; --l) {
        const bool top = (uint32_t)l == L;
        // H_l of the workgroup's tiles, 16 bytes per copy (zeros for tiles past the batch)
        for (uint32_t i = tid; i < SB * 8; i += SB) {
            const uint32_t t = i >> 9, c = (i >> 3) & 63u, j8 = 8 * (i & 7u);
            const uint32_t tile = b0 / 64 + t;
            V v;
            if (tile < tiles) {
                v = *reinterpret_cast<const V*>(hbuf + saved_tile((uint32_t)l, n, tile) + c * 64 + j8);
            } else {
#pragma unroll
                for (int k = 0; k < 8; ++k) v[k] = (T)0.0f;
            }
            *reinterpret_cast<V*>(&hf[c * CS + t * 64 + j8]) = v;
        }
        __syncthreads();
        // dW_l[out][in] = sum over the block's samples of dZ[s][out] H[s][in]
        const int otiles = top ? 1 : 4;
        const uint32_t goff = (uint32_t)l * kLayer;
        for (int t = (int)wave; t < otiles * 4; t += WPB) {
            const int mt = t >> 2, nt = t & 3;
            f4 acc = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < (int)(SB / 32); ++ks) {
                const V a = *reinterpret_cast<const V*>(&dzf[(mt * 16 + r16) * CS + ks * 32 + kq]);
                const V b = *reinterpret_cast<const V*>(&hf[(nt * 16 + r16) * CS + ks * 32 + kq]);
                acc = mfma(a, b, acc);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) gblk[goff + (mt * 16 + 4 * (lane >> 4) + i) * kWidth + nt * 16 + r16] = acc[i];
        }
        if (l == 0 && ENC != kEncHash) break;
        // dH_l = dZ_l W_l for this wavefront's 64 samples (K = 16, padded to one 32-step, for the top layer);
        // for l = 0 with the hash grid, the gradient of the encoded input
        const T* WT = wt + (size_t)l * kLayer;
        const uint32_t ldk = top ? kOutT : kWidth;
        const int ksteps = top ? 1 : 2;
        f4 acc[4][4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int ks = 0; ks < ksteps; ++ks) {
            V a[4];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                if constexpr (kSampleMajor) {
                    a[mt] = *reinterpret_cast<const V*>(&dzs[(wave * 64 + mt * 16 + r16) * kActStride + ks * 32 + kq]);
                } else {
#pragma unroll
                    for (int j = 0; j < 8; ++j) a[mt][j] = dzf[(ks * 32 + kq + j) * CS + wave * 64 + mt * 16 + r16];
                }
            }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const V b = *reinterpret_cast<const V*>(&WT[(nt * 16 + r16) * ldk + ks * 32 + kq]);
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) acc[mt][nt] = mfma(a[mt], b, acc[mt][nt]);
            }
        }
        __syncthreads();
        if (ENC == kEncHash && l == 0) {
            // dE for the 32 grid features, FP32, sample-major in the (now free) H tile
            float* dE = reinterpret_cast<float*>(hf);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        dE[(wave * 64 + mt * 16 + 4 * (lane >> 4) + i) * 32 + nt * 16 + r16] = acc[mt][nt][i];
            __syncthreads();
            // the rows of dE go out for grid_scatter_kernel, 16 bytes per lane and store
            for (uint32_t i = tid; i < SB * 8; i += SB) {
                const uint32_t sm = i >> 3;
                if (b0 + sm < n)
                    *reinterpret_cast<f4*>(dEout + (size_t)(b0 + sm) * 32 + 4 * (i & 7u)) =
                        *reinterpret_cast<const f4*>(dE + sm * 32 + 4 * (i & 7u));
            }
            break;
        }
        // ReLU mask: H_l is the post-ReLU output of layer l-1, so the derivative is [H_l > 0]
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t s = wave * 64 + mt * 16 + 4 * (lane >> 4) + i, c = nt * 16 + r16;
                    const float hv = (float)hf[c * CS + s];
                    const T v = (T)(hv > 0.0f ? acc[mt][nt][i] : 0.0f);
                    if constexpr (kSampleMajor) dzs[s * kActStride + c] = v;
                    dzf[c * CS + s] = v;
                }
        __syncthreads();
    }
}

// The gradient from its partials, and the finiteness check of the whole gradient.  Workgroups below
// mlp_groups reduce 64 MLP parameters each: lane (i, c) sums the per-workgroup partials of blocks
// c, c + 16, c + 32, ... (four loads in flight), and the 16 slices are added in a fixed order, so the
// result does not depend on scheduling.  The other workgroups sum 1,024 hash-table gradients each over
// the scatter's slices (entries P .. P_total; none without the hash grid).  Any non-finite sum sets *bad.
constexpr uint32_t kReduceSlices = 16;
constexpr uint32_t kReduceThreads = 64 * kReduceSlices;
__global__ __launch_bounds__(kReduceThreads) void grad_reduce_kernel(const float* __restrict__ gpart, uint32_t nblocks, uint32_t P,
                                                          uint32_t P_total, const float* __restrict__ grid_part,
                                                          uint32_t slices, float* __restrict__ grad,
                                                          uint32_t* __restrict__ bad)
{
    const uint32_t mlp_groups = (P + 63) / 64;
    bool nonfinite = false;
    if (blockIdx.x < mlp_groups) {
        __shared__ float part[kReduceSlices][64];
        const uint32_t il = threadIdx.x & 63u, c = threadIdx.x >> 6, i = blockIdx.x * 64 + il;
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
        if (i < P) {
            for (uint32_t b = c; b < nblocks; b += 4 * kReduceSlices) {
                a0 += gpart[(size_t)b * P + i];
                if (b + kReduceSlices < nblocks) a1 += gpart[(size_t)(b + kReduceSlices) * P + i];
                if (b + 2 * kReduceSlices < nblocks) a2 += gpart[(size_t)(b + 2 * kReduceSlices) * P + i];
                if (b + 3 * kReduceSlices < nblocks) a3 += gpart[(size_t)(b + 3 * kReduceSlices) * P + i];
            }
        }
        part[c][il] = (a0 + a1) + (a2 + a3);
        __syncthreads();
        if (c == 0 && i < P) {
            float g = part[0][il];
#pragma unroll
            for (uint32_t k = 1; k < kReduceSlices; ++k) g += part[k][il];
            grad[i] = g;
            nonfinite = !isfinite(g);
        }
    } else {
        const size_t j = (size_t)(blockIdx.x - mlp_groups) * kReduceThreads + threadIdx.x, p_grid = P_total - P;
        if (j < p_grid) {
            float sum = grid_part[j];
            for (uint32_t h = 1; h < slices; ++h) sum += grid_part[h * p_grid + j];
            grad[P + j] = sum;
            nonfinite = !isfinite(sum);
        }
    }
    if (__ballot(nonfinite) != 0 && (threadIdx.x & 63u) == 0) atomicOr(bad, 1u);
}

struct AdamArgs {
    float lr, beta1, beta2, eps, l2, decay;
    float bc1, bc2;       // 1 - beta^t
    float ema_bc;         // 1 - decay^t
    uint32_t P, L;        // P = all parameters (MLP, then hash tables)
    uint32_t P_mlp;
};

// Dynamic loss scaling (Micikevicius et al. 2018, "Mixed Precision Training"), in every precision:
// a state is [scale, finite steps since it last changed], the scale a power of two.  A step whose
// gradient is not finite (or whose scaled loss gradient left FP16's range) was skipped by the
// optimiser: the scale halves (not below 1).  After kLossScaleGrowth finite steps in a row it doubles
// (not above kLossScaleMax).  The state and the non-finite flag are double-buffered by step parity:
// step k reads state[k % 2] and flag[k % 2], and lane 0 of its optimiser kernel writes
// state[(k + 1) % 2] and clears flag[(k + 1) % 2], which no kernel of step k reads (a one-lane kernel
// after the optimiser took 5 us per step).
constexpr float kLossScaleMax = 16777216.0f;  // 2^24
constexpr float kLossScaleGrowth = 128.0f;
__device__ __forceinline__ void next_loss_scale(bool skipped, const float* __restrict__ cur, float* __restrict__ next)
{
    if (skipped) {
        next[0] = fmaxf(1.0f, cur[0] * 0.5f);
        next[1] = 0.0f;
    } else if (cur[1] + 1.0f >= kLossScaleGrowth) {
        next[0] = fminf(kLossScaleMax, cur[0] * 2.0f);
        next[1] = 0.0f;
    } else {
        next[0] = cur[0];
        next[1] = cur[1] + 1.0f;
    }
}

template <typename T>
__global__ void adam_kernel(const AdamArgs a, float* __restrict__ master, float* __restrict__ m, float* __restrict__ v,
                            float* __restrict__ ema, const float* __restrict__ grad, uint32_t* __restrict__ bad,
                            T* __restrict__ w, T* __restrict__ wt, T* __restrict__ winf, float* __restrict__ grid_inf,
                            const float* __restrict__ loss_scale, float* __restrict__ next_scale,
                            uint32_t* __restrict__ next_bad)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        next_loss_scale(*bad != 0u, loss_scale, next_scale);
        *next_bad = 0u;
    }
    if (i >= a.P) return;
    if (*bad) return;  // a non-finite gradient skips the whole step
    const float w0 = master[i];
    const float g = grad[i] / *loss_scale + a.l2 * w0;  // exact: the scale is a power of two
    const float mi = a.beta1 * m[i] + (1.0f - a.beta1) * g;
    const float vi = a.beta2 * v[i] + (1.0f - a.beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float w1 = w0 - a.lr * (mi / a.bc1) / (sqrtf(vi / a.bc2) + a.eps);
    master[i] = w1;
    const float e = a.decay * ema[i] + (1.0f - a.decay) * w1;
    ema[i] = e;
    if (i >= a.P_mlp) {  // hash table entry: the FP32 master is the training table
        grid_inf[i - a.P_mlp] = e / a.ema_bc;
        return;
    }
    w[i] = (T)w1;
    winf[i] = (T)(e / a.ema_bc);
    const uint32_t l = i / kLayer, r = i % kLayer, out = r / kWidth, in = r % kWidth;
    if (l < a.L)
        wt[(size_t)l * kLayer + in * kWidth + out] = (T)w1;
    else
        wt[(size_t)a.L * kLayer + in * kOutT + out] = (T)w1;
}

// g / loss scale, for the gradient handed out by rtc8x_nrc_gradient
__global__ void unscale_kernel(const float* __restrict__ g, uint32_t P, const float* __restrict__ state,
                               float* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P) out[i] = g[i] / state[0];
}

template <typename T>
__global__ void convert_kernel(const float* __restrict__ src, uint32_t P, uint32_t L, T* __restrict__ w, T* __restrict__ wt,
                               T* __restrict__ winf, uint32_t P_total, float* __restrict__ grid_inf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P_total) return;
    if (i >= P) {
        grid_inf[i - P] = src[i];
        return;
    }
    const T x = (T)src[i];
    w[i] = x;
    winf[i] = x;
    const uint32_t l = i / kLayer, r = i % kLayer, out = r / kWidth, in = r % kWidth;
    if (l < L)
        wt[(size_t)l * kLayer + in * kWidth + out] = x;
    else
        wt[(size_t)L * kLayer + in * kOutT + out] = x;
}

template <int ENC>
__global__ void encode_kernel(const rtc8x_radiance_query* __restrict__ in, uint32_t n, float* __restrict__ out,
                              const float* __restrict__ tables, const HashGeom geom)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float e[64];
    encode_any<ENC>(in[s], tables, geom, e);
    for (int j = 0; j < 64; ++j) out[(size_t)s * 64 + j] = e[j];
}

}  // namespace nrc
}  // namespace rtc8x

using namespace rtc8x;

// Parameter vector: the MLP's weights (P_mlp, [out][in] per layer), then the hash tables (P_grid floats,
// zero without the hash grid).  master / m / v / ema / grad hold all P = P_mlp + P_grid.
struct rtc8x_nrc {
    rtc8x_nrc_desc desc;
    uint32_t L = 0, P = 0, P_mlp = 0, P_grid = 0, PT = 0;
    int enc = RTC8X_NRC_TRIANGLE_WAVE;
    nrc::HashGeom geom{};
    size_t tsize = 2;
    uint64_t step = 0;
    float *master = nullptr, *m = nullptr, *v = nullptr, *ema = nullptr, *grad = nullptr, *gpart = nullptr;
    float* loss_part = nullptr;
    float* grid_inf = nullptr;  // EMA (debiased) hash tables for inference
    float* grid_de = nullptr;   // hash grid: the encoded input's gradient, 32 floats per training sample
    float* grid_part = nullptr; // hash grid: the tables' gradient per slice of the batch (kScatterSlices)
    uint32_t* bad = nullptr;
    float* loss_scale = nullptr;  // two [scale, finite steps since it changed] states, by step parity
    void *w = nullptr, *wt = nullptr, *winf = nullptr, *hbuf = nullptr, *dout = nullptr;
    uint8_t* w8 = nullptr;      // E4M3 inference weights (RTC8X_NRC_BF16_FP8)
    float* w8_inv = nullptr;    // per layer: 1 / the layer's weight scale
};

namespace {

// The loss-scale state the next training step reads (by step parity)
const float* scale_state(const rtc8x_nrc* h) { return h->loss_scale + 2 * (h->step & 1u); }

// Waves per workgroup of the training kernels for a batch of n samples (the loss parts and the
// per-workgroup weight gradients follow it: one loss part per wave, one gradient slot per workgroup).
uint32_t train_waves_per_block(uint32_t n) { return n <= nrc::kSmallBatch ? 1u : 4u; }
uint32_t train_blocks(uint32_t n) { return (n + 64 * train_waves_per_block(n) - 1) / (64 * train_waves_per_block(n)); }
// Sample slices of the hash tables' gradient scatter for a batch of n samples.
uint32_t scatter_slices(uint32_t n)
{
    return std::min(nrc::kScatterSlices, (n + nrc::kScatterMinSlice - 1) / nrc::kScatterMinSlice);
}

template <typename T, int ENC>
hipError_t launch_forward_enc(rtc8x_nrc* h, bool train, const void* w, const float* tables, hipStream_t s, const float* in,
                              uint32_t n, float* out, const float* target, const int32_t* count_dims,
                              const nrc::QueryMask& mask)
{
    if (train && train_waves_per_block(n) == 1)
        hipLaunchKernelGGL((nrc::forward_kernel<T, true, ENC, 1>), dim3(train_blocks(n)), dim3(64), 0, s, (const T*)w,
                           h->L, (const rtc8x_radiance_query*)in, n, out, target, (T*)h->hbuf, (T*)h->dout, h->loss_part,
                           tables, h->geom, (const int32_t*)nullptr, scale_state(h), nrc::QueryMask{});
    else if (train)
        hipLaunchKernelGGL((nrc::forward_kernel<T, true, ENC, 4>), dim3(train_blocks(n)), dim3(256), 0, s, (const T*)w,
                           h->L, (const rtc8x_radiance_query*)in, n, out, target, (T*)h->hbuf, (T*)h->dout, h->loss_part,
                           tables, h->geom, (const int32_t*)nullptr, scale_state(h), nrc::QueryMask{});
    else
        hipLaunchKernelGGL((nrc::forward_kernel<T, false, ENC, 4>), dim3((n + 255) / 256), dim3(256), 0, s, (const T*)w,
                           h->L, (const rtc8x_radiance_query*)in, n, out, target, (T*)nullptr, (T*)nullptr,
                           (float*)nullptr, tables, h->geom, count_dims, scale_state(h), mask);
    return hipGetLastError();
}

// train: the working weights and the FP32 master tables; inference: the EMA weights and tables
template <typename T>
hipError_t launch_forward(rtc8x_nrc* h, bool train, hipStream_t s, const float* in, uint32_t n, float* out,
                          const float* target, const int32_t* count_dims = nullptr,
                          const nrc::QueryMask& mask = nrc::QueryMask{})
{
    const void* w = train ? h->w : h->winf;
    const float* tables = train ? h->master + h->P_mlp : h->grid_inf;
    if (h->enc == RTC8X_NRC_HASH_GRID)
        return launch_forward_enc<T, RTC8X_NRC_HASH_GRID>(h, train, w, tables, s, in, n, out, target, count_dims, mask);
    return launch_forward_enc<T, RTC8X_NRC_TRIANGLE_WAVE>(h, train, w, tables, s, in, n, out, target, count_dims, mask);
}

template <typename T>
hipError_t launch_backward(rtc8x_nrc* h, hipStream_t s, const float* in, uint32_t n)
{
    const dim3 blocks(train_blocks(n));
    const bool one = train_waves_per_block(n) == 1;
    if (h->enc != RTC8X_NRC_HASH_GRID) {
        if (one)
            hipLaunchKernelGGL((nrc::backward_kernel<T, RTC8X_NRC_TRIANGLE_WAVE, 1>), blocks, dim3(64), 0, s,
                               (const T*)h->wt, h->L, n, (const T*)h->hbuf, (const T*)h->dout, h->gpart, (float*)nullptr);
        else
            hipLaunchKernelGGL((nrc::backward_kernel<T, RTC8X_NRC_TRIANGLE_WAVE, 4>), blocks, dim3(256), 0, s,
                               (const T*)h->wt, h->L, n, (const T*)h->hbuf, (const T*)h->dout, h->gpart, (float*)nullptr);
        return hipGetLastError();
    }
    if (one)
        hipLaunchKernelGGL((nrc::backward_kernel<T, RTC8X_NRC_HASH_GRID, 1>), blocks, dim3(64), 0, s, (const T*)h->wt,
                           h->L, n, (const T*)h->hbuf, (const T*)h->dout, h->gpart, h->grid_de);
    else
        hipLaunchKernelGGL((nrc::backward_kernel<T, RTC8X_NRC_HASH_GRID, 4>), blocks, dim3(256), 0, s, (const T*)h->wt,
                           h->L, n, (const T*)h->hbuf, (const T*)h->dout, h->gpart, h->grid_de);
    const uint32_t slices = scatter_slices(n);
    const uint32_t slice_len = (n + slices - 1) / slices;
    hipLaunchKernelGGL(nrc::grid_scatter_kernel, dim3(nrc::kHashLevels * nrc::kScatterParts, slices),
                       dim3(nrc::kScatterThreads), 0, s,
                       (const rtc8x_radiance_query*)in, n, (const float*)h->grid_de, h->geom, h->P_grid, slice_len,
                       h->grid_part);
    return hipGetLastError();
}

// step parity par: reads state and flag par, writes state and clears flag par ^ 1
template <typename T>
hipError_t launch_adam(rtc8x_nrc* h, hipStream_t s, const nrc::AdamArgs& a, uint32_t par)
{
    hipLaunchKernelGGL((nrc::adam_kernel<T>), dim3((h->P + 255) / 256), dim3(256), 0, s, a, h->master, h->m, h->v, h->ema,
                       h->grad, h->bad + par, (T*)h->w, (T*)h->wt, (T*)h->winf, h->grid_inf,
                       (const float*)(h->loss_scale + 2 * par), h->loss_scale + 2 * (par ^ 1u), h->bad + (par ^ 1u));
    return hipGetLastError();
}

template <typename T>
hipError_t launch_convert(rtc8x_nrc* h, hipStream_t s)
{
    hipLaunchKernelGGL((nrc::convert_kernel<T>), dim3((h->P + 255) / 256), dim3(256), 0, s, h->master, h->P_mlp, h->L,
                       (T*)h->w, (T*)h->wt, (T*)h->winf, h->P, h->grid_inf);
    return hipGetLastError();
}

// Calls f with a null pointer of the network's working type: FP16, FP32, or BF16 (also under FP8
// inference, whose training runs in BF16).
template <typename F>
hipError_t with_type(const rtc8x_nrc* h, F&& f)
{
    switch (h->desc.precision) {
    case RTC8X_NRC_FP16: return f(static_cast<_Float16*>(nullptr));
    case RTC8X_NRC_FP32: return f(static_cast<float*>(nullptr));
    default: return f(static_cast<__bf16*>(nullptr));
    }
}
#define RTC8X_WITH_TYPE(h, call) with_type(h, [&](auto tag) { using T = std::remove_pointer_t<decltype(tag)>; return call; })
bool is_fp8(const rtc8x_nrc* h) { return h->desc.precision == RTC8X_NRC_BF16_FP8; }

hipError_t launch_forward_fp8(rtc8x_nrc* h, hipStream_t s, const float* in, uint32_t n, float* out,
                              const int32_t* count_dims = nullptr, const nrc::QueryMask& mask = nrc::QueryMask{})
{
    hipLaunchKernelGGL(nrc::quantize_kernel, dim3(h->L + 1), dim3(256), 0, s, (const __bf16*)h->winf, h->L, h->w8,
                       h->w8_inv);
    const uint32_t blocks = (n + 255) / 256;
    if (h->enc == RTC8X_NRC_HASH_GRID)
        hipLaunchKernelGGL((nrc::forward_fp8_kernel<RTC8X_NRC_HASH_GRID>), dim3(blocks), dim3(256), 0, s, h->w8,
                           h->w8_inv, h->L, (const rtc8x_radiance_query*)in, n, out, h->grid_inf, h->geom, count_dims,
                           mask);
    else
        hipLaunchKernelGGL((nrc::forward_fp8_kernel<RTC8X_NRC_TRIANGLE_WAVE>), dim3(blocks), dim3(256), 0, s, h->w8,
                           h->w8_inv, h->L, (const rtc8x_radiance_query*)in, n, out, h->grid_inf, h->geom, count_dims,
                           mask);
    return hipGetLastError();
}

rtc8x_status reset_from_master(rtc8x_nrc* h)
{
    RTC8X_HIP(hipMemset(h->m, 0, sizeof(float) * h->P));
    RTC8X_HIP(hipMemset(h->v, 0, sizeof(float) * h->P));
    RTC8X_HIP(hipMemset(h->ema, 0, sizeof(float) * h->P));
    RTC8X_HIP(hipMemset(h->bad, 0, 2 * sizeof(uint32_t)));
    RTC8X_HIP(hipMemset(h->wt, 0, h->tsize * h->PT));
    RTC8X_HIP(RTC8X_WITH_TYPE(h, launch_convert<T>(h, nullptr)));
    const float s0 = h->desc.loss_scale > 0.0f ? h->desc.loss_scale : 128.0f;
    const float ls[4] = {s0, 0.0f, s0, 0.0f};
    RTC8X_HIP(hipMemcpy(h->loss_scale, ls, sizeof(ls), hipMemcpyHostToDevice));
    RTC8X_HIP(hipDeviceSynchronize());
    h->step = 0;
    return RTC8X_OK;
}

// forward (train) + backward + ordered reduction into h->grad; returns the loss sum in *loss_sum when set
rtc8x_status gradient_pass(rtc8x_nrc* h, hipStream_t s, const float* in, const float* target, uint32_t n)
{
    RTC8X_HIP(RTC8X_WITH_TYPE(h, launch_forward<T>(h, true, s, in, n, nullptr, target)));
    RTC8X_HIP(RTC8X_WITH_TYPE(h, launch_backward<T>(h, s, in, n)));
    // the step's flag is zero here: set so at creation, cleared by the previous step's optimiser kernel
    // (and after a gradient handed out by rtc8x_nrc_gradient)
    const uint32_t groups = (h->P_mlp + 63) / 64 + (h->P - h->P_mlp + nrc::kReduceThreads - 1) / nrc::kReduceThreads;
    hipLaunchKernelGGL(nrc::grad_reduce_kernel, dim3(groups), dim3(nrc::kReduceThreads), 0, s, h->gpart, train_blocks(n), h->P_mlp, h->P,
                       (const float*)h->grid_part, scatter_slices(n), h->grad, h->bad + (h->step & 1u));
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}

}  // namespace

extern "C" rtc8x_status rtc8x_nrc_destroy(rtc8x_nrc* h)
try {
    if (!h) return RTC8X_OK;
    void* ptrs[] = {h->master, h->m, h->v, h->ema, h->grad, h->gpart, h->loss_part, h->bad, h->w, h->wt, h->winf, h->hbuf, h->dout,
                    h->grid_inf, h->grid_de, h->grid_part, h->w8, h->w8_inv, h->loss_scale};
    for (void* p : ptrs)
        if (p) RTC8X_HIP(hipFree(p));
    delete h;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_destroy")

extern "C" rtc8x_status rtc8x_nrc_create(const rtc8x_nrc_desc* d, rtc8x_nrc** out)
try {
    RTC8X_REQUIRE(d && out, "rtc8x_nrc_create: NULL argument");
    RTC8X_REQUIRE(d->num_hidden_layers >= 1 && d->num_hidden_layers <= nrc::kMaxLayers,
                  "rtc8x_nrc_create: num_hidden_layers must be in [1, %u]", nrc::kMaxLayers);
    RTC8X_REQUIRE(d->precision == RTC8X_NRC_FP16 || d->precision == RTC8X_NRC_BF16 || d->precision == RTC8X_NRC_BF16_FP8 ||
                      d->precision == RTC8X_NRC_FP32,
                  "rtc8x_nrc_create: bad precision");
    RTC8X_REQUIRE(d->learning_rate > 0.0f && d->beta1 >= 0.0f && d->beta1 < 1.0f && d->beta2 >= 0.0f && d->beta2 < 1.0f &&
                      d->epsilon > 0.0f && d->ema_decay >= 0.0f && d->ema_decay < 1.0f && d->l2_reg >= 0.0f,
                  "rtc8x_nrc_create: bad optimiser parameters");
    RTC8X_REQUIRE(d->loss_scale == 0.0f || (d->loss_scale >= 1.0f && d->loss_scale <= nrc::kLossScaleMax &&
                                             std::exp2(std::floor(std::log2(d->loss_scale))) == d->loss_scale),
                  "rtc8x_nrc_create: loss_scale must be 0 or a power of two in [1, 2^24]");
    RTC8X_REQUIRE(d->encoding == RTC8X_NRC_TRIANGLE_WAVE || d->encoding == RTC8X_NRC_HASH_GRID,
                  "rtc8x_nrc_create: unknown encoding %d", d->encoding);
    *out = nullptr;
    rtc8x_nrc* h = new rtc8x_nrc();
    h->desc = *d;
    h->tsize = d->precision == RTC8X_NRC_FP32 ? 4 : 2;
    h->enc = d->encoding;
    h->L = d->num_hidden_layers;
    h->P_mlp = nrc::num_params(h->L);
    h->P_grid = h->enc == RTC8X_NRC_HASH_GRID ? nrc::hash_geometry(&h->geom) : 0u;
    h->P = h->P_mlp + h->P_grid;
    h->PT = nrc::num_params_t(h->L);
    const size_t P = h->P, P_mlp = h->P_mlp;
    const size_t max_blocks = std::max(nrc::kSmallBatch / 64, nrc::kMaxTrain / 256);  // gradient slots
    const size_t max_waves = nrc::kMaxTrain / 64;                                    // loss parts
    auto alloc = [&](void** p, size_t bytes) -> bool { return hipMalloc(p, bytes) == hipSuccess; };
    bool ok = alloc((void**)&h->master, 4 * P) && alloc((void**)&h->m, 4 * P) && alloc((void**)&h->v, 4 * P) &&
              alloc((void**)&h->ema, 4 * P) && alloc((void**)&h->grad, 4 * P) &&
              alloc((void**)&h->gpart, 4 * P_mlp * max_blocks) && alloc((void**)&h->loss_part, 4 * max_waves) &&
              alloc((void**)&h->bad, 8) && alloc((void**)&h->loss_scale, 16) && alloc(&h->w, h->tsize * P_mlp) && alloc(&h->wt, h->tsize * h->PT) &&
              alloc(&h->winf, h->tsize * P_mlp) && (h->P_grid == 0 || alloc((void**)&h->grid_inf, 4 * (size_t)h->P_grid)) &&
              (h->P_grid == 0 || alloc((void**)&h->grid_de, 4 * 32 * (size_t)nrc::kMaxTrain)) &&
              (h->P_grid == 0 || alloc((void**)&h->grid_part, 4 * (size_t)h->P_grid * nrc::kScatterSlices)) &&
              alloc(&h->hbuf, h->tsize * 64 * (size_t)(h->L + 1) * nrc::kMaxTrain) &&
              alloc(&h->dout, h->tsize * nrc::kOutPad * (size_t)nrc::kMaxTrain) &&
              (d->precision != RTC8X_NRC_BF16_FP8 ||
               (alloc((void**)&h->w8, P_mlp) && alloc((void**)&h->w8_inv, 4 * (size_t)(h->L + 1))));
    if (!ok) {
        rtc8x_nrc_destroy(h);
        return rtc8x::fail(RTC8X_ERR_OUT_OF_MEMORY, "rtc8x_nrc_create: out of device memory");
    }
    // Xavier-uniform initialisation from a PCG32 stream (tiny-cuda-nn's initialiser is not available).
    std::vector<float> init(P, 0.0f);
    uint64_t state = d->seed;
    auto uni = [&]() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + 1442695040888963407ULL;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u), rot = (uint32_t)(old >> 59u);
        uint32_t r = (xs >> rot) | (xs << ((0u - rot) & 31u));
        return (double)(r >> 8) * (1.0 / 16777216.0);
    };
    for (uint32_t l = 0; l <= h->L; ++l) {
        const bool last = l == h->L;
        const uint32_t rows = last ? RTC8X_NRC_OUTPUT_DIMS : nrc::kWidth;
        const double lim = std::sqrt(6.0 / (64.0 + (double)rows));
        for (uint32_t r = 0; r < rows; ++r)
            for (uint32_t c = 0; c < nrc::kWidth; ++c) init[(size_t)l * nrc::kLayer + r * nrc::kWidth + c] = (float)((2.0 * uni() - 1.0) * lim);
    }
    // hash tables: uniform in [-1e-4, 1e-4] (the initialisation the hash-encoding paper gives)
    for (size_t i = P_mlp; i < P; ++i) init[i] = (float)((2.0 * uni() - 1.0) * 1e-4);
    hipError_t e = hipMemcpy(h->master, init.data(), 4 * P, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        rtc8x_nrc_destroy(h);
        return rtc8x::fail(RTC8X_ERR_HIP, "rtc8x_nrc_create: %s", hipGetErrorString(e));
    }
    rtc8x_status st = reset_from_master(h);
    if (st != RTC8X_OK) {
        rtc8x_nrc_destroy(h);
        return st;
    }
    *out = h;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_create")

extern "C" uint32_t rtc8x_nrc_num_params(const rtc8x_nrc* h) { return h ? h->P : 0; }

extern "C" rtc8x_status rtc8x_nrc_get_loss_scale(const rtc8x_nrc* h, float* scale)
try {
    RTC8X_REQUIRE(h && scale, "rtc8x_nrc_get_loss_scale: NULL argument");
    RTC8X_HIP(hipDeviceSynchronize());
    RTC8X_HIP(hipMemcpy(scale, scale_state(h), sizeof(float), hipMemcpyDeviceToHost));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_get_loss_scale")

extern "C" rtc8x_status rtc8x_nrc_get_weights(const rtc8x_nrc* h, int32_t which, float* host, size_t count)
try {
    RTC8X_REQUIRE(h && host && count == h->P, "rtc8x_nrc_get_weights: bad arguments");
    if (which == 0 || h->step == 0) {
        RTC8X_HIP(hipMemcpy(host, h->master, 4 * h->P, hipMemcpyDeviceToHost));
        return RTC8X_OK;
    }
    RTC8X_HIP(hipMemcpy(host, h->ema, 4 * h->P, hipMemcpyDeviceToHost));
    const double bc = 1.0 - std::pow((double)h->desc.ema_decay, (double)h->step);
    const float fbc = (float)bc;
    for (size_t i = 0; i < count; ++i) host[i] = host[i] / fbc;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_get_weights")

extern "C" rtc8x_status rtc8x_nrc_set_weights(rtc8x_nrc* h, const float* host, size_t count)
try {
    RTC8X_REQUIRE(h && host && count == h->P, "rtc8x_nrc_set_weights: bad arguments");
    RTC8X_HIP(hipMemcpy(h->master, host, 4 * h->P, hipMemcpyHostToDevice));
    return reset_from_master(h);
}
RTC8X_CATCH_ALL("rtc8x_nrc_set_weights")

extern "C" rtc8x_status rtc8x_nrc_infer(rtc8x_nrc* h, void* stream, const float* in, uint32_t n, float* out)
try {
    RTC8X_REQUIRE(h && in && out, "rtc8x_nrc_infer: NULL argument");
    if (n == 0) return RTC8X_OK;
    hipStream_t s = (hipStream_t)stream;
    if (is_fp8(h))
        RTC8X_HIP(launch_forward_fp8(h, s, in, n, out));
    else
        RTC8X_HIP(RTC8X_WITH_TYPE(h, launch_forward<T>(h, false, s, in, n, out, nullptr)));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_infer")

extern "C" rtc8x_status rtc8x_nrc_infer_training(rtc8x_nrc* h, void* stream, const rtc8x_frame_params* p)
try {
    RTC8X_REQUIRE(h && p, "rtc8x_nrc_infer_training: NULL argument");
    const rtc8x_nrc_buffers& nb = p->nrc;
    RTC8X_REQUIRE(nb.inference_queries && nb.inferred_radiance && nb.train_image_size && nb.nrc_buffer_index < 2 &&
                      p->image_width > 0 && p->image_height > 0,
                  "rtc8x_nrc_infer_training: NRC buffers missing");
    // at most the image's queries; the training size itself is read on the device
    const uint32_t n = (uint32_t)(((size_t)p->image_width * p->image_height + 127u) & ~(size_t)127u);
    const int32_t* dims = nb.train_image_size + 2 * nb.nrc_buffer_index;
    hipStream_t s = (hipStream_t)stream;
    const float* in = reinterpret_cast<const float*>(nb.inference_queries);
    if (is_fp8(h))
        RTC8X_HIP(launch_forward_fp8(h, s, in, n, nb.inferred_radiance, dims));
    else
        RTC8X_HIP(RTC8X_WITH_TYPE(h, launch_forward<T>(h, false, s, in, n, nb.inferred_radiance, nullptr, dims)));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_infer_training")

extern "C" rtc8x_status rtc8x_nrc_infer_render(rtc8x_nrc* h, void* stream, const rtc8x_frame_params* p)
try {
    RTC8X_REQUIRE(h && p, "rtc8x_nrc_infer_render: NULL argument");
    const rtc8x_nrc_buffers& nb = p->nrc;
    RTC8X_REQUIRE(nb.inference_queries && nb.inferred_radiance && nb.inference_terminals && p->image_width > 0 &&
                      p->image_height > 0,
                  "rtc8x_nrc_infer_render: NRC buffers missing");
    const uint32_t npx = p->image_width * p->image_height;
    const uint32_t q_stride = (npx + 127u) & ~127u, k = nb.num_samples ? nb.num_samples : 1u;
    RTC8X_REQUIRE((uint64_t)q_stride * k <= 0xFFFFFFFFull, "rtc8x_nrc_infer_render: too many queries");
    const nrc::QueryMask mask{nb.inference_terminals, npx, q_stride};
    hipStream_t s = (hipStream_t)stream;
    const float* in = reinterpret_cast<const float*>(nb.inference_queries);
    if (is_fp8(h))
        RTC8X_HIP(launch_forward_fp8(h, s, in, q_stride * k, nb.inferred_radiance, nullptr, mask));
    else
        RTC8X_HIP(RTC8X_WITH_TYPE(h, launch_forward<T>(h, false, s, in, q_stride * k, nb.inferred_radiance, nullptr,
                                                       nullptr, mask)));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_infer_render")

extern "C" rtc8x_status rtc8x_nrc_gradient(rtc8x_nrc* h, void* stream, const float* in, const float* target, uint32_t n,
                                           float* grad)
try {
    RTC8X_REQUIRE(h && in && target && grad, "rtc8x_nrc_gradient: NULL argument");
    RTC8X_REQUIRE(n > 0 && n <= nrc::kMaxTrain, "rtc8x_nrc_gradient: n must be in [1, %u]", nrc::kMaxTrain);
    hipStream_t s = (hipStream_t)stream;
    rtc8x_status st = gradient_pass(h, s, in, target, n);
    if (st != RTC8X_OK) return st;
    hipLaunchKernelGGL(nrc::unscale_kernel, dim3((h->P + 255) / 256), dim3(256), 0, s, (const float*)h->grad, h->P,
                       scale_state(h), grad);
    RTC8X_HIP(hipGetLastError());
    RTC8X_HIP(hipMemsetAsync(h->bad + (h->step & 1u), 0, sizeof(uint32_t), s));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_gradient")

extern "C" rtc8x_status rtc8x_nrc_train(rtc8x_nrc* h, void* stream, const float* in, const float* target, uint32_t n,
                                        float* loss_on_cpu)
try {
    RTC8X_REQUIRE(h && in && target, "rtc8x_nrc_train: NULL argument");
    RTC8X_REQUIRE(n > 0 && n <= nrc::kMaxTrain, "rtc8x_nrc_train: n must be in [1, %u]", nrc::kMaxTrain);
    hipStream_t s = (hipStream_t)stream;
    const uint32_t par = (uint32_t)(h->step & 1u);
    rtc8x_status st = gradient_pass(h, s, in, target, n);
    if (st != RTC8X_OK) return st;
    h->step += 1;
    const rtc8x_nrc_desc& d = h->desc;
    nrc::AdamArgs a;
    a.lr = d.learning_rate;
    a.beta1 = d.beta1;
    a.beta2 = d.beta2;
    a.eps = d.epsilon;
    a.l2 = d.l2_reg;
    a.decay = d.ema_decay;
    a.bc1 = (float)(1.0 - std::pow((double)d.beta1, (double)h->step));
    a.bc2 = (float)(1.0 - std::pow((double)d.beta2, (double)h->step));
    a.ema_bc = (float)(1.0 - std::pow((double)d.ema_decay, (double)h->step));
    a.P = h->P;
    a.P_mlp = h->P_mlp;
    a.L = h->L;
    RTC8X_HIP(RTC8X_WITH_TYPE(h, launch_adam<T>(h, s, a, par)));
    if (loss_on_cpu) {
        const uint32_t nw = train_waves_per_block(n) * train_blocks(n);
        std::vector<float> parts(nw);
        RTC8X_HIP(hipMemcpyAsync(parts.data(), h->loss_part, 4 * nw, hipMemcpyDeviceToHost, s));
        RTC8X_HIP(hipStreamSynchronize(s));
        double sum = 0.0;
        for (float p : parts) sum += p;
        *loss_on_cpu = (float)(sum / (3.0 * (double)n));
    }
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_train")

extern "C" rtc8x_status rtc8x_nrc_encode(void* stream, const float* in, uint32_t n, float* out)
try {
    RTC8X_REQUIRE(in && out, "rtc8x_nrc_encode: NULL argument");
    if (n == 0) return RTC8X_OK;
    nrc::HashGeom g{};
    hipLaunchKernelGGL(nrc::encode_kernel<RTC8X_NRC_TRIANGLE_WAVE>, dim3((n + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, (const rtc8x_radiance_query*)in, n, out, (const float*)nullptr, g);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_encode")

extern "C" rtc8x_status rtc8x_nrc_encode_net(const rtc8x_nrc* h, void* stream, const float* in, uint32_t n, float* out,
                                             int32_t which)
try {
    RTC8X_REQUIRE(h && in && out && (which == 0 || which == 1), "rtc8x_nrc_encode_net: bad arguments");
    if (n == 0) return RTC8X_OK;
    const float* tables = which == 0 ? h->master + h->P_mlp : h->grid_inf;
    if (h->enc == RTC8X_NRC_HASH_GRID)
        hipLaunchKernelGGL(nrc::encode_kernel<RTC8X_NRC_HASH_GRID>, dim3((n + 255) / 256), dim3(256), 0,
                           (hipStream_t)stream, (const rtc8x_radiance_query*)in, n, out, tables, h->geom);
    else
        hipLaunchKernelGGL(nrc::encode_kernel<RTC8X_NRC_TRIANGLE_WAVE>, dim3((n + 255) / 256), dim3(256), 0,
                           (hipStream_t)stream, (const rtc8x_radiance_query*)in, n, out, tables, h->geom);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nrc_encode_net")

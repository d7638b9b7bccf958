// rtc8x_post.hip -- the post-process pass and image files.
//
//   tone_map_kernel   applyToneMap (renderer/gpu_kernels/post_process_kernels.cu:9-40) over the mean of
//                     the FP64 accumulator, with the SDR conversion of saveImage
//                     (common/common_host.cpp:515-570) fused in; one thread per pixel, 24 B read and
//                     16 + 4 B written per pixel (HBM-bound, about 100 MB per 1080p frame).
//   rtc8x_sdr_convert the same SDR conversion on the host, for any float image.
//   PNG / BMP / EXR   encoders into memory (host).
//
// The per-pixel arithmetic uses the deterministic exp / log / pow of rtc8x_detmath.h, so the device and
// the host give the same bits.
#include <hip/hip_runtime.h>
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rtc8x_common.hpp"
#include "rtc8x_detmath.h"

namespace rtc8x {
namespace post {

struct Rgb {
    float r, g, b;
};

// RGBSpectrum::luminance (common/basic_types.h:2981-2985)
__host__ __device__ inline float luminance(Rgb c)
{
    return 0.2126729f * c.r + 0.7151522f * c.g + 0.0721750f * c.b;
}

// applyToneMap: per-channel power, then applySimpleToneMap (post_process_kernels.cu:9-18, 28-33)
__host__ __device__ inline Rgb tone_map(Rgb c, float exponent, float brightness)
{
    c.r = rtc8x_powf(c.r, exponent);
    c.g = rtc8x_powf(c.g, exponent);
    c.b = rtc8x_powf(c.b, exponent);
    const float lum = luminance(c);
    if (lum > 0.0f) {
        const float lum_t = 1.0f - rtc8x_expf(-brightness * lum);
        const float s = lum_t / lum;
        c.r = c.r * s;
        c.g = c.g * s;
        c.b = c.b * s;
    }
    return c;
}

__host__ __device__ inline bool finite(float v) { return v - v == 0.0f; }
__host__ __device__ inline float clamp0(float v) { return v < 0.0f ? 0.0f : v; }

// sRGB_gamma_s (common/basic_types.h:3180-3186)
__host__ __device__ inline float srgb_gamma(float v)
{
    if (v <= 0.0031308f) return 12.92f * v;
    return 1.055f * rtc8x_powf(v, (float)(1.0 / 2.4)) - 0.055f;
}

// min(uint32(v * 255), 255) with NaN and negatives -> 0
__host__ __device__ inline uint32_t quantise(float v)
{
    const float s = v * 255.0f;
    if (!(s >= 1.0f)) return 0u;
    if (s >= 255.0f) return 255u;
    return (uint32_t)s;
}

// one pixel of saveImage's loop (common_host.cpp:530-566)
__host__ __device__ inline uint32_t sdr_pixel(Rgb v, float alpha, const rtc8x_sdr_config& cfg)
{
    if (cfg.apply_tone_map) {
        if (!(finite(v.r) && finite(v.g) && finite(v.b))) v = Rgb{0.0f, 0.0f, 0.0f};
        v = Rgb{clamp0(v.r), clamp0(v.g), clamp0(v.b)};
        const float lum = luminance(v);
        const float lum_t = 1.0f - rtc8x_expf(-(cfg.brightness_scale * lum));
        const float s = lum > 0.0f ? lum_t / lum : 0.0f;
        v.r = v.r * s;
        v.g = v.g * s;
        v.b = v.b * s;
    }
    if (cfg.apply_srgb_gamma) {
        v = Rgb{clamp0(v.r), clamp0(v.g), clamp0(v.b)};
        v = Rgb{srgb_gamma(v.r), srgb_gamma(v.g), srgb_gamma(v.b)};
    }
    if (cfg.alpha_override >= 0.0f) alpha = cfg.alpha_override;
    return quantise(v.r) | (quantise(v.g) << 8) | (quantise(v.b) << 16) | (quantise(alpha) << 24);
}

__global__ void __launch_bounds__(256) tone_map_kernel(rtc8x_tone_map_params p)
{
    const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u);
    const uint32_t y = blockIdx.y * 16u + (threadIdx.x >> 4);
    if (x >= p.width || y >= p.height) return;
    const size_t i = (size_t)y * p.width + x;
    const double inv = 1.0 / (double)p.samples;
    const double* a = p.accum + 3 * i;
    Rgb c{(float)(a[0] * inv), (float)(a[1] * inv), (float)(a[2] * inv)};
    c = tone_map(c, p.exponent, p.brightness);
    if (p.out_rgba32f) {
        float4 o;
        o.x = c.r;
        o.y = c.g;
        o.z = c.b;
        o.w = 1.0f;
        reinterpret_cast<float4*>(p.out_rgba32f)[i] = o;
    }
    if (p.out_rgba8) {
        const uint32_t sy = p.sdr.flip_y ? p.height - 1u - y : y;
        p.out_rgba8[(size_t)sy * p.width + x] = sdr_pixel(c, 1.0f, p.sdr);
    }
}

// ---------------------------------------------------------------------------------------------------
// PNG (ISO/IEC 15948): signature, IHDR, IDAT (zlib stream of filtered rows), IEND; CRC-32 per chunk.
// ---------------------------------------------------------------------------------------------------

void put_be32(std::vector<uint8_t>& v, uint32_t x)
{
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void put_chunk(std::vector<uint8_t>& out, const char* type, const uint8_t* data, size_t n)
{
    put_be32(out, (uint32_t)n);
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data, data + n);
    put_be32(out, (uint32_t)crc32(crc32(0L, Z_NULL, 0), out.data() + at, (uInt)(n + 4)));
}

uint8_t paeth(int a, int b, int c)
{
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    return (uint8_t)(pb <= pc ? b : c);
}

// Rows filtered with the type that minimises the sum of absolute signed residuals (the heuristic the
// PNG specification suggests, section 12.8).
std::vector<uint8_t> filtered_rows(const uint8_t* px, uint32_t w, uint32_t h)
{
    const size_t stride = (size_t)w * 4;
    std::vector<uint8_t> out;
    out.reserve((stride + 1) * h);
    std::vector<uint8_t> zero(stride, 0), cand(stride);
    std::vector<uint8_t> best(stride);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* row = px + y * stride;
        const uint8_t* up = y ? px + (y - 1) * stride : zero.data();
        uint64_t best_cost = UINT64_MAX;
        uint8_t best_type = 0;
        for (uint8_t type = 0; type < 5; ++type) {
            uint64_t cost = 0;
            for (size_t i = 0; i < stride; ++i) {
                const int a = i >= 4 ? row[i - 4] : 0, b = up[i], c = i >= 4 ? up[i - 4] : 0;
                uint8_t pred = 0;
                switch (type) {
                    case 1: pred = (uint8_t)a; break;
                    case 2: pred = (uint8_t)b; break;
                    case 3: pred = (uint8_t)((a + b) >> 1); break;
                    case 4: pred = paeth(a, b, c); break;
                    default: break;
                }
                cand[i] = (uint8_t)(row[i] - pred);
                cost += (uint64_t)std::abs((int)(int8_t)cand[i]);
            }
            if (cost < best_cost) {
                best_cost = cost;
                best_type = type;
                best.swap(cand);
            }
        }
        out.push_back(best_type);
        out.insert(out.end(), best.begin(), best.end());
    }
    return out;
}

void put_le16(uint8_t*& p, uint16_t v)
{
    std::memcpy(p, &v, 2);
    p += 2;
}
void put_le32(uint8_t*& p, uint32_t v)
{
    std::memcpy(p, &v, 4);
    p += 4;
}

// IEEE binary32 -> binary16, round to nearest even, with subnormals, overflow to inf and quiet NaN.
uint16_t to_half(float f)
{
    const uint32_t x = rtc8x_f2u(f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds to >= 65520: inf
    if (ax < 0x38800000u) {                                     // below 2^-14: subnormal half (or 0)
        if (ax < 0x33000000u) return (uint16_t)sign;             // below 2^-25: rounds to 0
        const uint32_t m = (ax & 0x007fffffu) | 0x00800000u;
        const int shift = 126 - (int)(ax >> 23);                 // 14..24
        const uint32_t q = m >> shift, rem = m & ((1u << shift) - 1u), halfway = 1u << (shift - 1);
        const uint32_t r = q + ((rem > halfway || (rem == halfway && (q & 1u))) ? 1u : 0u);
        return (uint16_t)(sign | r);
    }
    const uint32_t e = (ax >> 23) - 112u;  // rebias 127 -> 15
    const uint32_t m = ax & 0x007fffffu;
    uint32_t h = (e << 10) | (m >> 13);
    const uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;  // may carry into the exponent: correct
    return (uint16_t)(sign | h);
}

}  // namespace post
}  // namespace rtc8x

using namespace rtc8x;
using namespace rtc8x::post;

extern "C" rtc8x_status rtc8x_tone_map(void* stream, const rtc8x_tone_map_params* p)
try {
    RTC8X_REQUIRE(p && p->accum, "rtc8x_tone_map: NULL argument");
    RTC8X_REQUIRE(p->samples > 0, "rtc8x_tone_map: samples must be positive");
    RTC8X_REQUIRE(p->out_rgba32f || p->out_rgba8, "rtc8x_tone_map: no output");
    RTC8X_REQUIRE(((uintptr_t)p->out_rgba32f & 15u) == 0, "rtc8x_tone_map: out_rgba32f must be 16-byte aligned");
    if (p->width == 0 || p->height == 0) return RTC8X_OK;
    RTC8X_REQUIRE(p->width <= 65536u && p->height <= 65536u, "rtc8x_tone_map: %ux%u is too large", p->width,
                  p->height);
    dim3 grid((p->width + 15u) / 16u, (p->height + 15u) / 16u);
    hipLaunchKernelGGL(tone_map_kernel, grid, dim3(256), 0, (hipStream_t)stream, *p);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_tone_map")

extern "C" rtc8x_status rtc8x_sdr_convert(const float* data, uint32_t w, uint32_t h, uint32_t nch,
                                          const rtc8x_sdr_config* cfg, uint32_t* out)
try {
    RTC8X_REQUIRE(cfg && out && (data || (size_t)w * h == 0), "rtc8x_sdr_convert: NULL argument");
    RTC8X_REQUIRE(nch >= 1 && nch <= 4, "rtc8x_sdr_convert: %u channels (1..4 supported)", nch);
    for (uint32_t y = 0; y < h; ++y) {
        const uint32_t sy = cfg->flip_y ? h - 1 - y : y;
        for (uint32_t x = 0; x < w; ++x) {
            const float* d = data + ((size_t)sy * w + x) * nch;
            Rgb v{d[0], nch >= 2 ? d[1] : 0.0f, nch >= 3 ? d[2] : 0.0f};
            const float alpha = nch == 4 ? d[3] : 1.0f;
            out[(size_t)y * w + x] = sdr_pixel(v, alpha, *cfg);
        }
    }
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_sdr_convert")

extern "C" rtc8x_status rtc8x_png_encode(const uint32_t* px, uint32_t w, uint32_t h, void* out, size_t cap,
                                         size_t* size)
try {
    RTC8X_REQUIRE(size && (px || (size_t)w * h == 0), "rtc8x_png_encode: NULL argument");
    RTC8X_REQUIRE(w > 0 && h > 0 && w <= (1u << 24) && h <= (1u << 24), "rtc8x_png_encode: bad size %ux%u", w, h);
    const size_t raw = ((size_t)w * 4 + 1) * h;
    const size_t bound = 8 + 25 + 12 + (size_t)compressBound((uLong)raw) + 12;
    if (!out) {
        *size = bound;
        return RTC8X_OK;
    }
    std::vector<uint8_t> rows = filtered_rows(reinterpret_cast<const uint8_t*>(px), w, h);
    uLongf zn = compressBound((uLong)rows.size());
    std::vector<uint8_t> z(zn);
    const int zs = compress2(z.data(), &zn, rows.data(), (uLong)rows.size(), 6);
    RTC8X_REQUIRE(zs == Z_OK, "rtc8x_png_encode: zlib failed (%d)", zs);
    std::vector<uint8_t> f = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8 bits, RGBA, deflate, adaptive filtering, no interlace
    put_chunk(f, "IHDR", ihdr.data(), ihdr.size());
    put_chunk(f, "IDAT", z.data(), zn);
    put_chunk(f, "IEND", nullptr, 0);
    RTC8X_REQUIRE(cap >= f.size(), "rtc8x_png_encode: capacity %zu is below %zu", cap, f.size());
    std::memcpy(out, f.data(), f.size());
    *size = f.size();
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_png_encode")

extern "C" rtc8x_status rtc8x_bmp_encode(const uint32_t* px, uint32_t w, uint32_t h, void* out, size_t cap,
                                         size_t* size)
try {
    RTC8X_REQUIRE(size && (px || (size_t)w * h == 0), "rtc8x_bmp_encode: NULL argument");
    RTC8X_REQUIRE(w <= (1u << 16) && h <= (1u << 16), "rtc8x_bmp_encode: bad size %ux%u", w, h);
    const uint32_t pad = (uint32_t)(-(int64_t)w * 3) & 3u;
    const size_t total = 14 + 40 + ((size_t)w * 3 + pad) * h;
    *size = total;
    if (!out) return RTC8X_OK;
    RTC8X_REQUIRE(cap >= total, "rtc8x_bmp_encode: capacity %zu is below %zu", cap, total);
    uint8_t* p = (uint8_t*)out;
    *p++ = 'B';
    *p++ = 'M';
    put_le32(p, (uint32_t)total);
    put_le16(p, 0);
    put_le16(p, 0);
    put_le32(p, 14 + 40);
    put_le32(p, 40);
    put_le32(p, w);
    put_le32(p, h);
    put_le16(p, 1);
    put_le16(p, 24);
    for (int i = 0; i < 6; ++i) put_le32(p, 0);
    static const int bg[3] = {255, 0, 255};
    for (int64_t y = (int64_t)h - 1; y >= 0; --y) {  // bottom-up
        for (uint32_t x = 0; x < w; ++x) {
            const uint32_t v = px[(size_t)y * w + x];
            const int d[4] = {(int)(v & 255u), (int)((v >> 8) & 255u), (int)((v >> 16) & 255u), (int)(v >> 24)};
            int c[3];
            for (int k = 0; k < 3; ++k) c[k] = bg[k] + ((d[k] - bg[k]) * d[3]) / 255;  // C division truncates
            *p++ = (uint8_t)c[2];
            *p++ = (uint8_t)c[1];
            *p++ = (uint8_t)c[0];
        }
        for (uint32_t k = 0; k < pad; ++k) *p++ = 0;
    }
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_bmp_encode")

extern "C" rtc8x_status rtc8x_exr_encode(const float* data, uint32_t w, uint32_t h, uint32_t nch, float scale,
                                         int32_t flip_y, void* out, size_t cap, size_t* size)
try {
    RTC8X_REQUIRE(size && data, "rtc8x_exr_encode: NULL argument");
    RTC8X_REQUIRE(nch >= 1 && nch <= 4, "rtc8x_exr_encode: %u channels (1..4 supported)", nch);
    RTC8X_REQUIRE(w > 0 && h > 0 && w <= (1u << 24) && h <= (1u << 24), "rtc8x_exr_encode: bad size %ux%u", w, h);
    // header attributes: name, type, size, value (OpenEXR file layout, single-part scanline)
    std::vector<uint8_t> hdr;
    auto put_str = [&](const char* s) { hdr.insert(hdr.end(), s, s + std::strlen(s) + 1); };
    auto put_i32 = [&](int32_t v) {
        uint8_t b[4];
        std::memcpy(b, &v, 4);
        hdr.insert(hdr.end(), b, b + 4);
    };
    auto put_f32 = [&](float v) {
        uint8_t b[4];
        std::memcpy(b, &v, 4);
        hdr.insert(hdr.end(), b, b + 4);
    };
    const char* names[4] = {"A", "B", "G", "R"};
    put_str("channels");
    put_str("chlist");
    put_i32(4 * (2 + 16) + 1);
    for (const char* n : names) {
        put_str(n);
        put_i32(1);                                   // HALF
        hdr.insert(hdr.end(), {0, 0, 0, 0});          // pLinear, reserved
        put_i32(1);                                   // x sampling
        put_i32(1);                                   // y sampling
    }
    hdr.push_back(0);
    put_str("compression");
    put_str("compression");
    put_i32(1);
    hdr.push_back(0);                                 // NO_COMPRESSION
    for (const char* win : {"dataWindow", "displayWindow"}) {
        put_str(win);
        put_str("box2i");
        put_i32(16);
        put_i32(0);
        put_i32(0);
        put_i32((int32_t)w - 1);
        put_i32((int32_t)h - 1);
    }
    put_str("lineOrder");
    put_str("lineOrder");
    put_i32(1);
    hdr.push_back(0);                                 // INCREASING_Y
    put_str("pixelAspectRatio");
    put_str("float");
    put_i32(4);
    put_f32(1.0f);
    put_str("screenWindowCenter");
    put_str("v2f");
    put_i32(8);
    put_f32(0.0f);
    put_f32(0.0f);
    put_str("screenWindowWidth");
    put_str("float");
    put_i32(4);
    put_f32(1.0f);
    hdr.push_back(0);
    const size_t line_bytes = 8 + (size_t)w * 2 * 4;
    const size_t table_at = 8 + hdr.size();
    const size_t total = table_at + (size_t)h * 8 + (size_t)h * line_bytes;
    *size = total;
    if (!out) return RTC8X_OK;
    RTC8X_REQUIRE(cap >= total, "rtc8x_exr_encode: capacity %zu is below %zu", cap, total);
    uint8_t* o = (uint8_t*)out;
    const uint32_t magic = 20000630u, version = 2u;
    std::memcpy(o, &magic, 4);
    std::memcpy(o + 4, &version, 4);
    std::memcpy(o + 8, hdr.data(), hdr.size());
    uint8_t* table = o + table_at;
    uint8_t* p = table + (size_t)h * 8;
    for (uint32_t y = 0; y < h; ++y) {
        const uint64_t at = (uint64_t)(p - o);
        std::memcpy(table + (size_t)y * 8, &at, 8);
        const int32_t yy = (int32_t)y, bytes = (int32_t)(w * 2 * 4);
        std::memcpy(p, &yy, 4);
        std::memcpy(p + 4, &bytes, 4);
        p += 8;
        // saveImageHDR writes source row y to row (flip ? h-1-y : y): output row y reads source row
        // (flip ? h-1-y : y) as well
        const uint32_t sy = flip_y ? h - 1 - y : y;
        for (int c = 3; c >= 0; --c) {  // A, B, G, R
            for (uint32_t x = 0; x < w; ++x) {
                const float* d = data + ((size_t)sy * w + x) * nch;
                const float v = (uint32_t)c < nch ? scale * d[c] : 0.0f;
                const uint16_t hv = to_half(v);
                std::memcpy(p, &hv, 2);
                p += 2;
            }
        }
    }
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_exr_encode")

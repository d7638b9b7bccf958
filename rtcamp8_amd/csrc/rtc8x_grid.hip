// rtc8x_grid.hip -- the density grid: procedural fBm build on the device, upload of host tables,
// cell majorants, and the parity probe.  Replaces the volume part of Scene::allocateVolumeGrid /
// setVolumeGrid (renderer/renderer_scene.cpp:491-545).
//
// HBM layout (all library-owned):
//   root_tiles      num_root_tiles x 16 B
//   upper_tables    num_upper x 32768 x u32   (lower-node index or tile)
//   lower_tables    num_lower x 4096 x u32    (leaf index or tile)
//   leaf_values     num_leaves x 512 x f32    (2 KiB per leaf, NanoVDB voxel order x-major), + 1 f32 padding
//   tile_values     one f32 per tile value, + 1 f32 padding (lookups read voxel pairs z, z+1 unconditionally)
//   leaf_origins    num_leaves x 3 x i32
//   cell_majorants  one f32 per 8^3 cell of the index bbox (largest voxel over the cell and a one-voxel rim)
//   region_majorants one byte per region of region_cells^3 cells (quantised majorant or empty-space distance)
//   leaf_map        one u32 per 8^3 block of the index bbox plus a one-block margin (the tree walk's result)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "rtc8x_common.hpp"
#include "rtc8x_device.hpp"

namespace rtc8x {
namespace dev {

// fBm field: value noise on an integer lattice (lowbias32 hash), smoothstep weights, summed octaves.
// The same expressions, in the same order, as the CPU oracle's builder.
__device__ __forceinline__ uint32_t fbm_hash(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ float fbm_lattice(uint32_t s, int32_t ix, int32_t iy, int32_t iz)
{
    uint32_t h = fbm_hash(s + (uint32_t)ix);
    h = fbm_hash(h + (uint32_t)iy);
    h = fbm_hash(h + (uint32_t)iz);
    return (float)(h >> 8) * 0x1p-23f - 1.0f;
}
__device__ __forceinline__ float fbm_lerp(float a, float b, float w) { return a + w * (b - a); }
__device__ float fbm_noise(float px, float py, float pz, uint32_t s)
{
    float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    float tx = px - fx, ty = py - fy, tz = pz - fz;
    float wx = tx * tx * (3.0f - 2.0f * tx);
    float wy = ty * ty * (3.0f - 2.0f * ty);
    float wz = tz * tz * (3.0f - 2.0f * tz);
    float c000 = fbm_lattice(s, ix, iy, iz), c100 = fbm_lattice(s, ix + 1, iy, iz);
    float c010 = fbm_lattice(s, ix, iy + 1, iz), c110 = fbm_lattice(s, ix + 1, iy + 1, iz);
    float c001 = fbm_lattice(s, ix, iy, iz + 1), c101 = fbm_lattice(s, ix + 1, iy, iz + 1);
    float c011 = fbm_lattice(s, ix, iy + 1, iz + 1), c111 = fbm_lattice(s, ix + 1, iy + 1, iz + 1);
    float a00 = fbm_lerp(c000, c100, wx), a10 = fbm_lerp(c010, c110, wx);
    float a01 = fbm_lerp(c001, c101, wx), a11 = fbm_lerp(c011, c111, wx);
    float b0 = fbm_lerp(a00, a10, wy), b1 = fbm_lerp(a01, a11, wy);
    return fbm_lerp(b0, b1, wz);
}

__device__ float fbm_value(const rtc8x_fbm_desc& d, int32_t i, int32_t j, int32_t k)
{
    float inv_n = 1.0f / (float)d.resolution;
    float qx = ((float)i + 0.5f) * inv_n;
    float qy = ((float)j + 0.5f) * inv_n;
    float qz = ((float)k + 0.5f) * inv_n;
    float sum = 0.0f, amp = 1.0f, freq = d.base_frequency;
    uint32_t s_lo = (uint32_t)d.seed, s_hi = (uint32_t)(d.seed >> 32);
    for (uint32_t o = 0; o < d.octaves; ++o) {
        uint32_t s = fbm_hash(s_lo ^ fbm_hash(s_hi + o));
        sum = sum + amp * fbm_noise(qx * freq, qy * freq, qz * freq, s);
        amp = amp * d.gain;
        freq = freq * d.lacunarity;
    }
    if (!d.sparse) return sum;
    float dx = qx - 0.5f, dy = qy - 0.5f, dz = qz - 0.5f;
    float r = sqrtf(dx * dx + dy * dy + dz * dz);
    float shape = 1.0f - 2.0f * r;
    if (shape < 0.0f) shape = 0.0f;
    return shape + 0.6f * sum;
}

// One 512-lane workgroup per 8^3 leaf slot: max of the unnormalised field.
__global__ __launch_bounds__(512) void leaf_max_kernel(const rtc8x_fbm_desc d, uint32_t nl, float* __restrict__ leaf_max)
{
    const uint32_t slot = blockIdx.x;
    const uint32_t lz = slot % nl, ly = (slot / nl) % nl, lx = slot / (nl * nl);
    const uint32_t v = threadIdx.x;
    float g = fbm_value(d, (int32_t)(lx * 8 + (v >> 6)), (int32_t)(ly * 8 + ((v >> 3) & 7)), (int32_t)(lz * 8 + (v & 7)));
    for (int o = 32; o > 0; o >>= 1) g = fmaxf(g, __shfl_xor(g, o, 64));
    __shared__ float part[8];
    if ((v & 63) == 0) part[v >> 6] = g;
    __syncthreads();
    if (v == 0) {
        float m = part[0];
        for (int w = 1; w < 8; ++w) m = fmaxf(m, part[w]);
        leaf_max[slot] = m;
    }
}

// One 512-lane workgroup per active leaf: density = max(g - sub, 0) * inv_range.
__global__ __launch_bounds__(512) void fill_leaves_kernel(const rtc8x_fbm_desc d, const int32_t* __restrict__ origins,
                                                          float sub, float inv_range, float* __restrict__ values)
{
    const uint32_t leaf = blockIdx.x;
    const uint32_t v = threadIdx.x;
    const int32_t ox = origins[3 * leaf], oy = origins[3 * leaf + 1], oz = origins[3 * leaf + 2];
    float gv = fbm_value(d, ox + (int32_t)(v >> 6), oy + (int32_t)((v >> 3) & 7), oz + (int32_t)(v & 7));
    float dv = gv - sub;
    if (dv < 0.0f) dv = 0.0f;
    dv = dv * inv_range;
    dv = dv * d.contrast;
    if (dv > 1.0f) dv = 1.0f;
    values[(size_t)leaf * 512 + v] = dv;
}

// One lane per block of the dense leaf map: the tree walk's result for that block.
__global__ void leaf_map_kernel(const rtc8x_grid g, uint32_t* __restrict__ map)
{
    const size_t n = (size_t)g.leaf_map_dims[0] * g.leaf_map_dims[1] * g.leaf_map_dims[2];
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint32_t bz = (uint32_t)(q % (uint32_t)g.leaf_map_dims[2]);
    const size_t r = q / (uint32_t)g.leaf_map_dims[2];
    const uint32_t by = (uint32_t)(r % (uint32_t)g.leaf_map_dims[1]);
    const uint32_t bx = (uint32_t)(r / (uint32_t)g.leaf_map_dims[1]);
    map[q] = find_leaf(g, g.leaf_map_origin[0] + 8 * (int32_t)bx, g.leaf_map_origin[1] + 8 * (int32_t)by,
                       g.leaf_map_origin[2] + 8 * (int32_t)bz);
}

// One wavefront per 8^3 cell c (cell_origin + 8c): the largest voxel value, through the tree (leaves,
// tiles and background alike, never below 0), over index [8c - 1, 8c + 9]^3.  A point of the cell reads
// its trilinear taps from [8c, 8c + 8]; the rim of one more voxel on every side covers a sample position
// that rounding puts just outside the cell's span of the ray.
constexpr int32_t kCellSpan = 11;
__global__ __launch_bounds__(256) void cell_max_kernel(const rtc8x_grid g, float* __restrict__ out)
{
    const uint32_t dy = (uint32_t)g.cell_dims[1], dz = (uint32_t)g.cell_dims[2];
    const size_t n = (size_t)g.cell_dims[0] * dy * dz;
    const size_t cell = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    if (cell >= n) return;  // uniform over the wavefront
    const uint32_t cz = (uint32_t)(cell % dz), cy = (uint32_t)((cell / dz) % dy), cx = (uint32_t)(cell / ((size_t)dz * dy));
    const int32_t x0 = g.cell_origin[0] + 8 * (int32_t)cx - 1, y0 = g.cell_origin[1] + 8 * (int32_t)cy - 1,
                  z0 = g.cell_origin[2] + 8 * (int32_t)cz - 1;
    float m = 0.0f;
    for (int32_t v = (int32_t)lane; v < kCellSpan * kCellSpan * kCellSpan; v += 64) {
        const int32_t i = x0 + v / (kCellSpan * kCellSpan), j = y0 + (v / kCellSpan) % kCellSpan, k = z0 + v % kCellSpan;
        m = fmaxf(m, *voxel_ptr(g, block_entry(g, i, j, k), i, j, k));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if (lane == 0) out[cell] = m;
}

__global__ void probe_kernel(const rtc8x_grid g, uint32_t n, const float* __restrict__ xyz, int32_t* __restrict__ out_ijk,
                             uint32_t* __restrict__ out_off, float* __restrict__ out_val)
{
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    V3 idx = world_to_index(g, xyz[3 * q], xyz[3 * q + 1], xyz[3 * q + 2]);
    float v = trilinear(g, idx);
    int32_t i = (int32_t)floorf(idx.x), j = (int32_t)floorf(idx.y), k = (int32_t)floorf(idx.z);
    out_ijk[3 * q] = i;
    out_ijk[3 * q + 1] = j;
    out_ijk[3 * q + 2] = k;
    out_off[3 * q] = upper_offset(i, j, k);
    out_off[3 * q + 1] = lower_offset(i, j, k);
    out_off[3 * q + 2] = leaf_offset(i, j, k);
    out_val[q] = v;
}

}  // namespace dev

// Host-side table builder shared by the fBm build and the upload path.
namespace {

struct HostTables {
    std::vector<rtc8x_root_tile> root;
    std::vector<uint32_t> upper, lower;
    std::vector<int32_t> origins;
    std::vector<float> leaf_max;   // per stored leaf
    uint32_t num_lower = 0, num_leaves = 0;
};

// Region table from the cell majorants (the layout rtc8x.h defines for rtc8x_grid): region_cells as
// requested (0: RTC8X_DEFAULT_REGION_CELLS, doubled until there are at most RTC8X_MAX_REGIONS regions);
// a unit of (largest cell majorant) / 127 raised until 127 units cover it; per region with density the
// least q whose q units cover its largest cell, per empty region RTC8X_REGION_EMPTY + the number of empty
// layers around it (a Chebyshev distance transform by repeated 26-neighbour dilation of the regions with
// density).
rtc8x_status region_majorants(const std::vector<float>& cells, const int32_t cell_dims[3], uint32_t requested,
                              int32_t region_dims[3], uint32_t* region_cells, float* region_unit,
                              std::vector<uint8_t>& out)
{
    RTC8X_REQUIRE(requested <= 4096 && (requested & (requested - 1)) == 0,
                  "region_cells %u must be 0 or a power of two up to 4096", requested);
    uint32_t rc = requested ? requested : RTC8X_DEFAULT_REGION_CELLS;
    for (;;) {
        size_t n = 1;
        for (int a = 0; a < 3; ++a) {
            region_dims[a] = (cell_dims[a] + (int32_t)rc - 1) / (int32_t)rc;
            n *= (size_t)region_dims[a];
        }
        if (n <= RTC8X_MAX_REGIONS) break;
        RTC8X_REQUIRE(requested == 0, "region_cells %u gives %zu regions, more than %u", requested, n,
                      (unsigned)RTC8X_MAX_REGIONS);
        rc *= 2;
    }
    const int32_t rx = region_dims[0], ry = region_dims[1], rz = region_dims[2];
    const size_t nr = (size_t)rx * ry * rz;
    std::vector<float> rmax(nr, 0.0f);
    float top = 0.0f;
    for (int32_t x = 0; x < cell_dims[0]; ++x)
        for (int32_t y = 0; y < cell_dims[1]; ++y)
            for (int32_t z = 0; z < cell_dims[2]; ++z) {
                const float m = cells[((size_t)x * cell_dims[1] + y) * cell_dims[2] + z];
                float& r = rmax[((size_t)(x / (int32_t)rc) * ry + y / (int32_t)rc) * rz + z / (int32_t)rc];
                r = std::max(r, m);
                top = std::max(top, m);
            }
    float unit = top / 127.0f;
    while (127.0f * unit < top) unit = std::nextafter(unit, INFINITY);
    // distance (in regions) to the nearest region with density; 0 for those regions themselves
    const uint32_t kFar = 255;
    std::vector<uint8_t> dist(nr, (uint8_t)kFar);
    for (size_t r = 0; r < nr; ++r)
        if (rmax[r] > 0.0f) dist[r] = 0;
    for (uint32_t k = 1; k <= 128; ++k) {
        bool grew = false;
        std::vector<uint8_t> next = dist;
        for (int32_t x = 0; x < rx; ++x)
            for (int32_t y = 0; y < ry; ++y)
                for (int32_t z = 0; z < rz; ++z) {
                    const size_t r = ((size_t)x * ry + y) * rz + z;
                    if (dist[r] != kFar) continue;
                    bool near = false;
                    for (int dx = -1; dx <= 1 && !near; ++dx)
                        for (int dy = -1; dy <= 1 && !near; ++dy)
                            for (int dz = -1; dz <= 1 && !near; ++dz) {
                                const int32_t a = x + dx, b = y + dy, c = z + dz;
                                if (a < 0 || b < 0 || c < 0 || a >= rx || b >= ry || c >= rz) continue;
                                near = dist[((size_t)a * ry + b) * rz + c] == k - 1;
                            }
                    if (near) {
                        next[r] = (uint8_t)k;
                        grew = true;
                    }
                }
        dist.swap(next);
        if (!grew) break;
    }
    // a multiple of 16 bytes: the tracer stages the table in LDS 16 bytes at a time
    out.assign((nr + 15) & ~(size_t)15, 0);
    for (size_t r = 0; r < nr; ++r) {
        if (rmax[r] > 0.0f) {
            uint32_t q = 1;
            while (q < 127 && (float)q * unit < rmax[r]) ++q;
            out[r] = (uint8_t)q;
        } else {
            const uint32_t k = dist[r] == kFar ? 127u : std::min<uint32_t>(dist[r] - 1u, 127u);
            out[r] = (uint8_t)(RTC8X_REGION_EMPTY + k);
        }
    }
    *region_cells = rc;
    *region_unit = unit;
    return RTC8X_OK;
}

template <typename T>
rtc8x_status upload_vec(const std::vector<T>& v, T** dst)
{
    *dst = nullptr;
    if (v.empty()) return RTC8X_OK;
    RTC8X_HIP(hipMalloc((void**)dst, v.size() * sizeof(T)));
    RTC8X_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return RTC8X_OK;
}

#define RTC8X_TRY(expr)                        \
    do {                                       \
        rtc8x_status st_ = (expr);             \
        if (st_ != RTC8X_OK) return st_;       \
    } while (0)

// Dense leaf map over the index bbox's blocks plus one block on every side (a trilinear stencil at a
// position inside the world bbox reaches at most one voxel past the index bbox), built on the device
// from the uploaded tables.  Skipped (leaf_map = NULL) beyond 2^26 blocks.
rtc8x_status build_leaf_map(rtc8x_grid& g, hipStream_t s)
{
    g.leaf_map = nullptr;
    size_t n = 1;
    for (int a = 0; a < 3; ++a) {
        const int32_t lo = (int32_t)std::floor((float)g.index_bbox_min[a] / 8.0f) - 1;
        const int32_t hi = (int32_t)std::floor((float)g.index_bbox_max[a] / 8.0f) + 1;
        g.leaf_map_origin[a] = lo * 8;
        g.leaf_map_dims[a] = hi - lo + 1;
        n *= (size_t)g.leaf_map_dims[a];
    }
    if (n > ((size_t)1 << 26)) {
        for (int a = 0; a < 3; ++a) g.leaf_map_dims[a] = 0;
        return RTC8X_OK;
    }
    uint32_t* d_map = nullptr;
    RTC8X_HIP(hipMalloc((void**)&d_map, n * sizeof(uint32_t)));
    hipLaunchKernelGGL(dev::leaf_map_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, d_map);
    RTC8X_HIP(hipGetLastError());
    RTC8X_HIP(hipStreamSynchronize(s));
    g.leaf_map = d_map;
    return RTC8X_OK;
}

// The cell majorants of a device grid (cell grid over the index bbox, from cell_max_kernel) and its region
// table (region_cells as region_majorants takes it).  Grids without leaves, or whose index bbox spans more
// than 2^26 cells, get neither (local-majorant tracking then refuses them).
rtc8x_status build_majorant_tables(rtc8x_grid& g, uint32_t region_cells, hipStream_t s)
{
    g.cell_majorants = nullptr;
    g.region_majorants = nullptr;
    for (int a = 0; a < 3; ++a) {
        g.cell_origin[a] = g.cell_dims[a] = 0;
        g.region_dims[a] = 0;
    }
    g.region_cells = 0;
    g.region_unit = 0.0f;
    if (g.num_leaves == 0) return RTC8X_OK;
    size_t n = 1;
    for (int a = 0; a < 3; ++a) {
        const int32_t lo = (int32_t)std::floor((float)g.index_bbox_min[a] / 8.0f);
        const int32_t hi = (int32_t)std::floor((float)g.index_bbox_max[a] / 8.0f);
        g.cell_origin[a] = lo * 8;
        g.cell_dims[a] = hi - lo + 1;
        n *= (size_t)g.cell_dims[a];
    }
    if (n > ((size_t)1 << 26)) {
        for (int a = 0; a < 3; ++a) g.cell_origin[a] = g.cell_dims[a] = 0;
        return RTC8X_OK;
    }
    float* d_cm = nullptr;
    RTC8X_HIP(hipMalloc((void**)&d_cm, n * sizeof(float)));
    g.cell_majorants = d_cm;  // owned by g from here on
    hipLaunchKernelGGL(dev::cell_max_kernel, dim3((unsigned)((n * 64 + 255) / 256)), dim3(256), 0, s, g, d_cm);
    RTC8X_HIP(hipGetLastError());
    std::vector<float> cm(n);
    RTC8X_HIP(hipMemcpyAsync(cm.data(), d_cm, n * sizeof(float), hipMemcpyDeviceToHost, s));
    RTC8X_HIP(hipStreamSynchronize(s));
    std::vector<uint8_t> rm;
    int32_t rdims[3];
    uint32_t rc = 0;
    float unit = 0.0f;
    RTC8X_TRY(region_majorants(cm, g.cell_dims, region_cells, rdims, &rc, &unit, rm));
    uint8_t* d_rm = nullptr;
    RTC8X_TRY(upload_vec(rm, &d_rm));
    g.region_majorants = d_rm;
    for (int a = 0; a < 3; ++a) g.region_dims[a] = rdims[a];
    g.region_cells = rc;
    g.region_unit = unit;
    return RTC8X_OK;
}

}  // namespace
}  // namespace rtc8x

using namespace rtc8x;

extern "C" rtc8x_status rtc8x_grid_build_regions(rtc8x_grid* g, uint32_t region_cells)
try {
    RTC8X_REQUIRE(g != nullptr, "rtc8x_grid_build_regions: grid is NULL");
    const size_t n = (size_t)std::max(g->cell_dims[0], 0) * std::max(g->cell_dims[1], 0) * std::max(g->cell_dims[2], 0);
    RTC8X_REQUIRE(g->cell_majorants && n > 0, "rtc8x_grid_build_regions: the grid has no cell majorants");
    std::vector<float> cm(n);
    RTC8X_HIP(hipDeviceSynchronize());
    RTC8X_HIP(hipMemcpy(cm.data(), g->cell_majorants, n * sizeof(float), hipMemcpyDeviceToHost));
    std::vector<uint8_t> rm;
    int32_t rdims[3];
    uint32_t rc = 0;
    float unit = 0.0f;
    RTC8X_TRY(region_majorants(cm, g->cell_dims, region_cells, rdims, &rc, &unit, rm));
    uint8_t* d_rm = nullptr;
    RTC8X_TRY(upload_vec(rm, &d_rm));
    if (g->region_majorants) {
        const hipError_t e = hipFree(const_cast<uint8_t*>(g->region_majorants));
        if (e != hipSuccess) {
            (void)hipFree(d_rm);
            RTC8X_HIP(e);
        }
    }
    g->region_majorants = d_rm;
    for (int a = 0; a < 3; ++a) g->region_dims[a] = rdims[a];
    g->region_cells = rc;
    g->region_unit = unit;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_grid_build_regions")

extern "C" rtc8x_status rtc8x_grid_destroy(rtc8x_grid* g)
try {
    RTC8X_REQUIRE(g != nullptr, "rtc8x_grid_destroy: grid is NULL");
    const void* ptrs[] = {g->root_tiles, g->upper_tables, g->lower_tables, g->leaf_values,     g->tile_values,
                          g->leaf_origins, g->cell_majorants, g->leaf_map,    g->region_majorants};
    for (const void* p : ptrs)
        if (p) RTC8X_HIP(hipFree(const_cast<void*>(p)));
    std::memset(g, 0, sizeof(*g));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_grid_destroy")

extern "C" rtc8x_status rtc8x_grid_build_fbm(const rtc8x_fbm_desc* d, rtc8x_grid* out, void* stream)
try {
    RTC8X_REQUIRE(d && out, "rtc8x_grid_build_fbm: NULL argument");
    const uint32_t n = d->resolution;
    RTC8X_REQUIRE(n >= 8 && n <= 4096 && (n % 8) == 0, "rtc8x_grid_build_fbm: resolution %u must be a multiple of 8 in [8, 4096]", n);
    RTC8X_REQUIRE(d->octaves >= 1 && d->octaves <= 16, "rtc8x_grid_build_fbm: octaves %u out of range", d->octaves);
    RTC8X_REQUIRE(std::isfinite(d->base_frequency) && std::isfinite(d->lacunarity) && std::isfinite(d->gain),
                  "rtc8x_grid_build_fbm: non-finite parameters");
    RTC8X_REQUIRE(!d->sparse || (d->active_leaf_fraction > 0.0f && d->active_leaf_fraction <= 1.0f),
                  "rtc8x_grid_build_fbm: active_leaf_fraction must be in (0, 1]");
    RTC8X_REQUIRE(std::isfinite(d->contrast) && d->contrast >= 1.0f, "rtc8x_grid_build_fbm: contrast must be >= 1");
    std::memset(out, 0, sizeof(*out));
    hipStream_t s = (hipStream_t)stream;
    const uint32_t nl = n / 8, num_slots = nl * nl * nl;

    // pass 1: per-slot maxima of the unnormalised field
    float* d_leaf_max = nullptr;
    RTC8X_HIP(hipMalloc(&d_leaf_max, sizeof(float) * num_slots));
    hipLaunchKernelGGL(dev::leaf_max_kernel, dim3(num_slots), dim3(512), 0, s, *d, nl, d_leaf_max);
    RTC8X_HIP(hipGetLastError());
    std::vector<float> slot_max(num_slots);
    RTC8X_HIP(hipMemcpyAsync(slot_max.data(), d_leaf_max, sizeof(float) * num_slots, hipMemcpyDeviceToHost, s));
    RTC8X_HIP(hipStreamSynchronize(s));
    RTC8X_HIP(hipFree(d_leaf_max));

    float gmax = -FLT_MAX;
    for (float m : slot_max) gmax = std::max(gmax, m);
    float thr = 0.0f;
    bool keep_all = !d->sparse;
    if (d->sparse) {
        uint32_t n_keep = (uint32_t)(d->active_leaf_fraction * (float)num_slots);
        if (n_keep >= num_slots) {
            keep_all = true;
            thr = -FLT_MAX;
        } else {
            std::vector<float> sorted(slot_max);
            std::nth_element(sorted.begin(), sorted.begin() + n_keep, sorted.end(), std::greater<float>());
            thr = sorted[n_keep];
        }
    }
    const float range = d->sparse ? gmax - thr : gmax;
    RTC8X_REQUIRE(range > 0.0f, "rtc8x_grid_build_fbm: degenerate field (max %g, threshold %g)", gmax, thr);
    const float inv_range = 1.0f / range;
    float sub = d->sparse ? thr : 0.0f;
    if (d->sparse && keep_all) sub = 0.0f;
    auto active = [&](uint32_t lx, uint32_t ly, uint32_t lz) {
        return keep_all || slot_max[(lx * nl + ly) * nl + lz] > thr;
    };

    // tables in NanoVDB traversal order (lower nodes by upper offset, leaves by lower offset)
    std::vector<uint32_t> lower_of_uoff(RTC8X_UPPER_TABLE_SIZE, 0xFFFFFFFFu);
    const uint32_t nb = (nl + 15) / 16;
    uint32_t num_lower = 0;
    for (uint32_t bx = 0; bx < nb; ++bx)
        for (uint32_t by = 0; by < nb; ++by)
            for (uint32_t bz = 0; bz < nb; ++bz) {
                bool any = false;
                for (uint32_t o = 0; o < 4096 && !any; ++o) {
                    uint32_t lx = bx * 16 + (o >> 8), ly = by * 16 + ((o >> 4) & 15), lz = bz * 16 + (o & 15);
                    if (lx < nl && ly < nl && lz < nl && active(lx, ly, lz)) any = true;
                }
                if (any) lower_of_uoff[(bx << 10) | (by << 5) | bz] = num_lower++;
            }
    std::vector<uint32_t> upper(RTC8X_UPPER_TABLE_SIZE), lower((size_t)RTC8X_LOWER_TABLE_SIZE * num_lower, RTC8X_TILE_BIT);
    for (uint32_t u = 0; u < RTC8X_UPPER_TABLE_SIZE; ++u)
        upper[u] = lower_of_uoff[u] == 0xFFFFFFFFu ? RTC8X_TILE_BIT : lower_of_uoff[u];
    std::vector<int32_t> origins;
    std::vector<float> leaf_max;
    uint32_t num_leaves = 0;
    int32_t bbmin[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, bbmax[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (uint32_t u = 0; u < RTC8X_UPPER_TABLE_SIZE; ++u) {
        uint32_t li = lower_of_uoff[u];
        if (li == 0xFFFFFFFFu) continue;
        uint32_t bx = u >> 10, by = (u >> 5) & 31, bz = u & 31;
        for (uint32_t o = 0; o < 4096; ++o) {
            uint32_t lx = bx * 16 + (o >> 8), ly = by * 16 + ((o >> 4) & 15), lz = bz * 16 + (o & 15);
            if (lx >= nl || ly >= nl || lz >= nl || !active(lx, ly, lz)) continue;
            lower[(size_t)li * 4096 + o] = num_leaves++;
            int32_t org[3] = {(int32_t)lx * 8, (int32_t)ly * 8, (int32_t)lz * 8};
            for (int a = 0; a < 3; ++a) {
                origins.push_back(org[a]);
                bbmin[a] = std::min(bbmin[a], org[a]);
                bbmax[a] = std::max(bbmax[a], org[a] + 7);
            }
            float dv = slot_max[(lx * nl + ly) * nl + lz] - sub;
            if (dv < 0.0f) dv = 0.0f;
            dv = dv * inv_range;
            dv = dv * d->contrast;
            if (dv > 1.0f) dv = 1.0f;
            leaf_max.push_back(dv);
        }
    }
    RTC8X_REQUIRE(num_leaves > 0, "rtc8x_grid_build_fbm: no active leaves");

    float vmax = 0.0f;
    for (float m : leaf_max) vmax = std::max(vmax, m);

    rtc8x_root_tile tile;
    tile.key = 0;  // root key of the upper node at the origin
    tile.entry = 0;
    tile.state = 0;
    std::vector<rtc8x_root_tile> root(1, tile);
    std::vector<float> tile_values(2, 0.0f);  // the background tile and the lookup's padding float

    rtc8x_grid g;
    std::memset(&g, 0, sizeof(g));
    for (int a = 0; a < 9; ++a) g.inv_mat[a] = 0.0f;
    g.inv_mat[0] = g.inv_mat[4] = g.inv_mat[8] = (float)n;
    for (int a = 0; a < 3; ++a) {
        g.vec[a] = 0.0f;
        g.index_bbox_min[a] = bbmin[a];
        g.index_bbox_max[a] = bbmax[a];
        g.world_bbox_min[a] = (float)bbmin[a] / (float)n;
        g.world_bbox_max[a] = (float)(bbmax[a] + 1) / (float)n;
    }
    g.max_value = vmax;
    g.background = 0.0f;
    g.num_root_tiles = 1;
    g.num_upper = 1;
    g.num_lower = num_lower;
    g.num_leaves = num_leaves;
    rtc8x_root_tile* d_root;
    uint32_t *d_upper, *d_lower;
    int32_t* d_origins;
    float *d_tiles, *d_values;
    RTC8X_TRY(upload_vec(root, &d_root));
    g.root_tiles = d_root;
    RTC8X_TRY(upload_vec(upper, &d_upper));
    g.upper_tables = d_upper;
    RTC8X_TRY(upload_vec(lower, &d_lower));
    g.lower_tables = d_lower;
    RTC8X_TRY(upload_vec(origins, &d_origins));
    g.leaf_origins = d_origins;
    RTC8X_TRY(upload_vec(tile_values, &d_tiles));
    g.tile_values = d_tiles;
    RTC8X_HIP(hipMalloc(&d_values, ((size_t)num_leaves * 512 + 1) * sizeof(float)));
    RTC8X_HIP(hipMemsetAsync(d_values + (size_t)num_leaves * 512, 0, sizeof(float), s));  // lookup padding
    g.leaf_values = d_values;
    hipLaunchKernelGGL(dev::fill_leaves_kernel, dim3(num_leaves), dim3(512), 0, s, *d, d_origins, sub, inv_range, d_values);
    RTC8X_HIP(hipGetLastError());
    RTC8X_HIP(hipStreamSynchronize(s));
    RTC8X_TRY(build_leaf_map(g, s));
    RTC8X_TRY(build_majorant_tables(g, 0, s));
    *out = g;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_grid_build_fbm")

namespace rtc8x {

rtc8x_status upload_host_grid(const rtc8x_grid* h, rtc8x_grid* out, bool keep_max)
{
    RTC8X_REQUIRE(h && out, "rtc8x_grid_upload: NULL argument");
    RTC8X_REQUIRE(h->num_root_tiles > 0 && h->root_tiles && h->tile_values, "rtc8x_grid_upload: root missing");
    RTC8X_REQUIRE(h->num_upper == 0 || h->upper_tables, "rtc8x_grid_upload: upper tables missing");
    RTC8X_REQUIRE(h->num_lower == 0 || h->lower_tables, "rtc8x_grid_upload: lower tables missing");
    RTC8X_REQUIRE(h->num_leaves == 0 || h->leaf_values, "rtc8x_grid_upload: leaf values missing");
    // validate every table entry against the node counts
    uint32_t num_tiles_values = 1;
    for (uint32_t t = 0; t < h->num_root_tiles; ++t) {
        uint32_t e = h->root_tiles[t].entry;
        if (e & RTC8X_TILE_BIT) num_tiles_values = std::max(num_tiles_values, (e & ~RTC8X_TILE_BIT) + 1);
        else RTC8X_REQUIRE(e < h->num_upper, "rtc8x_grid_upload: root tile %u -> upper %u out of range", t, e);
    }
    for (size_t i = 0; i < (size_t)h->num_upper * RTC8X_UPPER_TABLE_SIZE; ++i) {
        uint32_t e = h->upper_tables[i];
        if (e & RTC8X_TILE_BIT) num_tiles_values = std::max(num_tiles_values, (e & ~RTC8X_TILE_BIT) + 1);
        else RTC8X_REQUIRE(e < h->num_lower, "rtc8x_grid_upload: upper entry %zu -> lower %u out of range", i, e);
    }
    for (size_t i = 0; i < (size_t)h->num_lower * RTC8X_LOWER_TABLE_SIZE; ++i) {
        uint32_t e = h->lower_tables[i];
        if (e & RTC8X_TILE_BIT) num_tiles_values = std::max(num_tiles_values, (e & ~RTC8X_TILE_BIT) + 1);
        else RTC8X_REQUIRE(e < h->num_leaves, "rtc8x_grid_upload: lower entry %zu -> leaf %u out of range", i, e);
    }
    // leaf origins from the tables (RootData::KeyToCoord, NanoVDB.h:2707-2713, and the node offsets)
    std::vector<int32_t> origins((size_t)3 * h->num_leaves, 0);
    std::vector<uint8_t> seen(h->num_leaves, 0);
    for (uint32_t t = 0; t < h->num_root_tiles; ++t) {
        uint32_t ue = h->root_tiles[t].entry;
        if (ue & RTC8X_TILE_BIT) continue;
        uint64_t key = h->root_tiles[t].key;
        const uint64_t mask = (1u << 21) - 1;
        int32_t uo[3] = {(int32_t)(((key >> 42) & mask) << 12), (int32_t)(((key >> 21) & mask) << 12),
                         (int32_t)((key & mask) << 12)};
        for (uint32_t u = 0; u < RTC8X_UPPER_TABLE_SIZE; ++u) {
            uint32_t le = h->upper_tables[(size_t)ue * RTC8X_UPPER_TABLE_SIZE + u];
            if (le & RTC8X_TILE_BIT) continue;
            int32_t lo3[3] = {uo[0] + (int32_t)((u >> 10) << 7), uo[1] + (int32_t)(((u >> 5) & 31) << 7),
                              uo[2] + (int32_t)((u & 31) << 7)};
            for (uint32_t o = 0; o < RTC8X_LOWER_TABLE_SIZE; ++o) {
                uint32_t leaf = h->lower_tables[(size_t)le * RTC8X_LOWER_TABLE_SIZE + o];
                if (leaf & RTC8X_TILE_BIT) continue;
                origins[3 * leaf] = lo3[0] + (int32_t)((o >> 8) << 3);
                origins[3 * leaf + 1] = lo3[1] + (int32_t)(((o >> 4) & 15) << 3);
                origins[3 * leaf + 2] = lo3[2] + (int32_t)((o & 15) << 3);
                seen[leaf] = 1;
            }
        }
    }
    float vmax = -FLT_MAX;
    for (size_t v = 0; v < (size_t)h->num_leaves * 512; ++v) vmax = std::max(vmax, h->leaf_values[v]);
    for (uint32_t t = 0; t < num_tiles_values; ++t) vmax = std::max(vmax, h->tile_values[t]);
    RTC8X_REQUIRE(std::isfinite(vmax), "rtc8x_grid_upload: non-finite voxel values");
    RTC8X_REQUIRE(!keep_max || std::isfinite(h->max_value), "rtc8x_grid_upload: non-finite majorant");

    rtc8x_grid g = *h;
    g.max_value = keep_max ? h->max_value : vmax;
    std::vector<rtc8x_root_tile> root(h->root_tiles, h->root_tiles + h->num_root_tiles);
    std::vector<uint32_t> upper(h->upper_tables, h->upper_tables + (size_t)h->num_upper * RTC8X_UPPER_TABLE_SIZE);
    std::vector<uint32_t> lower(h->lower_tables, h->lower_tables + (size_t)h->num_lower * RTC8X_LOWER_TABLE_SIZE);
    std::vector<float> values(h->leaf_values, h->leaf_values + (size_t)h->num_leaves * 512);
    std::vector<float> tiles(h->tile_values, h->tile_values + num_tiles_values);
    values.push_back(0.0f);  // padding: the lookup reads voxel pairs (z, z+1) at any voxel and tile
    tiles.push_back(0.0f);
    rtc8x_root_tile* d_root;
    uint32_t *d_upper, *d_lower;
    int32_t* d_origins;
    float *d_values, *d_tiles;
    RTC8X_TRY(upload_vec(root, &d_root));
    RTC8X_TRY(upload_vec(upper, &d_upper));
    RTC8X_TRY(upload_vec(lower, &d_lower));
    RTC8X_TRY(upload_vec(values, &d_values));
    RTC8X_TRY(upload_vec(tiles, &d_tiles));
    RTC8X_TRY(upload_vec(origins, &d_origins));
    g.root_tiles = d_root;
    g.upper_tables = d_upper;
    g.lower_tables = d_lower;
    g.leaf_values = d_values;
    g.tile_values = d_tiles;
    g.leaf_origins = d_origins;
    g.cell_majorants = nullptr;
    g.region_majorants = nullptr;
    g.leaf_map = nullptr;
    g.leaf_masks = g.lower_masks = g.upper_masks = nullptr;  // host-side activity is not uploaded
    if (h->num_root_tiles > 0 && h->num_leaves > 0) RTC8X_TRY(build_leaf_map(g, nullptr));
    RTC8X_TRY(build_majorant_tables(g, 0, nullptr));
    *out = g;
    return RTC8X_OK;
}

}  // namespace rtc8x

extern "C" rtc8x_status rtc8x_grid_upload(const rtc8x_grid* h, rtc8x_grid* out)
try {
    return rtc8x::upload_host_grid(h, out, false);
}
RTC8X_CATCH_ALL("rtc8x_grid_upload")

extern "C" rtc8x_status rtc8x_grid_download(const rtc8x_grid* d, rtc8x_grid* out)
try {
    RTC8X_REQUIRE(d && out, "rtc8x_grid_download: NULL argument");
    RTC8X_REQUIRE(d->num_root_tiles > 0 && d->root_tiles && d->tile_values, "rtc8x_grid_download: grid has no root");
    rtc8x::OwnedHostGrid t;
    t.root.resize(d->num_root_tiles);
    t.upper.resize((size_t)d->num_upper * RTC8X_UPPER_TABLE_SIZE);
    t.lower.resize((size_t)d->num_lower * RTC8X_LOWER_TABLE_SIZE);
    t.values.resize((size_t)d->num_leaves * RTC8X_LEAF_SIZE);
    RTC8X_HIP(hipDeviceSynchronize());
    RTC8X_HIP(hipMemcpy(t.root.data(), d->root_tiles, t.root.size() * sizeof(rtc8x_root_tile), hipMemcpyDeviceToHost));
    if (!t.upper.empty())
        RTC8X_HIP(hipMemcpy(t.upper.data(), d->upper_tables, t.upper.size() * 4, hipMemcpyDeviceToHost));
    if (!t.lower.empty())
        RTC8X_HIP(hipMemcpy(t.lower.data(), d->lower_tables, t.lower.size() * 4, hipMemcpyDeviceToHost));
    if (!t.values.empty())
        RTC8X_HIP(hipMemcpy(t.values.data(), d->leaf_values, t.values.size() * 4, hipMemcpyDeviceToHost));
    // the tile-value table is as long as the largest index a table entry names
    uint32_t num_tiles = 1;
    auto scan = [&](uint32_t e) {
        if (e & RTC8X_TILE_BIT) num_tiles = std::max(num_tiles, (e & ~RTC8X_TILE_BIT) + 1);
    };
    for (const rtc8x_root_tile& r : t.root) scan(r.entry);
    for (uint32_t e : t.upper) scan(e);
    for (uint32_t e : t.lower) scan(e);
    t.tiles.resize(num_tiles);
    RTC8X_HIP(hipMemcpy(t.tiles.data(), d->tile_values, num_tiles * 4, hipMemcpyDeviceToHost));
    if (d->leaf_origins && d->num_leaves) {
        t.origins.resize((size_t)3 * d->num_leaves);
        RTC8X_HIP(hipMemcpy(t.origins.data(), d->leaf_origins, t.origins.size() * 4, hipMemcpyDeviceToHost));
    }
    const size_t ncells = (size_t)std::max(d->cell_dims[0], 0) * std::max(d->cell_dims[1], 0) * std::max(d->cell_dims[2], 0);
    if (d->cell_majorants && ncells) {
        t.cells.resize(ncells);
        RTC8X_HIP(hipMemcpy(t.cells.data(), d->cell_majorants, ncells * 4, hipMemcpyDeviceToHost));
    }
    const size_t nregions =
        (size_t)std::max(d->region_dims[0], 0) * std::max(d->region_dims[1], 0) * std::max(d->region_dims[2], 0);
    if (d->region_majorants && nregions) {
        t.regions.resize(nregions);
        RTC8X_HIP(hipMemcpy(t.regions.data(), d->region_majorants, nregions, hipMemcpyDeviceToHost));
    }
    rtc8x_grid g = *d;
    g.leaf_map = nullptr;
    for (int a = 0; a < 3; ++a) {
        g.leaf_map_origin[a] = 0;
        g.leaf_map_dims[a] = 0;
    }
    return rtc8x::adopt_host_grid(std::move(t), g, out);
}
RTC8X_CATCH_ALL("rtc8x_grid_download")

extern "C" rtc8x_status rtc8x_grid_probe(const rtc8x_grid* g, uint32_t n, const float* world_xyz, int32_t* out_ijk,
                                         uint32_t* out_offsets, float* out_value, void* stream)
try {
    RTC8X_REQUIRE(g && world_xyz && out_ijk && out_offsets && out_value, "rtc8x_grid_probe: NULL argument");
    RTC8X_REQUIRE(g->num_root_tiles > 0 && g->root_tiles && g->tile_values, "rtc8x_grid_probe: grid has no root");
    if (n == 0) return RTC8X_OK;
    hipLaunchKernelGGL(dev::probe_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, *g, n, world_xyz,
                       out_ijk, out_offsets, out_value);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_grid_probe")

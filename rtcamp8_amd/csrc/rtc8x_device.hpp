// rtc8x_device.hpp -- device functions of the volumetric path (gfx950).
//
// Every arithmetic expression on the path mirrors the reference expression it replaces, in the same
// evaluation order, with IEEE single precision, no contraction (-ffp-contract=off) and the
// deterministic transcendental functions of rtc8x_detmath.h.  This is what makes a gfx950 launch
// reproduce a scalar host run of the same algorithm bit for bit on the same PCG streams.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtc8x.h"
#include "rtc8x_detmath.h"

namespace rtc8x {
namespace dev {

struct V3 {
    float x, y, z;
};
struct RGB {
    float r, g, b;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ RGB rgb(float r, float g, float b) { return RGB{r, g, b}; }
__device__ __forceinline__ RGB mul(RGB a, RGB b) { return RGB{a.r * b.r, a.g * b.g, a.b * b.b}; }
__device__ __forceinline__ RGB scale(RGB a, float s) { return RGB{a.r * s, a.g * s, a.b * s}; }
__device__ __forceinline__ RGB add(RGB a, RGB b) { return RGB{a.r + b.r, a.g + b.g, a.b + b.b}; }
// RGBSpectrum::luminance (common/basic_types.h:2981-2986)
__device__ __forceinline__ float luminance(RGB c)
{
    return (float)0.2126729 * c.r + (float)0.7151522 * c.g + (float)0.0721750 * c.b;
}
__device__ __forceinline__ bool all_nonneg_finite(RGB c)
{
    return c.r >= 0.0f && c.g >= 0.0f && c.b >= 0.0f && isfinite(c.r) && isfinite(c.g) && isfinite(c.b);
}
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// normalize(): multiply by 1/length (common/basic_types.h:817-820, 850-853, 1871-1875)
__device__ __forceinline__ V3 normalize(V3 v)
{
    float l = sqrtf(dot(v, v));
    float r = 1.0f / l;
    return V3{v.x * r, v.y * r, v.z * r};
}

// PCG32 XSH-RR, increment 1 (common/common_renderer_types.h:101-125)
struct Pcg {
    uint64_t state;
    __device__ __forceinline__ uint32_t next()
    {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + 1ULL;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
    }
    __device__ __forceinline__ float uni()
    {
        uint32_t bits = (next() >> 9) | 0x3f800000u;
        return __uint_as_float(bits) - 1.0f;
    }
};

struct Frame {
    V3 t, b, n;
};
// ReferenceFrame(normal) via makeCoordinateSystem (basic_types.h:857-864)
__device__ __forceinline__ Frame make_frame(V3 n)
{
    Frame f;
    float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    f.t = V3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    f.b = V3{b, sign + n.y * n.y * a, -n.y};
    f.n = n;
    return f;
}
__device__ __forceinline__ V3 to_local(const Frame& f, V3 v) { return V3{dot(f.t, v), dot(f.b, v), dot(f.n, v)}; }
__device__ __forceinline__ V3 from_local(const Frame& f, V3 v)
{
    return V3{dot(V3{f.t.x, f.b.x, f.n.x}, v), dot(V3{f.t.y, f.b.y, f.n.y}, v), dot(V3{f.t.z, f.b.z, f.n.z}, v)};
}

// ------------------------------------------------------------------ grid access
// NanoVDB node offsets (NanoVDB.h:2699-2706, 3217-3227, 3821-3828).
__device__ __forceinline__ uint64_t root_key(int32_t i, int32_t j, int32_t k)
{
    return ((uint64_t)((uint32_t)k >> 12)) | ((uint64_t)((uint32_t)j >> 12) << 21) |
           ((uint64_t)((uint32_t)i >> 12) << 42);
}
__device__ __forceinline__ uint32_t upper_offset(int32_t i, int32_t j, int32_t k)
{
    return ((((uint32_t)i & 4095u) >> 7) << 10) | ((((uint32_t)j & 4095u) >> 7) << 5) | (((uint32_t)k & 4095u) >> 7);
}
__device__ __forceinline__ uint32_t lower_offset(int32_t i, int32_t j, int32_t k)
{
    return ((((uint32_t)i & 127u) >> 3) << 8) | ((((uint32_t)j & 127u) >> 3) << 4) | (((uint32_t)k & 127u) >> 3);
}
__device__ __forceinline__ uint32_t leaf_offset(int32_t i, int32_t j, int32_t k)
{
    return (((uint32_t)i & 7u) << 6) | (((uint32_t)j & 7u) << 3) | ((uint32_t)k & 7u);
}

// Walks root -> upper -> lower; returns the leaf index, or RTC8X_TILE_BIT | tile value index.
__device__ __forceinline__ uint32_t find_leaf(const rtc8x_grid& g, int32_t i, int32_t j, int32_t k)
{
    uint64_t key = root_key(i, j, k);
    uint32_t entry = RTC8X_TILE_BIT;
    for (uint32_t t = 0; t < g.num_root_tiles; ++t) {
        if (g.root_tiles[t].key == key) {
            entry = g.root_tiles[t].entry;
            break;
        }
    }
    if (entry & RTC8X_TILE_BIT) return entry;
    uint32_t e1 = g.upper_tables[(size_t)entry * RTC8X_UPPER_TABLE_SIZE + upper_offset(i, j, k)];
    if (e1 & RTC8X_TILE_BIT) return e1;
    return g.lower_tables[(size_t)e1 * RTC8X_LOWER_TABLE_SIZE + lower_offset(i, j, k)];
}

// Tree walk through the dense leaf map when the block is inside it (the map caches find_leaf).
__device__ __forceinline__ uint32_t block_entry(const rtc8x_grid& g, int32_t i, int32_t j, int32_t k)
{
    const int32_t bx = (i >> 3) - (g.leaf_map_origin[0] >> 3), by = (j >> 3) - (g.leaf_map_origin[1] >> 3),
                  bz = (k >> 3) - (g.leaf_map_origin[2] >> 3);
    if (g.leaf_map && (uint32_t)bx < (uint32_t)g.leaf_map_dims[0] && (uint32_t)by < (uint32_t)g.leaf_map_dims[1] &&
        (uint32_t)bz < (uint32_t)g.leaf_map_dims[2])
        return g.leaf_map[((size_t)bx * (uint32_t)g.leaf_map_dims[1] + (uint32_t)by) * (uint32_t)g.leaf_map_dims[2] +
                          (uint32_t)bz];
    return find_leaf(g, i, j, k);
}

__device__ __forceinline__ const float* voxel_ptr(const rtc8x_grid& g, uint32_t e, int32_t i, int32_t j, int32_t k)
{
    return (e & RTC8X_TILE_BIT) ? g.tile_values + (e & ~RTC8X_TILE_BIT)
                                : g.leaf_values + (size_t)e * RTC8X_LEAF_SIZE + leaf_offset(i, j, k);
}

// Map::applyInverseMapF with matMult (NanoVDB.h:1409-1414, 2029-2032)
__device__ __forceinline__ V3 world_to_index(const rtc8x_grid& g, float px, float py, float pz)
{
    float x = px - g.vec[0], y = py - g.vec[1], z = pz - g.vec[2];
    const float* m = g.inv_mat;
    return V3{fmaf(x, m[0], fmaf(y, m[1], z * m[2])), fmaf(x, m[3], fmaf(y, m[4], z * m[5])),
              fmaf(x, m[6], fmaf(y, m[7], z * m[8]))};
}

struct F2 {
    float x, y;
};

// SampleFromVoxels<Acc,1,false>::operator() (util/SampleFromVoxels.h:419-460) with Floor (:58-65), for
// S index-space positions at once.  All loads are issued in two rounds: first every position's leaf-map
// entries (the stencil's block, and the other seven only for a stencil that crosses a block face), then
// every position's voxels, so the lookups of a batch overlap instead of each waiting on the one before.
// Positions outside the map take the tree walk.  The lerps are the reference's, in its order.
template <uint32_t S>
__device__ __forceinline__ void trilinear_batch(const rtc8x_grid& g, const V3 (&idx)[S], float (&out)[S])
{
    const int32_t ob0 = g.leaf_map_origin[0] >> 3, ob1 = g.leaf_map_origin[1] >> 3, ob2 = g.leaf_map_origin[2] >> 3;
    const int32_t d0 = g.leaf_map_dims[0], d1 = g.leaf_map_dims[1], d2 = g.leaf_map_dims[2];
    int32_t i[S], j[S], k[S];
    float x[S], y[S], z[S];
    bool cross[S];
    uint32_t e[S][8];  // taps 000, 001, 010, 011, 100, 101, 110, 111 (x y z)
#pragma unroll
    for (uint32_t s = 0; s < S; ++s) {
        const float fx = floorf(idx[s].x), fy = floorf(idx[s].y), fz = floorf(idx[s].z);
        x[s] = idx[s].x - fx;
        y[s] = idx[s].y - fy;
        z[s] = idx[s].z - fz;
        i[s] = (int32_t)fx;
        j[s] = (int32_t)fy;
        k[s] = (int32_t)fz;
        const int32_t bx0 = (i[s] >> 3) - ob0, bx1 = ((i[s] + 1) >> 3) - ob0;
        const int32_t by0 = (j[s] >> 3) - ob1, by1 = ((j[s] + 1) >> 3) - ob1;
        const int32_t bz0 = (k[s] >> 3) - ob2, bz1 = ((k[s] + 1) >> 3) - ob2;
        const bool cx = (i[s] & 7) == 7, cy = (j[s] & 7) == 7, cz = (k[s] & 7) == 7;
        cross[s] = cx || cy || cz;
#pragma unroll
        for (int t = 0; t < 8; ++t) e[s][t] = 0u;
        if (g.leaf_map && bx0 >= 0 && by0 >= 0 && bz0 >= 0 && bx1 < d0 && by1 < d1 && bz1 < d2) {
            const uint32_t* m = g.leaf_map + ((size_t)bx0 * (uint32_t)d1 + (uint32_t)by0) * (uint32_t)d2 + (uint32_t)bz0;
            e[s][0] = m[0];
            if (cross[s]) {
                const size_t sz = cz ? 1u : 0u, sy = cy ? (size_t)(uint32_t)d2 : 0u,
                             sx = cx ? (size_t)(uint32_t)d1 * (uint32_t)d2 : 0u;
                e[s][1] = m[sz];
                e[s][2] = m[sy];
                e[s][3] = m[sy + sz];
                e[s][4] = m[sx];
                e[s][5] = m[sx + sz];
                e[s][6] = m[sx + sy];
                e[s][7] = m[sx + sy + sz];
            }
        } else {
            cross[s] = true;
#pragma unroll
            for (int t = 0; t < 8; ++t) e[s][t] = find_leaf(g, i[s] + (t >> 2), j[s] + ((t >> 1) & 1), k[s] + (t & 1));
        }
    }
    // Voxel pairs (z, z+1) of the four (x, y) columns: one 8-byte load per column from the column's
    // lower block (the grid keeps one float of padding after the leaf and tile values, so the second
    // word is always readable); a stencil crossing a z face also loads its upper voxels one by one.
    F2 pr[S][4];
    float up[S][4];
#pragma unroll
    for (uint32_t s = 0; s < S; ++s) {
        const bool cz = (k[s] & 7) == 7;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int32_t ii = i[s] + (c >> 1), jj = j[s] + (c & 1);
            const uint32_t lo = cross[s] ? e[s][2 * c] : e[s][0];
            const float* pl = (lo & RTC8X_TILE_BIT) ? g.tile_values + (lo & ~RTC8X_TILE_BIT)
                                                    : g.leaf_values + (size_t)lo * RTC8X_LEAF_SIZE + leaf_offset(ii, jj, k[s]);
            __builtin_memcpy(&pr[s][c], pl, sizeof(F2));
            up[s][c] = 0.0f;
            if (cz) {
                const uint32_t hi = e[s][2 * c + 1];
                up[s][c] = *((hi & RTC8X_TILE_BIT) ? g.tile_values + (hi & ~RTC8X_TILE_BIT)
                                                   : g.leaf_values + (size_t)hi * RTC8X_LEAF_SIZE + leaf_offset(ii, jj, k[s] + 1));
            }
        }
    }
#pragma unroll
    for (uint32_t s = 0; s < S; ++s) {
        const bool cz = (k[s] & 7) == 7;
        float v[8];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint32_t lo = cross[s] ? e[s][2 * c] : e[s][0];
            v[2 * c] = pr[s][c].x;
            v[2 * c + 1] = cz ? up[s][c] : ((lo & RTC8X_TILE_BIT) ? pr[s][c].x : pr[s][c].y);
        }
        const float zz = z[s], yy = y[s];
        float vy = v[0] + zz * (v[1] - v[0]);
        float vy1 = v[2] + zz * (v[3] - v[2]);
        const float vx = vy + yy * (vy1 - vy);
        vy1 = v[6] + zz * (v[7] - v[6]);
        vy = v[4] + zz * (v[5] - v[4]);
        const float vx1 = vy + yy * (vy1 - vy);
        out[s] = vx + x[s] * (vx1 - vx);
    }
}

__device__ __forceinline__ float trilinear(const rtc8x_grid& g, V3 idx)
{
    const V3 in[1] = {idx};
    float out[1];
    trilinear_batch<1>(g, in, out);
    return out[0];
}

struct Medium {
    float coeff;
    float majorant;
    uint32_t lookups;
};

// Density coefficient times the trilinear grid value at o + t d (no lookup counting).
__device__ __forceinline__ float density_at(const rtc8x_grid& g, float coeff, V3 o, V3 d, float t)
{
    V3 idx = world_to_index(g, fmaf(t, d.x, o.x), fmaf(t, d.y, o.y), fmaf(t, d.z, o.z));
    return coeff * trilinear(g, idx);
}

// Density coefficient times the grid value at o + t[s] d for S distances, looked up together.
template <uint32_t S>
__device__ __forceinline__ void density_batch(const rtc8x_grid& g, float coeff, V3 o, V3 d, const float (&t)[S],
                                              float (&sig)[S])
{
    V3 idx[S];
#pragma unroll
    for (uint32_t s = 0; s < S; ++s) idx[s] = world_to_index(g, fmaf(t[s], d.x, o.x), fmaf(t[s], d.y, o.y), fmaf(t[s], d.z, o.z));
    trilinear_batch<S>(g, idx, sig);
#pragma unroll
    for (uint32_t s = 0; s < S; ++s) sig[s] = coeff * sig[s];
}

__device__ __forceinline__ float sample_density(const rtc8x_grid& g, Medium& m, V3 o, V3 d, float t)
{
    V3 idx = world_to_index(g, fmaf(t, d.x, o.x), fmaf(t, d.y, o.y), fmaf(t, d.z, o.z));
    m.lookups++;
    return m.coeff * trilinear(g, idx);
}

// nanovdb::Ray::clip against the float world bbox (util/Ray.h:63-73, 427-455, 512-520)
__device__ __forceinline__ bool clip(const rtc8x_grid& g, V3 o, V3 d, float tmax, float& t0o, float& t1o)
{
    float inv[3] = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    float e[3] = {o.x, o.y, o.z};
    float t0 = 0.0f, t1 = tmax;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float lo = g.world_bbox_min[a], hi = g.world_bbox_max[a];
        if (lo >= hi) return false;
        lo = (lo - e[a]) * inv[a];
        hi = (hi - e[a]) * inv[a];
        if (lo > hi) {
            float t = lo;
            lo = hi;
            hi = t;
        }
        if (lo > t0) t0 = lo;
        if (hi < t1) t1 = hi;
        if (t0 > t1) return false;
    }
    t0o = t0;
    t1o = t1;
    return true;
}

// -------------------------------------------------------------- local majorants
// 3D DDA over the majorant regions (region_cells^3 cells of 8^3 voxels) in index space; boundaries
// recomputed from the region index.  Per-axis state is held in named scalars and selected by axis (no
// indexed arrays, so it stays in registers).
__device__ __forceinline__ int32_t region_size(const rtc8x_grid& g) { return 8 * (int32_t)g.region_cells; }

struct Dda {
    int32_t c0, c1, c2;      // cell
    int32_t s0, s1, s2;      // step per axis (-1, 0, +1)
    float io0, io1, io2;     // ray origin in index space
    float iv0, iv1, iv2;     // 1 / index-space direction
    float tn0, tn1, tn2;     // ray parameter of the next cell boundary per axis
    static __device__ __forceinline__ float self(int a, float x, float y, float z) { return a == 0 ? x : (a == 1 ? y : z); }
    static __device__ __forceinline__ int32_t seli(int a, int32_t x, int32_t y, int32_t z)
    {
        return a == 0 ? x : (a == 1 ? y : z);
    }
    __device__ __forceinline__ int32_t cell(int a) const { return seli(a, c0, c1, c2); }
    __device__ __forceinline__ float tnext(int a) const { return self(a, tn0, tn1, tn2); }
    static __device__ __forceinline__ float boundary_of(int32_t origin, int32_t size, int32_t c, int32_t st, float io,
                                                        float iv)
    {
        int32_t b = origin + size * (c + (st > 0 ? 1 : 0));
        return ((float)b - io) * iv;
    }
    // the region holding the ray at t (clamped to the grid) and the axis' next boundary
    static __device__ __forceinline__ void axis_cell(float t, float io, float id, float iv, int32_t st, int32_t origin,
                                                     int32_t size, int32_t dims, int32_t& c, float& tn)
    {
        float pa = fmaf(t, id, io);
        float cf = floorf((pa - (float)origin) * (1.0f / (float)size));
        c = cf < 0.0f ? 0 : (cf > (float)(dims - 1) ? dims - 1 : (int32_t)cf);
        tn = st != 0 ? boundary_of(origin, size, c, st, io, iv) : __builtin_huge_valf();
    }
    static __device__ __forceinline__ void axis_init(float t, float io, float id, int32_t origin, int32_t size,
                                                     int32_t dims, int32_t& c, int32_t& st, float& iv, float& tn)
    {
        st = id > 0.0f ? 1 : (id < 0.0f ? -1 : 0);
        iv = 1.0f / id;
        axis_cell(t, io, id, iv, st, origin, size, dims, c, tn);
    }
    static __device__ __forceinline__ V3 index_dir(const rtc8x_grid& g, V3 d)
    {
        const float* mm = g.inv_mat;
        return V3{fmaf(d.x, mm[0], fmaf(d.y, mm[1], d.z * mm[2])), fmaf(d.x, mm[3], fmaf(d.y, mm[4], d.z * mm[5])),
                  fmaf(d.x, mm[6], fmaf(d.y, mm[7], d.z * mm[8]))};
    }
    __device__ __forceinline__ void init(const rtc8x_grid& g, V3 o, V3 d, float t)
    {
        const float* mm = g.inv_mat;
        float x = o.x - g.vec[0], y = o.y - g.vec[1], z = o.z - g.vec[2];
        io0 = fmaf(x, mm[0], fmaf(y, mm[1], z * mm[2]));
        io1 = fmaf(x, mm[3], fmaf(y, mm[4], z * mm[5]));
        io2 = fmaf(x, mm[6], fmaf(y, mm[7], z * mm[8]));
        const V3 id = index_dir(g, d);
        const int32_t size = region_size(g);
        axis_init(t, io0, id.x, g.cell_origin[0], size, g.region_dims[0], c0, s0, iv0, tn0);
        axis_init(t, io1, id.y, g.cell_origin[1], size, g.region_dims[1], c1, s1, iv1, tn1);
        axis_init(t, io2, id.z, g.cell_origin[2], size, g.region_dims[2], c2, s2, iv2, tn2);
    }
    // init(g, o, d, t) for the same ray at a new t, bit for bit, without the divisions: the index-space
    // origin, steps and reciprocals are those init computed; only the regions and boundaries are redone.
    __device__ __forceinline__ void jump(const rtc8x_grid& g, V3 d, float t)
    {
        const V3 id = index_dir(g, d);
        const int32_t size = region_size(g);
        axis_cell(t, io0, id.x, iv0, s0, g.cell_origin[0], size, g.region_dims[0], c0, tn0);
        axis_cell(t, io1, id.y, iv1, s1, g.cell_origin[1], size, g.region_dims[1], c1, tn1);
        axis_cell(t, io2, id.z, iv2, s2, g.cell_origin[2], size, g.region_dims[2], c2, tn2);
    }
    __device__ __forceinline__ int axis() const
    {
        int a = 0;
        float best = tn0;
        if (tn1 < best) {
            a = 1;
            best = tn1;
        }
        if (tn2 < best) a = 2;
        return a;
    }
    // The ray parameter at which the ray leaves the cube of regions within k of the current one: the
    // first crossing of its faces (boundaries computed as the steps compute them).
    __device__ __forceinline__ float skip_exit(const rtc8x_grid& g, int32_t k) const
    {
        const int32_t size = region_size(g);
        float tx = __builtin_huge_valf();
        if (s0 != 0) tx = fminf(tx, boundary_of(g.cell_origin[0], size, c0 + s0 * k, s0, io0, iv0));
        if (s1 != 0) tx = fminf(tx, boundary_of(g.cell_origin[1], size, c1 + s1 * k, s1, io1, iv1));
        if (s2 != 0) tx = fminf(tx, boundary_of(g.cell_origin[2], size, c2 + s2 * k, s2, io2, iv2));
        return tx;
    }
    // Steps the cell along axis a; false when it leaves the cell grid.  Updates that axis' boundary.
    // Written with value selects and unconditional stores: per-axis branches that each store a field
    // get merged into one store through a selected field address, which keeps the state in memory.
    __device__ __forceinline__ bool step(const rtc8x_grid& g, int a)
    {
        const int32_t st = seli(a, s0, s1, s2);
        const int32_t c = seli(a, c0, c1, c2) + st;
        if (c < 0 || c >= seli(a, g.region_dims[0], g.region_dims[1], g.region_dims[2])) return false;
        const float tn = boundary_of(seli(a, g.cell_origin[0], g.cell_origin[1], g.cell_origin[2]), region_size(g), c,
                                     st, self(a, io0, io1, io2), self(a, iv0, iv1, iv2));
        c0 = a == 0 ? c : c0;
        c1 = a == 1 ? c : c1;
        c2 = a == 2 ? c : c2;
        tn0 = a == 0 ? tn : tn0;
        tn1 = a == 1 ? tn : tn1;
        tn2 = a == 2 ? tn : tn2;
        return true;
    }
};

// Region codes: 1..127 is q for a region with density, majorant densityCoeff x (q x region_unit);
// RTC8X_REGION_EMPTY + k marks an empty region whose neighbours up to k regions away (Chebyshev distance)
// are empty too.  `table` is the grid's region bytes or a copy of them staged in LDS.
__device__ __forceinline__ uint32_t region_code(const rtc8x_grid& g, const Dda& s, const uint8_t* __restrict__ table)
{
    return table[((uint32_t)s.c0 * (uint32_t)g.region_dims[1] + (uint32_t)s.c1) * (uint32_t)g.region_dims[2] +
                 (uint32_t)s.c2];
}
__device__ __forceinline__ float region_majorant(const rtc8x_grid& g, const Medium& m, uint32_t code)
{
    return code < RTC8X_REGION_EMPTY ? m.coeff * ((float)code * g.region_unit) : 0.0f;
}

// One tracking or shadow segment of a path, resumable: Seg<TRACKING> per tracking mode (the tracer
// defines the global-majorant one next to its lookup batching).
template <int TRACKING>
struct Seg;

// Regions a local-majorant lane walks per stepping pass before it yields to the wavefront's lookups.
constexpr uint32_t kRegionBudget = 6;
enum : int { kAdvanceEnd = 0, kAdvanceCollision = 1, kAdvanceBudget = 2 };

// Local majorants over the grid's regions, the loops of track_local / ratio_local made resumable; the
// region bytes come from `table` (a copy staged in LDS by the single-image kernel).
template <>
struct Seg<RTC8X_TRACKING_LOCAL_MAJORANT> {
    V3 o, d;
    float t, t1, tau;
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
        return true;
    }
    // first call after init draws tau (as track_local does right after the DDA set-up)
    __device__ __forceinline__ void first_tau(Pcg& rng) { tau = -rtc8x_logf(1.0f - rng.uni()); }
    // Walks the regions up to the next tentative collision: kAdvanceCollision with t and mu set,
    // kAdvanceEnd when the segment ends first, kAdvanceBudget after kRegionBudget region visits without
    // either (the walk resumes where it stopped on the next call, so every lane of a wavefront returns to
    // the density lookups after a bounded number of region steps).  A visit derives everything it needs
    // (the region's code and majorant, the axis and end of its span) from the walk's position, so a call
    // that resumes inside a region after a null collision recomputes them instead of keeping them.
    __device__ __forceinline__ int advance(const rtc8x_grid& g, const Medium& m, Pcg&, float& mu,
                                           const uint8_t* __restrict__ table)
    {
        int res = kAdvanceBudget;
        float mu_r = 0.0f;
#pragma unroll 1
        for (uint32_t visit = 0; visit < kRegionBudget; ++visit) {
            const uint32_t code = region_code(g, s, table);
            if (code > RTC8X_REGION_EMPTY) {
                // an empty region with k empty layers around it: leave their cube in one jump
                const float tx = s.skip_exit(g, (int32_t)(code - RTC8X_REGION_EMPTY));
                if (tx > t) {
                    if (!(tx < t1)) {
                        res = kAdvanceEnd;
                        break;
                    }
                    t = tx;
                    s.jump(g, d, t);
                    continue;
                }
            }
            const int axis = s.axis();
            const float tna = s.tnext(axis);
            const float tend = tna < t1 ? tna : t1;
            mu_r = region_majorant(g, m, code);
            if (mu_r > 0.0f) {
                // the optical depth left in the region decides; the division only for a collision
                const float need = mu_r * (tend - t);
                if (tau < need) {
                    res = kAdvanceCollision;
                    break;
                }
                tau = tau - need;
            }
            if (!(tna < t1)) {
                res = kAdvanceEnd;
                break;
            }
            t = tend;
            if (!s.step(g, axis)) {
                res = kAdvanceEnd;
                break;
            }
        }
        if (res == kAdvanceCollision) {
            t = t + tau / mu_r;
            mu = mu_r;
        }
        return res;
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


// RadianceQuery of a medium vertex (optix_pathtracing_kernels.cu:14-29): AABB-local position, the
// reference's toPolarYUp(&phi, &theta) argument order (polar angle stored in *_phi), roughness 0,
// diffuse reflectance = albedo.
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

// -------------------------------------------------------------- phase functions
// SchlickPhaseFunction (common_renderer_types.h:2063-2097) and Henyey-Greenstein.
__device__ __forceinline__ float phase_value(int kind, float k, float cos_t)
{
    if (kind == RTC8X_PHASE_SCHLICK) {
        float dterm = 1.0f - k * cos_t;
        return (1.0f - k * k) / (4.0f * RTC8X_PI_F * (dterm * dterm));
    }
    float denom = 1.0f + k * k - 2.0f * k * cos_t;
    return (1.0f - k * k) / (4.0f * RTC8X_PI_F * (denom * sqrtf(denom)));
}

__device__ __forceinline__ float phase_sample(int kind, float k, float u0, float u1, V3& dir_local)
{
    float cos_t;
    if (kind == RTC8X_PHASE_SCHLICK) {
        cos_t = (2.0f * u1 + k - 1.0f) / (2.0f * k * u1 - k + 1.0f);
    } else {
        if ((k < 0.0f ? -k : k) < (float)1e-3) {
            cos_t = 1.0f - 2.0f * u1;
        } else {
            float sq = (1.0f - k * k) / (1.0f - k + 2.0f * k * u1);
            cos_t = (1.0f + k * k - sq * sq) / (2.0f * k);
        }
    }
    cos_t = cos_t < -1.0f ? -1.0f : (cos_t > 1.0f ? 1.0f : cos_t);
    float phi = 2.0f * RTC8X_PI_F * u0;
    float value = phase_value(kind, k, cos_t);
    float sin_t = sqrtf(1.0f - cos_t * cos_t);
    float sp, cp;
    rtc8x_sincosf(phi, &sp, &cp);
    dir_local = V3{cp * sin_t, sp * sin_t, -cos_t};
    return value;
}

// -------------------------------------------------------------- environment light
__device__ __forceinline__ RGB env_texel(const rtc8x_env_light& e, float u, float v)
{
    uint32_t x = u >= 0.0f ? (uint32_t)(u * (float)e.width) : 0u;
    uint32_t y = v >= 0.0f ? (uint32_t)(v * (float)e.height) : 0u;
    if (x >= e.width) x = e.width - 1;
    if (y >= e.height) y = e.height - 1;
    const float4 t = *reinterpret_cast<const float4*>(e.texels + ((size_t)y * e.width + x) * 4);
    return RGB{t.x, t.y, t.z};
}

__device__ __forceinline__ float env_texel_prob(const rtc8x_env_light& e, uint32_t px, uint32_t py)
{
    if (e.sampler == RTC8X_ENV_SAMPLER_ALIAS) {
        if (px >= e.width) px = e.width - 1;
        if (py >= e.height) py = e.height - 1;
        return e.alias_weights[(size_t)py * e.width + px] / e.alias_integral;
    }
    if (px >= e.map_dim_x) px = e.map_dim_x - 1;
    if (py >= e.map_dim_y) py = e.map_dim_y - 1;
    return e.mips[(size_t)py * e.map_dim_x + px] / e.map_integral;
}

// EnvironmentalLight::evaluate + ImageBasedEnvironmentalLight::evaluate (common_renderer_types.h:1002-1013, 2442-2458)
__device__ __forceinline__ RGB env_evaluate(const rtc8x_env_light& e, V3 dir, float& pdf)
{
    if (e.kind == RTC8X_ENV_CONSTANT) {
        pdf = 1.0f / (4.0f * RTC8X_PI_F);
        return scale(RGB{e.constant_radiance[0], e.constant_radiance[1], e.constant_radiance[2]}, e.power_coeff);
    }
    const float two_pi = 2.0f * RTC8X_PI_F;
    float y = dir.y < -1.0f ? -1.0f : (dir.y > 1.0f ? 1.0f : dir.y);
    float theta = rtc8x_acosf(y);
    float pos_phi = rtc8x_fmodf(rtc8x_atan2f(-dir.x, dir.z) + two_pi, two_pi);
    float phi = pos_phi + e.rotation;
    phi = phi - floorf(phi / two_pi) * two_pi;
    float u = phi / two_pi, v = theta / RTC8X_PI_F;
    RGB tex = env_texel(e, u, v);
    uint32_t px = (uint32_t)((float)e.width * u), py = (uint32_t)((float)e.height * v);
    float prob = env_texel_prob(e, px, py);
    float uv_pd = (float)(e.width * e.height) * prob;
    float sp, cp;
    rtc8x_sincosf(theta, &sp, &cp);
    pdf = uv_pd / (2.0f * (RTC8X_PI_F * RTC8X_PI_F) * sp);
    if (!isfinite(pdf)) {
        pdf = 0.0f;
        return RGB{0.0f, 0.0f, 0.0f};
    }
    return scale(tex, e.power_coeff);
}

// HierarchicalImportanceMap::sample (common_renderer_types.h:479-548) on explicit mip arrays.
__device__ __forceinline__ void map_sample(const rtc8x_env_light& e, float u0, float u1, uint32_t& ox, uint32_t& oy,
                                           float& ru0, float& ru1, float& prob)
{
    uint32_t ix = 0, iy = 0;
    uint32_t nlev = e.num_mip_levels;
    uint32_t cdx = 2, cdy = e.map_dim_x > e.map_dim_y ? 1 : 2;
    float pr = 1.0f;
    // offset of level lev = sum of the sizes of levels 0..lev-1
    size_t off_top = 0;
    for (uint32_t l = 0; l + 2 < nlev; ++l) {
        uint32_t dx = max(e.map_dim_x >> l, 1u), dy = max(e.map_dim_y >> l, 1u);
        off_top += (size_t)dx * dy;
    }
    size_t off = off_top;
    for (int32_t lev = (int32_t)nlev - 2; lev >= 0; --lev) {
        ix *= 2;
        iy *= 2;
        uint32_t dx = max(e.map_dim_x >> lev, 1u);
        const float* mip = e.mips + off;
        float vul = mip[(size_t)iy * dx + ix];
        float vur = (ix + 1 < cdx) ? mip[(size_t)iy * dx + ix + 1] : 0.0f;
        float vll = (iy + 1 < cdy) ? mip[(size_t)(iy + 1) * dx + ix] : 0.0f;
        float vlr = (ix + 1 < cdx && iy + 1 < cdy) ? mip[(size_t)(iy + 1) * dx + ix + 1] : 0.0f;
        float left = vll + vul, right = vlr + vur;
        u0 *= (left + right);
        pr *= 1.0f / (left + right);
        float upper, lower;
        if (u0 < left) {
            u0 = (u0 - 0.0f) / left;
            upper = vul;
            lower = vll;
        } else {
            ix += 1;
            u0 = (u0 - left) / right;
            upper = vur;
            lower = vlr;
        }
        u1 *= (upper + lower);
        if (u1 < upper) {
            pr *= upper;
            u1 = (u1 - 0.0f) / upper;
        } else {
            iy += 1;
            pr *= lower;
            u1 = (u1 - upper) / lower;
        }
        cdx *= 2;
        cdy *= 2;
        if (lev > 0) {
            uint32_t pdx = max(e.map_dim_x >> (lev - 1), 1u), pdy = max(e.map_dim_y >> (lev - 1), 1u);
            off -= (size_t)pdx * pdy;
        }
    }
    ox = ix;
    oy = iy;
    ru0 = u0;
    ru1 = u1;
    prob = pr;
}

// DiscreteDistribution1DTemplate<.., true>::sample (common_renderer_types.h:243-259)
__device__ __forceinline__ uint32_t alias_pick(const rtc8x_alias_entry* table, uint32_t n, float u, float& remapped)
{
    uint32_t idx = (uint32_t)(u * (float)n);
    if (idx > n - 1) idx = n - 1;
    float ur = u * (float)n - (float)idx;
    rtc8x_alias_entry en = table[idx];
    float pf = en.prob_to_pick_first;
    if (ur < pf) {
        remapped = ur / pf;
        return idx;
    }
    remapped = (ur - pf) / (1.0f - pf);
    return en.second_index;
}

// EnvironmentalLight::sample + ImageBasedEnvironmentalLight::sample (common_renderer_types.h:988-1000, 2419-2440)
__device__ __forceinline__ RGB env_sample(const rtc8x_env_light& e, float u0, float u1, V3& dir, float& pdf)
{
    if (e.kind == RTC8X_ENV_CONSTANT) {
        float z = 1.0f - 2.0f * u0;
        float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        float phi = 2.0f * RTC8X_PI_F * u1;
        float sp, cp;
        rtc8x_sincosf(phi, &sp, &cp);
        dir = V3{r * cp, r * sp, z};
        pdf = 1.0f / (4.0f * RTC8X_PI_F);
        return scale(RGB{e.constant_radiance[0], e.constant_radiance[1], e.constant_radiance[2]}, e.power_coeff);
    }
    uint32_t px, py;
    float ru0, ru1, prob;
    if (e.sampler == RTC8X_ENV_SAMPLER_ALIAS) {
        py = alias_pick(e.alias_rows, e.height, u1, ru1);
        px = alias_pick(e.alias_cols + (size_t)py * e.width, e.width, u0, ru0);
        prob = e.alias_weights[(size_t)py * e.width + px] / e.alias_integral;
    } else {
        map_sample(e, u0, u1, px, py, ru0, ru1, prob);
    }
    float u = ((float)px + ru0) / (float)e.width;
    float v = ((float)py + ru1) / (float)e.height;
    float uv_pd = (float)(e.width * e.height) * prob;
    const float two_pi = 2.0f * RTC8X_PI_F;
    float phi = two_pi * u;
    float theta = RTC8X_PI_F * v;
    float st, ct;
    rtc8x_sincosf(theta, &st, &ct);
    pdf = uv_pd / (2.0f * (RTC8X_PI_F * RTC8X_PI_F) * st);
    if (!isfinite(pdf)) {
        pdf = 0.0f;
        return RGB{0.0f, 0.0f, 0.0f};
    }
    RGB radiance = scale(env_texel(e, u, v), e.power_coeff);
    float pos_phi = phi - e.rotation;
    pos_phi = pos_phi - floorf(pos_phi / two_pi) * two_pi;
    float sp, cp;
    rtc8x_sincosf(pos_phi, &sp, &cp);
    dir = V3{-sp * st, ct, cp * st};  // Vector3D::fromPolarYUp (basic_types.h:900-907)
    return radiance;
}

}  // namespace dev
}  // namespace rtc8x

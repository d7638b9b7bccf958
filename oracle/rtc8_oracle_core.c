/*
 * rtc8_oracle_core.c -- CPU oracle (TEST INFRASTRUCTURE ONLY): random number generators, the
 * procedural fBm volume, the NanoVDB-layout tree and its sampler.
 *
 * Used only by the test-suite (as the checker) and by bench.py's cpu_baseline leg.
 */
#include "rtc8_oracle.h"
#include "../include/rtc8x_detmath.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- a1: PCG32
 * common/common_renderer_types.h:111-124: state' = state * 6364136223846793005 + 1; the output is
 * XSH-RR of the old state; getFloat0cTo1o builds a float in [1,2) from the top 23 bits and subtracts 1. */
uint32_t oracle_pcg_next(uint64_t* state)
{
    uint64_t old = *state;
    *state = old * 6364136223846793005ULL + 1ULL;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
}

float oracle_pcg_float(uint64_t* state)
{
    uint32_t bits = (oracle_pcg_next(state) >> 9) | 0x3f800000u;
    float f;
    memcpy(&f, &bits, 4);
    return f - 1.0f;
}

void oracle_pcg_fill(uint64_t state, uint32_t n, uint32_t* out_u32, float* out_f32, uint64_t* out_state)
{
    uint64_t s = state;
    for (uint32_t i = 0; i < n; ++i) {
        if (out_u32) {
            out_u32[i] = oracle_pcg_next(&s);
        } else {
            out_f32[i] = oracle_pcg_float(&s);
        }
    }
    if (out_state) *out_state = s;
}

/* ---------------------------------------------------------------- mt19937_64
 * The C++ standard's std::mt19937_64 (Matsumoto & Nishimura's 64-bit Mersenne Twister), used by
 * renderer/renderer_main.cpp:87-92 to seed one PCG state per pixel in row-major order. */
void oracle_mt19937_64(uint64_t seed, uint32_t n, uint64_t* out)
{
    enum { NN = 312, MM = 156 };
    uint64_t mt[NN];
    const uint64_t upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL;
    mt[0] = seed;
    for (int i = 1; i < NN; ++i)
        mt[i] = 6364136223846793005ULL * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (uint64_t)i;
    int idx = NN;
    for (uint32_t c = 0; c < n; ++c) {
        if (idx >= NN) {
            for (int i = 0; i < NN; ++i) {
                uint64_t y = (mt[i] & upper) | (mt[(i + 1) % NN] & lower);
                uint64_t v = mt[(i + MM) % NN] ^ (y >> 1);
                if (y & 1ULL) v ^= 0xB5026F5AA96619E9ULL;
                mt[i] = v;
            }
            idx = 0;
        }
        uint64_t y = mt[idx++];
        y ^= (y >> 29) & 0x5555555555555555ULL;
        y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
        y ^= (y << 37) & 0xFFF7EEE000000000ULL;
        y ^= (y >> 43);
        out[c] = y;
    }
}

void oracle_seed_rng(uint64_t* states, uint32_t width, uint32_t height, uint64_t seed)
{
    oracle_mt19937_64(seed, width * height, states);
}

/* ---------------------------------------------------------------- a2: LCG
 * renderer/renderer_shared.h:92-109 (a = 1103515245, c = 12345, m = 2^31; the product wraps in
 * 32 bits before the modulo); seeded per slot as renderer_main.cpp:220-229. */
uint32_t oracle_lcg_next(uint32_t* state)
{
    *state = ((*state * 1103515245u) + 12345u) % (1u << 31);
    return *state;
}

void oracle_lcg_seeds(uint32_t seed, uint32_t n, uint32_t* out)
{
    uint32_t s = seed;
    for (uint32_t i = 0; i < n; ++i) {
        oracle_lcg_next(&s);
        out[i] = s;
    }
}

/* ---------------------------------------------------------------- fBm volume
 * Procedural density (the benchmark input; the reference loads a .nvdb file instead,
 * renderer/renderer_scene.cpp:2467-2470).  Value noise on an integer lattice with smoothstep
 * interpolation, summed over octaves.  Integer hash: "lowbias32" (x ^= x>>16; x *= 0x7feb352d;
 * x ^= x>>15; x *= 0x846ca68b; x ^= x>>16). */
static uint32_t fbm_hash(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

static float fbm_lattice(uint32_t s, int32_t ix, int32_t iy, int32_t iz)
{
    uint32_t h = fbm_hash(s + (uint32_t)ix);
    h = fbm_hash(h + (uint32_t)iy);
    h = fbm_hash(h + (uint32_t)iz);
    return (float)(h >> 8) * 0x1p-23f - 1.0f;
}

static float fbm_lerp(float a, float b, float w) { return a + w * (b - a); }

static float fbm_noise(float px, float py, float pz, uint32_t s)
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

/* Unnormalised field g at voxel (i,j,k): fbm for dense volumes, an fbm-perturbed sphere
 * (shape + 0.6 fbm, shape = max(0, 1 - 2 r)) for sparse clouds. */
float oracle_fbm_value(const rtc8x_fbm_desc* d, int32_t i, int32_t j, int32_t k)
{
    float inv_n = 1.0f / (float)d->resolution;
    float qx = ((float)i + 0.5f) * inv_n;
    float qy = ((float)j + 0.5f) * inv_n;
    float qz = ((float)k + 0.5f) * inv_n;
    float sum = 0.0f, amp = 1.0f, freq = d->base_frequency;
    uint32_t s_lo = (uint32_t)d->seed, s_hi = (uint32_t)(d->seed >> 32);
    for (uint32_t o = 0; o < d->octaves; ++o) {
        uint32_t s = fbm_hash(s_lo ^ fbm_hash(s_hi + o));
        sum = sum + amp * fbm_noise(qx * freq, qy * freq, qz * freq, s);
        amp = amp * d->gain;
        freq = freq * d->lacunarity;
    }
    if (!d->sparse) return sum;
    float dx = qx - 0.5f, dy = qy - 0.5f, dz = qz - 0.5f;
    float r = sqrtf(dx * dx + dy * dy + dz * dz);
    float shape = 1.0f - 2.0f * r;
    if (shape < 0.0f) shape = 0.0f;
    return shape + 0.6f * sum;
}

static int cmp_float_desc(const void* a, const void* b)
{
    float fa = *(const float*)a, fb = *(const float*)b;
    return (fa < fb) - (fa > fb);
}

/* NanoVDB offsets: InternalNode::CoordToOffset (NanoVDB.h:3217-3227) with LOG2DIM 5 / child TOTAL 7
 * for the upper node and 4 / 3 for the lower node; LeafNode::CoordToOffset (NanoVDB.h:3821-3828). */
static uint32_t upper_offset(int32_t i, int32_t j, int32_t k)
{
    return ((((uint32_t)i & 4095u) >> 7) << 10) | ((((uint32_t)j & 4095u) >> 7) << 5) |
           (((uint32_t)k & 4095u) >> 7);
}
static uint32_t lower_offset(int32_t i, int32_t j, int32_t k)
{
    return ((((uint32_t)i & 127u) >> 3) << 8) | ((((uint32_t)j & 127u) >> 3) << 4) |
           (((uint32_t)k & 127u) >> 3);
}
static uint32_t leaf_offset(int32_t i, int32_t j, int32_t k)
{
    return (((uint32_t)i & 7u) << 6) | (((uint32_t)j & 7u) << 3) | ((uint32_t)k & 7u);
}
/* RootData::CoordToKey with 64-bit keys (NanoVDB.h:2699-2706), child TOTAL = 12. */
static uint64_t root_key(int32_t i, int32_t j, int32_t k)
{
    return ((uint64_t)((uint32_t)k >> 12)) | ((uint64_t)((uint32_t)j >> 12) << 21) |
           ((uint64_t)((uint32_t)i >> 12) << 42);
}

int oracle_grid_build_fbm(const rtc8x_fbm_desc* d, rtc8x_grid* g)
{
    memset(g, 0, sizeof(*g));
    uint32_t n = d->resolution;
    if (n == 0 || (n & 7u) || n > 4096) return -1;
    uint32_t nl = n / 8, num_slots = nl * nl * nl;
    float* leaf_max = (float*)malloc(sizeof(float) * num_slots);
    /* pass 1: per-leaf maximum of g */
    for (uint32_t lx = 0; lx < nl; ++lx)
        for (uint32_t ly = 0; ly < nl; ++ly)
            for (uint32_t lz = 0; lz < nl; ++lz) {
                float m = -FLT_MAX;
                for (int32_t v = 0; v < 512; ++v) {
                    int32_t i = (int32_t)lx * 8 + (v >> 6), j = (int32_t)ly * 8 + ((v >> 3) & 7),
                            k = (int32_t)lz * 8 + (v & 7);
                    float val = oracle_fbm_value(d, i, j, k);
                    if (val > m) m = val;
                }
                leaf_max[(lx * nl + ly) * nl + lz] = m;
            }
    float gmax = -FLT_MAX;
    for (uint32_t s = 0; s < num_slots; ++s) if (leaf_max[s] > gmax) gmax = leaf_max[s];
    float thr = 0.0f;
    int keep_all = !d->sparse;
    if (d->sparse) {
        uint32_t n_keep = (uint32_t)(d->active_leaf_fraction * (float)num_slots);
        if (n_keep >= num_slots) {
            keep_all = 1;
            thr = -FLT_MAX;
        } else {
            float* sorted = (float*)malloc(sizeof(float) * num_slots);
            memcpy(sorted, leaf_max, sizeof(float) * num_slots);
            qsort(sorted, num_slots, sizeof(float), cmp_float_desc);
            thr = sorted[n_keep];
            free(sorted);
        }
    }
    float range = d->sparse ? gmax - thr : gmax;
    if (!(range > 0.0f)) { free(leaf_max); return -2; }
    float inv_range = 1.0f / range;
    float sub = d->sparse ? thr : 0.0f;
    if (d->sparse && keep_all) sub = 0.0f;

    /* leaves in NanoVDB traversal order: lower node by upper offset, leaf by lower offset */
    uint32_t* lower_of_uoff = (uint32_t*)malloc(sizeof(uint32_t) * RTC8X_UPPER_TABLE_SIZE);
    for (uint32_t u = 0; u < RTC8X_UPPER_TABLE_SIZE; ++u) lower_of_uoff[u] = 0xFFFFFFFFu;
    uint32_t num_lower = 0, num_leaves = 0;
    /* iterate upper offsets in increasing order: uoff = (lx>>4)<<10 | (ly>>4)<<5 | (lz>>4) */
    uint32_t nb = (nl + 15) / 16;
    for (uint32_t bx = 0; bx < nb; ++bx)
        for (uint32_t by = 0; by < nb; ++by)
            for (uint32_t bz = 0; bz < nb; ++bz) {
                int any = 0;
                for (uint32_t o = 0; o < 4096 && !any; ++o) {
                    uint32_t lx = bx * 16 + (o >> 8), ly = by * 16 + ((o >> 4) & 15), lz = bz * 16 + (o & 15);
                    if (lx >= nl || ly >= nl || lz >= nl) continue;
                    float m = leaf_max[(lx * nl + ly) * nl + lz];
                    if (keep_all || m > thr) any = 1;
                }
                if (any) lower_of_uoff[(bx << 10) | (by << 5) | bz] = num_lower++;
            }
    uint32_t* upper = (uint32_t*)malloc(sizeof(uint32_t) * RTC8X_UPPER_TABLE_SIZE);
    uint32_t* lower = (uint32_t*)malloc(sizeof(uint32_t) * RTC8X_LOWER_TABLE_SIZE * (num_lower ? num_lower : 1));
    for (uint32_t u = 0; u < RTC8X_UPPER_TABLE_SIZE; ++u)
        upper[u] = lower_of_uoff[u] == 0xFFFFFFFFu ? RTC8X_TILE_BIT : lower_of_uoff[u];
    for (uint32_t e = 0; e < RTC8X_LOWER_TABLE_SIZE * num_lower; ++e) lower[e] = RTC8X_TILE_BIT;
    /* count leaves */
    for (uint32_t u = 0; u < RTC8X_UPPER_TABLE_SIZE; ++u) {
        uint32_t li = lower_of_uoff[u];
        if (li == 0xFFFFFFFFu) continue;
        uint32_t bx = u >> 10, by = (u >> 5) & 31, bz = u & 31;
        for (uint32_t o = 0; o < 4096; ++o) {
            uint32_t lx = bx * 16 + (o >> 8), ly = by * 16 + ((o >> 4) & 15), lz = bz * 16 + (o & 15);
            if (lx >= nl || ly >= nl || lz >= nl) continue;
            float m = leaf_max[(lx * nl + ly) * nl + lz];
            if (keep_all || m > thr) lower[li * 4096 + o] = num_leaves++;
        }
    }
    float* values = (float*)malloc(sizeof(float) * RTC8X_LEAF_SIZE * (num_leaves ? num_leaves : 1));
    int32_t* origins = (int32_t*)malloc(sizeof(int32_t) * 3 * (num_leaves ? num_leaves : 1));
    int32_t bbmin[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, bbmax[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    float vmax = 0.0f;
    for (uint32_t u = 0; u < RTC8X_UPPER_TABLE_SIZE; ++u) {
        uint32_t li = lower_of_uoff[u];
        if (li == 0xFFFFFFFFu) continue;
        uint32_t bx = u >> 10, by = (u >> 5) & 31, bz = u & 31;
        for (uint32_t o = 0; o < 4096; ++o) {
            uint32_t e = lower[li * 4096 + o];
            if (e & RTC8X_TILE_BIT) continue;
            int32_t ox = (int32_t)(bx * 16 + (o >> 8)) * 8, oy = (int32_t)(by * 16 + ((o >> 4) & 15)) * 8,
                    oz = (int32_t)(bz * 16 + (o & 15)) * 8;
            origins[3 * e + 0] = ox; origins[3 * e + 1] = oy; origins[3 * e + 2] = oz;
            if (ox < bbmin[0]) bbmin[0] = ox;
            if (oy < bbmin[1]) bbmin[1] = oy;
            if (oz < bbmin[2]) bbmin[2] = oz;
            if (ox + 7 > bbmax[0]) bbmax[0] = ox + 7;
            if (oy + 7 > bbmax[1]) bbmax[1] = oy + 7;
            if (oz + 7 > bbmax[2]) bbmax[2] = oz + 7;
            for (int32_t v = 0; v < 512; ++v) {
                float gv = oracle_fbm_value(d, ox + (v >> 6), oy + ((v >> 3) & 7), oz + (v & 7));
                float dv = gv - sub;
                if (dv < 0.0f) dv = 0.0f;
                dv = dv * inv_range;
                dv = dv * d->contrast;
                if (dv > 1.0f) dv = 1.0f;
                values[(size_t)e * 512 + (size_t)v] = dv;
                if (dv > vmax) vmax = dv;
            }
        }
    }
    free(leaf_max);
    free(lower_of_uoff);

    rtc8x_root_tile* tiles = (rtc8x_root_tile*)malloc(sizeof(rtc8x_root_tile));
    tiles[0].key = root_key(0, 0, 0);
    tiles[0].entry = 0;
    tiles[0].state = 0;
    float* tile_values = (float*)malloc(sizeof(float));
    tile_values[0] = 0.0f;

    for (int a = 0; a < 9; ++a) g->inv_mat[a] = 0.0f;
    g->inv_mat[0] = g->inv_mat[4] = g->inv_mat[8] = (float)n;
    for (int a = 0; a < 3; ++a) {
        g->vec[a] = 0.0f;
        g->index_bbox_min[a] = num_leaves ? bbmin[a] : 0;
        g->index_bbox_max[a] = num_leaves ? bbmax[a] : -1;
        g->world_bbox_min[a] = (float)g->index_bbox_min[a] / (float)n;
        g->world_bbox_max[a] = (float)(g->index_bbox_max[a] + 1) / (float)n;
    }
    g->max_value = vmax;
    g->background = 0.0f;
    g->num_root_tiles = 1;
    g->num_upper = 1;
    g->num_lower = num_lower;
    g->num_leaves = num_leaves;
    g->root_tiles = tiles;
    g->upper_tables = upper;
    g->lower_tables = lower;
    g->leaf_values = values;
    g->tile_values = tile_values;
    g->leaf_origins = origins;
    oracle_compute_cell_majorants(g, 0);
    return 0;
}

void oracle_grid_free(rtc8x_grid* g)
{
    free((void*)g->root_tiles);
    free((void*)g->upper_tables);
    free((void*)g->lower_tables);
    free((void*)g->leaf_values);
    free((void*)g->tile_values);
    free((void*)g->leaf_origins);
    free((void*)g->cell_majorants);
    free((void*)g->region_majorants);
    memset(g, 0, sizeof(*g));
}

/* ReadAccessor::getValue (NanoVDB.h:4432-4560) without the node cache: root tile linear search
 * (RootNode::findTile, NanoVDB.h:2921-2945) -> upper table -> lower table -> leaf value. */
float oracle_grid_get_value(const rtc8x_grid* g, int32_t i, int32_t j, int32_t k, uint32_t offsets[3])
{
    uint64_t key = root_key(i, j, k);
    uint32_t entry = RTC8X_TILE_BIT;
    int found = 0;
    for (uint32_t t = 0; t < g->num_root_tiles; ++t)
        if (g->root_tiles[t].key == key) { entry = g->root_tiles[t].entry; found = 1; break; }
    uint32_t uo = upper_offset(i, j, k), lo = lower_offset(i, j, k), vo = leaf_offset(i, j, k);
    if (offsets) { offsets[0] = uo; offsets[1] = lo; offsets[2] = vo; }
    if (!found) return g->background;
    if (entry & RTC8X_TILE_BIT) return g->tile_values[entry & ~RTC8X_TILE_BIT];
    uint32_t e1 = g->upper_tables[(size_t)entry * RTC8X_UPPER_TABLE_SIZE + uo];
    if (e1 & RTC8X_TILE_BIT) return g->tile_values[e1 & ~RTC8X_TILE_BIT];
    uint32_t e2 = g->lower_tables[(size_t)e1 * RTC8X_LOWER_TABLE_SIZE + lo];
    if (e2 & RTC8X_TILE_BIT) return g->tile_values[e2 & ~RTC8X_TILE_BIT];
    return g->leaf_values[(size_t)e2 * RTC8X_LEAF_SIZE + vo];
}

/* The tree walk of getValue stopped at the leaf: the leaf index holding voxel (i, j, k), or UINT32_MAX with
 * *tile = the tile (or background) value that covers it. */
static uint32_t leaf_of(const rtc8x_grid* g, int32_t i, int32_t j, int32_t k, float* tile)
{
    uint64_t key = root_key(i, j, k);
    for (uint32_t t = 0; t < g->num_root_tiles; ++t) {
        if (g->root_tiles[t].key != key) continue;
        uint32_t e = g->root_tiles[t].entry;
        if (!(e & RTC8X_TILE_BIT)) e = g->upper_tables[(size_t)e * RTC8X_UPPER_TABLE_SIZE + upper_offset(i, j, k)];
        if (!(e & RTC8X_TILE_BIT)) e = g->lower_tables[(size_t)e * RTC8X_LOWER_TABLE_SIZE + lower_offset(i, j, k)];
        if (!(e & RTC8X_TILE_BIT)) return e;
        *tile = g->tile_values[e & ~RTC8X_TILE_BIT];
        return UINT32_MAX;
    }
    *tile = g->background;
    return UINT32_MAX;
}

/* Map::applyInverseMapF (NanoVDB.h:2029-2032) with matMult (NanoVDB.h:1409-1414). */
static void world_to_index_f(const rtc8x_grid* g, const float p[3], float out[3])
{
    float x = p[0] - g->vec[0], y = p[1] - g->vec[1], z = p[2] - g->vec[2];
    const float* m = g->inv_mat;
    out[0] = fmaf(x, m[0], fmaf(y, m[1], z * m[2]));
    out[1] = fmaf(x, m[3], fmaf(y, m[4], z * m[5]));
    out[2] = fmaf(x, m[6], fmaf(y, m[7], z * m[8]));
}

/* SampleFromVoxels<Acc,1,false>::operator() (util/SampleFromVoxels.h:419-460) with Floor (:58-65). */
static float sample_trilinear_f32(const rtc8x_grid* g, const float idx[3], int32_t ijk_out[3])
{
    float fx = floorf(idx[0]), fy = floorf(idx[1]), fz = floorf(idx[2]);
    float x = idx[0] - fx, y = idx[1] - fy, z = idx[2] - fz;
    int32_t i = (int32_t)fx, j = (int32_t)fy, k = (int32_t)fz;
    if (ijk_out) { ijk_out[0] = i; ijk_out[1] = j; ijk_out[2] = k; }
    float v000 = oracle_grid_get_value(g, i, j, k, NULL);
    float v001 = oracle_grid_get_value(g, i, j, k + 1, NULL);
    float vy = v000 + z * (v001 - v000);
    float v011 = oracle_grid_get_value(g, i, j + 1, k + 1, NULL);
    float v010 = oracle_grid_get_value(g, i, j + 1, k, NULL);
    float vy1 = v010 + z * (v011 - v010);
    float vx = vy + y * (vy1 - vy);
    float v110 = oracle_grid_get_value(g, i + 1, j + 1, k, NULL);
    float v111 = oracle_grid_get_value(g, i + 1, j + 1, k + 1, NULL);
    vy1 = v110 + z * (v111 - v110);
    float v101 = oracle_grid_get_value(g, i + 1, j, k + 1, NULL);
    float v100 = oracle_grid_get_value(g, i + 1, j, k, NULL);
    vy = v100 + z * (v101 - v100);
    float vx1 = vy + y * (vy1 - vy);
    return vx + x * (vx1 - vx);
}

float oracle_sample_world_f32(const rtc8x_grid* g, const float p[3])
{
    float idx[3];
    world_to_index_f(g, p, idx);
    return sample_trilinear_f32(g, idx, NULL);
}

void oracle_grid_probe(const rtc8x_grid* g, uint32_t n, const float* world_xyz,
                       int32_t* out_ijk, uint32_t* out_offsets, float* out_value)
{
    for (uint32_t q = 0; q < n; ++q) {
        float idx[3];
        world_to_index_f(g, &world_xyz[3 * q], idx);
        int32_t ijk[3];
        float v = sample_trilinear_f32(g, idx, ijk);
        uint32_t offs[3];
        oracle_grid_get_value(g, ijk[0], ijk[1], ijk[2], offs);
        for (int a = 0; a < 3; ++a) {
            out_ijk[3 * q + a] = ijk[a];
            out_offsets[3 * q + a] = offs[a];
        }
        out_value[q] = v;
    }
}

/* Per-cell majorants for local-majorant tracking (the grid's definition in rtc8x.h): cell c covers index
 * [8c, 8c+8)^3 from cell_origin, over the cells of the index bbox.  A point of the cell reads its
 * trilinear taps from [8c, 8c+8]; a sample position computed in floating point can land a rounding error
 * outside the cell's t-range, so the bound is the largest value getValue returns (leaves, tiles,
 * background; at least 0) over index [8c-1, 8c+9]^3.  Found here block by block: per axis that range is
 * the last voxel of block c-1, all of block c and the first two of block c+1; a block under a tile
 * contributes the tile's value, a leaf the voxels of its part of the range.  No tables when the index
 * bbox spans more than 2^26 cells. */
void oracle_compute_cell_majorants(rtc8x_grid* g, uint32_t region_cells)
{
    free((void*)g->cell_majorants);
    g->cell_majorants = NULL;
    free((void*)g->region_majorants);
    g->region_majorants = NULL;
    g->region_cells = 0;
    g->region_unit = 0.0f;
    for (int a = 0; a < 3; ++a) { g->cell_origin[a] = 0; g->cell_dims[a] = 0; g->region_dims[a] = 0; }
    if (g->num_leaves == 0) return;
    int32_t lo[3], dims[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = (int32_t)floorf((float)g->index_bbox_min[a] / 8.0f);
        int32_t hi = (int32_t)floorf((float)g->index_bbox_max[a] / 8.0f);
        dims[a] = hi - lo[a] + 1;
        g->cell_origin[a] = lo[a] * 8;
        g->cell_dims[a] = dims[a];
    }
    if ((size_t)dims[0] * dims[1] * dims[2] > ((size_t)1 << 26)) {  /* too large an index bbox: no tables */
        for (int a = 0; a < 3; ++a) { g->cell_origin[a] = 0; g->cell_dims[a] = 0; }
        return;
    }
    float* cm = (float*)malloc(sizeof(float) * (size_t)dims[0] * dims[1] * dims[2]);
    for (int32_t x = 0; x < dims[0]; ++x)
        for (int32_t y = 0; y < dims[1]; ++y)
            for (int32_t z = 0; z < dims[2]; ++z) {
                const int32_t c[3] = {g->cell_origin[0] + 8 * x, g->cell_origin[1] + 8 * y, g->cell_origin[2] + 8 * z};
                /* per axis and block b = -1, 0, +1: the voxel range [first, last] inside the block */
                static const int32_t first[3] = {7, 0, 0}, last[3] = {7, 7, 1};
                float m = 0.0f;
                for (int bx = 0; bx < 3; ++bx)
                    for (int by = 0; by < 3; ++by)
                        for (int bz = 0; bz < 3; ++bz) {
                            int32_t o[3] = {c[0] + 8 * (bx - 1), c[1] + 8 * (by - 1), c[2] + 8 * (bz - 1)};
                            float tile = 0.0f;
                            uint32_t leaf = leaf_of(g, o[0], o[1], o[2], &tile);
                            if (leaf == UINT32_MAX) {
                                if (tile > m) m = tile;
                                continue;
                            }
                            const float* lv = g->leaf_values + (size_t)leaf * RTC8X_LEAF_SIZE;
                            for (int32_t i = first[bx]; i <= last[bx]; ++i)
                                for (int32_t j = first[by]; j <= last[by]; ++j)
                                    for (int32_t k = first[bz]; k <= last[bz]; ++k) {
                                        float v = lv[(i << 6) | (j << 3) | k];
                                        if (v > m) m = v;
                                    }
                        }
                cm[((size_t)x * dims[1] + y) * dims[2] + z] = m;
            }
    g->cell_majorants = cm;
    oracle_compute_region_majorants(g, region_cells);
}

/* Region table over groups of region_cells^3 cells (the grid's definition in rtc8x.h): region_cells as
 * given, or for 0 RTC8X_DEFAULT_REGION_CELLS doubled until there are at most RTC8X_MAX_REGIONS regions;
 * region_unit = the largest cell majorant / 127, raised float by float until 127 * region_unit covers it;
 * per region with density the least q in 1..127 with q * region_unit >= its largest cell majorant; per
 * empty region RTC8X_REGION_EMPTY + (its Chebyshev distance to the nearest region with density) - 1, at
 * most 127, found here by a brute-force search over the regions with density. */
void oracle_compute_region_majorants(rtc8x_grid* g, uint32_t region_cells)
{
    free((void*)g->region_majorants);
    g->region_majorants = NULL;
    g->region_cells = 0;
    g->region_unit = 0.0f;
    for (int a = 0; a < 3; ++a) g->region_dims[a] = 0;
    if (!g->cell_majorants) return;
    const int32_t* cd = g->cell_dims;
    uint32_t rc = region_cells ? region_cells : RTC8X_DEFAULT_REGION_CELLS;
    int32_t rd[3];
    for (;;) {
        size_t n = 1;
        for (int a = 0; a < 3; ++a) {
            rd[a] = (cd[a] + (int32_t)rc - 1) / (int32_t)rc;
            n *= (size_t)rd[a];
        }
        if (n <= RTC8X_MAX_REGIONS || region_cells) break;
        rc *= 2;
    }
    size_t nr = (size_t)rd[0] * rd[1] * rd[2];
    float* rmax = (float*)calloc(nr, sizeof(float));
    float top = 0.0f;
    for (int32_t x = 0; x < cd[0]; ++x)
        for (int32_t y = 0; y < cd[1]; ++y)
            for (int32_t z = 0; z < cd[2]; ++z) {
                float m = g->cell_majorants[((size_t)x * cd[1] + y) * cd[2] + z];
                size_t r = ((size_t)(x / (int32_t)rc) * rd[1] + y / (int32_t)rc) * rd[2] + z / (int32_t)rc;
                if (m > rmax[r]) rmax[r] = m;
                if (m > top) top = m;
            }
    float unit = top / 127.0f;
    while (127.0f * unit < top) unit = nextafterf(unit, INFINITY);
    /* the regions with density, for the distance search */
    size_t nd = 0;
    int32_t* dense = (int32_t*)malloc(sizeof(int32_t) * 3 * (nr ? nr : 1));
    for (int32_t x = 0; x < rd[0]; ++x)
        for (int32_t y = 0; y < rd[1]; ++y)
            for (int32_t z = 0; z < rd[2]; ++z)
                if (rmax[((size_t)x * rd[1] + y) * rd[2] + z] > 0.0f) {
                    dense[3 * nd] = x;
                    dense[3 * nd + 1] = y;
                    dense[3 * nd + 2] = z;
                    ++nd;
                }
    uint8_t* q = (uint8_t*)calloc((nr + 15) & ~(size_t)15, 1);
    for (int32_t x = 0; x < rd[0]; ++x)
        for (int32_t y = 0; y < rd[1]; ++y)
            for (int32_t z = 0; z < rd[2]; ++z) {
                size_t r = ((size_t)x * rd[1] + y) * rd[2] + z;
                uint32_t v;
                if (rmax[r] > 0.0f) {
                    v = 1;
                    while (v < 127 && (float)v * unit < rmax[r]) ++v;
                } else {
                    int32_t best = 128;  /* no region with density within 127 layers */
                    for (size_t i = 0; i < nd; ++i) {
                        int32_t dx = abs(dense[3 * i] - x), dy = abs(dense[3 * i + 1] - y), dz = abs(dense[3 * i + 2] - z);
                        int32_t c = dx > dy ? dx : dy;
                        c = c > dz ? c : dz;
                        if (c < best) best = c;
                    }
                    v = RTC8X_REGION_EMPTY + (uint32_t)(best - 1 < 127 ? best - 1 : 127);
                }
                q[r] = (uint8_t)v;
            }
    free(dense);
    free(rmax);
    for (int a = 0; a < 3; ++a) g->region_dims[a] = rd[a];
    g->region_cells = rc;
    g->region_unit = unit;
    g->region_majorants = q;
}

/* nanovdb::Ray::intersects(const BBox<Vec3f>&, t0, t1) and clip (util/Ray.h:63-73, 427-455, 512-520):
 * inverse direction 1/d, per-axis slab, [0, tmax] time span. */
int oracle_ray_clip(const float eye[3], const float dir[3], float tmax, const float bmin[3],
                    const float bmax[3], float* t0_out, float* t1_out)
{
    float t0 = 0.0f, t1 = tmax;
    for (int i = 0; i < 3; ++i) {
        float inv = 1.0f / dir[i];
        float a = bmin[i], b = bmax[i];
        if (a >= b) return 0;
        a = (a - eye[i]) * inv;
        b = (b - eye[i]) * inv;
        if (a > b) { float t = a; a = b; b = t; }
        if (a > t0) t0 = a;
        if (b < t1) t1 = b;
        if (t0 > t1) return 0;
    }
    *t0_out = t0;
    *t1_out = t1;
    return 1;
}

float oracle_logf(float x) { return rtc8x_logf(x); }
void oracle_sincosf(float x, float* s, float* c) { rtc8x_sincosf(x, s, c); }
float oracle_atan2f(float y, float x) { return rtc8x_atan2f(y, x); }
float oracle_acosf(float x) { return rtc8x_acosf(x); }

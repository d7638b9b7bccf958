// rtc8x_nvdb.hip -- NanoVDB grid buffers and .nvdb files (host code).
//
// Reads what the reference loads with nanovdb::io::readGrid (renderer/renderer_scene.cpp:2467-2470)
// into this library's grid tables, and writes those tables back as NanoVDB 32.x FloatGrids, so volumes
// move between the two programs unchanged.  Layouts follow NanoVDB 32.3 byte for byte (offsets below);
// the code is this library's own.
//
//   GridData  672 B  magic u64 @0, checksum u64 @8, version u32 @16, flags u32 @20, grid index/count
//                    @24/@28, grid size u64 @32, name char[256] @40, Map @296 (MatF[9] @+0, InvMatF[9]
//                    @+36, VecF[3] @+72, TaperF @+84, MatD[9] @+88, InvMatD[9] @+160, VecD[3] @+232,
//                    TaperD @+256), world bbox f64[6] @560, voxel size f64[3] @608, class u32 @632,
//                    type u32 @636, blind offset i64 @640, blind count u32 @648   (NanoVDB.h:2185-2205)
//   TreeData   64 B  node offsets u64[4] (leaf, lower, upper, root; from the tree) @0, node counts
//                    u32[3] @32, active tile counts u32[3] @44, active voxel count u64 @56 (:2501-2508)
//   RootData   64 B  bbox i32[6] @0, table size u32 @24, background/min/max/avg/stddev f32 @28..@44,
//                    then tiles of 32 B: key u64 @0, child i64 @8 (from the root), state u32 @16,
//                    value f32 @20                                                       (:2687-2760)
//   Internal         bbox i32[6] @0, flags u64 @24, value mask @32, child mask, min/max/avg/stddev f32,
//                    table of 8-byte entries (child i64 from the node, or value f32) aligned to 32:
//                    upper (32^3): masks 4096 B, stats @8224, table @8256, 270,400 B;
//                    lower (16^3): masks 512 B, stats @1056, table @1088, 33,856 B     (:3043-3072)
//   LeafData         bbox min i32[3] @0, bbox extent u8[3] @12, flags u8 @15, value mask 64 B @16,
//                    min/max/avg/stddev f32 @80..@92, values f32[512] @96; 2,144 B       (:3355-3373)
//   .nvdb segment    header 16 B (magic u64, version u32, grid count u16, codec u16), then per grid a
//                    176-byte MetaData record and its NUL-terminated name, then the grids' bytes
//                    (util/IO.h:104-160); ZIP data is a u64 byte count and a zlib stream (:254-344).
#include <hip/hip_runtime.h>
#include <zlib.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "rtc8x_common.hpp"

namespace rtc8x {
namespace {

constexpr uint64_t kMagic = 0x304244566f6e614eull;  // "NanoVDB0", little-endian
constexpr uint32_t kMajor = 32, kMinor = 3, kPatch = 3;
constexpr uint32_t kVersion = (kMajor << 21) | (kMinor << 10) | kPatch;
constexpr uint32_t kGridTypeFloat = 1, kGridClassFogVolume = 2;
constexpr uint32_t kFlagHasBBox = 1u << 1, kFlagHasMinMax = 1u << 2, kFlagHasAverage = 1u << 3,
                   kFlagHasStdDeviation = 1u << 4, kFlagIsBreadthFirst = 1u << 5;

constexpr size_t kGridBytes = 672, kTreeBytes = 64, kRootBytes = 64, kTileBytes = 32;
constexpr size_t kUpperBytes = 270400, kLowerBytes = 33856, kLeafBytes = 2144;
constexpr size_t kUpperStats = 8224, kUpperTable = 8256, kUpperMask = 4096;
constexpr size_t kLowerStats = 1056, kLowerTable = 1088, kLowerMask = 512;
constexpr size_t kLeafStats = 80, kLeafValues = 96;
constexpr size_t kHeaderBytes = 16, kMetaBytes = 176;

template <typename T>
T rd(const uint8_t* p)
{
    T v;
    std::memcpy(&v, p, sizeof(T));
    return v;
}
template <typename T>
void wr(uint8_t* p, T v)
{
    std::memcpy(p, &v, sizeof(T));
}
uint32_t fbits(float v)
{
    uint32_t b;
    std::memcpy(&b, &v, 4);
    return b;
}

const char* grid_type_name(uint32_t t)
{
    static const char* names[] = {"unknown", "float", "double", "int16", "int32", "int64", "Vec3f", "Vec3d",
                                  "mask", "half", "uint32", "bool", "RGBA8", "Fp4", "Fp8", "Fp16", "FpN",
                                  "Vec4f", "Vec4d"};
    return t < sizeof(names) / sizeof(names[0]) ? names[t] : "invalid";
}

// Owned host grids, keyed by their root-tile storage (never empty for a decoded grid).
std::mutex g_owned_mutex;
std::unordered_map<const void*, std::unique_ptr<OwnedHostGrid>>& owned()
{
    static std::unordered_map<const void*, std::unique_ptr<OwnedHostGrid>> m;
    return m;
}

// ---------------------------------------------------------------------------------------------------
// Decoding
// ---------------------------------------------------------------------------------------------------

// Root-key cube origin (RootData::KeyToCoord, NanoVDB.h:2707-2713).
void key_origin(uint64_t key, int64_t o[3])
{
    const uint64_t m = (1u << 21) - 1;
    o[0] = (int32_t)(uint32_t)(((key >> 42) & m) << 12);
    o[1] = (int32_t)(uint32_t)(((key >> 21) & m) << 12);
    o[2] = (int32_t)(uint32_t)((key & m) << 12);
}

struct Decoder {
    const uint8_t* b;
    uint64_t size;
    OwnedHostGrid t;
    std::unordered_map<uint32_t, uint32_t> tile_index;
    std::unordered_map<uint64_t, uint32_t> upper_at, lower_at, leaf_at;
    std::vector<uint64_t> upper_off, lower_off, leaf_off;
    std::vector<int64_t> upper_org, lower_org, leaf_org;  // 3 per node
    uint64_t nodes_begin = 0;

    uint32_t tile_entry(float v)
    {
        auto it = tile_index.emplace(fbits(v), (uint32_t)t.tiles.size());
        if (it.second) t.tiles.push_back(v);
        return RTC8X_TILE_BIT | it.first->second;
    }

    // A child link: the node must lie inside the buffer after the root, aligned to 32 bytes, and be
    // referenced once.
    rtc8x_status link(uint64_t parent, int64_t rel, size_t bytes, std::unordered_map<uint64_t, uint32_t>& at,
                      std::vector<uint64_t>& offs, std::vector<int64_t>& orgs, const int64_t org[3],
                      const char* what, uint32_t* index)
    {
        const int64_t off = (int64_t)parent + rel;
        if (off < (int64_t)nodes_begin || size < bytes || (uint64_t)off > size - bytes || (off & 31) != 0)
            return fail(RTC8X_ERR_INVALID_ARGUMENT, "nanovdb: %s link %lld from offset %llu is outside the buffer or "
                        "misaligned", what, (long long)rel, (unsigned long long)parent);
        auto it = at.emplace((uint64_t)off, (uint32_t)offs.size());
        if (!it.second)
            return fail(RTC8X_ERR_INVALID_ARGUMENT, "nanovdb: %s at offset %lld is linked twice", what, (long long)off);
        if (offs.size() >= RTC8X_TILE_BIT)
            return fail(RTC8X_ERR_NOT_SUPPORTED, "nanovdb: more than 2^31 %s nodes", what);
        offs.push_back((uint64_t)off);
        orgs.insert(orgs.end(), org, org + 3);
        *index = it.first->second;
        return RTC8X_OK;
    }

    rtc8x_status run(rtc8x_grid& g)
    {
        if (size < kGridBytes + kTreeBytes + kRootBytes)
            return fail(RTC8X_ERR_INVALID_ARGUMENT, "nanovdb: buffer of %llu bytes is too small for a grid",
                        (unsigned long long)size);
        if (rd<uint64_t>(b) != kMagic) return fail(RTC8X_ERR_INVALID_ARGUMENT, "nanovdb: bad grid magic number");
        const uint32_t version = rd<uint32_t>(b + 16);
        if ((version >> 21) != kMajor)
            return fail(RTC8X_ERR_NOT_SUPPORTED, "nanovdb: grid version %u.%u.%u, expected major version %u",
                        version >> 21, (version >> 10) & 2047u, version & 1023u, kMajor);
        const uint32_t type = rd<uint32_t>(b + 636);
        if (type != kGridTypeFloat)
            return fail(RTC8X_ERR_NOT_SUPPORTED, "nanovdb: grid type %u (%s) is not supported; the volume must be a "
                        "FloatGrid", type, grid_type_name(type));
        const uint64_t grid_size = rd<uint64_t>(b + 32);
        if (grid_size > size || grid_size < kGridBytes + kTreeBytes + kRootBytes)
            return fail(RTC8X_ERR_INVALID_ARGUMENT, "nanovdb: grid size %llu does not fit the %llu-byte buffer",
                        (unsigned long long)grid_size, (unsigned long long)size);
        size = grid_size;
        const uint8_t* tree = b + kGridBytes;
        const uint64_t root_rel = rd<uint64_t>(tree + 24);
        const uint64_t root = kGridBytes + root_rel;
        if (root_rel < kTreeBytes || root > size - kRootBytes || (root & 31) != 0)
            return fail(RTC8X_ERR_INVALID_ARGUMENT, "nanovdb: root offset %llu is invalid", (unsigned long long)root_rel);
        const uint32_t table_size = rd<uint32_t>(b + root + 24);
        if (table_size == 0) return fail(RTC8X_ERR_INVALID_ARGUMENT, "nanovdb: the grid is empty (no root tiles)");
        if ((uint64_t)table_size * kTileBytes > size - root - kRootBytes)
            return fail(RTC8X_ERR_INVALID_ARGUMENT, "nanovdb: %u root tiles overrun the buffer", table_size);
        nodes_begin = root + kRootBytes + (uint64_t)table_size * kTileBytes;
        const float background = rd<float>(b + root + 28);
        tile_entry(background);  // tile_values[0] is the background

        for (uint32_t i = 0; i < table_size; ++i) {
            const uint8_t* tile = b + root + kRootBytes + (size_t)i * kTileBytes;
            const uint64_t key = rd<uint64_t>(tile);
            const int64_t child = rd<int64_t>(tile + 8);
            rtc8x_root_tile rt{key, 0u, rd<uint32_t>(tile + 16)};
            if (child != 0) {
                int64_t org[3];
                key_origin(key, org);
                uint32_t idx;
                rtc8x_status st = link(root, child, kUpperBytes, upper_at, upper_off, upper_org, org, "upper node", &idx);
                if (st != RTC8X_OK) return st;
                rt.entry = idx;
            } else {
                rt.entry = tile_entry(rd<float>(tile + 20));
            }
            t.root.push_back(rt);
        }
        // upper nodes -> lower links, in traversal order
        t.upper.resize(upper_off.size() * (size_t)RTC8X_UPPER_TABLE_SIZE);
        t.upper_masks.resize(upper_off.size() * (kUpperMask / 8));
        for (size_t u = 0; u < upper_off.size(); ++u) {
            const uint8_t* node = b + upper_off[u];
            const uint8_t* cmask = node + 32 + kUpperMask;
            std::memcpy(&t.upper_masks[u * (kUpperMask / 8)], node + 32, kUpperMask);
            for (uint32_t n = 0; n < RTC8X_UPPER_TABLE_SIZE; ++n) {
                const uint8_t* e = node + kUpperTable + (size_t)n * 8;
                uint32_t entry;
                if ((rd<uint64_t>(cmask + (n >> 6) * 8) >> (n & 63)) & 1u) {
                    const int64_t org[3] = {upper_org[3 * u] + ((int64_t)(n >> 10) << 7),
                                            upper_org[3 * u + 1] + ((int64_t)((n >> 5) & 31) << 7),
                                            upper_org[3 * u + 2] + ((int64_t)(n & 31) << 7)};
                    rtc8x_status st = link(upper_off[u], rd<int64_t>(e), kLowerBytes, lower_at, lower_off, lower_org,
                                           org, "lower node", &entry);
                    if (st != RTC8X_OK) return st;
                } else {
                    entry = tile_entry(rd<float>(e));
                }
                t.upper[u * RTC8X_UPPER_TABLE_SIZE + n] = entry;
            }
        }
        t.lower.resize(lower_off.size() * (size_t)RTC8X_LOWER_TABLE_SIZE);
        t.lower_masks.resize(lower_off.size() * (kLowerMask / 8));
        for (size_t l = 0; l < lower_off.size(); ++l) {
            const uint8_t* node = b + lower_off[l];
            const uint8_t* cmask = node + 32 + kLowerMask;
            std::memcpy(&t.lower_masks[l * (kLowerMask / 8)], node + 32, kLowerMask);
            for (uint32_t n = 0; n < RTC8X_LOWER_TABLE_SIZE; ++n) {
                const uint8_t* e = node + kLowerTable + (size_t)n * 8;
                uint32_t entry;
                if ((rd<uint64_t>(cmask + (n >> 6) * 8) >> (n & 63)) & 1u) {
                    const int64_t org[3] = {lower_org[3 * l] + ((int64_t)(n >> 8) << 3),
                                            lower_org[3 * l + 1] + ((int64_t)((n >> 4) & 15) << 3),
                                            lower_org[3 * l + 2] + ((int64_t)(n & 15) << 3)};
                    rtc8x_status st = link(lower_off[l], rd<int64_t>(e), kLeafBytes, leaf_at, leaf_off, leaf_org, org,
                                           "leaf", &entry);
                    if (st != RTC8X_OK) return st;
                } else {
                    entry = tile_entry(rd<float>(e));
                }
                t.lower[l * RTC8X_LOWER_TABLE_SIZE + n] = entry;
            }
        }
        t.values.resize(leaf_off.size() * (size_t)RTC8X_LEAF_SIZE);
        t.origins.resize(leaf_off.size() * 3);
        t.leaf_masks.resize(leaf_off.size() * 8);
        for (size_t f = 0; f < leaf_off.size(); ++f) {
            const uint8_t* leaf = b + leaf_off[f];
            std::memcpy(&t.leaf_masks[f * 8], leaf + 16, 64);
            // LeafNode::origin() is the stored bbox minimum with the low three bits cleared (NanoVDB.h:3745)
            for (int a = 0; a < 3; ++a) {
                const int32_t stored = rd<int32_t>(leaf + 4 * a) & ~7;
                if ((int64_t)stored != leaf_org[3 * f + a])
                    return fail(RTC8X_ERR_INVALID_ARGUMENT, "nanovdb: leaf at offset %llu has origin %d on axis %d but "
                                "sits at %lld in the tree", (unsigned long long)leaf_off[f], stored, a,
                                (long long)leaf_org[3 * f + a]);
                t.origins[3 * f + a] = stored;
            }
            std::memcpy(&t.values[f * RTC8X_LEAF_SIZE], leaf + kLeafValues, RTC8X_LEAF_SIZE * 4);
        }
        const uint32_t counts[3] = {rd<uint32_t>(tree + 32), rd<uint32_t>(tree + 36), rd<uint32_t>(tree + 40)};
        if (counts[0] != leaf_off.size() || counts[1] != lower_off.size() || counts[2] != upper_off.size())
            return fail(RTC8X_ERR_INVALID_ARGUMENT, "nanovdb: the tree records %u/%u/%u leaf/lower/upper nodes but its "
                        "links reach %zu/%zu/%zu", counts[0], counts[1], counts[2], leaf_off.size(), lower_off.size(),
                        upper_off.size());

        std::memset(&g, 0, sizeof(g));
        const uint8_t* map = b + 296;
        for (int i = 0; i < 9; ++i) g.inv_mat[i] = rd<float>(map + 36 + 4 * i);
        for (int i = 0; i < 3; ++i) g.vec[i] = rd<float>(map + 72 + 4 * i);
        for (int i = 0; i < 9; ++i) g.map_mat_d[i] = rd<double>(map + 88 + 8 * i);
        for (int i = 0; i < 9; ++i) g.map_inv_mat_d[i] = rd<double>(map + 160 + 8 * i);
        for (int i = 0; i < 3; ++i) g.map_vec_d[i] = rd<double>(map + 232 + 8 * i);
        // Scene::allocateVolumeGrid keeps the world bbox as floats (renderer_scene.cpp:494,513-515)
        for (int i = 0; i < 3; ++i) {
            g.world_bbox_min[i] = (float)rd<double>(b + 560 + 8 * i);
            g.world_bbox_max[i] = (float)rd<double>(b + 584 + 8 * i);
            g.index_bbox_min[i] = rd<int32_t>(b + root + 4 * i);
            g.index_bbox_max[i] = rd<int32_t>(b + root + 12 + 4 * i);
        }
        g.background = background;
        g.num_root_tiles = (uint32_t)t.root.size();
        g.num_upper = (uint32_t)upper_off.size();
        g.num_lower = (uint32_t)lower_off.size();
        g.num_leaves = (uint32_t)leaf_off.size();
        return majorant(g);
    }

    // The reference's majorant: max of getValue over [min, max) of the root's index bbox on each axis
    // (renderer_scene.cpp:497-511), evaluated per node instead of per voxel.
    rtc8x_status majorant(rtc8x_grid& g)
    {
        int64_t lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = g.index_bbox_min[a];
            hi[a] = g.index_bbox_max[a];
            if (lo[a] >= hi[a])
                return fail(RTC8X_ERR_INVALID_ARGUMENT, "nanovdb: the index bbox [%d, %d] on axis %d leaves the "
                            "majorant loop empty", g.index_bbox_min[a], g.index_bbox_max[a], a);
        }
        auto overlaps = [&](const int64_t o[3], int64_t dim) {
            for (int a = 0; a < 3; ++a)
                if (o[a] + dim <= lo[a] || o[a] >= hi[a]) return false;
            return true;
        };
        float m = -INFINITY;
        auto take = [&](float v) { m = std::max(v, m); };
        for (size_t f = 0; f < leaf_off.size(); ++f) {
            const int64_t* o = &leaf_org[3 * f];
            if (!overlaps(o, 8)) continue;
            int64_t r0[3], r1[3];
            for (int a = 0; a < 3; ++a) {
                r0[a] = std::max(lo[a], o[a]) - o[a];
                r1[a] = std::min(hi[a], o[a] + 8) - o[a];
            }
            for (int64_t x = r0[0]; x < r1[0]; ++x)
                for (int64_t y = r0[1]; y < r1[1]; ++y)
                    for (int64_t z = r0[2]; z < r1[2]; ++z) take(t.values[f * RTC8X_LEAF_SIZE + ((x << 6) | (y << 3) | z)]);
        }
        auto tile_value = [&](uint32_t e) { return t.tiles[e & ~RTC8X_TILE_BIT]; };
        for (size_t l = 0; l < lower_off.size(); ++l)
            for (uint32_t n = 0; n < RTC8X_LOWER_TABLE_SIZE; ++n) {
                const uint32_t e = t.lower[l * RTC8X_LOWER_TABLE_SIZE + n];
                if (!(e & RTC8X_TILE_BIT)) continue;
                const int64_t o[3] = {lower_org[3 * l] + ((int64_t)(n >> 8) << 3),
                                      lower_org[3 * l + 1] + ((int64_t)((n >> 4) & 15) << 3),
                                      lower_org[3 * l + 2] + ((int64_t)(n & 15) << 3)};
                if (overlaps(o, 8)) take(tile_value(e));
            }
        for (size_t u = 0; u < upper_off.size(); ++u)
            for (uint32_t n = 0; n < RTC8X_UPPER_TABLE_SIZE; ++n) {
                const uint32_t e = t.upper[u * RTC8X_UPPER_TABLE_SIZE + n];
                if (!(e & RTC8X_TILE_BIT)) continue;
                const int64_t o[3] = {upper_org[3 * u] + ((int64_t)(n >> 10) << 7),
                                      upper_org[3 * u + 1] + ((int64_t)((n >> 5) & 31) << 7),
                                      upper_org[3 * u + 2] + ((int64_t)(n & 31) << 7)};
                if (overlaps(o, 128)) take(tile_value(e));
            }
        // root tiles, and the background wherever no root tile covers part of the range
        std::unordered_map<uint64_t, uint32_t> keys;
        uint64_t covering = 0;
        for (const rtc8x_root_tile& rt : t.root) {
            int64_t o[3];
            key_origin(rt.key, o);
            if (!overlaps(o, 4096)) continue;
            if (keys.emplace(rt.key, rt.entry).second) ++covering;
            if (rt.entry & RTC8X_TILE_BIT) take(tile_value(rt.entry));
        }
        double cubes = 1.0;
        int64_t c0[3], c1[3];
        for (int a = 0; a < 3; ++a) {
            c0[a] = lo[a] >> 12;
            c1[a] = (hi[a] - 1) >> 12;
            cubes *= (double)(c1[a] - c0[a] + 1);
        }
        bool uncovered = cubes > (double)covering;
        if (!uncovered) {
            for (int64_t x = c0[0]; x <= c1[0] && !uncovered; ++x)
                for (int64_t y = c0[1]; y <= c1[1] && !uncovered; ++y)
                    for (int64_t z = c0[2]; z <= c1[2] && !uncovered; ++z) {
                        const uint64_t key = ((uint64_t)((uint32_t)(z << 12) >> 12)) |
                                             ((uint64_t)((uint32_t)(y << 12) >> 12) << 21) |
                                             ((uint64_t)((uint32_t)(x << 12) >> 12) << 42);
                        if (!keys.count(key)) uncovered = true;
                    }
        }
        if (uncovered) take(g.background);
        if (!std::isfinite(m))
            return fail(RTC8X_ERR_INVALID_ARGUMENT, "nanovdb: the grid holds non-finite values in its index bbox");
        g.max_value = m;
        return RTC8X_OK;
    }
};

rtc8x_status decode_buffer(const void* buf, size_t size, rtc8x_grid* out)
{
    if (!buf || !out) return fail(RTC8X_ERR_INVALID_ARGUMENT, "rtc8x_nanovdb_decode: NULL argument");
    Decoder d;
    d.b = (const uint8_t*)buf;
    d.size = size;
    rtc8x_grid g;
    rtc8x_status st = d.run(g);
    if (st != RTC8X_OK) return st;
    return adopt_host_grid(std::move(d.t), g, out);
}

// ---------------------------------------------------------------------------------------------------
// Files
// ---------------------------------------------------------------------------------------------------

// Root-key hash of a grid name (nanovdb::io::stringHash, util/IO.h:733-745).
uint64_t string_hash(const char* s)
{
    uint64_t h = 0;
    for (const unsigned char* p = (const unsigned char*)s; *p; ++p) {
        const uint64_t overflow = h >> 56;
        h = h * 67u + *p + overflow;
    }
    return h;
}

rtc8x_status list_file(const uint8_t* f, uint64_t size, std::vector<rtc8x_nvdb_grid_info>& out)
{
    uint64_t off = 0;
    uint32_t segment = 0;
    while (off < size) {
        if (size - off < kHeaderBytes)
            return fail(RTC8X_ERR_INVALID_ARGUMENT, "nvdb: truncated segment header at offset %llu",
                        (unsigned long long)off);
        const uint64_t magic = rd<uint64_t>(f + off);
        if (magic != kMagic) {
            if (magic == __builtin_bswap64(kMagic))
                return fail(RTC8X_ERR_NOT_SUPPORTED, "nvdb: the file has reversed endianness");
            return fail(RTC8X_ERR_INVALID_ARGUMENT, "nvdb: bad magic number at offset %llu: not a .nvdb file",
                        (unsigned long long)off);
        }
        const uint32_t version = rd<uint32_t>(f + off + 8);
        if ((version >> 21) != kMajor)
            return fail(RTC8X_ERR_NOT_SUPPORTED, "nvdb: file version %u.%u.%u, expected major version %u", version >> 21,
                        (version >> 10) & 2047u, version & 1023u, kMajor);
        const uint32_t count = rd<uint16_t>(f + off + 12);
        const uint32_t codec = rd<uint16_t>(f + off + 14);
        off += kHeaderBytes;
        const size_t first = out.size();
        for (uint32_t i = 0; i < count; ++i) {
            if (size - off < kMetaBytes)
                return fail(RTC8X_ERR_INVALID_ARGUMENT, "nvdb: truncated grid record at offset %llu",
                            (unsigned long long)off);
            const uint8_t* m = f + off;
            rtc8x_nvdb_grid_info info;
            std::memset(&info, 0, sizeof(info));
            info.grid_size = rd<uint64_t>(m);
            info.file_size = rd<uint64_t>(m + 8);
            info.voxel_count = rd<uint64_t>(m + 24);
            info.grid_type = rd<uint32_t>(m + 32);
            info.grid_class = rd<uint32_t>(m + 36);
            for (int k = 0; k < 6; ++k) info.world_bbox[k] = rd<double>(m + 40 + 8 * k);
            for (int k = 0; k < 6; ++k) info.index_bbox[k] = rd<int32_t>(m + 88 + 4 * k);
            for (int k = 0; k < 3; ++k) info.voxel_size[k] = rd<double>(m + 112 + 8 * k);
            const uint32_t name_size = rd<uint32_t>(m + 136);
            for (int k = 0; k < 4; ++k) info.node_count[k] = rd<uint32_t>(m + 140 + 4 * k);
            for (int k = 0; k < 3; ++k) info.tile_count[k] = rd<uint32_t>(m + 156 + 4 * k);
            info.codec = codec;  // the segment header's codec governs the data (util/IO.h:601)
            info.version = version;
            info.segment = segment;
            off += kMetaBytes;
            if (size - off < name_size)
                return fail(RTC8X_ERR_INVALID_ARGUMENT, "nvdb: truncated grid name at offset %llu",
                            (unsigned long long)off);
            const size_t n = std::min<size_t>(name_size, sizeof(info.name) - 1);
            std::memcpy(info.name, f + off, n);
            info.name[n] = 0;
            off += name_size;
            out.push_back(info);
        }
        for (size_t i = first; i < out.size(); ++i) {
            if (size - off < out[i].file_size)
                return fail(RTC8X_ERR_INVALID_ARGUMENT, "nvdb: grid %zu (\"%s\") overruns the file", i, out[i].name);
            out[i].file_offset = off;
            off += out[i].file_size;
        }
        ++segment;
    }
    return RTC8X_OK;
}

// The grid buffer of one listed grid, decompressed as its codec requires (util/IO.h:301-344).
rtc8x_status grid_bytes(const uint8_t* f, const rtc8x_nvdb_grid_info& info, std::vector<uint8_t>& buf,
                        const uint8_t** data)
{
    if (info.codec == RTC8X_NVDB_CODEC_NONE) {
        if (info.file_size < info.grid_size)
            return fail(RTC8X_ERR_INVALID_ARGUMENT, "nvdb: grid \"%s\" stores %llu of its %llu bytes", info.name,
                        (unsigned long long)info.file_size, (unsigned long long)info.grid_size);
        *data = f + info.file_offset;
        return RTC8X_OK;
    }
    if (info.codec == RTC8X_NVDB_CODEC_ZIP) {
        if (info.file_size < 8) return fail(RTC8X_ERR_INVALID_ARGUMENT, "nvdb: truncated ZIP block");
        const uint64_t zbytes = rd<uint64_t>(f + info.file_offset);
        if (zbytes > info.file_size - 8) return fail(RTC8X_ERR_INVALID_ARGUMENT, "nvdb: ZIP block overruns the grid");
        // deflate expands at most 1032:1 (zlib's maximum ratio), so a header claiming more than that
        // is corrupt; it is rejected before the buffer is sized from it
        if (info.grid_size < kGridBytes + kTreeBytes + kRootBytes || info.grid_size / 1032 > zbytes)
            return fail(RTC8X_ERR_INVALID_ARGUMENT, "nvdb: grid \"%s\" claims %llu bytes from a %llu-byte ZIP block",
                        info.name, (unsigned long long)info.grid_size, (unsigned long long)zbytes);
        buf.resize(info.grid_size);
        uLongf n = (uLongf)info.grid_size;
        const int zs = uncompress(buf.data(), &n, f + info.file_offset + 8, (uLong)zbytes);
        if (zs != Z_OK || (uint64_t)n != info.grid_size)
            return fail(RTC8X_ERR_INVALID_ARGUMENT, "nvdb: ZIP data of grid \"%s\" does not inflate to %llu bytes "
                        "(zlib status %d)", info.name, (unsigned long long)info.grid_size, zs);
        *data = buf.data();
        return RTC8X_OK;
    }
    if (info.codec == RTC8X_NVDB_CODEC_BLOSC)
        return fail(RTC8X_ERR_NOT_SUPPORTED, "nvdb: grid \"%s\" is BLOSC-compressed; only NONE and ZIP are supported",
                    info.name);
    return fail(RTC8X_ERR_INVALID_ARGUMENT, "nvdb: unknown codec %u", info.codec);
}

rtc8x_status read_grid(const void* file, size_t size, uint32_t n, std::vector<uint8_t>& buf, const uint8_t** data,
                       uint64_t* grid_size)
{
    if (!file) return fail(RTC8X_ERR_INVALID_ARGUMENT, "nvdb: NULL file image");
    std::vector<rtc8x_nvdb_grid_info> infos;
    rtc8x_status st = list_file((const uint8_t*)file, size, infos);
    if (st != RTC8X_OK) return st;
    if (n >= infos.size())
        return fail(RTC8X_ERR_INVALID_ARGUMENT, "nvdb: grid index %u exceeds the %zu grids in the file", n, infos.size());
    *grid_size = infos[n].grid_size;
    return grid_bytes((const uint8_t*)file, infos[n], buf, data);
}

// ---------------------------------------------------------------------------------------------------
// Encoding
// ---------------------------------------------------------------------------------------------------

// Population statistics as nanovdb::Stats<float> keeps them: extrema plus a running mean and sum of
// squared deviations in double (Welford 1962), merged pairwise (Chan, Golub & LeVeque 1979), in the
// order nanovdb::GridStats visits values (util/GridStats.h:207-301, 583-708).
struct Stats {
    float mn = FLT_MAX, mx = -FLT_MAX;
    uint64_t n = 0;
    double avg = 0.0, aux = 0.0;
    void add(float v)
    {
        if (v < mn) mn = v;
        if (v > mx) mx = v;
        n += 1;
        const double delta = (double)v - avg;
        avg += delta / (double)n;
        aux += delta * ((double)v - avg);
    }
    void add(float v, uint64_t count)
    {
        const double denom = 1.0 / (double)(n + count);
        const double delta = (double)v - avg;
        avg += denom * delta * (double)count;
        aux += denom * delta * delta * (double)n * (double)count;
        if (v < mn) mn = v;
        if (v > mx) mx = v;
        n += count;
    }
    void add(const Stats& o)
    {
        if (o.mn < mn) mn = o.mn;
        if (o.mx > mx) mx = o.mx;
        if (o.n > 0) {
            const double denom = 1.0 / (double)(n + o.n);
            const double delta = o.avg - avg;
            avg += denom * delta * (double)o.n;
            aux += o.aux + denom * delta * delta * (double)n * (double)o.n;
            n += o.n;
        }
    }
    float stddev() const { return (float)std::sqrt(n < 2 ? 0.0 : aux / (double)n); }
};

struct NodeStats {
    Stats s;
    int32_t bmin[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, bmax[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    void add(const NodeStats& o)
    {
        s.add(o.s);
        for (int a = 0; a < 3; ++a) {
            bmin[a] = std::min(bmin[a], o.bmin[a]);
            bmax[a] = std::max(bmax[a], o.bmax[a]);
        }
    }
    void grow(const int32_t o[3], int32_t dim)
    {
        for (int a = 0; a < 3; ++a) {
            bmin[a] = std::min(bmin[a], o[a]);
            bmax[a] = std::max(bmax[a], o[a] + dim - 1);
        }
    }
    bool empty() const { return bmin[0] > bmax[0] || bmin[1] > bmax[1] || bmin[2] > bmax[2]; }
};

bool bit(const uint64_t* m, uint32_t n) { return (m[n >> 6] >> (n & 63)) & 1u; }

void write_bbox(uint8_t* p, const NodeStats& ns)
{
    for (int a = 0; a < 3; ++a) {
        wr<int32_t>(p + 4 * a, ns.bmin[a]);
        wr<int32_t>(p + 12 + 4 * a, ns.bmax[a]);
    }
}
void write_stats(uint8_t* p, const Stats& s)
{
    wr<float>(p, s.mn);
    wr<float>(p + 4, s.mx);
    wr<float>(p + 8, (float)s.avg);
    wr<float>(p + 12, s.stddev());
}

// matMult(mat, vec, xyz) in double (NanoVDB.h:1433-1438)
void apply_map(const double m[9], const double v[3], const double x[3], double out[3])
{
    for (int r = 0; r < 3; ++r) out[r] = std::fma(x[0], m[3 * r], std::fma(x[1], m[3 * r + 1], std::fma(x[2], m[3 * r + 2], v[r])));
}

rtc8x_status validate_tables(const rtc8x_grid* h, uint32_t* num_tiles)
{
    RTC8X_REQUIRE(h->num_root_tiles > 0 && h->root_tiles && h->tile_values, "rtc8x_nanovdb_encode: grid has no root");
    RTC8X_REQUIRE(h->num_upper == 0 || h->upper_tables, "rtc8x_nanovdb_encode: upper tables missing");
    RTC8X_REQUIRE(h->num_lower == 0 || h->lower_tables, "rtc8x_nanovdb_encode: lower tables missing");
    RTC8X_REQUIRE(h->num_leaves == 0 || h->leaf_values, "rtc8x_nanovdb_encode: leaf values missing");
    uint32_t nt = 1;
    auto check = [&](uint32_t e, uint32_t limit) {
        if (e & RTC8X_TILE_BIT) {
            nt = std::max(nt, (e & ~RTC8X_TILE_BIT) + 1);
            return true;
        }
        return e < limit;
    };
    for (uint32_t i = 0; i < h->num_root_tiles; ++i)
        RTC8X_REQUIRE(check(h->root_tiles[i].entry, h->num_upper), "rtc8x_nanovdb_encode: root tile %u out of range", i);
    for (size_t i = 0; i < (size_t)h->num_upper * RTC8X_UPPER_TABLE_SIZE; ++i)
        RTC8X_REQUIRE(check(h->upper_tables[i], h->num_lower), "rtc8x_nanovdb_encode: upper entry %zu out of range", i);
    for (size_t i = 0; i < (size_t)h->num_lower * RTC8X_LOWER_TABLE_SIZE; ++i)
        RTC8X_REQUIRE(check(h->lower_tables[i], h->num_leaves), "rtc8x_nanovdb_encode: lower entry %zu out of range", i);
    *num_tiles = nt;
    return RTC8X_OK;
}

// 3x3 inverse by cofactors, in double.
bool invert3(const double a[9], double inv[9])
{
    const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
    const double r = 1.0 / det;
    inv[0] = c00 * r;
    inv[1] = (a[2] * a[7] - a[1] * a[8]) * r;
    inv[2] = (a[1] * a[5] - a[2] * a[4]) * r;
    inv[3] = c01 * r;
    inv[4] = (a[0] * a[8] - a[2] * a[6]) * r;
    inv[5] = (a[2] * a[3] - a[0] * a[5]) * r;
    inv[6] = c02 * r;
    inv[7] = (a[1] * a[6] - a[0] * a[7]) * r;
    inv[8] = (a[0] * a[4] - a[1] * a[3]) * r;
    return true;
}

struct Encoder {
    const rtc8x_grid* h;
    std::vector<uint32_t> upper_order, lower_order, leaf_order;  // host index of the i-th node written
    std::vector<int32_t> upper_org, lower_org, leaf_org;
    uint64_t upper_base = 0, lower_base = 0, leaf_base = 0, total = 0;

    rtc8x_status plan()
    {
        std::vector<uint8_t> seen_u(h->num_upper, 0), seen_l(h->num_lower, 0), seen_f(h->num_leaves, 0);
        for (uint32_t i = 0; i < h->num_root_tiles; ++i) {
            const uint32_t e = h->root_tiles[i].entry;
            if (e & RTC8X_TILE_BIT) continue;
            RTC8X_REQUIRE(!seen_u[e], "rtc8x_nanovdb_encode: upper node %u is linked twice", e);
            seen_u[e] = 1;
            int64_t o[3];
            key_origin(h->root_tiles[i].key, o);
            upper_order.push_back(e);
            for (int a = 0; a < 3; ++a) upper_org.push_back((int32_t)o[a]);
        }
        for (size_t k = 0; k < upper_order.size(); ++k)
            for (uint32_t n = 0; n < RTC8X_UPPER_TABLE_SIZE; ++n) {
                const uint32_t e = h->upper_tables[(size_t)upper_order[k] * RTC8X_UPPER_TABLE_SIZE + n];
                if (e & RTC8X_TILE_BIT) continue;
                RTC8X_REQUIRE(!seen_l[e], "rtc8x_nanovdb_encode: lower node %u is linked twice", e);
                seen_l[e] = 1;
                lower_order.push_back(e);
                lower_org.push_back(upper_org[3 * k] + (int32_t)((n >> 10) << 7));
                lower_org.push_back(upper_org[3 * k + 1] + (int32_t)(((n >> 5) & 31) << 7));
                lower_org.push_back(upper_org[3 * k + 2] + (int32_t)((n & 31) << 7));
            }
        for (size_t k = 0; k < lower_order.size(); ++k)
            for (uint32_t n = 0; n < RTC8X_LOWER_TABLE_SIZE; ++n) {
                const uint32_t e = h->lower_tables[(size_t)lower_order[k] * RTC8X_LOWER_TABLE_SIZE + n];
                if (e & RTC8X_TILE_BIT) continue;
                RTC8X_REQUIRE(!seen_f[e], "rtc8x_nanovdb_encode: leaf %u is linked twice", e);
                seen_f[e] = 1;
                leaf_order.push_back(e);
                leaf_org.push_back(lower_org[3 * k] + (int32_t)((n >> 8) << 3));
                leaf_org.push_back(lower_org[3 * k + 1] + (int32_t)(((n >> 4) & 15) << 3));
                leaf_org.push_back(lower_org[3 * k + 2] + (int32_t)((n & 15) << 3));
            }
        // an active bit belongs to a tile, never to a child link
        for (uint32_t u = 0; h->upper_masks && u < h->num_upper; ++u)
            for (uint32_t n = 0; n < RTC8X_UPPER_TABLE_SIZE; ++n)
                RTC8X_REQUIRE(!bit(h->upper_masks + (size_t)u * 512, n) ||
                                  (h->upper_tables[(size_t)u * RTC8X_UPPER_TABLE_SIZE + n] & RTC8X_TILE_BIT),
                              "rtc8x_nanovdb_encode: upper node %u entry %u is a child but marked active", u, n);
        for (uint32_t l = 0; h->lower_masks && l < h->num_lower; ++l)
            for (uint32_t n = 0; n < RTC8X_LOWER_TABLE_SIZE; ++n)
                RTC8X_REQUIRE(!bit(h->lower_masks + (size_t)l * 64, n) ||
                                  (h->lower_tables[(size_t)l * RTC8X_LOWER_TABLE_SIZE + n] & RTC8X_TILE_BIT),
                              "rtc8x_nanovdb_encode: lower node %u entry %u is a child but marked active", l, n);
        upper_base = kGridBytes + kTreeBytes + kRootBytes + (uint64_t)h->num_root_tiles * kTileBytes;
        lower_base = upper_base + (uint64_t)upper_order.size() * kUpperBytes;
        leaf_base = lower_base + (uint64_t)lower_order.size() * kLowerBytes;
        total = leaf_base + (uint64_t)leaf_order.size() * kLeafBytes;
        return RTC8X_OK;
    }

    float tile_value(uint32_t e) const { return h->tile_values[e & ~RTC8X_TILE_BIT]; }

    rtc8x_status write(uint8_t* b, const char* name)
    {
        std::memset(b, 0, total);
        const uint64_t root = kGridBytes + kTreeBytes;
        // leaves: tight bbox and statistics of the active voxels (LeafNode::updateBBox, NanoVDB.h:3884-3921)
        static const uint64_t all_on[8] = {~0ull, ~0ull, ~0ull, ~0ull, ~0ull, ~0ull, ~0ull, ~0ull};
        uint64_t voxel_count = 0, tile_count[3] = {0, 0, 0};
        std::vector<NodeStats> leaf_stats(leaf_order.size());
        for (size_t k = 0; k < leaf_order.size(); ++k) {
            uint8_t* p = b + leaf_base + k * kLeafBytes;
            const float* v = h->leaf_values + (size_t)leaf_order[k] * RTC8X_LEAF_SIZE;
            const uint64_t* m = h->leaf_masks ? h->leaf_masks + (size_t)leaf_order[k] * 8 : all_on;
            NodeStats& ns = leaf_stats[k];
            const int32_t* o = &leaf_org[3 * k];
            for (uint32_t i = 0; i < RTC8X_LEAF_SIZE; ++i) {
                if (!bit(m, i)) continue;
                const int32_t c[3] = {o[0] + (int32_t)(i >> 6), o[1] + (int32_t)((i >> 3) & 7), o[2] + (int32_t)(i & 7)};
                ns.grow(c, 1);
                ns.s.add(v[i]);
            }
            voxel_count += ns.s.n;
            for (int a = 0; a < 3; ++a) {
                wr<int32_t>(p + 4 * a, ns.empty() ? o[a] : ns.bmin[a]);
                p[12 + a] = ns.empty() ? 0 : (uint8_t)(ns.bmax[a] - ns.bmin[a]);
            }
            std::memcpy(p + 16, m, 64);
            if (!ns.empty()) {
                p[15] = 2;  // holds active values
                write_stats(p + kLeafStats, ns.s);
            }
            std::memcpy(p + kLeafValues, v, RTC8X_LEAF_SIZE * 4);
        }
        // internal nodes: active tiles in mask order, then children in table order (GridStats.h:611-680)
        auto internal = [&](uint8_t* p, uint64_t node_off, const uint32_t* table, uint32_t count, size_t mask_bytes,
                            size_t stats_at, size_t table_at, std::vector<NodeStats>& child_stats,
                            const std::vector<uint64_t>& child_off, std::unordered_map<uint32_t, uint32_t>& slot,
                            const uint64_t* vmask, const int32_t* org, int log2dim, int32_t child_dim,
                            uint64_t& tiles) {
            NodeStats ns;
            uint8_t* cmask = p + 32 + mask_bytes;
            if (vmask) {
                std::memcpy(p + 32, vmask, mask_bytes);
                const uint64_t child_voxels = (uint64_t)child_dim * child_dim * child_dim;
                for (uint32_t n = 0; n < count; ++n) {
                    if (!bit(vmask, n)) continue;
                    const int32_t c[3] = {org[0] + (int32_t)(n >> (2 * log2dim)) * child_dim,
                                          org[1] + (int32_t)((n >> log2dim) & ((1u << log2dim) - 1)) * child_dim,
                                          org[2] + (int32_t)(n & ((1u << log2dim) - 1)) * child_dim};
                    ns.s.add(tile_value(table[n]), child_voxels);
                    ns.grow(c, child_dim);
                    ++tiles;
                }
            }
            for (uint32_t n = 0; n < count; ++n) {
                const uint32_t e = table[n];
                uint8_t* entry = p + table_at + (size_t)n * 8;
                if (e & RTC8X_TILE_BIT) {
                    wr<float>(entry, tile_value(e));
                } else {
                    const uint32_t c = slot.at(e);
                    cmask[n >> 3] |= (uint8_t)(1u << (n & 7));
                    wr<int64_t>(entry, (int64_t)child_off[c] - (int64_t)node_off);
                    ns.add(child_stats[c]);
                }
            }
            write_bbox(p, ns);
            if (!ns.empty()) {
                wr<uint64_t>(p + 24, 2u);
                write_stats(p + stats_at, ns.s);
            }
            return ns;
        };
        std::unordered_map<uint32_t, uint32_t> leaf_slot, lower_slot, upper_slot;
        std::vector<uint64_t> leaf_off(leaf_order.size()), lower_off(lower_order.size()), upper_off(upper_order.size());
        for (size_t k = 0; k < leaf_order.size(); ++k) {
            leaf_slot[leaf_order[k]] = (uint32_t)k;
            leaf_off[k] = leaf_base + k * kLeafBytes;
        }
        std::vector<NodeStats> lower_stats(lower_order.size());
        for (size_t k = 0; k < lower_order.size(); ++k) {
            lower_slot[lower_order[k]] = (uint32_t)k;
            lower_off[k] = lower_base + k * kLowerBytes;
            lower_stats[k] = internal(b + lower_off[k], lower_off[k],
                                      h->lower_tables + (size_t)lower_order[k] * RTC8X_LOWER_TABLE_SIZE,
                                      RTC8X_LOWER_TABLE_SIZE, kLowerMask, kLowerStats, kLowerTable, leaf_stats, leaf_off,
                                      leaf_slot, h->lower_masks ? h->lower_masks + (size_t)lower_order[k] * 64 : nullptr,
                                      &lower_org[3 * k], 4, 8, tile_count[0]);
        }
        std::vector<NodeStats> upper_stats(upper_order.size());
        for (size_t k = 0; k < upper_order.size(); ++k) {
            upper_slot[upper_order[k]] = (uint32_t)k;
            upper_off[k] = upper_base + k * kUpperBytes;
            upper_stats[k] = internal(b + upper_off[k], upper_off[k],
                                      h->upper_tables + (size_t)upper_order[k] * RTC8X_UPPER_TABLE_SIZE,
                                      RTC8X_UPPER_TABLE_SIZE, kUpperMask, kUpperStats, kUpperTable, lower_stats,
                                      lower_off, lower_slot,
                                      h->upper_masks ? h->upper_masks + (size_t)upper_order[k] * 512 : nullptr,
                                      &upper_org[3 * k], 5, 128, tile_count[1]);
        }
        // root
        NodeStats total_stats;
        for (uint32_t i = 0; i < h->num_root_tiles; ++i) {
            uint8_t* tile = b + root + kRootBytes + (size_t)i * kTileBytes;
            const uint32_t e = h->root_tiles[i].entry;
            wr<uint64_t>(tile, h->root_tiles[i].key);
            if (e & RTC8X_TILE_BIT) {
                const uint32_t state = h->root_tiles[i].state ? 1u : 0u;
                wr<uint32_t>(tile + 16, state);
                wr<float>(tile + 20, tile_value(e));
                if (state) {
                    int64_t o64[3];
                    key_origin(h->root_tiles[i].key, o64);
                    const int32_t o[3] = {(int32_t)o64[0], (int32_t)o64[1], (int32_t)o64[2]};
                    total_stats.s.add(tile_value(e), 1ull << 36);
                    total_stats.grow(o, 4096);
                    ++tile_count[2];
                }
            } else {
                const uint32_t c = upper_slot.at(e);
                wr<int64_t>(tile + 8, (int64_t)upper_off[c] - (int64_t)root);
                total_stats.add(upper_stats[c]);
            }
        }
        write_bbox(b + root, total_stats);
        wr<uint32_t>(b + root + 24, h->num_root_tiles);
        wr<float>(b + root + 28, h->tile_values[0]);
        write_stats(b + root + 32, total_stats.s);
        // tree
        uint8_t* tree = b + kGridBytes;
        wr<uint64_t>(tree + 0, leaf_order.empty() ? 0 : leaf_base - kGridBytes);
        wr<uint64_t>(tree + 8, lower_order.empty() ? 0 : lower_base - kGridBytes);
        wr<uint64_t>(tree + 16, upper_order.empty() ? 0 : upper_base - kGridBytes);
        wr<uint64_t>(tree + 24, root - kGridBytes);
        wr<uint32_t>(tree + 32, (uint32_t)leaf_order.size());
        wr<uint32_t>(tree + 36, (uint32_t)lower_order.size());
        wr<uint32_t>(tree + 40, (uint32_t)upper_order.size());
        for (int i = 0; i < 3; ++i) wr<uint32_t>(tree + 44 + 4 * i, (uint32_t)tile_count[i]);
        voxel_count += tile_count[0] * 512ull + tile_count[1] * (1ull << 21) + tile_count[2] * (1ull << 36);
        wr<uint64_t>(tree + 56, voxel_count);
        // grid
        wr<uint64_t>(b, kMagic);
        wr<uint32_t>(b + 16, kVersion);
        uint32_t flags = kFlagIsBreadthFirst | kFlagHasMinMax | kFlagHasAverage | kFlagHasStdDeviation;
        wr<uint32_t>(b + 28, 1u);
        wr<uint64_t>(b + 32, total);
        std::strncpy((char*)b + 40, name, 255);
        double inv[9], mat[9], vec[3];
        bool have_d = false;
        for (int i = 0; i < 9; ++i) have_d |= h->map_mat_d[i] != 0.0;
        if (have_d) {
            for (int i = 0; i < 9; ++i) {
                mat[i] = h->map_mat_d[i];
                inv[i] = h->map_inv_mat_d[i];
            }
            for (int i = 0; i < 3; ++i) vec[i] = h->map_vec_d[i];
        } else {
            for (int i = 0; i < 9; ++i) inv[i] = (double)h->inv_mat[i];
            for (int i = 0; i < 3; ++i) vec[i] = (double)h->vec[i];
            RTC8X_REQUIRE(invert3(inv, mat), "rtc8x_nanovdb_encode: the index-to-world map is singular");
        }
        uint8_t* map = b + 296;
        for (int i = 0; i < 9; ++i) {
            wr<float>(map + 4 * i, (float)mat[i]);
            wr<float>(map + 36 + 4 * i, h->inv_mat[i]);
            wr<double>(map + 88 + 8 * i, mat[i]);
            wr<double>(map + 160 + 8 * i, inv[i]);
        }
        for (int i = 0; i < 3; ++i) {
            wr<float>(map + 72 + 4 * i, h->vec[i]);
            wr<double>(map + 232 + 8 * i, vec[i]);
        }
        wr<float>(map + 84, 1.0f);
        wr<double>(map + 256, 1.0);
        // world bbox from the index bbox, max + 1, over its 8 corners (util/GridStats.h:536-562)
        if (!total_stats.empty()) {
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
            for (int c = 0; c < 8; ++c) {
                const double x[3] = {(double)(c & 4 ? (int64_t)total_stats.bmax[0] + 1 : total_stats.bmin[0]),
                                     (double)(c & 2 ? (int64_t)total_stats.bmax[1] + 1 : total_stats.bmin[1]),
                                     (double)(c & 1 ? (int64_t)total_stats.bmax[2] + 1 : total_stats.bmin[2])};
                double w[3];
                apply_map(mat, vec, x, w);
                for (int a = 0; a < 3; ++a) {
                    lo[a] = std::min(lo[a], w[a]);
                    hi[a] = std::max(hi[a], w[a]);
                }
            }
            for (int a = 0; a < 3; ++a) {
                wr<double>(b + 560 + 8 * a, lo[a]);
                wr<double>(b + 584 + 8 * a, hi[a]);
            }
            flags |= kFlagHasBBox;
        } else {
            // empty BBox<Vec3d>: min = +max double, max = -max double
            for (int a = 0; a < 3; ++a) {
                wr<double>(b + 560 + 8 * a, DBL_MAX);
                wr<double>(b + 584 + 8 * a, -DBL_MAX);
            }
        }
        wr<uint32_t>(b + 20, flags);
        const double one[3] = {1.0, 1.0, 1.0}, zero[3] = {0.0, 0.0, 0.0};
        double w1[3], w0[3];
        apply_map(mat, vec, one, w1);
        apply_map(mat, vec, zero, w0);
        for (int a = 0; a < 3; ++a) wr<double>(b + 608 + 8 * a, w1[a] - w0[a]);
        wr<uint32_t>(b + 632, kGridClassFogVolume);
        wr<uint32_t>(b + 636, kGridTypeFloat);
        // checksum, ChecksumMode::Partial (util/GridChecksum.h:175-203): the standard CRC-32 of the grid,
        // tree and root records past the checksum field in the low word; the high word stays empty
        const uint64_t head_end = root + kRootBytes + (uint64_t)h->num_root_tiles * kTileBytes;
        const uint32_t crc_head = (uint32_t)crc32(crc32(0L, Z_NULL, 0), b + 16, (uInt)(head_end - 16));
        wr<uint64_t>(b + 8, (uint64_t)crc_head | (0xFFFFFFFFull << 32));
        return RTC8X_OK;
    }
};

}  // namespace

rtc8x_status adopt_host_grid(OwnedHostGrid&& t, const rtc8x_grid& g0, rtc8x_grid* out)
{
    RTC8X_REQUIRE(!t.root.empty(), "host grid has no root tiles");
    auto owned_grid = std::make_unique<OwnedHostGrid>(std::move(t));
    OwnedHostGrid& o = *owned_grid;
    rtc8x_grid g = g0;
    g.root_tiles = o.root.data();
    g.upper_tables = o.upper.empty() ? nullptr : o.upper.data();
    g.lower_tables = o.lower.empty() ? nullptr : o.lower.data();
    g.leaf_values = o.values.empty() ? nullptr : o.values.data();
    g.tile_values = o.tiles.data();
    g.leaf_origins = o.origins.empty() ? nullptr : o.origins.data();
    g.cell_majorants = o.cells.empty() ? nullptr : o.cells.data();
    if (o.cells.empty())
        for (int a = 0; a < 3; ++a) g.cell_dims[a] = 0;
    g.region_majorants = o.regions.empty() ? nullptr : o.regions.data();
    if (o.regions.empty()) {
        for (int a = 0; a < 3; ++a) g.region_dims[a] = 0;
        g.region_cells = 0;
        g.region_unit = 0.0f;
    }
    g.leaf_map = nullptr;
    g.leaf_masks = o.leaf_masks.empty() ? nullptr : o.leaf_masks.data();
    g.lower_masks = o.lower_masks.empty() ? nullptr : o.lower_masks.data();
    g.upper_masks = o.upper_masks.empty() ? nullptr : o.upper_masks.data();
    {
        std::lock_guard<std::mutex> lock(g_owned_mutex);
        owned()[g.root_tiles] = std::move(owned_grid);
    }
    *out = g;
    return RTC8X_OK;
}

}  // namespace rtc8x

using namespace rtc8x;

extern "C" rtc8x_status rtc8x_grid_free_host(rtc8x_grid* g)
try {
    RTC8X_REQUIRE(g != nullptr, "rtc8x_grid_free_host: grid is NULL");
    {
        std::lock_guard<std::mutex> lock(g_owned_mutex);
        auto it = owned().find(g->root_tiles);
        RTC8X_REQUIRE(it != owned().end(), "rtc8x_grid_free_host: the grid's tables were not allocated by the library");
        owned().erase(it);
    }
    std::memset(g, 0, sizeof(*g));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_grid_free_host")

extern "C" rtc8x_status rtc8x_nanovdb_decode(const void* buf, size_t size, rtc8x_grid* out)
try {
    return decode_buffer(buf, size, out);
}
RTC8X_CATCH_ALL("rtc8x_nanovdb_decode")

extern "C" rtc8x_status rtc8x_nvdb_list(const void* file, size_t size, rtc8x_nvdb_grid_info* infos, uint32_t capacity,
                                        uint32_t* count)
try {
    RTC8X_REQUIRE(file && count, "rtc8x_nvdb_list: NULL argument");
    RTC8X_REQUIRE(capacity == 0 || infos, "rtc8x_nvdb_list: NULL infos with capacity %u", capacity);
    std::vector<rtc8x_nvdb_grid_info> v;
    rtc8x_status st = list_file((const uint8_t*)file, size, v);
    if (st != RTC8X_OK) return st;
    for (size_t i = 0; i < v.size() && i < capacity; ++i) infos[i] = v[i];
    *count = (uint32_t)v.size();
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nvdb_list")

extern "C" rtc8x_status rtc8x_nvdb_read_grid(const void* file, size_t size, uint32_t n, rtc8x_grid* out)
try {
    RTC8X_REQUIRE(out, "rtc8x_nvdb_read_grid: NULL output");
    std::vector<uint8_t> buf;
    const uint8_t* data = nullptr;
    uint64_t grid_size = 0;
    rtc8x_status st = read_grid(file, size, n, buf, &data, &grid_size);
    if (st != RTC8X_OK) return st;
    return decode_buffer(data, grid_size, out);
}
RTC8X_CATCH_ALL("rtc8x_nvdb_read_grid")

extern "C" rtc8x_status rtc8x_grid_upload_nanovdb(const void* buf, size_t size, rtc8x_grid* out)
try {
    RTC8X_REQUIRE(out, "rtc8x_grid_upload_nanovdb: NULL output");
    rtc8x_grid host;
    rtc8x_status st = decode_buffer(buf, size, &host);
    if (st != RTC8X_OK) return st;
    st = upload_host_grid(&host, out, true);
    rtc8x_grid_free_host(&host);
    return st;
}
RTC8X_CATCH_ALL("rtc8x_grid_upload_nanovdb")

extern "C" rtc8x_status rtc8x_grid_load_nvdb(const void* file, size_t size, uint32_t n, rtc8x_grid* out)
try {
    RTC8X_REQUIRE(out, "rtc8x_grid_load_nvdb: NULL output");
    rtc8x_grid host;
    rtc8x_status st = rtc8x_nvdb_read_grid(file, size, n, &host);
    if (st != RTC8X_OK) return st;
    st = upload_host_grid(&host, out, true);
    rtc8x_grid_free_host(&host);
    return st;
}
RTC8X_CATCH_ALL("rtc8x_grid_load_nvdb")

extern "C" rtc8x_status rtc8x_nanovdb_encode(const rtc8x_grid* h, const char* name, void* out, size_t capacity,
                                             size_t* size)
try {
    RTC8X_REQUIRE(h && size, "rtc8x_nanovdb_encode: NULL argument");
    const char* nm = name ? name : "";
    RTC8X_REQUIRE(std::strlen(nm) < 256, "rtc8x_nanovdb_encode: grid name longer than 255 bytes");
    uint32_t num_tiles = 0;
    rtc8x_status st = validate_tables(h, &num_tiles);
    if (st != RTC8X_OK) return st;
    Encoder e;
    e.h = h;
    st = e.plan();
    if (st != RTC8X_OK) return st;
    *size = e.total;
    if (!out) return RTC8X_OK;
    RTC8X_REQUIRE(capacity >= e.total, "rtc8x_nanovdb_encode: capacity %zu is below the %llu bytes needed", capacity,
                  (unsigned long long)e.total);
    return e.write((uint8_t*)out, nm);
}
RTC8X_CATCH_ALL("rtc8x_nanovdb_encode")

extern "C" rtc8x_status rtc8x_nvdb_write(const void* const* grids, const size_t* sizes, uint32_t count, uint32_t codec,
                                         void* out, size_t capacity, size_t* size)
try {
    RTC8X_REQUIRE(grids && sizes && size && count > 0, "rtc8x_nvdb_write: NULL argument or no grids");
    RTC8X_REQUIRE(count <= 0xFFFFu, "rtc8x_nvdb_write: %u grids exceed a segment's 65535", count);
    RTC8X_REQUIRE(codec == RTC8X_NVDB_CODEC_NONE || codec == RTC8X_NVDB_CODEC_ZIP,
                  "rtc8x_nvdb_write: codec %u is not supported (NONE and ZIP are)", codec);
    // metadata records (util/IO.h:403-425) and the data blocks
    std::vector<std::vector<uint8_t>> blocks(count);
    uint64_t head = kHeaderBytes, body = 0;
    std::vector<std::vector<uint8_t>> metas(count);
    for (uint32_t i = 0; i < count; ++i) {
        const uint8_t* g = (const uint8_t*)grids[i];
        RTC8X_REQUIRE(g && sizes[i] >= kGridBytes + kTreeBytes + kRootBytes, "rtc8x_nvdb_write: grid %u is too small", i);
        RTC8X_REQUIRE(rd<uint64_t>(g) == kMagic, "rtc8x_nvdb_write: grid %u has a bad magic number", i);
        const uint64_t gsize = rd<uint64_t>(g + 32);
        RTC8X_REQUIRE(gsize <= sizes[i], "rtc8x_nvdb_write: grid %u records %llu bytes but %zu are given", i,
                      (unsigned long long)gsize, sizes[i]);
        const char* name = (const char*)g + 40;
        const uint32_t name_size = (uint32_t)strnlen(name, 255) + 1;
        uint64_t fsize;
        if (codec == RTC8X_NVDB_CODEC_ZIP) {
            uLongf zb = compressBound((uLong)gsize);
            if (out) {
                blocks[i].resize(8 + zb);
                const int zs = compress(blocks[i].data() + 8, &zb, g, (uLong)gsize);
                RTC8X_REQUIRE(zs == Z_OK, "rtc8x_nvdb_write: zlib compress failed (%d)", zs);
                wr<uint64_t>(blocks[i].data(), (uint64_t)zb);
                blocks[i].resize(8 + zb);
            }
            fsize = 8 + zb;
        } else {
            fsize = gsize;
        }
        const uint8_t* tree = g + kGridBytes;
        const uint64_t root = kGridBytes + rd<uint64_t>(tree + 24);
        RTC8X_REQUIRE(root + kRootBytes <= gsize, "rtc8x_nvdb_write: grid %u has an invalid root offset", i);
        std::vector<uint8_t>& m = metas[i];
        m.assign(kMetaBytes + name_size, 0);
        wr<uint64_t>(m.data(), gsize);
        wr<uint64_t>(m.data() + 8, fsize);
        wr<uint64_t>(m.data() + 16, string_hash(name));
        wr<uint64_t>(m.data() + 24, rd<uint64_t>(tree + 56));
        wr<uint32_t>(m.data() + 32, rd<uint32_t>(g + 636));
        wr<uint32_t>(m.data() + 36, rd<uint32_t>(g + 632));
        std::memcpy(m.data() + 40, g + 560, 48);           // world bbox
        std::memcpy(m.data() + 88, g + root, 24);          // index bbox (the root's)
        std::memcpy(m.data() + 112, g + 608, 24);          // voxel size
        wr<uint32_t>(m.data() + 136, name_size);
        std::memcpy(m.data() + 140, tree + 32, 12);        // node counts
        wr<uint32_t>(m.data() + 152, 1u);                  // one root
        std::memcpy(m.data() + 156, tree + 44, 12);        // tile counts
        wr<uint16_t>(m.data() + 168, (uint16_t)codec);
        wr<uint32_t>(m.data() + 172, kVersion);
        std::memcpy(m.data() + kMetaBytes, name, name_size - 1);
        head += m.size();
        body += fsize;
    }
    *size = head + body;
    if (!out) return RTC8X_OK;
    RTC8X_REQUIRE(capacity >= *size, "rtc8x_nvdb_write: capacity %zu is below the %zu bytes needed", capacity, *size);
    uint8_t* o = (uint8_t*)out;
    wr<uint64_t>(o, kMagic);
    wr<uint32_t>(o + 8, kVersion);
    wr<uint16_t>(o + 12, (uint16_t)count);
    wr<uint16_t>(o + 14, (uint16_t)codec);
    uint64_t off = kHeaderBytes;
    for (const auto& m : metas) {
        std::memcpy(o + off, m.data(), m.size());
        off += m.size();
    }
    for (uint32_t i = 0; i < count; ++i) {
        if (codec == RTC8X_NVDB_CODEC_ZIP) {
            std::memcpy(o + off, blocks[i].data(), blocks[i].size());
            off += blocks[i].size();
        } else {
            const uint64_t gsize = rd<uint64_t>((const uint8_t*)grids[i] + 32);
            std::memcpy(o + off, grids[i], gsize);
            off += gsize;
        }
    }
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_nvdb_write")

// nvdb_probe.cpp -- NanoVDB ground truth for the .nvdb codec tests (test-side checker, built only where
// the rtcamp8 tree is mounted; see oracle/ref/Makefile).  Uses the NanoVDB that rtcamp8 vendors
// (ext/nanovdb, v32.3.3) with its ZIP codec enabled.
//
//   nvdb_probe write <dir>
//     Builds two FloatGrids and one DoubleGrid with nanovdb::GridBuilder and writes them with
//     nanovdb::io::writeGrids as <dir>/grids_none.nvdb (Codec::NONE) and <dir>/grids_zip.nvdb
//     (Codec::ZIP).  For each float grid it writes <dir>/grid<k>.bin:
//       f32 world_bbox[6] (as the reference keeps it), i32 index_bbox[6], f32 majorant (the reference's
//       loop, renderer/renderer_scene.cpp:497-511), f32 background, u32 node_count[3],
//       u32 n, n x { i32 ijk[3], f32 value }            (ReadAccessor::getValue)
//       u32 m, m x { f32 world[3], f32 value }          (createSampler<1,ReadAccessor,false> at worldToIndexF)
//   nvdb_probe check <file.nvdb> <in.bin> <out.bin>
//     Reads grid 0 of a file written by another program and reports what NanoVDB makes of it:
//     in.bin holds u32 n, n x i32 ijk[3], u32 m, m x f32 world[3]; out.bin receives
//       i32 valid (GridValidator, detailed), i32 checksum_ok (Partial, GridBuilder's default), u32 stats_bytes_changed
//       (bytes that nanovdb::gridStats(StatsMode::All) rewrites on a copy), u32 first_changed_offset,
//       f32 majorant (the reference's loop), then n getValue results and m trilinear samples.
#define NANOVDB_USE_ZIP
#include <nanovdb/NanoVDB.h>
#include <nanovdb/util/GridBuilder.h>
#include <nanovdb/util/GridChecksum.h>
#include <nanovdb/util/GridStats.h>
#include <nanovdb/util/GridValidator.h>
#include <nanovdb/util/IO.h>
#include <nanovdb/util/SampleFromVoxels.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

static uint32_t mix(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// The majorant exactly as Scene::allocateVolumeGrid computes it.
static float reference_majorant(const nanovdb::FloatGrid* grid)
{
    const nanovdb::CoordBBox bb = grid->indexBBox();
    const auto& acc = grid->tree().getAccessor();
    float m = -INFINITY;
    for (int ix = bb.min()[0]; ix < bb.max()[0]; ++ix)
        for (int iy = bb.min()[1]; iy < bb.max()[1]; ++iy)
            for (int iz = bb.min()[2]; iz < bb.max()[2]; ++iz) m = std::max(acc.getValue({ix, iy, iz}), m);
    return m;
}

static void put(FILE* f, const void* p, size_t n) { std::fwrite(p, 1, n, f); }

static int write_mode(const std::string& dir)
{
    std::vector<nanovdb::GridHandle<>> handles;
    // grid 0: sparse leaves around the origin (negative coordinates too), about half the voxels active
    {
        nanovdb::GridBuilder<float> b(0.0f, nanovdb::GridClass::FogVolume);
        auto acc = b.getAccessor();
        for (int x = -16; x < 24; x += 8)
            for (int y = 0; y < 16; y += 8)
                for (int z = -8; z < 16; z += 8) {
                    if ((mix((uint32_t)(x * 31 + y * 1009 + z * 7919 + 12345)) & 3u) == 0u) continue;
                    for (int v = 0; v < 512; ++v) {
                        uint32_t h = mix((uint32_t)((x + 64) * 4099 + (y + 64) * 131 + (z + 64)) * 512u + (uint32_t)v);
                        if (h & 1u) acc.setValue(nanovdb::Coord(x + (v >> 6), y + ((v >> 3) & 7), z + (v & 7)),
                                                 0.02f + (float)(h >> 24) * (1.0f / 256.0f));
                    }
                }
        handles.push_back(b.getHandle<>(1.0 / 32.0, nanovdb::Vec3d(0.0), "density"));
    }
    // grid 1: a non-zero background, a translated map with a non-power-of-two voxel size, a leaf in a
    // second root tile, and its maximum voxel on the index bbox's upper face
    {
        nanovdb::GridBuilder<float> b(0.25f, nanovdb::GridClass::FogVolume);
        auto acc = b.getAccessor();
        for (int x = 0; x < 24; ++x)
            for (int y = 0; y < 16; ++y)
                for (int z = 0; z < 20; ++z) {
                    uint32_t h = mix((uint32_t)(x * 977 + y * 31 + z) ^ 0x5bd1e995u);
                    if ((h & 7u) != 0u) acc.setValue(nanovdb::Coord(x, y, z), (float)(h >> 22) * (1.0f / 1024.0f));
                }
        acc.setValue(nanovdb::Coord(23, 15, 19), 7.5f);  // on the max face: excluded by the reference's loop
        acc.setValue(nanovdb::Coord(4100, 3, 3), 2.0f);  // second root tile
        handles.push_back(b.getHandle<>(0.05, nanovdb::Vec3d(-0.3, 0.1, 0.2), "shifted"));
    }
    // grid 2: not a FloatGrid
    {
        nanovdb::GridBuilder<double> b(0.0);
        auto acc = b.getAccessor();
        acc.setValue(nanovdb::Coord(1, 2, 3), 1.0);
        handles.push_back(b.getHandle<>(1.0, nanovdb::Vec3d(0.0), "doubles"));
    }
    nanovdb::io::writeGrids(dir + "/grids_none.nvdb", handles, nanovdb::io::Codec::NONE);
    nanovdb::io::writeGrids(dir + "/grids_zip.nvdb", handles, nanovdb::io::Codec::ZIP);

    std::mt19937_64 gen(7);
    for (int k = 0; k < 2; ++k) {
        const nanovdb::FloatGrid* grid = handles[k].grid<float>();
        FILE* f = std::fopen((dir + "/grid" + std::to_string(k) + ".bin").c_str(), "wb");
        if (!f) return 1;
        auto wb = grid->worldBBox();
        float bb[6] = {(float)wb.min()[0], (float)wb.min()[1], (float)wb.min()[2],
                       (float)wb.max()[0], (float)wb.max()[1], (float)wb.max()[2]};
        put(f, bb, sizeof(bb));
        const nanovdb::CoordBBox ib = grid->indexBBox();
        int32_t ibb[6] = {ib.min()[0], ib.min()[1], ib.min()[2], ib.max()[0], ib.max()[1], ib.max()[2]};
        put(f, ibb, sizeof(ibb));
        float maj = reference_majorant(grid);
        put(f, &maj, 4);
        float bg = grid->tree().background();
        put(f, &bg, 4);
        uint32_t nc[3] = {grid->tree().nodeCount(0), grid->tree().nodeCount(1), grid->tree().nodeCount(2)};
        put(f, nc, sizeof(nc));
        const auto& acc = grid->tree().getAccessor();
        std::uniform_int_distribution<int> dx(ib.min()[0] - 10, std::min(ib.max()[0], ib.min()[0] + 80) + 10),
            dy(ib.min()[1] - 10, ib.max()[1] + 10), dz(ib.min()[2] - 10, ib.max()[2] + 10);
        const uint32_t n = 20000;
        put(f, &n, 4);
        for (uint32_t i = 0; i < n; ++i) {
            int32_t c[3] = {dx(gen), dy(gen), dz(gen)};
            if (k == 1 && i % 10 == 0) c[0] = 4096 + (int)(mix(i) % 16u);
            float v = acc.getValue(nanovdb::Coord(c[0], c[1], c[2]));
            put(f, c, sizeof(c));
            put(f, &v, 4);
        }
        using AccT = nanovdb::DefaultReadAccessor<float>;
        const AccT racc = grid->getAccessor();
        const auto sampler = nanovdb::createSampler<1, AccT, false>(racc);
        std::uniform_real_distribution<float> u(0.0f, 1.0f);
        const uint32_t m = 5000;
        put(f, &m, 4);
        for (uint32_t i = 0; i < m; ++i) {
            nanovdb::Vec3f w;
            for (int a = 0; a < 3; ++a) w[a] = (float)(wb.min()[a] + (wb.max()[a] - wb.min()[a]) * (1.2 * u(gen) - 0.1));
            float v = sampler(grid->worldToIndexF(w));
            float wf[3] = {w[0], w[1], w[2]};
            put(f, wf, sizeof(wf));
            put(f, &v, 4);
        }
        std::fclose(f);
    }
    std::printf("nvdb_probe: wrote %zu grids\n", handles.size());
    return 0;
}

static int check_mode(const char* file, const char* in, const char* out)
{
    auto handle = nanovdb::io::readGrid(file, 0);
    nanovdb::FloatGrid* grid = handle.grid<float>();
    if (!grid) {
        std::fprintf(stderr, "nvdb_probe: grid 0 is not a FloatGrid\n");
        return 1;
    }
    int32_t valid = nanovdb::isValid(*grid, true, true) ? 1 : 0;
    int32_t checksum_ok = nanovdb::validateChecksum(*grid, nanovdb::ChecksumMode::Partial) &&
                                  nanovdb::GridChecksum(grid->checksum()).mode() == nanovdb::ChecksumMode::Partial
                              ? 1
                              : 0;
    std::vector<uint8_t> copy(handle.size());
    std::memcpy(copy.data(), handle.data(), handle.size());
    auto* g2 = reinterpret_cast<nanovdb::FloatGrid*>(copy.data());
    nanovdb::gridStats(*g2, nanovdb::StatsMode::All);
    uint32_t changed = 0, first = 0xFFFFFFFFu;
    const uint8_t* a = (const uint8_t*)handle.data();
    for (size_t i = 0; i < copy.size(); ++i)
        if (copy[i] != a[i]) {
            if (first == 0xFFFFFFFFu) first = (uint32_t)i;
            ++changed;
        }
    float maj = reference_majorant(grid);

    FILE* fi = std::fopen(in, "rb");
    FILE* fo = std::fopen(out, "wb");
    if (!fi || !fo) return 1;
    put(fo, &valid, 4);
    put(fo, &checksum_ok, 4);
    put(fo, &changed, 4);
    put(fo, &first, 4);
    put(fo, &maj, 4);
    const auto& acc = grid->tree().getAccessor();
    uint32_t n = 0;
    if (std::fread(&n, 4, 1, fi) != 1) return 1;
    for (uint32_t i = 0; i < n; ++i) {
        int32_t c[3];
        if (std::fread(c, 4, 3, fi) != 3) return 1;
        float v = acc.getValue(nanovdb::Coord(c[0], c[1], c[2]));
        put(fo, &v, 4);
    }
    using AccT = nanovdb::DefaultReadAccessor<float>;
    const AccT racc = grid->getAccessor();
    const auto sampler = nanovdb::createSampler<1, AccT, false>(racc);
    uint32_t m = 0;
    if (std::fread(&m, 4, 1, fi) != 1) return 1;
    for (uint32_t i = 0; i < m; ++i) {
        float w[3];
        if (std::fread(w, 4, 3, fi) != 3) return 1;
        float v = sampler(grid->worldToIndexF(nanovdb::Vec3f(w[0], w[1], w[2])));
        put(fo, &v, 4);
    }
    std::fclose(fi);
    std::fclose(fo);
    std::printf("nvdb_probe: valid %d checksum %d stats-changed bytes %u (first at %u)\n", valid, checksum_ok, changed,
                first);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc >= 3 && std::strcmp(argv[1], "write") == 0) return write_mode(argv[2]);
    if (argc >= 5 && std::strcmp(argv[1], "check") == 0) return check_mode(argv[2], argv[3], argv[4]);
    std::fprintf(stderr, "usage: nvdb_probe write <dir> | check <file.nvdb> <in.bin> <out.bin>\n");
    return 1;
}

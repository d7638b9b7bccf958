// nanovdb_probe.cpp -- writes NanoVDB ground truth for the grid-indexing parity tests (test-side
// checker, built only where the rtcamp8 tree is mounted; see oracle/ref/Makefile).
//
// It builds a small sparse FloatGrid with the NanoVDB that rtcamp8 vendors (ext/nanovdb, v32.3.3),
// with leaves spread over several root tiles (negative coordinates and coordinates beyond 4096
// included), and records for random world points what the reference's lookups compute:
// Grid::worldToIndexF, the Floor coordinate, the uncached trilinear sampler
// createSampler<1, ReadAccessor, false> (the one optix_pathtracing_kernels.cu:103 uses), the upper /
// lower / leaf table offsets, plus Ray<float>::clip against the grid's world bbox
// (optix_pathtracing_kernels.cu:106-111).
//
// Output (little-endian binary, argv[1]):
//   u32 magic 'NVPB', u32 num_leaves, f64 voxel_size, f32 world_bbox[6]
//   num_leaves x { i32 origin[3], f32 values[512] }
//   u32 num_points, num_points x { f32 world[3], f32 index[3], i32 ijk[3], u32 offsets[3], f32 value }
//   u32 num_rays, num_rays x { f32 eye[3], f32 dir[3], f32 tmax, i32 hit, f32 t0, f32 t1 }
#include <nanovdb/NanoVDB.h>
#include <nanovdb/util/GridBuilder.h>
#include <nanovdb/util/Ray.h>
#include <nanovdb/util/SampleFromVoxels.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
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

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: nanovdb_probe out.bin\n");
        return 1;
    }
    const double voxel = 1.0 / 64.0;
    std::vector<nanovdb::Coord> origins;
    // a dense-ish cluster of leaves near the origin (several lower nodes), some gaps
    for (int x = -16; x < 72; x += 8)
        for (int y = -8; y < 64; y += 8)
            for (int z = 0; z < 64; z += 8)
                if ((mix((uint32_t)(x * 7 + y * 131 + z * 1031 + 99991)) & 3u) != 0u) origins.emplace_back(x, y, z);
    // leaves in other root tiles (beyond 4096 and below -4096 voxels)
    origins.emplace_back(4096, 0, 8);
    origins.emplace_back(4104, 8, 8);
    origins.emplace_back(-4104, -8, 16);
    origins.emplace_back(8, 4096 + 128, -4096);

    nanovdb::GridBuilder<float> builder(0.0f);
    auto acc = builder.getAccessor();
    std::vector<float> values;
    for (size_t l = 0; l < origins.size(); ++l) {
        for (int v = 0; v < 512; ++v) {
            nanovdb::Coord ijk(origins[l][0] + (v >> 6), origins[l][1] + ((v >> 3) & 7), origins[l][2] + (v & 7));
            uint32_t h = mix((uint32_t)(l * 512 + v) ^ 0xA5A5A5A5u);
            float val = 0.05f + (float)(h >> 8) * (1.0f / 16777216.0f);  // never the background value
            acc.setValue(ijk, val);
            values.push_back(val);
        }
    }
    auto handle = builder.getHandle<>(voxel, nanovdb::Vec3d(0.0), "probe");
    const nanovdb::FloatGrid* grid = handle.grid<float>();
    if (!grid) {
        std::fprintf(stderr, "grid build failed\n");
        return 1;
    }
    auto wb = grid->worldBBox();
    nanovdb::BBox<nanovdb::Vec3f> fbbox(nanovdb::Vec3f(wb.min()[0], wb.min()[1], wb.min()[2]),
                                        nanovdb::Vec3f(wb.max()[0], wb.max()[1], wb.max()[2]));

    FILE* f = std::fopen(argv[1], "wb");
    if (!f) return 1;
    const uint32_t magic = 0x4250564e, nl = (uint32_t)origins.size();
    std::fwrite(&magic, 4, 1, f);
    std::fwrite(&nl, 4, 1, f);
    std::fwrite(&voxel, 8, 1, f);
    float bb[6] = {fbbox.min()[0], fbbox.min()[1], fbbox.min()[2], fbbox.max()[0], fbbox.max()[1], fbbox.max()[2]};
    std::fwrite(bb, 4, 6, f);
    for (size_t l = 0; l < origins.size(); ++l) {
        int32_t o[3] = {origins[l][0], origins[l][1], origins[l][2]};
        std::fwrite(o, 4, 3, f);
        std::fwrite(&values[l * 512], 4, 512, f);
    }

    using AccT = nanovdb::DefaultReadAccessor<float>;
    const AccT racc = grid->getAccessor();
    const auto sampler = nanovdb::createSampler<1, AccT, false>(racc);
    std::mt19937_64 gen(20221202);
    std::uniform_real_distribution<float> ux(-0.4f, 1.25f), uy(-0.25f, 1.1f), uz(-0.1f, 1.1f), u01(0.0f, 1.0f);
    const uint32_t np = 3000;
    std::fwrite(&np, 4, 1, f);
    for (uint32_t i = 0; i < np; ++i) {
        nanovdb::Vec3f w(ux(gen), uy(gen), uz(gen));
        if (i < 500) {  // exact voxel corners and near-integer index positions
            w = nanovdb::Vec3f(std::floor(w[0] * 64.0f) / 64.0f, std::floor(w[1] * 64.0f) / 64.0f, w[2]);
        } else if (i < 800) {  // points around the far root tiles
            w = nanovdb::Vec3f(4096.0f / 64.0f + u01(gen) * 0.4f, u01(gen) * 0.3f, u01(gen) * 0.4f);
        } else if (i < 1000) {
            w = nanovdb::Vec3f(-4104.0f / 64.0f + u01(gen) * 0.2f, -0.2f + u01(gen) * 0.2f, 0.2f + u01(gen) * 0.2f);
        }
        nanovdb::Vec3f idx = grid->worldToIndexF(w);
        nanovdb::Vec3f frac = idx;
        nanovdb::Coord ijk = nanovdb::Floor<nanovdb::Coord>(frac);
        float value = sampler(idx);
        uint32_t offs[3] = {nanovdb::NanoUpper<float>::CoordToOffset(ijk), nanovdb::NanoLower<float>::CoordToOffset(ijk),
                            nanovdb::NanoLeaf<float>::CoordToOffset(ijk)};
        float wf[3] = {w[0], w[1], w[2]}, xf[3] = {idx[0], idx[1], idx[2]};
        int32_t c[3] = {ijk[0], ijk[1], ijk[2]};
        std::fwrite(wf, 4, 3, f);
        std::fwrite(xf, 4, 3, f);
        std::fwrite(c, 4, 3, f);
        std::fwrite(offs, 4, 3, f);
        std::fwrite(&value, 4, 1, f);
    }
    const uint32_t nr = 2000;
    std::fwrite(&nr, 4, 1, f);
    std::uniform_real_distribution<float> ue(-2.0f, 3.0f), ud(-1.0f, 1.0f);
    for (uint32_t i = 0; i < nr; ++i) {
        nanovdb::Vec3f eye(ue(gen), ue(gen), ue(gen));
        nanovdb::Vec3f dir(ud(gen), ud(gen), ud(gen));
        if (i % 5 == 0) dir[i % 3] = 0.0f;  // axis-parallel rays exercise the infinite inverse direction
        float len = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        if (len == 0.0f) dir = nanovdb::Vec3f(0.0f, 1.0f, 0.0f), len = 1.0f;
        dir = nanovdb::Vec3f(dir[0] / len, dir[1] / len, dir[2] / len);
        float tmax = (i % 3 == 0) ? 1e10f : 0.5f + 4.0f * u01(gen);
        nanovdb::Ray<float> ray(eye, dir, 0.0f, tmax);
        int32_t hit = ray.clip(fbbox) ? 1 : 0;
        float t0 = ray.t0(), t1 = ray.t1();
        float e[3] = {eye[0], eye[1], eye[2]}, d[3] = {dir[0], dir[1], dir[2]};
        std::fwrite(e, 4, 3, f);
        std::fwrite(d, 4, 3, f);
        std::fwrite(&tmax, 4, 1, f);
        std::fwrite(&hit, 4, 1, f);
        std::fwrite(&t0, 4, 1, f);
        std::fwrite(&t1, 4, 1, f);
    }
    std::fclose(f);
    std::printf("nanovdb_probe: %u leaves, %u points, %u rays; world bbox (%g %g %g)-(%g %g %g)\n", nl, np, nr, bb[0],
                bb[1], bb[2], bb[3], bb[4], bb[5]);
    return 0;
}

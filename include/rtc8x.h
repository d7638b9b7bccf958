/*
 * rtc8x.h -- C ABI of the MI355X (gfx950) volumetric path tracer + neural radiance cache.
 *
 * This is the drop-in boundary for the hot path of rtcamp8's renderer:
 *   - the three OptiX ray-generation entry points (renderer/gpu_kernels/optix_pathtracing_kernels.cu:438-448)
 *     with their launch-parameter blocks (renderer/renderer_shared.h:113-172),
 *   - the NRC data-path kernels (renderer/gpu_kernels/nrc_set_up_kernels.cu),
 *   - the NeuralRadianceCache class (renderer/network_interface.h:17-31),
 *   - the volume upload / majorant (renderer/renderer_scene.cpp:491-545),
 *   - the sky bake and environment importance sampling (renderer/gpu_kernels/analytic_sky_kernels.cu,
 *     renderer/gpu_kernels/compute_light_probs.cu).
 *
 * Conventions
 *   - Every entry point returns rtc8x_status; nothing throws across the ABI.  rtc8x_last_error()
 *     returns a message for the calling thread's last failure.
 *   - Pointers inside the parameter structs are device pointers unless stated otherwise.  The caller
 *     owns every buffer it passes in (as in the reference, renderer/renderer_main.cpp:80-230); objects
 *     created by rtc8x_*_create / rtc8x_grid_build_* are owned by the library until destroyed.
 *   - Streams are hipStream_t passed as void*; NULL is the default stream.  Work is stream-ordered and
 *     asynchronous unless the function says it synchronises.
 *   - One GPU per process; the library uses the device that is current on the calling thread.
 */
#ifndef RTC8X_H
#define RTC8X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTC8X_ABI_VERSION 1

typedef enum rtc8x_status {
    RTC8X_OK = 0,
    RTC8X_ERR_INVALID_ARGUMENT = 1,
    RTC8X_ERR_HIP = 2,
    RTC8X_ERR_OUT_OF_MEMORY = 3,
    RTC8X_ERR_NOT_SUPPORTED = 4
} rtc8x_status;

/* ------------------------------------------------------------------------------------------------
 * Library, memory, streams and timers.
 * Stand-ins for the CUDA driver wrapper the reference uses (common/utils/cuda_util.h:175-300:
 * cudau::Buffer, cudau::Timer) so a host program can drive the path without any other GPU runtime.
 * ---------------------------------------------------------------------------------------------- */
int32_t rtc8x_abi_version(void);
const char* rtc8x_last_error(void);
/* sizeof of an ABI struct by its name ("rtc8x_grid", ...), 0 if unknown: lets bindings check layouts. */
size_t rtc8x_sizeof(const char* type_name);
rtc8x_status rtc8x_device_count(int32_t* count);
rtc8x_status rtc8x_set_device(int32_t device);
rtc8x_status rtc8x_device_name(char* buf, size_t buf_len);
rtc8x_status rtc8x_malloc(void** ptr, size_t bytes);
rtc8x_status rtc8x_free(void* ptr);
rtc8x_status rtc8x_memcpy_htod(void* dst, const void* src, size_t bytes);
rtc8x_status rtc8x_memcpy_dtoh(void* dst, const void* src, size_t bytes);
rtc8x_status rtc8x_memcpy_dtod(void* dst, const void* src, size_t bytes);
rtc8x_status rtc8x_memset(void* dst, int32_t value, size_t bytes);
rtc8x_status rtc8x_device_synchronize(void);
rtc8x_status rtc8x_stream_create(void** stream);
rtc8x_status rtc8x_stream_destroy(void* stream);
rtc8x_status rtc8x_stream_synchronize(void* stream);
/* HIP event pairs, as cudau::Timer (common/utils/cuda_util.h:257-300). */
rtc8x_status rtc8x_event_create(void** event);
rtc8x_status rtc8x_event_destroy(void* event);
rtc8x_status rtc8x_event_record(void* event, void* stream);
rtc8x_status rtc8x_event_elapsed_ms(float* ms, void* start, void* end);
/* Work queued on `stream` after this call waits for the work before the event's last record
 * (hipStreamWaitEvent; cudaStreamWaitEvent in CUDA).  The NRC renderer's host code uses it to trace the
 * next launch on a second stream while the network trains; no counterpart in the reference, whose
 * renderer runs one stream. */
rtc8x_status rtc8x_stream_wait_event(void* stream, void* event);

/* Measurement kernels (no counterpart in the reference).
 * rtc8x_copy_async: device-to-device copy by a plain 16-byte-per-lane kernel on `stream` (dst, src and
 *   bytes 16-byte aligned).  It reads and writes exactly `bytes` each, which calibrates HBM byte counters.
 * rtc8x_gather_probe: the memory-latency roofline of density lookups.  `threads` lanes (a multiple of
 *   256) each perform `steps` serially dependent steps of `width` (1, 2, 4) lookups issued together;
 *   a lookup is the tracer's trilinear lookup (leaf-map entry, then the 2x2x2 stencil's voxels) at a
 *   random position inside a random leaf of the device grid `grid`.  Occupancy is pinned to
 *   `waves_per_simd` (2..8) waves per SIMD through an LDS reservation.  out[threads] receives a per-lane
 *   checksum. */
struct rtc8x_grid;
rtc8x_status rtc8x_copy_async(void* dst, const void* src, size_t bytes, void* stream);
rtc8x_status rtc8x_gather_probe(void* stream, const struct rtc8x_grid* grid, uint32_t threads, uint32_t steps,
                                uint32_t width, uint32_t waves_per_simd, uint32_t seed, float* out);

struct rtc8x_frame_params;
/*
 * rtc8x_march_probe: the memory-latency roofline of the tracer's own lookups (no counterpart in the
 * reference).  A persistent launch, `waves_per_simd` (2..4, set with dynamic LDS) 256-lane workgroups per
 * CU (registers allotted for 3 waves per SIMD, as the tracer's render entries have, up to 3, and for 4
 * at 4), whose lanes take the frame's pixels in 8x8 tiles from a work counter as the tracer's do (the
 * tiles wrapping around the image; a pixel's PCG state is read, never written) and trace `spp` paths per
 * pixel: jittered camera rays, delta tracking as params->tracking selects it (the global majorant's
 * spacing with two tentative collisions looked up together, or the local tracker's region walk over the
 * grid's region table staged in LDS), and at every real collision a ratio-tracked shadow segment towards
 * `light_dir` and a sample of the frame's phase function, with none of the tracer's vertex shading.
 * Every lane stops after `budget` lookups (<= 2^24), so the launch has no tail.  stats[0] += lookups
 * made, stats[1] += paths ended; stats is a device array of two uint64.  With local tracking the 32 KiB
 * region table counts towards the LDS that sets the occupancy.
 */
rtc8x_status rtc8x_march_probe(void* stream, const struct rtc8x_frame_params* params, const float light_dir[3],
                               uint32_t spp, uint32_t budget, uint32_t waves_per_simd, uint64_t* stats);

/* ------------------------------------------------------------------------------------------------
 * Sparse density grid in the NanoVDB tree layout (root tiles -> 32^3 upper -> 16^3 lower -> 8^3 leaf;
 * ext/nanovdb/nanovdb/NanoVDB.h).  Node and voxel offsets follow NanoVDB exactly
 * (InternalNode::CoordToOffset NanoVDB.h:3217-3227, LeafNode::CoordToOffset :3821-3828,
 *  RootData::CoordToKey :2699-2706); the in-memory tables are this library's own (32-bit child indices).
 * ---------------------------------------------------------------------------------------------- */
#define RTC8X_UPPER_TABLE_SIZE 32768u   /* 32^3 */
#define RTC8X_LOWER_TABLE_SIZE 4096u    /* 16^3 */
#define RTC8X_LEAF_SIZE 512u            /* 8^3  */
#define RTC8X_TILE_BIT 0x80000000u      /* table entry is a tile: low 31 bits index tile_values */

typedef struct rtc8x_root_tile {
    uint64_t key;        /* NanoVDB 64-bit root key of the upper node's origin */
    uint32_t entry;      /* upper-node index, or RTC8X_TILE_BIT | tile value index */
    uint32_t state;      /* a tile's active state (NanoVDB RootData::Tile::state); not used by the tracer */
} rtc8x_root_tile;

typedef struct rtc8x_grid {
    /* Map subset used by Grid::worldToIndexF -> Map::applyInverseMapF (NanoVDB.h:2029-2032, 2384):
     * index = inv_mat (row-major 3x3) * (world - vec). */
    float inv_mat[9];
    float vec[3];
    /* World-space bounding box used to clip rays (Scene::setVolumeGrid, renderer_scene.cpp:517-519). */
    float world_bbox_min[3];
    float world_bbox_max[3];
    int32_t index_bbox_min[3];    /* inclusive index bbox of the active voxels */
    int32_t index_bbox_max[3];
    float max_value;              /* brute-force max of all voxel values (renderer_scene.cpp:499-511) */
    float background;
    uint32_t num_root_tiles;
    uint32_t num_upper;
    uint32_t num_lower;
    uint32_t num_leaves;
    const rtc8x_root_tile* root_tiles;   /* num_root_tiles */
    const uint32_t* upper_tables;        /* num_upper * 32768 entries */
    const uint32_t* lower_tables;        /* num_lower * 4096 entries */
    /* Device grids (from rtc8x_grid_build_fbm / rtc8x_grid_upload*) keep one float of padding after
     * leaf_values and after tile_values: the tracer reads voxel pairs (z, z+1) without a bounds test. */
    const float* leaf_values;            /* num_leaves * 512 voxel values */
    const float* tile_values;            /* tile_values[0] == background */
    const int32_t* leaf_origins;         /* num_leaves * 3 index-space origins (optional, may be NULL) */
    /* Local-majorant acceleration (library-built, optional, used by RTC8X_TRACKING_LOCAL_MAJORANT
     * through the region table below): one float per 8^3 cell over the index bbox, cell c spanning
     * index [cell_origin + 8c, cell_origin + 8c + 8): the largest voxel value (leaves, tiles and
     * background as lookups see them, and at least 0) over [cell_origin + 8c - 1, cell_origin + 8c + 9]
     * per axis.  A point of the cell reads its trilinear taps from [8c, 8c + 8]; the rim of one voxel
     * more on each side covers a sample position that rounding puts just outside the cell.  NULL (and
     * no region table) when the index bbox spans more than 2^26 cells. */
    int32_t cell_origin[3];              /* index-space origin of cell (0,0,0), a multiple of 8 */
    int32_t cell_dims[3];
    const float* cell_majorants;         /* cell_dims[0]*cell_dims[1]*cell_dims[2], x-major */
    /* Dense leaf map (library-built, optional): for every 8^3 block of index space in
     * [leaf_map_origin, leaf_map_origin + 8 * leaf_map_dims) the result of the root -> upper -> lower
     * walk (leaf index, or RTC8X_TILE_BIT | tile value index), x-major.  It is a cache of the tree:
     * lookups outside it walk the tree and get the same values.  NULL when the index bbox needs more
     * than 2^26 blocks. */
    int32_t leaf_map_origin[3];          /* a multiple of 8 */
    int32_t leaf_map_dims[3];
    const uint32_t* leaf_map;
    /* Activity (host tables only, optional; lookups ignore it): NanoVDB value masks, one bit per voxel of
     * a leaf (8 words) or per table entry of a lower (64 words) / upper (512 words) node, bit n of word
     * n / 64 for offset n.  A set bit on an internal table entry marks an active tile.  Decoded grids
     * carry them; the encoder writes them back, and treats NULL as every leaf voxel active and every
     * tile inactive.  Never uploaded. */
    const uint64_t* leaf_masks;          /* num_leaves x 8 */
    const uint64_t* lower_masks;         /* num_lower x 64 */
    const uint64_t* upper_masks;         /* num_upper x 512 */
    /* The index-to-world map in double precision (NanoVDB Map::mMatD, mInvMatD, mVecD; host tables only):
     * decoded grids keep it so that an encoded copy holds the same map and world bbox.  All zero: the
     * encoder derives it from inv_mat and vec. */
    double map_mat_d[9];
    double map_inv_mat_d[9];
    double map_vec_d[3];
    /* Region majorants (library-built with the cell majorants, optional; used by
     * RTC8X_TRACKING_LOCAL_MAJORANT, which walks regions, not cells): the cell grid is grouped into regions
     * of region_cells^3 cells from cell_origin (region_cells a power of two: RTC8X_DEFAULT_REGION_CELLS,
     * doubled until region_dims[0] * region_dims[1] * region_dims[2] <= RTC8X_MAX_REGIONS so the tracer can
     * stage the table in LDS, or the size rtc8x_grid_build_regions was given).  One byte per region,
     * x-major: for a region whose largest cell majorant m is > 0, the
     * least q in 1..127 with (float)q * region_unit >= m, the region's majorant being (float)q *
     * region_unit; for an empty region, RTC8X_REGION_EMPTY + k, k (<= 127) the number of region layers
     * around it that are empty too (its Chebyshev distance to the nearest region with density, minus one;
     * space outside the region grid counts as empty), which lets the tracker skip them at once.
     * region_unit is the largest cell majorant / 127, raised to the next float until 127 * region_unit
     * covers it.  The table's allocation is padded with zeros to a multiple of 16 bytes. */
    int32_t region_dims[3];
    uint32_t region_cells;
    float region_unit;
    const uint8_t* region_majorants;
} rtc8x_grid;

#define RTC8X_MAX_REGIONS 32768
#define RTC8X_REGION_EMPTY 128u
#define RTC8X_DEFAULT_REGION_CELLS 4u  /* regions of 32^3 voxels */

/* Procedural fBm volume description (the benchmark's synthetic input; the reference loads a .nvdb). */
typedef struct rtc8x_fbm_desc {
    uint32_t resolution;        /* N: voxels [0,N)^3 spanning the world cube [0,1]^3 (N % 8 == 0) */
    uint32_t octaves;           /* 4 */
    uint64_t seed;
    float base_frequency;       /* lattice cells across the cube at octave 0 */
    float lacunarity;           /* 2 */
    float gain;                 /* 0.5 */
    int32_t sparse;             /* 0: dense (every leaf stored); 1: thresholded cloud */
    float active_leaf_fraction; /* sparse only: fraction of the N^3/512 leaf slots kept (0.25) */
    float contrast;             /* density = min(contrast * (g - threshold) / (g_max - threshold), 1) */
} rtc8x_fbm_desc;

/* Builds the fBm grid on the device.  The returned struct's device tables are owned by the library
 * and released with rtc8x_grid_destroy.  Synchronises the stream. */
rtc8x_status rtc8x_grid_build_fbm(const rtc8x_fbm_desc* desc, rtc8x_grid* out, void* stream);
/* Uploads host tables (same layout as rtc8x_grid, host pointers in `host`) to a library-owned copy,
 * computing max_value and the cell majorants.  Replaces Scene::allocateVolumeGrid
 * (renderer/renderer_scene.cpp:491-537). */
rtc8x_status rtc8x_grid_upload(const rtc8x_grid* host, rtc8x_grid* out);
rtc8x_status rtc8x_grid_destroy(rtc8x_grid* grid);
/* Rebuilds a device grid's region table from its cell majorants with regions of region_cells^3 cells
 * (a power of two; 0: the default choice).  RTC8X_ERR_INVALID_ARGUMENT when the grid has no cell
 * majorants or the size gives more than RTC8X_MAX_REGIONS regions.  Synchronises the device.  The local
 * tracker's cost trades region steps (fewer with larger regions) against density lookups (fewer with
 * tighter majorants); callers tune it per volume.  No counterpart in the reference, which tracks with
 * the global majorant only (renderer_scene.cpp:499-511). */
rtc8x_status rtc8x_grid_build_regions(rtc8x_grid* grid, uint32_t region_cells);
/* Parity probe: for n world-space points, the index coordinate, node offsets and trilinear value
 * exactly as the tracer computes them.  out_ijk[3n], out_offsets[3n] = (upper, lower, leaf) offsets,
 * out_value[n]; all device pointers. */
rtc8x_status rtc8x_grid_probe(const rtc8x_grid* grid, uint32_t n, const float* world_xyz,
                              int32_t* out_ijk, uint32_t* out_offsets, float* out_value, void* stream);

/* ------------------------------------------------------------------------------------------------
 * NanoVDB files (.nvdb) and grid buffers, byte-compatible with NanoVDB 32.x (the version rtcamp8
 * vendors, ext/nanovdb/nanovdb/NanoVDB.h:104-106).  Host code; nothing here touches the GPU except
 * rtc8x_grid_load_nvdb / rtc8x_grid_upload_nanovdb / rtc8x_grid_download.
 *
 * File: one or more segments, each a 16-byte header (magic "NanoVDB0", packed version, grid count,
 * codec) followed by one 176-byte MetaData record plus the NUL-terminated name per grid, then the grids'
 * bytes in order (util/IO.h:104-160).  Codec NONE stores a grid buffer verbatim; ZIP stores a u64 byte
 * count and a zlib stream (util/IO.h:254-344).  BLOSC is reported as RTC8X_ERR_NOT_SUPPORTED.
 * Grid buffer: GridData (672 B) | TreeData (64 B) | RootData (64 B) + 32-byte root tiles | upper nodes
 * (270,400 B) | lower nodes (33,856 B) | leaves (2,144 B), child links as signed byte offsets
 * (NanoVDB.h:2185-2205, 2501-2508, 2687-2760, 3043-3072, 3355-3373).  Only FloatGrid (GridType 1) is
 * decoded: the reference reads its volume as nanovdb::FloatGrid (renderer/renderer_scene.cpp:491-495).
 * ---------------------------------------------------------------------------------------------- */
#define RTC8X_NVDB_CODEC_NONE 0
#define RTC8X_NVDB_CODEC_ZIP 1
#define RTC8X_NVDB_CODEC_BLOSC 2

/* One grid of a file: the MetaData record (util/IO.h:144-159) and where its bytes are. */
typedef struct rtc8x_nvdb_grid_info {
    uint64_t grid_size;        /* bytes of the decoded grid buffer */
    uint64_t file_size;        /* bytes the grid takes in the file (compressed size for ZIP) */
    uint64_t voxel_count;      /* active voxels */
    uint64_t file_offset;      /* offset of the grid's bytes in the file */
    uint32_t grid_type;        /* NanoVDB GridType: 1 = Float */
    uint32_t grid_class;       /* NanoVDB GridClass: 2 = FogVolume */
    uint32_t codec;            /* RTC8X_NVDB_CODEC_* */
    uint32_t version;          /* major << 21 | minor << 10 | patch */
    uint32_t node_count[4];    /* leaves, lower, upper, root (1) */
    uint32_t tile_count[3];    /* active tiles in lower nodes, upper nodes, the root */
    uint32_t segment;          /* index of the segment that holds it */
    double world_bbox[6];      /* min xyz, max xyz */
    int32_t index_bbox[6];     /* inclusive */
    double voxel_size[3];
    char name[256];            /* NUL-terminated; longer names are truncated */
} rtc8x_nvdb_grid_info;

/* Lists the grids of a file image.  Writes up to `capacity` records to `infos` (may be NULL with
 * capacity 0) and the total number of grids to *count.  Replaces readGridMetaData (util/IO.h:246-250). */
rtc8x_status rtc8x_nvdb_list(const void* file, size_t file_size, rtc8x_nvdb_grid_info* infos, uint32_t capacity,
                             uint32_t* count);
/* Decodes a FloatGrid buffer (what a GridHandle holds) into host tables of the layout of rtc8x_grid,
 * in memory the library allocates (release with rtc8x_grid_free_host).  max_value is the majorant as
 * the reference computes it after loading: the maximum of getValue over the half-open index range
 * [bbox.min, bbox.max) of the root's index bbox (Scene::allocateVolumeGrid,
 * renderer/renderer_scene.cpp:497-511); background is the root's background value.  Malformed buffers
 * (offsets out of range or misaligned, node counts inconsistent with the links) are rejected. */
rtc8x_status rtc8x_nanovdb_decode(const void* grid_buffer, size_t size, rtc8x_grid* out_host);
/* Reads grid `n` of a file image (counting across segments, as nanovdb::io::readGrid(fileName, n),
 * util/IO.h:576-612) and decodes it as rtc8x_nanovdb_decode does. */
rtc8x_status rtc8x_nvdb_read_grid(const void* file, size_t file_size, uint32_t n, rtc8x_grid* out_host);
/* Releases host tables returned by rtc8x_nanovdb_decode / rtc8x_nvdb_read_grid / rtc8x_grid_download. */
rtc8x_status rtc8x_grid_free_host(rtc8x_grid* host);
/* Device grid from a FloatGrid buffer / from grid n of a .nvdb file image: decode + upload, keeping the
 * decoded majorant (readGrid + Scene::allocateVolumeGrid, renderer/renderer_scene.cpp:2467-2470,
 * 491-537).  Release with rtc8x_grid_destroy. */
rtc8x_status rtc8x_grid_upload_nanovdb(const void* grid_buffer, size_t size, rtc8x_grid* out);
rtc8x_status rtc8x_grid_load_nvdb(const void* file, size_t file_size, uint32_t n, rtc8x_grid* out);
/* Host copy of a device grid's tables (library-allocated; release with rtc8x_grid_free_host).
 * Synchronises the device. */
rtc8x_status rtc8x_grid_download(const rtc8x_grid* device_grid, rtc8x_grid* out_host);
/* Encodes host tables as a NanoVDB FloatGrid buffer of class FogVolume named `name` (at most 255 bytes).
 * Activity comes from the tables' value masks and root-tile states (without masks every leaf voxel is
 * active and every tile inactive), the map from map_*_d (or inv_mat / vec); node bounding boxes, min/max/average/standard deviation and the world bbox are computed as
 * nanovdb::gridStats does (util/GridStats.h:530-708) and the checksum as ChecksumMode::Partial, the
 * default of nanovdb::GridBuilder (util/GridChecksum.h:33-37, 175-203).  Nodes are laid out breadth-first
 * in traversal order, so a grid built by GridBuilder and decoded re-encodes to the same bytes.  With
 * out == NULL only *size is set; otherwise capacity must be at least that size. */
rtc8x_status rtc8x_nanovdb_encode(const rtc8x_grid* host, const char* name, void* out, size_t capacity, size_t* size);
/* Writes `count` FloatGrid buffers as one segment of a .nvdb file image with the given codec (NONE or
 * ZIP), as nanovdb::io::writeGrids does (util/IO.h:548-566).  With out == NULL only *size is set (an
 * upper bound for ZIP; the exact size is returned by the writing call). */
rtc8x_status rtc8x_nvdb_write(const void* const* grid_buffers, const size_t* grid_sizes, uint32_t count,
                              uint32_t codec, void* out, size_t capacity, size_t* size);

/* ------------------------------------------------------------------------------------------------
 * Environment light and sun.
 * ---------------------------------------------------------------------------------------------- */
typedef enum rtc8x_env_kind {
    RTC8X_ENV_NONE = 0,
    RTC8X_ENV_CONSTANT = 1,     /* uniform radiance, uniform-sphere sampling */
    RTC8X_ENV_LATLONG = 2       /* Y-up lat-long RGBA32F texture, point-filtered
                                   (ImageBasedEnvironmentalLight, common/common_renderer_types.h:2419-2458) */
} rtc8x_env_kind;

typedef enum rtc8x_env_sampler {
    RTC8X_ENV_SAMPLER_HIERARCHICAL = 0,  /* HierarchicalImportanceMap (common_renderer_types.h:434-610) */
    RTC8X_ENV_SAMPLER_ALIAS = 1          /* Walker alias table over texels (common_renderer_types.h:226-281) */
} rtc8x_env_sampler;

typedef struct rtc8x_alias_entry {
    uint32_t second_index;
    float prob_to_pick_first;
} rtc8x_alias_entry;

typedef struct rtc8x_env_light {
    int32_t kind;
    int32_t sampler;
    float power_coeff;           /* EnvironmentalLight::powerCoeff */
    float rotation;              /* EnvironmentalLight::rotation (radians) */
    float constant_radiance[3];
    uint32_t width, height;      /* texture size */
    const float* texels;         /* width*height*4 floats, row-major, y = 0 at the zenith */
    /* Hierarchical importance map: mip 0 is map_dim_x x map_dim_y (powers of two), mip L is
     * max(dim >> L, 1); levels are stored back to back in `mips`. */
    uint32_t map_dim_x, map_dim_y, num_mip_levels, reserved0;
    const float* mips;
    float map_integral;
    /* Two-level alias tables (RegularConstantContinuousDistribution2D with alias tables,
     * common_renderer_types.h:338-418): a marginal over the `height` rows and one conditional
     * table per row over `width` texels.  Weights are the importance-map mip-0 values
     * (luminance * pi/2 * sin(theta)); the tables are built on the host in FP64. */
    float alias_integral;        /* sum of all weights */
    const rtc8x_alias_entry* alias_rows;   /* height entries */
    const rtc8x_alias_entry* alias_cols;   /* height * width entries */
    const float* alias_weights;  /* width*height texel weights */
} rtc8x_env_light;

/* Builds the two-level alias tables for `env` from its texels (FP64 on the host; device buffers are
 * library-owned and released with rtc8x_env_release).  Also builds the hierarchical importance map
 * (mips) on the device as computeImportanceMap does (compute_light_probs.cu:161-211). */
rtc8x_status rtc8x_env_build_sampler(rtc8x_env_light* env, void* stream);
rtc8x_status rtc8x_env_release(rtc8x_env_light* env);

typedef struct rtc8x_sun_light {
    int32_t enabled;
    float direction[3];          /* unit vector towards the sun */
    float irradiance[3];         /* radiance integrated over the disc, W/m^2 per channel */
    float select_prob;           /* probability that next-event estimation picks the sun */
} rtc8x_sun_light;

/* ------------------------------------------------------------------------------------------------
 * Hosek-Wilkie 2013 sky bake (renderer/gpu_kernels/analytic_sky_kernels.cu): 16 spectral bands
 * (CIE 1931 2-degree observer binned over 360-830 nm), each with its own model state, summed to
 * XYZ, converted to linear sRGB, saturation adjusted, solar disc inside solar_radius, and a sea
 * colour below the horizon.
 * ---------------------------------------------------------------------------------------------- */
#define RTC8X_SKY_BANDS 16

/* ArHosekSkyModelState in float (common/common_renderer_types.h:955-966). */
typedef struct rtc8x_hosek_band {
    float configs[11][9];
    float radiances[11];
    float turbidity;
    float solar_radius;
    float emission_correction_factor_sky[11];
    float emission_correction_factor_sun[11];
    float albedo;
    float elevation;
} rtc8x_hosek_band;

typedef struct rtc8x_hosek_sky {
    const rtc8x_hosek_band* bands;      /* device, RTC8X_SKY_BANDS entries */
    const float* solar_rows;            /* device, 11 wavelengths x 2 turbidity indices x 180 coefficients */
    const float* limb_darkening;        /* device, 11 x 6 */
    int32_t solar_turbidity_base;       /* turbidity index of solar_rows[.][0] (= int(turbidity) - 1) */
    float cmf_x[RTC8X_SKY_BANDS];       /* ArHosekSkyModelCMFSet (common_renderer_types.h:968-975) */
    float cmf_y[RTC8X_SKY_BANDS];
    float cmf_z[RTC8X_SKY_BANDS];
    float center_wavelengths[RTC8X_SKY_BANDS];
    float integral_cmf;
} rtc8x_hosek_sky;

/* generateArHosekSkyEnvironmentalTexture (analytic_sky_kernels.cu:220-276): writes width x height
 * RGBA32F texels (alpha 1) of the lat-long sky for the given unit sun direction. */
rtc8x_status rtc8x_sky_bake(void* stream, const rtc8x_hosek_sky* sky, const float sun_dir[3], uint32_t width,
                            uint32_t height, float* out_rgba);
/* The same per-texel evaluation at n explicit directions (phi[i], theta[i]) (Y-up polar angles);
 * out_rgb has 3n floats.  Used to check the bake against FP64 ground truth. */
rtc8x_status rtc8x_sky_eval(void* stream, const rtc8x_hosek_sky* sky, const float sun_dir[3], uint32_t n,
                            const float* phi, const float* theta, float* out_rgb);

/* The spectral Hosek-Wilkie 2013 model for any sun (the reference re-derives the sky state every frame
 * from the interpolated solar direction and turbidity: AnalyticSkyEnvironment::computeDistribution,
 * renderer/renderer_scene.cpp:555-602).  rtc8x_hosek_model_create reads the model's tables from the
 * image of rtcamp8_amd/data/hosek_model.bin (char[8] "RTC8XHW1"; f64 configs[11][1080],
 * radiances[11][120], solar[11][1800], limb_darkening[11][6]; f32 cmf_x, cmf_y, cmf_z,
 * center_wavelengths [16 each], integral_cmf; little-endian). */
typedef struct rtc8x_hosek_model rtc8x_hosek_model;
rtc8x_status rtc8x_hosek_model_create(const void* data, size_t size, rtc8x_hosek_model** out);
rtc8x_status rtc8x_hosek_model_destroy(rtc8x_hosek_model* model);
/* The 16 per-band states for a solar elevation in [0, pi/2] (radians) and turbidity in [1, 10], each band
 * with ground albedo 0.5 (1 - (b + 0.5) / 16): arhosekskymodelstate_alloc_init
 * (ext/hosek_wilkie_2013/ArHosekSkyModel.c:147-345) in double, cast to float field by field.  Host. */
rtc8x_status rtc8x_hosek_bands(const rtc8x_hosek_model* model, float solar_elevation, float turbidity,
                               rtc8x_hosek_band* out);
/* Uploads the band states and solar-disc rows for (solar_elevation, turbidity) to device memory the model
 * owns and fills `sky` for rtc8x_sky_bake / rtc8x_sky_eval.  Synchronises the stream first; the pointers
 * stay valid until the next call or rtc8x_hosek_model_destroy. */
rtc8x_status rtc8x_hosek_model_state(rtc8x_hosek_model* model, float solar_elevation, float turbidity,
                                     rtc8x_hosek_sky* sky, void* stream);

/* ------------------------------------------------------------------------------------------------
 * Camera: PerspectiveCamera (common/common_renderer_types.h:1019-1034) reduced to what the
 * ray generation uses (optix_pathtracing_kernels.cu:43-55).
 * ---------------------------------------------------------------------------------------------- */
typedef struct rtc8x_camera {
    float position[3];
    float orientation[9];        /* row-major, Quaternion::toMatrix3x3 (common/basic_types.h:1620-1628) */
    float vh;                    /* 2 tan(fovY / 2) */
    float vw;                    /* aspect * vh */
} rtc8x_camera;

/* Fills an rtc8x_camera from position, unit quaternion (x, y, z, w), fovY and aspect exactly as
 * the reference computes the same quantities. */
rtc8x_status rtc8x_camera_setup(rtc8x_camera* cam, const float position[3], const float quat_xyzw[4],
                                float fov_y, float aspect);

typedef enum rtc8x_phase_kind {
    RTC8X_PHASE_SCHLICK = 0,     /* SchlickPhaseFunction (common_renderer_types.h:2063-2097) */
    RTC8X_PHASE_HENYEY_GREENSTEIN = 1
} rtc8x_phase_kind;

typedef enum rtc8x_tracking {
    RTC8X_TRACKING_GLOBAL_MAJORANT = 0,   /* reference: delta/ratio tracking against one majorant */
    RTC8X_TRACKING_LOCAL_MAJORANT = 1     /* DDA over 8^3 cells with per-cell majorants (empty space skipped) */
} rtc8x_tracking;

typedef enum rtc8x_entry {
    RTC8X_ENTRY_PATH_TRACE = 0,              /* __raygen__pathTrace */
    RTC8X_ENTRY_GENERATE_TRAINING_DATA = 1,  /* __raygen__generateTrainingData */
    RTC8X_ENTRY_PATH_TRACE_WITH_NRC = 2      /* __raygen__pathTraceWithNRC */
} rtc8x_entry;

/* Counters accumulated by rtc8x_trace when stats != NULL (uint64 device array). */
enum {
    RTC8X_STAT_PATHS = 0,
    RTC8X_STAT_VERTICES = 1,       /* rays traced through the medium (the reference's pathLength) */
    RTC8X_STAT_LOOKUPS = 2,        /* trilinear density lookups (8 voxel taps each) */
    RTC8X_STAT_SCATTER = 3,        /* real collisions */
    RTC8X_STAT_COUNT = 8
};

/* RadianceQuery (renderer/renderer_shared.h:46-65): 14 floats.  Note that the reference stores the
 * polar angle theta in the *_phi fields and the azimuth phi in the *_theta fields because it calls
 * toPolarYUp(theta*, phi*) as toPolarYUp(&phi, &theta) (optix_pathtracing_kernels.cu:18-25). */
typedef struct rtc8x_radiance_query {
    float position[3];
    float normal_phi, normal_theta;
    float vout_phi, vout_theta;
    float roughness;
    float diffuse_reflectance[3];
    float specular_reflectance[3];
} rtc8x_radiance_query;

/* TerminalInfo (renderer_shared.h:67-72), unpacked. */
typedef struct rtc8x_terminal_info {
    float throughput[3];
    uint32_t has_query;          /* bit 0: has query; bits 8..15: path length */
} rtc8x_terminal_info;

/* TrainingVertexInfo (renderer_shared.h:76-82), unpacked. */
typedef struct rtc8x_training_vertex_info {
    float local_throughput[3];
    uint32_t prev_vertex_index;  /* 0x003FFFFF = none */
    uint32_t path_length;
    uint32_t is_unbiased;
} rtc8x_training_vertex_info;

/* TrainingSuffixTerminalInfo (renderer_shared.h:84-90), unpacked. */
typedef struct rtc8x_training_suffix_terminal {
    uint32_t prev_vertex_index;
    uint32_t has_query;
    uint32_t path_length;
    uint32_t is_unbiased;
} rtc8x_training_suffix_terminal;

#define RTC8X_INVALID_VERTEX_INDEX 0x003FFFFFu
#define RTC8X_NUM_TRAINING_DATA_PER_FRAME 65536u          /* renderer_shared.h:12 */
#define RTC8X_TRAIN_BUFFER_SIZE (2u * 65536u)             /* renderer_shared.h:13 */

/* NRC buffers of StaticPipelineLaunchParameters (renderer_shared.h:127-141). */
typedef struct rtc8x_nrc_buffers {
    uint32_t max_num_training_suffixes;
    uint32_t nrc_buffer_index;               /* PerFrame::nrcBufferIndex */
    uint32_t offset_to_select_unbiased_path; /* PerFrame::offsetToSelectUnbiasedPath */
    float radiance_scale;                    /* PerFrame::radianceScale */
    float path_termination_factor;           /* renderer_shared.h:11 (0.01) */
    /* Samples per PATH_TRACE_WITH_NRC launch (0 counts as 1; the launch's num_samples must equal it):
     * the render buffers below hold that many image slices.  Sample k of pixel i uses
     * inference_queries / inferred_radiance entry k * round_up(W*H, 128) + i and inference_terminals /
     * per_frame_contribution entry k * W*H + i, so one inference covers every sample of a launch;
     * rtc8x_nrc_accumulate adds the slices in sample order, which gives the sums of as many one-sample
     * launches. */
    uint32_t num_samples;
    uint32_t* num_training_data;             /* [2] */
    int32_t* train_image_size;               /* [2][2] */
    float* target_min_max;                   /* [2][2][3]: per buffer index, min RGB then max RGB */
    float* target_avg;                       /* [2][3] */
    rtc8x_radiance_query* inference_queries; /* image size (rounded up to 128) x num_samples */
    rtc8x_terminal_info* inference_terminals;/* image size x num_samples */
    rtc8x_training_suffix_terminal* train_suffix_terminals; /* max_num_training_suffixes */
    float* inferred_radiance;                /* image size (rounded up to 128) x 3 x num_samples */
    float* per_frame_contribution;           /* image size x 3 x num_samples */
    rtc8x_radiance_query* train_queries[2];  /* RTC8X_TRAIN_BUFFER_SIZE each */
    float* train_targets[2];                 /* RTC8X_TRAIN_BUFFER_SIZE x 3 each */
    rtc8x_training_vertex_info* train_vertex_info; /* RTC8X_TRAIN_BUFFER_SIZE */
    uint32_t* data_shufflers;                /* RTC8X_NUM_TRAINING_DATA_PER_FRAME LCG states */
    uint32_t* scratch;                       /* library work space, see rtc8x_nrc_scratch_bytes */
} rtc8x_nrc_buffers;

/* Everything one launch reads (StaticPipelineLaunchParameters + PerFramePipelineLaunchParameters,
 * renderer_shared.h:113-167, restricted to the volume-only path). */
typedef struct rtc8x_frame_params {
    uint32_t image_width, image_height;      /* render size (trace size for GENERATE_TRAINING_DATA
                                                comes from nrc.train_image_size) */
    rtc8x_camera camera;
    rtc8x_grid grid;
    float density_coeff;                     /* PerFrame::densityCoeff */
    float scattering_albedo;                 /* PerFrame::scatteringAlbedo */
    float scattering_forwardness;            /* Schlick k or HG g */
    int32_t phase;                           /* rtc8x_phase_kind */
    rtc8x_env_light env;
    rtc8x_sun_light sun;
    float world_aabb_min[3], world_aabb_max[3]; /* WorldDimInfo (common_renderer_types.h:1124-1128) */
    float world_radius;
    uint32_t rr_after_path_length;           /* Russian roulette when pathLength > this (reference: 0) */
    uint32_t max_path_length;                /* 0 = unbounded (reference) */
    int32_t tracking;                        /* rtc8x_tracking */
    uint64_t* rng_states;                    /* PCG32 state per pixel, row-major (rngBuffer) */
    double* accum;                           /* per-pixel RGB sums, FP64 (accumBuffer) */
    uint32_t* path_vertices;                 /* optional per-pixel vertex count sums */
    uint64_t* stats;                         /* optional RTC8X_STAT_COUNT counters */
    /* Optional scheduling (results do not depend on it): pixel_cost[i] receives the density lookups of
     * launch pixel i (row-major over the launch) in this launch; work_order, a permutation of the
     * launch's pixels, is the order in which lanes take them (see rtc8x_order_by_tile_cost).  Without
     * work_order lanes take pixel tiles (tile_shape below) in order (GENERATE_TRAINING_DATA: the
     * unbiased-suffix pixels first). */
    uint32_t* pixel_cost;
    const uint32_t* work_order;
    rtc8x_nrc_buffers nrc;
    /* Scheduling knobs (no counterpart in the reference; results do not depend on them), 0 = default:
     * event_eighths, 1..8: a wavefront runs its parked lanes' vertex events once that many eighths of its
     * active lanes are parked (default 5 for PATH_TRACE and PATH_TRACE_WITH_NRC with local-majorant
     * tracking, 4 with the global majorant, 3 for GENERATE_TRAINING_DATA); tile_shape: the pixel tiles lanes take without a work order, 0 = 8x8 (default), 1 = 16x4,
     * 2 = 32x2, 3 = 4x16. */
    uint32_t event_eighths;
    uint32_t tile_shape;
} rtc8x_frame_params;

/* One launch of an entry point over the image (optixPipeline.launch(stream, plp, w, h, 1),
 * renderer/renderer_main.cpp:1575-1577, 1622-1624, 1640-1642).  num_samples paths are traced per
 * pixel in that launch; for PATH_TRACE each is added to accum, for the NRC entry points num_samples
 * must be 1. */
rtc8x_status rtc8x_trace(void* stream, const rtc8x_frame_params* params, int32_t entry,
                         uint32_t num_samples);

/* Work order for the next launch from the last launch's pixel_cost: pixels sorted by decreasing
 * floor(log2(cost + 1)), row-major within a class (a stable counting sort on the device), so the
 * longest paths start first.  No counterpart in the reference (its OptiX launch has no such control).
 * scratch: rtc8x_order_scratch_bytes(n) bytes. */
rtc8x_status rtc8x_order_by_cost(void* stream, const uint32_t* cost, uint32_t n, uint32_t* order, void* scratch);
size_t rtc8x_order_scratch_bytes(uint32_t n);
/* The same for a width x height image by 8 x 8 tiles: tiles sorted by their summed cost S (saturating
 * at 2^32 - 1) in quarter-octave classes below the costliest tile's, k = min(31, q(max S) - q(S)) with
 * q(c) = 4 floor(log2 c) + the two bits below c's leading one (k = 31 for S = 0), stable in row-major
 * tile order; each tile's pixels row-major and contiguous in `order` (edge tiles hold their in-image
 * pixels only).  The lanes of a wavefront then take neighbouring pixels, whose
 * paths share cache lines, while the costliest tiles still start first.  No counterpart in the
 * reference.  scratch: rtc8x_order_tiles_scratch_bytes(width, height) bytes. */
rtc8x_status rtc8x_order_by_tile_cost(void* stream, const uint32_t* cost, uint32_t width, uint32_t height,
                                      uint32_t* order, void* scratch);
size_t rtc8x_order_tiles_scratch_bytes(uint32_t width, uint32_t height);

/* Several instances -- each its own volume, camera, sky, image buffers and, for the NRC entry, cache
 * buffers -- in one persistent launch of a full-image entry (PATH_TRACE or PATH_TRACE_WITH_NRC);
 * instance i traces num_samples[i] samples per pixel exactly as rtc8x_trace(params[i]) would.  The
 * device's workgroups are split among the instances in proportion to shares[i] (NULL: pixels x
 * samples; every instance gets at least one workgroup); each instance keeps its own work counter.
 * count <= 16; the instances share one tracking mode.  No counterpart in the reference (it launches
 * one image per OptiX launch). */
rtc8x_status rtc8x_trace_batch(void* stream, const rtc8x_frame_params* params, const uint32_t* num_samples,
                               const float* shares, uint32_t count, int32_t entry);

/* rtc8x_trace_batch for callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  No counterpart in the reference.  The same
 * launch and results as rtc8x_trace_batch, but the instances' frame parameters are staged in a fixed
 * device array of 16 entries (16 x sizeof(rtc8x_frame_params) bytes per GPU, held until the process
 * ends) instead of a per-call allocation: count above 16 is an error, and nothing is allocated or freed.
 * Every call on one GPU shares that GPU's array, so calls that overlap on one GPU (different streams, or
 * threads) race: order them on one stream. */
rtc8x_status rtc8x_trace_batch_static(void* stream, const rtc8x_frame_params* params, const uint32_t* num_samples,
                                      const float* shares, uint32_t count, int32_t entry);

/* ------------------------------------------------------------------------------------------------
 * NRC data-path kernels (renderer/gpu_kernels/nrc_set_up_kernels.cu).  They read the same
 * rtc8x_frame_params as rtc8x_trace (buffers in params->nrc, image size, accum).
 * ---------------------------------------------------------------------------------------------- */
/* prepareNRC (:4-44): sets train_image_size[buf] (image/4 for a new sequence, else the previous size
 * scaled by sqrt(65536 / previous record count), clamped to the image), zeroes num_training_data[buf],
 * resets the target statistics and clears the training-suffix terminals. */
rtc8x_status rtc8x_nrc_prepare(void* stream, const rtc8x_frame_params* params, int32_t is_new_sequence);
/* propagateRadianceValues (:46-86): completes training targets along each suffix's vertex chain. */
rtc8x_status rtc8x_nrc_propagate(void* stream, const rtc8x_frame_params* params);
/* shuffleTrainingData (:88-174): scatters 65536 records from train_queries/targets[0] into [1] through
 * the LCG permutation, zeroes invalid records, scales and clamps targets, and writes the target
 * min / max / mean (reduced in a fixed order, so the result is deterministic). */
rtc8x_status rtc8x_nrc_shuffle(void* stream, const rtc8x_frame_params* params);
/* accumulateInferredRadianceValues (:176-219): accum += direct + throughput * max(NRC, 0) / scale *
 * (rho_d + rho_s), in FP64. */
rtc8x_status rtc8x_nrc_accumulate(void* stream, const rtc8x_frame_params* params);
/* Bytes of params->nrc.scratch the library needs for an image of width x height. */
size_t rtc8x_nrc_scratch_bytes(uint32_t width, uint32_t height);

/* ------------------------------------------------------------------------------------------------
 * Neural radiance cache network: replaces class NeuralRadianceCache (renderer/network_interface.h:17-31,
 * network_interface.cu:50-159) and the tiny-cuda-nn objects behind it.
 *
 * Input: n RadianceQuery records (14 floats each, AoS).  Encoding (64 dims): TriangleWave with 12
 * frequencies on the 3 position dims (36; |2 frac(x 2^(f-1)) - 1|, period 2 at the base frequency), OneBlob with 4 bins on the 5 angle/roughness dims (20;
 * polar angles scaled by 1/pi, azimuths by 1/(2 pi)), identity on the 6 reflectance dims, then a
 * constant 1 (the first layer's bias) and one zero pad.  Network: num_hidden_layers ReLU layers of
 * width 64 and a linear layer to RGB, no biases (FullyFusedMLP).  Loss: relative L2 over luminance,
 * the mean over the batch's 3n outputs of (p_c - t_c)^2 / (lum(p)^2 + 0.01) with the normaliser held
 * constant (tiny-cuda-nn's RelativeL2Luminance).  Optimiser: Adam with L2
 * regularisation, wrapped by an exponential moving average of the weights that inference uses.
 * Weights and activations are FP16 or BF16 with FP32 accumulation (MFMA); master weights, moments and
 * the EMA are FP32.  The per-sample gradient of the mean loss enters the backward pass multiplied by
 * the loss scale (rtc8x_nrc_desc.loss_scale), which the optimiser step divides out.
 * ---------------------------------------------------------------------------------------------- */
typedef struct rtc8x_nrc rtc8x_nrc;

typedef enum rtc8x_nrc_precision {
    RTC8X_NRC_FP16 = 0,
    RTC8X_NRC_BF16 = 1,
    /* BF16 training; inference on E4M3 operands (v_mfma_f32_16x16x32_fp8_fp8, FP32 accumulation), the
     * EMA weights quantised with one power-of-two scale per layer and the activations with one per
     * 64-query tile and layer, each mapping the largest magnitude to at most 240 */
    RTC8X_NRC_BF16_FP8 = 2,
    /* FP32 weights and activations in training and inference (v_mfma_f32_16x16x4_f32, FP32
     * accumulation): the full-precision network the reduced-precision modes are compared against */
    RTC8X_NRC_FP32 = 3
} rtc8x_nrc_precision;

typedef enum rtc8x_nrc_encoding {
    RTC8X_NRC_TRIANGLE_WAVE = 0,  /* PositionEncoding::TriangleWave (network_interface.cu:76-98) */
    RTC8X_NRC_HASH_GRID = 1       /* PositionEncoding::HashGrid, the reference's default (network_interface.cu:99-123,
                                     renderer_scene.cpp:50) */
} rtc8x_nrc_encoding;

typedef struct rtc8x_nrc_desc {
    uint32_t num_hidden_layers;  /* renderer_scene.cpp:50 uses 2; the standalone benchmark 5 (1..8) */
    float learning_rate;         /* 1e-2 */
    float beta1, beta2;          /* 0.9, 0.99 */
    float epsilon;               /* 1e-8 with TriangleWave, 1e-15 with HashGrid (network_interface.cu:96,122) */
    float l2_reg;                /* 1e-6 */
    float ema_decay;             /* 0.99 */
    int32_t precision;           /* rtc8x_nrc_precision */
    uint64_t seed;               /* weight initialisation (Xavier-uniform from PCG32) */
    int32_t encoding;            /* rtc8x_nrc_encoding */
    /* The initial loss scale of training: 0 for the default (128, tiny-cuda-nn's for FP16), or a power of
     * two in [1, 2^24].  The per-sample gradient of the mean loss is multiplied by it before it is rounded
     * to the network's working type and divided out in the optimiser step.  The scale is dynamic
     * (Micikevicius et al. 2018), with the same rule in every precision: a step whose gradient is not
     * finite, or whose scaled per-sample loss gradient leaves FP16's finite range (|x| > 65504), is
     * skipped and the scale halves (not below 1); after 128 finite steps in a row it doubles (up to 2^24).
     * With the default the loss gradients of this cache stay far inside FP16's range and every step is
     * applied, as in the reference; a large initial scale skips a fresh network's first steps, those
     * whose gradients hold outliers, until the scale has backed off. */
    float loss_scale;
} rtc8x_nrc_desc;

/* HashGrid position encoding (multiresolution hash encoding, Mueller et al. 2022, with the reference's
 * settings, network_interface.cu:103-111): 16 levels x 2 features, 2^15 entries per level, base
 * resolution 16, resolution x2 per level.  Level l has scale s = 16 * 2^l - 1 and resolution
 * r = ceil(s) + 1; a position p in [0,1]^3 is trilinearly interpolated between the corners
 * floor(p s + 1/2) + {0,1} (coordinates 0..r).  A level whose (r + 1)^3 corners fit the table is
 * indexed densely, x + y (r + 1) + z (r + 1)^2; the others by the spatial hash
 * (x * 1) ^ (y * 2654435761) ^ (z * 805459861) mod 2^15.  The 32 grid features take the place of the
 * 36 triangle waves; the 64-feature input is then [grid 32 | OneBlob 20 | identity 6 | 1 | 0 x 5].
 * The tables (FP32, initialised uniformly in [-1e-4, 1e-4]) follow the MLP weights in the parameter
 * vector. */
#define RTC8X_NRC_HASH_LEVELS 16
#define RTC8X_NRC_HASH_FEATURES 2
#define RTC8X_NRC_HASH_LOG2_TABLE 15
#define RTC8X_NRC_HASH_BASE_RESOLUTION 16

#define RTC8X_NRC_INPUT_DIMS 14
#define RTC8X_NRC_ENCODED_DIMS 64
#define RTC8X_NRC_WIDTH 64
#define RTC8X_NRC_OUTPUT_DIMS 3
#define RTC8X_NRC_OUTPUT_PAD 16

/* initialize() (network_interface.cu:50-134) */
rtc8x_status rtc8x_nrc_create(const rtc8x_nrc_desc* desc, rtc8x_nrc** out);
/* finalize() (network_interface.cu:136-141) */
rtc8x_status rtc8x_nrc_destroy(rtc8x_nrc* nrc);
/* infer() (network_interface.cu:143-149): out[3n] from in[14n], EMA weights. */
rtc8x_status rtc8x_nrc_infer(rtc8x_nrc* nrc, void* stream, const float* in, uint32_t n, float* out);
/* The inference step of a training iteration (renderer_main.cpp:1579-1587: infer the training suffixes
 * queries): p->nrc.inference_queries -> p->nrc.inferred_radiance for round_up(w h, 128) queries, where
 * w x h is the current buffer's training size as prepareNRC left it on the device.  The count is read on
 * the device, so unlike the reference the host does not wait for it. */
rtc8x_status rtc8x_nrc_infer_training(rtc8x_nrc* nrc, void* stream, const struct rtc8x_frame_params* p);
/* The render inference (renderer_main.cpp:1626-1632: infer() over round_up(W H, 128) queries before
 * accumulateInferredRadianceValues), for the p->nrc.num_samples slices of a PATH_TRACE_WITH_NRC launch,
 * evaluated only at the slots whose p->nrc.inference_terminals entry has a query: the only inferred values
 * rtc8x_nrc_accumulate reads.  Those equal rtc8x_nrc_infer's over the whole batch (FP16, BF16, FP32; with
 * E4M3 inference the per-tile activation scales come from the queries alone); the other slots of
 * p->nrc.inferred_radiance are left as they are.  An escaped path has no query: on a sparse cloud most
 * slots are skipped. */
rtc8x_status rtc8x_nrc_infer_render(rtc8x_nrc* nrc, void* stream, const struct rtc8x_frame_params* p);
/* train() (network_interface.cu:151-159): one optimiser step on n samples (n <= 65536); if loss_on_cpu
 * is not NULL the mean loss is copied to it (synchronises the stream). */
rtc8x_status rtc8x_nrc_train(rtc8x_nrc* nrc, void* stream, const float* in, const float* target, uint32_t n,
                             float* loss_on_cpu);
/* Number of weights (layer matrices [out][in], the output layer padded to 16 rows), and host copies of
 * them: which = 0 master weights, 1 the (bias-corrected) EMA used for inference. */
uint32_t rtc8x_nrc_num_params(const rtc8x_nrc* nrc);
/* The current loss scale (waits for the device).  No counterpart in the reference. */
rtc8x_status rtc8x_nrc_get_loss_scale(const rtc8x_nrc* nrc, float* scale);
rtc8x_status rtc8x_nrc_get_weights(const rtc8x_nrc* nrc, int32_t which, float* host, size_t count);
/* Replaces the master weights (and resets optimiser state and EMA) from a host array. */
rtc8x_status rtc8x_nrc_set_weights(rtc8x_nrc* nrc, const float* host, size_t count);
/* Parity probes: the 64-dim encoding of n queries (out[64n]), and the gradient of the batch's mean loss
 * with respect to the weights, without an update (grad[num_params], device). */
rtc8x_status rtc8x_nrc_encode(void* stream, const float* in, uint32_t n, float* out);
/* The network's own input encoding (either kind), with its training (which = 0) or inference (1)
 * hash tables; out[64n]. */
rtc8x_status rtc8x_nrc_encode_net(const rtc8x_nrc* nrc, void* stream, const float* in, uint32_t n, float* out,
                                  int32_t which);
rtc8x_status rtc8x_nrc_gradient(rtc8x_nrc* nrc, void* stream, const float* in, const float* target, uint32_t n,
                                float* grad);

/* Seeds rng_states row-major from std::mt19937_64(seed) as initializeScreenRelatedBuffers does
 * (renderer/renderer_main.cpp:86-93).  Host-side, then uploads; `states` is a device pointer. */
rtc8x_status rtc8x_seed_rng(uint64_t* states, uint32_t width, uint32_t height, uint64_t seed);

/* ------------------------------------------------------------------------------------------------
 * Post-processing and image files.
 * The batch driver tone-maps the accumulated frame on the device (applyToneMap,
 * renderer/gpu_kernels/post_process_kernels.cu:9-40, launched at renderer_main.cpp:1646-1647) and
 * saves it through the SDR conversion (saveImage, common/common_host.cpp:515-570) as PNG; the GUI also
 * writes EXR (saveImageHDR, common/common_host.cpp:443-513) and BMP (common/common_host.cpp:422-441).
 * ---------------------------------------------------------------------------------------------- */
/* SDRImageSaverConfig (common/common_host.h:183-195). */
typedef struct rtc8x_sdr_config {
    float alpha_override;       /* >= 0: every pixel's alpha; < 0: the data's alpha (1 without one) */
    float brightness_scale;     /* multiplies the luminance fed to the tone curve */
    int32_t apply_tone_map;     /* colour *= (1 - exp(-brightness_scale * L)) / L, after zeroing non-finite
                                   pixels and clamping negatives */
    int32_t apply_srgb_gamma;   /* sRGB transfer curve per channel (negatives clamped first) */
    int32_t flip_y;             /* row y of the output comes from row height-1-y of the input */
} rtc8x_sdr_config;

typedef struct rtc8x_tone_map_params {
    const double* accum;        /* width*height*3 FP64 sums over `samples` samples (the tracer's accumulator) */
    uint32_t width, height;
    uint32_t samples;           /* > 0 */
    float exponent;             /* per-channel power before the tone curve: 1/0.6 in the reference */
    float brightness;           /* tone curve 1 - exp(-brightness * L): 1 in the reference */
    float* out_rgba32f;         /* optional: the tone-mapped frame, alpha 1 (the reference's output buffer) */
    uint32_t* out_rgba8;        /* optional: out_rgba32f through the SDR conversion with `sdr`, packed
                                   R | G << 8 | B << 16 | A << 24 (the saved image's pixels) */
    rtc8x_sdr_config sdr;
} rtc8x_tone_map_params;

/* The post-process pass on the device: mean of the accumulator, applyToneMap, and (optionally) the SDR
 * conversion fused into the same kernel.  Device pointers; asynchronous on `stream`. */
rtc8x_status rtc8x_tone_map(void* stream, const rtc8x_tone_map_params* params);
/* The SDR conversion of saveImage on the host: data holds width*height*num_channels floats (1..4
 * channels), out receives width*height packed RGBA8 pixels.  Channel values are quantised as
 * min(trunc(255 v), 255), with NaN and values below 0 giving 0. */
rtc8x_status rtc8x_sdr_convert(const float* data, uint32_t width, uint32_t height, uint32_t num_channels,
                               const rtc8x_sdr_config* config, uint32_t* out);
/* Image files in memory (host).  With out == NULL only *size is set (an upper bound for PNG; the exact
 * size is returned by the writing call).
 * PNG: 8-bit RGBA, per-row adaptive filters, zlib-deflated (the reference writes PNG with fpng).
 * BMP: the bytes stb_image_write v1.14's stbi_write_bmp writes for 4-channel data: 24-bit bottom-up BGR,
 *      alpha composited over (255, 0, 255), rows padded to 4 bytes.
 * EXR: scanline OpenEXR, no compression, half-float channels A, B, G, R, each value brightness_scale
 *      times the input (alpha included, as saveImageHDR scales it; 0 where the input has fewer channels),
 *      rounded to nearest even; flip_y as in rtc8x_sdr_config. */
rtc8x_status rtc8x_png_encode(const uint32_t* rgba8, uint32_t width, uint32_t height, void* out, size_t capacity,
                              size_t* size);
rtc8x_status rtc8x_bmp_encode(const uint32_t* rgba8, uint32_t width, uint32_t height, void* out, size_t capacity,
                              size_t* size);
rtc8x_status rtc8x_exr_encode(const float* data, uint32_t width, uint32_t height, uint32_t num_channels,
                              float brightness_scale, int32_t flip_y, void* out, size_t capacity, size_t* size);

#ifdef __cplusplus
}
#endif

#endif /* RTC8X_H */

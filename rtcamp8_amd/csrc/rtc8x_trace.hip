// rtc8x_trace.hip -- the path-tracing launch: the three OptiX ray-generation programs
// (pathTrace, generateTrainingData, pathTraceWithNRC; renderer/gpu_kernels/optix_pathtracing_kernels.cu
// :31-448) for volume-only scenes, as one persistent per-lookup state machine.
//
// A lane owns one pixel at a time (all of its samples, so the pixel's PCG stream is consumed in the
// reference's order) and steps its path through free-flight tracking towards the next scattering
// vertex, ratio tracking of the NEE shadow ray, and vertex events (camera ray, escape, scattering with
// Russian roulette and NRC termination, NEE set-up, phase sampling, training records, path end).  Each
// iteration of the main loop takes one tentative collision, with its density lookup, for every lane
// that is stepping a segment; a lane that reaches a vertex event parks, and the parked lanes run their
// events together once they are a quarter of the wavefront's active lanes, so that each event's code is
// issued once for many lanes.  A lane whose pixel is done takes the next one from a work counter
// (8x8 pixel tiles or a caller-given order, one atomic per wavefront).  Per pixel, the random numbers
// and floating-point operations are the reference's, so results are bit-identical to tracing each
// pixel's paths one after another.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>

#include "rtc8x_common.hpp"
#include "rtc8x_device.hpp"

namespace rtc8x {
namespace dev {

constexpr uint32_t kHardMaxPathLength = 4096u;
// Tentative collisions looked up together per lane (global majorant).  Deeper batches (3, 4) and a
// separate depth for ratio tracking or for the training launch measured slower on the 512^3 cloud.
constexpr uint32_t kSpeculate = 2;
// A wavefront with at least this many idle lanes evaluates its stepping lanes' segments together.
constexpr uint32_t kCoopIdle = 48;
constexpr uint32_t kTrainSpreadLanes = 64 - kCoopIdle;  // paths per wavefront of a spread training launch
constexpr uint32_t kCoopLocalPaths = 2;  // at most this many stepping lanes for the local walk's cooperative tail
constexpr uint32_t kCounterSlots = 1024;
constexpr uint32_t kMaxBatch = 16;  // instances of one batched launch
__device__ uint32_t g_work_counter[kCounterSlots];

__global__ void reset_counter_kernel(uint32_t slot) { g_work_counter[slot] = 0u; }
__global__ void reset_value_kernel(uint32_t* v) { *v = 0u; }

// Global majorant, the reference's loops (optix_pathtracing_kernels.cu:110-125,
// renderer_kernel_common.h:371-382).
template <>
struct Seg<RTC8X_TRACKING_GLOBAL_MAJORANT> {
    V3 o, d;
    float t, t1;
    __device__ __forceinline__ bool init(const rtc8x_grid& g, const Medium& m, V3 ro, V3 rd, float tmax)
    {
        float t0;
        if (!(m.majorant > 0.0f)) return false;
        if (!clip(g, ro, rd, tmax, t0, t1)) return false;
        t = fmaxf(0.0f, t0);
        o = ro;
        d = rd;
        return true;
    }
    __device__ __forceinline__ bool advance(const rtc8x_grid&, const Medium& m, Pcg& rng, float& mu, const uint8_t*)
    {
        t += -rtc8x_logf(1.0f - rng.uni()) / m.majorant;
        if (t > t1) return false;
        mu = m.majorant;
        return true;
    }
    // returns true when the tentative collision is real
    __device__ __forceinline__ bool after_track(Pcg& rng, float sigma, float mu) { return rng.uni() < sigma / mu; }
    __device__ __forceinline__ void after_shadow(Pcg&, float sigma, float mu, float& vis) { vis *= 1.0f - sigma / mu; }
};

enum : uint32_t {
    kModeWork = 0,      // needs a pixel
    kModeTrack = 1,     // free-flight tracking segment in progress
    kModeShadow = 2,    // NEE ratio-tracking segment in progress
    kModeIdle = 3,      // no work left
    kModeStart = 16,    // start a path (camera ray)
    kModeEscape = 33,   // tracking segment ended without a collision
    kModeNeeDone = 34,  // shadow segment ended
    kModeScatter = 49   // real collision at sg.t
};

struct Lane {
    uint32_t lx, ly, lin, sample, cost0, scat0;
    Pcg rng;
    V3 org, dir, pos;
    RGB tp, contribution, nee_f;
    float dir_pd, init_importance, vis, nee_k1, nee_k2, nee_k3;
    uint32_t path_length, mode;
    bool nee_sun;
    // NRC state (optix_pathtracing_kernels.cu:57-75, 200-214, 358-410)
    bool unbiased;
    float primary_spread, cur_sqrt_spread;
    uint32_t prev_idx, records, train_base;
    bool lost;
    RGB prev_local_tp, nee_c;
};

template <int TRACKING>
__device__ __forceinline__ bool start_segment(const rtc8x_grid& g, const Medium& med, Seg<TRACKING>& sg, Pcg& rng, V3 o,
                                              V3 d, float tmax)
{
    if (!sg.init(g, med, o, d, tmax)) return false;
    if constexpr (TRACKING == RTC8X_TRACKING_LOCAL_MAJORANT) sg.first_tau(rng);
    return true;
}

// S tentative collisions of a global-majorant segment at once (track: delta tracking, else ratio
// tracking).  The positions depend only on the random numbers, so their lookups are issued together.
// The numbers are drawn in the sequential algorithm's order (tracking: step, accept, step, accept, ...;
// ratio tracking: step, step, ...) and the results are consumed in order; after an acceptance or the
// segment's end the PCG state is rewound to where the sequential loop would stand, so later draws and
// the transmittance product are those of the sequential loop.
template <uint32_t S>
__device__ __forceinline__ void speculate_global(const rtc8x_grid& g, Medium& med,
                                                 Seg<RTC8X_TRACKING_GLOBAL_MAJORANT>& sg, Lane& L, const bool track)
{
    Pcg r = L.rng;
    float tk[S], ua[S];
    uint64_t st_after[S];
    uint32_t n = 0;
    bool ended = false;
    uint64_t st_end = 0;
    float t = sg.t;
#pragma unroll
    for (uint32_t k = 0; k < S; ++k) {
        tk[k] = sg.t;
        ua[k] = 0.0f;
        st_after[k] = 0;
        if (!ended) {
            t += -rtc8x_logf(1.0f - r.uni()) / med.majorant;
            if (t > sg.t1) {
                ended = true;
                st_end = r.state;
            } else {
                tk[k] = t;
                if (track) ua[k] = r.uni();
                st_after[k] = r.state;
                n = k + 1;
            }
        }
    }
    float sig[S];
    density_batch<S>(g, med.coeff, sg.o, sg.d, tk, sig);
    bool accepted = false;
    float t_acc = 0.0f;
    uint64_t st_acc = 0;
#pragma unroll
    for (uint32_t k = 0; k < S; ++k) {
        if (!accepted && k < n) {
            med.lookups++;
            if (track) {
                if (ua[k] < sig[k] / med.majorant) {
                    accepted = true;
                    t_acc = tk[k];
                    st_acc = st_after[k];
                }
            } else {
                L.vis *= 1.0f - sig[k] / med.majorant;
            }
        }
    }
    if (accepted) {
        sg.t = t_acc;
        L.rng.state = st_acc;
        L.mode = kModeScatter;
    } else if (ended) {
        L.rng.state = st_end;
        L.mode = track ? kModeEscape : kModeNeeDone;
    } else {
        sg.t = tk[S - 1];
        L.rng.state = st_after[S - 1];
    }
}

// GENERATE_TRAINING_DATA in one pass.  Each record takes the next slot of a staging area (query,
// target, vertex info, tag = launch pixel and the record's index on its path, and a snapshot of the path
// at the record) in whatever order lanes reach it, and each pixel leaves its record count; a prefix sum
// of the counts then gives every record its slot in pixel order, and stage_copy_kernel moves the
// records there.  A path whose records run past the buffer ends, in pixel-order indexing, at its first
// record beyond it: the copy rebuilds that end from the record and its snapshot.  When more records come
// than the staging area holds, the last ones find no slot; their pixels are traced again from their
// saved PCG states by a redo launch that records with the known offsets.
struct TrainStage {
    // staging launch (counter != NULL)
    uint32_t* counter;                   // records staged (may exceed kStageCapacity)
    rtc8x_radiance_query* queries;       // kStageCapacity each
    float* targets;
    rtc8x_training_vertex_info* info;
    uint32_t* tag;
    StageSnapshot* snap;
    uint32_t* path_length;               // per launch pixel: vertices, lookups and real collisions of its
    uint32_t* lookups;                   //   whole path
    uint32_t* scatters;
    uint32_t* lost;                      // per launch pixel: 1 when a record found no staging slot
    // redo launch (redo_lost != NULL): the pixels with a lost record, from their saved states
    const uint32_t* redo_lost;
    const uint64_t* rng_save;            // launch pixel order
    const uint32_t* staged;              // the staging launch's counter: no record was lost when it is at
                                         //   most kStageCapacity, and the redo launch then exits at once
};

// PCG32 jump-ahead: the state j draws later is mul[j] * state + add[j] (the LCG composed j times).
struct PcgJump {
    uint64_t mul[129], add[129];
};
constexpr PcgJump make_pcg_jump()
{
    PcgJump t{};
    uint64_t m = 1, a = 0;
    for (int j = 0; j <= 128; ++j) {
        t.mul[j] = m;
        t.add[j] = a;
        a = a * 6364136223846793005ULL + 1ULL;
        m = m * 6364136223846793005ULL;
    }
    return t;
}
__constant__ PcgJump kPcgJump = make_pcg_jump();

__device__ __forceinline__ float lane_f(float v, uint32_t lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)lane));
}
__device__ __forceinline__ uint64_t lane_u64(uint64_t v, uint32_t lane)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)lane);
    return ((uint64_t)hi << 32) | lo;
}

// Late in a launch most lanes of a wavefront have no work left while a few paths go on, each one
// segment at a time.  Then the whole wavefront takes the stepping lanes' segments one after the other:
// lane j evaluates the segment's j-th next tentative collision, with its random numbers taken from the
// PCG stream by jump-ahead, so 64 collisions cost one round of lookups.  The distances are summed and the
// transmittance factors multiplied in the sequential order (a uniform loop), and the source lane gets
// the state the sequential loop would end in: the same collision, PCG state, transmittance and lookup
// count.
__device__ __forceinline__ void coop_segments(const rtc8x_grid& g, Medium& med, Seg<RTC8X_TRACKING_GLOBAL_MAJORANT>& sg,
                                           Lane& L, uint64_t todo, uint32_t lane)
{
    const float maj = med.majorant;
    while (todo) {
        const uint32_t src = (uint32_t)__ffsll((long long)todo) - 1u;
        todo &= todo - 1;
        const bool track = (uint32_t)__builtin_amdgcn_readlane((int)L.mode, (int)src) == kModeTrack;
        const uint32_t k = track ? 2u : 1u;  // draws per tentative collision
        const V3 o{lane_f(sg.o.x, src), lane_f(sg.o.y, src), lane_f(sg.o.z, src)};
        const V3 d{lane_f(sg.d.x, src), lane_f(sg.d.y, src), lane_f(sg.d.z, src)};
        const float t1 = lane_f(sg.t1, src);
        float t = lane_f(sg.t, src);
        float vis = lane_f(L.vis, src);
        uint64_t st = lane_u64(L.rng.state, src);
        uint32_t looked = 0, mode = 0;
        for (;;) {
            Pcg r;
            r.state = kPcgJump.mul[lane * k] * st + kPcgJump.add[lane * k];
            const float u = r.uni();
            const uint64_t st_step = r.state;  // after the step's draw
            const float ua = track ? r.uni() : 0.0f;
            const uint64_t st_after = r.state;
            const float dj = -rtc8x_logf(1.0f - u) / maj;
            // t after collision j, summed in order; the segment ends at the first collision past t1
            float tj = INFINITY, run = t;
            uint32_t e = 64;
            for (uint32_t i = 0; i < 64; ++i) {
                run += lane_f(dj, i);
                if (i == lane) tj = run;
                if (run > t1) {
                    e = i;
                    break;
                }
            }
            const bool valid = lane < e;
            float sig = 0.0f;
            if (valid) sig = density_at(g, med.coeff, o, d, tj);
            if (track) {
                const uint64_t acc = __ballot(valid && ua < sig / maj);
                if (acc) {
                    const uint32_t a = (uint32_t)__ffsll((long long)acc) - 1u;
                    looked += a + 1;
                    t = lane_f(tj, a);
                    st = lane_u64(st_after, a);
                    mode = kModeScatter;
                    break;
                }
                looked += e;
                if (e < 64) {
                    st = lane_u64(st_step, e);
                    mode = kModeEscape;
                    break;
                }
            } else {
                const float f = 1.0f - sig / maj;
                for (uint32_t i = 0; i < e; ++i) vis *= lane_f(f, i);
                looked += e;
                if (e < 64) {
                    st = lane_u64(st_after, e);
                    mode = kModeNeeDone;
                    break;
                }
            }
            t = lane_f(tj, 63);
            st = lane_u64(st_after, 63);
        }
        if (lane == src) {
            if (mode == kModeScatter) sg.t = t;
            L.rng.state = st;
            L.vis = vis;
            L.mode = mode;
            med.lookups += looked;
        }
    }
}

// The cooperative tail of the local walk.  Its tentative collisions lie where the optical depth drawn for
// them runs out along the region walk, so lane j replays the sequential walk from the segment's state
// through the segment's j-th next collision (walk steps only, with the random numbers drawn in the
// sequential order: a collision's acceptance number when tracking, then the next optical depth), which
// gives the sequential positions bit for bit.  The 64 lookups then go out in one round and the results
// are consumed in order, as in coop_segments.  When all 64 collisions are null the wavefront goes on
// from the last one's walk state with its next optical depth drawn.
__device__ __forceinline__ void coop_segments_local(const rtc8x_grid& g, Medium& med,
                                                   Seg<RTC8X_TRACKING_LOCAL_MAJORANT>& sg, Lane& L, uint64_t todo,
                                                   uint32_t lane, const uint8_t* __restrict__ table)
{
    using SegL = Seg<RTC8X_TRACKING_LOCAL_MAJORANT>;
    auto take = [](const SegL& a, uint32_t from) {
        SegL b;
        b.o = V3{lane_f(a.o.x, from), lane_f(a.o.y, from), lane_f(a.o.z, from)};
        b.d = V3{lane_f(a.d.x, from), lane_f(a.d.y, from), lane_f(a.d.z, from)};
        b.t = lane_f(a.t, from);
        b.t1 = lane_f(a.t1, from);
        b.tau = lane_f(a.tau, from);
        b.s.c0 = __builtin_amdgcn_readlane(a.s.c0, (int)from);
        b.s.c1 = __builtin_amdgcn_readlane(a.s.c1, (int)from);
        b.s.c2 = __builtin_amdgcn_readlane(a.s.c2, (int)from);
        b.s.s0 = __builtin_amdgcn_readlane(a.s.s0, (int)from);
        b.s.s1 = __builtin_amdgcn_readlane(a.s.s1, (int)from);
        b.s.s2 = __builtin_amdgcn_readlane(a.s.s2, (int)from);
        b.s.io0 = lane_f(a.s.io0, from);
        b.s.io1 = lane_f(a.s.io1, from);
        b.s.io2 = lane_f(a.s.io2, from);
        b.s.iv0 = lane_f(a.s.iv0, from);
        b.s.iv1 = lane_f(a.s.iv1, from);
        b.s.iv2 = lane_f(a.s.iv2, from);
        b.s.tn0 = lane_f(a.s.tn0, from);
        b.s.tn1 = lane_f(a.s.tn1, from);
        b.s.tn2 = lane_f(a.s.tn2, from);
        return b;
    };
    while (todo) {
        const uint32_t src = (uint32_t)__ffsll((long long)todo) - 1u;
        todo &= todo - 1;
        const bool track = (uint32_t)__builtin_amdgcn_readlane((int)L.mode, (int)src) == kModeTrack;
        SegL w = take(sg, src);
        float vis = lane_f(L.vis, src);
        uint64_t st = lane_u64(L.rng.state, src);
        uint32_t looked = 0, mode = 0;
        float t_hit = 0.0f;
        for (;;) {
            SegL c = w;
            Pcg r;
            r.state = st;
            bool valid = true;
            float mu = 0.0f, ua = 0.0f;
            uint64_t st_acc = 0;  // after the collision's acceptance number
#pragma unroll 1
            for (uint32_t i = 0; i <= lane; ++i) {
                if (i > 0) c.tau = -rtc8x_logf(1.0f - r.uni());
                int res;
                do {
                    res = c.advance(g, med, r, mu, table);
                } while (res == kAdvanceBudget);
                if (res != kAdvanceCollision) {
                    valid = false;
                    break;
                }
                if (track) {
                    ua = r.uni();
                    st_acc = r.state;
                }
            }
            const float sig = valid ? density_at(g, med.coeff, c.o, c.d, c.t) : 0.0f;
            const uint32_t e = (uint32_t)__popcll(__ballot(valid));  // the valid lanes are 0 .. e-1
            if (track) {
                const uint64_t acc = __ballot(valid && ua < sig / mu);
                if (acc) {
                    const uint32_t a = (uint32_t)__ffsll((long long)acc) - 1u;
                    looked += a + 1;
                    t_hit = lane_f(c.t, a);
                    st = lane_u64(st_acc, a);
                    mode = kModeScatter;
                    break;
                }
            } else {
                const float f = valid ? 1.0f - sig / mu : 1.0f;
                for (uint32_t i = 0; i < e; ++i) vis *= lane_f(f, i);
            }
            looked += e;
            if (e < 64) {  // the segment ended: the state where its walk ran out
                st = lane_u64(r.state, e);
                mode = track ? kModeEscape : kModeNeeDone;
                break;
            }
            // all 64 null: on from the last collision, its next optical depth drawn
            w = take(c, 63);
            Pcg rn;
            rn.state = lane_u64(track ? st_acc : r.state, 63);
            w.tau = -rtc8x_logf(1.0f - rn.uni());
            st = rn.state;
        }
        if (lane == src) {
            if (mode == kModeScatter) sg.t = t_hit;
            L.rng.state = st;
            L.vis = vis;
            L.mode = mode;
            med.lookups += looked;
        }
    }
}

// The training launch's size, which prepareNRC left on the device for the current buffer, clamped to
// the image (nrc_set_up_kernels.cu:4-44 computes it; reading it here spares the host a synchronisation).
__device__ __forceinline__ void train_dims(const rtc8x_frame_params& p, uint32_t& tw, uint32_t& th)
{
    const int32_t* ts = p.nrc.train_image_size + 2 * p.nrc.nrc_buffer_index;
    tw = (uint32_t)min(max(ts[0], 0), (int32_t)p.image_width);
    th = (uint32_t)min(max(ts[1], 0), (int32_t)p.image_height);
}

// pathTrace_generic<withNRC, genData> (optix_pathtracing_kernels.cu:31-434) for volume-only scenes, as
// a per-lookup state machine over a persistent grid.  ENTRY selects the ray-generation program;
// GENERATE_TRAINING_DATA runs as the staging launch or the redo launch (TrainStage; the redo records at
// record_offsets).
template <int ENTRY, int TRACKING>
__device__ __forceinline__ void trace_body(const rtc8x_frame_params& p, uint32_t launch_w, uint32_t launch_h,
                                           uint32_t num_samples, uint32_t slot,
                                           const uint32_t* __restrict__ record_offsets,
                                           uint32_t* __restrict__ record_counts, const TrainStage& stage,
                                           const uint8_t* __restrict__ region_table)
{
    constexpr bool with_nrc = ENTRY != RTC8X_ENTRY_PATH_TRACE;
    constexpr bool gen_data = ENTRY == RTC8X_ENTRY_GENERATE_TRAINING_DATA;
    const rtc8x_nrc_buffers& nb = p.nrc;
    const bool staged = gen_data && stage.counter != nullptr;
    const bool redo = gen_data && stage.redo_lost != nullptr;
    float* const train_targets = staged ? stage.targets : nb.train_targets[0];
    // GENERATE_TRAINING_DATA is launched with a zero size and takes the training size from the device
    uint32_t W = launch_w, H = launch_h;
    if (gen_data && W == 0) train_dims(p, W, H);
    // pixel tiles of 64 lanes: 8x8, 16x4, 32x2 or 4x16
    const uint32_t tile_w = p.tile_shape == 1 ? 16u : (p.tile_shape == 2 ? 32u : (p.tile_shape == 3 ? 4u : 8u));
    const uint32_t tile_h = 64u / tile_w;
    const uint32_t tiles_x = (W + tile_w - 1) / tile_w;
    // GENERATE_TRAINING_DATA hands out the unbiased-suffix pixels first (the one pixel of each 4x4 block
    // with ((ly % 4) * 4 + lx % 4 + offset) % 16 == 0): their paths run to full length, so starting
    // them first keeps the launch from waiting on them at the end.
    const uint32_t first_c = (16u - nb.offset_to_select_unbiased_path % 16u) % 16u;
    const uint32_t first_rx = first_c % 4u, first_ry = first_c / 4u;
    const uint32_t first_nx = W > first_rx ? (W - first_rx + 3u) / 4u : 0u;
    const uint32_t first_ny = H > first_ry ? (H - first_ry + 3u) / 4u : 0u;
    const uint32_t n_first = (gen_data && !p.work_order) ? first_nx * first_ny : 0u;
    const uint32_t total = p.work_order ? W * H : n_first + tiles_x * ((H + tile_h - 1) / tile_h) * 64;
    // A training launch with few paths for the persistent grid (frame 0's quarter-size training image,
    // a small frame) is bound by its longest paths' chains of lookups, not by throughput.  Its
    // wavefronts then take at most `lanes` paths each, so that the paths spread over the whole grid and
    // every wavefront has the idle lanes that take its paths' segments cooperatively from the start.
    uint32_t lanes = 64;
    if constexpr (gen_data) {
        const uint32_t waves = gridDim.x * (blockDim.x >> 6);
        const uint32_t per = (total + waves - 1) / waves;
        if (per <= kTrainSpreadLanes) lanes = per > 0 ? per : 1u;
    }
    // default event thresholds, measured on the 512^3 cloud at 1080p: the render entries 5/8 with the
    // local walk (4/8 gave 35.1 against 34.7 ms per plain 16-spp frame, 6/8 34.7) and 4/8 with the global
    // one (3/8: 44.0 against 43.2 ms); the training entry 3/8 (no difference measured)
    constexpr uint32_t kEv8Default = ENTRY == RTC8X_ENTRY_GENERATE_TRAINING_DATA     ? 3u
                                     : TRACKING == RTC8X_TRACKING_LOCAL_MAJORANT ? 5u
                                                                                 : 4u;
    const uint32_t ev8 = (p.event_eighths >= 1 && p.event_eighths <= 8) ? p.event_eighths : kEv8Default;
    const uint32_t lane_id = threadIdx.x & 63;
    // per-sample slices of the render buffers (one launch may trace several samples into the cache)
    const size_t q_stride = ((size_t)W * H + 127u) & ~(size_t)127u, p_stride = (size_t)W * H;
    uint32_t max_len = (ENTRY == RTC8X_ENTRY_PATH_TRACE && p.max_path_length) ? p.max_path_length : kHardMaxPathLength;
    if (max_len > kHardMaxPathLength) max_len = kHardMaxPathLength;
    uint32_t rr_after = p.rr_after_path_length;
    if (gen_data && rr_after < 2) rr_after = 2;
    const float prob_env = (p.sun.enabled && p.env.kind != RTC8X_ENV_NONE) ? 1.0f - p.sun.select_prob : 1.0f;

    Medium med;
    med.coeff = p.density_coeff;
    med.majorant = p.density_coeff * p.grid.max_value;
    med.lookups = 0;
    uint32_t paths = 0, vertices = 0, scatters = 0;

    Lane L;
    L.mode = kModeWork;
    L.lin = 0;
    Seg<TRACKING> sg;
    float mu = 0.0f;
    for (;;) {
        // Vertex events are batched: a lane that reaches one parks, and once enough of the wavefront's
        // active lanes have parked they run their events together (each event block is issued once for
        // many lanes instead of once per lane), while the others keep stepping through lookups.  The
        // plain and training entries then run only the larger of the two vertex events (a real
        // collision: Russian roulette and NEE set-up; an NEE segment's end: phase sampling and the next
        // segment) together with the path starts and ends, from event_eighths / 8 of the active lanes;
        // the other event's lanes wait for a later pass.  The NRC render entry, whose paths are short,
        // runs every parked lane from event_eighths / 8.
        const uint64_t parked = __ballot(L.mode != kModeTrack && L.mode != kModeShadow && L.mode != kModeIdle);
        const uint64_t stepping = __ballot(L.mode == kModeTrack || L.mode == kModeShadow);
        if (parked == 0 && stepping == 0) break;
        const uint32_t n_parked = __popcll(parked), n_active = __popcll(parked | stepping);
        bool run_events;
        if constexpr (ENTRY == RTC8X_ENTRY_PATH_TRACE_WITH_NRC) {
            run_events = stepping == 0 || 8 * n_parked >= ev8 * n_active;
        } else {
            run_events = stepping == 0;
            if (!run_events && 8 * n_parked >= ev8 * n_active) {
                const uint64_t at_collision = __ballot(L.mode == kModeScatter), at_nee_end = __ballot(L.mode == kModeNeeDone);
                const uint64_t others = parked & ~(at_collision | at_nee_end);
                const uint64_t sel =
                    (__popcll(at_collision) >= __popcll(at_nee_end) ? at_collision : at_nee_end) | others;
                run_events = ((sel >> lane_id) & 1ull) != 0;
            }
        }
        // ---- events until this lane has a tracking or shadow segment under way, or no work left
        while (run_events && L.mode != kModeIdle && L.mode != kModeTrack && L.mode != kModeShadow) {
            if (L.mode == kModeWork) {
                if (lane_id >= lanes) {
                    L.mode = kModeIdle;
                    continue;
                }
                // the next pixel from the work counter, one atomic per wavefront
                const uint64_t ask = __ballot(1);
                const uint32_t leader = (uint32_t)__ffsll((long long)ask) - 1u;
                uint32_t base = 0;
                if (lane_id == leader) base = atomicAdd(&g_work_counter[slot], (uint32_t)__popcll(ask));
                base = __shfl(base, (int)leader, 64);
                const uint32_t w = base + (uint32_t)__popcll(ask & ((1ull << lane_id) - 1ull));
                if (w >= total) {
                    L.mode = kModeIdle;
                    continue;
                }
                if (p.work_order) {
                    const uint32_t lin = p.work_order[w];
                    if (lin >= W * H) continue;
                    L.lx = lin % W;
                    L.ly = lin / W;
                } else if (gen_data && w < n_first) {
                    L.lx = 4 * (w % first_nx) + first_rx;
                    L.ly = 4 * (w / first_nx) + first_ry;
                } else {
                    const uint32_t wt = gen_data ? w - n_first : w;
                    const uint32_t tile = wt >> 6, within = wt & 63u;
                    L.lx = (tile % tiles_x) * tile_w + within % tile_w;
                    L.ly = (tile / tiles_x) * tile_h + within / tile_w;
                    if (L.lx >= W || L.ly >= H) continue;
                    if (gen_data && (L.lx & 3u) == first_rx && (L.ly & 3u) == first_ry) continue;  // taken first
                }
                L.lin = L.ly * W + L.lx;
                if (redo && !stage.redo_lost[L.lin]) continue;  // all of its records were staged
                L.cost0 = med.lookups;
                L.scat0 = scatters;
                L.rng.state = redo ? stage.rng_save[L.lin] : p.rng_states[(size_t)L.ly * p.image_width + L.lx];
                L.sample = 0;
                if (gen_data) L.train_base = record_offsets ? record_offsets[L.lin] : 0u;
                L.mode = kModeStart;
            }
            if (L.mode == kModeStart) {
                // camera ray (optix_pathtracing_kernels.cu:43-55)
                float px = ((float)L.lx + L.rng.uni()) / (float)W;
                float py = ((float)L.ly + L.rng.uni()) / (float)H;
                const float vh = p.camera.vh, vw = p.camera.vw;
                const float sensor_area = vw * vh;
                L.org = V3{p.camera.position[0], p.camera.position[1], p.camera.position[2]};
                V3 ldir = normalize(V3{vw * (-0.5f + px), vh * (0.5f - py), -1.0f});
                const float* M = p.camera.orientation;
                V3 dir{dot(V3{M[0], M[1], M[2]}, ldir), dot(V3{M[3], M[4], M[5]}, ldir), dot(V3{M[6], M[7], M[8]}, ldir)};
                L.dir = normalize(dir);
                float az = ldir.z < 0.0f ? -ldir.z : ldir.z;
                L.dir_pd = 1.0f / (az * (az * az) * sensor_area);
                L.contribution = RGB{0.0f, 0.0f, 0.0f};
                float rinv = 1.0f / L.dir_pd;
                L.tp = RGB{rinv, rinv, rinv};
                L.init_importance = luminance(L.tp);
                L.path_length = 1;
                if (with_nrc) {
                    L.primary_spread = 0.0f;
                    L.cur_sqrt_spread = 0.0f;
                }
                if (gen_data) {
                    const uint32_t tl = (L.ly % 4u) * 4u + (L.lx % 4u);
                    L.unbiased = ((tl + nb.offset_to_select_unbiased_path) % 16u) == 0;
                    L.prev_idx = RTC8X_INVALID_VERTEX_INDEX;
                    L.prev_local_tp = RGB{0.0f, 0.0f, 0.0f};
                    L.records = 0;
                    L.lost = false;
                }
                L.mode = start_segment<TRACKING>(p.grid, med, sg, L.rng, L.org, L.dir, 1e10f) ? kModeTrack : kModeEscape;
            }
            bool after_nee = false, path_end = false;
            if (L.mode == kModeEscape) {
                // escaped to the environment (optix_pathtracing_kernels.cu:130-174)
                if (p.env.kind != RTC8X_ENV_NONE) {
                    float hyp_pd;
                    RGB radiance = env_evaluate(p.env, L.dir, hyp_pd);
                    float mis = 1.0f;
                    if (L.path_length > 1) {
                        float light_pd = prob_env * hyp_pd;
                        float bsdf_pd = L.dir_pd;
                        mis = bsdf_pd * bsdf_pd / (bsdf_pd * bsdf_pd + light_pd * light_pd);
                    }
                    const RGB implicit_cont = scale(radiance, mis);
                    if (!gen_data) L.contribution = add(L.contribution, mul(L.tp, implicit_cont));
                    if (gen_data && L.prev_idx != RTC8X_INVALID_VERTEX_INDEX) {
                        float* tgt = train_targets + 3 * (size_t)L.prev_idx;
                        const RGB a = mul(L.prev_local_tp, implicit_cont);
                        tgt[0] = tgt[0] + a.r;
                        tgt[1] = tgt[1] + a.g;
                        tgt[2] = tgt[2] + a.b;
                    }
                }
                if (gen_data) {
                    rtc8x_training_suffix_terminal ti;
                    ti.prev_vertex_index = L.prev_idx;
                    ti.has_query = 0;
                    ti.path_length = L.path_length & 0xFFu;
                    ti.is_unbiased = L.unbiased ? 1u : 0u;
                    nb.train_suffix_terminals[L.lin] = ti;
                } else if (with_nrc) {
                    rtc8x_terminal_info ti;
                    ti.throughput[0] = L.tp.r;
                    ti.throughput[1] = L.tp.g;
                    ti.throughput[2] = L.tp.b;
                    ti.has_query = (L.path_length & 0xFFu) << 8;
                    nb.inference_terminals[(size_t)L.sample * p_stride + L.lin] = ti;
                }
                path_end = true;
            } else if (L.mode == kModeNeeDone) {
                // end of the NEE shadow ray (performNextEventEstimation, renderer_kernel_common.h:297-403)
                RGB nee_c{0.0f, 0.0f, 0.0f};
                if (L.vis > 0.0f)
                    nee_c = L.nee_sun ? scale(L.nee_f, L.vis / L.nee_k1)
                                      : scale(L.nee_f, L.vis * L.nee_k1 * L.nee_k2 / L.nee_k3);
                if (!gen_data) L.contribution = add(L.contribution, mul(L.tp, nee_c));
                if (gen_data) L.nee_c = nee_c;
                after_nee = true;
            } else if (L.mode == kModeScatter) {
                // medium interaction at sg.t (optix_pathtracing_kernels.cu:176-268)
                scatters++;
                const float th = sg.t;
                L.pos = V3{L.org.x + th * L.dir.x, L.org.y + th * L.dir.y, L.org.z + th * L.dir.z};
                if (with_nrc) {
                    // path spread (optix_pathtracing_kernels.cu:200-210)
                    float ddx = L.pos.x - L.org.x, ddy = L.pos.y - L.org.y, ddz = L.pos.z - L.org.z;
                    float dist2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    if (L.path_length == 1)
                        L.primary_spread = dist2 / (4.0f * RTC8X_PI_F * 1.0f);
                    else
                        L.cur_sqrt_spread += sqrtf(dist2 / (L.dir_pd * 1.0f));
                }
                // Russian roulette (optix_pathtracing_kernels.cu:254-268)
                float continue_prob = 1.0f;
                bool rr_terminated = false;
                if (L.path_length > rr_after) {
                    float q = luminance(L.tp) / L.init_importance;
                    continue_prob = fminf(q, 1.0f);
                    if (L.rng.uni() >= continue_prob) rr_terminated = true;
                }
                if (!with_nrc && rr_terminated) path_end = true;
                if (with_nrc) {
                    // termination into the cache (optix_pathtracing_kernels.cu:282-330)
                    bool ends_with_cache = rr_terminated;
                    const float pt = nb.path_termination_factor;
                    if (gen_data) {
                        if (!L.unbiased)
                            ends_with_cache |= L.cur_sqrt_spread * L.cur_sqrt_spread > 2.0f * pt * L.primary_spread;
                    } else {
                        ends_with_cache |= L.cur_sqrt_spread * L.cur_sqrt_spread > pt * L.primary_spread;
                    }
                    if (ends_with_cache) {
                        const V3 vout{-L.dir.x, -L.dir.y, -L.dir.z};
                        make_query(p, L.pos, vout, &nb.inference_queries[(size_t)L.sample * q_stride + L.lin]);
                        if (gen_data) {
                            rtc8x_training_suffix_terminal ti;
                            ti.prev_vertex_index = L.prev_idx;
                            ti.has_query = 1;
                            ti.path_length = L.path_length & 0xFFu;
                            ti.is_unbiased = L.unbiased ? 1u : 0u;
                            nb.train_suffix_terminals[L.lin] = ti;
                        } else {
                            rtc8x_terminal_info ti;
                            ti.throughput[0] = L.tp.r;
                            ti.throughput[1] = L.tp.g;
                            ti.throughput[2] = L.tp.b;
                            ti.has_query = 1u | ((L.path_length & 0xFFu) << 8);
                            nb.inference_terminals[(size_t)L.sample * p_stride + L.lin] = ti;
                        }
                        path_end = true;
                    }
                }
                if (!path_end) {
                    L.tp = scale(L.tp, 1.0f / continue_prob);
                    // NEE set-up: light choice and sample, everything of the estimate except visibility
                    const bool have_env = p.env.kind != RTC8X_ENV_NONE, have_sun = p.sun.enabled != 0;
                    float u_light = L.rng.uni();
                    bool use_sun = false;
                    float prob_cur = 1.0f;
                    if (have_env && have_sun) {
                        if (u_light < p.sun.select_prob) {
                            use_sun = true;
                            prob_cur = p.sun.select_prob;
                        } else {
                            prob_cur = 1.0f - p.sun.select_prob;
                        }
                    } else if (have_sun) {
                        use_sun = true;
                    }
                    float u0 = L.rng.uni();
                    float u1 = L.rng.uni();
                    bool zero = !have_env && !have_sun;
                    V3 lpos{0.0f, 0.0f, 0.0f};
                    RGB emittance{0.0f, 0.0f, 0.0f};
                    float area_pd = 0.0f;
                    if (!zero) {
                        if (use_sun) {
                            lpos = V3{p.sun.direction[0], p.sun.direction[1], p.sun.direction[2]};
                            emittance = RGB{p.sun.irradiance[0], p.sun.irradiance[1], p.sun.irradiance[2]};
                        } else {
                            V3 ldir;
                            float ldir_pd;
                            RGB radiance = env_sample(p.env, u0, u1, ldir, ldir_pd);
                            emittance = scale(radiance, RTC8X_PI_F);
                            lpos = ldir;
                            area_pd = ldir_pd;
                            if (area_pd == 0.0f) zero = true;
                            area_pd *= prob_cur;
                        }
                    }
                    V3 sdir{-lpos.x, -lpos.y, -lpos.z};
                    float lp_cos = 0.0f;
                    if (!zero) {
                        V3 lnormal{-lpos.x, -lpos.y, -lpos.z};
                        lp_cos = dot(sdir, lnormal);
                        if (lp_cos <= 0.0f) zero = true;
                    }
                    if (zero) {
                        const RGB nee_c{0.0f, 0.0f, 0.0f};
                        if (!gen_data) L.contribution = add(L.contribution, mul(L.tp, nee_c));
                        if (gen_data) L.nee_c = nee_c;
                        after_nee = true;
                    } else {
                        const Frame fr = make_frame(V3{-L.dir.x, -L.dir.y, -L.dir.z});
                        V3 vin = to_local(fr, V3{-sdir.x, -sdir.y, -sdir.z});
                        float pval = phase_value(p.phase, p.scattering_forwardness, -vin.z);
                        float alb = p.scattering_albedo;
                        RGB fs{alb * pval, alb * pval, alb * pval};
                        L.nee_sun = use_sun;
                        if (use_sun) {
                            L.nee_f = mul(fs, emittance);
                            L.nee_k1 = prob_cur;
                        } else {
                            float bsdf_pd = pval * lp_cos / 1.0f;
                            if (!isfinite(bsdf_pd)) bsdf_pd = 0.0f;
                            float light_pd = area_pd;
                            L.nee_k1 = light_pd * light_pd / (bsdf_pd * bsdf_pd + light_pd * light_pd);
                            L.nee_f = mul(fs, scale(emittance, 1.0f / RTC8X_PI_F));
                            L.nee_k2 = lp_cos * 1.0f / 1.0f;
                            L.nee_k3 = area_pd;
                        }
                        const float trace_len = 2.0f * p.world_radius;
                        V3 lpoint{L.pos.x + trace_len * -sdir.x, L.pos.y + trace_len * -sdir.y,
                                  L.pos.z + trace_len * -sdir.z};
                        L.vis = 1.0f;
                        L.mode = start_segment<TRACKING>(p.grid, med, sg, L.rng, lpoint, sdir, trace_len) ? kModeShadow
                                                                                                        : kModeNeeDone;
                        continue;
                    }
                }
            }
            if (after_nee) {
                if (L.path_length >= max_len) {
                    path_end = true;
                } else {
                    // phase sampling (optix_pathtracing_kernels.cu:341-356)
                    float us0 = L.rng.uni();
                    float us1 = L.rng.uni();
                    V3 dl;
                    float pdf = phase_sample(p.phase, p.scattering_forwardness, us0, us1, dl);
                    if (pdf == 0.0f) {
                        path_end = true;
                    } else {
                        const V3 vout{-L.dir.x, -L.dir.y, -L.dir.z};
                        const Frame fr = make_frame(vout);
                        const float alb = p.scattering_albedo;
                        RGB fsv{alb * pdf, alb * pdf, alb * pdf};
                        L.dir = from_local(fr, dl);
                        const RGB local_tp = scale(fsv, 1.0f / pdf);
                        L.tp = mul(L.tp, local_tp);
                        L.org = L.pos;
                        L.dir_pd = pdf;
                        bool record_end = false;
                        if (gen_data) {
                            // training vertex record (optix_pathtracing_kernels.cu:358-410)
                            const uint32_t local = L.records;
                            L.records++;
                            rtc8x_radiance_query q;
                            make_query(p, L.pos, vout, &q);
                            // the record's slot: the next staging slot, or in pixel order (redo launch)
                            const uint32_t idx = staged ? atomicAdd(stage.counter, 1u) : L.train_base + local;
                            if (staged && idx >= kStageCapacity) {
                                // no staging slot left: the redo launch traces this pixel again
                                L.prev_idx = RTC8X_INVALID_VERTEX_INDEX;
                                L.lost = true;
                            } else if (staged || idx < RTC8X_TRAIN_BUFFER_SIZE) {
                                (staged ? stage.queries : nb.train_queries[0])[idx] = q;
                                rtc8x_training_vertex_info vi;
                                vi.local_throughput[0] = local_tp.r;
                                vi.local_throughput[1] = local_tp.g;
                                vi.local_throughput[2] = local_tp.b;
                                vi.prev_vertex_index = L.prev_idx;
                                vi.path_length = L.path_length & 0xFFu;
                                vi.is_unbiased = L.unbiased ? 1u : 0u;
                                if (staged) {
                                    stage.info[idx] = vi;
                                    stage.tag[2 * (size_t)idx] = L.lin;
                                    stage.tag[2 * (size_t)idx + 1] = local;
                                    StageSnapshot sn;
                                    sn.rng = L.rng.state;
                                    sn.path_length = L.path_length;
                                    sn.lookups = med.lookups - L.cost0;
                                    sn.scatters = scatters - L.scat0;
                                    sn.pad = 0;
                                    stage.snap[idx] = sn;
                                } else {
                                    nb.train_vertex_info[idx] = vi;
                                }
                                float* tgt = train_targets + 3 * (size_t)idx;
                                tgt[0] = L.nee_c.r;
                                tgt[1] = L.nee_c.g;
                                tgt[2] = L.nee_c.b;
                                L.prev_local_tp = local_tp;
                                L.prev_idx = idx;
                            } else {
                                nb.inference_queries[L.lin] = q;
                                rtc8x_training_suffix_terminal ti;
                                ti.prev_vertex_index = L.prev_idx;
                                ti.has_query = 1;
                                ti.path_length = L.path_length & 0xFFu;
                                ti.is_unbiased = L.unbiased ? 1u : 0u;
                                nb.train_suffix_terminals[L.lin] = ti;
                                record_end = true;
                            }
                        }
                        if (record_end) {
                            path_end = true;
                        } else {
                            ++L.path_length;
                            L.mode = start_segment<TRACKING>(p.grid, med, sg, L.rng, L.org, L.dir, 1e10f) ? kModeTrack
                                                                                                         : kModeEscape;
                            continue;
                        }
                    }
                }
            }
            if (path_end) {
                paths++;
                vertices += L.path_length;
                if (staged) record_counts[L.lin] = L.records;
                if (staged) {
                    stage.path_length[L.lin] = L.path_length;
                    stage.lookups[L.lin] = med.lookups - L.cost0;
                    stage.scatters[L.lin] = scatters - L.scat0;
                    stage.lost[L.lin] = L.lost ? 1u : 0u;
                }
                const bool ok = all_nonneg_finite(L.contribution);
                if (ENTRY == RTC8X_ENTRY_PATH_TRACE) {
                    if (ok) {
                        double* a = p.accum + 3 * (size_t)L.lin;
                        a[0] += (double)L.contribution.r;
                        a[1] += (double)L.contribution.g;
                        a[2] += (double)L.contribution.b;
                    }
                } else if (ENTRY == RTC8X_ENTRY_PATH_TRACE_WITH_NRC) {
                    float* pf = nb.per_frame_contribution + 3 * ((size_t)L.sample * p_stride + L.lin);
                    pf[0] = ok ? L.contribution.r : 0.0f;
                    pf[1] = ok ? L.contribution.g : 0.0f;
                    pf[2] = ok ? L.contribution.b : 0.0f;
                }
                if (p.path_vertices && !staged) p.path_vertices[L.lin] += L.path_length;
                if (++L.sample < num_samples) {
                    L.mode = kModeStart;
                } else {
                    p.rng_states[(size_t)L.ly * p.image_width + L.lx] = L.rng.state;
                    if (p.pixel_cost) p.pixel_cost[L.lin] = med.lookups - L.cost0;
                    L.mode = kModeWork;
                }
            }
        }
        // ---- tentative collisions, with their density lookups, for every stepping lane
        bool coop = false;
        if constexpr (TRACKING == RTC8X_TRACKING_GLOBAL_MAJORANT) {
            const uint64_t steppers = __ballot(L.mode == kModeTrack || L.mode == kModeShadow);
            coop = steppers != 0 && __popcll(__ballot(L.mode == kModeIdle)) >= kCoopIdle;
            if (coop) coop_segments(p.grid, med, sg, L, steppers, lane_id);
        } else if constexpr (ENTRY == RTC8X_ENTRY_GENERATE_TRAINING_DATA) {
            // a local path's 64 collisions cost one walk over their regions (tens of microseconds), while
            // the stepping lanes walk side by side: only the last few paths are taken this way
            const uint64_t steppers = __ballot(L.mode == kModeTrack || L.mode == kModeShadow);
            coop = steppers != 0 && __popcll(steppers) <= kCoopLocalPaths &&
                   __popcll(__ballot(L.mode == kModeIdle)) >= kCoopIdle;
            if (coop) coop_segments_local(p.grid, med, sg, L, steppers, lane_id, region_table);
        }
        if (coop) {
        } else if constexpr (TRACKING == RTC8X_TRACKING_GLOBAL_MAJORANT) {
            // one copy of the lookup code serves both segment kinds: the kernel's instruction footprint
            // matters (a copy per kind made the full-image launch half again as slow)
            if (L.mode == kModeTrack || L.mode == kModeShadow)
                speculate_global<kSpeculate>(p.grid, med, sg, L, L.mode == kModeTrack);
        } else if (L.mode == kModeTrack || L.mode == kModeShadow) {
            const bool track = L.mode == kModeTrack;
            const int r = sg.advance(p.grid, med, L.rng, mu, region_table);
            if (r == kAdvanceCollision) {
                // A speculative pair: the numbers the sequential loop draws after a null collision at
                // the first tentative collision (its acceptance number, then the next optical depth) are
                // drawn now, the walk goes on to the second, and both lookups are issued together.  The
                // results are consumed in order: an acceptance at the first rewinds the PCG state to
                // just after its acceptance number and drops the second (not counted as a lookup).
                const float t_first = sg.t, mu_first = mu;
                float u_first = 0.0f;
                uint64_t st_first = 0;
                if (track) {
                    u_first = L.rng.uni();
                    st_first = L.rng.state;
                }
                sg.tau = -rtc8x_logf(1.0f - L.rng.uni());
                float mu_second = 0.0f;
                const int r2 = sg.advance(p.grid, med, L.rng, mu_second, region_table);
                const float tk[2] = {t_first, r2 == kAdvanceCollision ? sg.t : t_first};
                float sig[2];
                density_batch<2>(p.grid, med.coeff, sg.o, sg.d, tk, sig);
                med.lookups++;
                if (track && u_first < sig[0] / mu_first) {
                    sg.t = t_first;
                    L.rng.state = st_first;
                    L.mode = kModeScatter;
                } else {
                    if (!track) L.vis *= 1.0f - sig[0] / mu_first;
                    if (r2 == kAdvanceCollision) {
                        med.lookups++;
                        if (track) {
                            if (sg.after_track(L.rng, sig[1], mu_second)) L.mode = kModeScatter;
                        } else {
                            sg.after_shadow(L.rng, sig[1], mu_second, L.vis);
                        }
                    } else if (r2 == kAdvanceEnd) {
                        L.mode = track ? kModeEscape : kModeNeeDone;
                    }
                }
            } else if (r == kAdvanceEnd) {
                L.mode = L.mode == kModeTrack ? kModeEscape : kModeNeeDone;
            }
        }
    }
    uint64_t* const stats = staged ? nullptr : p.stats;  // staged: from the per-pixel sums
    if (stats) {
        uint64_t v[4] = {paths, vertices, med.lookups, scatters};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            uint64_t x = v[i];
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
            v[i] = x;
        }
        if (lane_id == 0) {
            atomicAdd((unsigned long long*)&stats[RTC8X_STAT_PATHS], (unsigned long long)v[0]);
            atomicAdd((unsigned long long*)&stats[RTC8X_STAT_VERTICES], (unsigned long long)v[1]);
            atomicAdd((unsigned long long*)&stats[RTC8X_STAT_LOOKUPS], (unsigned long long)v[2]);
            atomicAdd((unsigned long long*)&stats[RTC8X_STAT_SCATTER], (unsigned long long)v[3]);
        }
    }
}

// Work counters of a launch: a block of kMaxBatch slots, taken in turn by successive launches. Launches
// on different streams count in different slots as long as fewer than kCounterSlots / kMaxBatch of
// them are in flight at once; the slots wrap after that many launches, so a caller that keeps more
// launches in flight concurrently must order them on one stream.
inline uint32_t take_counter_slots()
{
    static std::atomic<uint32_t> next{0};
    return (next.fetch_add(1) % (kCounterSlots / kMaxBatch)) * kMaxBatch;
}

// One launch of one image.
// Local tracking keeps a region walk per lane on top of the lane state: the render entries are held to
// three wavefronts per SIMD (168 VGPRs, a few spilled); the training entry, whose lane state is larger,
// would spill too much and keeps its registers.
template <int ENTRY, int TRACKING>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    TRACKING == RTC8X_TRACKING_LOCAL_MAJORANT && ENTRY != RTC8X_ENTRY_GENERATE_TRAINING_DATA ? 3 : 1)))
void wavefront_kernel(const rtc8x_frame_params p, uint32_t launch_w,
                                                        uint32_t launch_h, uint32_t num_samples, uint32_t slot,
                                                        const uint32_t* __restrict__ record_offsets,
                                                        uint32_t* __restrict__ record_counts, const TrainStage stage)
{
    if constexpr (ENTRY == RTC8X_ENTRY_GENERATE_TRAINING_DATA) {
        // every record found a staging slot: no pixel for the redo launch (a walk over the training image
        // that skips every pixel took about 0.1 ms)
        if (stage.redo_lost != nullptr && *stage.staged <= kStageCapacity) return;
    }
    if constexpr (TRACKING == RTC8X_TRACKING_LOCAL_MAJORANT) {
        // the region majorants staged in LDS: every DDA step reads them there
        __shared__ __align__(16) uint8_t regions[RTC8X_MAX_REGIONS];
        const uint32_t n16 = ((uint32_t)p.grid.region_dims[0] * (uint32_t)p.grid.region_dims[1] *
                                  (uint32_t)p.grid.region_dims[2] + 15u) / 16u;
        const uint4* src = reinterpret_cast<const uint4*>(p.grid.region_majorants);
        for (uint32_t i = threadIdx.x; i < n16; i += blockDim.x) reinterpret_cast<uint4*>(regions)[i] = src[i];
        __syncthreads();
        trace_body<ENTRY, TRACKING>(p, launch_w, launch_h, num_samples, slot, record_offsets, record_counts, stage,
                                    regions);
    } else {
        trace_body<ENTRY, TRACKING>(p, launch_w, launch_h, num_samples, slot, record_offsets, record_counts, stage,
                                    nullptr);
    }
}

// Several images (instances: their own volume, camera, sky, buffers and, with the NRC entry, cache
// buffers) in one persistent launch.  Workgroup b starts on the instance whose workgroup range
// [wg_end[i-1], wg_end[i]) holds b (the ranges are the caller's shares of the device,
// rtc8x_trace_batch) and takes pixels from that instance's work counter; a wavefront that has
// finished its instance's last pixels moves on to the following instances and helps with what is left
// of them.  The instance is uniform over a wavefront, so its frame parameters are read with scalar
// loads, as a single launch reads its kernel arguments.
struct BatchArgs {
    const rtc8x_frame_params* params;  // device copy, count entries
    uint32_t count, slot;              // instance i counts its work at g_work_counter[slot + i]
    uint32_t wg_end[kMaxBatch];
    uint32_t num_samples[kMaxBatch];
};

template <int ENTRY, int TRACKING>
__global__ __launch_bounds__(256) void batch_kernel(const BatchArgs b)
{
    uint32_t first = 0;
    while (first + 1 < b.count && blockIdx.x >= b.wg_end[first]) ++first;
    for (uint32_t k = 0; k < b.count; ++k) {
        const uint32_t inst = first + k < b.count ? first + k : first + k - b.count;
        const rtc8x_frame_params* __restrict__ pp = b.params + inst;
        trace_body<ENTRY, TRACKING>(*pp, pp->image_width, pp->image_height, b.num_samples[inst], b.slot + inst,
                                    nullptr, nullptr, TrainStage{}, pp->grid.region_majorants);
    }
}

__global__ void reset_counters_kernel(uint32_t slot, uint32_t count)
{
    if (threadIdx.x < count) g_work_counter[slot + threadIdx.x] = 0u;
}

// Exclusive prefix sum of the per-path record counts, in tiles of kPrefixTile pixels: record_tile_sum_kernel
// sums each tile, record_tile_scan_kernel adds the sums of the tiles before its own and scans the tile
// (16 counts per lane).  Sums are 64-bit; offsets and the total saturate at 2^32 - 1.
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t* lds4)
{
    for (int o = 32; o > 0; o >>= 1) v += (uint64_t)__shfl_xor((unsigned long long)v, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63u) == 0) lds4[threadIdx.x >> 6] = v;
    __syncthreads();
    return lds4[0] + lds4[1] + lds4[2] + lds4[3];
}

__global__ __launch_bounds__(256) void record_tile_sum_kernel(const rtc8x_frame_params p,
                                                              const uint32_t* __restrict__ counts,
                                                              uint64_t* __restrict__ tile_sums)
{
    __shared__ uint64_t lds4[4];
    uint32_t tw, th;
    train_dims(p, tw, th);
    const uint32_t n = tw * th, b0 = blockIdx.x * kPrefixTile;
    if (b0 >= n) return;  // uniform over the workgroup
    const uint32_t e = min(n, b0 + kPrefixTile);
    uint64_t s = 0;
    for (uint32_t i = b0 + threadIdx.x; i < e; i += 256) s += counts[i];
    s = block_sum_u64(s, lds4);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void record_tile_scan_kernel(const rtc8x_frame_params p,
                                                               const uint32_t* __restrict__ counts,
                                                               const uint64_t* __restrict__ tile_sums,
                                                               uint32_t* __restrict__ offsets,
                                                               uint32_t* __restrict__ total_out)
{
    __shared__ uint64_t lds4[4];
    __shared__ uint64_t wave_tot[4];
    uint32_t tw, th;
    train_dims(p, tw, th);
    const uint32_t n = tw * th, b0 = blockIdx.x * kPrefixTile;
    if (b0 >= n) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *total_out = 0u;  // no pixels
        return;
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint64_t before = 0;
    for (uint32_t j = threadIdx.x; j < blockIdx.x; j += 256) before += tile_sums[j];
    const uint64_t base = block_sum_u64(before, lds4);
    // lane t: counts [b0 + 16 t, b0 + 16 t + 16)
    uint32_t c[16];
    uint64_t mine = 0;
    const uint32_t i0 = b0 + 16u * threadIdx.x;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        c[k] = i0 + k < n ? counts[i0 + k] : 0u;
        mine += c[k];
    }
    // exclusive scan of the lanes' sums: within the wavefront, then over the four wavefronts
    uint64_t incl = mine;
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t v = (uint64_t)__shfl_up((unsigned long long)incl, o, 64);
        if (lane >= (uint32_t)o) incl += v;
    }
    if (lane == 63) wave_tot[wave] = incl;
    __syncthreads();
    uint64_t run = base + (incl - mine);
    for (uint32_t w = 0; w < wave; ++w) run += wave_tot[w];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (i0 + k < n) offsets[i0 + k] = run > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)run;
        run += c[k];
    }
    if (b0 + kPrefixTile >= n && threadIdx.x == 255) {
        const uint64_t tot = base + wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
        *total_out = tot > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)tot;
    }
}

// Final slot of staged record j: its pixel's offset plus its index on the path.
__device__ __forceinline__ uint32_t staged_slot(const TrainStage& st, const uint32_t* __restrict__ offsets, uint32_t j)
{
    return offsets[st.tag[2 * (size_t)j]] + st.tag[2 * (size_t)j + 1];
}

// Moves the staged records to their final slots (vertex links renamed), and finishes every pixel whose
// records were all staged: a pixel whose records all fit keeps its path's end (its suffix terminal's
// link renamed, its statistics added); a pixel whose records run past the buffer ends at its first
// record past it (pixel-order slot max(offset, capacity)), as in recordGeneration's full-buffer branch:
// that record's query becomes the suffix query, and the PCG state, path length and statistics are the
// record's snapshot.  Pixels with a lost record are left to the redo launch.  Slots are computed where
// they are needed (a kernel writing a slot map first took 5 us per training launch).
__global__ __launch_bounds__(256) void stage_copy_kernel(const rtc8x_frame_params p, const uint32_t* __restrict__ total,
                                                         const TrainStage st, const uint32_t* __restrict__ offsets,
                                                         const uint32_t* __restrict__ counts)
{
    const rtc8x_nrc_buffers& nb = p.nrc;
    uint32_t tw, th;
    train_dims(p, tw, th);
    const uint32_t npx = tw * th;
    const uint32_t n = (uint32_t)min((size_t)*total, kStageCapacity);
    const uint32_t stride = gridDim.x * blockDim.x;
    uint64_t sum[4] = {0, 0, 0, 0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t f = staged_slot(st, offsets, i);
        if (f < RTC8X_TRAIN_BUFFER_SIZE) {
            nb.train_queries[0][f] = st.queries[i];
            rtc8x_training_vertex_info vi = st.info[i];
            if (vi.prev_vertex_index != RTC8X_INVALID_VERTEX_INDEX)
                vi.prev_vertex_index = staged_slot(st, offsets, vi.prev_vertex_index);
            nb.train_vertex_info[f] = vi;
#pragma unroll
            for (int c = 0; c < 3; ++c) nb.train_targets[0][3 * (size_t)f + c] = st.targets[3 * (size_t)i + c];
            continue;
        }
        const uint32_t pix = st.tag[2 * (size_t)i], local = st.tag[2 * (size_t)i + 1];
        if (st.lost[pix] || f != max(offsets[pix], RTC8X_TRAIN_BUFFER_SIZE)) continue;
        // the cut: this record's vertex ends the path with a cache query
        const rtc8x_training_vertex_info vi = st.info[i];
        const StageSnapshot sn = st.snap[i];
        nb.inference_queries[pix] = st.queries[i];
        rtc8x_training_suffix_terminal ti;
        ti.prev_vertex_index = local ? f - 1u : RTC8X_INVALID_VERTEX_INDEX;
        ti.has_query = 1;
        ti.path_length = vi.path_length;
        ti.is_unbiased = vi.is_unbiased;
        nb.train_suffix_terminals[pix] = ti;
        p.rng_states[(size_t)(pix / tw) * p.image_width + pix % tw] = sn.rng;
        if (p.path_vertices) p.path_vertices[pix] += sn.path_length;
        sum[0] += 1;
        sum[1] += sn.path_length;
        sum[2] += sn.lookups;
        sum[3] += sn.scatters;
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npx; i += stride) {
        if (st.lost[i] || (uint64_t)offsets[i] + counts[i] > RTC8X_TRAIN_BUFFER_SIZE) continue;
        const uint32_t prev = nb.train_suffix_terminals[i].prev_vertex_index;
        if (prev != RTC8X_INVALID_VERTEX_INDEX) nb.train_suffix_terminals[i].prev_vertex_index = staged_slot(st, offsets, prev);
        if (p.path_vertices) p.path_vertices[i] += st.path_length[i];
        sum[0] += 1;
        sum[1] += st.path_length[i];
        sum[2] += st.lookups[i];
        sum[3] += st.scatters[i];
    }
    if (p.stats) {
        const uint32_t lane = threadIdx.x & 63;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint64_t x = sum[q];
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
            if (lane == 0 && x) atomicAdd((unsigned long long*)&p.stats[q], (unsigned long long)x);
        }
    }
}

// The training pixels' PCG states before the launch, in launch-pixel order (the redo launch's start).
__global__ __launch_bounds__(256) void save_rng_kernel(const rtc8x_frame_params p, uint64_t* __restrict__ out)
{
    uint32_t tw, th;
    train_dims(p, tw, th);
    const uint32_t n = tw * th;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = p.rng_states[(size_t)(i / tw) * p.image_width + i % tw];
}

// ---------------------------------------------------------------- cost-ordered work lists
// A stable counting sort of the pixels by class = 31 - floor(log2(cost + 1)) (costliest first):
// per-block class histograms, one exclusive scan over (class, block), and a stable scatter.
constexpr uint32_t kOrderBlock = 1024;  // pixels per workgroup (256 lanes x 4)
constexpr uint32_t kClasses = 32;

__device__ __forceinline__ uint32_t cost_class(uint32_t c)
{
    return (uint32_t)__clz((c < 0xFFFFFFFEu ? c : 0xFFFFFFFEu) + 1u);  // 0 (costliest) .. 31 (no lookups)
}

__global__ __launch_bounds__(256) void order_hist_kernel(const uint32_t* __restrict__ cost, uint32_t n,
                                                         uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[kClasses];
    if (threadIdx.x < kClasses) h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t b0 = blockIdx.x * kOrderBlock;
    for (uint32_t k = 0; k < kOrderBlock / 256; ++k) {
        const uint32_t i = b0 + k * 256 + threadIdx.x;
        if (i < n) atomicAdd(&h[cost_class(cost[i])], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kClasses) hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of hist[class][block] in class-major order, in place (one workgroup).
__global__ __launch_bounds__(1024) void order_scan_kernel(uint32_t* __restrict__ hist, uint32_t m)
{
    __shared__ uint32_t part[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t chunk = (m + 1023) / 1024;
    const uint32_t b = tid * chunk, e = min(m, b + chunk);
    uint32_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += hist[i];
    part[tid] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint32_t v = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint32_t run = part[tid] - s;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t c = hist[i];
        hist[i] = run;
        run += c;
    }
}

// Stable scatter: within a block, pixels keep their index order inside each class.
__global__ __launch_bounds__(256) void order_scatter_kernel(const uint32_t* __restrict__ cost, uint32_t n,
                                                            const uint32_t* __restrict__ offsets,
                                                            uint32_t* __restrict__ order)
{
    __shared__ uint32_t base[kClasses];
    __shared__ uint32_t cnt[kClasses][256];
    const uint32_t t = threadIdx.x;
    const uint32_t b0 = blockIdx.x * kOrderBlock + t * (kOrderBlock / 256);  // 4 consecutive pixels per lane
    uint32_t cls[kOrderBlock / 256];
    for (uint32_t c = 0; c < kClasses; ++c) cnt[c][t] = 0;
    for (uint32_t k = 0; k < kOrderBlock / 256; ++k) {
        const uint32_t i = b0 + k;
        cls[k] = i < n ? cost_class(cost[i]) : kClasses;
        if (cls[k] < kClasses) cnt[cls[k]][t] += 1;
    }
    if (t < kClasses) base[t] = offsets[t * gridDim.x + blockIdx.x];
    __syncthreads();
    // exclusive scan of cnt[c][0..255] for every class (one lane per class)
    if (t < kClasses) {
        uint32_t run = 0;
        for (uint32_t j = 0; j < 256; ++j) {
            const uint32_t v = cnt[t][j];
            cnt[t][j] = run;
            run += v;
        }
    }
    __syncthreads();
    for (uint32_t k = 0; k < kOrderBlock / 256; ++k) {
        if (cls[k] >= kClasses) continue;
        uint32_t before = 0;
        for (uint32_t q = 0; q < k; ++q) before += cls[q] == cls[k] ? 1u : 0u;
        order[base[cls[k]] + cnt[cls[k]][t] + before] = b0 + k;
    }
}

// Tile-grouped work order: the pixels' costs summed per kOrderTile^2 tile (saturating), the tiles
// sorted by quarter-octave class below the costliest tile (stable, with the pixels' counting sort
// above), then each tile's pixels written row-major at its offset
// (an exclusive scan of the tiles' pixel counts in sorted order; edge tiles hold fewer pixels).  A
// wavefront's lanes then take the pixels of one tile, whose paths cross nearby voxels.
constexpr uint32_t kOrderTile = 8;

__device__ __forceinline__ void tile_rect(uint32_t t, uint32_t width, uint32_t height, uint32_t& x0, uint32_t& y0,
                                          uint32_t& ew, uint32_t& eh)
{
    const uint32_t tx = (width + kOrderTile - 1) / kOrderTile;
    x0 = (t % tx) * kOrderTile;
    y0 = (t / tx) * kOrderTile;
    ew = min(kOrderTile, width - x0);
    eh = min(kOrderTile, height - y0);
}

// one wavefront per tile
__global__ __launch_bounds__(256) void tile_cost_kernel(const uint32_t* __restrict__ cost, uint32_t width,
                                                        uint32_t height, uint32_t n_tiles,
                                                        uint32_t* __restrict__ tile_cost)
{
    const uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
    if (t >= n_tiles) return;
    uint32_t x0, y0, ew, eh;
    tile_rect(t, width, height, x0, y0, ew, eh);
    const uint32_t dx = lane % kOrderTile, dy = lane / kOrderTile;
    uint64_t c = (dx < ew && dy < eh) ? cost[(size_t)(y0 + dy) * width + x0 + dx] : 0u;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane == 0) tile_cost[t] = c < 0xFFFFFFFFull ? (uint32_t)c : 0xFFFFFFFFu;
}

// Quarter-octave classes for the tiles, counted down from the costliest tile: class k holds the tiles
// whose cost lies k quarter octaves below the maximum (k <= 30), class 31 the cheaper ones and the
// tiles without lookups.  Each tile's cost is replaced by 2^(31 - k) - 1, whose cost_class is k, so
// that the pixels' counting sort orders the tiles.  One workgroup.
__device__ __forceinline__ uint32_t quarter_octave(uint32_t c)  // c >= 1; monotone in c
{
    const uint32_t e = 31u - (uint32_t)__clz(c);
    const uint32_t m = e >= 2 ? (c >> (e - 2)) & 3u : (c << (2 - e)) & 3u;
    return 4u * e + m;
}

__global__ __launch_bounds__(1024) void tile_class_kernel(uint32_t* __restrict__ tile_cost, uint32_t n_tiles)
{
    __shared__ uint32_t part[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    uint32_t m = 0;
    for (uint32_t i = tid; i < n_tiles; i += 1024) m = max(m, tile_cost[i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
    if (lane == 0) part[tid >> 6] = m;
    __syncthreads();
    m = 0;
    for (uint32_t w = 0; w < 16; ++w) m = max(m, part[w]);
    if (m == 0) return;  // no lookups anywhere: every tile keeps cost 0 (class 31)
    const uint32_t top = quarter_octave(m);
    for (uint32_t i = tid; i < n_tiles; i += 1024) {
        const uint32_t c = tile_cost[i];
        const uint32_t k = c == 0 ? 31u : min(31u, top - quarter_octave(c));
        tile_cost[i] = (1u << (31u - k)) - 1u;
    }
}

__global__ __launch_bounds__(256) void tile_count_kernel(const uint32_t* __restrict__ tile_rank, uint32_t width,
                                                         uint32_t height, uint32_t n_tiles,
                                                         uint32_t* __restrict__ counts)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_tiles) return;
    uint32_t x0, y0, ew, eh;
    tile_rect(tile_rank[r], width, height, x0, y0, ew, eh);
    counts[r] = ew * eh;
}

// one wavefront per tile rank; offsets[r]: exclusive scan of the counts in rank order
__global__ __launch_bounds__(256) void tile_expand_kernel(const uint32_t* __restrict__ tile_rank,
                                                          const uint32_t* __restrict__ offsets, uint32_t width,
                                                          uint32_t height, uint32_t n_tiles,
                                                          uint32_t* __restrict__ order)
{
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
    if (r >= n_tiles) return;
    uint32_t x0, y0, ew, eh;
    tile_rect(tile_rank[r], width, height, x0, y0, ew, eh);
    if (lane < ew * eh) order[offsets[r] + lane] = (y0 + lane / ew) * width + x0 + lane % ew;
}

// One persistent launch: as many 256-lane workgroups as fit on the device at once, a fresh work
// counter slot (cleared by a one-lane kernel on the same stream; launches on different streams take
// different slots).
template <int ENTRY, int TRACKING>
hipError_t launch(const rtc8x_frame_params& p, uint32_t w, uint32_t h, uint32_t num_samples, const uint32_t* offsets,
                  uint32_t* counts, const TrainStage& stage, hipStream_t stream)
{
    const uint32_t slot = take_counter_slots();
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, wavefront_kernel<ENTRY, TRACKING>, 256, 0);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, stream, slot);
    hipLaunchKernelGGL((wavefront_kernel<ENTRY, TRACKING>), dim3((uint32_t)(cus * per_cu)), dim3(256), 0,
                       stream, p, w, h, num_samples, slot, offsets, counts, stage);
    return hipGetLastError();
}

// One persistent launch over several instances; shares[i] is instance i's part of the device's
// workgroups (each gets at least one).
template <int ENTRY, int TRACKING>
hipError_t launch_batch(const rtc8x_frame_params* d_params, const uint32_t* num_samples, const double* shares,
                        uint32_t count, hipStream_t stream)
{
    const uint32_t slot = take_counter_slots();
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<ENTRY, TRACKING>, 256, 0);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    const uint32_t total = std::max<uint32_t>((uint32_t)(cus * per_cu), count);
    BatchArgs b{};
    b.params = d_params;
    b.count = count;
    b.slot = slot;
    double sum = 0.0;
    for (uint32_t i = 0; i < count; ++i) sum += shares[i];
    uint32_t end = 0;
    double acc = 0.0;
    for (uint32_t i = 0; i < count; ++i) {
        acc += shares[i];
        // cumulative rounding keeps the ranges summing to `total`; every instance keeps one workgroup
        uint32_t e_i = (uint32_t)std::llround(acc / sum * (total - count)) + i + 1;
        e_i = std::max(e_i, end + 1);
        b.wg_end[i] = end = std::min(e_i, total - (count - 1 - i));
        b.num_samples[i] = num_samples[i];
    }
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(kMaxBatch), 0, stream, slot, count);
    hipLaunchKernelGGL((batch_kernel<ENTRY, TRACKING>), dim3(end), dim3(256), 0, stream, b);
    return hipGetLastError();
}

template <int ENTRY>
hipError_t launch_tracking(const rtc8x_frame_params& p, uint32_t w, uint32_t h, uint32_t ns, const uint32_t* offsets,
                           uint32_t* counts, hipStream_t stream, const TrainStage& stage = TrainStage{})
{
    if (p.tracking == RTC8X_TRACKING_LOCAL_MAJORANT)
        return launch<ENTRY, RTC8X_TRACKING_LOCAL_MAJORANT>(p, w, h, ns, offsets, counts, stage, stream);
    return launch<ENTRY, RTC8X_TRACKING_GLOBAL_MAJORANT>(p, w, h, ns, offsets, counts, stage, stream);
}

}  // namespace dev
}  // namespace rtc8x

using namespace rtc8x;

extern "C" size_t rtc8x_order_scratch_bytes(uint32_t n)
{
    const size_t blocks = ((size_t)n + dev::kOrderBlock - 1) / dev::kOrderBlock;
    return blocks * dev::kClasses * sizeof(uint32_t);
}

extern "C" rtc8x_status rtc8x_order_by_cost(void* stream, const uint32_t* cost, uint32_t n, uint32_t* order,
                                            void* scratch)
try {
    RTC8X_REQUIRE(cost && order && scratch, "rtc8x_order_by_cost: NULL argument");
    if (n == 0) return RTC8X_OK;
    const uint32_t blocks = (n + dev::kOrderBlock - 1) / dev::kOrderBlock;
    hipStream_t s = (hipStream_t)stream;
    uint32_t* hist = (uint32_t*)scratch;
    hipLaunchKernelGGL(dev::order_hist_kernel, dim3(blocks), dim3(256), 0, s, cost, n, hist);
    RTC8X_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::order_scan_kernel, dim3(1), dim3(1024), 0, s, hist, blocks * dev::kClasses);
    RTC8X_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::order_scatter_kernel, dim3(blocks), dim3(256), 0, s, cost, n, hist, order);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_order_by_cost")

namespace {
struct TileOrderLayout {
    uint32_t n_tiles;
    size_t tile_cost, tile_rank, counts, hist, bytes;  // byte offsets into the scratch buffer
};
TileOrderLayout tile_order_layout(uint32_t width, uint32_t height)
{
    TileOrderLayout l{};
    const uint64_t tiles = (uint64_t)((width + dev::kOrderTile - 1) / dev::kOrderTile) *
                           ((height + dev::kOrderTile - 1) / dev::kOrderTile);
    l.n_tiles = (uint32_t)std::min<uint64_t>(tiles, 0xFFFFFFFFull);
    const size_t a = ((size_t)l.n_tiles * 4 + 255) & ~(size_t)255;
    l.tile_cost = 0;
    l.tile_rank = a;
    l.counts = 2 * a;
    l.hist = 3 * a;
    l.bytes = l.hist + std::max<size_t>(rtc8x_order_scratch_bytes(l.n_tiles), 4);
    return l;
}
}  // namespace

extern "C" size_t rtc8x_order_tiles_scratch_bytes(uint32_t width, uint32_t height)
{
    return tile_order_layout(width, height).bytes;
}

extern "C" rtc8x_status rtc8x_order_by_tile_cost(void* stream, const uint32_t* cost, uint32_t width, uint32_t height,
                                                 uint32_t* order, void* scratch)
try {
    RTC8X_REQUIRE(cost && order && scratch, "rtc8x_order_by_tile_cost: NULL argument");
    RTC8X_REQUIRE((uint64_t)width * height <= 0xFFFFFFFFull, "rtc8x_order_by_tile_cost: image too large");
    if (width == 0 || height == 0) return RTC8X_OK;
    const TileOrderLayout l = tile_order_layout(width, height);
    char* sb = (char*)scratch;
    uint32_t* tile_cost = (uint32_t*)(sb + l.tile_cost);
    uint32_t* tile_rank = (uint32_t*)(sb + l.tile_rank);
    uint32_t* counts = (uint32_t*)(sb + l.counts);
    hipStream_t s = (hipStream_t)stream;
    const uint32_t wave_blocks = (l.n_tiles + 3) / 4;
    hipLaunchKernelGGL(dev::tile_cost_kernel, dim3(wave_blocks), dim3(256), 0, s, cost, width, height, l.n_tiles,
                       tile_cost);
    RTC8X_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::tile_class_kernel, dim3(1), dim3(1024), 0, s, tile_cost, l.n_tiles);
    RTC8X_HIP(hipGetLastError());
    rtc8x_status st = rtc8x_order_by_cost(stream, tile_cost, l.n_tiles, tile_rank, sb + l.hist);
    if (st != RTC8X_OK) return st;
    hipLaunchKernelGGL(dev::tile_count_kernel, dim3((l.n_tiles + 255) / 256), dim3(256), 0, s, tile_rank, width,
                       height, l.n_tiles, counts);
    RTC8X_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::order_scan_kernel, dim3(1), dim3(1024), 0, s, counts, l.n_tiles);
    RTC8X_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::tile_expand_kernel, dim3(wave_blocks), dim3(256), 0, s, tile_rank, counts, width, height,
                       l.n_tiles, order);
    RTC8X_HIP(hipGetLastError());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_order_by_tile_cost")

static bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

namespace {

// The checks every launch of an entry point makes on its frame parameters.
rtc8x_status check_launch(const rtc8x_frame_params& p, int32_t entry, uint32_t num_samples)
{
    RTC8X_REQUIRE(entry >= RTC8X_ENTRY_PATH_TRACE && entry <= RTC8X_ENTRY_PATH_TRACE_WITH_NRC,
                  "rtc8x_trace: unknown entry %d", entry);
    RTC8X_REQUIRE(p.image_width > 0 && p.image_height > 0 && p.image_width <= 16384 && p.image_height <= 16384,
                  "rtc8x_trace: bad image size %ux%u", p.image_width, p.image_height);
    RTC8X_REQUIRE(p.rng_states != nullptr, "rtc8x_trace: rng_states is NULL");
    RTC8X_REQUIRE(num_samples >= 1, "rtc8x_trace: num_samples must be >= 1");
    RTC8X_REQUIRE(entry != RTC8X_ENTRY_GENERATE_TRAINING_DATA || num_samples == 1,
                  "rtc8x_trace: GENERATE_TRAINING_DATA traces one sample per launch");
    RTC8X_REQUIRE(entry != RTC8X_ENTRY_PATH_TRACE_WITH_NRC || num_samples == (p.nrc.num_samples ? p.nrc.num_samples : 1u),
                  "rtc8x_trace: PATH_TRACE_WITH_NRC traces nrc.num_samples (%u) samples per launch, not %u",
                  p.nrc.num_samples ? p.nrc.num_samples : 1u, num_samples);
    RTC8X_REQUIRE(entry != RTC8X_ENTRY_PATH_TRACE || p.accum != nullptr, "rtc8x_trace: accum is NULL");
    RTC8X_REQUIRE(finite3(p.camera.position) && std::isfinite(p.camera.vh) && std::isfinite(p.camera.vw),
                  "rtc8x_trace: camera is not finite");
    RTC8X_REQUIRE(p.phase == RTC8X_PHASE_SCHLICK || p.phase == RTC8X_PHASE_HENYEY_GREENSTEIN,
                  "rtc8x_trace: unknown phase function %d", p.phase);
    RTC8X_REQUIRE(std::fabs(p.scattering_forwardness) < 1.0f, "rtc8x_trace: |forwardness| must be < 1");
    RTC8X_REQUIRE(p.tracking == RTC8X_TRACKING_GLOBAL_MAJORANT || p.tracking == RTC8X_TRACKING_LOCAL_MAJORANT,
                  "rtc8x_trace: unknown tracking mode %d", p.tracking);
    RTC8X_REQUIRE(std::isfinite(p.world_radius) && p.world_radius >= 0.0f, "rtc8x_trace: bad world radius");
    const rtc8x_grid& g = p.grid;
    const float majorant = p.density_coeff * g.max_value;
    if (majorant > 0.0f) {
        RTC8X_REQUIRE(std::isfinite(majorant), "rtc8x_trace: majorant is not finite");
        RTC8X_REQUIRE(g.num_root_tiles > 0 && g.root_tiles && g.upper_tables && g.tile_values,
                      "rtc8x_trace: grid tables missing");
        RTC8X_REQUIRE(finite3(g.world_bbox_min) && finite3(g.world_bbox_max), "rtc8x_trace: grid bbox not finite");
        // Free-flight loops take about majorant * chord steps; refuse inputs whose loops cannot end in
        // reasonable time (the chord is bounded by the bbox diagonal).
        double dx = g.world_bbox_max[0] - g.world_bbox_min[0], dy = g.world_bbox_max[1] - g.world_bbox_min[1],
               dz = g.world_bbox_max[2] - g.world_bbox_min[2];
        double diag = std::sqrt(dx * dx + dy * dy + dz * dz);
        RTC8X_REQUIRE((double)majorant * diag < 1.0e6, "rtc8x_trace: majorant x bbox diagonal = %g is too large",
                      (double)majorant * diag);
        if (p.tracking == RTC8X_TRACKING_LOCAL_MAJORANT)
            RTC8X_REQUIRE(g.region_majorants && g.region_dims[0] > 0 && g.region_dims[1] > 0 && g.region_dims[2] > 0 &&
                              g.region_cells >= 1 && (g.region_cells & (g.region_cells - 1)) == 0 &&
                              std::isfinite(g.region_unit),
                          "rtc8x_trace: local-majorant tracking needs region majorants");
    }
    if (p.tracking == RTC8X_TRACKING_LOCAL_MAJORANT) {
        // the single-image kernel stages the whole region table in LDS
        const int64_t nr = (int64_t)std::max(g.region_dims[0], 0) * std::max(g.region_dims[1], 0) *
                           std::max(g.region_dims[2], 0);
        RTC8X_REQUIRE(nr <= RTC8X_MAX_REGIONS && (nr == 0 || g.region_majorants),
                      "rtc8x_trace: %lld regions (at most %d, with their table)", (long long)nr, RTC8X_MAX_REGIONS);
    }
    if (p.env.kind == RTC8X_ENV_LATLONG) {
        RTC8X_REQUIRE(p.env.texels && p.env.width > 0 && p.env.height > 0, "rtc8x_trace: env texture missing");
        if (p.env.sampler == RTC8X_ENV_SAMPLER_ALIAS)
            RTC8X_REQUIRE(p.env.alias_rows && p.env.alias_cols && p.env.alias_weights && p.env.alias_integral > 0.0f,
                          "rtc8x_trace: env alias tables missing");
        else
            RTC8X_REQUIRE(p.env.mips && p.env.num_mip_levels >= 1 && p.env.num_mip_levels <= 31 &&
                              p.env.map_integral > 0.0f,
                          "rtc8x_trace: env importance map missing");
    }
    if (p.sun.enabled) {
        RTC8X_REQUIRE(finite3(p.sun.direction) && finite3(p.sun.irradiance), "rtc8x_trace: sun not finite");
        if (p.env.kind != RTC8X_ENV_NONE)
            RTC8X_REQUIRE(p.sun.select_prob > 0.0f && p.sun.select_prob < 1.0f,
                          "rtc8x_trace: sun.select_prob must be in (0,1) when an environment is present");
    }
    RTC8X_REQUIRE(entry != RTC8X_ENTRY_GENERATE_TRAINING_DATA || (!p.work_order && !p.pixel_cost),
                  "rtc8x_trace: work_order / pixel_cost are for the full-image entries (the training size varies)");
    return RTC8X_OK;
}

}  // namespace

extern "C" rtc8x_status rtc8x_trace(void* stream, const rtc8x_frame_params* params, int32_t entry, uint32_t num_samples)
try {
    RTC8X_REQUIRE(params != nullptr, "rtc8x_trace: params is NULL");
    const rtc8x_frame_params& p = *params;
    {
        const rtc8x_status st = check_launch(p, entry, num_samples);
        if (st != RTC8X_OK) return st;
    }
    hipStream_t s = (hipStream_t)stream;
    if (entry == RTC8X_ENTRY_PATH_TRACE) {
        RTC8X_HIP((dev::launch_tracking<RTC8X_ENTRY_PATH_TRACE>(p, p.image_width, p.image_height, num_samples,
                                                                        nullptr, nullptr, s)));
        return RTC8X_OK;
    }
    const rtc8x_nrc_buffers& nb = p.nrc;
    RTC8X_REQUIRE(nb.inference_queries && nb.max_num_training_suffixes >= p.image_width * p.image_height,
                  "rtc8x_trace: NRC buffers missing");
    if (entry == RTC8X_ENTRY_PATH_TRACE_WITH_NRC) {
        RTC8X_REQUIRE(nb.inference_terminals && nb.per_frame_contribution, "rtc8x_trace: NRC buffers missing");
        RTC8X_HIP((dev::launch_tracking<RTC8X_ENTRY_PATH_TRACE_WITH_NRC>(p, p.image_width, p.image_height, num_samples,
                                                                                 nullptr, nullptr, s)));
        return RTC8X_OK;
    }
    // GENERATE_TRAINING_DATA: the launch size lives on the device (prepareNRC writes it); the kernels read
    // it there, so the host never waits for it
    RTC8X_REQUIRE(nb.train_image_size && nb.num_training_data && nb.train_queries[0] && nb.train_targets[0] &&
                      nb.train_vertex_info && nb.train_suffix_terminals && nb.scratch,
                  "rtc8x_trace: NRC training buffers missing");
    RTC8X_REQUIRE(nb.nrc_buffer_index < 2, "rtc8x_trace: nrc_buffer_index must be 0 or 1");
    const NrcScratchLayout lay = nrc_scratch_layout(p.image_width, p.image_height);
    char* const sb = reinterpret_cast<char*>(nb.scratch);
    uint32_t* counts = reinterpret_cast<uint32_t*>(sb + lay.counts);
    uint32_t* offsets = reinterpret_cast<uint32_t*>(sb + lay.offsets);
    uint32_t* total = nb.num_training_data + nb.nrc_buffer_index;
    // One pass: records are staged as the paths reach them, then moved to their pixel-order slots, paths
    // whose records run past the buffer are cut there, and the pixels with a record that found no
    // staging slot are traced again.  Record indices, and with them every result, are those of
    // pixel-order assignment with paths ending at the buffer's capacity.
    uint64_t* rng_save = reinterpret_cast<uint64_t*>(sb + lay.rng_save);
    constexpr uint32_t kCompactBlocks = 512;
    hipLaunchKernelGGL(dev::save_rng_kernel, dim3(kCompactBlocks), dim3(256), 0, s, p, rng_save);
    RTC8X_HIP(hipGetLastError());
    dev::TrainStage stage{};
    stage.counter = reinterpret_cast<uint32_t*>(sb + lay.counter);
    stage.queries = reinterpret_cast<rtc8x_radiance_query*>(sb + lay.queries);
    stage.targets = reinterpret_cast<float*>(sb + lay.targets);
    stage.info = reinterpret_cast<rtc8x_training_vertex_info*>(sb + lay.info);
    stage.tag = reinterpret_cast<uint32_t*>(sb + lay.tag);
    stage.snap = reinterpret_cast<StageSnapshot*>(sb + lay.snap);
    stage.path_length = reinterpret_cast<uint32_t*>(sb + lay.path_length);
    stage.lookups = reinterpret_cast<uint32_t*>(sb + lay.lookups);
    stage.scatters = reinterpret_cast<uint32_t*>(sb + lay.scatters);
    stage.lost = reinterpret_cast<uint32_t*>(sb + lay.lost);
    hipLaunchKernelGGL(dev::reset_value_kernel, dim3(1), dim3(1), 0, s, stage.counter);
    RTC8X_HIP(hipGetLastError());
    RTC8X_HIP((dev::launch_tracking<RTC8X_ENTRY_GENERATE_TRAINING_DATA>(p, 0, 0, 1, nullptr, counts, s, stage)));
    {
        // the training image is at most the frame's image: tiles over that bound, the extra ones exit
        const uint32_t tiles = (uint32_t)(((size_t)p.image_width * p.image_height + kPrefixTile - 1) / kPrefixTile);
        uint64_t* tile_sums = reinterpret_cast<uint64_t*>(sb + lay.tiles);
        hipLaunchKernelGGL(dev::record_tile_sum_kernel, dim3(tiles), dim3(256), 0, s, p, counts, tile_sums);
        RTC8X_HIP(hipGetLastError());
        hipLaunchKernelGGL(dev::record_tile_scan_kernel, dim3(tiles), dim3(256), 0, s, p, counts,
                           (const uint64_t*)tile_sums, offsets, total);
        RTC8X_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(dev::stage_copy_kernel, dim3(kCompactBlocks), dim3(256), 0, s, p, total, stage, offsets,
                       counts);
    RTC8X_HIP(hipGetLastError());
    dev::TrainStage redo{};
    redo.redo_lost = stage.lost;
    redo.rng_save = rng_save;
    redo.staged = stage.counter;
    RTC8X_HIP((dev::launch_tracking<RTC8X_ENTRY_GENERATE_TRAINING_DATA>(p, 0, 0, 1, offsets, nullptr, s, redo)));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_trace")

namespace {

// Checks a batch and turns its shares into the launch's weights.
rtc8x_status check_batch(const char* who, const rtc8x_frame_params* params, const uint32_t* num_samples,
                         const float* shares, uint32_t count, int32_t entry, double* w)
{
    RTC8X_REQUIRE(params && num_samples, "%s: NULL argument", who);
    RTC8X_REQUIRE(count >= 1 && count <= dev::kMaxBatch, "%s: count %u must be in [1, %u]", who, count, dev::kMaxBatch);
    RTC8X_REQUIRE(entry == RTC8X_ENTRY_PATH_TRACE || entry == RTC8X_ENTRY_PATH_TRACE_WITH_NRC,
                  "%s: entry %d: only the full-image entries run in a batch", who, entry);
    for (uint32_t i = 0; i < count; ++i) {
        const rtc8x_status st = check_launch(params[i], entry, num_samples[i]);
        if (st != RTC8X_OK) {
            const std::string why = rtc8x_last_error();
            return fail(st, "%s: instance %u: %s", who, i, why.c_str());
        }
        RTC8X_REQUIRE(params[i].tracking == params[0].tracking, "%s: instances must share the tracking mode", who);
        if (shares) {
            RTC8X_REQUIRE(std::isfinite(shares[i]) && shares[i] > 0.0f, "%s: shares must be positive", who);
            w[i] = shares[i];
        } else {
            w[i] = (double)params[i].image_width * params[i].image_height * num_samples[i];
        }
    }
    return RTC8X_OK;
}

// The batched launch over a device copy `d` of the instances' frame parameters.
hipError_t launch_batch_any(const rtc8x_frame_params* d, int32_t tracking, int32_t entry, const uint32_t* num_samples,
                            const double* w, uint32_t count, hipStream_t s)
{
    if (tracking == RTC8X_TRACKING_LOCAL_MAJORANT)
        return entry == RTC8X_ENTRY_PATH_TRACE
                   ? dev::launch_batch<RTC8X_ENTRY_PATH_TRACE, RTC8X_TRACKING_LOCAL_MAJORANT>(d, num_samples, w, count, s)
                   : dev::launch_batch<RTC8X_ENTRY_PATH_TRACE_WITH_NRC, RTC8X_TRACKING_LOCAL_MAJORANT>(d, num_samples, w,
                                                                                                      count, s);
    return entry == RTC8X_ENTRY_PATH_TRACE
               ? dev::launch_batch<RTC8X_ENTRY_PATH_TRACE, RTC8X_TRACKING_GLOBAL_MAJORANT>(d, num_samples, w, count, s)
               : dev::launch_batch<RTC8X_ENTRY_PATH_TRACE_WITH_NRC, RTC8X_TRACKING_GLOBAL_MAJORANT>(d, num_samples, w,
                                                                                                   count, s);
}

}  // namespace

namespace rtc8x {

rtc8x_status trace_batch_into(const char* who, void* stream, const rtc8x_frame_params* params,
                              const uint32_t* num_samples, const float* shares, uint32_t count, int32_t entry,
                              rtc8x_frame_params* table)
{
    double w[dev::kMaxBatch];
    {
        const rtc8x_status st = check_batch(who, params, num_samples, shares, count, entry, w);
        if (st != RTC8X_OK) return st;
    }
    RTC8X_REQUIRE(table != nullptr, "%s: no device table for the instances", who);
    hipStream_t s = (hipStream_t)stream;
    RTC8X_HIP(hipMemcpyAsync(table, params, sizeof(rtc8x_frame_params) * count, hipMemcpyHostToDevice, s));
    RTC8X_HIP(launch_batch_any(table, params[0].tracking, entry, num_samples, w, count, s));
    return RTC8X_OK;
}

}  // namespace rtc8x

extern "C" rtc8x_status rtc8x_trace_batch(void* stream, const rtc8x_frame_params* params, const uint32_t* num_samples,
                                          const float* shares, uint32_t count, int32_t entry)
try {
    RTC8X_REQUIRE(count >= 1 && count <= dev::kMaxBatch, "rtc8x_trace_batch: count %u must be in [1, %u]", count,
                  dev::kMaxBatch);
    hipStream_t s = (hipStream_t)stream;
    rtc8x_frame_params* d = nullptr;
    RTC8X_HIP(hipMallocAsync((void**)&d, sizeof(rtc8x_frame_params) * count, s));
    // the instance table is freed on every path, also when the checks, the copy or the launch fail
    const rtc8x_status st = trace_batch_into("rtc8x_trace_batch", stream, params, num_samples, shares, count, entry, d);
    const hipError_t freed = hipFreeAsync(d, s);
    if (st != RTC8X_OK) return st;
    RTC8X_HIP(freed);
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_trace_batch")

// rtc8x_common.hpp -- host-side error plumbing shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "rtc8x.h"

namespace rtc8x {

void set_error(const std::string& msg);
rtc8x_status fail(rtc8x_status st, const char* fmt, ...);

// Host copies of grid tables that the library owns until rtc8x_grid_free_host.
struct OwnedHostGrid {
    std::vector<rtc8x_root_tile> root;
    std::vector<uint32_t> upper, lower;
    std::vector<float> values, tiles, cells;
    std::vector<uint8_t> regions;
    std::vector<int32_t> origins;
    std::vector<uint64_t> leaf_masks, lower_masks, upper_masks;
};
// Takes ownership of `t`, points the table pointers of `g` at its storage and stores the result in *out.
rtc8x_status adopt_host_grid(OwnedHostGrid&& t, const rtc8x_grid& g, rtc8x_grid* out);
// rtc8x_grid_upload; with keep_max the host's max_value is kept instead of the brute-force maximum.
rtc8x_status upload_host_grid(const rtc8x_grid* h, rtc8x_grid* out, bool keep_max);

// rtc8x_trace_batch with the instances' frame parameters copied into `table`, a device array of at least
// `count` entries that the caller provides (`who` names the entry point in error messages).
rtc8x_status trace_batch_into(const char* who, void* stream, const rtc8x_frame_params* params,
                              const uint32_t* num_samples, const float* shares, uint32_t count, int32_t entry,
                              rtc8x_frame_params* table);

// Records the one-pass training launch can stage: beyond them, records find no slot and their pixels
// are traced again.
constexpr size_t kStageCapacity = 4 * (size_t)RTC8X_TRAIN_BUFFER_SIZE;

// A staged record's path state: what the path's end would be if it were cut at this record.
struct StageSnapshot {
    uint64_t rng;          // PCG state after the record's vertex
    uint32_t path_length;  // vertices so far
    uint32_t lookups;      // density lookups so far
    uint32_t scatters;     // real collisions so far
    uint32_t pad;
};

// Byte offsets of the parts of rtc8x_nrc_buffers.scratch for a width x height image: record counts and
// offsets of the training launch (one u32 per pixel each), the shuffle's per-block target statistics,
// then the one-pass training launch's staging: saved PCG states, per-pixel path lengths, lookups, real
// collisions and lost-record flags, the staged records (query, target, vertex info, tag, snapshot,
// final slot) and the staging counter.
struct NrcScratchLayout {
    size_t counts, offsets, shuffle, rng_save, path_length, lookups, scatters, lost;
    size_t queries, targets, info, tag, snap, counter, tiles, total;
};
constexpr uint32_t kPrefixTile = 4096;  // record counts per workgroup of the prefix sum
inline NrcScratchLayout nrc_scratch_layout(uint32_t width, uint32_t height)
{
    constexpr size_t kShuffleFloats = 9 * (RTC8X_NUM_TRAINING_DATA_PER_FRAME / 256);
    const size_t px = (size_t)width * height, cap = kStageCapacity;
    const auto up16 = [](size_t v) { return (v + 15) & ~(size_t)15; };
    NrcScratchLayout l;
    l.counts = 0;
    l.offsets = 4 * px;
    l.shuffle = 8 * px;
    l.rng_save = up16(l.shuffle + 4 * kShuffleFloats);
    l.path_length = l.rng_save + 8 * px;
    l.lookups = l.path_length + 4 * px;
    l.scatters = l.lookups + 4 * px;
    l.lost = l.scatters + 4 * px;
    l.queries = up16(l.lost + 4 * px);
    l.targets = l.queries + sizeof(rtc8x_radiance_query) * cap;
    l.info = up16(l.targets + 12 * cap);
    l.tag = l.info + sizeof(rtc8x_training_vertex_info) * cap;
    l.snap = l.tag + 8 * cap;
    l.counter = up16(l.snap + sizeof(StageSnapshot) * cap);
    l.tiles = l.counter + 16;
    l.total = l.tiles + 8 * ((px + kPrefixTile - 1) / kPrefixTile);
    return l;
}

}  // namespace rtc8x

// Evaluates a HIP call; on failure records the message and returns RTC8X_ERR_HIP from the caller.
#define RTC8X_HIP(call)                                                                          \
    do {                                                                                         \
        hipError_t err_ = (call);                                                                \
        if (err_ != hipSuccess) {                                                                \
            return rtc8x::fail(err_ == hipErrorOutOfMemory ? RTC8X_ERR_OUT_OF_MEMORY : RTC8X_ERR_HIP, \
                               "%s failed: %s (%s:%d)", #call, hipGetErrorString(err_), __FILE__,     \
                               __LINE__);                                                        \
        }                                                                                        \
    } while (0)

#define RTC8X_REQUIRE(cond, ...)                                              \
    do {                                                                      \
        if (!(cond)) return rtc8x::fail(RTC8X_ERR_INVALID_ARGUMENT, __VA_ARGS__); \
    } while (0)

// The handler of an entry point's function-try-block: no C++ exception crosses the C ABI. A failed host
// allocation (a std::vector sized from a file header, say) becomes RTC8X_ERR_OUT_OF_MEMORY, anything else
// RTC8X_ERR_INVALID_ARGUMENT, with the entry point's name in rtc8x_last_error().
#define RTC8X_CATCH_ALL(name)                                                                     \
    catch (const std::bad_alloc&)                                                                 \
    {                                                                                             \
        return rtc8x::fail(RTC8X_ERR_OUT_OF_MEMORY, "%s: out of host memory", name);               \
    }                                                                                             \
    catch (const std::exception& e_)                                                              \
    {                                                                                             \
        return rtc8x::fail(RTC8X_ERR_INVALID_ARGUMENT, "%s: %s", name, e_.what());                 \
    }                                                                                             \
    catch (...)                                                                                   \
    {                                                                                             \
        return rtc8x::fail(RTC8X_ERR_INVALID_ARGUMENT, "%s: unknown C++ exception", name);         \
    }

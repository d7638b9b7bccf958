// rtc8x_noalloc.hip -- the batched trace for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc): rtc8x_trace_batch with its
// instance table in a fixed device array instead of a per-call allocation.
#include <hip/hip_runtime.h>

#include <mutex>

#include "rtc8x_common.hpp"

namespace rtc8x {
namespace {

constexpr uint32_t kTableInstances = 16;  // the batched launch's maximum instance count
constexpr int kMaxDevices = 64;

// The instance table of rtc8x_trace_batch_static: 16 frame-parameter blocks (16 x sizeof(rtc8x_frame_params)
// bytes), one copy per GPU for the life of the process.  In host code this name is not a device address;
// the address of the current device's copy comes from hipGetSymbolAddress (table_ptr below).
static __device__ __align__(256) rtc8x_frame_params scratch[kTableInstances];

std::mutex g_table_mutex;
rtc8x_frame_params* g_table[kMaxDevices];

// The current device's copy of `scratch`, looked up once per device with that device current.
rtc8x_status table_ptr(rtc8x_frame_params** out)
{
    int dev = 0;
    RTC8X_HIP(hipGetDevice(&dev));
    RTC8X_REQUIRE(dev >= 0 && dev < kMaxDevices, "rtc8x_trace_batch_static: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_table_mutex);
    if (!g_table[dev]) {
        void* ptr = nullptr;
        RTC8X_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        RTC8X_REQUIRE(ptr != nullptr, "rtc8x_trace_batch_static: no device address for the instance table");
        g_table[dev] = static_cast<rtc8x_frame_params*>(ptr);
    }
    *out = g_table[dev];
    return RTC8X_OK;
}

}  // namespace
}  // namespace rtc8x

using namespace rtc8x;

extern "C" rtc8x_status rtc8x_trace_batch_static(void* stream, const rtc8x_frame_params* params,
                                                 const uint32_t* num_samples, const float* shares, uint32_t count,
                                                 int32_t entry)
try {
    RTC8X_REQUIRE(count >= 1 && count <= kTableInstances,
                  "rtc8x_trace_batch_static: count %u must be in [1, %u] (the fixed instance table holds %u)", count,
                  kTableInstances, kTableInstances);
    rtc8x_frame_params* table = nullptr;
    {
        const rtc8x_status st = table_ptr(&table);
        if (st != RTC8X_OK) return st;
    }
    return trace_batch_into("rtc8x_trace_batch_static", stream, params, num_samples, shares, count, entry, table);
}
RTC8X_CATCH_ALL("rtc8x_trace_batch_static")

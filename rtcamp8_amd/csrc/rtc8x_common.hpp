// rtc8x_common.hpp -- host-side error plumbing shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "rtc8x.h"

namespace rtc8x {

void set_error(const std::string& msg);
rtc8x_status fail(rtc8x_status st, const char* fmt, ...);

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

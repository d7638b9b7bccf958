// rtc8x_api.hip -- library plumbing of the C ABI: errors, device memory, streams, events, the camera
// helper and the per-pixel RNG seeding.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "rtc8x_common.hpp"

namespace rtc8x {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

rtc8x_status fail(rtc8x_status st, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return st;
}

}  // namespace rtc8x

using namespace rtc8x;

extern "C" {

int32_t rtc8x_abi_version(void) { return RTC8X_ABI_VERSION; }

const char* rtc8x_last_error(void) { return g_last_error.c_str(); }

rtc8x_status rtc8x_device_count(int32_t* count)
try {
    RTC8X_REQUIRE(count != nullptr, "rtc8x_device_count: count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_device_count")

rtc8x_status rtc8x_set_device(int32_t device)
try {
    RTC8X_HIP(hipSetDevice(device));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_set_device")

rtc8x_status rtc8x_device_name(char* buf, size_t buf_len)
try {
    RTC8X_REQUIRE(buf != nullptr && buf_len > 0, "rtc8x_device_name: bad buffer");
    int dev = 0;
    RTC8X_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    RTC8X_HIP(hipGetDeviceProperties(&prop, dev));
    snprintf(buf, buf_len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_device_name")

rtc8x_status rtc8x_malloc(void** ptr, size_t bytes)
try {
    RTC8X_REQUIRE(ptr != nullptr, "rtc8x_malloc: ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) return RTC8X_OK;
    RTC8X_HIP(hipMalloc(ptr, bytes));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_malloc")

rtc8x_status rtc8x_free(void* ptr)
try {
    if (ptr) RTC8X_HIP(hipFree(ptr));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_free")

rtc8x_status rtc8x_memcpy_htod(void* dst, const void* src, size_t bytes)
try {
    if (bytes == 0) return RTC8X_OK;
    RTC8X_REQUIRE(dst && src, "rtc8x_memcpy_htod: NULL pointer");
    RTC8X_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_memcpy_htod")

rtc8x_status rtc8x_memcpy_dtoh(void* dst, const void* src, size_t bytes)
try {
    if (bytes == 0) return RTC8X_OK;
    RTC8X_REQUIRE(dst && src, "rtc8x_memcpy_dtoh: NULL pointer");
    RTC8X_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_memcpy_dtoh")

rtc8x_status rtc8x_memcpy_dtod(void* dst, const void* src, size_t bytes)
try {
    if (bytes == 0) return RTC8X_OK;
    RTC8X_REQUIRE(dst && src, "rtc8x_memcpy_dtod: NULL pointer");
    RTC8X_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_memcpy_dtod")

rtc8x_status rtc8x_memset(void* dst, int32_t value, size_t bytes)
try {
    if (bytes == 0) return RTC8X_OK;
    RTC8X_REQUIRE(dst != nullptr, "rtc8x_memset: NULL pointer");
    RTC8X_HIP(hipMemset(dst, value, bytes));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_memset")

rtc8x_status rtc8x_device_synchronize(void)
try {
    RTC8X_HIP(hipDeviceSynchronize());
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_device_synchronize")

rtc8x_status rtc8x_stream_create(void** stream)
try {
    RTC8X_REQUIRE(stream != nullptr, "rtc8x_stream_create: stream is NULL");
    hipStream_t s;
    RTC8X_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_stream_create")

rtc8x_status rtc8x_stream_destroy(void* stream)
try {
    if (stream) RTC8X_HIP(hipStreamDestroy((hipStream_t)stream));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_stream_destroy")

rtc8x_status rtc8x_stream_synchronize(void* stream)
try {
    RTC8X_HIP(hipStreamSynchronize((hipStream_t)stream));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_stream_synchronize")

rtc8x_status rtc8x_event_create(void** event)
try {
    RTC8X_REQUIRE(event != nullptr, "rtc8x_event_create: event is NULL");
    hipEvent_t e;
    RTC8X_HIP(hipEventCreate(&e));
    *event = (void*)e;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_event_create")

rtc8x_status rtc8x_event_destroy(void* event)
try {
    if (event) RTC8X_HIP(hipEventDestroy((hipEvent_t)event));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_event_destroy")

rtc8x_status rtc8x_event_record(void* event, void* stream)
try {
    RTC8X_REQUIRE(event != nullptr, "rtc8x_event_record: event is NULL");
    RTC8X_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_event_record")

rtc8x_status rtc8x_stream_wait_event(void* stream, void* event)
try {
    RTC8X_REQUIRE(event != nullptr, "rtc8x_stream_wait_event: event is NULL");
    RTC8X_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_stream_wait_event")

rtc8x_status rtc8x_event_elapsed_ms(float* ms, void* start, void* end)
try {
    RTC8X_REQUIRE(ms && start && end, "rtc8x_event_elapsed_ms: NULL argument");
    RTC8X_HIP(hipEventSynchronize((hipEvent_t)end));
    RTC8X_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_event_elapsed_ms")

// PerspectiveCamera set-up: orientation = Quaternion::toMatrix3x3 (common/basic_types.h:1620-1628)
// stored row-major; vh = 2 tan(fovY/2), vw = aspect * vh (optix_pathtracing_kernels.cu:47-48).
rtc8x_status rtc8x_camera_setup(rtc8x_camera* cam, const float position[3], const float q[4], float fov_y,
                                float aspect)
try {
    RTC8X_REQUIRE(cam && position && q, "rtc8x_camera_setup: NULL argument");
    RTC8X_REQUIRE(fov_y > 0.0f && fov_y < 3.14159f && aspect > 0.0f, "rtc8x_camera_setup: bad fov/aspect");
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, yz = y * z, zx = z * x;
    const float xw = x * w, yw = y * w, zw = z * w;
    // columns c0, c1, c2 of the reference; row-major storage M[r][c] = c_c[r]
    const float c0[3] = {1 - 2 * (yy + zz), 2 * (xy + zw), 2 * (zx - yw)};
    const float c1[3] = {2 * (xy - zw), 1 - 2 * (xx + zz), 2 * (yz + xw)};
    const float c2[3] = {2 * (zx + yw), 2 * (yz - xw), 1 - 2 * (xx + yy)};
    for (int r = 0; r < 3; ++r) {
        cam->orientation[3 * r + 0] = c0[r];
        cam->orientation[3 * r + 1] = c1[r];
        cam->orientation[3 * r + 2] = c2[r];
        cam->position[r] = position[r];
    }
    cam->vh = 2.0f * std::tan(fov_y * 0.5f);
    cam->vw = aspect * cam->vh;
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_camera_setup")

// Per-pixel PCG states from std::mt19937_64 in row-major order (renderer/renderer_main.cpp:86-93).
rtc8x_status rtc8x_seed_rng(uint64_t* states, uint32_t width, uint32_t height, uint64_t seed)
try {
    RTC8X_REQUIRE(states != nullptr && width > 0 && height > 0, "rtc8x_seed_rng: bad arguments");
    std::vector<uint64_t> host((size_t)width * height);
    std::mt19937_64 gen(seed);
    for (auto& s : host) s = gen();
    RTC8X_HIP(hipMemcpy(states, host.data(), host.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    return RTC8X_OK;
}
RTC8X_CATCH_ALL("rtc8x_seed_rng")

}  // extern "C"

// Layout check for bindings: sizeof of the ABI structs by name (0 for an unknown name).
extern "C" size_t rtc8x_sizeof(const char* type_name)
{
    if (!type_name) return 0;
    struct Entry {
        const char* name;
        size_t size;
    };
    static const Entry table[] = {
        {"rtc8x_root_tile", sizeof(rtc8x_root_tile)},
        {"rtc8x_grid", sizeof(rtc8x_grid)},
        {"rtc8x_fbm_desc", sizeof(rtc8x_fbm_desc)},
        {"rtc8x_alias_entry", sizeof(rtc8x_alias_entry)},
        {"rtc8x_env_light", sizeof(rtc8x_env_light)},
        {"rtc8x_sun_light", sizeof(rtc8x_sun_light)},
        {"rtc8x_camera", sizeof(rtc8x_camera)},
        {"rtc8x_radiance_query", sizeof(rtc8x_radiance_query)},
        {"rtc8x_terminal_info", sizeof(rtc8x_terminal_info)},
        {"rtc8x_training_vertex_info", sizeof(rtc8x_training_vertex_info)},
        {"rtc8x_training_suffix_terminal", sizeof(rtc8x_training_suffix_terminal)},
        {"rtc8x_nrc_buffers", sizeof(rtc8x_nrc_buffers)},
        {"rtc8x_frame_params", sizeof(rtc8x_frame_params)},
        {"rtc8x_nrc_desc", sizeof(rtc8x_nrc_desc)},
        {"rtc8x_hosek_band", sizeof(rtc8x_hosek_band)},
        {"rtc8x_hosek_sky", sizeof(rtc8x_hosek_sky)},
        {"rtc8x_nvdb_grid_info", sizeof(rtc8x_nvdb_grid_info)},
        {"rtc8x_sdr_config", sizeof(rtc8x_sdr_config)},
        {"rtc8x_tone_map_params", sizeof(rtc8x_tone_map_params)},
    };
    for (const Entry& e : table)
        if (std::strcmp(e.name, type_name) == 0) return e.size;
    return 0;
}

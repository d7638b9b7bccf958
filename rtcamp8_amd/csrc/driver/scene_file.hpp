// scene_file.hpp -- the reference's scene-description language and its keyframed cameras and
// environments, for the batch renderer (renderer/renderer_scene.cpp:1610-2420 and
// renderer/renderer_scene.h:913-1330).  Host code, no GPU.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace rtc8x_driver {

struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
struct Quat {  // x, y, z imaginary; w real
    float x = 0.0f, y = 0.0f, z = 0.0f, w = 1.0f;
};

// KeyCameraState (renderer_scene.h:913-919): fovY in degrees
struct CameraKey {
    float time = 0.0f;
    Vec3 position, look_at, up;
    float fov_y = 0.0f;
};

struct CameraDesc {
    Vec3 next_position{0.0f, 1.0f, 5.0f};
    Vec3 next_look_at{0.0f, 0.0f, 0.0f};
    Vec3 next_up{0.0f, 1.0f, 0.0f};
    float next_fov_y = 60.0f * 3.14159265f / 180.0f;  // the reference's default, used as degrees later
    std::vector<CameraKey> keys;
};

// KeyEnvironmentState / KeyAnalyticSkyEnvironmentState (renderer_scene.h:1125-1244)
struct EnvironmentKey {
    float time = 0.0f;
    float coeff = 1.0f;
    float rotation = 0.0f;  // degrees
    Vec3 solar_direction;   // analytic sky only
    float turbidity = 0.0f;
};

enum class EnvKind { None, File, AnalyticSky };

struct EnvironmentDesc {
    EnvKind kind = EnvKind::None;
    std::string name;
    std::string file;  // EnvKind::File
    float next_coeff = 1.0f, next_rotation = 0.0f;
    Vec3 next_solar_direction;
    float next_turbidity = 4.0f;
    std::vector<EnvironmentKey> keys;
};

struct MeshDesc {
    bool rectangle = false;
    std::string file;
    float scale = 1.0f;
    float dim_x = 0.0f, dim_z = 0.0f;
    Vec3 emittance;
};

struct InstanceKey {
    float time = 0.0f;
    float scale = 1.0f;
    Quat orientation;
    Vec3 position;
};

struct InstanceDesc {
    std::string mesh;
    float next_scale = 1.0f;
    Quat next_orientation;
    Vec3 next_position;
    std::vector<InstanceKey> keys;
    bool cyclic = false;
};

struct ActiveCamera {
    float time = 0.0f;
    std::string name;
};

// RenderConfigs (renderer_scene.h:1335-1350) plus the volume file
struct Scene {
    uint32_t width = 0, height = 0;
    float time_begin = 0.0f, time_end = 0.0f;
    uint32_t fps = 0;
    std::string volume_file;
    EnvironmentDesc environment;
    std::map<std::string, CameraDesc> cameras;
    std::vector<ActiveCamera> active_cameras;
    std::map<std::string, MeshDesc> meshes;
    std::map<std::string, InstanceDesc> instances;
};

// Parses a scene file (loadScene, renderer_scene.cpp:2307-2420): relative paths resolve against the
// file's directory, key states are sorted by time.  Throws std::runtime_error "l.N: ..." on errors.
Scene load_scene(const std::string& path);
Scene parse_scene(const std::string& text, const std::string& base_dir);

// The device camera at a time point (Camera::setUpDeviceData, renderer_scene.h:929-985): Catmull-Rom
// position, slerped look-at orientation, interpolated field of view (returned in radians).
struct CameraState {
    Vec3 position;
    Quat orientation;  // as the device camera holds it (the conjugate of the look-at rotation)
    float fov_y = 0.0f;
};
CameraState camera_at(const CameraDesc& cam, float time_point);

// The environment at a time point (Environment / AnalyticSkyEnvironment::setUpDeviceData and
// computeDistribution, renderer_scene.h:1140-1165, 1263-1310, renderer_scene.cpp:555-570).
struct EnvironmentState {
    float coeff = 1.0f;
    float rotation = 0.0f;  // radians
    Vec3 solar_direction;
    float turbidity = 0.0f;
    float solar_elevation = 0.0f;
};
EnvironmentState environment_at(const EnvironmentDesc& env, float time_point);

// Vector3D::fromPolarYUp (common/basic_types.h:900-907)
Vec3 from_polar_y_up(float phi, float theta);
// qLookAt (common/basic_types.h:2650-2694)
Quat look_at_rotation(Vec3 dir, Vec3 up);

}  // namespace rtc8x_driver

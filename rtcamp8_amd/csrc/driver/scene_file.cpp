// scene_file.cpp -- parser of the reference's scene files and the keyframe evaluation of cameras and
// environments.  The command set, the token forms and the error texts are the reference's
// (renderer/renderer_scene.cpp:1610-2420); the code is this renderer's own.
#include "scene_file.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <functional>
#include <regex>
#include <sstream>

namespace rtc8x_driver {
namespace {

constexpr float kPi = 3.14159265358979323846f;

// ----------------------------------------------------------------------------------- small vector maths

Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec3 operator*(float s, Vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
Vec3 operator-(Vec3 a) { return {-a.x, -a.y, -a.z}; }
float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
Vec3 normalize(Vec3 a)
{
    const float l = std::sqrt(dot(a, a));
    return {a.x / l, a.y / l, a.z / l};
}

float dot(Quat a, Quat b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
Quat scale(Quat q, float s) { return {q.x * s, q.y * s, q.z * s, q.w * s}; }
Quat add(Quat a, Quat b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
Quat normalize(Quat q) { return scale(q, 1.0f / std::sqrt(dot(q, q))); }
Quat conjugate(Quat q) { return {-q.x, -q.y, -q.z, q.w}; }

float lerp(float a, float b, float t) { return (1.0f - t) * a + t * b; }
float safe_divide(float a, float b) { return b != 0.0f ? a / b : 0.0f; }
float clampf(float v, float lo, float hi) { return std::min(std::max(v, lo), hi); }

// slerp with the shorter-path option (common/basic_types.h:2697-2716)
Quat slerp(Quat a, Quat b, float t)
{
    float c = dot(a, b);
    if (c > 0.9995f) return normalize(add(scale(a, 1.0f - t), scale(b, t)));
    if (c < 0.0f) {
        b = scale(b, -1.0f);
        c = -c;
    }
    const float theta = std::acos(clampf(c, -1.0f, 1.0f));
    const Quat perp = normalize(add(b, scale(a, -c)));
    return add(scale(a, std::cos(theta * t)), scale(perp, std::sin(theta * t)));
}

// Rodrigues rotation of v by `angle` about `axis` (rotate3x3, common/basic_types.h:2196-2215)
Vec3 rotate(float angle, Vec3 axis, Vec3 v)
{
    const Vec3 n = normalize(axis);
    const float s = std::sin(angle), c = std::cos(angle), k = 1.0f - c;
    const float m00 = n.x * n.x * k + c, m01 = n.x * n.y * k - n.z * s, m02 = n.z * n.x * k + n.y * s;
    const float m10 = n.x * n.y * k + n.z * s, m11 = n.y * n.y * k + c, m12 = n.y * n.z * k - n.x * s;
    const float m20 = n.z * n.x * k - n.y * s, m21 = n.y * n.z * k + n.x * s, m22 = n.z * n.z * k + c;
    return {m00 * v.x + m01 * v.y + m02 * v.z, m10 * v.x + m11 * v.y + m12 * v.z, m20 * v.x + m21 * v.y + m22 * v.z};
}

// One vector of the orthonormal basis around a unit vector (Duff et al. 2017, as
// Vector3D::makeCoordinateSystem, common/basic_types.h:857-864)
Vec3 basis_x(Vec3 n)
{
    const float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + n.z);
    return {1.0f + sign * n.x * n.x * a, sign * (n.x * n.y * a), -sign * n.x};
}

// qRotate (common/basic_types.h:2622-2627)
Quat rotation_quat(float angle, Vec3 axis)
{
    const Vec3 n = normalize(axis);
    const float s = std::sin(angle / 2.0f), c = std::cos(angle / 2.0f);
    return {s * n.x, s * n.y, s * n.z, c};
}

Quat mul(Quat a, Quat b)
{
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}

// Largest index whose key time is at or before t (the power-of-two search of Camera::setUpDeviceData)
template <typename K>
int key_index(const std::vector<K>& keys, float t)
{
    int idx = 0;
    const int n = (int)keys.size();
    int d = 1;
    while (d < n) d <<= 1;
    for (d >>= 1; d >= 1; d >>= 1)
        if (idx + d < n && keys[idx + d].time <= t) idx += d;
    return idx;
}

// ----------------------------------------------------------------------------------------- the parser

const char* kInteger = R"(([+-]?\d+))";
const char* kReal = R"(([+-]?(?:\d+\.?\d*|\.\d+)))";
const char* kString = R"((\S+?))";
const char* kQuotedPath = R"*("(.+?)")*";

std::regex command_regex(std::initializer_list<const char*> tokens)
{
    std::string pattern;
    bool first = true;
    for (const char* t : tokens) {
        if (!first) pattern += R"([ \t]+?)";
        pattern += t;
        first = false;
    }
    return std::regex(pattern + "$");
}

[[noreturn]] void fail_at(uint32_t line, const char* fmt, ...)
{
    char msg[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    char full[1100];
    std::snprintf(full, sizeof(full), "l.%u: %s", line, msg);
    throw std::runtime_error(full);
}

struct Context {
    Scene scene;
    std::filesystem::path dir;
    uint32_t line_no = 0;
    std::string line;

    std::smatch match(const char* cmd, const std::regex& re)
    {
        std::smatch m;
        if (!std::regex_search(line, m, re)) fail_at(line_no, "failed to parse \"%s\" command: %s", cmd, line.c_str());
        return m;
    }
    void require(bool ok, const char* fmt, const std::string& arg = std::string())
    {
        if (!ok) fail_at(line_no, fmt, arg.c_str());
    }
    std::string resolve(const std::string& p) const
    {
        std::filesystem::path path(p);
        if (path.is_relative()) path = dir / path;
        return path.string();
    }
    bool exists(const std::string& p) const { return std::filesystem::exists(p); }
};

float f(const std::ssub_match& s) { return std::stof(s.str()); }

using Handler = std::function<void(Context&)>;

const std::map<std::string, Handler>& handlers()
{
    static const std::map<std::string, Handler> table = {
        {"image",
         [](Context& c) {
             static const std::regex re = command_regex({"image", kInteger, kInteger});
             auto m = c.match("image", re);
             const int w = std::stoi(m[1].str()), h = std::stoi(m[2].str());
             if (!(w > 0 && h > 0)) fail_at(c.line_no, "Invalid image size: %d x %d", w, h);
             c.scene.width = (uint32_t)w;
             c.scene.height = (uint32_t)h;
         }},
        {"time",
         [](Context& c) {
             static const std::regex re = command_regex({"time", kReal, kReal, kInteger});
             auto m = c.match("time", re);
             const float b = f(m[1]), e = f(m[2]);
             if (!(b >= 0.0f && b <= e)) fail_at(c.line_no, "Invalid timeBegin/End: %f - %f", b, e);
             c.scene.time_begin = b;
             c.scene.time_end = e;
             c.scene.fps = (uint32_t)std::stoi(m[3].str());
         }},
        {"envmap",
         [](Context& c) {
             static const std::regex re = command_regex({"envmap", kQuotedPath, kString});
             auto m = c.match("envmap", re);
             const std::string name = m[2].str();
             c.require(!c.scene.meshes.count(name), "Mesh %s has been already created.", name);
             const std::string path = c.resolve(m[1].str());
             c.require(c.exists(path), "Mesh %s does not exist.", path);
             EnvironmentDesc& e = c.scene.environment;
             e = EnvironmentDesc();
             e.kind = EnvKind::File;
             e.name = name;
             e.file = path;
         }},
        {"skyenv",
         [](Context& c) {
             static const std::regex re = command_regex({"skyenv", kString});
             auto m = c.match("skyenv", re);
             const std::string name = m[1].str();
             c.require(!c.scene.meshes.count(name), "Mesh %s has been already created.", name);
             EnvironmentDesc& e = c.scene.environment;
             e = EnvironmentDesc();
             e.kind = EnvKind::AnalyticSky;
             e.name = name;
             e.next_solar_direction = from_polar_y_up(0.5f * kPi, kPi / 3.0f);
             e.next_turbidity = 4.0f;
         }},
        {"env-coeff",
         [](Context& c) {
             static const std::regex re = command_regex({"env-coeff", kString, kReal});
             auto m = c.match("env-coeff", re);
             c.require(m[1].str() == c.scene.environment.name, "Environment %s does not exist.", m[1].str());
             const float v = f(m[2]);
             if (!(v >= 0.0f)) fail_at(c.line_no, "Invalid coeff: %f", v);
             c.scene.environment.next_coeff = v;
         }},
        {"skyenv-dir",
         [](Context& c) {
             static const std::regex re = command_regex({"skyenv-dir", kString, kReal, kReal});
             auto m = c.match("skyenv-dir", re);
             c.require(m[1].str() == c.scene.environment.name, "Environment %s does not exist.", m[1].str());
             c.require(c.scene.environment.kind == EnvKind::AnalyticSky,
                       "Elevation cannot be assigned to this environment.");
             c.scene.environment.next_solar_direction = from_polar_y_up(f(m[2]) * kPi / 180.0f, f(m[3]) * kPi / 180.0f);
         }},
        {"env-rotation",
         [](Context& c) {
             static const std::regex re = command_regex({"env-rotation", kString, kReal});
             auto m = c.match("env-rotation", re);
             c.require(m[1].str() == c.scene.environment.name, "Environment %s does not exist.", m[1].str());
             c.scene.environment.next_rotation = f(m[2]);
         }},
        {"env-addkey",
         [](Context& c) {
             static const std::regex re = command_regex({"env-addkey", kString, kReal});
             auto m = c.match("env-addkey", re);
             EnvironmentDesc& e = c.scene.environment;
             c.require(m[1].str() == e.name, "Environment %s does not exist.", m[1].str());
             const float t = f(m[2]);
             if (!(t >= 0.0f)) fail_at(c.line_no, "Invalid timePoint: %f", t);
             EnvironmentKey k;
             k.time = t;
             k.coeff = e.next_coeff;
             k.rotation = e.next_rotation;
             if (e.kind == EnvKind::AnalyticSky) {
                 k.solar_direction = e.next_solar_direction;
                 k.turbidity = e.next_turbidity;
             }
             e.keys.push_back(k);
         }},
        {"camera",
         [](Context& c) {
             static const std::regex re = command_regex({"camera", kString});
             auto m = c.match("camera", re);
             const std::string name = m[1].str();
             c.require(!c.scene.cameras.count(name), "Camera %s has been already created.", name);
             c.require(!c.scene.instances.count(name), "Instance with the same name %s exists.", name);
             c.scene.cameras[name] = CameraDesc();
         }},
        {"fovy",
         [](Context& c) {
             static const std::regex re = command_regex({"fovy", kString, kReal});
             auto m = c.match("fovy", re);
             c.require(c.scene.cameras.count(m[1].str()) > 0, "Camera %s does not exist.", m[1].str());
             const float v = f(m[2]);
             if (!(v > 0.0f)) fail_at(c.line_no, "Invalid fovY: %f", v);
             c.scene.cameras.at(m[1].str()).next_fov_y = v;
         }},
        {"lookat",
         [](Context& c) {
             static const std::regex re =
                 command_regex({"lookat", kString, kReal, kReal, kReal, kReal, kReal, kReal, kReal, kReal, kReal});
             auto m = c.match("lookat", re);
             c.require(c.scene.cameras.count(m[1].str()) > 0, "Camera %s does not exist.", m[1].str());
             CameraDesc& cam = c.scene.cameras.at(m[1].str());
             cam.next_position = {f(m[2]), f(m[3]), f(m[4])};
             cam.next_look_at = {f(m[5]), f(m[6]), f(m[7])};
             cam.next_up = {f(m[8]), f(m[9]), f(m[10])};
         }},
        {"cam-addkey",
         [](Context& c) {
             static const std::regex re = command_regex({"cam-addkey", kString, kReal, kString, kString});
             auto m = c.match("cam-addkey", re);
             c.require(c.scene.cameras.count(m[1].str()) > 0, "Camera %s does not exist.", m[1].str());
             const float t = f(m[2]);
             if (!std::isfinite(t)) fail_at(c.line_no, "Invalid timePoint: %f", t);
             CameraDesc& cam = c.scene.cameras.at(m[1].str());
             cam.keys.push_back(CameraKey{t, cam.next_position, cam.next_look_at, cam.next_up, cam.next_fov_y});
         }},
        {"active-cam",
         [](Context& c) {
             static const std::regex re = command_regex({"active-cam", kString, kReal});
             auto m = c.match("active-cam", re);
             c.require(c.scene.cameras.count(m[1].str()) > 0, "Camera %s does not exist.", m[1].str());
             const float t = f(m[2]);
             if (!std::isfinite(t)) fail_at(c.line_no, "Invalid timePoint: %f", t);
             c.scene.active_cameras.push_back(ActiveCamera{t, m[1].str()});
         }},
        {"mesh",
         [](Context& c) {
             static const std::regex re = command_regex({"mesh", kQuotedPath, kReal, kString});
             auto m = c.match("mesh", re);
             const std::string name = m[3].str();
             c.require(!c.scene.meshes.count(name), "Mesh %s has been already created.", name);
             const std::string path = c.resolve(m[1].str());
             c.require(c.exists(path), "Mesh %s does not exist.", path);
             const float s = f(m[2]);
             if (!(s > 0.0f)) fail_at(c.line_no, "Invalid mesh scale: %f", s);
             MeshDesc d;
             d.file = path;
             d.scale = s;
             c.scene.meshes[name] = d;
         }},
        {"rect",
         [](Context& c) {
             static const std::regex re = command_regex({"rect", kReal, kReal, kString});
             auto m = c.match("rect", re);
             const std::string name = m[3].str();
             c.require(!c.scene.meshes.count(name), "Mesh %s has been already created.", name);
             const float dx = f(m[1]), dz = f(m[2]);
             if (!(dx > 0.0f && dz > 0.0f)) fail_at(c.line_no, "Invalid rectangle size: %f x %f", dx, dz);
             MeshDesc d;
             d.rectangle = true;
             d.dim_x = dx;
             d.dim_z = dz;
             c.scene.meshes[name] = d;
         }},
        {"emittance",
         [](Context& c) {
             static const std::regex re = command_regex({"emittance", kString, kReal, kReal, kReal});
             auto m = c.match("emittance", re);
             c.require(c.scene.meshes.count(m[1].str()) > 0, "Mesh %s does not exist.", m[1].str());
             MeshDesc& d = c.scene.meshes.at(m[1].str());
             c.require(d.rectangle, "Emittance cannot be assigned to this mesh.");
             const Vec3 e{f(m[2]), f(m[3]), f(m[4])};
             if (!(e.x >= 0.0f && e.y >= 0.0f && e.z >= 0.0f && std::isfinite(e.x) && std::isfinite(e.y) &&
                   std::isfinite(e.z)))
                 fail_at(c.line_no, "Invalid emittance: (%f, %f, %f)", e.x, e.y, e.z);
             d.emittance = e;
         }},
        {"inst",
         [](Context& c) {
             static const std::regex re = command_regex({"inst", kString, kString});
             auto m = c.match("inst", re);
             c.require(c.scene.meshes.count(m[1].str()) > 0, "Mesh %s does not exist.", m[1].str());
             const std::string name = m[2].str();
             c.require(!c.scene.instances.count(name), "Instance %s has been already created.", name);
             c.require(!c.scene.cameras.count(name), "Camera with the same name %s exists.", name);
             InstanceDesc d;
             d.mesh = m[1].str();
             c.scene.instances[name] = d;
         }},
        {"volume",
         [](Context& c) {
             static const std::regex re = command_regex({"volume", kQuotedPath});
             auto m = c.match("volume", re);
             const std::string path = c.resolve(m[1].str());
             c.require(c.exists(path), "Volume %s does not exist.", path);
             c.scene.volume_file = path;
         }},
        {"scale",
         [](Context& c) {
             static const std::regex re = command_regex({"scale", kString, kReal});
             auto m = c.match("scale", re);
             c.require(c.scene.instances.count(m[1].str()) > 0, "Instance %s does not exist.", m[1].str());
             c.scene.instances.at(m[1].str()).next_scale = f(m[2]);
         }},
        {"rotate",
         [](Context& c) {
             static const std::regex re = command_regex({"rotate", kString, kReal, kReal, kReal, kReal});
             auto m = c.match("rotate", re);
             c.require(c.scene.instances.count(m[1].str()) > 0, "Instance %s does not exist.", m[1].str());
             c.scene.instances.at(m[1].str()).next_orientation =
                 rotation_quat(f(m[5]) * kPi / 180.0f, Vec3{f(m[2]), f(m[3]), f(m[4])});
         }},
        {"orient",
         [](Context& c) {
             static const std::regex re = command_regex({"orient", kString, kReal, kReal, kReal, kReal, kReal, kReal});
             auto m = c.match("orient", re);
             c.require(c.scene.instances.count(m[1].str()) > 0, "Instance %s does not exist.", m[1].str());
             const Vec3 dir = normalize(Vec3{f(m[2]), f(m[3]), f(m[4])});
             const Vec3 up = normalize(Vec3{f(m[5]), f(m[6]), f(m[7])});
             c.scene.instances.at(m[1].str()).next_orientation =
                 mul(conjugate(look_at_rotation(dir, up)), rotation_quat(kPi, Vec3{0.0f, 1.0f, 0.0f}));
         }},
        {"trans",
         [](Context& c) {
             static const std::regex re = command_regex({"trans", kString, kReal, kReal, kReal});
             auto m = c.match("trans", re);
             c.require(c.scene.instances.count(m[1].str()) > 0, "Instance%s does not exist.", m[1].str());
             c.scene.instances.at(m[1].str()).next_position = {f(m[2]), f(m[3]), f(m[4])};
         }},
        {"inst-addkey",
         [](Context& c) {
             static const std::regex re = command_regex({"inst-addkey", kString, kReal, kString, kString});
             auto m = c.match("inst-addkey", re);
             c.require(c.scene.instances.count(m[1].str()) > 0, "Instance %s does not exist.", m[1].str());
             const float t = f(m[2]);
             if (!std::isfinite(t)) fail_at(c.line_no, "Invalid timePoint: %f", t);
             InstanceDesc& d = c.scene.instances.at(m[1].str());
             d.keys.push_back(InstanceKey{t, d.next_scale, d.next_orientation, d.next_position});
         }},
        {"cyclic",
         [](Context& c) {
             static const std::regex re = command_regex({"cyclic", kString});
             auto m = c.match("cyclic", re);
             c.require(c.scene.instances.count(m[1].str()) > 0, "Instance %s does not exist.", m[1].str());
             c.scene.instances.at(m[1].str()).cyclic = true;
         }},
    };
    return table;
}

std::string trim(const std::string& s)
{
    const size_t b = s.find_first_not_of(" \t");
    if (b == std::string::npos) return std::string();
    const size_t e = s.find_last_not_of(" \t");
    return s.substr(b, e - b + 1);
}

}  // namespace

Vec3 from_polar_y_up(float phi, float theta)
{
    return {-std::sin(phi) * std::sin(theta), std::cos(theta), std::cos(phi) * std::sin(theta)};
}

Quat look_at_rotation(Vec3 dir, Vec3 up)
{
    const Vec3 d = normalize(-dir);
    const Vec3 s = normalize(cross(up, d));
    const Vec3 u = cross(d, s);
    const float trace = s.x + u.y + d.z;
    Quat q;
    // rotation matrix with rows s, u, d to a quaternion (Shoemake's method, largest-component branch)
    if (trace > 0.0f) {
        const float r = 0.5f / std::sqrt(trace + 1.0f);
        q = {(u.z - d.y) * r, (d.x - s.z) * r, (s.y - u.x) * r, 0.25f / r};
    } else if (s.x > u.y && s.x > d.z) {
        const float r = 2.0f * std::sqrt(1.0f + s.x - u.y - d.z);
        q = {0.25f * r, (u.x + s.y) / r, (d.x + s.z) / r, (u.z - d.y) / r};
    } else if (u.y > d.z) {
        const float r = 2.0f * std::sqrt(1.0f + u.y - s.x - d.z);
        q = {(u.x + s.y) / r, 0.25f * r, (d.y + u.z) / r, (d.x - s.z) / r};
    } else {
        const float r = 2.0f * std::sqrt(1.0f + d.z - s.x - u.y);
        q = {(d.x + s.z) / r, (d.y + u.z) / r, 0.25f * r, (s.y - u.x) / r};
    }
    return conjugate(q);  // the look-at matrix moves objects, not the camera
}

Scene parse_scene(const std::string& text, const std::string& base_dir)
{
    Context c;
    c.dir = base_dir;
    std::stringstream ss(text);
    std::string raw;
    while (std::getline(ss, raw, '\n')) {
        ++c.line_no;
        if (!raw.empty() && raw.back() == '\r') raw.pop_back();
        const std::string line = trim(raw);
        if (line.empty() || line.rfind("//", 0) == 0) continue;
        const size_t sp = line.find_first_of(" \t");
        const std::string token = sp == std::string::npos ? line : line.substr(0, sp);
        auto it = handlers().find(token);
        if (it == handlers().end()) fail_at(c.line_no, "Unknown token found: \"%s\"", token.c_str());
        c.line = line;
        it->second(c);
    }
    Scene& s = c.scene;
    if (s.active_cameras.empty() && !s.cameras.empty()) s.active_cameras.push_back(ActiveCamera{0.0f, s.cameras.begin()->first});
    auto by_time = [](const auto& a, const auto& b) { return a.time < b.time; };
    std::stable_sort(s.environment.keys.begin(), s.environment.keys.end(), by_time);
    for (auto& cam : s.cameras) std::stable_sort(cam.second.keys.begin(), cam.second.keys.end(), by_time);
    std::stable_sort(s.active_cameras.begin(), s.active_cameras.end(), by_time);
    return s;
}

Scene load_scene(const std::string& path)
{
    std::filesystem::path p(path);
    if (p.is_relative()) p = std::filesystem::current_path() / p;
    if (!std::filesystem::exists(p)) throw std::runtime_error("Scene file does not exist!");
    std::ifstream in(p, std::ios::binary);
    std::stringstream buf;
    buf << in.rdbuf();
    return parse_scene(buf.str(), p.parent_path().string());
}

CameraState camera_at(const CameraDesc& cam, float t)
{
    if (cam.keys.empty()) throw std::runtime_error("camera has no key state (cam-addkey)");
    const std::vector<CameraKey>& k = cam.keys;
    const int n = (int)k.size();
    t = clampf(t, k.front().time, k.back().time);
    const int i = key_index(k, t);
    const CameraKey& k1 = k[i];
    const CameraKey& k2 = k[std::min(i + 1, n - 1)];
    // neighbours for the Catmull-Rom tangents, mirrored at the ends
    const Vec3 p0 = i - 1 >= 0 ? k[i - 1].position : 2.0f * k1.position - k2.position;
    const Vec3 p3 = i + 2 < n ? k[i + 2].position : 2.0f * k2.position - k1.position;
    const Vec3 p1 = k1.position, p2 = k2.position;
    float u = safe_divide(t - k1.time, k2.time - k1.time);
    u = clampf(u, 0.0f, 1.0f);
    const Vec3 a4 = p1;
    const Vec3 a3 = 0.5f * (p2 - p0);
    const Vec3 a1 = 0.5f * (p3 - p1) - 2.0f * p2 + a3 + 2.0f * a4;
    const Vec3 a2 = 3.0f * p2 - 0.5f * (p3 - p1) - 2.0f * a3 - 3.0f * a4;
    CameraState out;
    out.position = (u * u * u) * a1 + (u * u) * a2 + u * a3 + a4;
    const Quat q0 = look_at_rotation(k1.look_at - k1.position, k1.up);
    const Quat q1 = look_at_rotation(k2.look_at - k2.position, k2.up);
    out.orientation = conjugate(slerp(q0, q1, u));
    out.fov_y = lerp(k1.fov_y, k2.fov_y, u) * kPi / 180.0f;
    return out;
}

EnvironmentState environment_at(const EnvironmentDesc& env, float t)
{
    if (env.keys.empty()) throw std::runtime_error("environment has no key state (env-addkey)");
    const std::vector<EnvironmentKey>& k = env.keys;
    const int i = key_index(k, t);
    const EnvironmentKey& k0 = k[i];
    const EnvironmentKey& k1 = k[std::min(i + 1, (int)k.size() - 1)];
    float u = safe_divide(t - k0.time, k1.time - k0.time);
    u = clampf(u, 0.0f, 1.0f);
    EnvironmentState s;
    s.coeff = lerp(k0.coeff, k1.coeff, u);
    s.rotation = lerp(k0.rotation, k1.rotation, u) * kPi / 180.0f;
    if (env.kind == EnvKind::AnalyticSky) {
        // rotate the first solar direction toward the second by the fraction u of their angle
        const float c = clampf(dot(k0.solar_direction, k1.solar_direction), -1.0f, 1.0f);
        const float angle = std::acos(c);
        const Vec3 axis = std::fabs(c) < 0.999f ? normalize(cross(k0.solar_direction, k1.solar_direction))
                                                : basis_x(k0.solar_direction);
        s.solar_direction = rotate(angle * u, axis, k0.solar_direction);
        s.turbidity = lerp(k0.turbidity, k1.turbidity, u);
        s.solar_elevation = 0.5f * kPi - std::acos(clampf(s.solar_direction.y, -1.0f, 1.0f));
    }
    return s;
}

}  // namespace rtc8x_driver

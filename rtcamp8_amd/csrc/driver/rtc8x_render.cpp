// rtc8x_render.cpp -- the batch renderer: the reference's no-GUI mode (runApp,
// renderer/renderer_main.cpp:1444-1667) over the rtc8x C ABI.
//
//   rtc8x_render -scene <file> [-spp N] [-t-mul X] [-no-gui] [-out DIR] [-max-frames N] [-no-nrc]
//                [-density-coeff C] [-albedo A] [-forwardness G] [-radiance-scale S]
//                [-encoding hash|triangle] [-ignore-surfaces] [-dump]
//
// For every time step t = time_begin + t_mul * frame / fps up to time_end: the active camera and the
// environment are evaluated at t, the sky is re-derived and re-baked when its sun or turbidity changed,
// the neural radiance cache trains (8 iterations on the first frame, then 1), N samples are traced into
// the cache and accumulated, and the frame is tone-mapped and written as DIR/%03u.png.  -no-nrc traces
// the plain unbiased path tracer instead.  -dump parses the scene and prints the evaluated cameras and
// environments per frame as JSON without touching the GPU.  The medium is the reference's batch setting
// (density coefficient 0.3, albedo 0.99, isotropic Schlick phase, renderer_main.cpp:1463-1467).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "rtc8x.h"
#include "scene_file.hpp"

using namespace rtc8x_driver;

namespace {

struct Options {
    std::string scene;
    uint32_t spp = 8;
    float t_mul = 1.0f;
    std::string out_dir = ".";
    int max_frames = -1;
    bool nrc = true;
    bool dump = false;
    bool ignore_surfaces = false;
    float density_coeff = 0.3f, albedo = 0.99f, forwardness = 0.0f, radiance_scale = 0.0f;
    int32_t encoding = RTC8X_NRC_HASH_GRID;
};

[[noreturn]] void usage(const char* msg)
{
    std::fprintf(stderr, "rtc8x_render: %s\n", msg);
    std::fprintf(stderr,
                 "usage: rtc8x_render -scene <file> [-spp N] [-t-mul X] [-out DIR] [-max-frames N] [-no-nrc]\n"
                 "                    [-density-coeff C] [-albedo A] [-forwardness G] [-radiance-scale S]\n"
                 "                    [-encoding hash|triangle] [-ignore-surfaces] [-dump] [-no-gui]\n");
    std::exit(2);
}

Options parse_args(int argc, char** argv)
{
    Options o;
    auto value = [&](int& i) -> const char* {
        if (i + 1 >= argc) usage((std::string("missing value after ") + argv[i]).c_str());
        return argv[++i];
    };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-no-gui") continue;
        else if (a == "-scene") o.scene = value(i);
        else if (a == "-spp") o.spp = (uint32_t)std::atoi(value(i));
        else if (a == "-t-mul") o.t_mul = (float)std::atof(value(i));
        else if (a == "-out") o.out_dir = value(i);
        else if (a == "-max-frames") o.max_frames = std::atoi(value(i));
        else if (a == "-no-nrc") o.nrc = false;
        else if (a == "-dump") o.dump = true;
        else if (a == "-ignore-surfaces") o.ignore_surfaces = true;
        else if (a == "-density-coeff") o.density_coeff = (float)std::atof(value(i));
        else if (a == "-albedo") o.albedo = (float)std::atof(value(i));
        else if (a == "-forwardness") o.forwardness = (float)std::atof(value(i));
        else if (a == "-radiance-scale") o.radiance_scale = (float)std::atof(value(i));
        else if (a == "-encoding") {
            const std::string e = value(i);
            if (e == "hash") o.encoding = RTC8X_NRC_HASH_GRID;
            else if (e == "triangle") o.encoding = RTC8X_NRC_TRIANGLE_WAVE;
            else usage("-encoding takes hash or triangle");
        } else usage(("unknown option " + a).c_str());
    }
    if (o.scene.empty()) usage("-scene is required");
    if (o.spp == 0) usage("-spp must be positive");
    if (!(o.t_mul > 0.0f)) usage("-t-mul must be positive");
    return o;
}

// ------------------------------------------------------------------------------------ GPU plumbing

void check(rtc8x_status st, const char* what)
{
    if (st != RTC8X_OK) {
        const char* msg = rtc8x_last_error();
        throw std::runtime_error(std::string(what) + ": " + (msg ? msg : "error"));
    }
}

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) : bytes(n)
    {
        check(rtc8x_malloc(&ptr, n), "rtc8x_malloc");
        check(rtc8x_memset(ptr, 0, n), "rtc8x_memset");
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            if (ptr) rtc8x_free(ptr);
            ptr = o.ptr;
            bytes = o.bytes;
            o.ptr = nullptr;
        }
        return *this;
    }
    ~DevBuf()
    {
        if (ptr) rtc8x_free(ptr);
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

std::vector<char> read_file(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    return std::vector<char>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

// The screen-size and training buffers of the NRC path (renderer_main.cpp:80-129, 197-230).
struct NrcBuffers {
    DevBuf num_training_data, train_image_size, target_min_max, target_avg, inference_queries,
        inference_terminals, train_suffix_terminals, inferred_radiance, per_frame_contribution, train_queries[2],
        train_targets[2], train_vertex_info, data_shufflers, scratch;

    // `samples`: image slices of the render buffers, the samples one render launch traces
    NrcBuffers(uint32_t w, uint32_t h, uint32_t samples)
    {
        const size_t n = (size_t)w * h, ninf = round_up(n, 128) * samples;
        num_training_data = DevBuf(8);
        train_image_size = DevBuf(16);
        target_min_max = DevBuf(48);
        target_avg = DevBuf(24);
        inference_queries = DevBuf(sizeof(rtc8x_radiance_query) * ninf);
        inference_terminals = DevBuf(sizeof(rtc8x_terminal_info) * n * samples);
        train_suffix_terminals = DevBuf(sizeof(rtc8x_training_suffix_terminal) * n);
        inferred_radiance = DevBuf(12 * ninf);
        per_frame_contribution = DevBuf(12 * n * samples);
        for (int i = 0; i < 2; ++i) {
            train_queries[i] = DevBuf(sizeof(rtc8x_radiance_query) * RTC8X_TRAIN_BUFFER_SIZE);
            train_targets[i] = DevBuf(12 * (size_t)RTC8X_TRAIN_BUFFER_SIZE);
        }
        train_vertex_info = DevBuf(sizeof(rtc8x_training_vertex_info) * RTC8X_TRAIN_BUFFER_SIZE);
        // LCG shufflers: slot i holds the state i + 1 steps after the seed (renderer_main.cpp:220-229)
        std::vector<uint32_t> states(RTC8X_NUM_TRAINING_DATA_PER_FRAME);
        uint32_t s = 471313181u;
        for (uint32_t& v : states) {
            s = (s * 1103515245u + 12345u) % (1u << 31);
            v = s;
        }
        data_shufflers = DevBuf(4 * states.size());
        check(rtc8x_memcpy_htod(data_shufflers.ptr, states.data(), 4 * states.size()), "upload shufflers");
        scratch = DevBuf(std::max<size_t>(rtc8x_nrc_scratch_bytes(w, h), 4));
    }

    void fill(rtc8x_nrc_buffers& nb, uint32_t w, uint32_t h) const
    {
        nb.max_num_training_suffixes = w * h;
        nb.num_training_data = num_training_data.as<uint32_t>();
        nb.train_image_size = train_image_size.as<int32_t>();
        nb.target_min_max = target_min_max.as<float>();
        nb.target_avg = target_avg.as<float>();
        nb.inference_queries = inference_queries.as<rtc8x_radiance_query>();
        nb.inference_terminals = inference_terminals.as<rtc8x_terminal_info>();
        nb.train_suffix_terminals = train_suffix_terminals.as<rtc8x_training_suffix_terminal>();
        nb.inferred_radiance = inferred_radiance.as<float>();
        nb.per_frame_contribution = per_frame_contribution.as<float>();
        for (int i = 0; i < 2; ++i) {
            nb.train_queries[i] = train_queries[i].as<rtc8x_radiance_query>();
            nb.train_targets[i] = train_targets[i].as<float>();
        }
        nb.train_vertex_info = train_vertex_info.as<rtc8x_training_vertex_info>();
        nb.data_shufflers = data_shufflers.as<uint32_t>();
        nb.scratch = scratch.as<uint32_t>();
    }
};

// -------------------------------------------------------------------------------------------- dump

std::string json_vec(Vec3 v)
{
    char b[128];
    std::snprintf(b, sizeof(b), "[%.9g, %.9g, %.9g]", v.x, v.y, v.z);
    return b;
}

// The active camera at time t (renderer_main.cpp:1518-1525): switches when t reaches the next entry.
struct CameraTrack {
    const Scene& scene;
    uint32_t index = 0;
    std::string current;
    explicit CameraTrack(const Scene& s) : scene(s), current(s.active_cameras.front().name) {}
    const CameraDesc& at(float t)
    {
        const uint32_t n = (uint32_t)scene.active_cameras.size();
        const ActiveCamera& next = scene.active_cameras[std::min(index + 1, n - 1)];
        if (t >= next.time) {
            if (next.name != current) ++index;
            current = next.name;
        }
        return scene.cameras.at(current);
    }
};

std::vector<float> time_points(const Scene& s, const Options& o)
{
    std::vector<float> ts;
    for (uint32_t i = 0;; ++i) {
        if (o.max_frames >= 0 && (int)i >= o.max_frames) break;
        const float t = s.time_begin + o.t_mul * (float)i / (float)s.fps;
        if (t > s.time_end) break;
        ts.push_back(t);
    }
    return ts;
}

void validate(const Scene& s, const Options& o)
{
    if (s.width == 0 || s.height == 0) throw std::runtime_error("the scene sets no image size (image)");
    if (s.fps == 0) throw std::runtime_error("the scene sets no frame rate (time)");
    if (s.cameras.empty()) throw std::runtime_error("the scene has no camera");
    if (!s.instances.empty() && !o.ignore_surfaces)
        throw std::runtime_error("the scene instantiates surface meshes; this renderer traces volumes only "
                                 "(pass -ignore-surfaces to render the volume and environment without them)");
    if (s.environment.kind != EnvKind::None && s.environment.keys.empty())
        throw std::runtime_error("the environment has no key state (env-addkey)");
}

int dump(const Scene& s, const Options& o)
{
    validate(s, o);
    std::printf("{\"image\": [%u, %u], \"time\": [%.9g, %.9g], \"fps\": %u, \"volume\": \"%s\", ", s.width, s.height,
                s.time_begin, s.time_end, s.fps, s.volume_file.c_str());
    const char* kinds[] = {"none", "file", "analytic_sky"};
    std::printf("\"environment\": \"%s\", \"cameras\": %zu, \"instances\": %zu, \"frames\": [",
                kinds[(int)s.environment.kind], s.cameras.size(), s.instances.size());
    CameraTrack track(s);
    const std::vector<float> ts = time_points(s, o);
    for (size_t i = 0; i < ts.size(); ++i) {
        const CameraDesc& cam = track.at(ts[i]);
        const CameraState c = camera_at(cam, ts[i]);
        std::printf("%s\n {\"frame\": %zu, \"time\": %.9g, \"camera\": \"%s\", \"position\": %s, "
                    "\"orientation\": [%.9g, %.9g, %.9g, %.9g], \"fov_y\": %.9g",
                    i ? "," : "", i, ts[i], track.current.c_str(), json_vec(c.position).c_str(), c.orientation.x,
                    c.orientation.y, c.orientation.z, c.orientation.w, c.fov_y);
        if (s.environment.kind != EnvKind::None) {
            const EnvironmentState e = environment_at(s.environment, ts[i]);
            std::printf(", \"env_coeff\": %.9g, \"env_rotation\": %.9g", e.coeff, e.rotation);
            if (s.environment.kind == EnvKind::AnalyticSky)
                std::printf(", \"sun\": %s, \"turbidity\": %.9g, \"solar_elevation\": %.9g",
                            json_vec(e.solar_direction).c_str(), e.turbidity, e.solar_elevation);
        }
        std::printf("}");
    }
    std::printf("]}\n");
    return 0;
}

// ------------------------------------------------------------------------------------------ render

constexpr uint32_t kSkyWidth = 3200, kSkyHeight = 1600;  // AnalyticSkyEnvironment(uint2(3200, 1600))

std::string exe_dir(const char* argv0)
{
    std::error_code ec;
    auto p = std::filesystem::canonical("/proc/self/exe", ec);
    if (ec) p = std::filesystem::absolute(argv0);
    return p.parent_path().string();
}

int render(const Scene& s, const Options& o, const std::string& data_dir)
{
    validate(s, o);
    if (s.volume_file.empty()) throw std::runtime_error("the scene has no volume (volume \"file.nvdb\")");
    if (s.environment.kind == EnvKind::File)
        throw std::runtime_error("image-based environments (envmap) are not supported; use skyenv");
    std::filesystem::create_directories(o.out_dir);
    const uint32_t W = s.width, H = s.height;
    const size_t npx = (size_t)W * H;
    void* stream = nullptr;
    check(rtc8x_stream_create(&stream), "stream");

    // volume (readGrid + allocateVolumeGrid, renderer_scene.cpp:2467-2470)
    std::vector<char> vol = read_file(s.volume_file);
    rtc8x_grid grid;
    std::memset(&grid, 0, sizeof(grid));
    check(rtc8x_grid_load_nvdb(vol.data(), vol.size(), 0, &grid), "load volume");
    vol.clear();
    vol.shrink_to_fit();

    rtc8x_frame_params p;
    std::memset(&p, 0, sizeof(p));
    p.image_width = W;
    p.image_height = H;
    p.grid = grid;
    p.density_coeff = o.density_coeff;
    p.scattering_albedo = o.albedo;
    p.scattering_forwardness = o.forwardness;
    p.phase = RTC8X_PHASE_SCHLICK;
    // WorldDimInfo of a volume-only scene: the volume's bbox and its half-diagonal
    float center[3], r2 = 0.0f;
    for (int a = 0; a < 3; ++a) {
        p.world_aabb_min[a] = grid.world_bbox_min[a];
        p.world_aabb_max[a] = grid.world_bbox_max[a];
        center[a] = (grid.world_bbox_min[a] + grid.world_bbox_max[a]) * 0.5f;
        const float d = grid.world_bbox_max[a] - center[a];
        r2 += d * d;
    }
    p.world_radius = std::sqrt(r2);
    p.rr_after_path_length = 0;
    p.max_path_length = 0;
    p.tracking = RTC8X_TRACKING_GLOBAL_MAJORANT;
    p.nrc.path_termination_factor = 0.01f;
    p.nrc.radiance_scale = o.radiance_scale;

    // per-pixel state
    DevBuf rng(8 * npx), accum(24 * npx), stats(8 * RTC8X_STAT_COUNT), rgba8(4 * npx), cost(4 * npx),
        order(4 * npx), order_scratch(std::max<size_t>(rtc8x_order_tiles_scratch_bytes(W, H), 4));
    check(rtc8x_seed_rng(rng.as<uint64_t>(), W, H, 591842031321323413ull), "seed");
    p.rng_states = rng.as<uint64_t>();
    p.accum = accum.as<double>();
    p.stats = stats.as<uint64_t>();

    // environment
    rtc8x_hosek_model* sky_model = nullptr;
    DevBuf sky_tex;
    bool have_sampler = false;
    float baked_elev = NAN, baked_turb = NAN;
    Vec3 baked_sun{NAN, NAN, NAN};
    if (s.environment.kind == EnvKind::AnalyticSky) {
        std::vector<char> blob = read_file(data_dir + "/data/hosek_model.bin");
        check(rtc8x_hosek_model_create(blob.data(), blob.size(), &sky_model), "sky model");
        sky_tex = DevBuf(16 * (size_t)kSkyWidth * kSkyHeight);
    }

    // neural radiance cache; a frame's samples go through the cache in launches of up to 16
    const uint32_t samples_per_launch = std::min<uint32_t>(o.spp, 16);
    rtc8x_nrc* cache = nullptr;
    std::unique_ptr<NrcBuffers> nb;
    if (o.nrc) {
        rtc8x_nrc_desc d;
        std::memset(&d, 0, sizeof(d));
        d.num_hidden_layers = 2;
        d.learning_rate = 1e-2f;
        d.beta1 = 0.9f;
        d.beta2 = 0.99f;
        d.epsilon = o.encoding == RTC8X_NRC_HASH_GRID ? 1e-15f : 1e-8f;
        d.l2_reg = 1e-6f;
        d.ema_decay = 0.99f;
        d.precision = RTC8X_NRC_FP16;
        d.seed = 0x9E3779B97F4A7C15ull;
        d.encoding = o.encoding;
        check(rtc8x_nrc_create(&d, &cache), "nrc create");
        nb = std::make_unique<NrcBuffers>(W, H, samples_per_launch);
        nb->fill(p.nrc, W, H);
    }
    std::mt19937 frame_rng(72139121u);
    uint32_t buf_idx = 1;
    bool have_cost = false;

    rtc8x_tone_map_params tm;
    std::memset(&tm, 0, sizeof(tm));
    tm.accum = accum.as<double>();
    tm.width = W;
    tm.height = H;
    tm.samples = o.spp;
    tm.exponent = 1.0f / 0.6f;
    tm.brightness = 1.0f;
    tm.out_rgba8 = rgba8.as<uint32_t>();
    tm.sdr.alpha_override = 1.0f;  // the batch saver's settings (renderer_main.cpp:1490-1495)
    tm.sdr.brightness_scale = 1.0f;
    tm.sdr.apply_tone_map = 0;
    tm.sdr.apply_srgb_gamma = 1;
    tm.sdr.flip_y = 0;

    CameraTrack track(s);
    const std::vector<float> ts = time_points(s, o);
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    std::vector<uint32_t> pixels(npx);
    std::vector<uint8_t> png;
    for (uint32_t frame = 0; frame < ts.size(); ++frame) {
        const float t = ts[frame];
        std::printf("Frame %u (Time: %.3f) (tp: %.3f [s])\n", frame, t, elapsed());
        std::fflush(stdout);
        const CameraState cam = camera_at(track.at(t), t);
        const float pos[3] = {cam.position.x, cam.position.y, cam.position.z};
        const float q[4] = {cam.orientation.x, cam.orientation.y, cam.orientation.z, cam.orientation.w};
        check(rtc8x_camera_setup(&p.camera, pos, q, cam.fov_y, (float)W / (float)H), "camera");

        if (s.environment.kind == EnvKind::AnalyticSky) {
            const EnvironmentState e = environment_at(s.environment, t);
            const float elev = std::min(std::max(e.solar_elevation, 0.0f), 1.5707963f);
            if (elev != baked_elev || e.turbidity != baked_turb || e.solar_direction.x != baked_sun.x ||
                e.solar_direction.y != baked_sun.y || e.solar_direction.z != baked_sun.z) {
                rtc8x_hosek_sky sky;
                std::memset(&sky, 0, sizeof(sky));
                check(rtc8x_hosek_model_state(sky_model, elev, e.turbidity, &sky, stream), "sky state");
                const float sun[3] = {e.solar_direction.x, e.solar_direction.y, e.solar_direction.z};
                check(rtc8x_sky_bake(stream, &sky, sun, kSkyWidth, kSkyHeight, sky_tex.as<float>()), "sky bake");
                if (have_sampler) check(rtc8x_env_release(&p.env), "env release");
                std::memset(&p.env, 0, sizeof(p.env));
                p.env.kind = RTC8X_ENV_LATLONG;
                p.env.sampler = RTC8X_ENV_SAMPLER_HIERARCHICAL;
                p.env.width = kSkyWidth;
                p.env.height = kSkyHeight;
                p.env.texels = sky_tex.as<float>();
                check(rtc8x_env_build_sampler(&p.env, stream), "env sampler");
                have_sampler = true;
                baked_elev = elev;
                baked_turb = e.turbidity;
                baked_sun = e.solar_direction;
            }
            p.env.power_coeff = e.coeff;
            p.env.rotation = e.rotation;
        } else {
            p.env.kind = RTC8X_ENV_NONE;
        }

        check(rtc8x_memset(accum.ptr, 0, accum.bytes), "clear accumulator");
        if (o.nrc) {
            const uint32_t iters = frame == 0 ? 8 : 1;
            for (uint32_t it = 0; it < iters; ++it) {
                buf_idx = (buf_idx + 1) % 2;
                p.nrc.nrc_buffer_index = buf_idx;
                p.nrc.offset_to_select_unbiased_path = frame_rng();
                p.work_order = nullptr;
                p.pixel_cost = nullptr;
                p.nrc.num_samples = 1;
                check(rtc8x_nrc_prepare(stream, &p, frame == 0 ? 1 : 0), "prepare");
                check(rtc8x_trace(stream, &p, RTC8X_ENTRY_GENERATE_TRAINING_DATA, 1), "training trace");
                // the training size stays on the device: no host synchronisation in the iteration
                check(rtc8x_nrc_infer_training(cache, stream, &p), "infer");
                check(rtc8x_nrc_propagate(stream, &p), "propagate");
                check(rtc8x_nrc_shuffle(stream, &p), "shuffle");
                const uint32_t batch = RTC8X_NUM_TRAINING_DATA_PER_FRAME / 4;
                for (uint32_t step = 0; step < 4; ++step)
                    check(rtc8x_nrc_train(cache, stream, (const float*)(nb->train_queries[1].as<rtc8x_radiance_query>() + step * batch),
                                          nb->train_targets[1].as<float>() + 3 * (size_t)step * batch, batch, nullptr),
                          "train");
            }
            for (uint32_t done = 0; done < o.spp;) {
                const uint32_t k = std::min(samples_per_launch, o.spp - done);
                p.nrc.num_samples = k;
                if (have_cost) {
                    check(rtc8x_order_by_tile_cost(stream, cost.as<uint32_t>(), W, H, order.as<uint32_t>(),
                                                   order_scratch.ptr), "order");
                    p.work_order = order.as<uint32_t>();
                } else {
                    p.work_order = nullptr;
                }
                p.pixel_cost = cost.as<uint32_t>();
                have_cost = true;
                check(rtc8x_trace(stream, &p, RTC8X_ENTRY_PATH_TRACE_WITH_NRC, k), "trace with NRC");
                check(rtc8x_nrc_infer_render(cache, stream, &p), "infer");
                check(rtc8x_nrc_accumulate(stream, &p), "accumulate");
                done += k;
            }
        } else {
            if (have_cost) {
                check(rtc8x_order_by_tile_cost(stream, cost.as<uint32_t>(), W, H, order.as<uint32_t>(),
                                               order_scratch.ptr), "order");
                p.work_order = order.as<uint32_t>();
            }
            p.pixel_cost = cost.as<uint32_t>();
            have_cost = true;
            check(rtc8x_trace(stream, &p, RTC8X_ENTRY_PATH_TRACE, o.spp), "trace");
        }

        check(rtc8x_tone_map(stream, &tm), "tone map");
        check(rtc8x_stream_synchronize(stream), "sync");
        check(rtc8x_memcpy_dtoh(pixels.data(), rgba8.ptr, 4 * npx), "download frame");
        size_t bound = 0, size = 0;
        check(rtc8x_png_encode(pixels.data(), W, H, nullptr, 0, &bound), "png size");
        png.resize(bound);
        check(rtc8x_png_encode(pixels.data(), W, H, png.data(), png.size(), &size), "png encode");
        char name[64];
        std::snprintf(name, sizeof(name), "%03u.png", frame);
        const std::string path = (std::filesystem::path(o.out_dir) / name).string();
        std::ofstream out(path, std::ios::binary);
        out.write((const char*)png.data(), (std::streamsize)size);
        if (!out) throw std::runtime_error("cannot write " + path);
    }
    std::printf("Done: %zu frames (tp: %.3f [s])\n", ts.size(), elapsed());

    if (cache) rtc8x_nrc_destroy(cache);
    if (have_sampler) rtc8x_env_release(&p.env);
    if (sky_model) rtc8x_hosek_model_destroy(sky_model);
    rtc8x_grid_destroy(&grid);
    rtc8x_stream_destroy(stream);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    const Options o = parse_args(argc, argv);
    try {
        const Scene s = load_scene(o.scene);
        if (o.dump) return dump(s, o);
        return render(s, o, exe_dir(argv[0]));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "rtc8x_render: %s\n", e.what());
        return 1;
    }
}

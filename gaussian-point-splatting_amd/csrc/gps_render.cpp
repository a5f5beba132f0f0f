// gps_render — headless frame driver over the C ABI.
//
// Keeps the reference's command-line surface (src/parse_args.cuh:21-127) and
// frame loop (src/main.cu:79-252, src/core/rendering/render_server.cu:86-196)
// without a window: frames are written as PPM (RGB8, bottom row first flipped
// as save_image_to_disk does) and PFM (linear colour).  Flags of the
// interactive viewer that have no headless meaning are accepted and ignored.
// Headless additions: --synthetic, --num-gaussians, --seed, --frames, --out,
// --format, --orbit, --eye/--center, --taa-current-depth,
// --frames-per-camera-keyframe, --total-test-keyframes, --screenshot-dir
// (base folder of the evaluation screenshots, default out/screenshots).
//
// Camera paths (--camera-path): the first camera sets the resolution (the
// width is kept, the height follows its aspect ratio) and every camera its
// fovy, unless --resolution or --fovy forces them (camera_loader.cu:73-88).
// --play-camera-path plays the keyframes 1/24 s apart at one keyframe per
// frame; --run-tests steps through the cameras in main.cu's order (the
// selection advances after each frame, clamped at the last camera, and
// restarts at 0 every total_test_keyframes frames; the first lap is not
// timed) and prints the per-camera timing JSON of main.cu:211-241.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <errno.h>
#include <sys/stat.h>

#include "../../include/gps.h"

namespace {

struct Options {
    // RenderSettings (src/render_settings.cuh:36-88)
    int width = 1920, height = 1080;
    int render_pass_count = 1;
    int supersampling_factor = 2;
    long long max_number_of_points = 250000000;
    bool enable_occlusion_culling = true;
    bool enable_hierarchical_culling = true;
    bool cull_small = false;
    bool reduce_point_count = true;
    bool taa = false;
    bool run_tests = false;
    float near_plane = 0.2f, far_plane = 10000.0f, fovy = 45.0f;
    float background[3] = {0, 0, 0};
    float rotation[3] = {0, 0, 0}, translation[3] = {0, 0, 0}, scaling[3] = {1, 1, 1};
    std::string model_path;
    bool sort_morton = false;
    std::string sorted_model_output_path;
    // headless
    std::string synthetic = "c1";
    long long num_gaussians = -1;
    unsigned long long seed = 0;
    int frames = 1;
    std::string out;
    int format = GPS_FORMAT_FP32;
    float orbit_radius = -1.0f, orbit_elevation = 20.0f, orbit_step = 3.0f;
    float eye[3] = {0, 0, -30}, center[3] = {0, 0, 0};
    bool taa_current_depth = false;
    int gaussians_per_block = 0;
    int accumulation_none = 0;
    // camera paths (render_settings.cuh:42-79)
    std::string camera_path;
    bool play_camera_path = false;
    bool override_camera_settings = false;
    int frames_per_camera_keyframe = 10;
    int total_test_keyframes = 0;
    // evaluation screenshots (main.cu:72-120, render_server.cu:263-315)
    bool create_evaluation_screenshots = false;
    int max_eval_count = 40;
    std::string screenshot_dir = "out/screenshots";
};

bool parse_vec(int& i, int argc, char** argv, float* out, int n) {
    if (i + n >= argc) return false;
    for (int k = 0; k < n; ++k) out[k] = std::stof(argv[++i]);
    return true;
}

Options parse(int argc, char** argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string { return (i + 1 < argc) ? std::string(argv[++i]) : std::string(); };
        if (a == "--background") parse_vec(i, argc, argv, o.background, 3);
        else if (a == "--rotation") parse_vec(i, argc, argv, o.rotation, 3);
        else if (a == "--translation") parse_vec(i, argc, argv, o.translation, 3);
        else if (a == "--scaling") parse_vec(i, argc, argv, o.scaling, 3);
        else if (a == "--fovy") {
            o.fovy = std::stof(next());
            o.override_camera_settings = true;
        }
        else if (a == "--near-view-plane" || a == "--near_view_plane") o.near_plane = std::stof(next());
        else if (a == "--far-view-plane" || a == "--far_view_plane") o.far_plane = std::stof(next());
        else if (a == "--samples-per-pixel" || a == "--samples_per_pixel") o.render_pass_count = std::stoi(next());
        else if (a == "--max-number-of-points" || a == "--max_number_of_points") o.max_number_of_points = std::stoll(next());
        else if (a == "--supersampling-factor" || a == "--supersampling_factor") o.supersampling_factor = std::stoi(next());
        else if (a == "--resolution") {
            float r[2];
            parse_vec(i, argc, argv, r, 2);
            o.width = (int)r[0];
            o.height = (int)r[1];
            o.override_camera_settings = true;
        } else if (a == "--set-target-width" || a == "--set_target_width") o.width = std::stoi(next());
        else if (a == "--cull-small") o.cull_small = true;
        else if (a == "--one_point_per_thread" || a == "--one-point-per-thread") o.reduce_point_count = false;
        else if (a == "--disable_hierarchical_culling" || a == "--disable-hierarchical-culling") o.enable_hierarchical_culling = false;
        else if (a == "--disable_occlusion_culling" || a == "--disable-occlusion-culling") o.enable_occlusion_culling = false;
        else if (a == "--taa") o.taa = true;
        else if (a == "--run_tests" || a == "--run-tests") o.run_tests = true;
        else if (a == "--model-path") o.model_path = next();
        else if (a == "--sort_morton_order" || a == "--sort-morton-order") o.sort_morton = true;
        else if (a == "--cameras-path" || a == "--camera-path") o.camera_path = next();
        else if (a == "--play_camera_path" || a == "--play-camera-path") o.play_camera_path = true;
        else if (a == "--frames-per-camera-keyframe") o.frames_per_camera_keyframe = std::stoi(next());
        else if (a == "--total-test-keyframes") o.total_test_keyframes = std::stoi(next());
        else if (a == "--sorted-model-output-path" || a == "--sorted_model_output_path")
            o.sorted_model_output_path = next();
        else if (a == "--max-eval-count" || a == "--max_eval_count") o.max_eval_count = std::stoi(next());
        else if (a == "--create-evaluation-screenshots" || a == "--create_evaluation_screenshots")
            o.create_evaluation_screenshots = true;
        else if (a == "--screenshot-dir") o.screenshot_dir = next();
        else if (a == "--freeze-culling" || a == "--enable_profiling" || a == "--disable_gui" || a == "--disable-gui" ||
                 a == "--snap_window_to_corner" || a == "--snap-window-to-corner") {
        }
        // headless additions
        else if (a == "--synthetic") o.synthetic = next();
        else if (a == "--num-gaussians") o.num_gaussians = std::stoll(next());
        else if (a == "--seed") o.seed = std::stoull(next());
        else if (a == "--frames") o.frames = std::stoi(next());
        else if (a == "--out") o.out = next();
        else if (a == "--format") {
            const std::string f = next();
            o.format = f == "half_fp8" ? GPS_FORMAT_HALF_FP8 : f == "quantized" ? GPS_FORMAT_QUANTIZED : GPS_FORMAT_FP32;
        }
        else if (a == "--orbit") {
            float r[2];
            parse_vec(i, argc, argv, r, 2);
            o.orbit_radius = r[0];
            o.orbit_elevation = r[1];
        } else if (a == "--eye") parse_vec(i, argc, argv, o.eye, 3);
        else if (a == "--center") parse_vec(i, argc, argv, o.center, 3);
        else if (a == "--taa-current-depth") o.taa_current_depth = true;
        else if (a == "--gaussians-per-block") o.gaussians_per_block = std::stoi(next());
        else if (a == "--accumulation-none") o.accumulation_none = 1;
        else std::printf("Unknown argument: %s\n", a.c_str());
    }
    return o;
}

#define CHECK(x)                                                                                   \
    do {                                                                                           \
        gps_status s_ = (x);                                                                       \
        if (s_ != GPS_OK) {                                                                        \
            std::fprintf(stderr, "%s failed: %s (%s)\n", #x, gps_status_str(s_), gps_last_error()); \
            std::exit(1);                                                                          \
        }                                                                                          \
    } while (0)

// Transform::getModelMatrix (src/world/transform.cpp:35-44) of the scene
// transform: translate * mat4_cast(quat(radians(euler))) * scale.
void scene_matrix(const Options& o, float out[16]) {
    const float rx = o.rotation[0] * 0.01745329251994329576923690768489f;
    const float ry = o.rotation[1] * 0.01745329251994329576923690768489f;
    const float rz = o.rotation[2] * 0.01745329251994329576923690768489f;
    // glm quat from euler angles (pitch, yaw, roll)
    const float cx = std::cos(rx * 0.5f), sx = std::sin(rx * 0.5f);
    const float cy = std::cos(ry * 0.5f), sy = std::sin(ry * 0.5f);
    const float cz = std::cos(rz * 0.5f), sz = std::sin(rz * 0.5f);
    const float w = cx * cy * cz + sx * sy * sz;
    const float x = sx * cy * cz - cx * sy * sz;
    const float y = cx * sy * cz + sx * cy * sz;
    const float z = cx * cy * sz - sx * sy * cz;
    const float R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y),
                        2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x),
                        2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)};
    std::memset(out, 0, 64);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) out[4 * c + r] = R[3 * c + r] * o.scaling[c];
    out[12] = o.translation[0];
    out[13] = o.translation[1];
    out[14] = o.translation[2];
    out[15] = 1.0f;
}

void write_ppm(const std::string& path, const std::vector<uint8_t>& rgb, int w, int h) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return;
    std::fprintf(f, "P6\n%d %d\n255\n", w, h);
    for (int y = h - 1; y >= 0; --y) std::fwrite(&rgb[(size_t)y * w * 3], 1, (size_t)w * 3, f);
    std::fclose(f);
}

void write_pfm(const std::string& path, const std::vector<float>& c, int w, int h, float scale) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return;
    std::fprintf(f, "PF\n%d %d\n-1.0\n", w, h);
    std::vector<float> row((size_t)w * 3);
    for (int y = 0; y < h; ++y) {
        for (int i = 0; i < w * 3; ++i) row[i] = c[(size_t)y * w * 3 + i] * scale;
        std::fwrite(row.data(), 4, row.size(), f);
    }
    std::fclose(f);
}

// PNG (RGB8, no interlace) with the image data in stored (uncompressed)
// deflate blocks: a valid zlib stream without a compressor.  Rows are
// written last row first, as save_image_to_disk flips the resolve's image
// (render_server.cu:268-277) before stbi_write_png.
uint32_t crc32_update(uint32_t c, const uint8_t* p, size_t n) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1u) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            table[i] = v;
        }
        init = true;
    }
    c = ~c;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return ~c;
}

bool write_png(const std::string& path, const std::vector<uint8_t>& rgb, int w, int h) {
    std::vector<uint8_t> raw;
    raw.reserve((size_t)h * ((size_t)w * 3 + 1));
    for (int y = h - 1; y >= 0; --y) {
        raw.push_back(0);  // filter: none
        raw.insert(raw.end(), rgb.begin() + (size_t)y * w * 3, rgb.begin() + (size_t)(y + 1) * w * 3);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    for (size_t off = 0; off < raw.size() || off == 0;) {
        const size_t len = std::min<size_t>(65535, raw.size() - off);
        const bool last = off + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)(len & 0xFF));
        z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)(~len & 0xFF));
        z.push_back((uint8_t)((~len >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + len);
        off += len;
        if (last) break;
    }
    uint32_t a = 1, b = 0;  // Adler-32 of the uncompressed data
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    const uint32_t adler = (b << 16) | a;
    for (int k = 3; k >= 0; --k) z.push_back((uint8_t)(adler >> (8 * k)));
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    auto be32 = [&](uint32_t v) {
        const uint8_t q[4] = {(uint8_t)(v >> 24), (uint8_t)(v >> 16), (uint8_t)(v >> 8), (uint8_t)v};
        std::fwrite(q, 1, 4, f);
    };
    auto chunk = [&](const char* type, const std::vector<uint8_t>& data) {
        be32((uint32_t)data.size());
        std::fwrite(type, 1, 4, f);
        if (!data.empty()) std::fwrite(data.data(), 1, data.size(), f);
        uint32_t c = crc32_update(0, reinterpret_cast<const uint8_t*>(type), 4);
        c = crc32_update(c, data.data(), data.size());
        be32(c);
    };
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    std::fwrite(sig, 1, 8, f);
    std::vector<uint8_t> ihdr = {(uint8_t)(w >> 24), (uint8_t)(w >> 16), (uint8_t)(w >> 8), (uint8_t)w,
                                 (uint8_t)(h >> 24), (uint8_t)(h >> 16), (uint8_t)(h >> 8), (uint8_t)h,
                                 8, 2, 0, 0, 0};
    chunk("IHDR", ihdr);
    chunk("IDAT", z);
    chunk("IEND", {});
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    return ok;
}

// RenderSettings::model_name (render_settings.cuh:155-198): a model
// directory's name or the PLY file's stem; synthetic scenes are named after
// their configuration.
std::string model_name(const Options& o) {
    if (o.model_path.empty()) return "synthetic_" + o.synthetic;
    std::string p = o.model_path;
    while (!p.empty() && (p.back() == '/' || p.back() == '\\')) p.pop_back();
    const size_t slash = p.find_last_of("/\\");
    std::string base = slash == std::string::npos ? p : p.substr(slash + 1);
    const size_t dot = base.find_last_of('.');
    if (dot != std::string::npos && dot > 0) base = base.substr(0, dot);
    return base;
}

bool make_dirs(const std::string& path) {
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty() && cur != ".") {
                if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return false;
            }
        }
        if (i < path.size()) cur.push_back(path[i]);
    }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    Options o = parse(argc, argv);
    // ---- scene: a 3DGS PLY file (gaussian_loader.cu) or a synthetic scene
    gps_synth_params sp;
    std::memset(&sp, 0, sizeof(sp));
    sp.seed = o.seed ? o.seed : 0x5EED0001ull;
    if (o.synthetic == "c2" || o.synthetic == "clustered") {
        sp.kind = 1;
        sp.num_gaussians = 10000000;
        sp.sh_degree = 3;
        sp.clusters = 2000;
        sp.half_extent = 100.0;
        sp.cluster_sigma = 2.0;
        sp.log_scale_mu = -3.0;
        sp.log_scale_sigma = 0.7;
        sp.opacity_sigma = 1.5;
        sp.dc_sigma = 0.3;
        sp.rest_sigma = 0.05;
        if (!o.seed) sp.seed = 0x5EED0002ull;
    } else {
        sp.kind = 0;
        sp.num_gaussians = 1000000;
        sp.sh_degree = 0;
        sp.half_extent = 10.0;
        sp.log_scale_mu = -4.0;
        sp.log_scale_sigma = 0.5;
        sp.opacity_sigma = 1.0;
        sp.dc_sigma = 0.3;
    }
    if (o.num_gaussians >= 0) sp.num_gaussians = (uint64_t)o.num_gaussians;
    gps_ply_info info{};
    if (!o.model_path.empty()) {
        CHECK(gps_ply_info_read(o.model_path.c_str(), &info));
        sp.num_gaussians = info.num_gaussians;
        sp.sh_degree = 3;
    }
    const uint64_t N = sp.num_gaussians;
    const int nc = (sp.sh_degree + 1) * (sp.sh_degree + 1);
    std::vector<float> means(3 * N), rots(4 * N), scales(3 * N), ops(N), shs((size_t)3 * nc * N);
    auto apply_perm = [&](const std::vector<uint32_t>& perm) {
        auto permute = [&](std::vector<float>& a, size_t w) {
            std::vector<float> t(a.size());
            for (uint64_t i = 0; i < N; ++i) std::memcpy(&t[w * i], &a[w * perm[i]], 4 * w);
            a.swap(t);
        };
        permute(means, 3);
        permute(rots, 4);
        permute(scales, 3);
        permute(ops, 1);
        permute(shs, (size_t)3 * nc);
    };
    auto fill_desc = [&](gps_scene_desc& d) {
        d.num_gaussians = N;
        d.sh_degree = sp.sh_degree;
        d.format = o.format;
        d.means = means.data();
        d.rotations = rots.data();
        d.scales = scales.data();
        d.opacities = ops.data();
        d.shs = shs.data();
    };
    if (!o.model_path.empty()) {
        uint64_t got = 0;
        if (!o.sorted_model_output_path.empty()) {
            // main.cu:266-273: load without Morton order, BVH reorder, save
            // with the "bvh_sorted_order" comment
            CHECK(gps_ply_load(o.model_path.c_str(), 0, N, means.data(), rots.data(), scales.data(), ops.data(),
                               shs.data(), &got));
            if (N > 1) {
                std::vector<uint32_t> perm(N);
                CHECK(gps_bvh_order(means.data(), rots.data(), scales.data(), N, perm.data()));
                apply_perm(perm);
            }
            gps_scene_desc sd;
            fill_desc(sd);
            CHECK(gps_ply_write(o.sorted_model_output_path.c_str(), &sd, 1));
            std::printf("Wrote BVH-sorted model: %s\n", o.sorted_model_output_path.c_str());
        }
        CHECK(gps_ply_load(o.model_path.c_str(), o.sort_morton ? 1 : 0, N, means.data(), rots.data(), scales.data(),
                           ops.data(), shs.data(), &got));
        // gaussian_point_splatting.cu:791-797: the BVH reorder unless the
        // scene was Morton-sorted on load or carries "bvh_sorted_order"
        if (!o.sort_morton && !info.bvh_sorted && N > 1) {
            std::vector<uint32_t> perm(N);
            CHECK(gps_bvh_order(means.data(), rots.data(), scales.data(), N, perm.data()));
            apply_perm(perm);
        }
        std::printf("is pre-sorted: %s\n", info.bvh_sorted ? "True" : "false");
    } else {
        CHECK(gps_synth_generate(&sp, 0, N, means.data(), rots.data(), scales.data(), ops.data(), shs.data()));
        if (o.sort_morton && N > 1) {  // gaussian_loader.cu:363-426
            std::vector<uint32_t> perm(N);
            CHECK(gps_morton_order(means.data(), N, perm.data()));
            apply_perm(perm);
        }
    }
    gps_scene_desc d;
    fill_desc(d);
    gps_scene* scene = nullptr;
    CHECK(gps_scene_create(&d, &scene));
    // camera path (RenderSettings::load_camera, render_settings.cuh:248-256)
    std::vector<gps_camera_state> cams;
    if (!o.camera_path.empty()) {
        int32_t n = 0;
        if (gps_camera_path_load(o.camera_path.c_str(), nullptr, 0, &n) == GPS_OK && n > 0) {
            cams.resize((size_t)n);
            CHECK(gps_camera_path_load(o.camera_path.c_str(), cams.data(), n, &n));
        } else {
            std::fprintf(stderr, "Error reading camera_states file: %s\n", o.camera_path.c_str());
        }
        if (!cams.empty() && !o.override_camera_settings) {
            int32_t w = 0, h = 0;
            CHECK(gps_camera_resolution(&cams[0], o.width, &w, &h));
            o.width = w;
            o.height = h;
        }
        if (o.total_test_keyframes <= 0) o.total_test_keyframes = (int)cams.size();
    }
    gps_render_settings rs;
    std::memset(&rs, 0, sizeof(rs));
    rs.width = o.width;
    rs.height = o.height;
    rs.render_pass_count = o.render_pass_count;
    rs.supersampling_factor = o.supersampling_factor;
    rs.max_number_of_points = o.max_number_of_points;
    rs.fully_disable_hierarchical_culling = 0;
    rs.gaussians_per_block = o.gaussians_per_block;
    gps_pass* pass = nullptr;
    CHECK(gps_pass_create(&rs, scene, &pass));

    const int W = o.width, H = o.height;
    const size_t npx = (size_t)W * H;
    void* stream = nullptr;
    CHECK(gps_stream_create(&stream));
    void *color[2], *depth[2], *hist[2], *rgb;
    for (int k = 0; k < 2; ++k) {
        CHECK(gps_device_malloc(&color[k], npx * 12));
        CHECK(gps_device_malloc(&depth[k], npx * 4));
        CHECK(gps_device_malloc(&hist[k], npx * 12));
        CHECK(gps_memset_d8(color[k], 0, npx * 12, stream));
        CHECK(gps_memset_d8(depth[k], 0, npx * 4, stream));
        CHECK(gps_memset_d8(hist[k], 0, npx * 12, stream));
    }
    CHECK(gps_device_malloc(&rgb, npx * 3));
    float scene_m[16];
    scene_matrix(o, scene_m);
    float prev_vp[16];
    std::memset(prev_vp, 0, sizeof(prev_vp));
    int progressive_frames = 0;
    std::vector<double> times;
    std::vector<uint8_t> h_rgb(npx * 3);
    std::vector<float> h_col(npx * 3);
    // --run-tests over a camera path: main.cu:196-243
    const bool cam_tests = o.run_tests && !cams.empty();
    const int n_test = std::max(1, o.total_test_keyframes);
    const int passes_per_camera = o.frames_per_camera_keyframe + 1;
    int n_frames = o.frames;
    if (cam_tests) n_frames = passes_per_camera * n_test + 1;
    else if (o.play_camera_path && !cams.empty()) n_frames = std::max(o.frames, (int)cams.size() - 1);  // to the end
    int selected = 0, prev_selected = -1;
    std::vector<double> test_times;
    // --create-evaluation-screenshots (main.cu:72-120): per camera, the
    // progressive image after 1, 4, 16, 64, 256 and 1024 accumulated frames,
    // saved as <image_name>_sppNNNN.png under <dir>/eval/<model name>/; then
    // the next camera, up to camera max_eval_count or the last one.
    static const int kSppTargets[6] = {1, 4, 16, 64, 256, 1024};
    const bool eval = o.create_evaluation_screenshots;
    int eval_idx = 0;
    bool eval_next_camera = false;
    const std::string eval_dir = o.screenshot_dir + "/eval/" + model_name(o);
    if (eval) {
        if (cams.empty()) {
            std::fprintf(stderr, "--create-evaluation-screenshots needs cameras (--camera-path)\n");
            return 1;
        }
        // the reference's harness counts progressive frames: with TAA the
        // count is never reset per camera and with --accumulation-none it
        // stays 1, so its next target is never reached
        if (o.taa || o.accumulation_none) {
            std::fprintf(stderr, "--create-evaluation-screenshots needs progressive accumulation (no --taa, "
                                 "no --accumulation-none)\n");
            return 1;
        }
        if (!make_dirs(eval_dir)) {
            std::fprintf(stderr, "ERROR: Failed to ensure directory structure for screenshots.\n");
            return 1;
        }
        n_frames = 0x7FFFFFFF;
    }
    for (int fi = 1; fi <= n_frames; ++fi) {
        std::string save_name;
        if (eval) {
            int next_spp = progressive_frames + 1;
            if (eval_next_camera) {
                if (selected >= (int)cams.size() - 1 || selected >= o.max_eval_count) break;
                selected = selected + 1;  // change_camera_index(1); the accumulation restarts
                eval_idx = 0;
                eval_next_camera = false;
                next_spp = 1;
            }
            if (eval_idx < 6 && next_spp == kSppTargets[eval_idx]) {
                char suffix[16];
                std::snprintf(suffix, sizeof(suffix), "_spp%04d", next_spp);
                save_name = std::string(cams[(size_t)selected].img_name) + suffix;
                ++eval_idx;
            }
            if (eval_idx >= 6) eval_next_camera = true;
        }
        gps_frame_params f;
        bool moved = o.orbit_radius > 0;
        float fovy = o.fovy;
        if (!cams.empty()) {
            if (o.play_camera_path) {
                CHECK(gps_camera_path_frame(cams.data(), (int32_t)cams.size(), (float)fi / 24.0f,
                                            o.override_camera_settings ? o.fovy : 0.0f, W, H, o.supersampling_factor,
                                            o.near_plane, o.far_plane, &f, nullptr));
                moved = true;
            } else {
                CHECK(gps_frame_params_from_camera(&cams[(size_t)selected], W, H,
                                                   o.override_camera_settings ? o.fovy : 0.0f, o.supersampling_factor,
                                                   o.near_plane, o.far_plane, &f));
                moved = selected != prev_selected && prev_selected >= 0;
                prev_selected = selected;
            }
            fovy = f.fovy_degrees;
        } else {
            float eye[3] = {o.eye[0], o.eye[1], o.eye[2]};
            if (o.orbit_radius > 0) {
                const float az = (float)(fi - 1) * o.orbit_step * 0.01745329251994329576923690768489f;
                const float el = o.orbit_elevation * 0.01745329251994329576923690768489f;
                eye[0] = o.orbit_radius * std::cos(el) * std::sin(az);
                eye[1] = o.orbit_radius * std::sin(el);
                eye[2] = -o.orbit_radius * std::cos(el) * std::cos(az);
            }
            const float up[3] = {0, 1, 0};
            gps_frame_params_init(&f, eye, o.center, up, o.fovy, W, H, o.supersampling_factor, o.near_plane,
                                  o.far_plane);
        }
        float view_scene[16];
        gps_mat4_mul(f.view, scene_m, view_scene);
        std::memcpy(f.view, view_scene, 64);
        gps_frustum_planes(fovy, f.view, float(W) / H, o.near_plane, o.far_plane, f.frustum_planes);
        f.frame = fi;
        f.enable_occlusion_culling = o.enable_occlusion_culling;
        f.enable_hierarchical_culling = o.enable_hierarchical_culling;
        f.cull_small_gaussians = o.cull_small;
        f.reduce_point_count = o.reduce_point_count;
        for (int k = 0; k < 3; ++k) f.background[k] = o.background[k];
        // render_server.cu:86-160: pass -> (color[a], depth[b]); resolve.
        const int a = fi % 2, b = 1 - a;
        const auto t0 = std::chrono::steady_clock::now();
        CHECK(gps_pass_render(pass, &f, (float*)color[a], (float*)depth[b], stream));
        if (o.taa) {
            float M[16];
            gps_taa_reprojection(prev_vp, f.view, f.proj, M);
            ++progressive_frames;
            // The reference hands TAA depth[a] (written the frame before) as the
            // current depth (render_server.cu:92,131-136); --taa-current-depth
            // passes this frame's depth instead.
            const float* cur_depth = (const float*)(o.taa_current_depth ? depth[b] : depth[a]);
            CHECK(gps_resolve_taa((const float*)color[a], cur_depth, (const float*)hist[b], (const float*)depth[b], M,
                                  0.9f, W, H, progressive_frames, (float*)hist[a], (uint8_t*)rgb, stream));
        } else {
            progressive_frames = (moved || o.accumulation_none) ? 1 : progressive_frames + 1;
            CHECK(gps_resolve_progressive((float*)color[a], (float*)color[b], W, H, progressive_frames, (uint8_t*)rgb,
                                          stream));
        }
        CHECK(gps_stream_synchronize(stream));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        times.push_back(ms);
        gps_mat4_mul(f.proj, f.view, prev_vp);
        if (cam_tests) {
            const int frame_index = fi - 1;
            selected = std::min(selected + 1, (int)cams.size() - 1);  // change_camera_index(1)
            if (frame_index >= n_test) test_times.push_back(ms);
            if (frame_index % n_test == 0 && frame_index > 0) selected = 0;  // load_camera_state(0)
        }
        if (!save_name.empty()) {
            CHECK(gps_memcpy_d2h(h_rgb.data(), rgb, npx * 3, stream));
            const std::string path = eval_dir + "/" + save_name + ".png";
            std::printf("Saving screenshot to: %s\n", path.c_str());
            if (!write_png(path, h_rgb, W, H)) {
                std::fprintf(stderr, "ERROR: Failed to save screenshot to %s\n", path.c_str());
                return 1;
            }
        }
        if (!o.out.empty()) {
            CHECK(gps_memcpy_d2h(h_rgb.data(), rgb, npx * 3, stream));
            char name[64];
            std::snprintf(name, sizeof(name), "_%04d", fi);
            write_ppm(o.out + name + ".ppm", h_rgb, W, H);
            CHECK(gps_memcpy_d2h(h_col.data(), o.taa ? hist[a] : color[a], npx * 12, stream));
            write_pfm(o.out + name + ".pfm", h_col, W, H, o.taa ? 1.0f : 1.0f / (float)progressive_frames);
        }
        gps_stats st;
        CHECK(gps_pass_stats(pass, &st, stream));
        if (!o.run_tests && !eval)
            std::printf("frame %d: %.3f ms, points %llu + %llu, samples %llu\n", fi, ms,
                        (unsigned long long)st.num_points[0], (unsigned long long)st.num_points[1],
                        (unsigned long long)st.samples);
    }
    if (cam_tests) {
        // src/main.cu:211-241, per camera
        std::printf("---BEGIN_TIMINGS_JSON---\n{\n  \"algorithm\": \"GaussianPointSplatting\",\n");
        std::printf("  \"model_path\": \"%s%s\",\n  \"timings\": [\n", o.model_path.empty() ? "synthetic:" : "",
                    o.model_path.empty() ? o.synthetic.c_str() : o.model_path.c_str());
        for (int cam = 0; cam < n_test; ++cam) {
            std::printf("    { \"camera\": %d, \"samples\": [", cam);
            for (int rep = 0; rep < o.frames_per_camera_keyframe; ++rep) {
                const size_t idx = (size_t)cam + (size_t)rep * n_test;
                std::printf("%s%.4f", rep ? ", " : "", idx < test_times.size() ? test_times[idx] : 0.0);
            }
            std::printf("] }%s\n", cam + 1 < n_test ? "," : "");
        }
        std::printf("  ]\n}\n---END_TIMINGS_JSON---\n");
    } else if (o.run_tests) {
        // src/main.cu:211-241 envelope
        std::printf("---BEGIN_TIMINGS_JSON---\n{\n  \"algorithm\": \"GaussianPointSplatting\",\n");
        std::printf("  \"model_path\": \"%s%s\",\n  \"timings\": [\n    { \"camera\": 0, \"samples\": [",
                    o.model_path.empty() ? "synthetic:" : "",
                    o.model_path.empty() ? o.synthetic.c_str() : o.model_path.c_str());
        for (size_t i = 0; i < times.size(); ++i) std::printf("%s%.4f", i ? ", " : "", times[i]);
        std::printf("] }\n  ]\n}\n---END_TIMINGS_JSON---\n");
    }
    for (int k = 0; k < 2; ++k) {
        gps_device_free(color[k]);
        gps_device_free(depth[k]);
        gps_device_free(hist[k]);
    }
    gps_device_free(rgb);
    gps_stream_destroy(stream);
    gps_pass_destroy(pass);
    gps_scene_destroy(scene);
    return 0;
}

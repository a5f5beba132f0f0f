// reference_types.h — the reference types the RenderPass adapter touches,
// restated with the same field names and memory layout so the adapter in
// gps_render_pass.h compiles and runs here unchanged:
//   glm::vec3 / vec4 / ivec2 / mat4 (column-major, m[column][row])
//   Statistics, RenderSettings (the fields render() and the constructor
//   read)                         src/render_settings.cuh:23-125
//   CPUGaussian                   src/core/gaussians/gaussian_types.cuh:89-100
//   RenderPass                    src/core/rendering/render_pass.cuh:17-24
// TEST INFRASTRUCTURE: only tests/adapter/adapter_demo.cpp includes it.
#pragma once
#include <array>
#include <cstdint>

namespace glm {
struct vec3 {
    float x, y, z;
};
struct vec4 {
    float x, y, z, w;
    float& operator[](int i) { return (&x)[i]; }
    const float& operator[](int i) const { return (&x)[i]; }
};
struct ivec2 {
    int x, y;
};
struct mat4 {
    vec4 c[4];
    vec4& operator[](int i) { return c[i]; }
    const vec4& operator[](int i) const { return c[i]; }
};
}  // namespace glm

struct Statistics {
    float totalWeight = 0.0f;
    int numPoints = 0;
    int numPointsOne = 0;
    int numPointsTwo = 0;
    int numTilesOne = 0;
    int numTilesTwo = 0;
    int numGaussiansOne = 0;
    int numGaussiansTwo = 0;
};

struct RenderSettings {
    Statistics* statistics = nullptr;
    glm::ivec2 resolution = {1920, 1080};
    int render_pass_count = 1;
    int supersampling_factor = 2;
    int occlusion_culling_min_reduce_count = 0;
    int max_number_of_points = 250000000;
    bool enable_occlusion_culling = true;
    bool enable_hierarchical_culling = true;
    bool fully_disable_hierarchical_culling = false;
    bool cull_small_gaussians = false;
    bool use_unbiased_2d_splatting = true;
    bool reduce_point_count = true;
    glm::vec3 backgroundColor{0.0f, 0.0f, 0.0f};
    float near_view_plane = 0.2f;
    float far_view_plane = 10000.0f;
    float fovy = 45.0f;
    glm::mat4 sceneMatrix{};
    glm::mat4 cameraViewMatrix{};
    glm::mat4 cameraProjectionMatrix{};
    std::array<glm::vec4, 6> frustumPlanes{};
    int frame = 0;
};

struct CPUGaussian {
    glm::vec3 mean;
    float sh[48];
    float opacity;
    glm::vec3 scale;
    glm::vec4 rotation;
};

class RenderPass {
public:
    virtual void render(glm::vec3* d_image, float* d_depth, const RenderSettings& renderSettings) = 0;
    RenderPass() = default;
    virtual ~RenderPass() = default;
    bool enabled = true;
};

// gps_render_pass.h — drop-in for GaussianPointSplatting on MI355X: a
// RenderPass (src/core/rendering/render_pass.cuh:17-24) that owns a gps_scene
// and a gps_pass and forwards render() through the C ABI of include/gps.h.
// It replaces the class of gaussian_point_splatting.cuh:34-71, constructed at
// render_server.cu:48.
#pragma once
#include <cstring>
#include <stdexcept>
#include <vector>

#include "gps.h"

class GaussianPointSplattingMI355X : public RenderPass {
public:
    GaussianPointSplattingMI355X(const RenderSettings& s, const std::vector<CPUGaussian>& g, int sh_degree) {
        // CPUGaussian (gaussian_types.cuh:89-100) -> planar host arrays
        const size_t n = g.size(), nc = (size_t)(sh_degree + 1) * (sh_degree + 1);
        std::vector<float> means(3 * n), rots(4 * n), scales(3 * n), opac(n), shs(3 * nc * n);
        for (size_t i = 0; i < n; ++i) {
            std::memcpy(&means[3 * i], &g[i].mean, 12);
            std::memcpy(&rots[4 * i], &g[i].rotation, 16);  // the vec4 holds (w, x, y, z)
            std::memcpy(&scales[3 * i], &g[i].scale, 12);
            opac[i] = g[i].opacity;
            std::memcpy(&shs[3 * nc * i], g[i].sh, 12 * nc);  // coefficient-major, rgb interleaved
        }
        gps_scene_desc d{};
        d.num_gaussians = n;
        d.sh_degree = sh_degree;
        d.format = GPS_FORMAT_FP32;  // or GPS_FORMAT_HALF_FP8 / GPS_FORMAT_QUANTIZED
        d.means = means.data();
        d.rotations = rots.data();
        d.scales = scales.data();
        d.opacities = opac.data();
        d.shs = shs.data();
        check(gps_scene_create(&d, &scene_));
        gps_render_settings rs{};
        rs.width = s.resolution.x;
        rs.height = s.resolution.y;
        rs.render_pass_count = s.render_pass_count;
        rs.supersampling_factor = s.supersampling_factor;
        rs.max_number_of_points = s.max_number_of_points;
        rs.fully_disable_hierarchical_culling = s.fully_disable_hierarchical_culling;
        rs.gaussians_per_block = 0;  // reference rule: max(4, N / 1e6)
        check(gps_pass_create(&rs, scene_, &pass_));
    }
    ~GaussianPointSplattingMI355X() override {
        gps_pass_destroy(pass_);
        gps_scene_destroy(scene_);
    }

    void render(glm::vec3* d_image, float* d_depth, const RenderSettings& s) override {
        gps_frame_params f{};
        // cameraViewMatrix already holds the scene transform (main.cu:148);
        // render() uses it as it is (gaussian_point_splatting.cu:605)
        std::memcpy(f.view, &s.cameraViewMatrix[0][0], 64);  // column-major, as glm stores it
        std::memcpy(f.proj, &s.cameraProjectionMatrix[0][0], 64);
        std::memcpy(f.frustum_planes, s.frustumPlanes.data(), 96);  // host copy, not d_frustumPlanes
        f.fovy_degrees = s.fovy;
        f.near_plane = s.near_view_plane;
        f.far_plane = s.far_view_plane;
        f.background[0] = s.backgroundColor.x;
        f.background[1] = s.backgroundColor.y;
        f.background[2] = s.backgroundColor.z;
        f.width = s.resolution.x;
        f.height = s.resolution.y;
        f.supersampling_factor = s.supersampling_factor;
        f.frame = s.frame;
        f.enable_occlusion_culling = s.enable_occlusion_culling;
        f.enable_hierarchical_culling = s.enable_hierarchical_culling;
        f.cull_small_gaussians = s.cull_small_gaussians;
        f.use_unbiased_2d_splatting = s.use_unbiased_2d_splatting;
        f.reduce_point_count = s.reduce_point_count;
        f.occlusion_culling_min_reduce_count = s.occlusion_culling_min_reduce_count;
        f.samples_per_pixel = s.render_pass_count;
        check(gps_pass_render(pass_, &f, reinterpret_cast<float*>(d_image), d_depth, /*stream*/ nullptr));
        // settings.statistics (gaussian_point_splatting.cu:675-685): points
        // per pass, the pass's Gaussian count (the distributor's weight
        // count, i.e. the scene size) and totalWeight = the last pass's points
        gps_stats st;
        if (s.statistics && gps_pass_stats(pass_, &st, nullptr) == GPS_OK) {
            s.statistics->numPointsOne = (int)st.num_points[0];
            s.statistics->numGaussiansOne = (int)st.num_gaussians;
            if (st.passes == 2) {
                s.statistics->numPointsTwo = (int)st.num_points[1];
                s.statistics->numGaussiansTwo = (int)st.num_gaussians;
            }
            s.statistics->totalWeight = (float)st.num_points[st.passes - 1];
        }
    }

private:
    static void check(gps_status e) {
        if (e != GPS_OK) throw std::runtime_error(gps_last_error());
    }
    gps_scene* scene_ = nullptr;
    gps_pass* pass_ = nullptr;
};

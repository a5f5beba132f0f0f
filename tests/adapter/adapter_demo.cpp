// adapter_demo.cpp — drives the RenderPass adapter (gps_render_pass.h) the way
// the reference's RenderServer does (render_server.cu:48,90-93): constructs it
// from CPUGaussians and RenderSettings, renders into device colour / depth
// buffers, and writes the images and Statistics for the parity test.
//
// Input file (little endian): int32 n, sh_degree, width, height, spp, ss,
// frame, occlusion, hierarchical, unbiased; float view[16], proj[16],
// planes[24], fovy, near, far; then n CPUGaussian records (59 floats).
// Output file: float colour[H*W*3], depth[H*W]; then Statistics (8 x 4 B).
// TEST INFRASTRUCTURE.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "reference_types.h"
#include "gps_render_pass.h"

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: adapter_demo IN OUT\n");
        return 2;
    }
    FILE* fi = std::fopen(argv[1], "rb");
    if (!fi) return 2;
    int32_t h[10];
    float m[16 + 16 + 24 + 3];
    if (std::fread(h, 4, 10, fi) != 10 || std::fread(m, 4, 59, fi) != 59) return 2;
    std::vector<CPUGaussian> g((size_t)h[0]);
    if (std::fread(g.data(), sizeof(CPUGaussian), g.size(), fi) != g.size()) return 2;
    std::fclose(fi);
    static_assert(sizeof(CPUGaussian) == 59 * 4, "CPUGaussian layout");

    Statistics stats;
    RenderSettings s;
    s.statistics = &stats;
    s.resolution = {h[2], h[3]};
    s.render_pass_count = h[4];
    s.supersampling_factor = h[5];
    s.frame = h[6];
    s.enable_occlusion_culling = h[7] != 0;
    s.enable_hierarchical_culling = h[8] != 0;
    s.fully_disable_hierarchical_culling = h[8] == 0;
    s.use_unbiased_2d_splatting = h[9] != 0;
    std::memcpy(&s.cameraViewMatrix[0][0], m, 64);
    std::memcpy(&s.cameraProjectionMatrix[0][0], m + 16, 64);
    std::memcpy(s.frustumPlanes.data(), m + 32, 96);
    s.fovy = m[56];
    s.near_view_plane = m[57];
    s.far_view_plane = m[58];

    const size_t px = (size_t)h[2] * (size_t)h[3];
    glm::vec3* d_image = nullptr;
    float* d_depth = nullptr;
    if (hipMalloc((void**)&d_image, px * 12) != hipSuccess || hipMalloc((void**)&d_depth, px * 4) != hipSuccess) return 3;
    std::vector<float> color(px * 3), depth(px);
    try {
        GaussianPointSplattingMI355X pass(s, g, h[1]);
        pass.render(d_image, d_depth, s);
        if (hipDeviceSynchronize() != hipSuccess) return 3;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "adapter: %s\n", e.what());
        return 4;
    }
    if (hipMemcpy(color.data(), d_image, px * 12, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(depth.data(), d_depth, px * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return 3;
    FILE* fo = std::fopen(argv[2], "wb");
    if (!fo) return 2;
    std::fwrite(color.data(), 4, color.size(), fo);
    std::fwrite(depth.data(), 4, depth.size(), fo);
    std::fwrite(&stats, sizeof(stats), 1, fo);
    std::fclose(fo);
    (void)hipFree(d_image);
    (void)hipFree(d_depth);
    std::printf("adapter ok: points %d / %d, gaussians %d, total weight %.0f\n", stats.numPointsOne,
                stats.numPointsTwo, stats.numGaussiansOne, stats.totalWeight);
    return 0;
}

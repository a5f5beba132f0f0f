// Runs the renderer's project_gaussian (gps_project.h, the source the GPU
// kernels compile) on the CPU for Gaussians read from a file, so a test can
// compare it bit for bit with the oracle's projection.
//
// Input (binary): gps_frame_params, uint64 N, N x 10 float (centre xyz,
// quaternion wxyz, scale xyz).  Output (binary): N x 12 float: valid, view z,
// pixel mean x y, ca, cb, cc (dilated 2D covariance), NDC z, tile bounds
// minx miny maxx maxy.
#include <cstdio>
#include <vector>

#include "gps.h"
#include "gps_project.h"

namespace gps {
void make_frame_consts(const gps_frame_params* fp, int spp, FrameConsts* out);
}
using namespace gps;

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 2;
    gps_frame_params fp;
    uint64_t n = 0;
    if (fread(&fp, sizeof(fp), 1, in) != 1 || fread(&n, 8, 1, in) != 1) return 2;
    std::vector<float> a(n * 10);
    if (fread(a.data(), 4, a.size(), in) != a.size()) return 2;
    fclose(in);
    FrameConsts f;
    make_frame_consts(&fp, 1, &f);
    std::vector<float> out(n * 12, 0.0f);
    for (uint64_t i = 0; i < n; ++i) {
        const float* g = &a[10 * i];
        Projected pr;
        float* o = &out[12 * i];
        if (!project_gaussian(f, g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], g[8], g[9], pr)) continue;
        const float v[12] = {1.0f,  pr.view_z, pr.pmx, pr.pmy, pr.ca, pr.cb, pr.cc, pr.pz,
                             (float)pr.minx, (float)pr.miny, (float)pr.maxx, (float)pr.maxy};
        for (int k = 0; k < 12; ++k) o[k] = v[k];
    }
    FILE* fo = fopen(argv[2], "wb");
    if (!fo || fwrite(out.data(), 4, out.size(), fo) != out.size()) return 2;
    fclose(fo);
    return 0;
}
